// rvk_internal.h -- host-side state behind the opaque C-ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "rvk.h"
#include "rvk_device.h"

struct rvk_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;  // created by rvk_ctx_create, default priority
  hipStream_t stream = nullptr;      // the stream all work is issued on
  // small device workspace: work-queue counters and reduction partials
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
};

struct rvk_volume {
  rvk_ctx* ctx = nullptr;
  int dims[3] = {0, 0, 0};
  double spacing[3] = {1, 1, 1};
  int dtype = RVK_U8;
  int n_frames = 0;
  int ncb[3] = {0, 0, 0};        // cell-brick grid: ceil((dims+2)/brick edge)
  size_t n_cells = 0;            // cells incl. brick padding (each 8 voxels)
  void* frames[RVK_MAX_FRAMES] = {};    // cell volumes, n_cells * 8 voxels
  uint8_t* gradmag[RVK_MAX_FRAMES] = {};  // u8 cell volumes
  float gradmag_scale[RVK_MAX_FRAMES] = {};
};

struct rvk_tf1d {
  rvk_ctx* ctx = nullptr;
  int n = 0;
  float* tables = nullptr;  // device: alpha[n], galpha[n], r[n], g[n], b[n]
  rvk_tf1d_params p{};
};

// One 2D TF cell, 32 bytes: the alpha test needs only the first float4.
struct Tf2dCell {
  float alpha, ambient, diffuse, specular;
  float spec_power, r, g, b;
};

struct rvk_tf2d {
  rvk_ctx* ctx = nullptr;
  int n = 0;
  Tf2dCell* cells = nullptr;   // n*n
  Tf2dCell* kcells = nullptr;  // keyhole TF or null
  rvk_tf2d_params p{};
};

namespace rvk {

// error reporting (thread-local message)
int fail(int status, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
#define RVK_HIP(call)                                  \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return rvk::hip_fail(_e, #call); \
  } while (0)

size_t dtype_size(int dtype);
VolGeom vol_geom(const rvk_volume* v);
ViewGeom view_geom(const rvk_view* v);
float min_spacing(const rvk_volume* v);
int check_view(const rvk_view* v);
int check_frame(const rvk_volume* v, int frame);
int ensure_scratch(rvk_ctx* ctx, size_t bytes);

}  // namespace rvk
