// rvk_internal.h -- host-side state behind the opaque C-ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "rvk.h"
#include "rvk_device.h"
#include "rvk_geometry.h"

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
  int layout = 0;                // cell order: kCellLayoutPlain, kCellLayoutBrick4 or kCellLayoutPair8
  int ncb[3] = {0, 0, 0};        // cell-brick grid: ceil((dims+2)/brick edge)
  size_t n_cells = 0;            // storage in units of 8 voxels (whole cells incl. brick padding, or
                                 // x-pair bricks: 162 units each)
  void* frames[RVK_MAX_FRAMES] = {};    // cell volumes, n_cells * 8 voxels
  uint8_t enc[RVK_MAX_FRAMES] = {};     // cell encoding of each frame: kEnc* (see Cells, Half16)
  uint8_t* gradmag[RVK_MAX_FRAMES] = {};  // u8 cell volumes
  float gradmag_scale[RVK_MAX_FRAMES] = {};
  // Value range (min, max) of every brick of kRangeEdge^3 cells, for the DRR
  // march's empty-space skipping; nrb = ceil((dims + 2) / kRangeEdge).
  float2* ranges[RVK_MAX_FRAMES] = {};
  int nrb[3] = {0, 0, 0};
  // bumped whenever the frames' cells move (rvk_volume_relayout): graphs
  // captured before hold stale cell pointers and are not replayed
  unsigned generation = 0;
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
  // The cells lit for the gradient-magnitude path, where the Phong-Lambert
  // factor is 1: (alpha, sat((ambient + diffuse) r + specular), ... g, ... b),
  // computed on the device with the caster's own operations.
  float4* lit = nullptr;
  float4* klit = nullptr;
  // Opacity bitmaps (bit c = alpha[c] > 0), staged in LDS by the 2D caster
  // when they fit (n <= kTf2dOccMaxN); words = ceil(n*n / 32) each.
  uint32_t* occ = nullptr;
  uint32_t* kocc = nullptr;
  int occ_words = 0;
  // Column bitmaps (bit i = some g has alpha[g*n + i] > 0), col_words =
  // ceil(n / 32) each: the TF, then the keyhole TF if any.  A sample whose
  // intensity column is empty cannot composite, so its gradient is not
  // needed.  Always staged in LDS (n <= 2048: at most 512 B).
  uint32_t* col = nullptr;
  int col_words = 0;
  rvk_tf2d_params p{};
  // In/Ex logic (rvk_tf2d_set_inex).  The caster only compares the table
  // with 0 and 0.5, so it keeps per cell bit 0 = (inEx > 0 || alpha > 0)
  // and bit 1 = (inEx > 0.5): as bytes, and for n <= kTf2dOccMaxN also as
  // two bitmaps (occ_words each) staged in LDS.
  uint8_t* inex_flags = nullptr;
  uint32_t* inex_occ = nullptr;
  uint32_t* inex_col = nullptr;  // column bitmap of flag bit 0 (col_words)
  std::vector<float> alpha_host;  // the opacity table, kept for the flags
};

// Largest 2D TF whose opacity bitmap the caster stages in LDS (8 KiB; 16 KiB
// with the keyhole TF).
constexpr int kTf2dOccMaxN = 256;

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

// Cell encodings of a frame: x-pairs stored as (a, b), as (a, b - a), or as
// (a as binary16, b - a as int16) (Half16; 16-bit frames within +-2048), or
// with both as binary16 (HalfPair; also every difference exact in binary16).
constexpr int kEncPlain = 0;
constexpr int kEncDiff = 1;
constexpr int kEncHalf = 2;
constexpr int kEncHalfPair = 3;

// Calls f(Cells<T, L, D>{nullptr}) for a frame's voxel type, cell order and
// encoding: float cells always store x-pair differences, 16-bit cells when
// the frame's value range allows, 8- and 32-bit integer cells never.  With
// HALF = false the binary16 encodings are not dispatched (callers convert
// such frames first, see frame_drop_half).
template <int L, bool HALF, typename F> int dispatch_cells_t(int dtype, int enc, F&& f) {
  switch (dtype) {
    case RVK_U8: return f(Cells<uint8_t, L>{nullptr});
    case RVK_I8: return f(Cells<int8_t, L>{nullptr});
    case RVK_U16:
    case RVK_I16:
      if (enc == kEncHalf || enc == kEncHalfPair) {
        if constexpr (HALF) {
          if (enc == kEncHalfPair) return f(Cells<HalfPair, L, true>{nullptr});
          return f(Cells<Half16, L, true>{nullptr});
        } else {
          return fail(RVK_ERR_INVALID_ARGUMENT, "binary16 cells are not dispatched here");
        }
      }
      if (dtype == RVK_U16) return enc ? f(Cells<uint16_t, L, true>{nullptr}) : f(Cells<uint16_t, L>{nullptr});
      return enc ? f(Cells<int16_t, L, true>{nullptr}) : f(Cells<int16_t, L>{nullptr});
    case RVK_U32: return f(Cells<uint32_t, L>{nullptr});
    case RVK_I32: return f(Cells<int32_t, L>{nullptr});
    case RVK_F32: return f(Cells<float, L, true>{nullptr});
    default: return fail(RVK_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
  }
}
template <bool HALF = true, typename F> int dispatch_cells(int dtype, int layout, int enc, F&& f) {
  if (layout == kCellLayoutBrick4) return dispatch_cells_t<kCellLayoutBrick4, HALF>(dtype, enc, f);
  if (layout == kCellLayoutPair8) return dispatch_cells_t<kCellLayoutPair8, HALF>(dtype, enc, f);
  return dispatch_cells_t<kCellLayoutPlain, HALF>(dtype, enc, f);
}
// The cells of one frame of a volume.
template <bool HALF = true, typename F> int dispatch_frame(const rvk_volume* v, int frame, F&& f) {
  return dispatch_cells<HALF>(v->dtype, v->layout, v->enc[frame], f);
}
// Re-encodes a binary16 frame (Half16 or HalfPair cells) as int16
// difference cells (same values), on stream s; a no-op for other frames.
int frame_drop_half(const rvk_volume* v, int frame, hipStream_t s);
VolGeom vol_geom(const rvk_volume* v);
ViewGeom view_geom(const rvk_view* v);
float min_spacing(const rvk_volume* v);
int check_view(const rvk_view* v);
int check_frame(const rvk_volume* v, int frame);
int ensure_scratch(rvk_ctx* ctx, size_t bytes);
// Camera transform P V (rvk_geometry.cpp) and voxel -> world matrix.
int camera_pv(const rvk_camera* cam, double aspect, double pv[16]);
void voxels_matrix(const double origin[3], const double spacing[3], double vox[16]);

inline int check_march(const rvk_march_params* m) {
  if (!m) return fail(RVK_ERR_INVALID_ARGUMENT, "march params are NULL");
  if (!(m->step_scale > 0.0f) || !(m->step_scale < __builtin_huge_valf()))
    return fail(RVK_ERR_INVALID_ARGUMENT, "step_scale must be positive and finite");
  return RVK_OK;
}

// Workgroups for a caster launch over a view: whole super-tiles of
// 16x16-pixel tiles; `stiles_x` receives the number of super-tile columns.
inline int launch_grid(const rvk_view* v, int& stiles_x) {
  const int e = 16 << kSuperLog2;
  stiles_x = (v->width + e - 1) / e;
  return stiles_x * ((v->height + e - 1) / e) << (2 * kSuperLog2);
}

}  // namespace rvk
