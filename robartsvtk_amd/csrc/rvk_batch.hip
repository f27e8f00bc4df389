// rvk_batch.hip -- batched multi-view DRR, NCC registration loss and the
// hand-written DRR adjoint (pose and volume gradients).
//
// No counterpart in the reference: it renders one view per Render() call and
// has no derivatives.  The forward march is the reference DRR caster
// (V/CUDA_vtkCudaDRRImageVolumeMapper_renderAlgo.cuh:11-117) shared with the
// single-view kernel through rvk_march.h.
//
// Work unit: an 8x8-pixel tile of one view, processed by one wave64.  Tiles
// are numbered view by view; inside a view, four consecutive tiles form a
// 16x16 block (same ray-offset phase as the single-view kernel).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rvk_internal.h"
#include "rvk_march.h"
#include "rvk_pose_math.h"

namespace rvk {

struct BatchView {
  ViewGeom v;
  float dm[96];      // d(view_to_voxels)/d pose_i, i = 0..5 (baked, row major)
  float inv[16];     // inverse of view_to_voxels: voxel position -> (vx, vy, z, 1) * w
  long long pix_off; // first pixel of this view in the packed image buffer
  int tiles16_x;
  int pad;
};

struct BatchArgs {
  VolGeom g;
  DrrMarchArgs m;
  float step_len;
  const void* vol;
  const BatchView* views;
  const int* item_begin;  // n_views + 1 prefix sums of 8x8 tiles
  int n_views;
  int n_items;
  float* out;
  float* jac;             // 6 x total pixels (SoA) d L_p / d pose_i, or NULL
  long long total;        // pixels in the packed images
  unsigned* queues;       // kNumXcd work-queue heads (persistent schedule)
  int shape;              // wave tile of the forward kernels (RVK_TILE_*)
  const int2* work;       // forward work order: (view, tile of the view) per work item
};

// Accumulates S = sum grad mu and T = sum (j+t) grad mu over the accumulated
// samples of one ray; A = ds*slope*S and B = ds*slope*T (see the pose adjoint
// below) are formed once per ray by pose_sums().
struct PoseVisit {
  static constexpr bool kFlatRuns = true;  // a zero gradient adds nothing
  float j;
  V3<float>& S;
  V3<float>& T;
  __device__ __forceinline__ void operator()(const V3<float>&, float t, float, const V3<float>& gr) const {
    const float w = j + t;
    S.x += gr.x; S.y += gr.y; S.z += gr.z;
    T.x = fmaf(w, gr.x, T.x); T.y = fmaf(w, gr.y, T.y); T.z = fmaf(w, gr.z, T.z);
  }
};
// Scales the sums by ncs = ds*slope; called only for rays that accumulated
// a sample (a ray that misses the volume can have a non-finite ds).
__device__ __forceinline__ void pose_sums(float ncs, V3<float>& A, V3<float>& B) {
  A.x *= ncs; A.y *= ncs; A.z *= ncs;
  B.x *= ncs; B.y *= ncs; B.z *= ncs;
}

__device__ __forceinline__ int find_view(const int* __restrict__ ib, int n, int item) {
  int lo = 0, hi = n;  // ib[lo] <= item < ib[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (ib[mid] <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void item_pixel(int tiles16_x, int local, int lane, int& x, int& y) {
  const int macro = local >> 2, q = local & 3;
  x = (macro % tiles16_x) * 16 + (q & 1) * 8 + (lane & 7);
  y = (macro / tiles16_x) * 16 + (q >> 1) * 8 + (lane >> 3);
}
// The forward kernels' wave tile (the autotuner's choice; the reductions
// keep the 8x8 mapping above, so their sums do not depend on it): the four
// items of a 16x16 block as 8x8 quadrants, 16x4 rows or 4x16 columns.
__device__ __forceinline__ void item_pixel_shaped(int shape, int tiles16_x, int local, int lane, int& x, int& y) {
  const int macro = local >> 2, q = local & 3;
  const int bx = (macro % tiles16_x) * 16, by = (macro / tiles16_x) * 16;
  if (shape == RVK_TILE_16X4) {
    x = bx + (lane & 15);
    y = by + q * 4 + (lane >> 4);
  } else if (shape == RVK_TILE_4X16) {
    x = bx + q * 4 + (lane & 3);
    y = by + (lane >> 2);
  } else {
    x = bx + (q & 1) * 8 + (lane & 7);
    y = by + (q >> 1) * 8 + (lane >> 3);
  }
}

// One pixel of one view.  With JAC the march also gathers trilinear gradients
// and the pixel's pose Jacobian dL/dpose_i = A . dstart_i + B . dinc_i is
// written next to the image (the registration step's single march).
template <typename CV, bool SKIP, bool PRECISE, bool JAC, bool TERM>
__device__ __forceinline__ void forward_item(const BatchArgs& a, int item, int lane) {
  // the forward writes only its own pixels, so it may take the tiles in any
  // order (the reductions keep the view-major one)
  const int2 vt = a.work[item];
  const int v = __builtin_amdgcn_readfirstlane(vt.x);
  const BatchView& bv = a.views[v];
  int x, y;
  item_pixel_shaped(a.shape, bv.tiles16_x, __builtin_amdgcn_readfirstlane(vt.y), lane, x, y);
  if (x >= bv.v.W || y >= bv.v.H) return;
  const Ray r = form_ray(a.g, bv.v, a.step_len, x, y);
  const long long pi = bv.pix_off + (long long)y * bv.v.W + x;
  if (!JAC) {
    a.out[pi] = drr_march<CV, SKIP, PRECISE, false, true, true, TERM>(cells_of<CV>(a.vol), a.g, a.m, r, x, y);
    return;
  }
  const float ds = sqrtf(r.inc.x * r.inc.x * a.g.sp[0] * a.g.sp[0] + r.inc.y * r.inc.y * a.g.sp[1] * a.g.sp[1] +
                         r.inc.z * r.inc.z * a.g.sp[2] * a.g.sp[2]);
  V3<float> A{0, 0, 0}, B{0, 0, 0};
  PoseVisit visit{ray_offset(a.m.jitter, x, y), A, B};
  a.out[pi] = drr_march<CV, SKIP, true, true, true, true, TERM>(cells_of<CV>(a.vol), a.g, a.m, r, x, y, visit);
  const bool any = A.x != 0.0f || A.y != 0.0f || A.z != 0.0f || B.x != 0.0f || B.y != 0.0f || B.z != 0.0f;
  if (any) pose_sums(ds * a.m.slope, A, B);
#pragma unroll 1
  for (int i = 0; i < 6; ++i) {
    float jv = 0.0f;
    if (any) {
      const RayTangent tg = form_ray_tangent(a.g, bv.v, bv.dm + 16 * i, a.step_len, x, y);
      jv = A.x * tg.dstart.x + A.y * tg.dstart.y + A.z * tg.dstart.z + B.x * tg.dinc.x + B.y * tg.dinc.y +
           B.z * tg.dinc.z;
    }
    a.jac[i * a.total + pi] = jv;
  }
}

// One wave per tile; workgroups XCD-remapped so neighbouring tiles share an L2.
template <typename CV, bool SKIP, bool PRECISE, bool JAC, bool TERM>
__global__ __launch_bounds__(256) void batch_grid_kernel(BatchArgs a) {
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int item = blk * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (item >= a.n_items) return;
  forward_item<CV, SKIP, PRECISE, JAC, TERM>(a, item, threadIdx.x & 63);
}

// Persistent: the tile range is cut into kNumXcd contiguous queues.  A wave
// drains the queue of the XCD its workgroup was dispatched to (blockIdx mod
// kNumXcd under round-robin dispatch), then the others in turn; every queue
// is finite, so every wave reaches the end.
template <typename CV, bool SKIP, bool PRECISE, bool JAC, bool TERM>
__global__ __launch_bounds__(256) void batch_persistent_kernel(BatchArgs a) {
  const int lane = threadIdx.x & 63;
  const int home = blockIdx.x % kNumXcd;
  for (int r = 0; r < kNumXcd; ++r) {
    const int c = (home + r) % kNumXcd;
    const int begin = (int)((long long)a.n_items * c / kNumXcd);
    const int end = (int)((long long)a.n_items * (c + 1) / kNumXcd);
    while (true) {
      int w = 0;
      if (lane == 0) w = (int)atomicAdd(&a.queues[c], 1u);
      w = __builtin_amdgcn_readfirstlane(w);
      const int item = begin + w;
      if (item >= end) break;
      forward_item<CV, SKIP, PRECISE, JAC, TERM>(a, item, lane);
    }
  }
}

// Queue heads are cleared by a kernel (not a memset node) so the sequence can
// live in a replayed hipGraph.
__global__ void reset_queues_kernel(unsigned* q) {
  if (threadIdx.x < kNumXcd) q[threadIdx.x] = 0u;
}

// ---------------------------------------------------------------------------
// NCC per view, zero-mean (Pearson) or the reference's uncentred
// sum(I T) / (sqrt(sum I^2) sqrt(sum T^2))
// (ImageProcessing/vtkRobartsRegistration/vtkImageNormalizedCrossCorrelation.cxx:116-128,203-218).
// Per tile: raw moments in double; per view: combine.
struct NccArgs {
  const BatchView* views;
  const int* item_begin;
  int n_views, n_items;
  int centred;
  const float* img;
  const float* tgt;
  double* part;    // n_items x 5
  double* stats;   // n_views x 8: mean_I, mean_T, inv_sqrt(Saa*Sbb), ncc/Saa, ncc, n, valid
  float* loss;     // n_views
  float* grad_img; // optional packed dloss/dI
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ float wave_sumf(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__global__ __launch_bounds__(256) void ncc_partial_kernel(NccArgs a) {
  const int item = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (item >= a.n_items) return;
  const int lane = threadIdx.x & 63;
  const int v = __builtin_amdgcn_readfirstlane(find_view(a.item_begin, a.n_views, item));
  const BatchView& bv = a.views[v];
  int x, y;
  item_pixel(bv.tiles16_x, item - a.item_begin[v], lane, x, y);
  double s[5] = {0, 0, 0, 0, 0};
  if (x < bv.v.W && y < bv.v.H) {
    const long long i = bv.pix_off + (long long)y * bv.v.W + x;
    const double I = a.img[i], T = a.tgt[i];
    s[0] = I; s[1] = T; s[2] = I * I; s[3] = T * T; s[4] = I * T;
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) s[k] = wave_sum(s[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 5; ++k) a.part[(long long)item * 5 + k] = s[k];
}

// One workgroup per view; a fixed reduction order keeps results reproducible.
__global__ __launch_bounds__(256) void ncc_view_kernel(NccArgs a) {
  const int v = blockIdx.x;
  __shared__ double red[5][256];
  double s[5] = {0, 0, 0, 0, 0};
  for (int it = a.item_begin[v] + threadIdx.x; it < a.item_begin[v + 1]; it += blockDim.x)
#pragma unroll
    for (int k = 0; k < 5; ++k) s[k] += a.part[(long long)it * 5 + k];
#pragma unroll
  for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double n = (double)a.views[v].v.W * (double)a.views[v].v.H;
    // the uncentred form is the centred one with both means taken as 0
    const double mi = a.centred ? red[0][0] / n : 0.0, mt = a.centred ? red[1][0] / n : 0.0;
    const double saa = red[2][0] - n * mi * mi, sbb = red[3][0] - n * mt * mt, sab = red[4][0] - n * mi * mt;
    const bool ok = saa > 0.0 && sbb > 0.0;
    const double inv = !ok ? 0.0 : a.centred ? 1.0 / sqrt(saa * sbb) : 1.0 / (sqrt(saa) * sqrt(sbb));
    const double ncc = ok ? sab * inv : 0.0;
    double* st = a.stats + 8 * v;
    st[0] = mi; st[1] = mt; st[2] = inv; st[3] = ok ? ncc / saa : 0.0; st[4] = ncc; st[5] = n; st[6] = ok;
    a.loss[v] = (float)(1.0 - ncc);
  }
}

// d loss / d I_p = -( (T_p - mean_T) / sqrt(Saa Sbb) - ncc (I_p - mean_I) / Saa )
__device__ __forceinline__ float dloss_dimage(const double* st, float I, float T) {
  return (float)(-((double)T - st[1]) * st[2] + st[3] * ((double)I - st[0]));
}

__global__ __launch_bounds__(256) void ncc_grad_image_kernel(NccArgs a) {
  const int item = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (item >= a.n_items) return;
  const int lane = threadIdx.x & 63;
  const int v = __builtin_amdgcn_readfirstlane(find_view(a.item_begin, a.n_views, item));
  const BatchView& bv = a.views[v];
  int x, y;
  item_pixel(bv.tiles16_x, item - a.item_begin[v], lane, x, y);
  if (x >= bv.v.W || y >= bv.v.H) return;
  const long long i = bv.pix_off + (long long)y * bv.v.W + x;
  a.grad_img[i] = dloss_dimage(a.stats + 8 * v, a.img[i], a.tgt[i]);
}

// ---------------------------------------------------------------------------
// Pose adjoint.  For one pixel, L = sum_m ds*mu(p_m) over accumulated samples
// p_m = start + (j + t_m) inc, so
//   dL/dpose_i = A . dstart_i + B . dinc_i,
//   A = sum_m ds*slope*grad mu(p_m),   B = sum_m (j + t_m) ds*slope*grad mu(p_m),
// with (dstart_i, dinc_i) the forward-mode tangent of the ray set-up along
// d(view_to_voxels)/d pose_i.  The ray is re-marched (PRECISE positions) with
// the forward's control flow; nothing is stored per sample.
struct AdjArgs {
  BatchArgs b;
  const float* img;   // rendered images (packed)
  const float* dldi;  // d loss / d image, or NULL: formed from the NCC state
  const double* stats;
  const float* tgt;
  float* part;  // n_items x 6
};

__device__ __forceinline__ float dloss_at(const float* dldi, const double* stats, const float* img, const float* tgt,
                                          int v, long long pi) {
  return dldi ? dldi[pi] : dloss_dimage(stats + 8 * v, img[pi], tgt[pi]);
}

template <typename CV, bool SKIP>
__global__ __launch_bounds__(256) void adjoint_pose_kernel(AdjArgs a) {
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int item = blk * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (item >= a.b.n_items) return;
  const int lane = threadIdx.x & 63;
  const int v = __builtin_amdgcn_readfirstlane(find_view(a.b.item_begin, a.b.n_views, item));
  const BatchView& bv = a.b.views[v];
  int x, y;
  item_pixel(bv.tiles16_x, item - a.b.item_begin[v], lane, x, y);
  float c[6] = {0, 0, 0, 0, 0, 0};
  if (x < bv.v.W && y < bv.v.H) {
    const VolGeom& g = a.b.g;
    const Ray r = form_ray(g, bv.v, a.b.step_len, x, y);
    const float ds = sqrtf(r.inc.x * r.inc.x * g.sp[0] * g.sp[0] + r.inc.y * r.inc.y * g.sp[1] * g.sp[1] +
                           r.inc.z * r.inc.z * g.sp[2] * g.sp[2]);
    V3<float> A{0, 0, 0}, B{0, 0, 0};
    const float jt = ray_offset(a.b.m.jitter, x, y);
    PoseVisit visit{jt, A, B};
    const float I = drr_march<CV, SKIP, true, true, true>(cells_of<CV>(a.b.vol), g, a.b.m, r, x, y, visit);
    const long long pi = bv.pix_off + (long long)y * bv.v.W + x;
    const float dLdI = dloss_at(a.dldi, a.stats, a.img, a.tgt, v, pi);
    // dI/dL = -I for exp(-L) output, 1 for L output
    const float coef = (a.b.m.output == RVK_DRR_F32_EXP_NEG_L) ? -dLdI * I : dLdI;
    if (coef != 0.0f && (A.x != 0.0f || A.y != 0.0f || A.z != 0.0f)) {
      pose_sums(ds * a.b.m.slope, A, B);
#pragma unroll 1
      for (int i = 0; i < 6; ++i) {
        const RayTangent tg = form_ray_tangent(g, bv.v, bv.dm + 16 * i, a.b.step_len, x, y);
        c[i] = coef * (A.x * tg.dstart.x + A.y * tg.dstart.y + A.z * tg.dstart.z + B.x * tg.dinc.x +
                       B.y * tg.dinc.y + B.z * tg.dinc.z);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) c[i] = wave_sumf(c[i]);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < 6; ++i) a.part[(long long)item * 6 + i] = c[i];
}

__global__ __launch_bounds__(256) void pose_reduce_kernel(const int* item_begin, const float* part, float* grad) {
  const int v = blockIdx.x;
  __shared__ double red[6][256];
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int it = item_begin[v] + threadIdx.x; it < item_begin[v + 1]; it += blockDim.x)
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] += part[(long long)it * 6 + k];
#pragma unroll
  for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 6) grad[6 * v + threadIdx.x] = (float)red[threadIdx.x][0];
}

// d loss / d pose from the stored per-pixel Jacobians (registration step).
struct JacGradArgs {
  const BatchView* views;
  const int* item_begin;
  int n_views, n_items;
  const float* img;
  const float* tgt;
  const float* jac;
  long long total;
  const double* stats;
  int output;
  float* part;
};

__global__ __launch_bounds__(256) void jac_grad_kernel(JacGradArgs a) {
  const int item = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (item >= a.n_items) return;
  const int lane = threadIdx.x & 63;
  const int v = __builtin_amdgcn_readfirstlane(find_view(a.item_begin, a.n_views, item));
  const BatchView& bv = a.views[v];
  int x, y;
  item_pixel(bv.tiles16_x, item - a.item_begin[v], lane, x, y);
  float c[6] = {0, 0, 0, 0, 0, 0};
  if (x < bv.v.W && y < bv.v.H) {
    const long long pi = bv.pix_off + (long long)y * bv.v.W + x;
    const float I = a.img[pi];
    const float dLdI = dloss_dimage(a.stats + 8 * v, I, a.tgt[pi]);
    const float coef = (a.output == RVK_DRR_F32_EXP_NEG_L) ? -dLdI * I : dLdI;
#pragma unroll
    for (int i = 0; i < 6; ++i) c[i] = coef * a.jac[i * a.total + pi];
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) c[i] = wave_sumf(c[i]);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < 6; ++i) a.part[(long long)item * 6 + i] = c[i];
}

// ---------------------------------------------------------------------------
// Volume adjoint (any voxel type, plain march): per pixel a ray record, then
// a brick-driven scatter.  d loss / d v_ijk (v the stored voxel value,
// alpha = cs v + ci) = sum_views sum_pixels coef_p *
// sum_m [alpha(p_m) < 0] w_ijk(p_m), w the trilinear weight of voxel ijk at
// the sample and coef_p = d loss / d L_p * ds * slope.  One workgroup owns an
// 8^3-voxel brick: it marches every ray that can meet the cells its voxels
// weigh in over that stretch only, and each accumulated sample adds its 8
// corner contributions to the brick's accumulators in LDS.  The accumulators
// are 64-bit fixed point (contributions rounded to multiples of 2^-k, with
// 2^k set per call from the largest |coef| so that one contribution stays
// below 2^30): integer sums do not depend on the order of the adds, so the
// result is the same whichever wave adds first, and LDS integer atomics are
// far cheaper than float ones.
struct alignas(16) RayRec {
  float p0x, p0y, p0z;  // start in cell coordinates (p0 + 0.5)
  float coef;           // d loss / d L_p * ds * slope
  float ix, iy, iz;     // step
  int n;                // samples
  float cs, ci;         // alpha = cs * value + ci
  float kw0, kw1;       // keyhole window: samples t in [kw0, kw1) are never accumulated
};
static_assert(sizeof(RayRec) == 48, "RayRec is three 16-byte vectors");

struct VolGradArgs {
  BatchArgs b;
  const float* img;
  const float* dldi;
  RayRec* rays;      // packed like the images
  float* coef_max;   // per 8x8 tile: the largest |coef| of its rays (+inf if one is not finite)
  int* scale_log2;   // k of the fixed-point accumulation, or kVgNonFinite
  float* grad;       // dims x-fastest
};

constexpr int kVgNonFinite = 1 << 30;

__global__ __launch_bounds__(256) void ray_record_kernel(VolGradArgs a) {
  const int item = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (item >= a.b.n_items) return;
  const int lane = threadIdx.x & 63;
  const int v = __builtin_amdgcn_readfirstlane(find_view(a.b.item_begin, a.b.n_views, item));
  const BatchView& bv = a.b.views[v];
  int x, y;
  item_pixel(bv.tiles16_x, item - a.b.item_begin[v], lane, x, y);
  float amax = 0.0f;
  if (x < bv.v.W && y < bv.v.H) {
    const VolGeom& g = a.b.g;
    const Ray r = form_ray(g, bv.v, a.b.step_len, x, y);
    const float j = ray_offset(a.b.m.jitter, x, y);
    const long long pi = bv.pix_off + (long long)y * bv.v.W + x;
    const float I = a.img[pi];
    const float dLdI = a.dldi[pi];
    const float ds = ray_length(g, r.inc);
    RayRec rr;
    rr.p0x = (r.start.x + j * r.inc.x) + 0.5f;
    rr.p0y = (r.start.y + j * r.inc.y) + 0.5f;
    rr.p0z = (r.start.z + j * r.inc.z) + 0.5f;
    rr.ix = r.inc.x; rr.iy = r.inc.y; rr.iz = r.inc.z;
    rr.n = step_count(r.num_steps - j);
    // the rays the precise forward march takes no samples of (drr_march)
    if (!samples_in_cells(g, V3<float>{rr.p0x, rr.p0y, rr.p0z}, r.inc, rr.n)) rr.n = 0;
    // the keyhole window exactly as drr_march forms it (t = kst - k)
    const int es = rr.n - (int)ceilf(r.exclude_start), ee = rr.n - (int)floorf(r.exclude_end);
    rr.kw0 = (float)(rr.n - es);
    rr.kw1 = (float)(rr.n - ee);
    rr.ci = a.b.m.intercept * -ds;
    rr.cs = a.b.m.slope * -ds;
    // d loss / d mu at a sample = (d loss / d L) * ds * slope, with dL/dI = -I for exp(-L)
    const float dLdL = (a.b.m.output == RVK_DRR_F32_EXP_NEG_L) ? -dLdI * I : dLdI;
    rr.coef = dLdL * ds * a.b.m.slope;
    a.rays[pi] = rr;
    // rays without samples add nothing; a non-finite coefficient poisons the call
    if (rr.n > 0) amax = isfinite(rr.coef) ? fabsf(rr.coef) : __builtin_huge_valf();
  }
  for (int o = 32; o > 0; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o, 64));
  if (lane == 0) a.coef_max[item] = amax;
}

// k = 30 - e for the largest |coef| = f 2^e (f in [0.5, 1)): every scaled
// contribution |coef w 2^k| (w <= 1) is below 2^30 and rounds into an int32.
__global__ __launch_bounds__(1024) void vg_scale_kernel(const float* __restrict__ coef_max, int n, int* scale_log2) {
  __shared__ float s_m[16];
  float m = 0.0f;
  for (int i = threadIdx.x; i < n; i += 1024) m = fmaxf(m, coef_max[i]);
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 16; ++w) m = fmaxf(m, s_m[w]);
    int e = 0;
    if (m > 0.0f) frexpf(m, &e);
    scale_log2[0] = isfinite(m) ? 30 - e : kVgNonFinite;
  }
}

// The LDS cube holds voxels [8b - 2, 8b + 9] per axis.  A sample's lower
// corner index, clamped into [0, 10], is the true one whenever either corner
// is a brick voxel; otherwise both land in the halo, which is dropped, except
// that at the volume's faces the clamped edge voxels are folded in (voxel -1
// holds voxel 0, voxels n and n + 1 hold voxel n - 1, as the edge cells do).
constexpr int kVgEdge = 8;
constexpr int kVgL = kVgEdge + 4;
constexpr int kVgCube = kVgL * kVgL * kVgL;
constexpr int kVgViews = 64;  // view boxes held in LDS at a time

__device__ __forceinline__ void vg_add(unsigned long long* s, float v) {
  const long long q = (long long)(int)__builtin_rintf(v);
  __hip_atomic_fetch_add(s, (unsigned long long)q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <typename CV>
__global__ __launch_bounds__(256) void volume_grad_kernel(VolGradArgs a) {
  __shared__ unsigned long long s_acc[kVgCube];
  __shared__ int s_box[kVgViews][3];
  __shared__ long long s_pre[kVgViews + 1];  // box ray counts can pass 2^31 (whole images)
  const VolGeom& g = a.b.g;
  const int nb[3] = {(g.n[0] + kVgEdge - 1) / kVgEdge, (g.n[1] + kVgEdge - 1) / kVgEdge,
                     (g.n[2] + kVgEdge - 1) / kVgEdge};
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  const int bi[3] = {bid % nb[0], (bid / nb[0]) % nb[1], bid / (nb[0] * nb[1])};
  const int o[3] = {bi[0] * kVgEdge, bi[1] * kVgEdge, bi[2] * kVgEdge};
  const int tid = threadIdx.x;
  const int k = a.scale_log2[0];
  const CV vol = cells_of<CV>(a.b.vol);
  // With mu = value (slope > 0, intercept 0) a sample is accumulated exactly
  // when its trilinear value is > 0; the samples weighing in this brick
  // interpolate voxels [8b - 1, 8b + 8], so if all of those are <= 0 none is.
  // (A voxel is read as the trilinear value at its own cell corner, weights
  // exactly 0, which is its value in every cell encoding.)
  bool live = true;
  if (a.b.m.intercept == 0.0f && a.b.m.slope > 0.0f) {
    float mx = -1.0f;
    for (int e = tid; e < 1000; e += 256) {
      const int xi = min(max(o[0] - 1 + e % 10, 0), g.n[0] - 1);
      const int yi = min(max(o[1] - 1 + (e / 10) % 10, 0), g.n[1] - 1);
      const int zi = min(max(o[2] - 1 + e / 100, 0), g.n[2] - 1);
      mx = fmaxf(mx, trilinear<false>(vol, g, (float)(xi + 1), (float)(yi + 1), (float)(zi + 1)));
    }
    live = __syncthreads_or(mx > 0.0f);
  }
  if (live && k != kVgNonFinite) {
    for (int e = tid; e < kVgCube; e += 256) s_acc[e] = 0ull;
    // the cells whose samples weigh in the brick's voxels: (8b, 8b + 9) per
    // axis, reaching below c = 1 and above c = n at the faces
    float clo[3], chi[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      clo[ax] = bi[ax] == 0 ? -1.0f : (float)o[ax];
      chi[ax] = bi[ax] == nb[ax] - 1 ? (float)(g.n[ax] + 2) : (float)(o[ax] + kVgEdge + 1);
    }
    for (int v0 = 0; v0 < a.b.n_views; v0 += kVgViews) {
      const int nv = min(kVgViews, a.b.n_views - v0);
      __syncthreads();
      if (tid < nv) {
        // pixel box of the region's 8 corners (sample coordinates p = c -
        // 0.5), a pixel of margin beyond the rays that can meet it
        const BatchView& bv = a.b.views[v0 + tid];
        const int W = bv.v.W, H = bv.v.H;
        float xmin = 1e30f, xmax = -1e30f, ymin = 1e30f, ymax = -1e30f;
        bool behind = false;
        for (int c = 0; c < 8; ++c) {
          const float px = ((c & 1) ? chi[0] : clo[0]) - 0.5f, py = ((c & 2) ? chi[1] : clo[1]) - 0.5f;
          const float pz = ((c & 4) ? chi[2] : clo[2]) - 0.5f;
          const float* M = bv.inv;
          const float hx = M[0] * px + M[1] * py + M[2] * pz + M[3];
          const float hy = M[4] * px + M[5] * py + M[6] * pz + M[7];
          const float hw = M[12] * px + M[13] * py + M[14] * pz + M[15];
          if (!(hw > 0.0f)) { behind = true; break; }
          const float vx = hx / hw, vy = hy / hw;
          const float xp = bv.v.flipped ? vx * (float)W : (1.0f - vx) * (float)W;
          const float yp = vy * (float)H;
          xmin = fminf(xmin, xp); xmax = fmaxf(xmax, xp);
          ymin = fminf(ymin, yp); ymax = fmaxf(ymax, yp);
        }
        int x0 = 0, x1 = W - 1, y0 = 0, y1 = H - 1;
        if (!behind) {
          x0 = max(0, (int)floorf(xmin) - 1); x1 = min(W - 1, (int)ceilf(xmax) + 1);
          y0 = max(0, (int)floorf(ymin) - 1); y1 = min(H - 1, (int)ceilf(ymax) + 1);
        }
        const int w = max(0, x1 - x0 + 1), h = max(0, y1 - y0 + 1);
        s_box[tid][0] = x0;
        s_box[tid][1] = y0;
        s_box[tid][2] = w;
        s_pre[tid + 1] = (long long)w * h;
      }
      __syncthreads();
      if (tid == 0) {
        s_pre[0] = 0;
        for (int v = 0; v < nv; ++v) s_pre[v + 1] += s_pre[v];
      }
      __syncthreads();
      // the boxes' rays, view after view, one per thread
      const long long total = s_pre[nv];
      int cv = 0;
      for (long long r = tid; r < total; r += 256) {
        while (r >= s_pre[cv + 1]) ++cv;
        const int local = (int)(r - s_pre[cv]);
        const int bw = s_box[cv][2];
        const int ly = local / bw;
        const int px = s_box[cv][0] + (local - ly * bw), py = s_box[cv][1] + ly;
        const BatchView& bv = a.b.views[v0 + cv];
        const float4* rec = reinterpret_cast<const float4*>(a.rays + bv.pix_off);
        const long long q = 3 * ((long long)py * bv.v.W + px);
        const float4 A = rec[q], B = rec[q + 1], K = rec[q + 2];
        const int n = __float_as_int(B.w);
        if (n <= 0 || A.w == 0.0f) continue;
        // sample indices m with c0 + m inc inside the open region: the entry
        // and exit times from v_rcp_f32 reciprocals differ from the
        // quotients by a few ulp, far less than the one-sample margin each
        // side, and a sample outside adds only to the halo
        float tlo = 0.0f, thi = (float)(n - 1);
        const float p0[3] = {A.x, A.y, A.z}, in[3] = {B.x, B.y, B.z};
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
          if (in[ax] != 0.0f) {
            const float rin = __builtin_amdgcn_rcpf(in[ax]);
            const float ta = (clo[ax] - p0[ax]) * rin, tb = (chi[ax] - p0[ax]) * rin;
            tlo = fmaxf(tlo, fminf(ta, tb));
            thi = fminf(thi, fmaxf(ta, tb));
          } else if (p0[ax] <= clo[ax] || p0[ax] >= chi[ax]) {
            thi = -1.0f;
          }
        }
        const int m0 = max(0, (int)floorf(tlo) - 1), m1 = min(n - 1, (int)ceilf(thi) + 1);
        const float coef = ldexpf(A.w, k);
        // each lane starts its stretch at its own sample (3 x its lane, mod
        // the length) and wraps round: a wave's rays enter the region
        // together, and marched in step they would add to the same voxels in
        // the same instructions (LDS address conflicts); the sum is the same
        // in any order
        const int cnt = m1 - m0 + 1;
        int mm = m0 + (cnt > 0 ? (int)(3 * (threadIdx.x & 63)) % cnt : 0);
        for (int i = 0; i < cnt; ++i, mm = (mm == m1) ? m0 : mm + 1) {
          const int m = mm;
          const float t = (float)m;
          // samples in the keyhole window are not accumulated by the forward
          if (t >= K.z && t < K.w) continue;
          const float qx = fmaf(t, B.x, A.x), qy = fmaf(t, B.y, A.y), qz = fmaf(t, B.z, A.z);
          // the forward's sample value (precise march: no clamp needed)
          const float alpha = fmaf(K.x, trilinear<false>(vol, g, qx, qy, qz), K.y);
          if (!(alpha < 0.0f)) continue;
          // corners ix - 1 (weight 1 - f) and ix (weight f) per axis
          const float fx = __builtin_amdgcn_fractf(qx), fy = __builtin_amdgcn_fractf(qy);
          const float fz = __builtin_amdgcn_fractf(qz);
          const int lx = min(max((int)qx - o[0] + 1, 0), kVgL - 2);
          const int ly2 = min(max((int)qy - o[1] + 1, 0), kVgL - 2);
          const int lz = min(max((int)qz - o[2] + 1, 0), kVgL - 2);
          const float cz0 = coef * (1.0f - fz), cz1 = coef * fz;
          const float c00 = cz0 * (1.0f - fy), c10 = cz0 * fy, c01 = cz1 * (1.0f - fy), c11 = cz1 * fy;
          const float wx0 = 1.0f - fx;
          unsigned long long* s = s_acc + (lz * kVgL + ly2) * kVgL + lx;
          vg_add(s, c00 * wx0);
          vg_add(s + 1, c00 * fx);
          vg_add(s + kVgL, c10 * wx0);
          vg_add(s + kVgL + 1, c10 * fx);
          vg_add(s + kVgL * kVgL, c01 * wx0);
          vg_add(s + kVgL * kVgL + 1, c01 * fx);
          vg_add(s + kVgL * kVgL + kVgL, c11 * wx0);
          vg_add(s + kVgL * kVgL + kVgL + 1, c11 * fx);
        }
      }
    }
    // fold the clamped edge voxels at the volume's faces, one axis at a time
    const int st[3] = {1, kVgL, kVgL * kVgL};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      __syncthreads();
      const int a1 = ax == 0 ? 1 : 0, a2 = ax == 2 ? 1 : 2;
      if (tid < kVgL * kVgL) {
        const int base = (tid % kVgL) * st[a1] + (tid / kVgL) * st[a2];
        if (bi[ax] == 0) s_acc[base + 2 * st[ax]] += s_acc[base + st[ax]];
        if (bi[ax] == nb[ax] - 1) {
          const int nl = g.n[ax] - o[ax] + 1;  // voxel n - 1
          s_acc[base + nl * st[ax]] += s_acc[base + (nl + 1) * st[ax]] + s_acc[base + (nl + 2) * st[ax]];
        }
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < kVgEdge * kVgEdge * kVgEdge; e += 256) {
    const int lx = e % kVgEdge, ly = (e / kVgEdge) % kVgEdge, lz = e / (kVgEdge * kVgEdge);
    const int x = o[0] + lx, y = o[1] + ly, z = o[2] + lz;
    if (x >= g.n[0] || y >= g.n[1] || z >= g.n[2]) continue;
    float val = 0.0f;
    if (live) {
      val = k == kVgNonFinite ? __builtin_nanf("")
                              : (float)ldexp((double)(long long)s_acc[((lz + 2) * kVgL + ly + 2) * kVgL + lx + 2], -k);
    }
    a.grad[((long long)z * g.n[1] + y) * g.n[0] + x] = val;
  }
}

// ---------------------------------------------------------------------------
// Device-resident poses: the registration loop's pose update and the views'
// matrices recomputed from the poses, so that one replayed graph holds the
// whole iteration (render, NCC, pose gradient, update).
struct PoseGeom {
  double pv[16];      // the view's camera transform P V (its aspect)
  double vox[16];     // voxel -> world
  double center[3];   // rotation centre (world mm)
  int slot;           // which pose drives this view
  int pad;
};

__global__ __launch_bounds__(64) void pose_views_kernel(const PoseGeom* __restrict__ geo, const double* poses,
                                                        BatchView* views, int n_views) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n_views) return;
  const PoseGeom& G = geo[v];
  float m[16], dm[96];
  // a singular posed matrix (no rigid pose gives one) leaves the view as it was
  if (!pose::posed_view(G.pv, G.vox, poses + 6 * G.slot, G.center, m, dm)) return;
  BatchView& b = views[v];
  for (int k = 0; k < 16; ++k) b.v.m[k] = m[k];
  for (int k = 0; k < 96; ++k) b.dm[k] = dm[k];
  pose::baked_inverse(m, b.inv);
}

struct PoseUpdateArgs {
  rvk_pose_optimizer o;
  int n_views, n_slots;
};

// One workgroup.  The gradient of slot s's pose is the sum of d loss / d pose
// over the views it drives (one view each with per-view poses).  state holds
// the Adam moments m[n_slots*6], v[n_slots*6] and the step count.
__global__ __launch_bounds__(256) void pose_update_kernel(PoseUpdateArgs a, const PoseGeom* __restrict__ geo,
                                                          const float* __restrict__ grad, double* poses,
                                                          double* state) {
  const int nd = a.n_slots * 6;
  const double t = state[2 * nd] + 1.0;
  for (int i = threadIdx.x; i < nd; i += blockDim.x) {
    const int slot = i / 6, d = i % 6;
    double g = 0.0;
    for (int v = 0; v < a.n_views; ++v)
      if (geo[v].slot == slot) g += (double)grad[6 * v + d];
    const double lr = d < 3 ? a.o.lr_rotation : a.o.lr_translation;
    if (a.o.kind == RVK_OPT_ADAM) {
      const double b1 = a.o.beta1, b2 = a.o.beta2;
      const double m = b1 * state[i] + (1.0 - b1) * g;
      const double s = b2 * state[nd + i] + (1.0 - b2) * g * g;
      state[i] = m;
      state[nd + i] = s;
      const double mh = m / (1.0 - pow(b1, t)), sh = s / (1.0 - pow(b2, t));
      poses[i] -= lr * mh / (sqrt(sh) + (double)a.o.eps);
    } else {
      poses[i] -= lr * g;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) state[2 * nd] = t;
}

}  // namespace rvk

using namespace rvk;

struct rvk_drr_plan {
  rvk_ctx* ctx = nullptr;
  const rvk_volume* vol = nullptr;
  int frame = 0;
  int n_views = 0;
  rvk_march_params march{};
  rvk_drr_params drr{};
  std::vector<int> item_begin;     // host copy
  std::vector<long long> pix_off;  // n_views + 1
  int n_items = 0;
  BatchView* d_views = nullptr;
  int* d_item_begin = nullptr;
  unsigned* d_queues = nullptr;
  double* d_part = nullptr;   // n_items x 5
  double* d_stats = nullptr;  // n_views x 8
  float* d_pose_part = nullptr;  // n_items x 6
  RayRec* d_rays = nullptr;      // per pixel, allocated by the first volume gradient
  float* d_vg_cmax = nullptr;    // per 8x8 tile, the volume gradient's largest |coef|, allocated with d_rays
  int* d_vg_scale = nullptr;     // its fixed-point scale exponent
  float* d_jac = nullptr;        // 6 x total pixels, allocated with the plan when it has pose derivatives
  int schedule = RVK_SCHED_GRID;
  int waves_per_simd = 4;
  int shape = RVK_TILE_8X8;       // wave tile of the render
  int jac_shape = RVK_TILE_8X8;   // and of the registration forward
  // forward work orders: view-major, or grouped by the volume region the
  // tile's central ray crosses (all views' tiles of one region together, so
  // an XCD's L2 serves one region's cells to every view)
  int2* d_work[2] = {nullptr, nullptr};
  int order = 0, jac_order = 0;
  // the fused registration forward (pose Jacobians) has its own schedule:
  // its per-wave work differs enough that the best choice can differ
  int jac_schedule = RVK_SCHED_GRID;
  int jac_waves_per_simd = 4;
  int n_cu = 256;
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // buffers the graph was captured with
  float* g_out = nullptr;
  const float* g_tgt = nullptr;
  float* g_loss = nullptr;
  float* g_grad = nullptr;
  int g_enc = 0;  // cell encoding of the frame when the graph was captured
  unsigned g_gen = 0;  // and the volume's generation (its cells' addresses and layout)
  bool g_step = false;  // the graph is a registration step with the pose update
  int ncc_mode = RVK_NCC_CENTERED;
  // device-resident poses (rvk_drr_plan_set_poses)
  PoseGeom* d_pose_geom = nullptr;  // per view
  double* d_poses = nullptr;        // n_slots x 6
  double* d_opt_state = nullptr;    // 2 x n_slots x 6 moments + the step count
  int n_slots = 0;
  rvk_pose_optimizer opt{};
};

namespace {

int fill_views(rvk_drr_plan* p, const rvk_view* views, const float* dview, std::vector<BatchView>& hv) {
  hv.resize(p->n_views);
  for (int v = 0; v < p->n_views; ++v) {
    int rc = check_view(&views[v]);
    if (rc) return rc;
    if (v < (int)p->pix_off.size() - 1 && p->item_begin.size() == (size_t)p->n_views + 1) {
      const long long npix = (long long)views[v].width * views[v].height;
      if (npix != p->pix_off[v + 1] - p->pix_off[v])
        return fail(RVK_ERR_INVALID_ARGUMENT, "view %d changed size", v);
    }
    BatchView& b = hv[v];
    std::memset(&b, 0, sizeof(b));
    b.v = view_geom(&views[v]);
    if (dview) std::memcpy(b.dm, dview + 96 * (size_t)v, sizeof(b.dm));
    // inverse of the baked matrix (Gauss-Jordan, partial pivoting)
    pose::baked_inverse(views[v].view_to_voxels, b.inv);
    b.tiles16_x = (views[v].width + 15) / 16;
  }
  return RVK_OK;
}

// Work orders of the forward.  Order 0 is view-major.  Order 1 sorts the
// 16x16 tiles (4 work items each, kept together) by the volume region their
// central ray crosses: the midpoint of the ray's chord through the volume
// box, on a 32^3 grid of regions in Morton order (c4 forward: 39.2 ms
// view-major, 35.3 ms on a 16^3 grid, 35.0 ms on 32^3, 36.3 ms on 8^3);
// tiles whose ray misses the volume go last.  Ties keep view-major order.
static uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  uint32_t k = 0;
  for (int b = 0; b < 5; ++b) k |= (((x >> b) & 1u) << (3 * b)) | (((y >> b) & 1u) << (3 * b + 1)) | (((z >> b) & 1u) << (3 * b + 2));
  return k;
}

static void work_orders(const rvk_drr_plan* p, const rvk_view* views, std::vector<int2>& by_view,
                        std::vector<int2>& by_region) {
  by_view.resize(p->n_items);
  for (int v = 0; v < p->n_views; ++v)
    for (int i = p->item_begin[v]; i < p->item_begin[v + 1]; ++i) by_view[i] = make_int2(v, i - p->item_begin[v]);
  const int* dims = p->vol->dims;
  const double cell = std::max(1.0, std::max(dims[0], std::max(dims[1], dims[2])) / 32.0);
  struct Tile { uint32_t key; int v, macro; };
  std::vector<Tile> tiles;
  tiles.reserve(p->n_items / 4);
  for (int v = 0; v < p->n_views; ++v) {
    const rvk_view& vw = views[v];
    const float* m = vw.view_to_voxels;
    const int tx = (vw.width + 15) / 16, n_macro = (p->item_begin[v + 1] - p->item_begin[v]) / 4;
    for (int mc = 0; mc < n_macro; ++mc) {
      const double x = std::min(vw.width - 1, (mc % tx) * 16 + 8), y = std::min(vw.height - 1, (mc / tx) * 16 + 8);
      const double vx = vw.flipped ? x / vw.width : 1.0 - x / vw.width, vy = y / vw.height;
      double pt[2][3];
      bool ok = true;
      for (int d = 0; d < 2; ++d) {
        const double w = vx * m[12] + vy * m[13] + d * m[14] + m[15];
        ok = ok && w != 0.0;
        for (int a = 0; a < 3; ++a) pt[d][a] = (vx * m[4 * a] + vy * m[4 * a + 1] + d * m[4 * a + 2] + m[4 * a + 3]) / (ok ? w : 1.0);
      }
      // the line's chord through the box [0, n - 1]^3 (slab clipping)
      double t0 = -1e300, t1 = 1e300;
      for (int a = 0; a < 3 && ok; ++a) {
        const double dir = pt[1][a] - pt[0][a];
        if (dir == 0.0) {
          ok = pt[0][a] >= 0.0 && pt[0][a] <= dims[a] - 1;
        } else {
          double ta = (0.0 - pt[0][a]) / dir, tb = (dims[a] - 1 - pt[0][a]) / dir;
          if (ta > tb) std::swap(ta, tb);
          t0 = std::max(t0, ta);
          t1 = std::min(t1, tb);
        }
      }
      uint32_t key = 0xffffffffu;
      if (ok && t0 <= t1) {
        const double tm = 0.5 * (t0 + t1);
        uint32_t g[3];
        for (int a = 0; a < 3; ++a)
          g[a] = (uint32_t)std::min(31.0, std::max(0.0, std::floor((pt[0][a] + tm * (pt[1][a] - pt[0][a])) / cell)));
        key = morton3(g[0], g[1], g[2]);
      }
      tiles.push_back(Tile{key, v, mc});
    }
  }
  std::stable_sort(tiles.begin(), tiles.end(), [](const Tile& a, const Tile& b) { return a.key < b.key; });
  by_region.resize(p->n_items);
  size_t i = 0;
  for (const Tile& tl : tiles)
    for (int q = 0; q < 4; ++q) by_region[i++] = make_int2(tl.v, 4 * tl.macro + q);
}

static int upload_work_orders(rvk_drr_plan* p, const rvk_view* views) {
  std::vector<int2> o[2];
  work_orders(p, views, o[0], o[1]);
  for (int k = 0; k < 2; ++k) {
    if (!p->d_work[k]) RVK_HIP(hipMalloc((void**)&p->d_work[k], sizeof(int2) * (size_t)p->n_items));
    RVK_HIP(hipMemcpy(p->d_work[k], o[k].data(), sizeof(int2) * (size_t)p->n_items, hipMemcpyHostToDevice));
  }
  return RVK_OK;
}

BatchArgs batch_args(const rvk_drr_plan* p, float* out) {
  BatchArgs a;
  a.g = vol_geom(p->vol);
  a.m.intercept = p->drr.ct_intercept;
  a.m.slope = p->drr.ct_slope;
  a.m.offset = p->drr.ct_offset;
  a.m.output = p->drr.output;
  a.m.ranges = p->vol->ranges[p->frame];
  a.m.rbx = p->vol->nrb[0];
  a.m.rby = p->vol->nrb[1];
  a.m.ref_term = p->drr.ref_termination;
  a.m.jitter = p->march.jitter;
  a.step_len = min_spacing(p->vol) * p->march.step_scale;
  a.vol = p->vol->frames[p->frame];
  a.views = p->d_views;
  a.item_begin = p->d_item_begin;
  a.n_views = p->n_views;
  a.n_items = p->n_items;
  a.out = out;
  a.jac = p->d_jac;
  a.total = p->pix_off[p->n_views];
  a.queues = p->d_queues;
  a.shape = p->shape;
  a.work = p->d_work[p->order];
  return a;
}

int launch_forward(rvk_drr_plan* p, float* out, hipStream_t s, bool jac = false) {
  BatchArgs a = batch_args(p, out);
  a.shape = jac ? p->jac_shape : p->shape;
  a.work = p->d_work[jac ? p->jac_order : p->order];
  const bool skip = p->march.skip_rule != 0, pr = p->march.precise != 0;
  return dispatch_frame(p->vol, p->frame, [&](auto tag) {
    using CV = decltype(tag);
    const int sched = jac ? p->jac_schedule : p->schedule;
    const int wps = jac ? p->jac_waves_per_simd : p->waves_per_simd;
    if (sched == RVK_SCHED_PERSISTENT) {
      hipLaunchKernelGGL(reset_queues_kernel, dim3(1), dim3(64), 0, s, p->d_queues);
      const int blocks = std::max(1, std::min(p->n_cu * wps, (p->n_items + 3) / 4));
#define RVK_LP(S, P, J)                                                                                     \
  do {                                                                                                      \
    if (p->drr.ref_termination)                                                                             \
      hipLaunchKernelGGL((batch_persistent_kernel<CV, S, P, J, true>), dim3(blocks), dim3(256), 0, s, a);    \
    else                                                                                                    \
      hipLaunchKernelGGL((batch_persistent_kernel<CV, S, P, J, false>), dim3(blocks), dim3(256), 0, s, a);   \
  } while (0)
      if (jac) { if (skip) RVK_LP(true, true, true); else RVK_LP(false, true, true); }
      else if (skip) { if (pr) RVK_LP(true, true, false); else RVK_LP(true, false, false); }
      else { if (pr) RVK_LP(false, true, false); else RVK_LP(false, false, false); }
#undef RVK_LP
    } else {
      const int blocks = (p->n_items + 3) / 4;
#define RVK_LG(S, P, J)                                                                                     \
  do {                                                                                                      \
    if (p->drr.ref_termination)                                                                             \
      hipLaunchKernelGGL((batch_grid_kernel<CV, S, P, J, true>), dim3(blocks), dim3(256), 0, s, a);          \
    else                                                                                                    \
      hipLaunchKernelGGL((batch_grid_kernel<CV, S, P, J, false>), dim3(blocks), dim3(256), 0, s, a);         \
  } while (0)
      if (jac) { if (skip) RVK_LG(true, true, true); else RVK_LG(false, true, true); }
      else if (skip) { if (pr) RVK_LG(true, true, false); else RVK_LG(true, false, false); }
      else { if (pr) RVK_LG(false, true, false); else RVK_LG(false, false, false); }
#undef RVK_LG
    }
    RVK_HIP(hipGetLastError());
    return (int)RVK_OK;
  });
}

NccArgs ncc_args(const rvk_drr_plan* p, const float* img, const float* tgt, float* loss, float* grad_img) {
  NccArgs n;
  n.views = p->d_views;
  n.item_begin = p->d_item_begin;
  n.n_views = p->n_views;
  n.n_items = p->n_items;
  n.img = img;
  n.tgt = tgt;
  n.part = p->d_part;
  n.stats = p->d_stats;
  n.loss = loss;
  n.grad_img = grad_img;
  n.centred = p->ncc_mode == RVK_NCC_CENTERED;
  return n;
}

// The registration step: one march per pixel producing the image and its
// pose Jacobian, then NCC (tile moments, per-view statistics) and the
// Jacobian-weighted reduction to d loss / d pose.
int launch_register(rvk_drr_plan* p, float* out, const float* tgt, float* loss, float* grad, hipStream_t s) {
  if (!p->d_jac) return fail(RVK_ERR_NOT_READY, "the plan was created without pose derivatives (dview)");
  int rc = launch_forward(p, out, s, true);
  if (rc) return rc;
  const int blocks = (p->n_items + 3) / 4;
  NccArgs n = ncc_args(p, out, tgt, loss, nullptr);
  hipLaunchKernelGGL(ncc_partial_kernel, dim3(blocks), dim3(256), 0, s, n);
  hipLaunchKernelGGL(ncc_view_kernel, dim3(p->n_views), dim3(256), 0, s, n);
  JacGradArgs ja;
  ja.views = p->d_views;
  ja.item_begin = p->d_item_begin;
  ja.n_views = p->n_views;
  ja.n_items = p->n_items;
  ja.img = out;
  ja.tgt = tgt;
  ja.jac = p->d_jac;
  ja.total = p->pix_off[p->n_views];
  ja.stats = p->d_stats;
  ja.output = p->drr.output;
  ja.part = p->d_pose_part;
  hipLaunchKernelGGL(jac_grad_kernel, dim3(blocks), dim3(256), 0, s, ja);
  hipLaunchKernelGGL(pose_reduce_kernel, dim3(p->n_views), dim3(256), 0, s, p->d_item_begin, p->d_pose_part, grad);
  RVK_HIP(hipGetLastError());
  return RVK_OK;
}

// A captured graph holds the frame's cell pointer, encoding and layout as
// they were at capture: replayed only while all three are unchanged.
bool graph_current(const rvk_drr_plan* p) {
  return p->exec && p->g_enc == p->vol->enc[p->frame] && p->g_gen == p->vol->generation;
}

void drop_graph(rvk_drr_plan* p) {
  if (p->exec) hipGraphExecDestroy(p->exec);
  if (p->graph) hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
  p->g_out = nullptr;
  p->g_tgt = nullptr;
  p->g_loss = nullptr;
  p->g_grad = nullptr;
  p->g_step = false;
}

// The pose update and the views' matrices from the new poses.
int launch_pose_update(rvk_drr_plan* p, const float* grad, hipStream_t s) {
  PoseUpdateArgs a;
  a.o = p->opt;
  a.n_views = p->n_views;
  a.n_slots = p->n_slots;
  hipLaunchKernelGGL(pose_update_kernel, dim3(1), dim3(256), 0, s, a, p->d_pose_geom, grad, p->d_poses,
                     p->d_opt_state);
  hipLaunchKernelGGL(pose_views_kernel, dim3((p->n_views + 63) / 64), dim3(64), 0, s, p->d_pose_geom, p->d_poses,
                     p->d_views, p->n_views);
  RVK_HIP(hipGetLastError());
  return RVK_OK;
}

int launch_step(rvk_drr_plan* p, float* out, const float* tgt, float* loss, float* grad, hipStream_t s) {
  int rc = launch_register(p, out, tgt, loss, grad, s);
  if (rc) return rc;
  return launch_pose_update(p, grad, s);
}

}  // namespace

extern "C" {

int rvk_drr_plan_create(rvk_ctx* ctx, const rvk_volume* vol, int frame, int n_views, const rvk_view* views,
                        const float* dview, const rvk_march_params* march, const rvk_drr_params* drr,
                        rvk_drr_plan** out) {
  if (!ctx || !views || !march || !drr || !out) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  int rc = check_frame(vol, frame);
  if (rc) return rc;
  if (n_views < 1 || n_views > 65536) return fail(RVK_ERR_INVALID_ARGUMENT, "n_views %d out of range", n_views);
  if (drr->output != RVK_DRR_F32_EXP_NEG_L && drr->output != RVK_DRR_F32_L)
    return fail(RVK_ERR_UNSUPPORTED, "batched DRR renders the float output modes only");
  if (!(march->step_scale > 0.0f) || !std::isfinite(march->step_scale))
    return fail(RVK_ERR_INVALID_ARGUMENT, "step_scale must be positive and finite");
  RVK_HIP(hipSetDevice(ctx->device));
  rvk_drr_plan* p = new rvk_drr_plan();
  p->ctx = ctx;
  p->vol = vol;
  p->frame = frame;
  p->n_views = n_views;
  p->march = *march;
  p->drr = *drr;
  p->item_begin.assign(n_views + 1, 0);
  p->pix_off.assign(n_views + 1, 0);
  long long items = 0;
  for (int v = 0; v < n_views; ++v) {
    p->item_begin[v] = (int)items;
    p->pix_off[v] = (v ? p->pix_off[v - 1] + (long long)views[v - 1].width * views[v - 1].height : 0);
    const long long t = 4LL * ((views[v].width + 15) / 16) * ((views[v].height + 15) / 16);
    items += t;
  }
  p->pix_off[n_views] = p->pix_off[n_views - 1] + (long long)views[n_views - 1].width * views[n_views - 1].height;
  if (items > (1LL << 30)) { delete p; return fail(RVK_ERR_INVALID_ARGUMENT, "too many pixels in one plan"); }
  p->item_begin[n_views] = (int)items;
  p->n_items = (int)items;
  std::vector<BatchView> hv;
  std::vector<long long> keep = p->pix_off;
  p->pix_off.clear();  // no size check on creation
  rc = fill_views(p, views, dview, hv);
  p->pix_off = keep;
  if (rc) { delete p; return rc; }
  for (int v = 0; v < n_views; ++v) hv[v].pix_off = p->pix_off[v];
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) p->n_cu = prop.multiProcessorCount;
  auto alloc = [&](void** d, size_t bytes) { return hipMalloc(d, bytes); };
  hipError_t e = alloc((void**)&p->d_views, sizeof(BatchView) * n_views);
  if (e == hipSuccess) e = alloc((void**)&p->d_item_begin, sizeof(int) * (n_views + 1));
  if (e == hipSuccess) e = alloc((void**)&p->d_queues, sizeof(unsigned) * kNumXcd);
  if (e == hipSuccess) e = alloc((void**)&p->d_part, sizeof(double) * 5 * (size_t)p->n_items);
  if (e == hipSuccess) e = alloc((void**)&p->d_stats, sizeof(double) * 8 * (size_t)n_views);
  if (e == hipSuccess) e = alloc((void**)&p->d_pose_part, sizeof(float) * 6 * (size_t)p->n_items);
  if (e == hipSuccess && dview) e = alloc((void**)&p->d_jac, sizeof(float) * 6 * (size_t)p->pix_off[n_views]);
  if (e == hipSuccess) e = hipMemcpy(p->d_views, hv.data(), sizeof(BatchView) * n_views, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(p->d_item_begin, p->item_begin.data(), sizeof(int) * (n_views + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    rvk_drr_plan_destroy(p);
    return hip_fail(e, "plan allocation");
  }
  rc = upload_work_orders(p, views);
  if (rc) {
    rvk_drr_plan_destroy(p);
    return rc;
  }
  *out = p;
  return RVK_OK;
}

int rvk_drr_plan_destroy(rvk_drr_plan* p) {
  if (!p) return RVK_OK;
  hipSetDevice(p->ctx->device);
  hipStreamSynchronize(p->ctx->stream);
  drop_graph(p);
  if (p->cap_stream) hipStreamDestroy(p->cap_stream);
  hipFree(p->d_views);
  hipFree(p->d_item_begin);
  hipFree(p->d_work[0]);
  hipFree(p->d_work[1]);
  hipFree(p->d_queues);
  hipFree(p->d_part);
  hipFree(p->d_stats);
  hipFree(p->d_pose_part);
  if (p->d_rays) hipFree(p->d_rays);
  if (p->d_vg_cmax) hipFree(p->d_vg_cmax);
  if (p->d_vg_scale) hipFree(p->d_vg_scale);
  if (p->d_jac) hipFree(p->d_jac);
  if (p->d_pose_geom) hipFree(p->d_pose_geom);
  if (p->d_poses) hipFree(p->d_poses);
  if (p->d_opt_state) hipFree(p->d_opt_state);
  delete p;
  return RVK_OK;
}

int rvk_drr_plan_update_views(rvk_drr_plan* p, const rvk_view* views, const float* dview) {
  if (!p || !views) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (p->d_jac && !dview) return fail(RVK_ERR_INVALID_ARGUMENT, "this plan carries pose derivatives: pass dview");
  std::vector<BatchView> hv;
  int rc = fill_views(p, views, dview, hv);
  if (rc) return rc;
  for (int v = 0; v < p->n_views; ++v) hv[v].pix_off = p->pix_off[v];
  RVK_HIP(hipSetDevice(p->ctx->device));
  RVK_HIP(hipMemcpyAsync(p->d_views, hv.data(), sizeof(BatchView) * p->n_views, hipMemcpyHostToDevice,
                         p->ctx->stream));
  RVK_HIP(hipStreamSynchronize(p->ctx->stream));
  // new views, new regions (a captured graph reads the arrays, not their contents)
  return upload_work_orders(p, views);
}

int rvk_drr_plan_offsets(const rvk_drr_plan* p, int64_t* offsets) {
  if (!p || !offsets) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  for (int v = 0; v <= p->n_views; ++v) offsets[v] = p->pix_off[v];
  return RVK_OK;
}

int rvk_drr_plan_set_schedule(rvk_drr_plan* p, int schedule, int wps) {
  if (!p) return fail(RVK_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (schedule != RVK_SCHED_GRID && schedule != RVK_SCHED_PERSISTENT)
    return fail(RVK_ERR_INVALID_ARGUMENT, "unknown schedule %d", schedule);
  if (wps < 1 || wps > 8) return fail(RVK_ERR_INVALID_ARGUMENT, "waves_per_simd must be within 1..8");
  if (schedule != p->schedule || wps != p->waves_per_simd || schedule != p->jac_schedule ||
      wps != p->jac_waves_per_simd)
    drop_graph(p);
  p->schedule = p->jac_schedule = schedule;
  p->waves_per_simd = p->jac_waves_per_simd = wps;
  return RVK_OK;
}

int rvk_drr_plan_get_schedule(const rvk_drr_plan* p, int* schedule, int* wps) {
  if (!p) return fail(RVK_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (schedule) *schedule = p->schedule;
  if (wps) *wps = p->waves_per_simd;
  return RVK_OK;
}

int rvk_drr_plan_autotune(rvk_drr_plan* p, float* out, int reps, float* best_ms) {
  if (!p || !out) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  reps = std::max(1, std::min(reps, 100));
  RVK_HIP(hipSetDevice(p->ctx->device));
  drop_graph(p);
  // schedule x waves per SIMD x wave tile x work order
  const int cand[][2] = {{RVK_SCHED_GRID, 4}, {RVK_SCHED_PERSISTENT, 2}, {RVK_SCHED_PERSISTENT, 4},
                         {RVK_SCHED_PERSISTENT, 8}};
  const int shapes[] = {RVK_TILE_8X8, RVK_TILE_16X4, RVK_TILE_4X16};
  hipEvent_t e0, e1;
  RVK_HIP(hipEventCreate(&e0));
  RVK_HIP(hipEventCreate(&e1));
  int rc = RVK_OK;
  // the plain render, and for plans with pose derivatives the registration
  // forward, each keep their fastest candidate
  for (int jac = 0; jac < ((p->d_jac && p->march.precise) ? 2 : 1) && !rc; ++jac) {
    int& sched = jac ? p->jac_schedule : p->schedule;
    int& wps = jac ? p->jac_waves_per_simd : p->waves_per_simd;
    int& shape = jac ? p->jac_shape : p->shape;
    int& order = jac ? p->jac_order : p->order;
    float best = 1e30f;
    int bs = sched, bw = wps, bsh = shape, bo = order;
    for (int oi = 0; oi < 2 * 3 && !rc; ++oi) {
      const int o = oi / 3, sh = shapes[oi % 3];
      for (const auto& c : cand) {
        sched = c[0];
        wps = c[1];
        shape = sh;
        order = o;
        rc = launch_forward(p, out, p->ctx->stream, jac);  // warm-up
        if (rc) break;
        hipEventRecord(e0, p->ctx->stream);
        for (int r = 0; r < reps && !rc; ++r) rc = launch_forward(p, out, p->ctx->stream, jac);
        hipEventRecord(e1, p->ctx->stream);
        hipEventSynchronize(e1);
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e0, e1);
        ms /= reps;
        if (ms < best) { best = ms; bs = c[0]; bw = c[1]; bsh = sh; bo = o; }
      }
    }
    sched = bs;
    wps = bw;
    shape = bsh;
    order = bo;
    if (!jac && best_ms) *best_ms = best;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return rc;
}

int rvk_drr_plan_autotune_layout(rvk_drr_plan* p, rvk_volume* vol, float* out, int reps, float* best_ms,
                                 int* layout) {
  if (!p || !vol || !out) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (vol != p->vol) return fail(RVK_ERR_INVALID_ARGUMENT, "vol is not the volume the plan renders");
  int rc = RVK_OK;
  struct Choice {
    int layout = -1, sched, wps, shape, order, jsched, jwps, jshape, jorder;
    float ms = 1e30f;
  } best;
  for (const int l : {RVK_CELLS_PLAIN, RVK_CELLS_BRICK4, RVK_CELLS_PAIR8}) {
    // a layout the volume does not fit (its address limits, or the device
    // memory for a second copy of its cells) is skipped; the volume keeps
    // its cells then
    if (rvk_volume_relayout(vol, l) != RVK_OK) continue;
    float ms = 0.0f;
    rc = rvk_drr_plan_autotune(p, out, reps, &ms);
    if (rc) return rc;
    if (ms < best.ms) {
      best.layout = l;
      best.ms = ms;
      best.sched = p->schedule;
      best.wps = p->waves_per_simd;
      best.shape = p->shape;
      best.order = p->order;
      best.jsched = p->jac_schedule;
      best.jwps = p->jac_waves_per_simd;
      best.jshape = p->jac_shape;
      best.jorder = p->jac_order;
    }
  }
  if (best.layout < 0) return fail(RVK_ERR_INVALID_ARGUMENT, "no cell layout fits the volume");
  rc = rvk_volume_relayout(vol, best.layout);
  if (rc) return rc;
  drop_graph(p);
  p->schedule = best.sched;
  p->waves_per_simd = best.wps;
  p->shape = best.shape;
  p->order = best.order;
  p->jac_schedule = best.jsched;
  p->jac_waves_per_simd = best.jwps;
  p->jac_shape = best.jshape;
  p->jac_order = best.jorder;
  if (best_ms) *best_ms = best.ms;
  if (layout) *layout = best.layout;
  return RVK_OK;
}

int rvk_drr_plan_set_work_order(rvk_drr_plan* p, int order) {
  if (!p) return fail(RVK_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (order != RVK_ORDER_VIEWS && order != RVK_ORDER_REGIONS)
    return fail(RVK_ERR_INVALID_ARGUMENT, "unknown work order %d", order);
  if (order != p->order || order != p->jac_order) drop_graph(p);
  p->order = p->jac_order = order;
  return RVK_OK;
}

int rvk_drr_plan_get_work_order(const rvk_drr_plan* p, int* order) {
  if (!p || !order) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  *order = p->order;
  return RVK_OK;
}

int rvk_drr_plan_set_tile_shape(rvk_drr_plan* p, int shape) {
  if (!p) return fail(RVK_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (shape != RVK_TILE_8X8 && shape != RVK_TILE_16X4 && shape != RVK_TILE_4X16)
    return fail(RVK_ERR_INVALID_ARGUMENT, "unknown tile shape %d", shape);
  if (shape != p->shape || shape != p->jac_shape) drop_graph(p);
  p->shape = p->jac_shape = shape;
  return RVK_OK;
}

int rvk_drr_plan_get_tile_shape(const rvk_drr_plan* p, int* shape) {
  if (!p || !shape) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  *shape = p->shape;
  return RVK_OK;
}

int rvk_drr_plan_render(rvk_drr_plan* p, float* out) {
  if (!p || !out) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  RVK_HIP(hipSetDevice(p->ctx->device));
  if (graph_current(p) && !p->g_step && p->g_out == out && !p->g_tgt) {
    RVK_HIP(hipGraphLaunch(p->exec, p->ctx->stream));
    return RVK_OK;
  }
  return launch_forward(p, out, p->ctx->stream);
}

int rvk_drr_plan_capture(rvk_drr_plan* p, float* out, const float* tgt, float* loss, float* grad) {
  if (!p) return fail(RVK_ERR_INVALID_ARGUMENT, "plan is NULL");
  RVK_HIP(hipSetDevice(p->ctx->device));
  drop_graph(p);
  if (!out) return RVK_OK;
  if (tgt && (!loss || !grad)) return fail(RVK_ERR_INVALID_ARGUMENT, "registration capture needs loss and grad");
  if (tgt && !p->march.precise) return fail(RVK_ERR_UNSUPPORTED, "the adjoint differentiates the precise march");
  // Capture on the plan's own stream (the context may use the legacy default
  // stream, which cannot be captured); replays go to the context stream.
  RVK_HIP(hipStreamSynchronize(p->ctx->stream));
  RVK_HIP(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
  int rc = tgt ? launch_register(p, out, tgt, loss, grad, p->cap_stream) : launch_forward(p, out, p->cap_stream);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
  if (rc) { if (g) hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  p->graph = g;
  e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) { drop_graph(p); return hip_fail(e, "hipGraphInstantiate"); }
  p->g_out = out;
  p->g_tgt = tgt;
  p->g_loss = loss;
  p->g_grad = grad;
  p->g_enc = p->vol->enc[p->frame];
  p->g_gen = p->vol->generation;
  return RVK_OK;
}

int rvk_drr_plan_register(rvk_drr_plan* p, float* out, const float* tgt, float* loss, float* grad) {
  if (!p || !out || !tgt || !loss || !grad) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!p->march.precise) return fail(RVK_ERR_UNSUPPORTED, "the adjoint differentiates the precise march");
  RVK_HIP(hipSetDevice(p->ctx->device));
  if (graph_current(p) && !p->g_step && p->g_out == out && p->g_tgt == tgt && p->g_loss == loss &&
      p->g_grad == grad) {
    RVK_HIP(hipGraphLaunch(p->exec, p->ctx->stream));
    return RVK_OK;
  }
  return launch_register(p, out, tgt, loss, grad, p->ctx->stream);
}

int rvk_drr_plan_ncc(rvk_drr_plan* p, const float* images, const float* tgt, float* loss, float* dldi) {
  if (!p || !images || !tgt || !loss) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  RVK_HIP(hipSetDevice(p->ctx->device));
  const int blocks = (p->n_items + 3) / 4;
  NccArgs n = ncc_args(p, images, tgt, loss, dldi);
  hipLaunchKernelGGL(ncc_partial_kernel, dim3(blocks), dim3(256), 0, p->ctx->stream, n);
  hipLaunchKernelGGL(ncc_view_kernel, dim3(p->n_views), dim3(256), 0, p->ctx->stream, n);
  if (dldi) hipLaunchKernelGGL(ncc_grad_image_kernel, dim3(blocks), dim3(256), 0, p->ctx->stream, n);
  RVK_HIP(hipGetLastError());
  return RVK_OK;
}

int rvk_drr_plan_backward(rvk_drr_plan* p, const float* images, const float* dldi, float* grad_pose,
                          float* grad_vol) {
  if (!p || !images || !dldi) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!p->march.precise) return fail(RVK_ERR_UNSUPPORTED, "the adjoint differentiates the precise march");
  // the pose tangents come from the plan's dview (d_jac exists exactly then)
  if (grad_pose && !p->d_jac) return fail(RVK_ERR_NOT_READY, "the plan was created without pose derivatives (dview)");
  RVK_HIP(hipSetDevice(p->ctx->device));
  hipStream_t s = p->ctx->stream;
  const int blocks = (p->n_items + 3) / 4;
  if (grad_pose) {
    AdjArgs aa;
    aa.b = batch_args(p, const_cast<float*>(images));
    aa.img = images;
    aa.dldi = dldi;
    aa.stats = nullptr;
    aa.tgt = nullptr;
    aa.part = p->d_pose_part;
    int rc = dispatch_frame(p->vol, p->frame, [&](auto tag) {
      using CV = decltype(tag);
      if (p->march.skip_rule) hipLaunchKernelGGL((adjoint_pose_kernel<CV, true>), dim3(blocks), dim3(256), 0, s, aa);
      else hipLaunchKernelGGL((adjoint_pose_kernel<CV, false>), dim3(blocks), dim3(256), 0, s, aa);
      return (int)RVK_OK;
    });
    if (rc) return rc;
    hipLaunchKernelGGL(pose_reduce_kernel, dim3(p->n_views), dim3(256), 0, s, p->d_item_begin, p->d_pose_part,
                       grad_pose);
    RVK_HIP(hipGetLastError());
  }
  if (grad_vol) {
    if (p->march.skip_rule || p->drr.ref_termination)
      return fail(RVK_ERR_UNSUPPORTED, "the volume gradient needs the plain march (no skip rule, no termination)");
    const long long total = p->pix_off[p->n_views];
    // each buffer on its own, set only once allocated: after a failed
    // allocation the next call allocates what is still missing
    auto need = [](void* have, size_t bytes, void** out) -> hipError_t {
      if (have) return hipSuccess;
      void* q = nullptr;
      const hipError_t e = hipMalloc(&q, bytes);
      if (e == hipSuccess) *out = q;
      return e;
    };
    RVK_HIP(need(p->d_rays, sizeof(RayRec) * (size_t)total, (void**)&p->d_rays));
    RVK_HIP(need(p->d_vg_cmax, sizeof(float) * (size_t)p->n_items, (void**)&p->d_vg_cmax));
    RVK_HIP(need(p->d_vg_scale, sizeof(int), (void**)&p->d_vg_scale));
    VolGradArgs va;
    va.b = batch_args(p, nullptr);
    va.img = images;
    va.dldi = dldi;
    va.rays = p->d_rays;
    va.coef_max = p->d_vg_cmax;
    va.scale_log2 = p->d_vg_scale;
    va.grad = grad_vol;
    hipLaunchKernelGGL(ray_record_kernel, dim3(blocks), dim3(256), 0, s, va);
    hipLaunchKernelGGL(vg_scale_kernel, dim3(1), dim3(1024), 0, s, (const float*)p->d_vg_cmax, p->n_items,
                       p->d_vg_scale);
    const int* d = p->vol->dims;
    const dim3 vgrid((unsigned)(((d[0] + kVgEdge - 1) / kVgEdge) * ((d[1] + kVgEdge - 1) / kVgEdge) *
                                ((d[2] + kVgEdge - 1) / kVgEdge)));
    int rc = dispatch_frame(p->vol, p->frame, [&](auto tag) {
      using CV = decltype(tag);
      hipLaunchKernelGGL(volume_grad_kernel<CV>, vgrid, dim3(256), 0, s, va);
      return (int)RVK_OK;
    });
    if (rc) return rc;
    RVK_HIP(hipGetLastError());
  }
  return RVK_OK;
}

int rvk_drr_plan_set_ncc(rvk_drr_plan* p, int mode) {
  if (!p) return fail(RVK_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (mode != RVK_NCC_CENTERED && mode != RVK_NCC_UNCENTERED)
    return fail(RVK_ERR_INVALID_ARGUMENT, "unknown NCC mode %d", mode);
  if (mode != p->ncc_mode) drop_graph(p);
  p->ncc_mode = mode;
  return RVK_OK;
}

int rvk_drr_plan_set_poses(rvk_drr_plan* p, const rvk_camera* cameras, const double center[3], const double origin[3],
                           const double* poses, const rvk_pose_optimizer* opt) {
  if (!p || !cameras || !center || !poses || !opt) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!p->d_jac) return fail(RVK_ERR_NOT_READY, "the plan was created without pose derivatives (dview)");
  if (!p->march.precise) return fail(RVK_ERR_UNSUPPORTED, "the adjoint differentiates the precise march");
  if (opt->kind != RVK_OPT_GD && opt->kind != RVK_OPT_ADAM)
    return fail(RVK_ERR_INVALID_ARGUMENT, "unknown optimizer kind %d", opt->kind);
  if (!std::isfinite(opt->lr_rotation) || !std::isfinite(opt->lr_translation) || opt->lr_rotation < 0.0f ||
      opt->lr_translation < 0.0f)
    return fail(RVK_ERR_INVALID_ARGUMENT, "learning rates must be finite and >= 0");
  if (opt->kind == RVK_OPT_ADAM && !(opt->beta1 >= 0.0f && opt->beta1 < 1.0f && opt->beta2 >= 0.0f &&
                                     opt->beta2 < 1.0f && opt->eps > 0.0f))
    return fail(RVK_ERR_INVALID_ARGUMENT, "Adam needs beta1, beta2 in [0, 1) and eps > 0");
  const int n_slots = opt->shared_pose ? 1 : p->n_views;
  const double zero[3] = {0.0, 0.0, 0.0};
  std::vector<PoseGeom> geo(p->n_views);
  std::vector<BatchView> hv(p->n_views);
  RVK_HIP(hipSetDevice(p->ctx->device));
  RVK_HIP(hipStreamSynchronize(p->ctx->stream));
  RVK_HIP(hipMemcpy(hv.data(), p->d_views, sizeof(BatchView) * p->n_views, hipMemcpyDeviceToHost));
  for (int v = 0; v < p->n_views; ++v) {
    PoseGeom& g = geo[v];
    std::memset(&g, 0, sizeof(g));
    const double aspect = (double)hv[v].v.W / (double)hv[v].v.H;
    if (camera_pv(&cameras[v], aspect, g.pv)) return fail(RVK_ERR_INVALID_ARGUMENT, "camera %d is invalid", v);
    voxels_matrix(origin ? origin : zero, p->vol->spacing, g.vox);
    for (int a = 0; a < 3; ++a) g.center[a] = center[a];
    g.slot = opt->shared_pose ? 0 : v;
    float m[16];
    if (!pose::posed_view(g.pv, g.vox, poses + 6 * g.slot, g.center, m, nullptr))
      return fail(RVK_ERR_INVALID_ARGUMENT, "view %d: singular posed matrix", v);
  }
  if (n_slots != p->n_slots) {
    if (p->d_poses) hipFree(p->d_poses);
    if (p->d_opt_state) hipFree(p->d_opt_state);
    p->d_poses = nullptr;
    p->d_opt_state = nullptr;
    p->n_slots = 0;
    // both or neither: the plan never holds poses without their optimiser state
    void* dp = nullptr;
    void* ds = nullptr;
    hipError_t e = hipMalloc(&dp, sizeof(double) * 6 * n_slots);
    if (e == hipSuccess) e = hipMalloc(&ds, sizeof(double) * (12 * (size_t)n_slots + 1));
    if (e != hipSuccess) {
      if (dp) hipFree(dp);
      return hip_fail(e, "hipMalloc(device poses)");
    }
    p->d_poses = (double*)dp;
    p->d_opt_state = (double*)ds;
    p->n_slots = n_slots;
  }
  if (!p->d_pose_geom) {
    void* q = nullptr;
    RVK_HIP(hipMalloc(&q, sizeof(PoseGeom) * p->n_views));
    p->d_pose_geom = (PoseGeom*)q;
  }
  drop_graph(p);
  p->opt = *opt;
  const std::vector<double> state(12 * (size_t)n_slots + 1, 0.0);
  RVK_HIP(hipMemcpy(p->d_pose_geom, geo.data(), sizeof(PoseGeom) * p->n_views, hipMemcpyHostToDevice));
  RVK_HIP(hipMemcpy(p->d_poses, poses, sizeof(double) * 6 * n_slots, hipMemcpyHostToDevice));
  RVK_HIP(hipMemcpy(p->d_opt_state, state.data(), sizeof(double) * state.size(), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(pose_views_kernel, dim3((p->n_views + 63) / 64), dim3(64), 0, p->ctx->stream, p->d_pose_geom,
                     p->d_poses, p->d_views, p->n_views);
  RVK_HIP(hipGetLastError());
  RVK_HIP(hipStreamSynchronize(p->ctx->stream));
  return RVK_OK;
}

int rvk_drr_plan_get_poses(rvk_drr_plan* p, double* poses) {
  if (!p || !poses) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!p->d_poses) return fail(RVK_ERR_NOT_READY, "the plan has no device poses (rvk_drr_plan_set_poses)");
  RVK_HIP(hipSetDevice(p->ctx->device));
  RVK_HIP(hipMemcpyAsync(poses, p->d_poses, sizeof(double) * 6 * p->n_slots, hipMemcpyDeviceToHost, p->ctx->stream));
  RVK_HIP(hipStreamSynchronize(p->ctx->stream));
  return RVK_OK;
}

int rvk_drr_plan_step(rvk_drr_plan* p, float* out, const float* tgt, float* loss, float* grad) {
  if (!p || !out || !tgt || !loss || !grad) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!p->d_poses) return fail(RVK_ERR_NOT_READY, "the plan has no device poses (rvk_drr_plan_set_poses)");
  RVK_HIP(hipSetDevice(p->ctx->device));
  if (graph_current(p) && p->g_step && p->g_out == out && p->g_tgt == tgt && p->g_loss == loss &&
      p->g_grad == grad) {
    RVK_HIP(hipGraphLaunch(p->exec, p->ctx->stream));
    return RVK_OK;
  }
  return launch_step(p, out, tgt, loss, grad, p->ctx->stream);
}

int rvk_drr_plan_capture_step(rvk_drr_plan* p, float* out, const float* tgt, float* loss, float* grad) {
  if (!p || !out || !tgt || !loss || !grad) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!p->d_poses) return fail(RVK_ERR_NOT_READY, "the plan has no device poses (rvk_drr_plan_set_poses)");
  RVK_HIP(hipSetDevice(p->ctx->device));
  drop_graph(p);
  RVK_HIP(hipStreamSynchronize(p->ctx->stream));
  RVK_HIP(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
  int rc = launch_step(p, out, tgt, loss, grad, p->cap_stream);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
  if (rc) { if (g) hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  p->graph = g;
  e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) { drop_graph(p); return hip_fail(e, "hipGraphInstantiate"); }
  p->g_out = out;
  p->g_tgt = tgt;
  p->g_loss = loss;
  p->g_grad = grad;
  p->g_enc = p->vol->enc[p->frame];
  p->g_gen = p->vol->generation;
  p->g_step = true;
  return RVK_OK;
}

}  // extern "C"
