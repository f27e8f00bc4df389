// rvk_march.h -- the DRR march and its shared pieces, used by the single-view
// caster and by the batched forward / adjoint kernels, so every path runs the
// same operations in the same order.
#pragma once
#include <type_traits>

#include "rvk_internal.h"

namespace rvk {

// The reference's 16x16 ray-offset table (V/vtkCudaVolumeMapper.cxx:59-93),
// indexed by (x mod 16) + 16 (y mod 16).  Parity depends on it bit for bit.
static __constant__ float kRayOffsets[256] = {
  0.70554f, 0.53342f, 0.57951f, 0.28956f, 0.30194f, 0.77474f, 0.01401f, 0.76072f,
  0.81449f, 0.70903f, 0.04535f, 0.41403f, 0.86261f, 0.79048f, 0.37353f, 0.96195f,
  0.87144f, 0.05623f, 0.94955f, 0.36401f, 0.52486f, 0.76711f, 0.0535f, 0.59245f,
  0.4687f, 0.29816f, 0.62269f, 0.64782f, 0.26379f, 0.27934f, 0.8298f, 0.8246f,
  0.58916f, 0.98609f, 0.91096f, 0.22686f, 0.69511f, 0.98f, 0.24393f, 0.53387f,
  0.10636f, 0.99941f, 0.67617f, 0.0157f, 0.57518f, 0.10005f, 0.10302f, 0.79888f,
  0.28448f, 0.04564f, 0.29577f, 0.38201f, 0.30097f, 0.94857f, 0.97982f, 0.40137f,
  0.27827f, 0.16044f, 0.16282f, 0.64658f, 0.41007f, 0.41276f, 0.71273f, 0.3262f,
  0.63317f, 0.20756f, 0.18601f, 0.58335f, 0.08071f, 0.45797f, 0.90572f, 0.26136f,
  0.78521f, 0.3789f, 0.28966f, 0.91937f, 0.63174f, 0.62764f, 0.42845f, 0.09797f,
  0.56104f, 0.69448f, 0.91371f, 0.83481f, 0.02262f, 0.54336f, 0.91616f, 0.43026f,
  0.67794f, 0.50245f, 0.51373f, 0.46298f, 0.35347f, 0.40483f, 0.26973f, 0.05559f,
  0.24384f, 0.97907f, 0.06091f, 0.39029f, 0.36499f, 0.48989f, 0.15566f, 0.47445f,
  0.25726f, 0.62875f, 0.54207f, 0.1563f, 0.93854f, 0.65449f, 0.50608f, 0.39047f,
  0.10737f, 0.78399f, 0.45964f, 0.75368f, 0.59609f, 0.83273f, 0.01875f, 0.21036f,
  0.07395f, 0.10545f, 0.33169f, 0.12824f, 0.00024f, 0.53679f, 0.65705f, 0.54401f,
  0.82741f, 0.08189f, 0.19192f, 0.67891f, 0.4542f, 0.35702f, 0.14998f, 0.70439f,
  0.92878f, 0.53021f, 0.08964f, 0.75772f, 0.40184f, 0.46187f, 0.49216f, 0.20762f,
  0.32973f, 0.09542f, 0.58979f, 0.16987f, 0.92761f, 0.09792f, 0.44386f, 0.27294f,
  0.87254f, 0.75068f, 0.27294f, 0.67364f, 0.25662f, 0.08989f, 0.03095f, 0.32271f,
  0.79012f, 0.29725f, 0.23528f, 0.48047f, 0.2546f, 0.3406f, 0.04493f, 0.48242f,
  0.20601f, 0.86453f, 0.58862f, 0.7549f, 0.92788f, 0.33101f, 0.54294f, 0.08069f,
  0.63437f, 0.41003f, 0.96042f, 0.11462f, 0.92344f, 0.6202f, 0.34772f, 0.14924f,
  0.47997f, 0.2194f, 0.99373f, 0.13042f, 0.02888f, 0.34539f, 0.54766f, 0.92295f,
  0.53824f, 0.40642f, 0.84724f, 0.82622f, 0.67242f, 0.72189f, 0.99677f, 0.3398f,
  0.49521f, 0.41296f, 0.69528f, 0.17908f, 0.42291f, 0.54317f, 0.81466f, 0.54091f,
  0.42753f, 0.50906f, 0.22778f, 0.61918f, 0.48983f, 0.68081f, 0.8866f, 0.37051f,
  0.30249f, 0.29286f, 0.15031f, 0.52982f, 0.22326f, 0.58452f, 0.36345f, 0.87597f,
  0.47801f, 0.19063f, 0.68406f, 0.74741f, 0.61393f, 0.78213f, 0.16174f, 0.80777f,
  0.20261f, 0.95676f, 0.06585f, 0.06152f, 0.79319f, 0.3796f, 0.46358f, 0.11954f,
  0.11547f, 0.17377f, 0.04811f, 0.71481f, 0.53302f, 0.561f, 0.21673f, 0.468f,
  0.74635f, 0.75231f, 0.39893f, 0.90309f, 0.746f, 0.08855f, 0.63457f, 0.71302f};

__device__ __forceinline__ float ray_offset(int jitter, int x, int y) {
  return jitter ? kRayOffsets[(x & 15) + 16 * (y & 15)] : 0.0f;
}

__device__ __forceinline__ float ray_length(const VolGeom& g, const V3<float>& inc) {
  return sqrtf(inc.x * inc.x * g.sp[0] * g.sp[0] + inc.y * inc.y * g.sp[1] * g.sp[1] +
               inc.z * inc.z * g.sp[2] * g.sp[2]);
}

__device__ __forceinline__ void step(V3<float>& p, const V3<float>& inc) {
  p.x += inc.x; p.y += inc.y; p.z += inc.z;
}

// Sample coordinate in cell units (c = p + 0.5, what trilinear() takes).
// Reference mode keeps the incrementally updated p (p += inc per step, as the
// CUDA kernels do) and adds 0.5; PRECISE derives it from the step index,
// c = (p0 + 0.5) + (k_start - k) inc with one rounding per axis, which removes
// the drift repeated FP32 additions build up.
template <bool PRECISE>
__device__ __forceinline__ V3<float> sample_pos(const V3<float>& p, const V3<float>& p0c, const V3<float>& inc,
                                                float t) {
  if (!PRECISE) return V3<float>{p.x + 0.5f, p.y + 0.5f, p.z + 0.5f};
  return V3<float>{fmaf(t, inc.x, p0c.x), fmaf(t, inc.y, p0c.y), fmaf(t, inc.z, p0c.z)};
}
template <bool PRECISE>
__device__ __forceinline__ V3<float> sample_pos(const V3<float>& p, const V3<float>& p0c, const V3<float>& inc, int kst,
                                                int k) {
  return sample_pos<PRECISE>(p, p0c, inc, (float)(kst - k));
}

__device__ __forceinline__ V3<float> cell_coord(const V3<float>& p) {
  return V3<float>{p.x + 0.5f, p.y + 0.5f, p.z + 0.5f};
}

// The precise DRR march samples at c = p0c + t inc for t in [0, kst - 1], one
// fma per axis, which is monotone in t.  So when c at t = 0 and at t = kst
// lies in the cell range [0, n + 1] on every axis, every sample does and
// trilinear() needs no clamp.  Rays that pass the volume clip have both ends
// within [0, n - 1] in voxels (V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:297-311),
// so their samples lie in [0.5, n - 0.5] and always pass; a ray that fails
// (a non-finite set-up) takes no samples, in the oracle as here.  (The
// compositing casters keep the clamp: their +-1 gradient taps reach
// [-0.5, n + 0.5].)
__device__ __forceinline__ bool samples_in_cells(const VolGeom& g, const V3<float>& p0c, const V3<float>& inc,
                                                 int kst) {
  const float t = (float)kst;
  const float s[3] = {p0c.x, p0c.y, p0c.z};
  const float e[3] = {fmaf(t, inc.x, p0c.x), fmaf(t, inc.y, p0c.y), fmaf(t, inc.z, p0c.z)};
  bool ok = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float hi = (float)(g.n[a] + 1);
    ok = ok && s[a] >= 0.0f && s[a] <= hi && e[a] >= 0.0f && e[a] <= hi;
  }
  return ok;
}

// ===========================================================================
// DRR march (V/CUDA_vtkCudaDRRImageVolumeMapper_renderAlgo.cuh:11-117)
struct DrrMarchArgs {
  float intercept, slope, offset;
  int output, ref_term, jitter;
  const float2* ranges;  // per range brick (min, max) of the frame, or NULL
  int rbx, rby;          // range-brick grid extent in x and y
};

// Run skipping of the precise march.  Each range brick (kRangeEdge^3 cells)
// has the (min, max) of every value its cells can interpolate.  After a
// sample at cell coordinate q:
//  - if no value of q's brick can contribute (alpha = fma(cs, v, ci) >= 0 at
//    the bound that minimises it: alpha is monotone in v), the samples the
//    march takes next inside the brick do not contribute either;
//  - if the brick is constant (min == max: every corner equal, so every
//    sample in it interpolates that value exactly) and q contributed, the
//    next samples contribute the same alpha.
// Either way the march's state after such a run is known in closed form, so
// the run's gathers are skipped (a contributing run still adds its alphas
// one by one, in order: the sum is the same to the bit).
struct RangeWalk {
  float hx, hy, hz;  // the brick face a ray leaves through: kRangeEdge (ahead) or 0
  float rx, ry, rz;  // 1 / inc, +inf where inc = 0
};
__device__ __forceinline__ RangeWalk range_walk(const V3<float>& inc) {
  const float e = (float)kRangeEdge, inf = __builtin_huge_valf();
  return RangeWalk{inc.x >= 0.0f ? e : 0.0f, inc.y >= 0.0f ? e : 0.0f, inc.z >= 0.0f ? e : 0.0f,
                   inc.x != 0.0f ? 1.0f / inc.x : inf, inc.y != 0.0f ? 1.0f / inc.y : inf,
                   inc.z != 0.0f ? 1.0f / inc.z : inf};
}

// How many of the samples t1, t1 + stride, ... before lim lie in range brick
// (bx, by, bz), which holds the sample before t1.  The estimate from the
// brick's exit time is confirmed on its last sample, at the exact position
// the march computes: positions are monotone in t per axis, so when the
// sample before t1 and the last one lie in the (box-shaped) brick, all
// between do.  t1 and lim are integers, so the cap at lim is exact.
__device__ __forceinline__ int brick_run(int bx, int by, int bz, const V3<float>& p0c, const V3<float>& inc,
                                         const RangeWalk& w, float t1, float lim, float stride) {
  const float e = (float)kRangeEdge;
  const float tx = (fmaf((float)bx, e, w.hx) - p0c.x) * w.rx;
  const float ty = (fmaf((float)by, e, w.hy) - p0c.y) * w.ry;
  const float tz = (fmaf((float)bz, e, w.hz) - p0c.z) * w.rz;
  const float te = fminf(fminf(tx, ty), fminf(tz, lim));
  const float inv = stride == 1.0f ? 1.0f : 0.5f;
  int m = (int)ceilf((te - t1) * inv);
  for (int k = 0; k < 2 && m > 0; ++k, --m) {
    const float tl = fmaf(stride, (float)(m - 1), t1);
    const int cx = (int)fmaf(tl, inc.x, p0c.x), cy = (int)fmaf(tl, inc.y, p0c.y), cz = (int)fmaf(tl, inc.z, p0c.z);
    if ((cx >> kRangeLog2) == bx && (cy >> kRangeLog2) == by && (cz >> kRangeLog2) == bz) return m;
  }
  return 0;
}

__device__ __forceinline__ float2 range_brick(const DrrMarchArgs& a, const V3<float>& q, int& bx, int& by, int& bz) {
  bx = (int)q.x >> kRangeLog2;
  by = (int)q.y >> kRangeLog2;
  bz = (int)q.z >> kRangeLog2;
  return a.ranges[(bz * a.rby + by) * a.rbx + bx];
}

// Non-contributing samples to pass over after the one at q (the next at t1).
__device__ __forceinline__ int empty_run(const DrrMarchArgs& a, const V3<float>& q, const V3<float>& p0c,
                                         const V3<float>& inc, const RangeWalk& w, float cs, float ci, float t1,
                                         float lim, float stride) {
  int bx, by, bz;
  const float2 r = range_brick(a, q, bx, by, bz);
  if (!(fmaf(cs, cs <= 0.0f ? r.y : r.x, ci) >= 0.0f)) return 0;
  return brick_run(bx, by, bz, p0c, inc, w, t1, lim, stride);
}

// Contributing run after the sample at q (alpha; next sample at t): adds the
// run's alphas to acc and advances t.  Returns true when acc reached stop_at
// inside the run (the march ends there, as it would have).
template <typename Acc>
__device__ __forceinline__ bool flat_run(const DrrMarchArgs& a, const V3<float>& q, const V3<float>& p0c,
                                         const V3<float>& inc, const RangeWalk& w, float cs, float ci, float alpha,
                                         float& t, float lim, Acc& acc, Acc stop_at) {
  int bx, by, bz;
  const float2 r = range_brick(a, q, bx, by, bz);
  if (!(r.x == r.y && fmaf(cs, r.x, ci) == alpha)) return false;
  const int m = brick_run(bx, by, bz, p0c, inc, w, t, lim, 1.0f);
  for (int i = 0; i < m; ++i) {
    acc += (Acc)alpha;
    if (acc <= stop_at) return true;
  }
  t += (float)m;
  return false;
}

// Visitors declare whether a sample with a zero gradient may go unvisited
// (kFlatRuns): samples of constant bricks have gradient +0.
struct NoVisit {
  static constexpr bool kFlatRuns = true;
  __device__ __forceinline__ void operator()(const V3<float>&, float, float, const V3<float>&) const {}
};

// Line integral along one ray.  The accumulator starts at ln(256*offset) in
// the reference RGBA8 mode and at 0 in the float modes.  `visit(c, t, alpha,
// grad)` is called for every accumulated sample, with c its sample cell
// coordinate (c = p0 + 0.5 + t*inc in PRECISE mode); with VG the trilinear
// gradient at c comes from the same gathers as the value (the adjoint marches
// through this very control flow).  RUNS (precise mode only) takes the runs
// above when the volume frame has brick ranges (a.ranges); the result is the
// same to the bit.  Runs pay where the rays of a wave skip together (many
// views in one launch); a single view's launch is ordered for L2 locality,
// which runs unbalance, and marches without them.
template <typename CV, bool SKIP, bool PRECISE, bool VG = false, bool RUNS = false, bool KEY = true,
          bool TERM = true, typename F = NoVisit>
__device__ __forceinline__ float drr_march(CV vol, const VolGeom& g, const DrrMarchArgs& a,
                                           const Ray& r, int x, int y, F&& visit = F()) {
  using FV = typename std::remove_reference<F>::type;
  const bool ref_mode = a.output == RVK_DRR_RGBA8_REF;
  const float lg = logf(256.0f * a.offset);
  // PRECISE sums the line integral in double: a converted f32 sample and one
  // f64 add cost less than a compensated f32 sum and carry more precision
  // (the f64 add runs at the f32 rate).  The reference mode sums in f32.
  using Acc = typename std::conditional<PRECISE, double, float>::type;
  Acc acc = ref_mode ? (Acc)lg : (Acc)0;
  // termination at acc <= stop_at only with ref_term (otherwise never)
  const Acc stop_at = !a.ref_term ? (Acc)-__builtin_huge_valf() : (ref_mode ? (Acc)0 : (Acc)-lg);
  const float j = ray_offset(a.jitter, x, y);
  V3<float> p{r.start.x + j * r.inc.x, r.start.y + j * r.inc.y, r.start.z + j * r.inc.z};
  const V3<float> inc = r.inc;
  const V3<float> p0c = cell_coord(p);
  int kst = step_count(r.num_steps - j);
  if (PRECISE && !samples_in_cells(g, p0c, inc, kst)) kst = 0;
  const float ds = ray_length(g, inc);
  const float ci = a.intercept * -ds;
  const float cs = a.slope * -ds;
  // The march counts in t = kst - k (k = steps left), exactly, as a float:
  // sample t sits at p0 + t inc, the ray ends at t = kst, and the keyhole
  // window es >= k > ee is t in [kst - es, kst - ee).
  const int es = kst - (int)ceilf(r.exclude_start);
  const int ee = kst - (int)floorf(r.exclude_end);
  const float tend = (float)kst, kw0 = (float)(kst - es), kw1 = (float)(kst - ee);
  float t = 0.0f;
  // skip-rule state z = sb - sa in {-1, 0, 1}: sa = the sample is a
  // back-step, sb = the last step skipped a sample (never both), so each arm
  // updates one float in place
  float z = 0.0f;
  const bool runs = RUNS && PRECISE && a.ranges != nullptr;
  const RangeWalk w = range_walk(inc);
  // a contributing run must end before the keyhole window
  const float flat_lim = fminf(kw0, tend);
  // a repeated alpha is the cheap hint of a constant brick
  float alpha_prev = __builtin_nanf("");
  while (t < tend) {
    // KEY = false: the view has no keyhole planes, so no ray has a window
    const bool in_key = KEY && (t >= kw0 && t < kw1);
    const V3<float> q = sample_pos<PRECISE>(p, p0c, inc, t);
    V3<float> gr{0.0f, 0.0f, 0.0f};
    const float alpha = fmaf(cs, VG ? trilinear_vg<!PRECISE>(vol, g, q.x, q.y, q.z, gr)
                                    : trilinear<!PRECISE>(vol, g, q.x, q.y, q.z), ci);
    const bool repeat = alpha == alpha_prev;
    alpha_prev = alpha;
    const float ts = t;
    if (SKIP) {
      // t stays an exact integer, so one add of +-1 or 2 equals the
      // reference's single steps; a back-stepped sample adds +0 (acc is
      // never -0), which leaves the sum unchanged without a select
      if (alpha < 0.0f && !in_key) {
        // sa' = sb, sb' = 0: z' = -sb = median(-z, -1, 0), and sa' = -z'
        z = __builtin_amdgcn_fmed3f(-z, -1.0f, 0.0f);
        if (!PRECISE) {
          if (z != 0.0f) { p.x = p.x - inc.x; p.y = p.y - inc.y; p.z = p.z - inc.z; }
          else { p.x = p.x + inc.x; p.y = p.y + inc.y; p.z = p.z + inc.z; }
        }
        t += fmaf(2.0f, z, 1.0f);  // 1 - 2 sa'
        if (VG && z == 0.0f) visit(q, ts, alpha, gr);
        // alpha - alpha * sa': alpha, or +0 for a back-stepped sample
        acc += (Acc)fmaf(alpha, z, alpha);
        if (TERM && acc <= stop_at) break;
        // each further sample of a contributing run steps 1 (sa = sb = 0)
        if (FV::kFlatRuns && runs && repeat && z == 0.0f) {
          if (flat_run(a, q, p0c, inc, w, cs, ci, alpha, t, t >= kw1 ? tend : flat_lim, acc, stop_at)) break;
        }
      } else {
        if (!PRECISE) {
          if (z >= 0.0f) step(p, inc);  // sa == 0
          step(p, inc);
        }
        // sa' = 0, sb' = 1 - sa = min(z + 1, 1) (z + 1 >= 0); the step is 2 - sa
        z = __builtin_amdgcn_fmed3f(z + 1.0f, 0.0f, 1.0f);
        t += 1.0f + z;
        // each further non-contributing sample steps 2 and leaves sb = 1
        if (runs) {
          const int m = empty_run(a, q, p0c, inc, w, cs, ci, t, tend, 2.0f);
          if (m > 0) {
            t += 2.0f * (float)m;
            z = 1.0f;
          }
        }
      }
    } else {
      if (!PRECISE) step(p, inc);
      t += 1.0f;
      if (alpha < 0.0f && !in_key) {
        acc += (Acc)alpha;
        if (VG) visit(q, ts, alpha, gr);
        if (TERM && acc <= stop_at) break;
        if (FV::kFlatRuns && runs && repeat) {
          if (flat_run(a, q, p0c, inc, w, cs, ci, alpha, t, t >= kw1 ? tend : flat_lim, acc, stop_at)) break;
        }
      } else if (runs) {
        t += (float)empty_run(a, q, p0c, inc, w, cs, ci, t, tend, 1.0f);
      }
    }
  }
  const float L = (float)acc;
  if (ref_mode) {
    const float o = 256.0f - expf(L);
    return o < 0.0f ? 0.0f : o;
  }
  if (a.output == RVK_DRR_F32_EXP_NEG_L) return expf(L);
  return -L;
}

}  // namespace rvk
