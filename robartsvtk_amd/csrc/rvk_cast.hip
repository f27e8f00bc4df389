// rvk_cast.hip -- single-view gfx950 ray casters (DRR, 1D TF, 2D TF).
//
// One thread per output pixel; a 256-thread workgroup covers a 16x16 tile and
// each wave64 an 8x8 quadrant, so a wave's 64 samples at one step fall in a
// few neighbouring 4x4x4 volume bricks.  formRays is fused into each kernel's
// prologue (ray state stays in registers instead of 9 SoA arrays), and all
// parameters are kernel arguments.
//
// Reference kernels restated (V/ = Visualization/vtkCudaVisualization):
//   DRR  V/CUDA_vtkCudaDRRImageVolumeMapper_renderAlgo.cuh:11-171
//   1D   V/CUDA_vtkCuda1DVolumeMapper_renderAlgo.cuh:27-227
//   2D   V/CUDA_vtkCuda2DVolumeMapper_renderAlgo.cuh:30-470
//   post V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:449-512
#include <cmath>
#include <mutex>

#include "rvk_internal.h"
#include "rvk_march.h"

namespace rvk {

struct DrrArgs {
  VolGeom g;
  ViewGeom v;
  float step_len;
  DrrMarchArgs m;
  uint32_t tint;
  const void* vol;
  void* out;
  int tiles_x;
};

// KEY = false: the view has no keyhole planes; TERM = false: no termination
// (the accumulated sum is never compared).
template <typename CV, bool SKIP, bool PRECISE, bool KEY, bool TERM>
__global__ __launch_bounds__(256) void drr_kernel(DrrArgs a) {
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  int x, y;
  tile_pixel(tile, a.tiles_x, x, y);
  if (x >= a.v.W || y >= a.v.H) return;
  const Ray r = form_ray(a.g, a.v, a.step_len, x, y);
  const float v = drr_march<CV, SKIP, PRECISE, false, false, KEY, TERM>(cells_of<CV>(a.vol), a.g, a.m, r, x, y);
  const long idx = (long)y * a.v.W + x;
  if (a.m.output == RVK_DRR_RGBA8_REF) {
    // V/...DRR...cuh:163-166
    const float tw = (float)(a.tint >> 24);
    const float q = tw / (tw + v);
    const uint32_t A = to_u8(tw + (1.0f - tw / 255.0f) * v);
    const uint32_t R = to_u8(q * (float)(a.tint & 255u));
    const uint32_t G = to_u8(q * (float)((a.tint >> 8) & 255u));
    const uint32_t B = to_u8(q * (float)((a.tint >> 16) & 255u));
    ((uint32_t*)a.out)[idx] = R | (G << 8) | (B << 16) | (A << 24);
  } else {
    ((float*)a.out)[idx] = v;
  }
}

// ===========================================================================
// 1D transfer function
struct Cast1dArgs {
  VolGeom g;
  ViewGeom v;
  float step_len, term_T;
  int jitter;
  int n;
  rvk_tf1d_params p;
  const float* tables;
  const void* vol;
  uint32_t* out;
  float* depth;
  float* min_depth;
  float* max_depth;
  int tiles_x;
};

// V/...1D...cuh:27-172.  GRAD=false is taken only when the gradient cannot
// change the result: diffuse = specular = 0, specular power >= 0 and no
// gradient-opacity table (non-finite multiplier); then shadeD = ambient and
// shadeS = 0 exactly, whatever the Phong-Lambert factor.
template <typename CV, bool SKIP, bool GRAD, bool PRECISE>
__global__ __launch_bounds__(256) void cast1d_kernel(Cast1dArgs a) {
  extern __shared__ float lds[];
  const int n = a.n;
  for (int i = threadIdx.x; i < 5 * n; i += blockDim.x) lds[i] = a.tables[i];
  __syncthreads();
  const float* t_alpha = lds;
  const float* t_galpha = lds + n;
  const float* t_r = lds + 2 * n;
  const float* t_g = lds + 3 * n;
  const float* t_b = lds + 4 * n;

  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  int x, y;
  tile_pixel(tile, a.tiles_x, x, y);
  if (x >= a.v.W || y >= a.v.H) return;
  const Ray r = form_ray(a.g, a.v, a.step_len, x, y);
  const CV vol = cells_of<CV>(a.vol);
  const VolGeom& g = a.g;

  float Tr = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
  const float lo = a.p.intensity_low, mult = a.p.intensity_multiplier;
  const float glo = a.p.gradient_low, gmult = a.p.gradient_multiplier;
  const float amb = a.p.ambient, dif = a.p.diffuse, spc = a.p.specular, spw = a.p.specular_power;
  const bool use_galpha = isfinite(gmult);
  const float term = a.term_T;
  const float j = ray_offset(a.jitter, x, y);
  V3<float> p{r.start.x + j * r.inc.x, r.start.y + j * r.inc.y, r.start.z + j * r.inc.z};
  const V3<float> inc = r.inc;
  const V3<float> p0c = cell_coord(p);
  int k = step_count(r.num_steps - j);
  const int k0 = k;
  const float rlen = ray_length(g, inc);
  const float sx = g.sp[0], sy = g.sp[1], sz = g.sp[2];
  const int es = (int)ceilf(r.exclude_start), ee = (int)floorf(r.exclude_end);
  const int kst = k;
  int sa = 0, sb = 0;

  auto accumulate = [&](float alpha, float u) {
    float shd, shs;
    if (GRAD) {
      V3<float> gr;
      const V3<float> q = sample_pos<PRECISE>(p, p0c, inc, kst, k);
      gr.x = (trilinear(vol, g, q.x + 1.0f, q.y, q.z) - trilinear(vol, g, q.x - 1.0f, q.y, q.z)) * g.srecip[0];
      gr.y = (trilinear(vol, g, q.x, q.y + 1.0f, q.z) - trilinear(vol, g, q.x, q.y - 1.0f, q.z)) * g.srecip[1];
      gr.z = (trilinear(vol, g, q.x, q.y, q.z + 1.0f) - trilinear(vol, g, q.x, q.y, q.z - 1.0f)) * g.srecip[2];
      const float gm = sqrtf(gr.x * gr.x + gr.y * gr.y + gr.z * gr.z);
      alpha *= use_galpha ? tf1d_lookup(t_galpha, n, gmult * (gm - glo)) : 1.0f;
      const float pl = satf(fabsf(gr.x * inc.x * sx + gr.y * inc.y * sy + gr.z * inc.z * sz) / (gm * rlen));
      shd = amb + dif * pl;
      shs = spc * powf(pl, spw);
    } else {
      shd = amb;
      shs = 0.0f;
    }
    const float m = Tr * alpha;
    Tr *= (1.0f - alpha);
    cr += m * satf(shd * tf1d_lookup(t_r, n, u) + shs);
    cg += m * satf(shd * tf1d_lookup(t_g, n, u) + shs);
    cb += m * satf(shd * tf1d_lookup(t_b, n, u) + shs);
  };

  while (k > 0) {
    // the 1D caster compares unconverted exclusion bounds (V/...1D...cuh:74)
    if (es > k && ee < k) {
      p.x += (float)(k - ee) * inc.x; p.y += (float)(k - ee) * inc.y; p.z += (float)(k - ee) * inc.z;
      k = ee;
      sa = 0;
      sb = 0;
      continue;
    }
    const V3<float> q = sample_pos<PRECISE>(p, p0c, inc, kst, k);
    // (no register cell cache here: c1's cells stay in the caches, and the
    // compare and branch cost more than the loads they save: 0.094 -> 0.102 ms)
    const float u = mult * (trilinear(vol, g, q.x, q.y, q.z) - lo);
    const float alpha = tf1d_lookup(t_alpha, n, u);
    if (SKIP) {
      if (alpha > 0.0f) {
        sa = sb;
        sb = 0;
        if (sa) { p.x = p.x - inc.x; p.y = p.y - inc.y; p.z = p.z - inc.z; k += 1; }
        else { p.x = p.x + inc.x; p.y = p.y + inc.y; p.z = p.z + inc.z; k -= 1; }
        if (!sa) accumulate(alpha, u);
        if (Tr < term) { Tr = 0.0f; break; }
      } else {
        if (!sa) { step(p, inc); k--; }
        sb = !sa;
        step(p, inc);
        k--;
        sa = 0;
      }
    } else {
      step(p, inc);
      k--;
      if (alpha > 0.0f) {
        accumulate(alpha, u);
        if (Tr < term) { Tr = 0.0f; break; }
      }
    }
  }
  const long idx = (long)y * a.v.W + x;
  const uint32_t R = to_u8(255.0f * satf(cr)), G = to_u8(255.0f * satf(cg)), B = to_u8(255.0f * satf(cb));
  const uint32_t A = to_u8(255.0f * (1.0f - Tr));
  a.out[idx] = R | (G << 8) | (B << 16) | (A << 24);
  if (a.depth) a.depth[idx] = (float)(k0 - k);
  if (a.min_depth) a.min_depth[idx] = r.min_depth;
  if (a.max_depth) a.max_depth[idx] = r.max_depth;
}

// ===========================================================================
// 2D transfer function
struct Cast2dArgs {
  VolGeom g;
  ViewGeom v;
  float step_len, term_T;
  int jitter;
  int n;
  rvk_tf2d_params p;
  const Tf2dCell* cells;
  const Tf2dCell* kcells;
  const float4* lit;     // the cells lit with a Phong-Lambert factor of 1 (GMAG)
  const float4* klit;
  const uint32_t* occ;   // opacity bitmaps staged in LDS, or NULL
  const uint32_t* kocc;
  int occ_words;
  const uint32_t* col;   // column bitmaps (TF, then keyhole TF), staged in LDS
  int col_words;
  const void* vol;
  const uint8_t* gradmag;
  float gradmag_scale;
  uint32_t* out;
  float* depth;
  float* min_depth;
  float* max_depth;
  int tiles_x;
};

// V/...2D...cuh:30-351.  DUAL selects the keyhole-TF variant (WithDual).
// GMAG reads |grad| from the precomputed u8 volume instead of 6 extra gathers;
// without a direction the Phong-Lambert factor is taken as 1, so a table
// cell's whole contribution (alpha and the saturated shaded colour) is
// precomputed and one 16-byte read serves a composited sample.
// With an opacity bitmap in LDS the empty test (alpha > 0 is false) needs no
// table fetch; the cell is read only for samples that contribute.  The value
// and gradient-magnitude cells are kept in registers between samples (a new
// load only when a ray moves to another cell).
template <typename CV, bool SKIP, bool DUAL, bool GMAG, bool PRECISE>
__global__ __launch_bounds__(256) void cast2d_kernel(Cast2dArgs a) {
  extern __shared__ uint32_t s_lds[];
  const int cw = a.col_words;
  for (int i = threadIdx.x; i < (DUAL ? 2 : 1) * cw; i += 256) s_lds[i] = a.col[i];
  uint32_t* s_occ = s_lds + (DUAL ? 2 : 1) * cw;
  const bool use_occ = a.occ != nullptr;
  if (use_occ) {
    const int w = a.occ_words;
    for (int i = threadIdx.x; i < w; i += 256) s_occ[i] = a.occ[i];
    if (DUAL)
      for (int i = threadIdx.x; i < w; i += 256) s_occ[w + i] = a.kocc[i];
  }
  __syncthreads();
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  int x, y;
  tile_pixel(tile, a.tiles_x, x, y);
  if (x >= a.v.W || y >= a.v.H) return;
  const Ray r = form_ray(a.g, a.v, a.step_len, x, y);
  const CV vol = cells_of<CV>(a.vol);
  const VolGeom& g = a.g;
  const int n = a.n;
  CellCache<typename CV::value_type, CV::kLog2> vcache;
  CellCache<uint8_t, CV::kLog2> gcache;

  float Tr = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
  const float lo = a.p.intensity_low, mult = a.p.intensity_multiplier;
  const float glo = a.p.gradient_low, gmult = a.p.gradient_multiplier, goff = a.p.gradient_offset;
  const float term = a.term_T;
  const float j = ray_offset(a.jitter, x, y);
  V3<float> p{r.start.x + j * r.inc.x, r.start.y + j * r.inc.y, r.start.z + j * r.inc.z};
  const V3<float> inc = r.inc;
  const V3<float> p0c = cell_coord(p);
  int k = step_count(r.num_steps - j);
  const int k0 = k;
  const float rlen = ray_length(g, inc);
  const float sx = g.sp[0], sy = g.sp[1], sz = g.sp[2];
  // WithDual rounds the exclusion bounds up/down, WithoutDual down/up
  int es = DUAL ? (int)ceilf(r.exclude_start) : (int)floorf(r.exclude_start);
  int ee = DUAL ? (int)floorf(r.exclude_end) : (int)ceilf(r.exclude_end);
  es = k - es;
  ee = k - ee;
  const int kst = k;
  int sa = 0, sb = 0;
  while (k > 0) {
    const bool in_key = (es >= k && ee < k);
    if (!DUAL && in_key) {
      p.x += (float)(k - ee) * inc.x; p.y += (float)(k - ee) * inc.y; p.z += (float)(k - ee) * inc.z;
      k = ee;
      sa = 0;
      sb = 0;
      continue;
    }
    const V3<float> q = sample_pos<PRECISE>(p, p0c, inc, kst, k);
    const float u = mult * (vcache.sample(vol, g, q.x, q.y, q.z) - lo);
    const Tf2dCell* __restrict__ cells = (DUAL && in_key) ? a.kcells : a.cells;
    const float4* __restrict__ lits = (DUAL && in_key) ? a.klit : a.lit;
    V3<float> gr{0.0f, 0.0f, 0.0f};
    float gmag = 0.0f;
    int ti = 0;
    float4 c0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // a sample in a TF column without opacity cannot composite, whatever its
    // gradient: its gradient gathers and table fetch are skipped
    if (tf2d_column_live(s_lds + ((DUAL && in_key) ? cw : 0), tf2d_axis(n, u))) {
      float gidx;
      if (GMAG) {
        gmag = gcache.sample(Cells<uint8_t, CV::kLog2>{a.gradmag}, g, q.x, q.y, q.z) * a.gradmag_scale;
        gidx = gmult * (log2f(gmag * gmag + goff) + glo);
      } else {
        gr.x = (trilinear(vol, g, q.x + 1.0f, q.y, q.z) - trilinear(vol, g, q.x - 1.0f, q.y, q.z)) * g.srecip[0];
        gr.y = (trilinear(vol, g, q.x, q.y + 1.0f, q.z) - trilinear(vol, g, q.x, q.y - 1.0f, q.z)) * g.srecip[1];
        gr.z = (trilinear(vol, g, q.x, q.y, q.z + 1.0f) - trilinear(vol, g, q.x, q.y, q.z - 1.0f)) * g.srecip[2];
        const float gm2 = gr.x * gr.x + gr.y * gr.y + gr.z * gr.z;
        gidx = gmult * (log2f(gm2 + goff) + glo);
        gmag = sqrtf(gm2);
      }
      ti = tf2d_index(n, u, gidx);
      if (use_occ) {
        const uint32_t* occ = s_occ + ((DUAL && in_key) ? a.occ_words : 0);
        if ((occ[ti >> 5] >> (ti & 31)) & 1u) c0 = GMAG ? lits[ti] : *(const float4*)&cells[ti].alpha;
      } else {
        c0 = GMAG ? lits[ti] : *(const float4*)&cells[ti].alpha;
      }
    }
    const float alpha = c0.x;
    auto accumulate = [&]() {
      if (GMAG) {
        // c0 = (alpha, lit r, g, b): the same values the general arm forms
        const float m = Tr * alpha;
        Tr *= (1.0f - alpha);
        cr += m * c0.y;
        cg += m * c0.z;
        cb += m * c0.w;
        return;
      }
      const float4 c1 = *(const float4*)&cells[ti].spec_power;
      const float pl = GMAG ? 1.0f
                            : satf(fabsf(gr.x * inc.x * sx + gr.y * inc.y * sy + gr.z * inc.z * sz) / (gmag * rlen));
      const float shd = c0.y + c0.z * pl;
      const float shs = c0.w * powf(pl, c1.x);
      const float m = Tr * alpha;
      Tr *= (1.0f - alpha);
      cr += m * satf(shd * c1.y + shs);
      cg += m * satf(shd * c1.z + shs);
      cb += m * satf(shd * c1.w + shs);
    };
    if (SKIP) {
      if (alpha > 0.0f) {
        sa = sb;
        sb = 0;
        if (sa) { p.x = p.x - inc.x; p.y = p.y - inc.y; p.z = p.z - inc.z; k += 1; }
        else { p.x = p.x + inc.x; p.y = p.y + inc.y; p.z = p.z + inc.z; k -= 1; }
        if (!sa) accumulate();
        if (Tr < term) { Tr = 0.0f; break; }
      } else {
        if (!sa) { step(p, inc); k--; }
        sb = !sa;
        step(p, inc);
        k--;
        sa = 0;
      }
    } else {
      step(p, inc);
      k--;
      if (alpha > 0.0f) {
        accumulate();
        if (Tr < term) { Tr = 0.0f; break; }
      }
    }
  }
  const long idx = (long)y * a.v.W + x;
  const uint32_t R = to_u8(255.0f * cr), G = to_u8(255.0f * cg), B = to_u8(255.0f * cb);
  const uint32_t A = to_u8(255.0f * (1.0f - Tr));
  a.out[idx] = R | (G << 8) | (B << 16) | (A << 24);
  if (a.depth) a.depth[idx] = (float)(k0 - k);
  if (a.min_depth) a.min_depth[idx] = r.min_depth;
  if (a.max_depth) a.max_depth[idx] = r.max_depth;
}

// ===========================================================================
// Post passes (V/...renderAlgo.cu:449-512), fused into one kernel per pixel.
// The depth normalisation of a pixel's neighbours is recomputed locally, so
// no pass ordering between pixels is needed.  INEX is the In/Ex-logic
// mapper's cel shade (V/CUDA_vtkCuda2DInExLogicVolumeMapper_renderAlgo.cuh:182-239):
// the squared depth difference without the square root, and pixels with zero
// alpha get alpha = (char)(255 - 255 g), a float-to-signed-char conversion
// (truncating, clamped to [-128, 127], NaN -> 0) stored as a byte.
template <bool INEX>
__global__ __launch_bounds__(256) void post_shade_kernel(int W, int H, rvk_shade_params s, const float* depth,
                                                         const float* mind, const float* maxd, uint32_t* rgba,
                                                         float* depth_out) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= W || y >= H) return;
  auto norm = [&](long i) {
    const float c = depth[i], mx = maxd[i], mn = mind[i];
    return (mx > 0.0f) ? (c + mn) / mx : 1.0f;
  };
  const long i = (long)y * W + x;
  const float d0 = norm(i);
  const float dy = norm((long)min(y + 1, H - 1) * W + x);
  const float dx = norm((long)y * W + min(x + 1, W - 1));
  float gm = (dx - d0) * (dx - d0) + (dy - d0) * (dy - d0);
  if (!INEX) gm = sqrt_rz(gm);
  gm = 1.0f - s.celr * satf((gm - s.cela) * s.celc);
  gm *= 1.0f - s.disr * satf((d0 - s.disa) * s.disc);
  const uint32_t c = rgba[i];
  const uint32_t R = to_u8(gm * (float)(c & 255u)), G = to_u8(gm * (float)((c >> 8) & 255u));
  const uint32_t B = to_u8(gm * (float)((c >> 16) & 255u));
  uint32_t A = c & 0xff000000u;
  if (INEX && A == 0u) {
    const float v = 255.0f - 255.0f * gm;
    const int t = (v == v) ? (int)fmaxf(-128.0f, fminf(127.0f, truncf(v))) : 0;
    A = (uint32_t)(uint8_t)(int8_t)t << 24;
  }
  rgba[i] = R | (G << 8) | (B << 16) | A;
  if (depth_out) depth_out[i] = d0;
}

// ===========================================================================
// Ray-sample count of the plain march without opacity termination: the sum
// over pixels of floor(num_steps - offset).  This is the DRR's metric
// denominator; it depends only on geometry, not on voxel values.
__global__ __launch_bounds__(256) void count_steps_kernel(VolGeom g, ViewGeom v, float step_len, int jitter,
                                                          int tiles_x, unsigned long long* out) {
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  int x, y;
  tile_pixel(tile, tiles_x, x, y);
  unsigned long long c = 0;
  if (x < v.W && y < v.H) {
    const Ray r = form_ray(g, v, step_len, x, y);
    c = (unsigned long long)step_count(r.num_steps - ray_offset(jitter, x, y));
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

// ===========================================================================
}  // namespace rvk

using namespace rvk;

extern "C" {

int rvk_render_drr(rvk_ctx* ctx, const rvk_volume* vol, int frame, const rvk_view* view, const rvk_march_params* march,
                   const rvk_drr_params* drr, void* out) {
  if (!ctx || !drr || !out) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  int rc = check_frame(vol, frame);
  if (!rc) rc = check_view(view);
  if (!rc) rc = check_march(march);
  if (rc) return rc;
  if (drr->output < RVK_DRR_RGBA8_REF || drr->output > RVK_DRR_F32_L)
    return fail(RVK_ERR_INVALID_ARGUMENT, "unknown DRR output mode %d", drr->output);
  if (!(drr->ct_offset > 0.0f)) return fail(RVK_ERR_INVALID_ARGUMENT, "ct_offset must be positive");
  RVK_HIP(hipSetDevice(ctx->device));
  DrrArgs a;
  a.g = vol_geom(vol);
  a.v = view_geom(view);
  a.step_len = min_spacing(vol) * march->step_scale;
  a.m.jitter = march->jitter;
  a.m.ref_term = drr->ref_termination;
  a.m.intercept = drr->ct_intercept;
  a.m.slope = drr->ct_slope;
  a.m.offset = drr->ct_offset;
  a.m.output = drr->output;
  a.tint = (uint32_t)drr->tint[0] | ((uint32_t)drr->tint[1] << 8) | ((uint32_t)drr->tint[2] << 16) |
           ((uint32_t)drr->tint[3] << 24);
  a.vol = vol->frames[frame];
  a.out = out;
  const int nblk = launch_grid(view, a.tiles_x);
  return dispatch_frame(vol, frame, [&](auto tag) {
    using CV = decltype(tag);
    const bool pr = march->precise != 0, skip = march->skip_rule != 0;
#define RVK_LDRR2(S, P, K)                                                                                 \
  do {                                                                                                     \
    if (drr->ref_termination)                                                                              \
      hipLaunchKernelGGL((drr_kernel<CV, S, P, K, true>), dim3(nblk), dim3(256), 0, ctx->stream, a);       \
    else                                                                                                   \
      hipLaunchKernelGGL((drr_kernel<CV, S, P, K, false>), dim3(nblk), dim3(256), 0, ctx->stream, a);      \
  } while (0)
#define RVK_LDRR(S, P)                                                                                     \
  do {                                                                                                     \
    if (view->n_keyhole_planes) RVK_LDRR2(S, P, true);                                                     \
    else RVK_LDRR2(S, P, false);                                                                           \
  } while (0)
    if (skip && pr) RVK_LDRR(true, true);
    else if (skip) RVK_LDRR(true, false);
    else if (pr) RVK_LDRR(false, true);
    else RVK_LDRR(false, false);
#undef RVK_LDRR
#undef RVK_LDRR2
    RVK_HIP(hipGetLastError());
    return (int)RVK_OK;
  });
}

int rvk_render_1d(rvk_ctx* ctx, const rvk_volume* vol, int frame, const rvk_tf1d* tf, const rvk_view* view,
                  const rvk_march_params* march, uint8_t* out, const rvk_aux_out* aux) {
  if (!ctx || !tf || !out) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  int rc = check_frame(vol, frame);
  if (!rc) rc = check_view(view);
  if (!rc) rc = check_march(march);
  if (rc) return rc;
  RVK_HIP(hipSetDevice(ctx->device));
  Cast1dArgs a;
  a.g = vol_geom(vol);
  a.v = view_geom(view);
  a.step_len = min_spacing(vol) * march->step_scale;
  a.term_T = march->term_T;
  a.jitter = march->jitter;
  a.n = tf->n;
  a.p = tf->p;
  a.tables = tf->tables;
  a.vol = vol->frames[frame];
  a.out = (uint32_t*)out;
  a.depth = aux ? aux->depth : nullptr;
  a.min_depth = aux ? aux->min_depth : nullptr;
  a.max_depth = aux ? aux->max_depth : nullptr;
  const int nblk = launch_grid(view, a.tiles_x);
  const size_t lds = sizeof(float) * 5 * (size_t)tf->n;
  const bool grad = !(tf->p.diffuse == 0.0f && tf->p.specular == 0.0f && tf->p.specular_power >= 0.0f &&
                      !std::isfinite(tf->p.gradient_multiplier));
  const bool skip = march->skip_rule != 0, pr = march->precise != 0;
  return dispatch_frame(vol, frame, [&](auto tag) {
    using CV = decltype(tag);
#define RVK_L1D(S, G, P) \
  hipLaunchKernelGGL((cast1d_kernel<CV, S, G, P>), dim3(nblk), dim3(256), lds, ctx->stream, a)
    if (pr) {
      if (skip) { if (grad) RVK_L1D(true, true, true); else RVK_L1D(true, false, true); }
      else { if (grad) RVK_L1D(false, true, true); else RVK_L1D(false, false, true); }
    } else {
      if (skip) { if (grad) RVK_L1D(true, true, false); else RVK_L1D(true, false, false); }
      else { if (grad) RVK_L1D(false, true, false); else RVK_L1D(false, false, false); }
    }
#undef RVK_L1D
    RVK_HIP(hipGetLastError());
    return (int)RVK_OK;
  });
}

int rvk_render_2d(rvk_ctx* ctx, const rvk_volume* vol, int frame, const rvk_tf2d* tf, const rvk_view* view,
                  const rvk_march_params* march, int use_gradmag, uint8_t* out, const rvk_aux_out* aux) {
  if (!ctx || !tf || !out) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  int rc = check_frame(vol, frame);
  if (!rc) rc = check_view(view);
  if (!rc) rc = check_march(march);
  if (rc) return rc;
  if (use_gradmag && !vol->gradmag[frame])
    return fail(RVK_ERR_NOT_READY, "use_gradmag set but frame %d has no gradient-magnitude volume", frame);
  RVK_HIP(hipSetDevice(ctx->device));
  Cast2dArgs a;
  a.g = vol_geom(vol);
  a.v = view_geom(view);
  a.step_len = min_spacing(vol) * march->step_scale;
  a.term_T = march->term_T;
  a.jitter = march->jitter;
  a.n = tf->n;
  a.p = tf->p;
  a.cells = tf->cells;
  a.kcells = tf->kcells;
  a.lit = tf->lit;
  a.klit = tf->klit;
  a.occ = tf->occ;
  a.kocc = tf->kocc;
  a.occ_words = tf->occ_words;
  a.col = tf->col;
  a.col_words = tf->col_words;
  // the keyhole bitmap is needed exactly when the keyhole TF is
  if (tf->kcells && !tf->kocc) a.occ = nullptr;
  const int ntf = tf->kcells ? 2 : 1;
  const size_t lds = sizeof(uint32_t) * ((size_t)tf->col_words * ntf + (a.occ ? (size_t)a.occ_words * ntf : 0));
  a.vol = vol->frames[frame];
  a.gradmag = use_gradmag ? vol->gradmag[frame] : nullptr;
  a.gradmag_scale = use_gradmag ? vol->gradmag_scale[frame] : 0.0f;
  a.out = (uint32_t*)out;
  a.depth = aux ? aux->depth : nullptr;
  a.min_depth = aux ? aux->min_depth : nullptr;
  a.max_depth = aux ? aux->max_depth : nullptr;
  const int nblk = launch_grid(view, a.tiles_x);
  const bool skip = march->skip_rule != 0, dual = tf->kcells != nullptr, gm = use_gradmag != 0;
  const bool pr = march->precise != 0;
  return dispatch_frame(vol, frame, [&](auto tag) {
    using CV = decltype(tag);
#define RVK_L2D(S, D, G)                                                                              \
  do {                                                                                                \
    if (pr) hipLaunchKernelGGL((cast2d_kernel<CV, S, D, G, true>), dim3(nblk), dim3(256), lds, ctx->stream, a);  \
    else hipLaunchKernelGGL((cast2d_kernel<CV, S, D, G, false>), dim3(nblk), dim3(256), lds, ctx->stream, a);    \
  } while (0)
    if (skip) {
      if (dual) { if (gm) RVK_L2D(true, true, true); else RVK_L2D(true, true, false); }
      else { if (gm) RVK_L2D(true, false, true); else RVK_L2D(true, false, false); }
    } else {
      if (dual) { if (gm) RVK_L2D(false, true, true); else RVK_L2D(false, true, false); }
      else { if (gm) RVK_L2D(false, false, true); else RVK_L2D(false, false, false); }
    }
#undef RVK_L2D
    RVK_HIP(hipGetLastError());
    return (int)RVK_OK;
  });
}

static int post_shade(rvk_ctx* ctx, bool inex, int width, int height, const rvk_shade_params* shade, float* depth,
                      const float* min_depth, const float* max_depth, uint8_t* rgba) {
  if (!ctx || !shade || !depth || !min_depth || !max_depth || !rgba)
    return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (width <= 0 || height <= 0) return fail(RVK_ERR_INVALID_ARGUMENT, "bad image size");
  RVK_HIP(hipSetDevice(ctx->device));
  dim3 grid((width + 15) / 16, (height + 15) / 16);
  // depth is normalised in place only after every pixel has read its
  // neighbours, so the normalised copy goes to a second buffer first.
  float* tmp = nullptr;
  RVK_HIP(hipMallocAsync((void**)&tmp, sizeof(float) * (size_t)width * height, ctx->stream));
  if (inex)
    hipLaunchKernelGGL(post_shade_kernel<true>, grid, dim3(256), 0, ctx->stream, width, height, *shade, depth,
                       min_depth, max_depth, (uint32_t*)rgba, tmp);
  else
    hipLaunchKernelGGL(post_shade_kernel<false>, grid, dim3(256), 0, ctx->stream, width, height, *shade, depth,
                       min_depth, max_depth, (uint32_t*)rgba, tmp);
  RVK_HIP(hipGetLastError());
  RVK_HIP(hipMemcpyAsync(depth, tmp, sizeof(float) * (size_t)width * height, hipMemcpyDeviceToDevice, ctx->stream));
  RVK_HIP(hipFreeAsync(tmp, ctx->stream));
  return RVK_OK;
}

int rvk_post_shade(rvk_ctx* ctx, int width, int height, const rvk_shade_params* shade, float* depth,
                   const float* min_depth, const float* max_depth, uint8_t* rgba) {
  return post_shade(ctx, false, width, height, shade, depth, min_depth, max_depth, rgba);
}

// The allocation-free post pass: its normalised-depth staging (the buffer
// rvk_post_shade allocates per call) is a file-scope device array, one copy
// per GPU, held until the process ends.  Host code reaches a GPU's copy only
// through the address hipGetSymbolAddress returns on that device (the
// array's name on the host is a stand-in, not a device address).
static __device__ __align__(256) float scratch[RVK_POST_SHADE_NOALLOC_MAX_PIXELS];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch[kMaxScratchDevices] = {};

// This device's copy of `scratch`, looked up on the first call on it (the
// device must be current).
int scratch_on(int device, float** out) {
  if (device < 0 || device >= kMaxScratchDevices) return fail(RVK_ERR_INVALID_ARGUMENT, "device %d", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    float* p = nullptr;
    RVK_HIP(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)));
    g_scratch[device] = p;
  }
  *out = g_scratch[device];
  return RVK_OK;
}
}  // namespace

int rvk_post_shade_noalloc(rvk_ctx* ctx, int width, int height, const rvk_shade_params* shade, float* depth,
                           const float* min_depth, const float* max_depth, uint8_t* rgba) {
  if (!ctx || !shade || !depth || !min_depth || !max_depth || !rgba)
    return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (width <= 0 || height <= 0) return fail(RVK_ERR_INVALID_ARGUMENT, "bad image size");
  if ((long long)width * height > (long long)RVK_POST_SHADE_NOALLOC_MAX_PIXELS)
    return fail(RVK_ERR_INVALID_ARGUMENT, "%d x %d pixels exceed the allocation-free post pass's %d", width, height,
                RVK_POST_SHADE_NOALLOC_MAX_PIXELS);
  RVK_HIP(hipSetDevice(ctx->device));
  float* tmp = nullptr;
  int rc = scratch_on(ctx->device, &tmp);
  if (rc) return rc;
  dim3 grid((width + 15) / 16, (height + 15) / 16);
  // every pixel writes its normalised depth to tmp before the copy back
  hipLaunchKernelGGL(post_shade_kernel<false>, grid, dim3(256), 0, ctx->stream, width, height, *shade, depth,
                     min_depth, max_depth, (uint32_t*)rgba, tmp);
  RVK_HIP(hipGetLastError());
  RVK_HIP(hipMemcpyAsync(depth, tmp, sizeof(float) * (size_t)width * height, hipMemcpyDeviceToDevice, ctx->stream));
  return RVK_OK;
}

int rvk_post_shade_inex(rvk_ctx* ctx, int width, int height, const rvk_shade_params* shade, float* depth,
                        const float* min_depth, const float* max_depth, uint8_t* rgba) {
  return post_shade(ctx, true, width, height, shade, depth, min_depth, max_depth, rgba);
}

}  // extern "C"

extern "C" int rvk_count_ray_steps(rvk_ctx* ctx, const rvk_volume* vol, const rvk_view* view,
                                   const rvk_march_params* march, int64_t* out) {
  if (!ctx || !vol || !out) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  int rc = check_view(view);
  if (!rc) rc = check_march(march);
  if (rc) return rc;
  RVK_HIP(hipSetDevice(ctx->device));
  rc = ensure_scratch(ctx, 4096);
  if (rc) return rc;
  unsigned long long* d = (unsigned long long*)ctx->scratch;
  RVK_HIP(hipMemsetAsync(d, 0, sizeof(*d), ctx->stream));
  int tiles_x;
  const int nblk = launch_grid(view, tiles_x);
  hipLaunchKernelGGL(count_steps_kernel, dim3(nblk), dim3(256), 0, ctx->stream, vol_geom(vol), view_geom(view),
                     min_spacing(vol) * march->step_scale, march->jitter, tiles_x, d);
  RVK_HIP(hipGetLastError());
  unsigned long long h = 0;
  RVK_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  RVK_HIP(hipStreamSynchronize(ctx->stream));
  *out = (int64_t)h;
  return RVK_OK;
}
