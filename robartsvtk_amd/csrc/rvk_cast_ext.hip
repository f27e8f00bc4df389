// rvk_cast_ext.hip -- the sibling compositing casters: the 2D-TF caster with
// in/ex logic, and the dual-image caster.
//
// Both keep the layout of the 2D caster: one thread per pixel, 16x16-pixel
// workgroup tiles in super-tiles, formRays fused into the prologue, the TF as
// 32-byte cells, and the tests that decide whether a sample composites
// answered from bitmaps in LDS when the table has at most 256x256 entries.
//
// Reference kernels restated (V/ = Visualization/vtkCudaVisualization):
//   In/Ex  V/CUDA_vtkCuda2DInExLogicVolumeMapper_renderAlgo.cuh:24-179, 241-302
//   dual   V/CUDA_vtkCudaDualImageVolumeMapper_renderAlgo.cuh:33-380, 383-517
#include <cmath>
#include <type_traits>

#include "rvk_internal.h"
#include "rvk_march.h"

namespace rvk {

// ===========================================================================
// 2D TF with in/ex logic
struct InExArgs {
  VolGeom g;
  ViewGeom v;
  float step_len, term_T;
  int jitter;
  int n;
  int black_keyhole;
  rvk_tf2d_params p;
  const Tf2dCell* cells;
  const uint8_t* flags;  // per cell: bit 0 composites, bit 1 excludes
  const uint32_t* occ;   // the same as two bitmaps (occ_words each), or NULL
  int occ_words;
  const uint32_t* col;   // column bitmap of bit 0 (col_words), staged in LDS
  int col_words;
  const void* vol;
  uint32_t* out;
  float* depth;
  float* min_depth;
  float* max_depth;
  int tiles_x;
};

// V/...InExLogic...cuh:24-179.  Against the 2D caster: the step count is
// floor(numSteps) (the ray offset moves the start only), a sample composites
// when inEx > 0 or alpha > 0, the first composited sample outside the
// keyhole window [ee, es] with inEx > 0.5 ends the ray before it is added,
// the colour is saturated, and with the black keyhole a ray that composited
// anything is opaque.
template <typename CV, bool SKIP, bool PRECISE>
__global__ __launch_bounds__(256) void inex_kernel(InExArgs a) {
  extern __shared__ uint32_t s_lds[];
  const int cw = a.col_words;
  for (int i = threadIdx.x; i < cw; i += 256) s_lds[i] = a.col[i];
  uint32_t* s_occ = s_lds + cw;
  const bool use_occ = a.occ != nullptr;
  const int w = a.occ_words;
  if (use_occ)
    for (int i = threadIdx.x; i < 2 * w; i += 256) s_occ[i] = a.occ[i];
  __syncthreads();
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  int x, y;
  tile_pixel(tile, a.tiles_x, x, y);
  if (x >= a.v.W || y >= a.v.H) return;
  const Ray r = form_ray(a.g, a.v, a.step_len, x, y);
  const CV vol = cells_of<CV>(a.vol);
  const VolGeom& g = a.g;
  const int n = a.n;

  float Tr = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
  const float lo = a.p.intensity_low, mult = a.p.intensity_multiplier;
  const float glo = a.p.gradient_low, gmult = a.p.gradient_multiplier, goff = a.p.gradient_offset;
  const float term = a.term_T;
  const float j = ray_offset(a.jitter, x, y);
  V3<float> p{r.start.x + j * r.inc.x, r.start.y + j * r.inc.y, r.start.z + j * r.inc.z};
  const V3<float> p0c = cell_coord(p);
  int k = step_count(r.num_steps);
  const float depth0 = j + (float)k;
  const int es = k - (int)ceilf(r.exclude_start);
  const int ee = k - (int)floorf(r.exclude_end);
  const V3<float> inc = r.inc;
  const float rlen = ray_length(g, inc);
  const float sx = g.sp[0], sy = g.sp[1], sz = g.sp[2];
  const int kst = k;
  bool skip = false, back = false, special = false;
  while (k > 0) {
    const V3<float> q = sample_pos<PRECISE>(p, p0c, inc, kst, k);
    const float u = mult * (trilinear(vol, g, q.x, q.y, q.z) - lo);
    V3<float> gr{0.0f, 0.0f, 0.0f};
    float gmag = 0.0f;
    int ti = 0;
    uint32_t fl = 0u;
    // no cell of this column composites: the gradient is not needed
    if (tf2d_column_live(s_lds, tf2d_axis(n, u))) {
      gr.x = (trilinear(vol, g, q.x + 1.0f, q.y, q.z) - trilinear(vol, g, q.x - 1.0f, q.y, q.z)) * g.srecip[0];
      gr.y = (trilinear(vol, g, q.x, q.y + 1.0f, q.z) - trilinear(vol, g, q.x, q.y - 1.0f, q.z)) * g.srecip[1];
      gr.z = (trilinear(vol, g, q.x, q.y, q.z + 1.0f) - trilinear(vol, g, q.x, q.y, q.z - 1.0f)) * g.srecip[2];
      const float gm2 = gr.x * gr.x + gr.y * gr.y + gr.z * gr.z;
      const float gidx = gmult * (log2f(gm2 + goff) + glo);
      gmag = sqrtf(gm2);
      ti = tf2d_index(n, u, gidx);
      if (use_occ) {
        const uint32_t wi = (uint32_t)ti >> 5, b = (uint32_t)ti & 31u;
        fl = ((s_occ[wi] >> b) & 1u) | (((s_occ[w + wi] >> b) & 1u) << 1);
      } else {
        fl = a.flags[ti];
      }
    }
    // false: the ray ends here, nothing added
    auto accumulate = [&]() -> bool {
      if (!special && !(es >= k && ee <= k) && (fl & 2u)) return false;
      special = true;
      const float4 c0 = *(const float4*)&a.cells[ti].alpha;
      const float4 c1 = *(const float4*)&a.cells[ti].spec_power;
      const float alpha = c0.x;
      const float pl = satf(fabsf(gr.x * inc.x * sx + gr.y * inc.y * sy + gr.z * inc.z * sz) / (gmag * rlen));
      const float shd = c0.y + c0.z * pl;
      const float shs = c0.w * powf(pl, c1.x);
      const float m = Tr * alpha;
      Tr *= (1.0f - alpha);
      cr += m * satf(shd * c1.y + shs);
      cg += m * satf(shd * c1.z + shs);
      cb += m * satf(shd * c1.w + shs);
      return true;
    };
    if (SKIP) {
      if (fl & 1u) {
        back = skip;
        skip = false;
        if (back) { p.x = p.x - inc.x; p.y = p.y - inc.y; p.z = p.z - inc.z; k += 1; }
        else { p.x = p.x + inc.x; p.y = p.y + inc.y; p.z = p.z + inc.z; k -= 1; }
        if (!back && !accumulate()) break;
        if (Tr < term) { Tr = 0.0f; break; }
      } else {
        if (!back) { step(p, inc); k--; }
        skip = !back;
        step(p, inc);
        k--;
        back = false;
      }
    } else {
      step(p, inc);
      k--;
      if (fl & 1u) {
        if (!accumulate()) break;
        if (Tr < term) { Tr = 0.0f; break; }
      }
    }
  }
  const long idx = (long)y * a.v.W + x;
  const float A = (a.black_keyhole && special) ? 1.0f : 1.0f - Tr;
  const uint32_t R = to_u8(255.0f * satf(cr)), G = to_u8(255.0f * satf(cg)), B = to_u8(255.0f * satf(cb));
  a.out[idx] = R | (G << 8) | (B << 16) | (to_u8(255.0f * satf(A)) << 24);
  if (a.depth) a.depth[idx] = depth0 - (float)k;
  if (a.min_depth) a.min_depth[idx] = r.min_depth;
  if (a.max_depth) a.max_depth[idx] = r.max_depth;
}

// ===========================================================================
// Dual-image caster
struct DualArgs {
  VolGeom g;
  ViewGeom v;
  float step_len, term_T;
  int jitter;
  int n;
  rvk_dual_params p;
  const Tf2dCell* cells;
  const Tf2dCell* kcells;
  const uint32_t* occ;  // opacity bitmaps staged in LDS, or NULL
  const uint32_t* kocc;
  int occ_words;
  const uint32_t* col;  // column bitmaps (TF, then keyhole TF), staged in LDS
  int col_words;
  const void* vol1;
  const void* vol2;
  uint32_t* out;
  float* depth;
  float* min_depth;
  float* max_depth;
  int tiles_x;
};

// Ray samples of the dual caster: floor(numSteps) less the ray offset,
// truncated (V/...DualImage...cuh:66-67 store an int minus a float back in an
// int); rays the 2D caster would not march march nothing here either.
__device__ __forceinline__ int dual_step_count(float num_steps, float j) {
  if (!(num_steps > 0.0f && num_steps < 16777216.0f)) return 0;
  return (int)((float)(int)floorf(num_steps) - j);
}

// V/...DualImage...cuh:33-380.  The table is indexed by the two components
// (component 1 the column, component 2 the row); shading uses the gradient of
// component 1.  KEYTF selects the keyhole-TF variant; without it the march
// jumps over the keyhole window.  Both round the window bounds up/down.
template <typename CV, typename CV2, bool SKIP, bool KEYTF, bool PRECISE>
__global__ __launch_bounds__(256) void dual_kernel(DualArgs a) {
  extern __shared__ uint32_t s_lds[];
  const int cw = a.col_words;
  for (int i = threadIdx.x; i < (KEYTF ? 2 : 1) * cw; i += 256) s_lds[i] = a.col[i];
  uint32_t* s_occ = s_lds + (KEYTF ? 2 : 1) * cw;
  const bool use_occ = a.occ != nullptr;
  if (use_occ) {
    const int w = a.occ_words;
    for (int i = threadIdx.x; i < w; i += 256) s_occ[i] = a.occ[i];
    if (KEYTF)
      for (int i = threadIdx.x; i < w; i += 256) s_occ[w + i] = a.kocc[i];
  }
  __syncthreads();
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  int x, y;
  tile_pixel(tile, a.tiles_x, x, y);
  if (x >= a.v.W || y >= a.v.H) return;
  const Ray r = form_ray(a.g, a.v, a.step_len, x, y);
  const CV vol1 = cells_of<CV>(a.vol1);
  const CV2 vol2 = cells_of<CV2>(a.vol2);
  const VolGeom& g = a.g;
  const int n = a.n;

  float Tr = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
  const float lo1 = a.p.intensity1_low, m1 = a.p.intensity1_multiplier;
  const float lo2 = a.p.intensity2_low, m2 = a.p.intensity2_multiplier;
  const float term = a.term_T;
  const float j = ray_offset(a.jitter, x, y);
  V3<float> p{r.start.x + j * r.inc.x, r.start.y + j * r.inc.y, r.start.z + j * r.inc.z};
  const V3<float> p0c = cell_coord(p);
  int k = dual_step_count(r.num_steps, j);
  const int k0 = k;
  const int es = k - (int)ceilf(r.exclude_start);
  const int ee = k - (int)floorf(r.exclude_end);
  const V3<float> inc = r.inc;
  const float rlen = ray_length(g, inc);
  const float sx = g.sp[0], sy = g.sp[1], sz = g.sp[2];
  const int kst = k;
  int sa = 0, sb = 0;
  while (k > 0) {
    const bool in_key = (es >= k && ee < k);
    if (!KEYTF && in_key) {
      p.x += (float)(k - ee) * inc.x; p.y += (float)(k - ee) * inc.y; p.z += (float)(k - ee) * inc.z;
      k = ee;
      sa = 0;
      sb = 0;
      continue;
    }
    const V3<float> q = sample_pos<PRECISE>(p, p0c, inc, kst, k);
    const float u1 = m1 * (trilinear(vol1, g, q.x, q.y, q.z) - lo1);
    const bool kt = KEYTF && in_key;
    const Tf2dCell* __restrict__ cells = kt ? a.kcells : a.cells;
    V3<float> gr{0.0f, 0.0f, 0.0f};
    float gmag = 0.0f;
    int ti = 0;
    float4 c0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // component 1 picks the column: an empty one needs neither component 2
    // nor the gradient
    if (tf2d_column_live(s_lds + (kt ? cw : 0), tf2d_axis(n, u1))) {
      const float u2 = m2 * (trilinear(vol2, g, q.x, q.y, q.z) - lo2);
      gr.x = (trilinear(vol1, g, q.x + 1.0f, q.y, q.z) - trilinear(vol1, g, q.x - 1.0f, q.y, q.z)) * g.srecip[0];
      gr.y = (trilinear(vol1, g, q.x, q.y + 1.0f, q.z) - trilinear(vol1, g, q.x, q.y - 1.0f, q.z)) * g.srecip[1];
      gr.z = (trilinear(vol1, g, q.x, q.y, q.z + 1.0f) - trilinear(vol1, g, q.x, q.y, q.z - 1.0f)) * g.srecip[2];
      gmag = sqrtf(gr.x * gr.x + gr.y * gr.y + gr.z * gr.z);
      ti = tf2d_index(n, u1, u2);
      if (use_occ) {
        const uint32_t* occ = s_occ + (kt ? a.occ_words : 0);
        if ((occ[ti >> 5] >> (ti & 31)) & 1u) c0 = *(const float4*)&cells[ti].alpha;
      } else {
        c0 = *(const float4*)&cells[ti].alpha;
      }
    }
    const float alpha = c0.x;
    auto accumulate = [&]() {
      const float4 c1 = *(const float4*)&cells[ti].spec_power;
      const float pl = satf(fabsf(gr.x * inc.x * sx + gr.y * inc.y * sy + gr.z * inc.z * sz) / (gmag * rlen));
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

}  // namespace rvk

using namespace rvk;

extern "C" {

int rvk_render_2d_inex(rvk_ctx* ctx, const rvk_volume* vol, int frame, const rvk_tf2d* tf, const rvk_view* view,
                       const rvk_march_params* march, int black_keyhole, uint8_t* out, const rvk_aux_out* aux) {
  if (!ctx || !tf || !out) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  int rc = check_frame(vol, frame);
  if (!rc) rc = check_view(view);
  if (!rc) rc = check_march(march);
  if (rc) return rc;
  if (!tf->inex_flags) return fail(RVK_ERR_NOT_READY, "the 2D TF has no in/ex table (rvk_tf2d_set_inex)");
  RVK_HIP(hipSetDevice(ctx->device));
  InExArgs a;
  a.g = vol_geom(vol);
  a.v = view_geom(view);
  a.step_len = min_spacing(vol) * march->step_scale;
  a.term_T = march->term_T;
  a.jitter = march->jitter;
  a.n = tf->n;
  a.black_keyhole = black_keyhole != 0;
  a.p = tf->p;
  a.cells = tf->cells;
  a.flags = tf->inex_flags;
  a.occ = tf->inex_occ;
  a.occ_words = tf->occ_words;
  a.col = tf->inex_col;
  a.col_words = tf->col_words;
  const size_t lds = sizeof(uint32_t) * ((size_t)tf->col_words + (a.occ ? 2 * (size_t)a.occ_words : 0));
  a.vol = vol->frames[frame];
  a.out = (uint32_t*)out;
  a.depth = aux ? aux->depth : nullptr;
  a.min_depth = aux ? aux->min_depth : nullptr;
  a.max_depth = aux ? aux->max_depth : nullptr;
  const int nblk = launch_grid(view, a.tiles_x);
  const bool skip = march->skip_rule != 0, pr = march->precise != 0;
  return dispatch_frame(vol, frame, [&](auto tag) {
    using CV = decltype(tag);
#define RVK_LIE(S, P) hipLaunchKernelGGL((inex_kernel<CV, S, P>), dim3(nblk), dim3(256), lds, ctx->stream, a)
    if (skip) { if (pr) RVK_LIE(true, true); else RVK_LIE(true, false); }
    else { if (pr) RVK_LIE(false, true); else RVK_LIE(false, false); }
#undef RVK_LIE
    RVK_HIP(hipGetLastError());
    return (int)RVK_OK;
  });
}

int rvk_render_dual(rvk_ctx* ctx, const rvk_volume* vol1, const rvk_volume* vol2, int frame, const rvk_tf2d* tf,
                    const rvk_dual_params* params, const rvk_view* view, const rvk_march_params* march, uint8_t* out,
                    const rvk_aux_out* aux) {
  if (!ctx || !tf || !params || !out) return fail(RVK_ERR_INVALID_ARGUMENT, "NULL argument");
  int rc = check_frame(vol1, frame);
  if (!rc) rc = check_frame(vol2, frame);
  if (!rc) rc = check_view(view);
  if (!rc) rc = check_march(march);
  if (rc) return rc;
  if (vol1->dtype != vol2->dtype || vol1->layout != vol2->layout || vol1->dims[0] != vol2->dims[0] ||
      vol1->dims[1] != vol2->dims[1] || vol1->dims[2] != vol2->dims[2])
    return fail(RVK_ERR_INVALID_ARGUMENT, "the two component volumes differ in size, type or cell layout");
  RVK_HIP(hipSetDevice(ctx->device));
  DualArgs a;
  a.g = vol_geom(vol1);
  a.v = view_geom(view);
  a.step_len = min_spacing(vol1) * march->step_scale;
  a.term_T = march->term_T;
  a.jitter = march->jitter;
  a.n = tf->n;
  a.p = *params;
  a.cells = tf->cells;
  a.kcells = tf->kcells;
  a.occ = tf->occ;
  a.kocc = tf->kocc;
  a.occ_words = tf->occ_words;
  a.col = tf->col;
  a.col_words = tf->col_words;
  if (tf->kcells && !tf->kocc) a.occ = nullptr;
  const int ntf = tf->kcells ? 2 : 1;
  const size_t lds = sizeof(uint32_t) * ((size_t)tf->col_words * ntf + (a.occ ? (size_t)a.occ_words * ntf : 0));
  a.vol1 = vol1->frames[frame];
  a.vol2 = vol2->frames[frame];
  a.out = (uint32_t*)out;
  a.depth = aux ? aux->depth : nullptr;
  a.min_depth = aux ? aux->min_depth : nullptr;
  a.max_depth = aux ? aux->max_depth : nullptr;
  const int nblk = launch_grid(view, a.tiles_x);
  const bool skip = march->skip_rule != 0, key = tf->kcells != nullptr, pr = march->precise != 0;
  // the two components share voxel type and cell order, but each frame has
  // its own cell encoding; binary16-base frames are re-encoded as plain
  // differences first (same values), which keeps the kernel pairs to those
  // of one voxel type
  rc = frame_drop_half(vol1, frame, ctx->stream);
  if (!rc) rc = frame_drop_half(vol2, frame, ctx->stream);
  if (rc) return rc;
  return dispatch_frame<false>(vol1, frame, [&](auto tag1) {
    using CV = decltype(tag1);
    return dispatch_frame<false>(vol2, frame, [&](auto tag2) {
      using CV2 = decltype(tag2);
      if constexpr (std::is_same<typename CV::value_type, typename CV2::value_type>::value &&
                    CV::kLog2 == CV2::kLog2) {
#define RVK_LDU(S, K)                                                                                               \
  do {                                                                                                              \
    if (pr) hipLaunchKernelGGL((dual_kernel<CV, CV2, S, K, true>), dim3(nblk), dim3(256), lds, ctx->stream, a);    \
    else hipLaunchKernelGGL((dual_kernel<CV, CV2, S, K, false>), dim3(nblk), dim3(256), lds, ctx->stream, a);      \
  } while (0)
        if (skip) { if (key) RVK_LDU(true, true); else RVK_LDU(true, false); }
        else { if (key) RVK_LDU(false, true); else RVK_LDU(false, false); }
#undef RVK_LDU
        RVK_HIP(hipGetLastError());
        return (int)RVK_OK;
      } else {
        return fail(RVK_ERR_INVALID_ARGUMENT, "the two component volumes differ in type or cell layout");
      }
    });
  });
}

}  // extern "C"
