/*
 * rvk_oracle.cpp -- CPU restatement of RobartsVTK's CUDA ray casters.
 *
 * TEST INFRASTRUCTURE ONLY (what the HIP casters are tested against, and the
 * reported CPU baseline).  Nothing in the product library links or loads this file.
 *
 * The reference kernels cannot run here (they need nvcc, legacy CUDA texture
 * references and VTK 7), so this is a transcription written from the source
 * text.  Texture hardware is restated from its documented behaviour:
 *   - 3D volume: unnormalised coordinates, linear filter, clamp.  A sample at
 *     p reads the trilinear value at p - 0.5 (texel centres at i + 0.5), with
 *     exact (not 9-bit fixed-point) weights; each linear interpolation is
 *     a + t (b - a) with one rounding (fma), as the GPU casters evaluate it.
 *   - 1D TF: normalised coordinates, linear filter, clamp: x = u*N - 0.5.
 *   - 2D TF and z-buffer: normalised coordinates, point filter, clamp.
 * Directed-rounding square roots (__fsqrt_ru/__fsqrt_rz) are emulated exactly
 * in FP32 from a correctly rounded sqrtf and the sign of the FMA residual.
 *
 * Parity pins (tests/test_oracle_*.py): jitter table against the golden
 * fixture copied from V/vtkCudaVolumeMapper.cxx:59-93; ray setup against
 * closed-form camera geometry; trilinear against exact interpolation of
 * linear fields; DRR against analytic chord lengths through boxes; 1D/2D
 * compositing against (1-a)^n on uniform volumes, first-hit colour, zero-
 * opacity TFs and a naive independent restatement on tiny volumes; FP32 vs
 * FP64 agreement.  NVIDIA's fixed-point texture weights, VTK's GetTable edge
 * behaviour and VTK's camera matrices have no pin: parity unpinned there.
 */
#include "rvk_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// V/vtkCudaVolumeMapper.cxx:59-93 (the only data table the reference holds).
// The table is data that parity depends on bit for bit, so it is reproduced.
const float kJitter[256] = {
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

template <typename T> struct V3 { T x, y, z; };

template <typename T> inline T sat(T v) { return std::fmin(std::fmax(v, T(0)), T(1)); }

// __fsqrt_ru / __fsqrt_rz (V/...renderAlgo.cu:372,383-386,414).
inline float sqrt_ru(float x) {
  float r = std::sqrt(x);
  if (std::fma(r, r, -x) < 0.0f) r = std::nextafter(r, INFINITY);
  return r;
}
inline float sqrt_rz(float x) {
  float r = std::sqrt(x);
  if (std::fma(r, r, -x) > 0.0f) r = std::nextafter(r, 0.0f);
  return r;
}
inline double sqrt_ru(double x) { return std::sqrt(x); }
inline double sqrt_rz(double x) { return std::sqrt(x); }

// float -> unsigned char as the CUDA conversion does it: truncate, saturate, NaN -> 0.
template <typename T> inline uint8_t to_u8(T v) {
  if (!(v > T(0))) return 0;
  if (v >= T(255)) return 255;
  return (uint8_t)(int)v;
}

// V/vtkCudaVolumeInformationHandler.cxx:80-108
template <typename T> struct VolInfo {
  int n[3];
  T bounds[6];
  T srecip[3];
  T sp[3];
  T minsp;
};

template <typename T> VolInfo<T> vol_info(const orc_volume* v) {
  VolInfo<T> vi;
  for (int a = 0; a < 3; ++a) {
    vi.n[a] = v->dims[a];
    vi.srecip[a] = (T)(0.5 / v->spacing[a]);
    vi.sp[a] = (T)v->spacing[a];
    vi.bounds[2 * a] = T(0);
    vi.bounds[2 * a + 1] = (T)v->dims[a] - T(1);
  }
  double m = std::min(v->spacing[0], std::min(v->spacing[1], v->spacing[2]));
  vi.minsp = (T)m;
  return vi;
}

// ---------------------------------------------------------------------------
// Volume sampling: CUDA 3D texture, unnormalised, linear, clamp.
template <typename VT> struct Grid {
  const VT* d;
  int nx, ny, nz;
};

// a + t (b - a) with one rounding of the product-sum
template <typename T> inline T lerp(T a, T b, T t) { return std::fma(t, b - a, a); }

// Trilinear value at the cell coordinate c = p + 0.5 (the marches carry c;
// sampling at p means passing p + 0.5).  Texel centres sit at i + 0.5, so the
// corners along an axis are voxels floor(c) - 1 and floor(c), clamped.
template <typename T, typename VT> inline T trilinear(const Grid<VT>& g, T cx, T cy, T cz) {
  // clamping the coordinate to [0, n + 1] leaves clamp-addressed results unchanged
  T x = std::fmin(std::fmax(cx, T(0)), (T)(g.nx + 1));
  T y = std::fmin(std::fmax(cy, T(0)), (T)(g.ny + 1));
  T z = std::fmin(std::fmax(cz, T(0)), (T)(g.nz + 1));
  T fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
  T ax = x - fx, ay = y - fy, az = z - fz;
  int ix = (int)fx - 1, iy = (int)fy - 1, iz = (int)fz - 1;
  int x0 = std::min(std::max(ix, 0), g.nx - 1), x1 = std::min(std::max(ix + 1, 0), g.nx - 1);
  int y0 = std::min(std::max(iy, 0), g.ny - 1), y1 = std::min(std::max(iy + 1, 0), g.ny - 1);
  int z0 = std::min(std::max(iz, 0), g.nz - 1), z1 = std::min(std::max(iz + 1, 0), g.nz - 1);
  const long sy = g.nx, sz = (long)g.nx * g.ny;
  auto at = [&](int i, int j, int k) { return (T)g.d[i + j * sy + k * sz]; };
  T v000 = at(x0, y0, z0), v100 = at(x1, y0, z0), v010 = at(x0, y1, z0), v110 = at(x1, y1, z0);
  T v001 = at(x0, y0, z1), v101 = at(x1, y0, z1), v011 = at(x0, y1, z1), v111 = at(x1, y1, z1);
  T c00 = lerp(v000, v100, ax), c10 = lerp(v010, v110, ax);
  T c01 = lerp(v001, v101, ax), c11 = lerp(v011, v111, ax);
  T c0 = lerp(c00, c10, ay), c1 = lerp(c01, c11, ay);
  return lerp(c0, c1, az);
}

// The precise DRR march samples at c = (p0 + 0.5) + t inc, t in [0, k - 1]
// (one fma per axis, monotone in t).  The GPU DRR tests c at t = 0 and t = k
// against the cell range [0, n + 1] per axis and then drops the per-sample
// clamp; a ray that fails takes no samples.  Rays that pass the volume clip
// sample within [0.5, n - 0.5] and always pass (tests/test_oracle.py counts
// the rays that fail on every scene).
static long long g_unsafe_rays = 0;
template <typename T, typename VT>
bool samples_in_cells(const Grid<VT>& g, const V3<T>& p, const V3<T>& inc, int k) {
  const T t = (T)k;
  const T s[3] = {p.x + T(0.5), p.y + T(0.5), p.z + T(0.5)};
  const T d[3] = {inc.x, inc.y, inc.z};
  const int n[3] = {g.nx, g.ny, g.nz};
  bool ok = true;
  for (int a = 0; a < 3; ++a) {
    const T e = std::fma(t, d[a], s[a]);
    ok = ok && s[a] >= T(0) && s[a] <= (T)(n[a] + 1) && e >= T(0) && e <= (T)(n[a] + 1);
  }
  if (!ok && k > 0) __atomic_add_fetch(&g_unsafe_rays, 1, __ATOMIC_RELAXED);
  return ok;
}

// 1D TF: tex1D, normalised, linear, clamp (V/...1D...cuh:248-267).
template <typename T> inline T tf1d(const float* t, int n, T u) {
  T x = u * (T)n - T(0.5);
  x = std::fmin(std::fmax(x, T(-1)), (T)n);
  T fx = std::floor(x);
  T a = x - fx;
  int i = (int)fx;
  int i0 = std::min(std::max(i, 0), n - 1), i1 = std::min(std::max(i + 1, 0), n - 1);
  return lerp((T)t[i0], (T)t[i1], a);
}

// 2D TF: tex2D, normalised, point, clamp (V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:21-25).
template <typename T> inline int tf2d_index(int n, T u, T v) {
  T x = std::fmin(std::fmax(u * (T)n, T(-1)), (T)n);
  T y = std::fmin(std::fmax(v * (T)n, T(-1)), (T)n);
  int i = std::min(std::max((int)std::floor(x), 0), n - 1);
  int j = std::min(std::max((int)std::floor(y), 0), n - 1);
  return j * n + i;
}

// ---------------------------------------------------------------------------
// Ray setup: SetRayEnds + clipping + formRays (V/...renderAlgo.cu:130-447).
template <typename T> struct Ray {
  V3<T> start, inc;
  T num_steps;
  T exclude_start, exclude_end;
  T min_depth, max_depth;
};

// V/...renderAlgo.cu:130-197
template <typename T> void clip_planes(const orc_view* v, V3<T>& s, V3<T>& e, V3<T>& d) {
  const int np = v->n_clip;
  if (!np) return;
  int flag = 0;
  for (int i = 0; i < np; ++i) {
    d.x = e.x - s.x; d.y = e.y - s.y; d.z = e.z - s.z;
    const T px = v->clip[4 * i], py = v->clip[4 * i + 1], pz = v->clip[4 * i + 2], pw = v->clip[4 * i + 3];
    const T dp = px * d.x + py * d.y + pz * d.z;
    const T t = -(px * s.x + py * s.y + pz * s.z + pw) / dp;
    const T p0 = s.x + t * d.x, p1 = s.y + t * d.y, p2 = s.z + t * d.z;
    if (t > T(0) && t < T(1)) {
      if (dp > T(0)) { s.x = p0; s.y = p1; s.z = p2; }
      else { e.x = p0; e.y = p1; e.z = p2; }
    }
    flag |= (dp > T(0) && t > T(1));
    flag |= (dp < T(0) && t < T(0));
  }
  if (flag) s = e;
}

// V/...renderAlgo.cu:199-318
template <typename T> void clip_volume(const VolInfo<T>& vi, V3<T>& s, V3<T>& e, V3<T>& d) {
  d.x = e.x - s.x; d.y = e.y - s.y; d.z = e.z - s.z;
  const T b0 = vi.bounds[0] + T(1), b1 = vi.bounds[1] - T(1);
  const T b2 = vi.bounds[2] + T(1), b3 = vi.bounds[3] - T(1);
  const T b4 = vi.bounds[4] + T(1), b5 = vi.bounds[5] - T(1);
  T ds, de;
  auto crop = [&](T dS, T dE) {
    if (std::isfinite(dS)) {
      s.x += d.x * dS; s.y += d.y * dS; s.z += d.z * dS;
      e.x += d.x * dE; e.y += d.y * dE; e.z += d.z * dE;
    }
  };
  if (d.x > T(0)) { ds = s.x < b0 ? b0 - s.x : T(0); de = e.x > b1 ? b1 - e.x : T(0); }
  else { ds = s.x > b1 ? b1 - s.x : T(0); de = e.x < b0 ? b0 - e.x : T(0); }
  ds /= d.x; de /= d.x;
  crop(ds, de);
  if (d.y > T(0)) { ds = s.y < b2 ? b2 - s.y : T(0); de = e.y > b3 ? b3 - e.y : T(0); }
  else { ds = s.y > b3 ? b3 - s.y : T(0); de = e.y < b2 ? b2 - e.y : T(0); }
  ds /= d.y; de /= d.y;
  crop(ds, de);
  if (d.z > T(0)) { ds = s.z < b4 ? b4 - s.z : T(0); de = e.z > b5 ? b5 - e.z : T(0); }
  else { ds = s.z > b5 ? b5 - s.z : T(0); de = e.z < b4 ? b4 - e.z : T(0); }
  ds /= d.z; de /= d.z;
  crop(ds, de);
  if (e.x > b1 + T(1) || e.y > b3 + T(1) || e.z > b5 + T(1) || e.x < b0 - T(1) || e.y < b2 - T(1) ||
      e.z < b4 - T(1) || s.x > b1 + T(1) || s.y > b3 + T(1) || s.z > b5 + T(1) || s.x < b0 - T(1) ||
      s.y < b2 - T(1) || s.z < b4 - T(1)) {
    s = e;
  }
  d.x = e.x - s.x; d.y = e.y - s.y; d.z = e.z - s.z;
}

// V/...renderAlgo.cu:38-128
template <typename T>
void keyhole_values(const orc_view* v, V3<T> s, const V3<T>& inc, T num_steps, T& es, T& ee) {
  const int np = v->n_keyhole;
  const V3<T> old = s;
  V3<T> d{num_steps * inc.x, num_steps * inc.y, num_steps * inc.z};
  V3<T> e{s.x + d.x, s.y + d.y, s.z + d.z};
  es = T(1);
  ee = T(-1);
  if (!np) return;
  int flag = 0;
  for (int i = 0; i < np; ++i) {
    d.x = e.x - s.x; d.y = e.y - s.y; d.z = e.z - s.z;
    const T px = v->keyhole[4 * i], py = v->keyhole[4 * i + 1], pz = v->keyhole[4 * i + 2],
            pw = v->keyhole[4 * i + 3];
    const T dp = px * d.x + py * d.y + pz * d.z;
    const T t = -(px * s.x + py * s.y + pz * s.z + pw) / dp;
    const T p0 = s.x + t * d.x, p1 = s.y + t * d.y, p2 = s.z + t * d.z;
    if (t > T(0) && t < T(1)) {
      if (dp > T(0)) { s.x = p0; s.y = p1; s.z = p2; }
      else { e.x = p0; e.y = p1; e.z = p2; }
    }
    flag |= (dp > T(0) && t > T(1));
    flag |= (dp < T(0) && t < T(0));
  }
  s.x -= old.x; s.y -= old.y; s.z -= old.z;
  e.x -= old.x; e.y -= old.y; e.z -= old.z;
  const T inv = T(1) / (inc.x * inc.x + inc.y * inc.y + inc.z * inc.z);
  es = flag ? T(-1) : (s.x * inc.x + s.y * inc.y + s.z * inc.z) * inv;
  ee = flag ? T(-1) : (e.x * inc.x + e.y * inc.y + e.z * inc.z) * inv;
}

// V/...renderAlgo.cu:320-447
template <typename T>
Ray<T> form_ray(const VolInfo<T>& vi, const orc_view* v, T step_len, int x, int y) {
  Ray<T> r;
  T m[16];
  for (int i = 0; i < 16; ++i) m[i] = (v->m64 && sizeof(T) == sizeof(double)) ? (T)v->m64[i] : (T)v->m[i];
  const T vx = v->flipped ? (T)x / (T)v->width : T(1) - (T)x / (T)v->width;
  const T vy = (T)y / (T)v->height;
  T end_depth = T(1);
  if (v->zbuffer) {
    const T u = T(1) - vx;
    int zi = (int)std::floor(std::fmin(std::fmax(u * (T)v->zb_width, T(-1)), (T)v->zb_width));
    int zj = (int)std::floor(std::fmin(std::fmax(vy * (T)v->zb_height, T(-1)), (T)v->zb_height));
    zi = std::min(std::max(zi, 0), v->zb_width - 1);
    zj = std::min(std::max(zj, 0), v->zb_height - 1);
    end_depth = (T)v->zbuffer[(long)zj * v->zb_width + zi];
  }
  V3<T> s, e, f;
  s.x = vx * m[0] + vy * m[1] + m[3];
  s.y = vx * m[4] + vy * m[5] + m[7];
  s.z = vx * m[8] + vy * m[9] + m[11];
  const T sn = vx * m[12] + vy * m[13] + m[15];
  e.x = s.x + end_depth * m[2];
  e.y = s.y + end_depth * m[6];
  e.z = s.z + end_depth * m[10];
  const T en = sn + end_depth * m[14];
  f.x = s.x + m[2];
  f.y = s.y + m[6];
  f.z = s.z + m[10];
  const T fn = sn + m[14];
  s.x /= sn; s.y /= sn; s.z /= sn;
  e.x /= en; e.y /= en; e.z /= en;
  f.x /= fn; f.y /= fn; f.z /= fn;
  const V3<T> old = s;
  V3<T> d{f.x - s.x, f.y - s.y, f.z - s.z};
  r.max_depth = sqrt_rz(d.x * d.x + d.y * d.y + d.z * d.z);
  clip_planes(v, s, e, d);
  clip_volume(vi, s, e, d);
  const T len = sqrt_rz(d.x * d.x + d.y * d.y + d.z * d.z);
  const T md = sqrt_rz((s.x - old.x) * (s.x - old.x) + (s.y - old.y) * (s.y - old.y) +
                       (s.z - old.z) * (s.z - old.z));
  r.min_depth = (len > T(0)) ? md : r.max_depth;
  r.num_steps = sqrt_ru(d.x * d.x * vi.sp[0] * vi.sp[0] + d.y * d.y * vi.sp[1] * vi.sp[1] +
                        d.z * d.z * vi.sp[2] * vi.sp[2]) / step_len;
  r.inc.x = d.x / r.num_steps;
  r.inc.y = d.y / r.num_steps;
  r.inc.z = d.z / r.num_steps;
  r.start = s;
  keyhole_values(v, s, r.inc, r.num_steps, r.exclude_start, r.exclude_end);
  return r;
}

template <typename T> inline T step_length(const VolInfo<T>& vi, const orc_march* mp) {
  return vi.minsp * (T)mp->step_scale;
}

template <typename T> inline T jitter_of(const orc_march* mp, int x, int y) {
  return mp->jitter ? (T)kJitter[(x % 16) + 16 * (y % 16)] : T(0);
}

inline int fl2i_rd(double v) { return (int)std::floor(v); }
// floor(num_steps - offset); non-finite or >= 2^24-step rays march nothing
inline int step_count(double v) { return (v > 0.0 && v < 16777216.0) ? (int)std::floor(v) : 0; }
inline int fl2i_ru(double v) { return (int)std::ceil(v); }

// Branch control of one ray.  `rec` records the decisions an FP32 march takes
// (contributes? / terminates?) and its integer set-up; `play` makes an FP64
// march replay them, so FP64 measures the arithmetic error of the very path
// the FP32 kernel follows (a decision such as mu > 0 in exact-zero air can
// otherwise flip between precisions and change which samples are visited).
struct Ctl {
  std::vector<uint8_t>* rec = nullptr;
  const std::vector<uint8_t>* play = nullptr;
  size_t pos = 0;
  int k0 = 0, es = 0, ee = 0;
  long long* count = nullptr;
  void ints(int& k, int& es_, int& ee_) {
    if (rec) { k0 = k; es = es_; ee = ee_; }
    if (play) { k = k0; es_ = es; ee_ = ee; }
  }
  bool decide(bool d) {
    if (rec) rec->push_back((uint8_t)d);
    if (play) return pos < play->size() ? (*play)[pos++] != 0 : d;
    return d;
  }
};

// ---------------------------------------------------------------------------
// Threading: rows are handed out dynamically to a pool of std::threads.
template <typename F> void parallel_rows(int H, int nthreads, F&& fn) {
  if (nthreads <= 1) {
    for (int y = 0; y < H; ++y) fn(y);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < nthreads; ++t)
    pool.emplace_back([&]() {
      for (int y = next.fetch_add(1); y < H; y = next.fetch_add(1)) fn(y);
    });
  for (auto& th : pool) th.join();
}

template <typename F> int with_grid(const orc_volume* v, F&& fn) {
  switch (v->dtype) {
    case ORC_U8: return fn(Grid<uint8_t>{(const uint8_t*)v->data, v->dims[0], v->dims[1], v->dims[2]});
    case ORC_I8: return fn(Grid<int8_t>{(const int8_t*)v->data, v->dims[0], v->dims[1], v->dims[2]});
    case ORC_U16: return fn(Grid<uint16_t>{(const uint16_t*)v->data, v->dims[0], v->dims[1], v->dims[2]});
    case ORC_I16: return fn(Grid<int16_t>{(const int16_t*)v->data, v->dims[0], v->dims[1], v->dims[2]});
    case ORC_U32: return fn(Grid<uint32_t>{(const uint32_t*)v->data, v->dims[0], v->dims[1], v->dims[2]});
    case ORC_I32: return fn(Grid<int32_t>{(const int32_t*)v->data, v->dims[0], v->dims[1], v->dims[2]});
    case ORC_F32: return fn(Grid<float>{(const float*)v->data, v->dims[0], v->dims[1], v->dims[2]});
    default: return 1;
  }
}

// ---------------------------------------------------------------------------
// DRR caster: V/CUDA_vtkCudaDRRImageVolumeMapper_renderAlgo.cuh:11-171
template <typename T, typename VT>
T drr_ray(const Grid<VT>& g, const VolInfo<T>& vi, const orc_march* mp, const orc_drr* dp, const Ray<T>& r,
          int x, int y, Ctl* ctl) {
  Ctl none;
  if (!ctl) ctl = &none;
  const bool ref_mode = dp->output == ORC_DRR_RGBA8_REF;
  const T lg = std::log(T(256) * (T)dp->offset);
  T acc = ref_mode ? lg : T(0);
  const T stop_at = ref_mode ? T(0) : -lg;
  const T j = jitter_of<T>(mp, x, y);
  int k = step_count(r.num_steps - j);
  V3<T> p{r.start.x + j * r.inc.x, r.start.y + j * r.inc.y, r.start.z + j * r.inc.z};
  if (mp->precise && !samples_in_cells(g, p, r.inc, k)) k = 0;
  const V3<T>& inc = r.inc;
  const T ds = std::sqrt(inc.x * inc.x * vi.sp[0] * vi.sp[0] + inc.y * inc.y * vi.sp[1] * vi.sp[1] +
                         inc.z * inc.z * vi.sp[2] * vi.sp[2]);
  const T ci = (T)dp->intercept * -ds;
  const T cs = (T)dp->slope * -ds;
  int es = k - fl2i_ru(r.exclude_start);
  int ee = k - fl2i_rd(r.exclude_end);
  ctl->ints(k, es, ee);
  const int kst = k;
  const V3<T> p0c{p.x + T(0.5), p.y + T(0.5), p.z + T(0.5)};
  // Sample coordinates are carried as cell coordinates (c = p + 0.5).
  // Precise mode: c follows from the step index, (p0 + 0.5) + t inc with one
  // rounding per axis (no drift from repeated p += inc); reference mode: the
  // incrementally updated p, plus 0.5.
  auto cur = [&]() -> V3<T> {
    if (!mp->precise) return V3<T>{p.x + T(0.5), p.y + T(0.5), p.z + T(0.5)};
    const T t = (T)(kst - k);
    return V3<T>{std::fma(t, inc.x, p0c.x), std::fma(t, inc.y, p0c.y), std::fma(t, inc.z, p0c.z)};
  };
  // Precise mode sums the line integral in double, the reference mode in T.
  double accd = (double)acc;
  auto add = [&](T v) {
    if (mp->precise) accd += (double)v;
    else acc += v;
  };
  auto stop = [&]() { return mp->precise ? accd <= (double)stop_at : acc <= stop_at; };
  int a = 0, b = 0;
  while (k > 0) {
    const bool in_key = (es >= k && ee < k);
    if (ctl->count) ++*ctl->count;
    const V3<T> q = cur();
    const T alpha = std::fma(cs, trilinear<T>(g, q.x, q.y, q.z), ci);
    if (mp->skip_rule) {
      if (ctl->decide(alpha < T(0) && !in_key)) {
        a = b;
        b = 0;
        if (a) { p.x = p.x - inc.x; p.y = p.y - inc.y; p.z = p.z - inc.z; k += 1; }
        else { p.x = p.x + inc.x; p.y = p.y + inc.y; p.z = p.z + inc.z; k -= 1; }
        if (!a) add(alpha);
        if (ctl->decide(dp->ref_termination && stop())) break;
      } else {
        if (!a) { p.x += inc.x; p.y += inc.y; p.z += inc.z; k--; }
        b = !a;
        p.x += inc.x; p.y += inc.y; p.z += inc.z; k--;
        a = 0;
      }
    } else {
      p.x += inc.x; p.y += inc.y; p.z += inc.z; k--;
      if (ctl->decide(alpha < T(0) && !in_key)) {
        add(alpha);
        if (ctl->decide(dp->ref_termination && stop())) break;
      }
    }
  }
  if (mp->precise) acc = (T)accd;
  if (ref_mode) {
    T out = T(256) - std::exp(acc);
    return out < T(0) ? T(0) : out;
  }
  if (dp->output == ORC_DRR_F32_EXP) return std::exp(acc);
  return -acc;
}

template <typename T> int drr_impl(const orc_volume* vol, const orc_view* view, const orc_march* mp,
                                   const orc_drr* dp, void* out, int nthreads) {
  if (!vol || !view || !mp || !dp || !out) return 1;
  const VolInfo<T> vi = vol_info<T>(vol);
  const T sl = step_length(vi, mp);
  const VolInfo<float> vi32 = vol_info<float>(vol);
  const float sl32 = step_length(vi32, mp);
  const bool shadow = mp->f32_control && sizeof(T) != sizeof(float);
  return with_grid(vol, [&](auto g) {
    parallel_rows(view->height, nthreads, [&](int y) {
      std::vector<uint8_t> dec;
      for (int x = 0; x < view->width; ++x) {
        Ctl ctl;
        if (shadow) {
          dec.clear();
          ctl.rec = &dec;
          drr_ray<float>(g, vi32, mp, dp, form_ray(vi32, view, sl32, x, y), x, y, &ctl);
          ctl.rec = nullptr;
          ctl.play = &dec;
        }
        const Ray<T> r = form_ray(vi, view, sl, x, y);
        const T v = drr_ray<T>(g, vi, mp, dp, r, x, y, &ctl);
        const long idx = (long)y * view->width + x;
        if (dp->output == ORC_DRR_RGBA8_REF) {
          const T tw = (T)dp->tint[3];
          uint8_t* o = (uint8_t*)out + 4 * idx;
          o[3] = to_u8(tw + (T(1) - tw / T(255)) * v);
          o[0] = to_u8(tw / (tw + v) * (T)dp->tint[0]);
          o[1] = to_u8(tw / (tw + v) * (T)dp->tint[1]);
          o[2] = to_u8(tw / (tw + v) * (T)dp->tint[2]);
        } else {
          ((T*)out)[idx] = v;
        }
      }
    });
    return 0;
  });
}

// ---------------------------------------------------------------------------
// 1D caster: V/CUDA_vtkCuda1DVolumeMapper_renderAlgo.cuh:27-227
template <typename T> struct RGBAT { T r, g, b, a, depth; };

template <typename T, typename VT>
RGBAT<T> cast1d_ray(const Grid<VT>& g, const VolInfo<T>& vi, const orc_tf1d* tf, const orc_march* mp,
                    const Ray<T>& r, int x, int y, Ctl* ctl) {
  Ctl none;
  if (!ctl) ctl = &none;
  T Tr = T(1), cr = T(0), cg = T(0), cb = T(0);
  const T lo = (T)tf->intensity_low, mult = (T)tf->intensity_mult;
  const T glo = (T)tf->gradient_low, gmult = (T)tf->gradient_mult;
  const T amb = (T)tf->ambient, dif = (T)tf->diffuse, spc = (T)tf->specular, spw = (T)tf->spec_power;
  const T term = (T)mp->term_T;
  const T j = jitter_of<T>(mp, x, y);
  int k = step_count(r.num_steps - j);
  V3<T> p{r.start.x + j * r.inc.x, r.start.y + j * r.inc.y, r.start.z + j * r.inc.z};
  const V3<T>& inc = r.inc;
  const T sx = vi.sp[0], sy = vi.sp[1], sz = vi.sp[2];
  const T ray_len = std::sqrt(inc.x * inc.x * sx * sx + inc.y * inc.y * sy * sy + inc.z * inc.z * sz * sz);
  // the 1D caster compares unconverted exclusion bounds (V/...1D...cuh:74)
  int es = fl2i_ru(r.exclude_start), ee = fl2i_rd(r.exclude_end);
  int k0 = k;
  ctl->ints(k, es, ee);
  k0 = k;
  const int kst = k;
  const V3<T> p0c{p.x + T(0.5), p.y + T(0.5), p.z + T(0.5)};
  // Sample coordinates are carried as cell coordinates (c = p + 0.5).
  // Precise mode: c follows from the step index, (p0 + 0.5) + t inc with one
  // rounding per axis (no drift from repeated p += inc); reference mode: the
  // incrementally updated p, plus 0.5.
  auto cur = [&]() -> V3<T> {
    if (!mp->precise) return V3<T>{p.x + T(0.5), p.y + T(0.5), p.z + T(0.5)};
    const T t = (T)(kst - k);
    return V3<T>{std::fma(t, inc.x, p0c.x), std::fma(t, inc.y, p0c.y), std::fma(t, inc.z, p0c.z)};
  };
  int a = 0, b = 0;
  auto accumulate = [&](T alpha, T u) {
    V3<T> gr;
    const V3<T> q = cur();
    gr.x = (trilinear<T>(g, q.x + T(1), q.y, q.z) - trilinear<T>(g, q.x - T(1), q.y, q.z)) * vi.srecip[0];
    gr.y = (trilinear<T>(g, q.x, q.y + T(1), q.z) - trilinear<T>(g, q.x, q.y - T(1), q.z)) * vi.srecip[1];
    gr.z = (trilinear<T>(g, q.x, q.y, q.z + T(1)) - trilinear<T>(g, q.x, q.y, q.z - T(1))) * vi.srecip[2];
    const T gm = std::sqrt(gr.x * gr.x + gr.y * gr.y + gr.z * gr.z);
    alpha *= std::isfinite(gmult) ? tf1d<T>(tf->galpha, tf->n, gmult * (gm - glo)) : T(1);
    const T pl = sat(std::fabs(gr.x * inc.x * sx + gr.y * inc.y * sy + gr.z * inc.z * sz) / (gm * ray_len));
    const T shd = amb + dif * pl;
    const T shs = spc * std::pow(pl, spw);
    const T m = Tr * alpha;
    Tr *= (T(1) - alpha);
    cr += m * sat(shd * tf1d<T>(tf->r, tf->n, u) + shs);
    cg += m * sat(shd * tf1d<T>(tf->g, tf->n, u) + shs);
    cb += m * sat(shd * tf1d<T>(tf->b, tf->n, u) + shs);
  };
  while (k > 0) {
    if (es > k && ee < k) {
      p.x += (T)(k - ee) * inc.x; p.y += (T)(k - ee) * inc.y; p.z += (T)(k - ee) * inc.z;
      k = ee;
      a = 0;
      b = 0;
      continue;
    }
    if (ctl->count) ++*ctl->count;
    const V3<T> q = cur();
    const T u = mult * (trilinear<T>(g, q.x, q.y, q.z) - lo);
    const T alpha = tf1d<T>(tf->alpha, tf->n, u);
    if (mp->skip_rule) {
      if (ctl->decide(alpha > T(0))) {
        a = b;
        b = 0;
        if (a) { p.x = p.x - inc.x; p.y = p.y - inc.y; p.z = p.z - inc.z; k += 1; }
        else { p.x = p.x + inc.x; p.y = p.y + inc.y; p.z = p.z + inc.z; k -= 1; }
        if (!a) accumulate(alpha, u);
        if (ctl->decide(Tr < term)) { Tr = T(0); break; }
      } else {
        if (!a) { p.x += inc.x; p.y += inc.y; p.z += inc.z; k--; }
        b = !a;
        p.x += inc.x; p.y += inc.y; p.z += inc.z; k--;
        a = 0;
      }
    } else {
      p.x += inc.x; p.y += inc.y; p.z += inc.z; k--;
      if (ctl->decide(alpha > T(0))) {
        accumulate(alpha, u);
        if (ctl->decide(Tr < term)) { Tr = T(0); break; }
      }
    }
  }
  RGBAT<T> o;
  o.a = T(1) - Tr;
  o.r = sat(cr);
  o.g = sat(cg);
  o.b = sat(cb);
  o.depth = (T)(k0 - k);
  return o;
}

// ---------------------------------------------------------------------------
// 2D caster: V/CUDA_vtkCuda2DVolumeMapper_renderAlgo.cuh:30-470
template <typename T, typename VT>
RGBAT<T> cast2d_ray(const Grid<VT>& g, const Grid<uint8_t>* gm_grid, const VolInfo<T>& vi, const orc_tf2d* tf,
                    const orc_march* mp, const Ray<T>& r, int x, int y, Ctl* ctl) {
  Ctl none;
  if (!ctl) ctl = &none;
  T Tr = T(1), cr = T(0), cg = T(0), cb = T(0);
  const T lo = (T)tf->intensity_low, mult = (T)tf->intensity_mult;
  const T glo = (T)tf->gradient_low, gmult = (T)tf->gradient_mult, goff = (T)tf->gradient_offset;
  const T term = (T)mp->term_T;
  const bool dual = tf->k_alpha != nullptr;
  const T j = jitter_of<T>(mp, x, y);
  int k = step_count(r.num_steps - j);
  V3<T> p{r.start.x + j * r.inc.x, r.start.y + j * r.inc.y, r.start.z + j * r.inc.z};
  const V3<T>& inc = r.inc;
  const T sx = vi.sp[0], sy = vi.sp[1], sz = vi.sp[2];
  const T ray_len = std::sqrt(inc.x * inc.x * sx * sx + inc.y * inc.y * sy * sy + inc.z * inc.z * sz * sz);
  // WithDual rounds the exclusion bounds up/down, WithoutDual down/up (V/...2D...cuh:388-390, 448-450)
  int es = dual ? fl2i_ru(r.exclude_start) : fl2i_rd(r.exclude_start);
  int ee = dual ? fl2i_rd(r.exclude_end) : fl2i_ru(r.exclude_end);
  es = k - es;
  ee = k - ee;
  ctl->ints(k, es, ee);
  const int k0 = k;
  const int kst = k;
  const V3<T> p0c{p.x + T(0.5), p.y + T(0.5), p.z + T(0.5)};
  // Sample coordinates are carried as cell coordinates (c = p + 0.5).
  // Precise mode: c follows from the step index, (p0 + 0.5) + t inc with one
  // rounding per axis (no drift from repeated p += inc); reference mode: the
  // incrementally updated p, plus 0.5.
  auto cur = [&]() -> V3<T> {
    if (!mp->precise) return V3<T>{p.x + T(0.5), p.y + T(0.5), p.z + T(0.5)};
    const T t = (T)(kst - k);
    return V3<T>{std::fma(t, inc.x, p0c.x), std::fma(t, inc.y, p0c.y), std::fma(t, inc.z, p0c.z)};
  };
  int a = 0, b = 0;
  while (k > 0) {
    const bool in_key = (es >= k && ee < k);
    if (!dual && in_key) {
      p.x += (T)(k - ee) * inc.x; p.y += (T)(k - ee) * inc.y; p.z += (T)(k - ee) * inc.z;
      k = ee;
      a = 0;
      b = 0;
      continue;
    }
    if (ctl->count) ++*ctl->count;
    const V3<T> q = cur();
    const T u = mult * (trilinear<T>(g, q.x, q.y, q.z) - lo);
    V3<T> gr{T(0), T(0), T(0)};
    T gidx, gmag;
    if (gm_grid) {
      gmag = trilinear<T>(*gm_grid, q.x, q.y, q.z) * (T)tf->gradmag_scale;
      gidx = gmult * (std::log2(gmag * gmag + goff) + glo);
    } else {
      gr.x = (trilinear<T>(g, q.x + T(1), q.y, q.z) - trilinear<T>(g, q.x - T(1), q.y, q.z)) * vi.srecip[0];
      gr.y = (trilinear<T>(g, q.x, q.y + T(1), q.z) - trilinear<T>(g, q.x, q.y - T(1), q.z)) * vi.srecip[1];
      gr.z = (trilinear<T>(g, q.x, q.y, q.z + T(1)) - trilinear<T>(g, q.x, q.y, q.z - T(1))) * vi.srecip[2];
      const T gm2 = gr.x * gr.x + gr.y * gr.y + gr.z * gr.z;
      gidx = gmult * (std::log2(gm2 + goff) + glo);
      gmag = std::sqrt(gm2);
    }
    const int ti = tf2d_index<T>(tf->n, u, gidx);
    const bool kt = dual && in_key;
    const T alpha = (T)(kt ? tf->k_alpha : tf->alpha)[ti];
    // no gradient direction in precomputed-magnitude mode: head-light shading (PL = 1)
    const T pl = gm_grid ? T(1)
                         : sat(std::fabs(gr.x * inc.x * sx + gr.y * inc.y * sy + gr.z * inc.z * sz) / (gmag * ray_len));
    const T shd = (T)(kt ? tf->k_ambient : tf->ambient)[ti] + (T)(kt ? tf->k_diffuse : tf->diffuse)[ti] * pl;
    const T shs = (T)(kt ? tf->k_specular : tf->specular)[ti] * std::pow(pl, (T)(kt ? tf->k_spec_power : tf->spec_power)[ti]);
    auto accumulate = [&]() {
      const T m = Tr * alpha;
      Tr *= (T(1) - alpha);
      cr += m * sat(shd * (T)(kt ? tf->k_r : tf->r)[ti] + shs);
      cg += m * sat(shd * (T)(kt ? tf->k_g : tf->g)[ti] + shs);
      cb += m * sat(shd * (T)(kt ? tf->k_b : tf->b)[ti] + shs);
    };
    if (mp->skip_rule) {
      if (ctl->decide(alpha > T(0))) {
        a = b;
        b = 0;
        if (a) { p.x = p.x - inc.x; p.y = p.y - inc.y; p.z = p.z - inc.z; k += 1; }
        else { p.x = p.x + inc.x; p.y = p.y + inc.y; p.z = p.z + inc.z; k -= 1; }
        if (!a) accumulate();
        if (ctl->decide(Tr < term)) { Tr = T(0); break; }
      } else {
        if (!a) { p.x += inc.x; p.y += inc.y; p.z += inc.z; k--; }
        b = !a;
        p.x += inc.x; p.y += inc.y; p.z += inc.z; k--;
        a = 0;
      }
    } else {
      p.x += inc.x; p.y += inc.y; p.z += inc.z; k--;
      if (ctl->decide(alpha > T(0))) {
        accumulate();
        if (ctl->decide(Tr < term)) { Tr = T(0); break; }
      }
    }
  }
  RGBAT<T> o;
  o.a = T(1) - Tr;
  o.r = cr;
  o.g = cg;
  o.b = cb;
  o.depth = (T)(k0 - k);
  return o;
}

template <typename T>
void store_rgba(const RGBAT<T>& o, long idx, uint8_t* rgba8, T* rgbaf, T* depth) {
  if (rgba8) {
    rgba8[4 * idx + 0] = to_u8(T(255) * o.r);
    rgba8[4 * idx + 1] = to_u8(T(255) * o.g);
    rgba8[4 * idx + 2] = to_u8(T(255) * o.b);
    rgba8[4 * idx + 3] = to_u8(T(255) * o.a);
  }
  if (rgbaf) {
    rgbaf[4 * idx + 0] = o.r;
    rgbaf[4 * idx + 1] = o.g;
    rgbaf[4 * idx + 2] = o.b;
    rgbaf[4 * idx + 3] = o.a;
  }
  if (depth) depth[idx] = o.depth;
}

template <typename T>
int cast1d_impl(const orc_volume* vol, const orc_tf1d* tf, const orc_view* view, const orc_march* mp,
                uint8_t* rgba8, T* rgbaf, T* depth, int nthreads) {
  if (!vol || !tf || !view || !mp || tf->n < 1) return 1;
  const VolInfo<T> vi = vol_info<T>(vol);
  const T sl = step_length(vi, mp);
  const VolInfo<float> vi32 = vol_info<float>(vol);
  const float sl32 = step_length(vi32, mp);
  const bool shadow = mp->f32_control && sizeof(T) != sizeof(float);
  return with_grid(vol, [&](auto g) {
    parallel_rows(view->height, nthreads, [&](int y) {
      std::vector<uint8_t> dec;
      for (int x = 0; x < view->width; ++x) {
        Ctl ctl;
        if (shadow) {
          dec.clear();
          ctl.rec = &dec;
          cast1d_ray<float>(g, vi32, tf, mp, form_ray(vi32, view, sl32, x, y), x, y, &ctl);
          ctl.rec = nullptr;
          ctl.play = &dec;
        }
        const Ray<T> r = form_ray(vi, view, sl, x, y);
        store_rgba(cast1d_ray<T>(g, vi, tf, mp, r, x, y, &ctl), (long)y * view->width + x, rgba8, rgbaf, depth);
      }
    });
    return 0;
  });
}

template <typename T>
int cast2d_impl(const orc_volume* vol, const orc_tf2d* tf, const orc_view* view, const orc_march* mp,
                uint8_t* rgba8, T* rgbaf, T* depth, int nthreads) {
  if (!vol || !tf || !view || !mp || tf->n < 1) return 1;
  const VolInfo<T> vi = vol_info<T>(vol);
  const T sl = step_length(vi, mp);
  Grid<uint8_t> gmg{tf->gradmag, vol->dims[0], vol->dims[1], vol->dims[2]};
  const Grid<uint8_t>* gmp = tf->gradmag ? &gmg : nullptr;
  const VolInfo<float> vi32 = vol_info<float>(vol);
  const float sl32 = step_length(vi32, mp);
  const bool shadow = mp->f32_control && sizeof(T) != sizeof(float);
  return with_grid(vol, [&](auto g) {
    parallel_rows(view->height, nthreads, [&](int y) {
      std::vector<uint8_t> dec;
      for (int x = 0; x < view->width; ++x) {
        Ctl ctl;
        if (shadow) {
          dec.clear();
          ctl.rec = &dec;
          cast2d_ray<float>(g, gmp, vi32, tf, mp, form_ray(vi32, view, sl32, x, y), x, y, &ctl);
          ctl.rec = nullptr;
          ctl.play = &dec;
        }
        const Ray<T> r = form_ray(vi, view, sl, x, y);
        store_rgba(cast2d_ray<T>(g, gmp, vi, tf, mp, r, x, y, &ctl), (long)y * view->width + x, rgba8, rgbaf,
                   depth);
      }
    });
    return 0;
  });
}

// ---------------------------------------------------------------------------
// In/Ex-logic 2D caster: V/CUDA_vtkCuda2DInExLogicVolumeMapper_renderAlgo.cuh:24-179 (march) and
// :241-302 (composite).  Differences from the 2D caster: the step count is
// floor(numSteps) (the offset moves the start only), a sample contributes when
// inEx > 0 or alpha > 0, the first accumulated sample outside the keyhole
// range [ee, es] with inEx > 0.5 ends the ray with nothing accumulated, the
// output is saturated, and with the black keyhole a ray that accumulated is
// opaque.
template <typename T, typename VT>
RGBAT<T> inex_ray(const Grid<VT>& g, const VolInfo<T>& vi, const orc_tf2d* tf, const float* inex, int black,
                  const orc_march* mp, const Ray<T>& r, int x, int y, Ctl* ctl) {
  Ctl none;
  if (!ctl) ctl = &none;
  T Tr = T(1), cr = T(0), cg = T(0), cb = T(0);
  const T lo = (T)tf->intensity_low, mult = (T)tf->intensity_mult;
  const T glo = (T)tf->gradient_low, gmult = (T)tf->gradient_mult, goff = (T)tf->gradient_offset;
  const T term = (T)mp->term_T;
  const T j = jitter_of<T>(mp, x, y);
  V3<T> p{r.start.x + j * r.inc.x, r.start.y + j * r.inc.y, r.start.z + j * r.inc.z};
  int k = step_count(r.num_steps);
  int es = k - fl2i_ru(r.exclude_start);
  int ee = k - fl2i_rd(r.exclude_end);
  ctl->ints(k, es, ee);
  const T depth0 = j + (T)k;
  const int kst = k;
  const V3<T>& inc = r.inc;
  const T sx = vi.sp[0], sy = vi.sp[1], sz = vi.sp[2];
  const T ray_len = std::sqrt(inc.x * inc.x * sx * sx + inc.y * inc.y * sy * sy + inc.z * inc.z * sz * sz);
  const V3<T> p0c{p.x + T(0.5), p.y + T(0.5), p.z + T(0.5)};
  auto cur = [&]() -> V3<T> {
    if (!mp->precise) return V3<T>{p.x + T(0.5), p.y + T(0.5), p.z + T(0.5)};
    const T t = (T)(kst - k);
    return V3<T>{std::fma(t, inc.x, p0c.x), std::fma(t, inc.y, p0c.y), std::fma(t, inc.z, p0c.z)};
  };
  bool skip = false, back = false, special = false;
  while (k > 0) {
    if (ctl->count) ++*ctl->count;
    const V3<T> q = cur();
    const T u = mult * (trilinear<T>(g, q.x, q.y, q.z) - lo);
    V3<T> gr;
    gr.x = (trilinear<T>(g, q.x + T(1), q.y, q.z) - trilinear<T>(g, q.x - T(1), q.y, q.z)) * vi.srecip[0];
    gr.y = (trilinear<T>(g, q.x, q.y + T(1), q.z) - trilinear<T>(g, q.x, q.y - T(1), q.z)) * vi.srecip[1];
    gr.z = (trilinear<T>(g, q.x, q.y, q.z + T(1)) - trilinear<T>(g, q.x, q.y, q.z - T(1))) * vi.srecip[2];
    const T gm2 = gr.x * gr.x + gr.y * gr.y + gr.z * gr.z;
    const T gidx = gmult * (std::log2(gm2 + goff) + glo);
    const T gmag = std::sqrt(gm2);
    const int ti = tf2d_index<T>(tf->n, u, gidx);
    const T ie = (T)inex[ti], alpha = (T)tf->alpha[ti];
    const T pl = sat(std::fabs(gr.x * inc.x * sx + gr.y * inc.y * sy + gr.z * inc.z * sz) / (gmag * ray_len));
    const T shd = (T)tf->ambient[ti] + (T)tf->diffuse[ti] * pl;
    const T shs = (T)tf->specular[ti] * std::pow(pl, (T)tf->spec_power[ti]);
    // accumulate after the step; false: the ray ends here (excluded)
    auto accumulate = [&]() -> bool {
      if (ctl->decide(!special && !(es >= k && ee <= k) && ie > T(0.5))) return false;
      special = true;
      const T m = Tr * alpha;
      Tr *= (T(1) - alpha);
      cr += m * sat(shd * (T)tf->r[ti] + shs);
      cg += m * sat(shd * (T)tf->g[ti] + shs);
      cb += m * sat(shd * (T)tf->b[ti] + shs);
      return true;
    };
    const bool contributes = ctl->decide(ie > T(0) || alpha > T(0));
    if (mp->skip_rule) {
      if (contributes) {
        back = skip;
        skip = false;
        if (back) { p.x = p.x - inc.x; p.y = p.y - inc.y; p.z = p.z - inc.z; k += 1; }
        else { p.x = p.x + inc.x; p.y = p.y + inc.y; p.z = p.z + inc.z; k -= 1; }
        if (!back && !accumulate()) break;
        if (ctl->decide(Tr < term)) { Tr = T(0); break; }
      } else {
        if (!back) { p.x += inc.x; p.y += inc.y; p.z += inc.z; k--; }
        skip = !back;
        p.x += inc.x; p.y += inc.y; p.z += inc.z; k--;
        back = false;
      }
    } else {
      p.x += inc.x; p.y += inc.y; p.z += inc.z; k--;
      if (contributes) {
        if (!accumulate()) break;
        if (ctl->decide(Tr < term)) { Tr = T(0); break; }
      }
    }
  }
  RGBAT<T> o;
  o.a = (black && special) ? T(1) : T(1) - Tr;
  o.r = cr;
  o.g = cg;
  o.b = cb;
  o.depth = depth0 - (T)k;
  return o;
}

// ---------------------------------------------------------------------------
// Dual-image caster: V/CUDA_vtkCudaDualImageVolumeMapper_renderAlgo.cuh:33-208 (no keyhole TF),
// :210-380 (keyhole TF), :383-517 (composites).  The 2D table is indexed by
// the two components, shading uses the gradient of component 1, and the step
// count is trunc(floor(numSteps) - offset) (an int minus a float stored back
// in an int).  The exclusion bounds round up/down in both variants.
template <typename T> inline int dual_step_count(T num_steps, T j) {
  if (!(num_steps > T(0) && num_steps < T(16777216))) return 0;
  const T f = (T)(int)std::floor(num_steps) - j;
  return (int)f;
}

template <typename T, typename VT>
RGBAT<T> dual_ray(const Grid<VT>& g1, const Grid<VT>& g2, const VolInfo<T>& vi, const orc_tf2d* tf,
                  const orc_march* mp, const Ray<T>& r, int x, int y, Ctl* ctl) {
  Ctl none;
  if (!ctl) ctl = &none;
  T Tr = T(1), cr = T(0), cg = T(0), cb = T(0);
  const T lo1 = (T)tf->intensity_low, m1 = (T)tf->intensity_mult;
  const T lo2 = (T)tf->gradient_low, m2 = (T)tf->gradient_mult;
  const T term = (T)mp->term_T;
  const bool dual = tf->k_alpha != nullptr;
  const T j = jitter_of<T>(mp, x, y);
  V3<T> p{r.start.x + j * r.inc.x, r.start.y + j * r.inc.y, r.start.z + j * r.inc.z};
  int k = dual_step_count<T>(r.num_steps, j);
  int es = k - fl2i_ru(r.exclude_start);
  int ee = k - fl2i_rd(r.exclude_end);
  ctl->ints(k, es, ee);
  const int k0 = k, kst = k;
  const V3<T>& inc = r.inc;
  const T sx = vi.sp[0], sy = vi.sp[1], sz = vi.sp[2];
  const T ray_len = std::sqrt(inc.x * inc.x * sx * sx + inc.y * inc.y * sy * sy + inc.z * inc.z * sz * sz);
  const V3<T> p0c{p.x + T(0.5), p.y + T(0.5), p.z + T(0.5)};
  auto cur = [&]() -> V3<T> {
    if (!mp->precise) return V3<T>{p.x + T(0.5), p.y + T(0.5), p.z + T(0.5)};
    const T t = (T)(kst - k);
    return V3<T>{std::fma(t, inc.x, p0c.x), std::fma(t, inc.y, p0c.y), std::fma(t, inc.z, p0c.z)};
  };
  int a = 0, b = 0;
  while (k > 0) {
    const bool in_key = (es >= k && ee < k);
    if (!dual && in_key) {
      p.x += (T)(k - ee) * inc.x; p.y += (T)(k - ee) * inc.y; p.z += (T)(k - ee) * inc.z;
      k = ee;
      a = 0;
      b = 0;
      continue;
    }
    if (ctl->count) ++*ctl->count;
    const V3<T> q = cur();
    const T u1 = m1 * (trilinear<T>(g1, q.x, q.y, q.z) - lo1);
    const T u2 = m2 * (trilinear<T>(g2, q.x, q.y, q.z) - lo2);
    V3<T> gr;
    gr.x = (trilinear<T>(g1, q.x + T(1), q.y, q.z) - trilinear<T>(g1, q.x - T(1), q.y, q.z)) * vi.srecip[0];
    gr.y = (trilinear<T>(g1, q.x, q.y + T(1), q.z) - trilinear<T>(g1, q.x, q.y - T(1), q.z)) * vi.srecip[1];
    gr.z = (trilinear<T>(g1, q.x, q.y, q.z + T(1)) - trilinear<T>(g1, q.x, q.y, q.z - T(1))) * vi.srecip[2];
    const T gmag = std::sqrt(gr.x * gr.x + gr.y * gr.y + gr.z * gr.z);
    const int ti = tf2d_index<T>(tf->n, u1, u2);
    const bool kt = dual && in_key;
    const T alpha = (T)(kt ? tf->k_alpha : tf->alpha)[ti];
    const T pl = sat(std::fabs(gr.x * inc.x * sx + gr.y * inc.y * sy + gr.z * inc.z * sz) / (gmag * ray_len));
    const T shd = (T)(kt ? tf->k_ambient : tf->ambient)[ti] + (T)(kt ? tf->k_diffuse : tf->diffuse)[ti] * pl;
    const T shs = (T)(kt ? tf->k_specular : tf->specular)[ti] * std::pow(pl, (T)(kt ? tf->k_spec_power : tf->spec_power)[ti]);
    auto accumulate = [&]() {
      const T m = Tr * alpha;
      Tr *= (T(1) - alpha);
      cr += m * sat(shd * (T)(kt ? tf->k_r : tf->r)[ti] + shs);
      cg += m * sat(shd * (T)(kt ? tf->k_g : tf->g)[ti] + shs);
      cb += m * sat(shd * (T)(kt ? tf->k_b : tf->b)[ti] + shs);
    };
    if (mp->skip_rule) {
      if (ctl->decide(alpha > T(0))) {
        a = b;
        b = 0;
        if (a) { p.x = p.x - inc.x; p.y = p.y - inc.y; p.z = p.z - inc.z; k += 1; }
        else { p.x = p.x + inc.x; p.y = p.y + inc.y; p.z = p.z + inc.z; k -= 1; }
        if (!a) accumulate();
        if (ctl->decide(Tr < term)) { Tr = T(0); break; }
      } else {
        if (!a) { p.x += inc.x; p.y += inc.y; p.z += inc.z; k--; }
        b = !a;
        p.x += inc.x; p.y += inc.y; p.z += inc.z; k--;
        a = 0;
      }
    } else {
      p.x += inc.x; p.y += inc.y; p.z += inc.z; k--;
      if (ctl->decide(alpha > T(0))) {
        accumulate();
        if (ctl->decide(Tr < term)) { Tr = T(0); break; }
      }
    }
  }
  RGBAT<T> o;
  o.a = T(1) - Tr;
  o.r = cr;
  o.g = cg;
  o.b = cb;
  o.depth = (T)(k0 - k);
  return o;
}

template <typename T>
int inex_impl(const orc_volume* vol, const orc_tf2d* tf, const float* inex, int black, const orc_view* view,
              const orc_march* mp, uint8_t* rgba8, T* rgbaf, T* depth, int nthreads) {
  if (!vol || !tf || !inex || !view || !mp || tf->n < 1) return 1;
  const VolInfo<T> vi = vol_info<T>(vol);
  const T sl = step_length(vi, mp);
  const VolInfo<float> vi32 = vol_info<float>(vol);
  const float sl32 = step_length(vi32, mp);
  const bool shadow = mp->f32_control && sizeof(T) != sizeof(float);
  return with_grid(vol, [&](auto g) {
    parallel_rows(view->height, nthreads, [&](int y) {
      std::vector<uint8_t> dec;
      for (int x = 0; x < view->width; ++x) {
        Ctl ctl;
        if (shadow) {
          dec.clear();
          ctl.rec = &dec;
          inex_ray<float>(g, vi32, tf, inex, black, mp, form_ray(vi32, view, sl32, x, y), x, y, &ctl);
          ctl.rec = nullptr;
          ctl.play = &dec;
        }
        const Ray<T> r = form_ray(vi, view, sl, x, y);
        RGBAT<T> o = inex_ray<T>(g, vi, tf, inex, black, mp, r, x, y, &ctl);
        o.r = sat(o.r);
        o.g = sat(o.g);
        o.b = sat(o.b);
        o.a = sat(o.a);
        store_rgba(o, (long)y * view->width + x, rgba8, rgbaf, depth);
      }
    });
    return 0;
  });
}

template <typename T>
int dual_impl(const orc_volume* v1, const orc_volume* v2, const orc_tf2d* tf, const orc_view* view,
              const orc_march* mp, uint8_t* rgba8, T* rgbaf, T* depth, int nthreads) {
  if (!v1 || !v2 || !tf || !view || !mp || tf->n < 1) return 1;
  if (v1->dtype != v2->dtype || v1->dims[0] != v2->dims[0] || v1->dims[1] != v2->dims[1] ||
      v1->dims[2] != v2->dims[2])
    return 1;
  const VolInfo<T> vi = vol_info<T>(v1);
  const T sl = step_length(vi, mp);
  const VolInfo<float> vi32 = vol_info<float>(v1);
  const float sl32 = step_length(vi32, mp);
  const bool shadow = mp->f32_control && sizeof(T) != sizeof(float);
  return with_grid(v1, [&](auto g1) {
    using G = decltype(g1);
    const G g2{(decltype(g1.d))v2->data, v2->dims[0], v2->dims[1], v2->dims[2]};
    parallel_rows(view->height, nthreads, [&](int y) {
      std::vector<uint8_t> dec;
      for (int x = 0; x < view->width; ++x) {
        Ctl ctl;
        if (shadow) {
          dec.clear();
          ctl.rec = &dec;
          dual_ray<float>(g1, g2, vi32, tf, mp, form_ray(vi32, view, sl32, x, y), x, y, &ctl);
          ctl.rec = nullptr;
          ctl.play = &dec;
        }
        const Ray<T> r = form_ray(vi, view, sl, x, y);
        store_rgba(dual_ray<T>(g1, g2, vi, tf, mp, r, x, y, &ctl), (long)y * view->width + x, rgba8, rgbaf, depth);
      }
    });
    return 0;
  });
}

template <typename T>
int form_rays_impl(const orc_volume* vol, const orc_view* view, const orc_march* mp, T* out9, T* mind, T* maxd) {
  if (!vol || !view || !mp) return 1;
  const VolInfo<T> vi = vol_info<T>(vol);
  const T sl = step_length(vi, mp);
  const long n = (long)view->width * view->height;
  for (int y = 0; y < view->height; ++y)
    for (int x = 0; x < view->width; ++x) {
      const Ray<T> r = form_ray(vi, view, sl, x, y);
      const long i = (long)y * view->width + x;
      if (out9) {
        out9[0 * n + i] = r.start.x; out9[1 * n + i] = r.start.y; out9[2 * n + i] = r.start.z;
        out9[3 * n + i] = r.inc.x; out9[4 * n + i] = r.inc.y; out9[5 * n + i] = r.inc.z;
        out9[6 * n + i] = r.num_steps; out9[7 * n + i] = r.exclude_start; out9[8 * n + i] = r.exclude_end;
      }
      if (mind) mind[i] = r.min_depth;
      if (maxd) maxd[i] = r.max_depth;
    }
  return 0;
}

}  // namespace

extern "C" {

const float* orc_jitter_table(void) { return kJitter; }

int orc_form_rays_f32(const orc_volume* vol, const orc_view* view, const orc_march* mp, float* out9,
                      float* min_depth, float* max_depth) {
  return form_rays_impl<float>(vol, view, mp, out9, min_depth, max_depth);
}
int orc_form_rays_f64(const orc_volume* vol, const orc_view* view, const orc_march* mp, double* out9,
                      double* min_depth, double* max_depth) {
  return form_rays_impl<double>(vol, view, mp, out9, min_depth, max_depth);
}

int orc_drr_f32(const orc_volume* vol, const orc_view* view, const orc_march* mp, const orc_drr* dp, void* out,
                int nthreads) {
  return drr_impl<float>(vol, view, mp, dp, out, nthreads);
}
int orc_drr_f64(const orc_volume* vol, const orc_view* view, const orc_march* mp, const orc_drr* dp, void* out,
                int nthreads) {
  return drr_impl<double>(vol, view, mp, dp, out, nthreads);
}

int orc_1d_f32(const orc_volume* vol, const orc_tf1d* tf, const orc_view* view, const orc_march* mp, uint8_t* rgba8,
               float* rgbaf, float* depth, int nthreads) {
  return cast1d_impl<float>(vol, tf, view, mp, rgba8, rgbaf, depth, nthreads);
}
int orc_1d_f64(const orc_volume* vol, const orc_tf1d* tf, const orc_view* view, const orc_march* mp, uint8_t* rgba8,
               double* rgbaf, double* depth, int nthreads) {
  return cast1d_impl<double>(vol, tf, view, mp, rgba8, rgbaf, depth, nthreads);
}

int orc_2d_f32(const orc_volume* vol, const orc_tf2d* tf, const orc_view* view, const orc_march* mp, uint8_t* rgba8,
               float* rgbaf, float* depth, int nthreads) {
  return cast2d_impl<float>(vol, tf, view, mp, rgba8, rgbaf, depth, nthreads);
}
int orc_2d_f64(const orc_volume* vol, const orc_tf2d* tf, const orc_view* view, const orc_march* mp, uint8_t* rgba8,
               double* rgbaf, double* depth, int nthreads) {
  return cast2d_impl<double>(vol, tf, view, mp, rgba8, rgbaf, depth, nthreads);
}

// V/...renderAlgo.cu:449-512.  The reference reads depth[idx+W] and depth[idx+1]
// past the last row/column; here the neighbour index is clamped to the image.
int orc_post_shade_f32(const orc_shade* sp, int W, int H, float* depth, const float* min_depth,
                       const float* max_depth, uint8_t* rgba8) {
  if (!sp || !depth || !min_depth || !max_depth || !rgba8) return 1;
  const long n = (long)W * H;
  for (long i = 0; i < n; ++i) {
    const float cur = depth[i], mx = max_depth[i], mn = min_depth[i];
    depth[i] = (mx > 0.0f) ? (cur + mn) / mx : 1.0f;
  }
  std::vector<uint8_t> src(rgba8, rgba8 + 4 * n);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const long i = (long)y * W + x;
      const long iy = (long)std::min(y + 1, H - 1) * W + x;
      const long ix = (long)y * W + std::min(x + 1, W - 1);
      const float d0 = depth[i], dy = depth[iy], dx = depth[ix];
      float gm = sqrt_rz((dx - d0) * (dx - d0) + (dy - d0) * (dy - d0));
      gm = 1.0f - sp->celr * sat((gm - sp->cela) * sp->celc);
      gm *= 1.0f - sp->disr * sat((d0 - sp->disa) * sp->disc);
      for (int c = 0; c < 3; ++c) rgba8[4 * i + c] = to_u8(gm * (float)src[4 * i + c]);
    }
  return 0;
}

int orc_inex_f32(const orc_volume* vol, const orc_tf2d* tf, const float* inex, int black_keyhole,
                 const orc_view* view, const orc_march* mp, uint8_t* rgba8, float* rgbaf, float* depth, int nthreads) {
  return inex_impl<float>(vol, tf, inex, black_keyhole, view, mp, rgba8, rgbaf, depth, nthreads);
}
int orc_inex_f64(const orc_volume* vol, const orc_tf2d* tf, const float* inex, int black_keyhole,
                 const orc_view* view, const orc_march* mp, uint8_t* rgba8, double* rgbaf, double* depth,
                 int nthreads) {
  return inex_impl<double>(vol, tf, inex, black_keyhole, view, mp, rgba8, rgbaf, depth, nthreads);
}
int orc_dual_f32(const orc_volume* vol1, const orc_volume* vol2, const orc_tf2d* tf, const orc_view* view,
                 const orc_march* mp, uint8_t* rgba8, float* rgbaf, float* depth, int nthreads) {
  return dual_impl<float>(vol1, vol2, tf, view, mp, rgba8, rgbaf, depth, nthreads);
}
int orc_dual_f64(const orc_volume* vol1, const orc_volume* vol2, const orc_tf2d* tf, const orc_view* view,
                 const orc_march* mp, uint8_t* rgba8, double* rgbaf, double* depth, int nthreads) {
  return dual_impl<double>(vol1, vol2, tf, view, mp, rgba8, rgbaf, depth, nthreads);
}

// The In/Ex-logic cel shade (V/CUDA_vtkCuda2DInExLogicVolumeMapper_renderAlgo.cuh:182-239) after
// normBuffer: the squared depth difference (no square root) and, for pixels
// with zero alpha, alpha = (signed char)(255 - 255 g) stored as an unsigned
// byte (float-to-char conversion clamps to [-128, 127]).
int orc_post_shade_inex_f32(const orc_shade* sp, int W, int H, float* depth, const float* min_depth,
                            const float* max_depth, uint8_t* rgba8) {
  if (!sp || !depth || !min_depth || !max_depth || !rgba8) return 1;
  const long n = (long)W * H;
  for (long i = 0; i < n; ++i) {
    const float cur = depth[i], mx = max_depth[i], mn = min_depth[i];
    depth[i] = (mx > 0.0f) ? (cur + mn) / mx : 1.0f;
  }
  std::vector<uint8_t> src(rgba8, rgba8 + 4 * n);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const long i = (long)y * W + x;
      const long iy = (long)std::min(y + 1, H - 1) * W + x;
      const long ix = (long)y * W + std::min(x + 1, W - 1);
      const float d0 = depth[i], dy = depth[iy], dx = depth[ix];
      float gm = (dx - d0) * (dx - d0) + (dy - d0) * (dy - d0);
      gm = 1.0f - sp->celr * sat((gm - sp->cela) * sp->celc);
      gm *= 1.0f - sp->disr * sat((d0 - sp->disa) * sp->disc);
      for (int c = 0; c < 3; ++c) rgba8[4 * i + c] = to_u8(gm * (float)src[4 * i + c]);
      if (src[4 * i + 3] == 0) {
        const float v = 255.0f - 255.0f * gm;
        int t = std::isnan(v) ? 0 : (int)std::fmax(-128.0f, std::fmin(127.0f, std::trunc(v)));
        rgba8[4 * i + 3] = (uint8_t)(int8_t)t;
      }
    }
  return 0;
}

// Rays the precise marches dropped since the last reset (see samples_in_cells).
long long orc_unsafe_rays(int reset) {
  const long long v = __atomic_load_n(&g_unsafe_rays, __ATOMIC_RELAXED);
  if (reset) __atomic_store_n(&g_unsafe_rays, 0, __ATOMIC_RELAXED);
  return v;
}

long long orc_count_samples(const orc_volume* vol, const orc_view* view, const orc_march* mp, const orc_tf1d* tf1d,
                            const orc_tf2d* tf2d, int nthreads) {
  if (!vol || !view || !mp) return -1;
  orc_march plain = *mp;
  plain.skip_rule = 0;
  const VolInfo<double> vi = vol_info<double>(vol);
  const double sl = step_length(vi, &plain);
  std::vector<long long> per_row(view->height, 0);
  orc_drr dp{0.0f, 0.0f, 1.0f, ORC_DRR_F32_L, 0, {0, 0, 0, 0}};
  Grid<uint8_t> gmg{tf2d ? tf2d->gradmag : nullptr, vol->dims[0], vol->dims[1], vol->dims[2]};
  const Grid<uint8_t>* gmp = (tf2d && tf2d->gradmag) ? &gmg : nullptr;
  int rc = with_grid(vol, [&](auto g) {
    parallel_rows(view->height, nthreads, [&](int y) {
      long long c = 0;
      Ctl ctl;
      ctl.count = &c;
      for (int x = 0; x < view->width; ++x) {
        const Ray<double> r = form_ray(vi, view, sl, x, y);
        if (tf1d) cast1d_ray<double>(g, vi, tf1d, &plain, r, x, y, &ctl);
        else if (tf2d) cast2d_ray<double>(g, gmp, vi, tf2d, &plain, r, x, y, &ctl);
        else drr_ray<double>(g, vi, &plain, &dp, r, x, y, &ctl);
      }
      per_row[y] = c;
    });
    return 0;
  });
  if (rc) return -1;
  long long tot = 0;
  for (long long c : per_row) tot += c;
  return tot;
}

long long orc_count_drr_evaluated(const orc_volume* vol, const orc_view* view, const orc_march* mp,
                                  const orc_drr* dp, int nthreads) {
  if (!vol || !view || !mp || !dp) return -1;
  const VolInfo<float> vi = vol_info<float>(vol);
  const float sl = step_length(vi, mp);
  std::vector<long long> per_row(view->height, 0);
  int rc = with_grid(vol, [&](auto g) {
    parallel_rows(view->height, nthreads, [&](int y) {
      long long c = 0;
      Ctl ctl;
      ctl.count = &c;
      for (int x = 0; x < view->width; ++x) drr_ray<float>(g, vi, mp, dp, form_ray(vi, view, sl, x, y), x, y, &ctl);
      per_row[y] = c;
    });
    return 0;
  });
  if (rc) return -1;
  long long tot = 0;
  for (long long c : per_row) tot += c;
  return tot;
}

double orc_sample_f64(const orc_volume* vol, double px, double py, double pz) {
  double out = 0.0;
  with_grid(vol, [&](auto g) { out = trilinear<double>(g, px + 0.5, py + 0.5, pz + 0.5); return 0; });
  return out;
}
float orc_sample_f32(const orc_volume* vol, float px, float py, float pz) {
  float out = 0.0f;
  with_grid(vol, [&](auto g) { out = trilinear<float>(g, px + 0.5f, py + 0.5f, pz + 0.5f); return 0; });
  return out;
}

float orc_tf1d_lookup_f32(const float* table, int n, float u) { return tf1d<float>(table, n, u); }
float orc_tf2d_lookup_f32(const float* table, int n, float u, float v) { return table[tf2d_index<float>(n, u, v)]; }

}  // extern "C"
