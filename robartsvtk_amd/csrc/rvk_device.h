// rvk_device.h -- device-side building blocks shared by the gfx950 casters.
//
// Semantics follow the reference kernels (paths relative to the reference
// root, V/ = Visualization/vtkCudaVisualization):
//   ray setup .......... V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:38-447
//   volume texture ..... unnormalised / linear / clamp (V/CUDA_vtkCuda1DVolumeMapper_renderAlgo.cuh:305-316)
//   1D TF texture ...... normalised / linear / clamp  (V/CUDA_vtkCuda1DVolumeMapper_renderAlgo.cuh:248-267)
//   2D TF texture ...... normalised / point / clamp   (V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:21-25)
// The texture unit is replaced by one vector load per sample from a
// native-type cell volume (every 2x2x2 trilinear footprint stored whole) with
// exact FP32 weights.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace rvk {

// ---------------------------------------------------------------------------
// Parameters every caster receives by value (no __constant__ module state).
struct VolGeom {
  int n[3];        // dims
  uint32_t cys, czs;  // cell-brick row / slab strides in cells
  uint32_t crows;     // cell-brick rows per slab (czs = crows * cys)
  float srecip[3]; // 0.5/spacing  (V/vtkCudaVolumeInformationHandler.cxx:91-93)
  float sp[3];     // spacing
  float bounds[6]; // [0, dim-1] per axis
};

struct ViewGeom {
  float m[16];
  int W, H, flipped;
  int n_clip, n_key;
  float clip[24];
  float key[24];
  const float* zbuf;
  int zbw, zbh;
};

struct MarchArgs {
  float step_len;  // MinSpacing * step_scale
  float term_T;
  int jitter;
};

template <typename T> struct V3 { T x, y, z; };

__device__ __forceinline__ float satf(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

// __fsqrt_ru / __fsqrt_rz from the correctly rounded sqrtf and the sign of the
// exact FMA residual.
__device__ __forceinline__ float sqrt_ru(float x) {
  float r = sqrtf(x);
  if (fmaf(r, r, -x) < 0.0f) r = nextafterf(r, __builtin_huge_valf());
  return r;
}
__device__ __forceinline__ float sqrt_rz(float x) {
  float r = sqrtf(x);
  if (fmaf(r, r, -x) > 0.0f) r = nextafterf(r, 0.0f);
  return r;
}

// Steps a ray takes: floor(num_steps - offset) as __float2int_rd gives it for
// any real ray.  Non-finite or absurd counts (>= 2^24 steps) march nothing,
// so no input geometry can keep a wave looping without bound.
__device__ __forceinline__ int step_count(float v) { return (v > 0.0f && v < 16777216.0f) ? (int)floorf(v) : 0; }

// CUDA float->uchar conversion: truncate toward zero, saturate, NaN -> 0.
__device__ __forceinline__ uint32_t to_u8(float v) {
  if (!(v > 0.0f)) return 0u;
  if (v >= 255.0f) return 255u;
  return (uint32_t)(int)v;
}

// ---------------------------------------------------------------------------
// Cell layout.  A trilinear reconstruction reads the 2x2x2 voxels of one cell,
// so every cell is stored whole: cell (cx,cy,cz) holds the 8 voxels
// v[clamp(cx-1+dx)][clamp(cy-1+dy)][clamp(cz-1+dz)], dx,dy,dz in {0,1}, in
// the order dx + 2 dy + 4 dz, as one 8/16/32-byte vector (u8/16-bit/32-bit
// voxels).  Cell indices run over [0, n+1] per axis (the sample coordinate is
// clamped to [-1, n] first), so edge clamping is baked into the cells and a
// sample costs one vector load and no index clamps.  The volume grows 8x.
//
// Cell order: x-fastest (brick edge 1) or 4x4x4-cell bricks, chosen per
// volume; both are kernel template instances.  Plain order suits rays that
// run along the x rows and dense sampling (c2: 0.83 ms against 0.92 with
// bricks); bricks suit large volumes seen from arbitrary directions (c3:
// 6.6 ms against 8.2).
template <int LOG2> struct CellOrder {
  static constexpr int kLog2 = LOG2;
  static constexpr int kEdge = 1 << LOG2;
  static constexpr int kCells = kEdge * kEdge * kEdge;
  static __device__ __forceinline__ uint32_t x(int c) {
    return (uint32_t)(((c >> LOG2) * kCells) | (c & (kEdge - 1)));
  }
  // The row and slab terms use the full-rate 24-bit multiply: the slab
  // stride is < 2^24 (checked when a volume is created), the brick row/slab
  // index < 2^13, and the slab term stays below the cell count (< 2^32), so
  // the low 32 bits of the 24-bit products are exact.
  static __device__ __forceinline__ uint32_t y(int c, uint32_t cys) {
    return __umul24((uint32_t)(c >> LOG2), cys) + (uint32_t)((c & (kEdge - 1)) << LOG2);
  }
  static __device__ __forceinline__ uint32_t z(int c, uint32_t czs) {
    return __umul24((uint32_t)(c >> LOG2), czs) + (uint32_t)((c & (kEdge - 1)) << (2 * LOG2));
  }
  static __device__ __forceinline__ uint32_t index(const VolGeom& g, int cx, int cy, int cz) {
    if constexpr (LOG2 == 0) {
      // (cz * rows + cy) * cys + cx: two 24-bit multiply-adds (the row count
      // of a volume, (ny + 2)(nz + 2), is below 2^24 like the slab stride)
      return __umul24(__umul24((uint32_t)cz, g.crows) + (uint32_t)cy, g.cys) + (uint32_t)cx;
    } else {
      // x(cx) + y(cy) + z(cz) with c & (E - 1) = c - E (c >> L) regrouped:
      // cx + E cy + E^2 cz + (cx >> L)(E^3 - E) + (cy >> L)(cys - E^2)
      // + (cz >> L)(czs - E^3): shifts and multiply-adds instead of masks
      // and ors.  Every term is >= 0 (cys, czs >= E^3) and
      // the sum is the cell index (< 2^32), so the 24-bit products and the
      // 32-bit partial sums are exact.
      constexpr uint32_t E = kEdge;
      const uint32_t s = ((uint32_t)cz << (2 * LOG2)) + ((uint32_t)cy << LOG2) + (uint32_t)cx;
      return __umul24((uint32_t)(cz >> LOG2), g.czs - E * E * E) + __umul24((uint32_t)(cy >> LOG2), g.cys - E * E) +
             __umul24((uint32_t)(cx >> LOG2), E * E * E - E) + s;
    }
  }
};
constexpr int kCellLayoutPlain = 0;
constexpr int kCellLayoutBrick4 = 2;

// X-pair bricks (layout 3): the cells of an 8x8x8 cell brick are not stored
// whole; the brick keeps every x-pair (v[x], v[x+1]) its cells use once, as
// 9 z-slabs x 9 y-rows x 8 pairs (the 9 voxel rows 8b-1 .. 8b+7 a brick's
// cells reach per axis, clamped at the volume edges).  A cell's 4 pairs are
// then pairs p, p + 8 (next row), p + 72 (next slab) and p + 80 of its
// brick: one address and four loads with constant offsets, and 2.53x the
// native voxel bytes instead of the 8x of whole cells.
constexpr int kCellLayoutPair8 = 3;
constexpr int kPairRow = 8;                // pairs per brick row
constexpr int kPairSlab = 9 * kPairRow;    // pairs per brick slab
constexpr int kPairBrick = 9 * kPairSlab;  // 648 pairs per brick

template <> struct CellOrder<kCellLayoutPair8> {
  static constexpr int kLog2 = kCellLayoutPair8;
  // index() returns the cell's first pair (pair units: 2 voxels)
  static __device__ __forceinline__ uint32_t index(const VolGeom& g, int cx, int cy, int cz) {
    // brick (cx >> 3, cy >> 3, cz >> 3) of a grid cys x crows x ..., then
    // slab, row and pair inside it; bricks and in-brick offsets stay < 2^24
    const uint32_t b = __umul24(__umul24((uint32_t)(cz >> 3), g.crows) + (uint32_t)(cy >> 3), g.cys) +
                       (uint32_t)(cx >> 3);
    const uint32_t in = __umul24((uint32_t)(cz & 7), (uint32_t)kPairSlab) + (uint32_t)((cy & 7) * kPairRow) +
                        (uint32_t)(cx & 7);
    return __umul24(b, (uint32_t)kPairBrick) + in;
  }
};
// Voxels per unit of CellOrder<L>::index(): a whole cell (8) or a pair (2).
template <int L> struct CellStride { static constexpr int k = 8; };
template <> struct CellStride<kCellLayoutPair8> { static constexpr int k = 2; };

// Range bricks: kRangeEdge^3 cells (cells [8b, 8b + 8) per axis, i.e. the
// voxels [8b - 1, 8b + 7] their corners reach).
constexpr int kRangeLog2 = 3;
constexpr int kRangeEdge = 1 << kRangeLog2;

// A cell volume of voxel type T in cell order LOG2 (what the kernels are
// templated on; `d` is the device array of cells).  DIFF cells store each
// x-pair as (a, b - a) instead of (a, b): the first lerp level then needs no
// subtraction.  The difference is exact as stored (16-bit frames whose value
// range is below 32768; f32 stores the rounded difference the lerp would
// compute), so values are bit-identical to those of plain cells.
template <typename T, int LOG2, bool DIFF = false> struct Cells {
  using value_type = T;
  using order = CellOrder<LOG2>;
  static constexpr int kLog2 = LOG2;
  static constexpr bool kDiff = DIFF;
  const T* __restrict__ d;
};

// A cell's raw bits (`Raw`, what one vector load returns) and their 8 slots
// as floats (`unpack`, in the order dx + 2 dy + 4 dz).  `unpack_ad` gives the
// four x-pairs as a = v(dx=0) and d = v(dx=1) - v(dx=0) for either encoding.
template <typename VT> struct CellLoad;
// raw_pairs(d, p) assembles the same Raw from the four x-pairs of an x-pair
// brick cell (pairs p, p + 8, p + 72, p + 80).
template <typename P> __device__ __forceinline__ P pair_at(const void* d, uint32_t p) {
  return reinterpret_cast<const P*>(d)[p];
}
template <> struct CellLoad<uint8_t> {
  using Raw = uint2;
  static __device__ __forceinline__ Raw raw(const uint8_t* __restrict__ d, uint32_t c) {
    return *reinterpret_cast<const uint2*>(d + (size_t)c * 8);
  }
  static __device__ __forceinline__ Raw raw_pairs(const uint8_t* __restrict__ d, uint32_t p) {
    const uint32_t a = pair_at<uint16_t>(d, p), b = pair_at<uint16_t>(d, p + kPairRow);
    const uint32_t c = pair_at<uint16_t>(d, p + kPairSlab), e = pair_at<uint16_t>(d, p + kPairSlab + kPairRow);
    return Raw{a | (b << 16), c | (e << 16)};
  }
  static __device__ __forceinline__ void unpack(Raw w, float v[8]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i] = (float)((w.x >> (8 * i)) & 0xffu);
      v[4 + i] = (float)((w.y >> (8 * i)) & 0xffu);
    }
  }  template <bool DIFF> static __device__ __forceinline__ void unpack_ad(Raw w, float a[4], float d[4]) {
    static_assert(!DIFF, "8-bit cells are not difference-encoded");
    float v[8];
    unpack(w, v);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = v[2 * i];
      d[i] = v[2 * i + 1] - v[2 * i];
    }
  }
};
template <> struct CellLoad<int8_t> {
  using Raw = int2;
  static __device__ __forceinline__ Raw raw(const int8_t* __restrict__ d, uint32_t c) {
    return *reinterpret_cast<const int2*>(d + (size_t)c * 8);
  }
  static __device__ __forceinline__ Raw raw_pairs(const int8_t* __restrict__ d, uint32_t p) {
    const uint32_t a = pair_at<uint16_t>(d, p), b = pair_at<uint16_t>(d, p + kPairRow);
    const uint32_t c = pair_at<uint16_t>(d, p + kPairSlab), e = pair_at<uint16_t>(d, p + kPairSlab + kPairRow);
    return Raw{(int)(a | (b << 16)), (int)(c | (e << 16))};
  }
  static __device__ __forceinline__ void unpack(Raw w, float v[8]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i] = (float)((int)((uint32_t)w.x << (24 - 8 * i)) >> 24);
      v[4 + i] = (float)((int)((uint32_t)w.y << (24 - 8 * i)) >> 24);
    }
  }  template <bool DIFF> static __device__ __forceinline__ void unpack_ad(Raw w, float a[4], float d[4]) {
    static_assert(!DIFF, "8-bit cells are not difference-encoded");
    float v[8];
    unpack(w, v);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = v[2 * i];
      d[i] = v[2 * i + 1] - v[2 * i];
    }
  }
};
// 16-bit voxels: a pair is one 32-bit word.
__device__ __forceinline__ uint4 raw_pair_words(const void* d, uint32_t p) {
  return uint4{pair_at<uint32_t>(d, p), pair_at<uint32_t>(d, p + kPairRow), pair_at<uint32_t>(d, p + kPairSlab),
               pair_at<uint32_t>(d, p + kPairSlab + kPairRow)};
}
template <> struct CellLoad<uint16_t> {
  using Raw = uint4;
  static __device__ __forceinline__ Raw raw(const uint16_t* __restrict__ d, uint32_t c) {
    return *reinterpret_cast<const uint4*>(d + (size_t)c * 8);
  }
  static __device__ __forceinline__ Raw raw_pairs(const uint16_t* __restrict__ d, uint32_t p) {
    return raw_pair_words(d, p);
  }
  static __device__ __forceinline__ void unpack(Raw w, float v[8]) {
    const uint32_t q[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[2 * i] = (float)(q[i] & 0xffffu);
      v[2 * i + 1] = (float)(q[i] >> 16);
    }
  }
  template <bool DIFF> static __device__ __forceinline__ void unpack_ad(Raw w, float a[4], float d[4]) {
    const uint32_t q[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = (float)(q[i] & 0xffffu);
      // a DIFF slot holds the signed 16-bit difference
      d[i] = DIFF ? (float)((int)q[i] >> 16) : (float)(q[i] >> 16) - a[i];
    }
  }
};
template <> struct CellLoad<int16_t> {
  using Raw = int4;
  static __device__ __forceinline__ Raw raw(const int16_t* __restrict__ d, uint32_t c) {
    return *reinterpret_cast<const int4*>(d + (size_t)c * 8);
  }
  static __device__ __forceinline__ Raw raw_pairs(const int16_t* __restrict__ d, uint32_t p) {
    const uint4 w = raw_pair_words(d, p);
    return Raw{(int)w.x, (int)w.y, (int)w.z, (int)w.w};
  }
  static __device__ __forceinline__ void unpack(Raw w, float v[8]) {
    const int q[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[2 * i] = (float)((int)((uint32_t)q[i] << 16) >> 16);
      v[2 * i + 1] = (float)(q[i] >> 16);
    }
  }
  template <bool DIFF> static __device__ __forceinline__ void unpack_ad(Raw w, float a[4], float d[4]) {
    const int q[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = (float)((int)((uint32_t)q[i] << 16) >> 16);
      d[i] = DIFF ? (float)(q[i] >> 16) : (float)(q[i] >> 16) - a[i];
    }
  }
};
template <typename W, typename V4> struct CellLoad32 {
  struct Raw {
    V4 a, b;
  };
  static __device__ __forceinline__ Raw raw(const W* __restrict__ d, uint32_t c) {
    const V4* p = reinterpret_cast<const V4*>(d + (size_t)c * 8);
    return Raw{p[0], p[1]};
  }
  static __device__ __forceinline__ Raw raw_pairs(const W* __restrict__ d, uint32_t p) {
    const uint2 a = pair_at<uint2>(d, p), b = pair_at<uint2>(d, p + kPairRow);
    const uint2 c = pair_at<uint2>(d, p + kPairSlab), e = pair_at<uint2>(d, p + kPairSlab + kPairRow);
    using E = decltype(V4{}.x);
    return Raw{V4{__builtin_bit_cast(E, a.x), __builtin_bit_cast(E, a.y), __builtin_bit_cast(E, b.x),
                  __builtin_bit_cast(E, b.y)},
               V4{__builtin_bit_cast(E, c.x), __builtin_bit_cast(E, c.y), __builtin_bit_cast(E, e.x),
                  __builtin_bit_cast(E, e.y)}};
  }
  static __device__ __forceinline__ void unpack(const Raw& w, float v[8]) {
    v[0] = (float)w.a.x; v[1] = (float)w.a.y; v[2] = (float)w.a.z; v[3] = (float)w.a.w;
    v[4] = (float)w.b.x; v[5] = (float)w.b.y; v[6] = (float)w.b.z; v[7] = (float)w.b.w;
  }
  template <bool DIFF> static __device__ __forceinline__ void unpack_ad(const Raw& w, float a[4], float d[4]) {
    float v[8];
    unpack(w, v);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = v[2 * i];
      d[i] = DIFF ? v[2 * i + 1] : v[2 * i + 1] - v[2 * i];
    }
  }
};
template <> struct CellLoad<float> : CellLoad32<float, float4> {};

// 16-bit frames whose values all lie in [-2048, 2048] store each x-pair's
// base as an IEEE binary16 (exact for those integers) and its difference as
// int16 (always difference-encoded).  The first lerp level then converts the
// base inside its fma (v_fma_mix_f32) instead of with a separate convert;
// the values are the same as those of int16 or uint16 cells.
struct Half16 {
  uint16_t bits;
  __device__ __forceinline__ operator float() const { return (float)__builtin_bit_cast(_Float16, bits); }
};
template <> struct CellLoad<Half16> {
  using Raw = uint4;
  static __device__ __forceinline__ Raw raw(const Half16* __restrict__ d, uint32_t c) {
    return *reinterpret_cast<const uint4*>(d + (size_t)c * 8);
  }
  static __device__ __forceinline__ Raw raw_pairs(const Half16* __restrict__ d, uint32_t p) {
    return raw_pair_words(d, p);
  }
  static __device__ __forceinline__ float base(uint32_t q) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(q & 0xffffu));
  }
  static __device__ __forceinline__ void unpack(Raw w, float v[8]) {
    const uint32_t q[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[2 * i] = base(q[i]);
      v[2 * i + 1] = v[2 * i] + (float)((int)q[i] >> 16);
    }
  }
  template <bool DIFF> static __device__ __forceinline__ void unpack_ad(Raw w, float a[4], float d[4]) {
    static_assert(DIFF, "binary16-base cells are difference-encoded");
    const uint32_t q[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = base(q[i]);
      d[i] = (float)((int)q[i] >> 16);
    }
  }
};
// When, in addition, every x-pair difference of the frame is an integer that
// binary16 holds exactly (|b - a| <= 2048, or even up to 4096), the
// difference is stored as binary16 too, and the first lerp level is one
// v_fma_mix_f32 reading both halves of the word.  Same values again.
struct HalfPair {
  uint16_t bits;
  __device__ __forceinline__ operator float() const { return (float)__builtin_bit_cast(_Float16, bits); }
};
template <> struct CellLoad<HalfPair> {
  using Raw = uint4;
  static __device__ __forceinline__ Raw raw(const HalfPair* __restrict__ d, uint32_t c) {
    return *reinterpret_cast<const uint4*>(d + (size_t)c * 8);
  }
  static __device__ __forceinline__ Raw raw_pairs(const HalfPair* __restrict__ d, uint32_t p) {
    return raw_pair_words(d, p);
  }
  static __device__ __forceinline__ float lo(uint32_t q) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(q & 0xffffu));
  }
  static __device__ __forceinline__ float hi(uint32_t q) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(q >> 16)); }
  static __device__ __forceinline__ void unpack(Raw w, float v[8]) {
    const uint32_t q[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[2 * i] = lo(q[i]);
      v[2 * i + 1] = v[2 * i] + hi(q[i]);
    }
  }
  template <bool DIFF> static __device__ __forceinline__ void unpack_ad(Raw w, float a[4], float d[4]) {
    static_assert(DIFF, "binary16-pair cells are difference-encoded");
    const uint32_t q[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = lo(q[i]);
      d[i] = hi(q[i]);
    }
  }
};
template <> struct CellLoad<int32_t> : CellLoad32<int32_t, int4> {};
template <> struct CellLoad<uint32_t> : CellLoad32<uint32_t, uint4> {};

template <typename CV> __device__ __forceinline__ CV cells_of(const void* p) {
  return CV{static_cast<const typename CV::value_type*>(p)};
}

// The raw bits of the cell at CellOrder<L>::index() c.
template <typename T, int L>
__device__ __forceinline__ typename CellLoad<T>::Raw fetch_cell(const T* __restrict__ d, uint32_t c) {
  if constexpr (L == kCellLayoutPair8) return CellLoad<T>::raw_pairs(d, c);
  else return CellLoad<T>::raw(d, c);
}

// Voxel (i,j,k) (in range) = corner 000 of cell (i+1, j+1, k+1).
template <typename T, int L, bool D>
__device__ __forceinline__ float load_voxel(Cells<T, L, D> v, const VolGeom& g, int i, int j, int k) {
  return (float)v.d[(size_t)CellOrder<L>::index(g, i + 1, j + 1, k + 1) * CellStride<L>::k];
}

// Cell and weights of tex3D(unnormalised, linear, clamp) at p, given the
// cell coordinate c = p + 0.5.  Texel centres sit at i + 0.5, so the
// reconstruction at p interpolates voxels floor(p - 0.5) and the next one
// along each axis: the corners of cell floor(c), since cell i holds voxels
// i - 1 and i (clamped).  Clamping c to [0, n + 1] leaves every
// clamp-addressed result unchanged; on c >= 0 truncation is floor and the
// fraction c - floor(c) is exact, so the cell is one v_cvt_i32_f32 and the
// weight one v_fract_f32 per axis (sample coordinates are finite: rays with
// non-finite steps take no samples).
struct CellPos {
  uint32_t cell;
  float ax, ay, az;
};
// CLAMP = false when the caller has checked that c lies in [0, n + 1].
template <int L, bool CLAMP = true>
__device__ __forceinline__ CellPos cell_pos(const VolGeom& g, float cx, float cy, float cz) {
  const float x = CLAMP ? __builtin_amdgcn_fmed3f(cx, 0.0f, (float)(g.n[0] + 1)) : cx;
  const float y = CLAMP ? __builtin_amdgcn_fmed3f(cy, 0.0f, (float)(g.n[1] + 1)) : cy;
  const float z = CLAMP ? __builtin_amdgcn_fmed3f(cz, 0.0f, (float)(g.n[2] + 1)) : cz;
  CellPos c;
  c.ax = __builtin_amdgcn_fractf(x);
  c.ay = __builtin_amdgcn_fractf(y);
  c.az = __builtin_amdgcn_fractf(z);
  const int ix = (int)x, iy = (int)y, iz = (int)z;
  c.cell = CellOrder<L>::index(g, ix, iy, iz);
  return c;
}

// a + t (b - a) as one rounding of the product-sum
__device__ __forceinline__ float lerp(float a, float b, float t) { return fmaf(t, b - a, a); }

// Trilinear value at the cell coordinate c = p + 0.5 (CLAMP as for cell_pos).
template <bool CLAMP = true, typename T, int L, bool D>
__device__ __forceinline__ float trilinear(Cells<T, L, D> vol, const VolGeom& g, float cx, float cy, float cz) {
  const CellPos c = cell_pos<L, CLAMP>(g, cx, cy, cz);
  float a[4], d[4];
  CellLoad<T>::template unpack_ad<D>(fetch_cell<T, L>(vol.d, c.cell), a, d);
  // lerp(a, b, t) = fma(t, b - a, a) with b - a = d
  const float c00 = fmaf(c.ax, d[0], a[0]), c10 = fmaf(c.ax, d[1], a[1]);
  const float c01 = fmaf(c.ax, d[2], a[2]), c11 = fmaf(c.ax, d[3], a[3]);
  const float c0 = lerp(c00, c10, c.ay), c1 = lerp(c01, c11, c.ay);
  return lerp(c0, c1, c.az);
}

// trilinear() with the cell's raw bits kept in registers: consecutive
// samples of a ray (half a voxel apart) mostly fall in the cell of the one
// before, and then no load is issued (only the lanes that moved to another
// cell load, so the L1 sees fewer accesses).  Same operations as trilinear().
template <typename T, int L> struct CellCache {
  uint32_t idx = 0xffffffffu;
  typename CellLoad<T>::Raw raw{};
  template <bool CLAMP = true, bool D>
  __device__ __forceinline__ float sample(Cells<T, L, D> vol, const VolGeom& g, float cx, float cy, float cz) {
    const CellPos c = cell_pos<L, CLAMP>(g, cx, cy, cz);
    if (c.cell != idx) {
      raw = fetch_cell<T, L>(vol.d, c.cell);
      idx = c.cell;
    }
    float a[4], d[4];
    CellLoad<T>::template unpack_ad<D>(raw, a, d);
    const float c00 = fmaf(c.ax, d[0], a[0]), c10 = fmaf(c.ax, d[1], a[1]);
    const float c01 = fmaf(c.ax, d[2], a[2]), c11 = fmaf(c.ax, d[3], a[3]);
    const float c0 = lerp(c00, c10, c.ay), c1 = lerp(c01, c11, c.ay);
    return lerp(c0, c1, c.az);
  }
};

// Trilinear value and its gradient with respect to p (voxel index units) on
// the same clamped cell, at the cell coordinate c = p + 0.5.  The value is
// computed with exactly the operations of trilinear(), so branch decisions
// made on it match the plain caster.
// d/dx = lerp_z(lerp_y(v100-v000, v110-v010), lerp_y(v101-v001, v111-v011)),
// d/dy = lerp_z(c10-c00, c11-c01), d/dz = c1-c0.  Along an axis whose two
// corners clamp to one voxel (the edge cells 0, n and n + 1, where every
// clamped coordinate also lands) the corners are equal, so these
// differences, and the derivative, are exactly +0 with no test.
template <bool CLAMP = true, typename T, int L, bool D>
__device__ __forceinline__ float trilinear_vg(Cells<T, L, D> vol, const VolGeom& g, float cx, float cy, float cz,
                                              V3<float>& gr) {
  const CellPos c = cell_pos<L, CLAMP>(g, cx, cy, cz);
  float a[4], d[4];
  CellLoad<T>::template unpack_ad<D>(fetch_cell<T, L>(vol.d, c.cell), a, d);
  const float c00 = fmaf(c.ax, d[0], a[0]), c10 = fmaf(c.ax, d[1], a[1]);
  const float c01 = fmaf(c.ax, d[2], a[2]), c11 = fmaf(c.ax, d[3], a[3]);
  const float c0 = lerp(c00, c10, c.ay), c1 = lerp(c01, c11, c.ay);
  const float dx0 = lerp(d[0], d[1], c.ay), dx1 = lerp(d[2], d[3], c.ay);
  gr.x = lerp(dx0, dx1, c.az);
  gr.y = lerp(c10 - c00, c11 - c01, c.az);
  gr.z = c1 - c0;
  return lerp(c0, c1, c.az);
}

// tex1D(normalised, linear, clamp) on an n-entry table.
__device__ __forceinline__ float tf1d_lookup(const float* t, int n, float u) {
  float x = u * (float)n - 0.5f;
  x = fminf(fmaxf(x, -1.0f), (float)n);
  const float fx = floorf(x);
  const float a = x - fx;
  const int i = (int)fx;
  const int i0 = min(max(i, 0), n - 1), i1 = min(max(i + 1, 0), n - 1);
  return lerp(t[i0], t[i1], a);
}

// tex2D(normalised, point, clamp) cell index g*n + i.
__device__ __forceinline__ int tf2d_axis(int n, float u) {
  const float x = fminf(fmaxf(u * (float)n, -1.0f), (float)n);
  return min(max((int)floorf(x), 0), n - 1);
}
__device__ __forceinline__ int tf2d_index(int n, float u, float v) { return tf2d_axis(n, v) * n + tf2d_axis(n, u); }
// Whether column i of a TF's column bitmap has any opacity.
__device__ __forceinline__ bool tf2d_column_live(const uint32_t* col, int i) { return (col[i >> 5] >> (i & 31)) & 1u; }

// ---------------------------------------------------------------------------
// Ray setup (formRays, fused into every caster's prologue).
struct Ray {
  V3<float> start, inc;
  float num_steps;
  float exclude_start, exclude_end;
  float min_depth, max_depth;
};

// V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:130-197
__device__ __forceinline__ void clip_planes(const ViewGeom& v, V3<float>& s, V3<float>& e, V3<float>& d) {
  const int np = v.n_clip;
  if (!np) return;
  int flag = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    if (i >= np) break;
    d.x = e.x - s.x; d.y = e.y - s.y; d.z = e.z - s.z;
    const float px = v.clip[4 * i], py = v.clip[4 * i + 1], pz = v.clip[4 * i + 2], pw = v.clip[4 * i + 3];
    const float dp = px * d.x + py * d.y + pz * d.z;
    const float t = -(px * s.x + py * s.y + pz * s.z + pw) / dp;
    const float p0 = s.x + t * d.x, p1 = s.y + t * d.y, p2 = s.z + t * d.z;
    if (t > 0.0f && t < 1.0f) {
      if (dp > 0.0f) { s.x = p0; s.y = p1; s.z = p2; }
      else { e.x = p0; e.y = p1; e.z = p2; }
    }
    flag |= (dp > 0.0f && t > 1.0f);
    flag |= (dp < 0.0f && t < 0.0f);
  }
  if (flag) s = e;
}

// V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:199-318
__device__ __forceinline__ void clip_volume(const VolGeom& g, V3<float>& s, V3<float>& e, V3<float>& d) {
  d.x = e.x - s.x; d.y = e.y - s.y; d.z = e.z - s.z;
  const float b0 = g.bounds[0] + 1.0f, b1 = g.bounds[1] - 1.0f;
  const float b2 = g.bounds[2] + 1.0f, b3 = g.bounds[3] - 1.0f;
  const float b4 = g.bounds[4] + 1.0f, b5 = g.bounds[5] - 1.0f;
  float ds, de;
#define RVK_CROP()                                                   \
  if (isfinite(ds)) {                                                \
    s.x += d.x * ds; s.y += d.y * ds; s.z += d.z * ds;               \
    e.x += d.x * de; e.y += d.y * de; e.z += d.z * de;               \
  }
  if (d.x > 0.0f) { ds = s.x < b0 ? b0 - s.x : 0.0f; de = e.x > b1 ? b1 - e.x : 0.0f; }
  else { ds = s.x > b1 ? b1 - s.x : 0.0f; de = e.x < b0 ? b0 - e.x : 0.0f; }
  ds /= d.x; de /= d.x;
  RVK_CROP()
  if (d.y > 0.0f) { ds = s.y < b2 ? b2 - s.y : 0.0f; de = e.y > b3 ? b3 - e.y : 0.0f; }
  else { ds = s.y > b3 ? b3 - s.y : 0.0f; de = e.y < b2 ? b2 - e.y : 0.0f; }
  ds /= d.y; de /= d.y;
  RVK_CROP()
  if (d.z > 0.0f) { ds = s.z < b4 ? b4 - s.z : 0.0f; de = e.z > b5 ? b5 - e.z : 0.0f; }
  else { ds = s.z > b5 ? b5 - s.z : 0.0f; de = e.z < b4 ? b4 - e.z : 0.0f; }
  ds /= d.z; de /= d.z;
  RVK_CROP()
#undef RVK_CROP
  if (e.x > b1 + 1.0f || e.y > b3 + 1.0f || e.z > b5 + 1.0f || e.x < b0 - 1.0f || e.y < b2 - 1.0f ||
      e.z < b4 - 1.0f || s.x > b1 + 1.0f || s.y > b3 + 1.0f || s.z > b5 + 1.0f || s.x < b0 - 1.0f ||
      s.y < b2 - 1.0f || s.z < b4 - 1.0f) {
    s = e;
  }
  d.x = e.x - s.x; d.y = e.y - s.y; d.z = e.z - s.z;
}

// V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:38-128
__device__ __forceinline__ void keyhole_values(const ViewGeom& v, V3<float> s, const V3<float>& inc, float ns,
                                               float& es, float& ee) {
  es = 1.0f;
  ee = -1.0f;
  const int np = v.n_key;
  if (!np) return;
  const V3<float> old = s;
  V3<float> d{ns * inc.x, ns * inc.y, ns * inc.z};
  V3<float> e{s.x + d.x, s.y + d.y, s.z + d.z};
  int flag = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    if (i >= np) break;
    d.x = e.x - s.x; d.y = e.y - s.y; d.z = e.z - s.z;
    const float px = v.key[4 * i], py = v.key[4 * i + 1], pz = v.key[4 * i + 2], pw = v.key[4 * i + 3];
    const float dp = px * d.x + py * d.y + pz * d.z;
    const float t = -(px * s.x + py * s.y + pz * s.z + pw) / dp;
    const float p0 = s.x + t * d.x, p1 = s.y + t * d.y, p2 = s.z + t * d.z;
    if (t > 0.0f && t < 1.0f) {
      if (dp > 0.0f) { s.x = p0; s.y = p1; s.z = p2; }
      else { e.x = p0; e.y = p1; e.z = p2; }
    }
    flag |= (dp > 0.0f && t > 1.0f);
    flag |= (dp < 0.0f && t < 0.0f);
  }
  s.x -= old.x; s.y -= old.y; s.z -= old.z;
  e.x -= old.x; e.y -= old.y; e.z -= old.z;
  const float inv = 1.0f / (inc.x * inc.x + inc.y * inc.y + inc.z * inc.z);
  es = flag ? -1.0f : (s.x * inc.x + s.y * inc.y + s.z * inc.z) * inv;
  ee = flag ? -1.0f : (e.x * inc.x + e.y * inc.y + e.z * inc.z) * inv;
}

// V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:320-447
__device__ __forceinline__ Ray form_ray(const VolGeom& g, const ViewGeom& v, float step_len, int x, int y) {
  Ray r;
  const float vx = v.flipped ? (float)x / (float)v.W : 1.0f - (float)x / (float)v.W;
  const float vy = (float)y / (float)v.H;
  float end_depth = 1.0f;
  if (v.zbuf) {
    const float u = 1.0f - vx;
    int zi = (int)floorf(fminf(fmaxf(u * (float)v.zbw, -1.0f), (float)v.zbw));
    int zj = (int)floorf(fminf(fmaxf(vy * (float)v.zbh, -1.0f), (float)v.zbh));
    zi = min(max(zi, 0), v.zbw - 1);
    zj = min(max(zj, 0), v.zbh - 1);
    end_depth = v.zbuf[(long)zj * v.zbw + zi];
  }
  const float* m = v.m;
  V3<float> s, e, f;
  s.x = vx * m[0] + vy * m[1] + m[3];
  s.y = vx * m[4] + vy * m[5] + m[7];
  s.z = vx * m[8] + vy * m[9] + m[11];
  const float sn = vx * m[12] + vy * m[13] + m[15];
  e.x = s.x + end_depth * m[2];
  e.y = s.y + end_depth * m[6];
  e.z = s.z + end_depth * m[10];
  const float en = sn + end_depth * m[14];
  f.x = s.x + m[2];
  f.y = s.y + m[6];
  f.z = s.z + m[10];
  const float fn = sn + m[14];
  s.x /= sn; s.y /= sn; s.z /= sn;
  e.x /= en; e.y /= en; e.z /= en;
  f.x /= fn; f.y /= fn; f.z /= fn;
  const V3<float> old = s;
  V3<float> d{f.x - s.x, f.y - s.y, f.z - s.z};
  r.max_depth = sqrt_rz(d.x * d.x + d.y * d.y + d.z * d.z);
  clip_planes(v, s, e, d);
  clip_volume(g, s, e, d);
  const float len = sqrt_rz(d.x * d.x + d.y * d.y + d.z * d.z);
  const float md = sqrt_rz((s.x - old.x) * (s.x - old.x) + (s.y - old.y) * (s.y - old.y) +
                           (s.z - old.z) * (s.z - old.z));
  r.min_depth = (len > 0.0f) ? md : r.max_depth;
  r.num_steps = sqrt_ru(d.x * d.x * g.sp[0] * g.sp[0] + d.y * d.y * g.sp[1] * g.sp[1] +
                        d.z * d.z * g.sp[2] * g.sp[2]) / step_len;
  r.inc.x = d.x / r.num_steps;
  r.inc.y = d.y / r.num_steps;
  r.inc.z = d.z / r.num_steps;
  r.start = s;
  keyhole_values(v, s, r.inc, r.num_steps, r.exclude_start, r.exclude_end);
  return r;
}

// ---------------------------------------------------------------------------
// Forward-mode tangent of the ray setup: given dM (the derivative of the
// baked matrix along one pose direction), the derivatives of the clipped
// start point and of the step vector.  Values follow form_ray operation for
// operation, so every branch (clip face, reject) is the one the forward took.
struct Dl {
  float v, d;
};
__device__ __forceinline__ Dl operator+(Dl a, Dl b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dl operator-(Dl a, Dl b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dl operator*(Dl a, Dl b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dl operator*(float a, Dl b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dl operator/(Dl a, Dl b) {
  const float q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dl dconst(float c) { return {c, 0.0f}; }

struct RayTangent {
  V3<float> dstart, dinc;
};

__device__ __forceinline__ RayTangent form_ray_tangent(const VolGeom& g, const ViewGeom& v, const float* dm,
                                                       float step_len, int x, int y) {
  const float vx = v.flipped ? (float)x / (float)v.W : 1.0f - (float)x / (float)v.W;
  const float vy = (float)y / (float)v.H;
  float end_depth = 1.0f;
  if (v.zbuf) {
    const float u = 1.0f - vx;
    int zi = (int)floorf(fminf(fmaxf(u * (float)v.zbw, -1.0f), (float)v.zbw));
    int zj = (int)floorf(fminf(fmaxf(vy * (float)v.zbh, -1.0f), (float)v.zbh));
    zi = min(max(zi, 0), v.zbw - 1);
    zj = min(max(zj, 0), v.zbh - 1);
    end_depth = v.zbuf[(long)zj * v.zbw + zi];
  }
  Dl M[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) M[i] = Dl{v.m[i], dm[i]};
  V3<Dl> s, e, f;
  s.x = vx * M[0] + vy * M[1] + M[3];
  s.y = vx * M[4] + vy * M[5] + M[7];
  s.z = vx * M[8] + vy * M[9] + M[11];
  const Dl sn = vx * M[12] + vy * M[13] + M[15];
  e.x = s.x + end_depth * M[2];
  e.y = s.y + end_depth * M[6];
  e.z = s.z + end_depth * M[10];
  const Dl en = sn + end_depth * M[14];
  f.x = s.x + M[2];
  f.y = s.y + M[6];
  f.z = s.z + M[10];
  const Dl fn = sn + M[14];
  s.x = s.x / sn; s.y = s.y / sn; s.z = s.z / sn;
  e.x = e.x / en; e.y = e.y / en; e.z = e.z / en;
  f.x = f.x / fn; f.y = f.y / fn; f.z = f.z / fn;
  V3<Dl> d{f.x - s.x, f.y - s.y, f.z - s.z};
  // clipping planes (constant planes)
  if (v.n_clip) {
    int flag = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      if (i >= v.n_clip) break;
      d.x = e.x - s.x; d.y = e.y - s.y; d.z = e.z - s.z;
      const float px = v.clip[4 * i], py = v.clip[4 * i + 1], pz = v.clip[4 * i + 2], pw = v.clip[4 * i + 3];
      const Dl dp = px * d.x + py * d.y + pz * d.z;
      const Dl num = px * s.x + py * s.y + pz * s.z + dconst(pw);
      const Dl t = Dl{-num.v, -num.d} / dp;
      const Dl p0 = s.x + t * d.x, p1 = s.y + t * d.y, p2 = s.z + t * d.z;
      if (t.v > 0.0f && t.v < 1.0f) {
        if (dp.v > 0.0f) { s.x = p0; s.y = p1; s.z = p2; }
        else { e.x = p0; e.y = p1; e.z = p2; }
      }
      flag |= (dp.v > 0.0f && t.v > 1.0f);
      flag |= (dp.v < 0.0f && t.v < 0.0f);
    }
    if (flag) s = e;
  }
  // volume box
  d.x = e.x - s.x; d.y = e.y - s.y; d.z = e.z - s.z;
  const float b0 = g.bounds[0] + 1.0f, b1 = g.bounds[1] - 1.0f;
  const float b2 = g.bounds[2] + 1.0f, b3 = g.bounds[3] - 1.0f;
  const float b4 = g.bounds[4] + 1.0f, b5 = g.bounds[5] - 1.0f;
  auto crop = [&](Dl ds, Dl de) {
    if (isfinite(ds.v)) {
      s.x = s.x + d.x * ds; s.y = s.y + d.y * ds; s.z = s.z + d.z * ds;
      e.x = e.x + d.x * de; e.y = e.y + d.y * de; e.z = e.z + d.z * de;
    }
  };
  Dl ds, de;
  if (d.x.v > 0.0f) { ds = s.x.v < b0 ? dconst(b0) - s.x : dconst(0.0f); de = e.x.v > b1 ? dconst(b1) - e.x : dconst(0.0f); }
  else { ds = s.x.v > b1 ? dconst(b1) - s.x : dconst(0.0f); de = e.x.v < b0 ? dconst(b0) - e.x : dconst(0.0f); }
  crop(ds / d.x, de / d.x);
  if (d.y.v > 0.0f) { ds = s.y.v < b2 ? dconst(b2) - s.y : dconst(0.0f); de = e.y.v > b3 ? dconst(b3) - e.y : dconst(0.0f); }
  else { ds = s.y.v > b3 ? dconst(b3) - s.y : dconst(0.0f); de = e.y.v < b2 ? dconst(b2) - e.y : dconst(0.0f); }
  crop(ds / d.y, de / d.y);
  if (d.z.v > 0.0f) { ds = s.z.v < b4 ? dconst(b4) - s.z : dconst(0.0f); de = e.z.v > b5 ? dconst(b5) - e.z : dconst(0.0f); }
  else { ds = s.z.v > b5 ? dconst(b5) - s.z : dconst(0.0f); de = e.z.v < b4 ? dconst(b4) - e.z : dconst(0.0f); }
  crop(ds / d.z, de / d.z);
  if (e.x.v > b1 + 1.0f || e.y.v > b3 + 1.0f || e.z.v > b5 + 1.0f || e.x.v < b0 - 1.0f || e.y.v < b2 - 1.0f ||
      e.z.v < b4 - 1.0f || s.x.v > b1 + 1.0f || s.y.v > b3 + 1.0f || s.z.v > b5 + 1.0f || s.x.v < b0 - 1.0f ||
      s.y.v < b2 - 1.0f || s.z.v < b4 - 1.0f) {
    s = e;
  }
  d.x = e.x - s.x; d.y = e.y - s.y; d.z = e.z - s.z;
  const Dl q = d.x * d.x * dconst(g.sp[0]) * dconst(g.sp[0]) + d.y * d.y * dconst(g.sp[1]) * dconst(g.sp[1]) +
               d.z * d.z * dconst(g.sp[2]) * dconst(g.sp[2]);
  const float rq = sqrt_ru(q.v);
  const Dl ns{rq / step_len, (rq > 0.0f ? q.d / (2.0f * rq) : 0.0f) / step_len};
  RayTangent t;
  t.dstart = V3<float>{s.x.d, s.y.d, s.z.d};
  const Dl ix = d.x / ns, iy = d.y / ns, iz = d.z / ns;
  t.dinc = V3<float>{ix.d, iy.d, iz.d};
  return t;
}

// ---------------------------------------------------------------------------
// Pixel tiling: a 256-thread workgroup covers a 16x16 pixel tile, each wave64
// an 8x8 quadrant.  Workgroups are dispatched round-robin over the 8 XCDs;
// the remap gives each XCD a contiguous run of tiles (whole tile rows), so
// rays sharing volume bricks share that XCD's L2.
constexpr int kNumXcd = 8;

__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  const int per = nblk / kNumXcd, rem = nblk % kNumXcd;
  const int xcd = bid % kNumXcd, local = bid / kNumXcd;
  return (xcd < rem) ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// 16x16-pixel workgroup tiles are ordered in super-tiles of 2^kSuperLog2 x
// 2^kSuperLog2 tiles (row-major over super-tiles, row-major inside), so the
// tiles an XCD runs at one time cover a compact image region (and volume
// footprint) rather than a few full-width rows.  `stiles_x` is the number of
// super-tile columns; tiles past the image edge return at once.
#ifndef RVK_SUPER_LOG2
#define RVK_SUPER_LOG2 2
#endif
constexpr int kSuperLog2 = RVK_SUPER_LOG2;
__device__ __forceinline__ void tile_pixel(int tile, int stiles_x, int& px, int& py) {
  const int t = threadIdx.x;
  const int w = t >> 6, l = t & 63;
  const int st = tile >> (2 * kSuperLog2), in = tile & ((1 << (2 * kSuperLog2)) - 1);
  const int tx = (st % stiles_x) * (1 << kSuperLog2) + (in & ((1 << kSuperLog2) - 1));
  const int ty = (st / stiles_x) * (1 << kSuperLog2) + (in >> kSuperLog2);
  px = tx * 16 + (w & 1) * 8 + (l & 7);
  py = ty * 16 + (w >> 1) * 8 + (l >> 3);
}

}  // namespace rvk
