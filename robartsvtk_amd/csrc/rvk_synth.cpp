// rvk_synth.cpp -- seeded synthetic volumes (host, multithreaded over z).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "rvk_synth.h"

namespace {

inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

struct Xoshiro256ss {
  uint64_t s[4];
  explicit Xoshiro256ss(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) v = splitmix64(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
};

// Counter-based per-voxel randoms: independent of traversal order.
inline double hash_uniform(uint64_t seed, uint64_t idx) {
  uint64_t s = seed ^ (idx * 0xd1b54a32d192ed03ULL);
  return (double)(splitmix64(s) >> 11) * 0x1.0p-53;
}
inline double hash_normal(uint64_t seed, uint64_t idx) {
  // Box-Muller on two counter-based uniforms
  const double u1 = 1.0 - hash_uniform(seed, 2 * idx);
  const double u2 = hash_uniform(seed, 2 * idx + 1);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

template <typename F> void parallel_z(int nz, F&& fn) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t)
    pool.emplace_back([&, t]() {
      for (int z = (int)t; z < nz; z += (int)nt) fn(z);
    });
  for (auto& th : pool) th.join();
}

struct Blob {
  double c[3], amp, inv2s2;
};

std::vector<Blob> make_blobs(Xoshiro256ss& rng, const int32_t dims[3], int n, double alo, double ahi, double slo,
                             double shi) {
  std::vector<Blob> b(n);
  for (auto& bl : b) {
    bl.amp = rng.uniform(alo, ahi);
    const double s = rng.uniform(slo, shi);
    bl.inv2s2 = 1.0 / (2.0 * s * s);
    for (int a = 0; a < 3; ++a) bl.c[a] = rng.uniform(0.15, 0.85) * dims[a];
  }
  return b;
}

inline double blob_sum(const std::vector<Blob>& b, double x, double y, double z) {
  double v = 0.0;
  for (const auto& bl : b) {
    const double dx = x - bl.c[0], dy = y - bl.c[1], dz = z - bl.c[2];
    v += bl.amp * std::exp(-(dx * dx + dy * dy + dz * dz) * bl.inv2s2);
  }
  return v;
}

// Improved noise (K. Perlin, "Improving Noise", SIGGRAPH 2002): quintic fade,
// gradients from the 12 cube-edge directions chosen by a hashed permutation.
struct Perlin {
  int p[512];
  explicit Perlin(Xoshiro256ss& rng) {
    int q[256];
    for (int i = 0; i < 256; ++i) q[i] = i;
    for (int i = 255; i > 0; --i) {
      const int j = (int)(rng.next() % (uint64_t)(i + 1));
      std::swap(q[i], q[j]);
    }
    for (int i = 0; i < 512; ++i) p[i] = q[i & 255];
  }
  static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
  static double lerp(double t, double a, double b) { return a + t * (b - a); }
  static double grad(int h, double x, double y, double z) {
    h &= 15;
    const double u = h < 8 ? x : y;
    const double v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
    return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
  }
  double noise(double x, double y, double z) const {
    const double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    const int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
    x -= fx; y -= fy; z -= fz;
    const double u = fade(x), v = fade(y), w = fade(z);
    const int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z;
    const int B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;
    return lerp(w,
                lerp(v, lerp(u, grad(p[AA], x, y, z), grad(p[BA], x - 1, y, z)),
                     lerp(u, grad(p[AB], x, y - 1, z), grad(p[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(p[AA + 1], x, y, z - 1), grad(p[BA + 1], x - 1, y, z - 1)),
                     lerp(u, grad(p[AB + 1], x, y - 1, z - 1), grad(p[BB + 1], x - 1, y - 1, z - 1))));
  }
};

bool bad_dims(const int32_t d[3]) { return !d || d[0] < 1 || d[1] < 1 || d[2] < 1; }

}  // namespace

extern "C" {

void rvk_synth_uniform(uint64_t seed, int64_t n, double* out) {
  Xoshiro256ss rng(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = rng.uniform();
}

int rvk_synth_blobs(const int32_t dims[3], uint64_t seed, int n_blobs, double alo, double ahi, double slo,
                    double shi, double noise, double vmax, int out_is_u16, void* out) {
  if (bad_dims(dims) || !out || n_blobs < 0) return 1;
  Xoshiro256ss rng(seed);
  const auto blobs = make_blobs(rng, dims, n_blobs, alo, ahi, slo, shi);
  const uint64_t nseed = rng.next();
  const long sy = dims[0], sz = (long)dims[0] * dims[1];
  parallel_z(dims[2], [&](int z) {
    for (int y = 0; y < dims[1]; ++y)
      for (int x = 0; x < dims[0]; ++x) {
        const long i = x + y * sy + z * sz;
        double v = blob_sum(blobs, x, y, z) + noise * hash_uniform(nseed, (uint64_t)i);
        v = std::min(std::max(v, 0.0), vmax);
        const double r = std::nearbyint(v);
        if (out_is_u16) ((uint16_t*)out)[i] = (uint16_t)r;
        else ((uint8_t*)out)[i] = (uint8_t)r;
      }
  });
  return 0;
}

int rvk_synth_ct_phantom(const int32_t dims[3], const double spacing[3], uint64_t seed, int n_bones, int16_t* out) {
  if (bad_dims(dims) || !spacing || !out || n_bones < 0 || n_bones > 64) return 1;
  Xoshiro256ss rng(seed);
  double ext[3], c[3], body[3];
  for (int a = 0; a < 3; ++a) {
    ext[a] = dims[a] * spacing[a];
    c[a] = 0.5 * ext[a] + rng.uniform(-0.02, 0.02) * ext[a];
  }
  body[0] = rng.uniform(0.36, 0.42) * ext[0];
  body[1] = rng.uniform(0.26, 0.34) * ext[1];
  body[2] = rng.uniform(0.38, 0.44) * ext[2];
  struct Ell { double c[3], r[3]; double hu; };
  std::vector<Ell> bones(n_bones);
  for (auto& b : bones) {
    for (int a = 0; a < 3; ++a) {
      b.c[a] = c[a] + rng.uniform(-0.55, 0.55) * body[a];
      b.r[a] = rng.uniform(0.04, 0.12) * ext[a];
    }
    b.hu = rng.uniform(400.0, 1500.0);
  }
  const uint64_t nseed = rng.next();
  const long sy = dims[0], sz = (long)dims[0] * dims[1];
  parallel_z(dims[2], [&](int z) {
    for (int y = 0; y < dims[1]; ++y)
      for (int x = 0; x < dims[0]; ++x) {
        const long i = x + y * sy + z * sz;
        const double w[3] = {(x + 0.5) * spacing[0], (y + 0.5) * spacing[1], (z + 0.5) * spacing[2]};
        double q = 0.0;
        for (int a = 0; a < 3; ++a) q += (w[a] - c[a]) * (w[a] - c[a]) / (body[a] * body[a]);
        double hu = -1000.0;
        if (q <= 1.0) {
          hu = 40.0 + 20.0 * hash_normal(nseed, (uint64_t)i);
          for (const auto& b : bones) {
            double qb = 0.0;
            for (int a = 0; a < 3; ++a) qb += (w[a] - b.c[a]) * (w[a] - b.c[a]) / (b.r[a] * b.r[a]);
            if (qb <= 1.0) { hu = b.hu; break; }
          }
        }
        out[i] = (int16_t)std::nearbyint(std::min(std::max(hu, -1024.0), 3071.0));
      }
  });
  return 0;
}

int rvk_synth_blobs_perlin_u16(const int32_t dims[3], uint64_t seed, int n_blobs, double base_period, int octaves,
                               double vmax, uint16_t* out) {
  if (bad_dims(dims) || !out || !(base_period > 0) || octaves < 1 || octaves > 8) return 1;
  Xoshiro256ss rng(seed);
  const auto blobs = make_blobs(rng, dims, n_blobs, 0.5, 1.0, dims[0] * 0.03, dims[0] * 0.12);
  const Perlin perlin(rng);
  const long sy = dims[0], sz = (long)dims[0] * dims[1];
  const long n = sz * dims[2];
  std::vector<float> f((size_t)n);
  parallel_z(dims[2], [&](int z) {
    for (int y = 0; y < dims[1]; ++y)
      for (int x = 0; x < dims[0]; ++x) {
        double nsum = 0.0, amp = 0.5, per = base_period;
        for (int o = 0; o < octaves; ++o) {
          nsum += amp * perlin.noise(x / per, y / per, z / per);
          amp *= 0.5;
          per *= 0.5;
        }
        f[x + y * sy + z * sz] = (float)(blob_sum(blobs, x, y, z) + 0.6 * nsum);
      }
  });
  float lo = f[0], hi = f[0];
  for (long i = 0; i < n; ++i) { lo = std::min(lo, f[i]); hi = std::max(hi, f[i]); }
  const double scale = (hi > lo) ? vmax / ((double)hi - lo) : 0.0;
  parallel_z(dims[2], [&](int z) {
    for (long i = z * sz; i < (z + 1) * sz; ++i)
      out[i] = (uint16_t)std::nearbyint(std::min(std::max(((double)f[i] - lo) * scale, 0.0), vmax));
  });
  return 0;
}

}  // extern "C"
