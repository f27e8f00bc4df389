/*
 * rvk_synth.h -- seeded synthetic inputs (host) for the benchmark and the
 * tests.  No counterpart in the reference, which ships no data; these build
 * the volumes its callers would load (CT, blob and noise phantoms) from a
 * project PRNG (SplitMix64 seeding xoshiro256**), so every run and every
 * language sees the same bits.  Voxel values depend only on (seed, index),
 * never on the thread count.
 */
#ifndef RVK_SYNTH_H
#define RVK_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sum of n isotropic Gaussian blobs (amplitude U[amp_lo,amp_hi], sigma
 * U[sig_lo,sig_hi] voxels, centres U[0.15,0.85]*dim) plus U[0,noise] noise,
 * clamped to [0,vmax] and rounded.  Writes u8 (vmax <= 255) or u16. */
int rvk_synth_blobs(const int32_t dims[3], uint64_t seed, int n_blobs, double amp_lo, double amp_hi, double sig_lo,
                    double sig_hi, double noise, double vmax, int out_is_u16, void* out);

/* CT-like phantom in HU (int16): air -1000, an ellipsoidal body of soft
 * tissue N(40,20) HU per voxel, and n_bones bone ellipsoids inside the body
 * with constant U[400,1500] HU each.  spacing in mm. */
int rvk_synth_ct_phantom(const int32_t dims[3], const double spacing[3], uint64_t seed, int n_bones, int16_t* out);

/* Improved Perlin noise (Perlin 2002 construction, seeded permutation),
 * `octaves` octaves starting at `base_period` voxels, added to Gaussian
 * blobs, rescaled to [0, vmax] and rounded to u16. */
int rvk_synth_blobs_perlin_u16(const int32_t dims[3], uint64_t seed, int n_blobs, double base_period, int octaves,
                               double vmax, uint16_t* out);

/* Uniform doubles in [0,1) from the project PRNG (n values). */
void rvk_synth_uniform(uint64_t seed, int64_t n, double* out);

#ifdef __cplusplus
}
#endif
#endif
