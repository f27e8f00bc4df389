/*
 * rvk_oracle.h -- CPU restatement of RobartsVTK's CUDA ray-casting kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP casters are tested against this
 * library: tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it; the product library (robartsvtk_amd/csrc) never
 * does.  Every function is templated on the working precision and exported
 * in an FP32 and an FP64 flavour.
 *
 * Reference files followed (paths relative to the reference root,
 * V/ = Visualization/vtkCudaVisualization/):
 *   ray setup ........ V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:38-447
 *   1D caster ........ V/CUDA_vtkCuda1DVolumeMapper_renderAlgo.cuh:27-227
 *   2D caster ........ V/CUDA_vtkCuda2DVolumeMapper_renderAlgo.cuh:30-470
 *   DRR caster ....... V/CUDA_vtkCudaDRRImageVolumeMapper_renderAlgo.cuh:11-171
 *   post passes ...... V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:449-512
 *   volume info ...... V/vtkCudaVolumeInformationHandler.cxx:80-108
 *
 * Parity status of each function is recorded in the header comment of
 * rvk_oracle.cpp.
 */
#ifndef RVK_ORACLE_H
#define RVK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* voxel scalar types (same codes as the product's rvk_dtype) */
enum { ORC_U8 = 0, ORC_I8 = 1, ORC_U16 = 2, ORC_I16 = 3, ORC_U32 = 4, ORC_I32 = 5, ORC_F32 = 6 };

/* DRR output modes */
enum { ORC_DRR_RGBA8_REF = 0, ORC_DRR_F32_EXP = 1, ORC_DRR_F32_L = 2 };

typedef struct orc_volume {
  const void* data;    /* x fastest, then y, then z (VTK point order) */
  int dtype;
  int dims[3];
  double spacing[3];
} orc_volume;

typedef struct orc_view {
  float m[16];         /* pre-baked ViewToVoxels matrix, row major */
  int width, height;   /* output resolution in pixels */
  int flipped;
  const float* zbuffer; /* nullable; zb_width x zb_height, row major */
  int zb_width, zb_height;
  int n_clip;          /* 0..6 clipping planes in voxel space (a,b,c,d) */
  float clip[24];
  int n_keyhole;       /* 0..6 keyhole planes in voxel space */
  float keyhole[24];
  const double* m64;   /* nullable: the f64 flavour uses this matrix instead of m
                          (finite differences of the pose need it unrounded) */
} orc_view;

typedef struct orc_march {
  float step_scale;    /* sample distance in units of MinSpacing (reference: 1) */
  float term_T;        /* stop when remaining transmittance < term_T */
  int jitter;          /* 1: 16x16 ray-offset table; 0: none */
  int skip_rule;       /* 1: two-step skip / back-step; 0: plain march */
  int precise;         /* 1: positions from the step index + compensated DRR sum */
  int f32_control;     /* f64 flavour only: replay the branch decisions of the
                          f32 march (arithmetic error along the FP32 path) */
} orc_march;

typedef struct orc_drr {
  float intercept, slope, offset;
  int output;          /* ORC_DRR_* */
  int ref_termination; /* 1: stop once L >= ln(256*offset) */
  unsigned char tint[4];
} orc_drr;

typedef struct orc_tf1d {
  int n;
  const float *alpha, *galpha, *r, *g, *b;
  float intensity_low, intensity_mult, gradient_low, gradient_mult;
  float ambient, diffuse, specular, spec_power;
} orc_tf1d;

typedef struct orc_tf2d {
  int n;                 /* tables are n x n, index = g*n + i */
  const float *alpha, *r, *g, *b, *ambient, *diffuse, *specular, *spec_power;
  const float *k_alpha, *k_r, *k_g, *k_b, *k_ambient, *k_diffuse, *k_specular, *k_spec_power; /* keyhole TF or NULL */
  float intensity_low, intensity_mult, gradient_low, gradient_mult, gradient_offset;
  const uint8_t* gradmag; /* optional precomputed gradient magnitude volume (same dims) */
  float gradmag_scale;    /* |grad| = gradmag * gradmag_scale */
} orc_tf2d;

typedef struct orc_shade {
  float celr, cela, celc;  /* cel shading: darkness, start, 1/(stop-start) */
  float disr, disa, disc;  /* distance shading */
} orc_shade;

/* The reference's 256-entry ray-offset table (V/vtkCudaVolumeMapper.cxx:59-93). */
const float* orc_jitter_table(void);

/* Per-pixel ray setup (formRays).  Outputs 9 SoA arrays of W*H (any may be
 * NULL): start xyz, inc xyz, numSteps, exclude start/end, plus min/max depth. */
int orc_form_rays_f32(const orc_volume* vol, const orc_view* view, const orc_march* mp,
                      float* out9, float* min_depth, float* max_depth);
int orc_form_rays_f64(const orc_volume* vol, const orc_view* view, const orc_march* mp,
                      double* out9, double* min_depth, double* max_depth);

/* DRR caster. out: W*H floats for F32 modes, W*H*4 bytes for RGBA8.  */
int orc_drr_f32(const orc_volume* vol, const orc_view* view, const orc_march* mp,
                const orc_drr* dp, void* out, int nthreads);
int orc_drr_f64(const orc_volume* vol, const orc_view* view, const orc_march* mp,
                const orc_drr* dp, void* out, int nthreads);

/* 1D-TF compositing caster. rgba8: W*H*4; rgbaf (nullable): W*H*4 floats
 * before quantisation; depth (nullable): W*H steps taken. */
int orc_1d_f32(const orc_volume* vol, const orc_tf1d* tf, const orc_view* view, const orc_march* mp,
               uint8_t* rgba8, float* rgbaf, float* depth, int nthreads);
int orc_1d_f64(const orc_volume* vol, const orc_tf1d* tf, const orc_view* view, const orc_march* mp,
               uint8_t* rgba8, double* rgbaf, double* depth, int nthreads);

/* 2D-TF compositing caster (uses the keyhole tables when k_alpha != NULL). */
int orc_2d_f32(const orc_volume* vol, const orc_tf2d* tf, const orc_view* view, const orc_march* mp,
               uint8_t* rgba8, float* rgbaf, float* depth, int nthreads);
int orc_2d_f64(const orc_volume* vol, const orc_tf2d* tf, const orc_view* view, const orc_march* mp,
               uint8_t* rgba8, double* rgbaf, double* depth, int nthreads);

/* In/Ex-logic 2D caster (vtkCuda2DInExLogicVolumeMapper): tf as for the 2D
 * caster (no keyhole tables), inex the n x n in/ex table, black_keyhole the
 * UseBlackKeyhole flag.  Output saturated. */
int orc_inex_f32(const orc_volume* vol, const orc_tf2d* tf, const float* inex, int black_keyhole,
                 const orc_view* view, const orc_march* mp, uint8_t* rgba8, float* rgbaf, float* depth, int nthreads);
int orc_inex_f64(const orc_volume* vol, const orc_tf2d* tf, const float* inex, int black_keyhole,
                 const orc_view* view, const orc_march* mp, uint8_t* rgba8, double* rgbaf, double* depth,
                 int nthreads);
/* Dual-image caster (vtkCudaDualImageVolumeMapper): two co-registered volumes
 * (same dims and dtype); the 2D table is indexed by intensity_mult*(v1 -
 * intensity_low) and gradient_mult*(v2 - gradient_low); keyhole tables
 * optional; gradmag fields unused. */
int orc_dual_f32(const orc_volume* vol1, const orc_volume* vol2, const orc_tf2d* tf, const orc_view* view,
                 const orc_march* mp, uint8_t* rgba8, float* rgbaf, float* depth, int nthreads);
int orc_dual_f64(const orc_volume* vol1, const orc_volume* vol2, const orc_tf2d* tf, const orc_view* view,
                 const orc_march* mp, uint8_t* rgba8, double* rgbaf, double* depth, int nthreads);
/* The In/Ex-logic mapper's post pass (normBuffer then its cel shade). */
int orc_post_shade_inex_f32(const orc_shade* sp, int W, int H, float* depth, const float* min_depth,
                            const float* max_depth, uint8_t* rgba8);

/* Post passes: normBuffer then cel/distance shading, in place. */
int orc_post_shade_f32(const orc_shade* sp, int W, int H, float* depth, const float* min_depth,
                       const float* max_depth, uint8_t* rgba8);

/* Metric sampler: number of trilinear positions the FP64 plain march visits
 * (no skip rule) from the jittered clipped entry to the ray exit.  For the
 * compositing casters the march also stops at the transmittance threshold;
 * pass tf1d or tf2d (or neither, for the DRR). */
/* Rays the precise DRR march took no samples of (their ends left the cell
 * range [0, n + 1]) since the last reset; 0 for every volume-clipped ray. */
long long orc_unsafe_rays(int reset);
long long orc_count_samples(const orc_volume* vol, const orc_view* view, const orc_march* mp,
                            const orc_tf1d* tf1d, const orc_tf2d* tf2d, int nthreads);

/* Trilinear samples the FP32 DRR march takes as configured (with the skip
 * rule on, the samples it evaluates, fewer than the plain march's). */
long long orc_count_drr_evaluated(const orc_volume* vol, const orc_view* view, const orc_march* mp,
                                  const orc_drr* dp, int nthreads);

/* Trilinear sample at position p (voxel index space, half-texel convention). */
double orc_sample_f64(const orc_volume* vol, double px, double py, double pz);
float orc_sample_f32(const orc_volume* vol, float px, float py, float pz);

/* Linear (1D, normalized, clamp) and point (2D, normalized, clamp) TF lookups. */
float orc_tf1d_lookup_f32(const float* table, int n, float u);
float orc_tf2d_lookup_f32(const float* table, int n, float u, float v);

#ifdef __cplusplus
}
#endif
#endif
