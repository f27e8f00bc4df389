/*
 * rvk.h -- C-ABI of the MI355X (gfx950) ray-casting library librvkhip.so.
 *
 * It replaces the nvcc-compiled "renderAlgo" boundary of RobartsVTK's CUDA
 * volume mappers, and the CUDA device/stream handling under it.  Paths below
 * are relative to the reference root; V/ = Visualization/vtkCudaVisualization,
 * C/ = Common/vtkCudaCommon.
 *
 * Conventions (differences from the reference are deliberate):
 *   - every entry point returns an rvk_status (0 = OK) and never throws; the
 *     reference returned `cudaGetLastError()==0` as a bool
 *     (e.g. V/CUDA_vtkCuda1DVolumeMapper_renderAlgo.cuh:302);
 *   - there is no process-global device state: per-render parameters are
 *     kernel arguments, not __constant__ symbols and texture references
 *     (V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:12-19), so calls on different
 *     contexts need no mutex;
 *   - the library owns device allocations behind opaque handles; render
 *     outputs are caller-owned device pointers written asynchronously on the
 *     context's stream;
 *   - volumes keep their native scalar type on the device (no float widening).
 */
#ifndef RVK_H
#define RVK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RVK_VERSION 100

/* At most this many frames per volume (V/vtkCudaVolumeMapper.h:59). */
#define RVK_MAX_FRAMES 30

typedef enum rvk_status {
  RVK_OK = 0,
  RVK_ERR_INVALID_ARGUMENT = 1,
  RVK_ERR_NO_DEVICE = 2,
  RVK_ERR_HIP = 3,
  RVK_ERR_OUT_OF_MEMORY = 4,
  RVK_ERR_NOT_READY = 5,
  RVK_ERR_UNSUPPORTED = 6
} rvk_status;

/* Voxel scalar types accepted by SetInputData (V/vtkCuda1DVolumeMapper.cxx:107-187). */
typedef enum rvk_dtype {
  RVK_U8 = 0, RVK_I8 = 1, RVK_U16 = 2, RVK_I16 = 3, RVK_U32 = 4, RVK_I32 = 5, RVK_F32 = 6
} rvk_dtype;

/* DRR output modes.  RGBA8_REF is the reference's tinted image whose alpha
 * carries 256*(1 - offset*exp(-L)) (V/CUDA_vtkCudaDRRImageVolumeMapper_renderAlgo.cuh:114-115,163-166). */
typedef enum rvk_drr_output {
  RVK_DRR_RGBA8_REF = 0,
  RVK_DRR_F32_EXP_NEG_L = 1, /* float exp(-L) */
  RVK_DRR_F32_L = 2          /* float L, the line integral */
} rvk_drr_output;

typedef struct rvk_ctx rvk_ctx;
typedef struct rvk_volume rvk_volume;
typedef struct rvk_tf1d rvk_tf1d;
typedef struct rvk_tf2d rvk_tf2d;

/* One view.  view_to_voxels is the pre-baked screen->voxel matrix of
 * vtkCudaRendererInformationHandler::SetViewToVoxelsMatrix
 * (V/vtkCudaRendererInformationHandler.cxx:112-139), row major: pixel (x,y)
 * maps through (vx, vy, z, 1) with vx = flipped ? x/W : 1-x/W, vy = y/H.
 * Planes are (a,b,c,d) in voxel space as FigurePlanes makes them
 * (V/vtkCudaRendererInformationHandler.cxx:216-260). */
typedef struct rvk_view {
  float view_to_voxels[16];
  int32_t width, height;
  int32_t flipped;
  int32_t n_clip_planes;     /* 0..6 */
  float clip_planes[24];
  int32_t n_keyhole_planes;  /* 0..6 */
  float keyhole_planes[24];
  const float* zbuffer;      /* device pointer or NULL (= depth 1, no geometry) */
  int32_t zb_width, zb_height;
} rvk_view;

/* March parameters.  Reference defaults: step_scale 1 (one MinSpacing per
 * step, V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:414-419), term_T 1/64 (1D,
 * V/CUDA_vtkCuda1DVolumeMapper_renderAlgo.cuh:135) or 1/32 (2D,
 * V/CUDA_vtkCuda2DVolumeMapper_renderAlgo.cuh:313), jitter on, skip rule on. */
typedef struct rvk_march_params {
  float step_scale;  /* sample distance in units of the smallest voxel spacing */
  float term_T;      /* stop a ray once its transmittance drops below term_T */
  int32_t jitter;    /* 1: per-pixel 16x16 ray-offset table, 0: none */
  int32_t skip_rule; /* 1: reference two-step skip/back-step, 0: plain march */
  int32_t precise;   /* 0: reference FP32 march (p += inc each step); 1: sample
                        positions from the step index and a compensated DRR
                        line integral (no drift over long rays) */
} rvk_march_params;

/* DRR parameters: mu = slope*value + intercept per mm; reference defaults
 * 0.022 / 0.000022 / 1.0 (V/vtkCudaDRRImageVolumeMapper.cxx:21-23). */
typedef struct rvk_drr_params {
  float ct_intercept, ct_slope, ct_offset;
  int32_t output;          /* rvk_drr_output */
  int32_t ref_termination; /* 1: stop once L >= ln(256*ct_offset) */
  uint8_t tint[4];         /* RGBA tint for RVK_DRR_RGBA8_REF */
} rvk_drr_params;

/* 1D TF scalars: the POD half of cuda1DTransferFunctionInformation
 * (V/CUDA_container1DTransferFunctionInformation.h:20-39). */
typedef struct rvk_tf1d_params {
  float intensity_low, intensity_multiplier;
  float gradient_low, gradient_multiplier; /* non-finite multiplier: no gradient opacity */
  float ambient, diffuse, specular, specular_power;
} rvk_tf1d_params;

/* 2D TF scalars (V/CUDA_container2DTransferFunctionInformation.h:20-27). */
typedef struct rvk_tf2d_params {
  float intensity_low, intensity_multiplier;
  float gradient_low, gradient_multiplier, gradient_offset;
} rvk_tf2d_params;

/* Dual-image TF scalars: each component's low end and 1/range
 * (V/CUDA_containerDualImageTransferFunctionInformation.h:20-30); component 1
 * picks the table column, component 2 the row. */
typedef struct rvk_dual_params {
  float intensity1_low, intensity1_multiplier;
  float intensity2_low, intensity2_multiplier;
} rvk_dual_params;

/* Optional per-pixel side outputs of the compositing casters (device
 * pointers, each W*H floats, any may be NULL): steps taken, and the min/max
 * ray depths formRays writes (V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:373,387). */
typedef struct rvk_aux_out {
  float* depth;
  float* min_depth;
  float* max_depth;
} rvk_aux_out;

/* Cel and distance shading constants as the renderer handler stores them:
 * darkness, start, 1/(stop-start) (V/vtkCudaRendererInformationHandler.cxx:79-99). */
typedef struct rvk_shade_params {
  float celr, cela, celc;
  float disr, disa, disc;
} rvk_shade_params;

/* ---- errors -------------------------------------------------------------- */
/* Message of the last failing call on this thread ("" if none). */
const char* rvk_last_error_string(void);
const char* rvk_status_string(int status);
int rvk_version(void);

/* ---- device / stream context ----------------------------------------------
 * Replaces CudaObject + vtkCudaDeviceManager (C/CudaObject.h:44-51,
 * C/vtkCudaDeviceManager.cxx:249-296,365-381): one context = one device and
 * one stream.  rvk_ctx_create makes its own stream (default priority);
 * rvk_ctx_create_on_stream borrows a caller's hipStream_t (NULL = the
 * device's default stream). */
int rvk_device_count(int* count);
int rvk_ctx_create(int device, rvk_ctx** out);
int rvk_ctx_create_on_stream(int device, void* hip_stream, rvk_ctx** out);
int rvk_ctx_destroy(rvk_ctx* ctx);
int rvk_ctx_device(const rvk_ctx* ctx, int* device);
int rvk_ctx_stream(const rvk_ctx* ctx, void** hip_stream);
int rvk_ctx_set_stream(rvk_ctx* ctx, void* hip_stream); /* NULL: back to the context's own stream */
int rvk_ctx_sync(rvk_ctx* ctx);                          /* CallSyncThreads */
/* Device memory on the context's device (the output and z-buffer handling the
 * reference's output handler did with raw CUDA calls,
 * V/vtkCudaOutputImageInformationHandler.cxx:316-399).  Copies are ordered on
 * the context stream; to-host copies return once the data has arrived. */
int rvk_device_alloc(rvk_ctx* ctx, size_t bytes, void** out);
int rvk_device_free(rvk_ctx* ctx, void* ptr);
int rvk_copy_to_device(rvk_ctx* ctx, void* dst, const void* src, size_t bytes);
int rvk_copy_to_host(rvk_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- volumes --------------------------------------------------------------
 * Replaces CUDA_castBuffer<T,float> + *_loadImageInfo / changeFrame /
 * clearImageArray (V/CUDA_vtkCudaVolumeMapper_renderAlgo.h:30-35,
 * V/CUDA_vtkCuda1DVolumeMapper_renderAlgo.h:74-77).  Frames keep their native
 * type in a device cell layout (see rvk_volume_set_layout); src is x-fastest
 * (VTK order) host or device memory, borrowed only for the duration of the
 * call.  Float cells store each x-pair of voxels as (a, b - a); so do 16-bit
 * cells when the frame's values span less than 32768 (checked on the device
 * at upload, which then waits for the check), with a stored as an IEEE
 * binary16 when all values lie in [-2048, 2048], and b - a too when every
 * x-neighbour difference is exact in binary16 (|b - a| <= 2048, or even).
 * Renders are bit-identical in every encoding. */
int rvk_volume_create(rvk_ctx* ctx, const int32_t dims[3], const double spacing[3], int dtype,
                      int n_frames, rvk_volume** out);
int rvk_volume_set_frame(rvk_volume* vol, int frame, const void* src, int src_on_device);
/* u8 gradient-magnitude companion volume, |grad| = value*scale, computed on
 * the device by central differences (0.5/spacing), scale = max_gradient/255. */
int rvk_volume_compute_gradmag(rvk_volume* vol, int frame, float max_gradient);
int rvk_volume_destroy(rvk_volume* vol);
/* Cell order of the device layout (no reference counterpart: the reference
 * keeps a float cudaArray whose tiling the driver picks).  Each trilinear
 * 2x2x2 footprint is stored whole; RVK_CELLS_PLAIN orders the cells x-fastest,
 * RVK_CELLS_BRICK4 groups them in 4x4x4 bricks.  RVK_CELLS_PAIR8 stores, per
 * 8x8x8 brick of cells, each x-pair of voxels its cells use once (9 x 9 rows
 * of 8 pairs: 2.53x the native voxel bytes instead of 8x) and reads a cell as
 * four pairs.  The default is BRICK4 when the cells take more than 1 GiB (or
 * PLAIN would exceed its limit of 2^24 cell rows, (ny + 2)(nz + 2)), PLAIN
 * below; set_layout must come before the first rvk_volume_set_frame.
 * Renders give identical results in every layout. */
enum { RVK_CELLS_PLAIN = 0, RVK_CELLS_BRICK4 = 1, RVK_CELLS_PAIR8 = 2 };
int rvk_volume_set_layout(rvk_volume* vol, int layout);
int rvk_volume_get_layout(const rvk_volume* vol, int* layout);
/* Re-encode every loaded frame (and gradient-magnitude volume) of vol in
 * another cell layout on the device; each frame keeps its encoding, renders
 * stay identical.  Waits for the context stream; graphs captured on the
 * volume before are not replayed afterwards (the next call launches
 * directly).  Needs device memory for the new cells and one native copy of
 * the frames beside the old cells; when it cannot allocate them it fails
 * with RVK_ERR_OUT_OF_MEMORY and leaves the volume unchanged. */
int rvk_volume_relayout(rvk_volume* vol, int layout);
int rvk_volume_dims(const rvk_volume* vol, int32_t dims[3], double spacing[3], int* dtype, int* n_frames);

/* ---- transfer functions -----------------------------------------------------
 * 1D: replaces CUDA_vtkCuda1DVolumeMapper_renderAlgo_loadTextures
 * (V/CUDA_vtkCuda1DVolumeMapper_renderAlgo.h:61-64).  Tables of n floats
 * (n <= 2048; the five tables are staged in LDS); galpha may be NULL (all ones). */
int rvk_tf1d_create(rvk_ctx* ctx, int n, const float* alpha, const float* galpha, const float* red,
                    const float* green, const float* blue, const rvk_tf1d_params* params, rvk_tf1d** out);
int rvk_tf1d_destroy(rvk_tf1d* tf);
/* 2D: replaces CUDA_vtkCuda2DVolumeMapper_renderAlgo_loadTextures
 * (V/CUDA_vtkCuda2DVolumeMapper_renderAlgo.h:30-80).  tables[8] are n*n
 * floats indexed g*n+i, in the order alpha, red, green, blue, ambient,
 * diffuse, specular, specular power.  keyhole_tables may be NULL. */
int rvk_tf2d_create(rvk_ctx* ctx, int n, const float* const tables[8], const float* const keyhole_tables[8],
                    const rvk_tf2d_params* params, rvk_tf2d** out);
int rvk_tf2d_destroy(rvk_tf2d* tf);
/* In/Ex logic table for rvk_render_2d_inex: n*n floats indexed like the
 * other tables (the reference loads it beside them,
 * V/CUDA_vtkCuda2DInExLogicVolumeMapper_renderAlgo.h:35-45).  The caster only
 * tests inEx > 0 and inEx > 0.5, so the table is kept as those two bits per
 * cell.  NULL removes it. */
int rvk_tf2d_set_inex(rvk_tf2d* tf, const float* inex_table);

/* ---- renders ----------------------------------------------------------------
 * Each fuses formRays into the caster (V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:391-447).
 * rvk_render_1d replaces CUDA_vtkCuda1DVolumeMapper_renderAlgo_doRender
 * (V/CUDA_vtkCuda1DVolumeMapper_renderAlgo.h:29-33); out is W*H RGBA8. */
int rvk_render_1d(rvk_ctx* ctx, const rvk_volume* vol, int frame, const rvk_tf1d* tf, const rvk_view* view,
                  const rvk_march_params* march, uint8_t* out_rgba, const rvk_aux_out* aux);
/* Replaces CUDA_vtkCuda2DVolumeMapper_renderAlgo_doRender
 * (V/CUDA_vtkCuda2DVolumeMapper_renderAlgo.h:30-34).  use_gradmag = 1 reads
 * the gradient magnitude from rvk_volume_compute_gradmag's volume (shading
 * then takes the Phong-Lambert factor as 1). */
int rvk_render_2d(rvk_ctx* ctx, const rvk_volume* vol, int frame, const rvk_tf2d* tf, const rvk_view* view,
                  const rvk_march_params* march, int use_gradmag, uint8_t* out_rgba, const rvk_aux_out* aux);
/* Replaces CUDA_vtkCuda2DInExLogicVolumeMapper_renderAlgo_doRender
 * (V/CUDA_vtkCuda2DInExLogicVolumeMapper_renderAlgo.h:29-33): the 2D caster
 * with in/ex logic (tf needs rvk_tf2d_set_inex; its keyhole tables, if any,
 * are not used).  black_keyhole: rays that composited anything are opaque. */
int rvk_render_2d_inex(rvk_ctx* ctx, const rvk_volume* vol, int frame, const rvk_tf2d* tf, const rvk_view* view,
                       const rvk_march_params* march, int black_keyhole, uint8_t* out_rgba, const rvk_aux_out* aux);
/* Replaces CUDA_vtkCudaDualImageVolumeMapper_renderAlgo_doRender
 * (V/CUDA_vtkCudaDualImageVolumeMapper_renderAlgo.h:29-33): the two components
 * of the reference's 2-component image as two volumes of one size, type and
 * cell layout (frame applies to both).  tf's own params are not used; its
 * keyhole tables select the keyhole-TF variant.  A component frame stored
 * with binary16 bases is re-encoded in place, on ctx's stream, with int16
 * bases (the same values): work on other streams that reads that frame must
 * not overlap the call. */
int rvk_render_dual(rvk_ctx* ctx, const rvk_volume* vol1, const rvk_volume* vol2, int frame, const rvk_tf2d* tf,
                    const rvk_dual_params* params, const rvk_view* view, const rvk_march_params* march,
                    uint8_t* out_rgba, const rvk_aux_out* aux);
/* Replaces CUDA_vtkCudaDRRImageVolumeMapper_renderAlgo_doRender
 * (V/CUDA_vtkCudaDRRImageVolumeMapper_renderAlgo.h:30-35).  out is W*H
 * RGBA8 or W*H floats according to drr->output. */
int rvk_render_drr(rvk_ctx* ctx, const rvk_volume* vol, int frame, const rvk_view* view,
                   const rvk_march_params* march, const rvk_drr_params* drr, void* out);
/* normBuffer + doCelShade (V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:449-512)
 * in place on an RGBA8 image; depth/min/max from rvk_aux_out.  Neighbours
 * past the last row/column are clamped (the reference reads past them). */
int rvk_post_shade(rvk_ctx* ctx, int width, int height, const rvk_shade_params* shade, float* depth,
                   const float* min_depth, const float* max_depth, uint8_t* rgba);
/* rvk_post_shade without any allocation, for callers that cannot allocate
 * device memory (code under hipGraph capture, tight serving loops, hosts that
 * forbid it).  No counterpart in the reference.  Same arguments and results
 * as rvk_post_shade; the per-call staging buffer of the normalised depth is
 * instead a fixed device array of RVK_POST_SHADE_NOALLOC_MAX_PIXELS floats
 * (16 MiB) per GPU, held until the process ends; larger images fail with
 * RVK_ERR_INVALID_ARGUMENT.  Calls share that array, so calls that overlap on
 * one GPU (different streams or threads) race. */
#define RVK_POST_SHADE_NOALLOC_MAX_PIXELS 4194304
int rvk_post_shade_noalloc(rvk_ctx* ctx, int width, int height, const rvk_shade_params* shade, float* depth,
                           const float* min_depth, const float* max_depth, uint8_t* rgba);
/* The In/Ex-logic mapper's cel shade (V/CUDA_vtkCuda2DInExLogicVolumeMapper_renderAlgo.cuh:182-239):
 * squared depth differences, and pixels with zero alpha get
 * alpha = (signed char)(255 - 255 g) as a byte. */
int rvk_post_shade_inex(rvk_ctx* ctx, int width, int height, const rvk_shade_params* shade, float* depth,
                        const float* min_depth, const float* max_depth, uint8_t* rgba);

/* ---- batched DRR and its adjoint (no counterpart in the reference) ----------
 * A plan holds many views of one volume frame -- rigid poses and detectors of
 * different sizes -- and renders all of them in one launch.  Outputs are one
 * packed float buffer: view v occupies W_v*H_v floats starting at
 * rvk_drr_plan_offsets()[v].  Only the float output modes are batched.
 *
 * Work is cut into 8x8-pixel tiles, one per wave64.  Schedules:
 *   RVK_SCHED_GRID        one wave per tile, tiles XCD-remapped;
 *   RVK_SCHED_PERSISTENT  a grid sized to the device; each wave pulls tiles
 *                         from per-XCD work queues and then steals from the
 *                         other XCDs' queues until all are empty.
 * rvk_drr_plan_autotune times the candidates on this plan and keeps the
 * fastest; rvk_drr_plan_capture records the launch sequence in a hipGraph
 * that rvk_drr_plan_render then replays. */
typedef struct rvk_drr_plan rvk_drr_plan;

typedef enum rvk_schedule { RVK_SCHED_GRID = 1, RVK_SCHED_PERSISTENT = 2 } rvk_schedule;

/* dview (nullable): per view the 6 x 16 pose derivatives of view_to_voxels
 * (rvk_posed_view's dm_out), needed by the adjoint. */
int rvk_drr_plan_create(rvk_ctx* ctx, const rvk_volume* vol, int frame, int n_views, const rvk_view* views,
                        const float* dview, const rvk_march_params* march, const rvk_drr_params* drr,
                        rvk_drr_plan** out);
int rvk_drr_plan_destroy(rvk_drr_plan* plan);
/* Replace the views' matrices (and derivatives) keeping their sizes; a
 * captured graph stays valid (it reads the views from device memory). */
int rvk_drr_plan_update_views(rvk_drr_plan* plan, const rvk_view* views, const float* dview);
int rvk_drr_plan_offsets(const rvk_drr_plan* plan, int64_t* offsets /* n_views + 1 */);
/* set_schedule applies to the render and to the registration forward;
 * get_schedule reports the render's. */
int rvk_drr_plan_set_schedule(rvk_drr_plan* plan, int schedule, int waves_per_simd);
int rvk_drr_plan_get_schedule(const rvk_drr_plan* plan, int* schedule, int* waves_per_simd);
/* Wave tile of the batched forward kernels: the 64 rays of a wave as an 8x8,
 * 16x4 or 4x16 pixel block (results do not depend on it). */
enum { RVK_TILE_8X8 = 0, RVK_TILE_16X4 = 1, RVK_TILE_4X16 = 2 };
int rvk_drr_plan_set_tile_shape(rvk_drr_plan* plan, int shape);
int rvk_drr_plan_get_tile_shape(const rvk_drr_plan* plan, int* shape);
/* Order in which the batched forward takes the views' tiles (results do not
 * depend on it): view after view, or grouped by the volume region each
 * tile's central ray crosses, all views' tiles of a region together, so one
 * XCD's L2 serves that region's cells to every view (recomputed when the
 * views are updated from the host). */
enum { RVK_ORDER_VIEWS = 0, RVK_ORDER_REGIONS = 1 };
int rvk_drr_plan_set_work_order(rvk_drr_plan* plan, int order);
int rvk_drr_plan_get_work_order(const rvk_drr_plan* plan, int* order);
/* Time every candidate -- schedule x waves per SIMD x wave tile x work order -- `reps`
 * times on this plan (writes into out) and keep the fastest; *best_ms
 * (nullable) receives its time.  A plan with pose derivatives also tunes the
 * registration step's forward (the march that produces the pose Jacobians)
 * on its own. */
int rvk_drr_plan_autotune(rvk_drr_plan* plan, float* out, int reps, float* best_ms);
/* rvk_drr_plan_autotune in every cell layout (brick size) the volume fits,
 * by rvk_volume_relayout of vol -- the plan's own volume, passed again
 * because it changes for every plan that renders it -- then vol in the
 * fastest layout and the plan with that layout's best schedule and tile.
 * *layout (nullable) receives the RVK_CELLS_* chosen. */
int rvk_drr_plan_autotune_layout(rvk_drr_plan* plan, rvk_volume* vol, float* out, int reps, float* best_ms,
                                 int* layout);
/* Render every view into the packed buffer `out` on the context stream. */
int rvk_drr_plan_render(rvk_drr_plan* plan, float* out);
/* Capture render(out) (and, with targets, the registration step below) into
 * a hipGraph; later rvk_drr_plan_render / rvk_drr_plan_register calls with
 * the same buffers replay it (as long as the frame keeps the cell encoding
 * it had at capture; a re-uploaded 16-bit frame whose value range crossed
 * 32767 is rendered by a direct launch instead).  Passing NULL drops the
 * graph. */
int rvk_drr_plan_capture(rvk_drr_plan* plan, float* out, const float* targets, float* loss, float* grad_pose);

/* Zero-mean normalised cross-correlation of every view with packed target
 * images: loss_v = 1 - NCC_v (n_views floats) and, if dloss_dimage is not
 * NULL, d loss_v / d image (packed like the images).  The NCC is the centred
 * form of vtkImageNormalizedCrossCorrelation::GetResult
 * (ImageProcessing/vtkRobartsRegistration/vtkImageNormalizedCrossCorrelation.cxx:203-218,
 * which is uncentred); a constant image or target gives NCC 0 and no gradient. */
int rvk_drr_plan_ncc(rvk_drr_plan* plan, const float* images, const float* targets, float* loss,
                     float* dloss_dimage);
/* Which NCC the plan's loss uses (rvk_drr_plan_ncc, _register and the
 * registration step): RVK_NCC_CENTERED (default, zero-mean) or
 * RVK_NCC_UNCENTERED, the reference's own
 * sum(S T) / (sqrt(sum S^2) sqrt(sum T^2)) with S the rendered image and T
 * the target (vtkImageNormalizedCrossCorrelation::GetResult,
 * ImageProcessing/vtkRobartsRegistration/vtkImageNormalizedCrossCorrelation.cxx:203-218).
 * Drops a captured graph when the mode changes. */
enum { RVK_NCC_CENTERED = 0, RVK_NCC_UNCENTERED = 1 };
int rvk_drr_plan_set_ncc(rvk_drr_plan* plan, int mode);
/* Adjoint of the batched render: given the rendered images and d loss / d
 * image (packed), write d loss / d pose (n_views x 6 floats, pose as in
 * rvk_pose_matrix; RVK_ERR_NOT_READY when the plan has no dview) and/or
 * d loss / d volume (float32 per voxel, x fastest: the derivative with respect
 * to the stored voxel value, before the plan's intercept and slope, for every
 * voxel type; plain march only; samples in a view's keyhole window are
 * excluded, as in the forward).  For the pose gradient rays are re-marched with the forward's
 * control flow; nothing per sample is stored.
 * The volume gradient is brick-driven: each 8^3-voxel brick re-marches the
 * stretch of every ray that meets the cells its voxels weigh in and adds each
 * accumulated sample's trilinear weights times the ray's coefficient into
 * 64-bit fixed-point accumulators (a multiple of 2^-k, 2^k set per call so
 * that the largest |coefficient| is below 2^30), so the result is the same on
 * every call.  A non-finite d loss / d image on a ray with samples makes every
 * voxel whose neighbourhood can contribute NaN.  Its first call allocates a
 * per-pixel ray record buffer (48 bytes per pixel) and 4 bytes per 8x8 pixel
 * tile inside the plan.  Gradients are those of the precise march. */
int rvk_drr_plan_backward(rvk_drr_plan* plan, const float* images, const float* dloss_dimage, float* grad_pose,
                          float* grad_volume);
/* The registration step, fused: render every view into out, NCC against
 * targets (loss, n_views floats) and the pose gradient of the loss (n_views x
 * 6 floats), with d loss / d image formed inside the adjoint kernel. */
int rvk_drr_plan_register(rvk_drr_plan* plan, float* out, const float* targets, float* loss, float* grad_pose);

/* Streaming-read probe (no counterpart in the reference): launches one
 * coalesced pass that reads every byte of buf (16-byte aligned, a multiple of
 * 16 bytes) on the context stream; time it with events.  It measures the
 * device's attainable HBM read bandwidth, the denominator the gather
 * roofline is checked against.  sink (nullable, device) is written only in
 * a case no real buffer produces. */
int rvk_stream_read(rvk_ctx* ctx, const void* buf, size_t bytes, uint32_t* sink);

/* Ray-sample count of a plain march without opacity termination: sum over
 * pixels of floor(num_steps - ray offset) (the DRR's sample count, from the
 * geometry alone).  Synchronous; *out is a host pointer. */
int rvk_count_ray_steps(rvk_ctx* ctx, const rvk_volume* vol, const rvk_view* view, const rvk_march_params* march,
                        int64_t* out);

#ifdef __cplusplus
}
#endif
#endif
