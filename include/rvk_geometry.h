/*
 * rvk_geometry.h -- host-side camera/volume matrix helpers of librvkhip.so.
 *
 * Replaces vtkCudaVolumeMapper::ComputeMatrices (V/vtkCudaVolumeMapper.cxx:349-447)
 * and vtkCudaRendererInformationHandler::SetViewToVoxelsMatrix
 * (V/vtkCudaRendererInformationHandler.cxx:112-139), with VTK's camera,
 * transform and matrix classes (absent here) restated from their published
 * math: an OpenGL frustum/ortho projection remapped to depth [0,1]
 * (vtkCamera::GetProjectionTransformMatrix(aspect, 0, 1)), a look-at view
 * matrix (vtkCamera::GetViewTransformMatrix), and a 4x4 inverse.
 * V/ = Visualization/vtkCudaVisualization.  All matrices are row major.
 */
#ifndef RVK_GEOMETRY_H
#define RVK_GEOMETRY_H

#include "rvk.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The vtkCamera state ComputeMatrices reads. */
typedef struct rvk_camera {
  double position[3];
  double focal_point[3];
  double view_up[3];
  double view_angle;          /* degrees; vertical unless use_horizontal_view_angle */
  double clipping_range[2];   /* near, far (> 0 for perspective) */
  double window_center[2];    /* VTK default (0, 0) */
  int32_t parallel_projection;
  double parallel_scale;
  int32_t use_horizontal_view_angle;
} rvk_camera;

/* Fill view->view_to_voxels (pre-baked) and width/height/flipped; planes and
 * z-buffer are cleared.  user_matrix is the volume's vtkProp3D::UserMatrix
 * (NULL = identity).  origin/spacing are those of the vtkImageData, with the
 * extent origin folded into origin.  aspect = width/height. */
int rvk_compute_view(const rvk_camera* cam, const double user_matrix[16], const double origin[3],
                     const double spacing[3], int width, int height, int flipped, rvk_view* view);

/* The un-baked double matrices: view_to_voxels = inverse(P * V * U * Vox). */
int rvk_compute_matrices(const rvk_camera* cam, const double user_matrix[16], const double origin[3],
                         const double spacing[3], double aspect, double view_to_voxels[16],
                         double voxels_to_view[16]);

/* Pre-bake a ViewToVoxels matrix for pixel coordinates (the reference's
 * SetViewToVoxelsMatrix): M[r][3] += M[r][0]-M[r][1]; M[r][0] *= -2; M[r][1] *= 2,
 * carried out in float as the reference stores it. */
void rvk_prebake_view_to_voxels(const double m[16], float out[16]);

/* Rigid pose of a volume about a centre (world mm): U = T(c + t) Rz Ry Rx T(-c),
 * pose = (rx, ry, rz [rad], tx, ty, tz [mm]).  d_out (nullable) receives
 * dU/dpose_i for i = 0..5, 6 x 16 doubles. */
void rvk_pose_matrix(const double pose[6], const double center[3], double out[16], double* d_out);

/* ViewToVoxels of a posed volume and its pose derivatives, both pre-baked:
 * m_out = bake(inverse(P V U(pose) Vox)), dm_out[i] = bake(d m / d pose_i)
 * (6 x 16 floats, nullable).  This is the per-view geometry the DRR adjoint
 * differentiates through. */
int rvk_posed_view(const rvk_camera* cam, const double pose[6], const double center[3], const double origin[3],
                   const double spacing[3], int width, int height, int flipped, rvk_view* view, float* dm_out);

/* ---- device-resident registration loop --------------------------------------
 * Replaces the derivative-free vtkPowellMinimizer loop around the registration
 * metric (ImageProcessing/vtkRobartsRegistration/vtkPowellMinimizer.h:30-117)
 * with a gradient step that stays on the device: the plan keeps the poses in
 * device memory and recomputes its views' matrices (and pose derivatives)
 * from them with the posed-view geometry above, so that one step -- render,
 * NCC, pose gradient, optimizer update, new views -- is a sequence of kernels
 * that rvk_drr_plan_capture_step records in one hipGraph.
 *
 * RVK_OPT_GD:   pose -= lr * g
 * RVK_OPT_ADAM: Adam (bias-corrected moments), pose -= lr * m^ / (sqrt(v^) + eps)
 * lr is lr_rotation for the three angles (rad) and lr_translation for the
 * three translations (mm).  shared_pose = 1 drives every view with one pose
 * (the volume's pose seen from several cameras; its gradient is the sum over
 * the views), 0 gives each view its own pose. */
enum { RVK_OPT_GD = 0, RVK_OPT_ADAM = 1 };
typedef struct rvk_pose_optimizer {
  int32_t kind;
  int32_t shared_pose;
  float lr_rotation, lr_translation;
  float beta1, beta2, eps;
} rvk_pose_optimizer;

/* Attach device poses to a plan created with pose derivatives: cameras (one
 * per view; the views keep their sizes and flip), the rotation centre (world
 * mm), origin (nullable: 0) and poses (6 doubles per pose: 1 pose when
 * shared_pose, else one per view).  Resets the optimizer state and sets the
 * views from the poses; drops a captured graph. */
int rvk_drr_plan_set_poses(rvk_drr_plan* plan, const rvk_camera* cameras, const double center[3],
                           const double origin[3], const double* poses, const rvk_pose_optimizer* opt);
/* Current device poses (synchronous). */
int rvk_drr_plan_get_poses(rvk_drr_plan* plan, double* poses);
/* One registration step: rvk_drr_plan_register at the current poses (out,
 * loss and grad_pose are those of the poses before the update), then the
 * optimizer update and the views at the new poses.  Replays the graph of
 * rvk_drr_plan_capture_step when called with its buffers. */
int rvk_drr_plan_step(rvk_drr_plan* plan, float* out, const float* targets, float* loss, float* grad_pose);
int rvk_drr_plan_capture_step(rvk_drr_plan* plan, float* out, const float* targets, float* loss, float* grad_pose);

/* 4x4 helpers (row major). */
void rvk_mat4_mul(const double a[16], const double b[16], double out[16]);
int rvk_mat4_invert(const double a[16], double out[16]);

#ifdef __cplusplus
}
#endif
#endif
