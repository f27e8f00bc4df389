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

/* 4x4 helpers (row major). */
void rvk_mat4_mul(const double a[16], const double b[16], double out[16]);
int rvk_mat4_invert(const double a[16], double out[16]);

#ifdef __cplusplus
}
#endif
#endif
