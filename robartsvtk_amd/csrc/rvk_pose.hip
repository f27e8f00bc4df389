// rvk_pose.hip -- rigid pose matrices and posed views (host entry points).
//
// The geometry itself is rvk_pose_math.h, shared with the plan's on-device
// pose updates so that host and device produce the same matrices.
#include <cstring>

#include "rvk_internal.h"
#include "rvk_pose_math.h"

extern "C" {

void rvk_pose_matrix(const double pose[6], const double center[3], double out[16], double* d_out) {
  rvk::pose::pose_matrix(pose, center, out, d_out);
}

int rvk_posed_view(const rvk_camera* cam, const double pose[6], const double center[3], const double origin[3],
                   const double spacing[3], int width, int height, int flipped, rvk_view* view, float* dm_out) {
  if (!cam || !pose || !center || !origin || !spacing || !view || width <= 0 || height <= 0)
    return RVK_ERR_INVALID_ARGUMENT;
  double pv[16], vox[16];
  int rc = rvk::camera_pv(cam, (double)width / (double)height, pv);
  if (rc) return rc;
  rvk::voxels_matrix(origin, spacing, vox);
  std::memset(view, 0, sizeof(*view));
  if (!rvk::pose::posed_view(pv, vox, pose, center, view->view_to_voxels, dm_out)) return RVK_ERR_INVALID_ARGUMENT;
  view->width = width;
  view->height = height;
  view->flipped = flipped;
  return RVK_OK;
}

}  // extern "C"
