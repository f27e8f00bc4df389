// rvk_geometry.cpp -- camera, volume and pose matrices (host).
//
// Restates vtkCudaVolumeMapper::ComputeMatrices (V/vtkCudaVolumeMapper.cxx:349-447,
// V/ = Visualization/vtkCudaVisualization) without VTK:
//   PerspectiveTransform = P(aspect, 0, 1) * ViewTransform          (:374-376)
//   VoxelsToWorld        = UserMatrix * Translate(origin) * Scale(spacing) (:403-422)
//   ViewToVoxels         = inverse(PerspectiveTransform * VoxelsToWorld) (:437-442)
// and the pixel pre-bake of V/vtkCudaRendererInformationHandler.cxx:112-139.
#include <cmath>
#include <cstring>

#include "rvk_geometry.h"

namespace {

void identity(double m[16]) {
  std::memset(m, 0, 16 * sizeof(double));
  m[0] = m[5] = m[10] = m[15] = 1.0;
}

void normalize(double v[3]) {
  const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (n > 0) { v[0] /= n; v[1] /= n; v[2] /= n; }
}

void cross(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// vtkCamera view transform (look-at): rows are sideways, up, view-plane normal.
void view_matrix(const rvk_camera* c, double m[16]) {
  double n[3] = {c->position[0] - c->focal_point[0], c->position[1] - c->focal_point[1],
                 c->position[2] - c->focal_point[2]};
  normalize(n);
  double s[3];
  cross(c->view_up, n, s);
  normalize(s);
  double u[3];
  cross(n, s, u);
  identity(m);
  for (int j = 0; j < 3; ++j) {
    m[0 + j] = s[j];
    m[4 + j] = u[j];
    m[8 + j] = n[j];
  }
  for (int i = 0; i < 3; ++i)
    m[4 * i + 3] = -(m[4 * i] * c->position[0] + m[4 * i + 1] * c->position[1] + m[4 * i + 2] * c->position[2]);
}

// vtkCamera projection with the depth range remapped from [-1,1] to [0,1].
void projection_matrix(const rvk_camera* c, double aspect, double m[16]) {
  double p[16];
  std::memset(p, 0, sizeof(p));
  const double n = c->clipping_range[0], f = c->clipping_range[1];
  double xmin, xmax, ymin, ymax;
  if (c->parallel_projection) {
    const double w = c->parallel_scale * aspect, h = c->parallel_scale;
    xmin = (c->window_center[0] - 1.0) * w; xmax = (c->window_center[0] + 1.0) * w;
    ymin = (c->window_center[1] - 1.0) * h; ymax = (c->window_center[1] + 1.0) * h;
    p[0] = 2.0 / (xmax - xmin);
    p[5] = 2.0 / (ymax - ymin);
    p[10] = -2.0 / (f - n);
    p[3] = -(xmin + xmax) / (xmax - xmin);
    p[7] = -(ymin + ymax) / (ymax - ymin);
    p[11] = -(n + f) / (f - n);
    p[15] = 1.0;
  } else {
    const double t = std::tan(c->view_angle * M_PI / 360.0);
    double w, h;
    if (c->use_horizontal_view_angle) { w = n * t; h = n * t / aspect; }
    else { w = n * t * aspect; h = n * t; }
    xmin = (c->window_center[0] - 1.0) * w; xmax = (c->window_center[0] + 1.0) * w;
    ymin = (c->window_center[1] - 1.0) * h; ymax = (c->window_center[1] + 1.0) * h;
    p[0] = 2.0 * n / (xmax - xmin);
    p[2] = (xmin + xmax) / (xmax - xmin);
    p[5] = 2.0 * n / (ymax - ymin);
    p[6] = (ymin + ymax) / (ymax - ymin);
    p[10] = -(n + f) / (f - n);
    p[11] = -2.0 * n * f / (f - n);
    p[14] = -1.0;
  }
  // z' = (z + 1) / 2 : row2 <- 0.5 row2 + 0.5 row3
  for (int j = 0; j < 4; ++j) p[8 + j] = 0.5 * p[8 + j] + 0.5 * p[12 + j];
  std::memcpy(m, p, sizeof(p));
}

void voxels_to_world(const double* user, const double origin[3], const double spacing[3], double m[16]) {
  double vox[16];
  identity(vox);
  for (int a = 0; a < 3; ++a) {
    vox[5 * a] = spacing[a];
    vox[4 * a + 3] = origin[a];
  }
  if (user) rvk_mat4_mul(user, vox, m);
  else std::memcpy(m, vox, sizeof(vox));
}

}  // namespace

namespace rvk {

// P V of a camera (the posed-view geometry composes it with the pose and the
// voxel matrix, see rvk_pose_math.h).
int camera_pv(const rvk_camera* cam, double aspect, double pv[16]) {
  if (!cam || !(aspect > 0.0)) return RVK_ERR_INVALID_ARGUMENT;
  if (!cam->parallel_projection && !(cam->clipping_range[0] > 0.0)) return RVK_ERR_INVALID_ARGUMENT;
  if (!(cam->clipping_range[1] > cam->clipping_range[0])) return RVK_ERR_INVALID_ARGUMENT;
  double p[16], v[16];
  projection_matrix(cam, aspect, p);
  view_matrix(cam, v);
  rvk_mat4_mul(p, v, pv);
  return RVK_OK;
}

void voxels_matrix(const double origin[3], const double spacing[3], double vox[16]) {
  voxels_to_world(nullptr, origin, spacing, vox);
}

}  // namespace rvk

extern "C" {

void rvk_mat4_mul(const double a[16], const double b[16], double out[16]) {
  double r[16];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += a[4 * i + k] * b[4 * k + j];
      r[4 * i + j] = s;
    }
  std::memcpy(out, r, sizeof(r));
}

// Gauss-Jordan elimination with partial pivoting.
int rvk_mat4_invert(const double a[16], double out[16]) {
  double m[4][8];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      m[i][j] = a[4 * i + j];
      m[i][4 + j] = (i == j) ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; ++c) {
    int piv = c;
    for (int r = c + 1; r < 4; ++r)
      if (std::fabs(m[r][c]) > std::fabs(m[piv][c])) piv = r;
    if (m[piv][c] == 0.0) return RVK_ERR_INVALID_ARGUMENT;
    if (piv != c)
      for (int j = 0; j < 8; ++j) { double t = m[c][j]; m[c][j] = m[piv][j]; m[piv][j] = t; }
    const double inv = 1.0 / m[c][c];
    for (int j = 0; j < 8; ++j) m[c][j] *= inv;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      const double f = m[r][c];
      if (f != 0.0)
        for (int j = 0; j < 8; ++j) m[r][j] -= f * m[c][j];
    }
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out[4 * i + j] = m[i][4 + j];
  return RVK_OK;
}

void rvk_prebake_view_to_voxels(const double m[16], float out[16]) {
  for (int i = 0; i < 16; ++i) out[i] = (float)m[i];
  for (int r = 0; r < 4; ++r) {
    out[4 * r + 3] += out[4 * r + 0] - out[4 * r + 1];
    out[4 * r + 0] *= -2.0f;
    out[4 * r + 1] *= 2.0f;
  }
}

int rvk_compute_matrices(const rvk_camera* cam, const double user[16], const double origin[3],
                         const double spacing[3], double aspect, double v2vox[16], double vox2v[16]) {
  if (!cam || !origin || !spacing || !(aspect > 0.0)) return RVK_ERR_INVALID_ARGUMENT;
  if (!cam->parallel_projection && !(cam->clipping_range[0] > 0.0)) return RVK_ERR_INVALID_ARGUMENT;
  if (!(cam->clipping_range[1] > cam->clipping_range[0])) return RVK_ERR_INVALID_ARGUMENT;
  double p[16], v[16], pv[16], w[16], full[16];
  projection_matrix(cam, aspect, p);
  view_matrix(cam, v);
  rvk_mat4_mul(p, v, pv);
  voxels_to_world(user, origin, spacing, w);
  rvk_mat4_mul(pv, w, full);
  if (vox2v) std::memcpy(vox2v, full, sizeof(full));
  if (v2vox) return rvk_mat4_invert(full, v2vox);
  return RVK_OK;
}

int rvk_compute_view(const rvk_camera* cam, const double user[16], const double origin[3], const double spacing[3],
                     int width, int height, int flipped, rvk_view* view) {
  if (!view || width <= 0 || height <= 0) return RVK_ERR_INVALID_ARGUMENT;
  double m[16];
  int rc = rvk_compute_matrices(cam, user, origin, spacing, (double)width / (double)height, m, nullptr);
  if (rc) return rc;
  std::memset(view, 0, sizeof(*view));
  rvk_prebake_view_to_voxels(m, view->view_to_voxels);
  view->width = width;
  view->height = height;
  view->flipped = flipped;
  return RVK_OK;
}

}  // extern "C"
