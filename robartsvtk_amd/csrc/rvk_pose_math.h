// rvk_pose_math.h -- the posed view geometry, one implementation for the
// host (rvk_posed_view) and the device (a plan's on-device pose updates).
//
// A volume at rigid pose (rx, ry, rz [rad], tx, ty, tz [mm]) about a centre c
// has the user matrix U = T(c + t) Rz Ry Rx T(-c); the view's screen->voxel
// matrix is M = inverse(PV U Vox) (vtkCudaVolumeMapper::ComputeMatrices,
// V/vtkCudaVolumeMapper.cxx:349-447, V/ = Visualization/vtkCudaVisualization),
// pre-baked for pixel coordinates as in
// V/vtkCudaRendererInformationHandler.cxx:112-139, and its pose derivatives
// dM_i = -M (PV dU_i Vox) M, baked the same way (the bake is linear).
// Everything is double precision until the baked float matrices, with the
// same operations in the same order on both sides.
#pragma once
#include <hip/hip_runtime.h>

#include <math.h>

namespace rvk {
namespace pose {

__host__ __device__ inline void identity(double m[16]) {
  for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0 : 0.0;
}

__host__ __device__ inline void mul(const double a[16], const double b[16], double out[16]) {
  double r[16];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += a[4 * i + k] * b[4 * k + j];
      r[4 * i + j] = s;
    }
  for (int i = 0; i < 16; ++i) out[i] = r[i];
}

// Gauss-Jordan elimination with partial pivoting; false if singular.
__host__ __device__ inline bool invert(const double a[16], double out[16]) {
  double m[4][8];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      m[i][j] = a[4 * i + j];
      m[i][4 + j] = (i == j) ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; ++c) {
    int piv = c;
    for (int r = c + 1; r < 4; ++r)
      if (fabs(m[r][c]) > fabs(m[piv][c])) piv = r;
    if (m[piv][c] == 0.0) return false;
    if (piv != c)
      for (int j = 0; j < 8; ++j) {
        const double t = m[c][j];
        m[c][j] = m[piv][j];
        m[piv][j] = t;
      }
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
  return true;
}

// Rotation about one axis and its derivative with respect to the angle.
__host__ __device__ inline void rot(int axis, double ang, double m[16], double dm[16]) {
  identity(m);
  for (int i = 0; i < 16; ++i) dm[i] = 0.0;
  const double c = cos(ang), s = sin(ang);
  const int a = (axis + 1) % 3, b = (axis + 2) % 3;
  m[5 * a] = c; m[4 * a + b] = -s; m[4 * b + a] = s; m[5 * b] = c;
  dm[5 * a] = -s; dm[4 * a + b] = -c; dm[4 * b + a] = c; dm[5 * b] = -s;
}

// U = T(c + t) Rz Ry Rx T(-c) and (du != nullptr) dU/dpose_i, 6 x 16.
__host__ __device__ inline void pose_matrix(const double pose[6], const double c[3], double out[16], double* du) {
  double rx[16], ry[16], rz[16], drx[16], dry[16], drz[16];
  rot(0, pose[0], rx, drx);
  rot(1, pose[1], ry, dry);
  rot(2, pose[2], rz, drz);
  double tc[16], tmc[16];
  identity(tc);
  identity(tmc);
  for (int a = 0; a < 3; ++a) {
    tc[4 * a + 3] = c[a] + pose[3 + a];
    tmc[4 * a + 3] = -c[a];
  }
  auto chain = [&](const double* a, const double* b, const double* cc, double* o) {
    double t1[16], t2[16], t3[16];
    mul(a, b, t1);
    mul(t1, cc, t2);
    mul(tc, t2, t3);
    mul(t3, tmc, o);
  };
  chain(rz, ry, rx, out);
  if (du) {
    chain(rz, ry, drx, du + 0);
    chain(rz, dry, rx, du + 16);
    chain(drz, ry, rx, du + 32);
    // translation: d/dt_a of T(c+t) R T(-c) = E_a3 (R T(-c) has last row e4)
    for (int a = 0; a < 3; ++a) {
      double* d = du + 16 * (3 + a);
      for (int i = 0; i < 16; ++i) d[i] = 0.0;
      d[4 * a + 3] = 1.0;
    }
  }
}

// The reference's pixel pre-bake, in float: M[r][3] += M[r][0] - M[r][1];
// M[r][0] *= -2; M[r][1] *= 2.
__host__ __device__ inline void bake(const double m[16], float out[16]) {
  for (int i = 0; i < 16; ++i) out[i] = (float)m[i];
  for (int r = 0; r < 4; ++r) {
    out[4 * r + 3] += out[4 * r + 0] - out[4 * r + 1];
    out[4 * r + 0] *= -2.0f;
    out[4 * r + 1] *= 2.0f;
  }
}

// Baked view matrix (m_out) and its baked pose derivatives (dm_out, 6 x 16,
// nullable) for the camera transform pv = P V, the voxel->world matrix vox
// and the pose about c.  False if the posed matrix is singular.
__host__ __device__ inline bool posed_view(const double pv[16], const double vox[16], const double pose[6],
                                           const double c[3], float m_out[16], float* dm_out) {
  double u[16], du[96];
  pose_matrix(pose, c, u, dm_out ? du : nullptr);
  double t[16], a[16], m[16];
  mul(pv, u, t);
  mul(t, vox, a);
  if (!invert(a, m)) return false;
  bake(m, m_out);
  if (dm_out) {
    for (int i = 0; i < 6; ++i) {
      double t1[16], da[16], t2[16], dm[16];
      mul(pv, du + 16 * i, t1);
      mul(t1, vox, da);
      mul(m, da, t2);
      mul(t2, m, dm);
      for (int k = 0; k < 16; ++k) dm[k] = -dm[k];
      bake(dm, dm_out + 16 * i);
    }
  }
  return true;
}

// Inverse of a baked float matrix (screen->voxel): voxel position ->
// (vx, vy, z, 1) * w, as floats; zero if singular.
__host__ __device__ inline void baked_inverse(const float mb[16], float out[16]) {
  double m[16], mi[16];
  for (int k = 0; k < 16; ++k) m[k] = mb[k];
  const bool ok = invert(m, mi);
  for (int k = 0; k < 16; ++k) out[k] = ok ? (float)mi[k] : 0.0f;
}

}  // namespace pose
}  // namespace rvk
