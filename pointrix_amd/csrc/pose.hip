// Camera pose composition of pointrix's CameraModel (camera_model.py:92-175
// with unitquat_to_rotmat, utils/pose.py:40-84): a wxyz quaternion q,
// normalised as F.normalize does (q / max(|q|, 1e-12)), and a translation t
// give the world-to-camera matrix E = [R | t; 0 0 0 1] and the camera centre
// -R^T t.  Forward and backward, one thread per camera: one launch each
// instead of the few dozen scalar tensor operations (and their backward
// nodes) of the unfused composition.
#include "gsr_internal.h"

using namespace gsr;

namespace {

struct Quat {
  float w, x, y, z;
};

__device__ __forceinline__ void quat_rotmat(const Quat &q, float (&R)[3][3]) {
  const float x2 = q.x * q.x, y2 = q.y * q.y, z2 = q.z * q.z, w2 = q.w * q.w;
  const float xy = q.x * q.y, zw = q.z * q.w, xz = q.x * q.z, yw = q.y * q.w, yz = q.y * q.z,
              xw = q.x * q.w;
  R[0][0] = x2 - y2 - z2 + w2;
  R[1][0] = 2.0f * (xy + zw);
  R[2][0] = 2.0f * (xz - yw);
  R[0][1] = 2.0f * (xy - zw);
  R[1][1] = -x2 + y2 - z2 + w2;
  R[2][1] = 2.0f * (yz + xw);
  R[0][2] = 2.0f * (xz + yw);
  R[1][2] = 2.0f * (yz - xw);
  R[2][2] = -x2 - y2 + z2 + w2;
}

__device__ __forceinline__ Quat load_unit(const float *q, float &norm) {
  const float a = q[0], b = q[1], c = q[2], d = q[3];
  norm = sqrtf(((a * a + b * b) + c * c) + d * d);
  const float s = fmaxf(norm, 1e-12f);
  return Quat{a / s, b / s, c / s, d / s};
}

__global__ void pose_fwd_kernel(const float *__restrict__ q, const float *__restrict__ t, int C,
                                float *__restrict__ E, float *__restrict__ center) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= C) return;
  float n;
  const Quat u = load_unit(q + 4 * i, n);
  float R[3][3];
  quat_rotmat(u, R);
  const float tv[3] = {t[3 * i], t[3 * i + 1], t[3 * i + 2]};
  float *e = E + 16 * i;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
#pragma unroll
    for (int c = 0; c < 3; ++c) e[4 * r + c] = R[r][c];
    e[4 * r + 3] = tv[r];
  }
  e[12] = 0.0f;
  e[13] = 0.0f;
  e[14] = 0.0f;
  e[15] = 1.0f;
  if (center != nullptr) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
      center[3 * i + c] = -((R[0][c] * tv[0] + R[1][c] * tv[1]) + R[2][c] * tv[2]);
  }
}

// dR -> d(unit quaternion): the partial derivatives of the nine entries of
// quat_rotmat with respect to (w, x, y, z).
__device__ __forceinline__ void rotmat_vjp(const Quat &q, const float (&g)[3][3], float (&gq)[4]) {
  const float w = 2.0f * q.w, x = 2.0f * q.x, y = 2.0f * q.y, z = 2.0f * q.z;
  gq[0] = g[0][0] * w - g[0][1] * z + g[0][2] * y + g[1][0] * z + g[1][1] * w - g[1][2] * x -
          g[2][0] * y + g[2][1] * x + g[2][2] * w;
  gq[1] = g[0][0] * x + g[0][1] * y + g[0][2] * z + g[1][0] * y - g[1][1] * x - g[1][2] * w +
          g[2][0] * z + g[2][1] * w - g[2][2] * x;
  gq[2] = -g[0][0] * y + g[0][1] * x + g[0][2] * w + g[1][0] * x + g[1][1] * y + g[1][2] * z -
          g[2][0] * w + g[2][1] * z - g[2][2] * y;
  gq[3] = -g[0][0] * z - g[0][1] * w + g[0][2] * x + g[1][0] * w - g[1][1] * z + g[1][2] * y +
          g[2][0] * x + g[2][1] * y + g[2][2] * z;
}

__global__ void pose_bwd_kernel(const float *__restrict__ q, const float *__restrict__ t, int C,
                                const float *__restrict__ dE, const float *__restrict__ dcenter,
                                float *__restrict__ dq, float *__restrict__ dt) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= C) return;
  float n;
  const Quat u = load_unit(q + 4 * i, n);
  float R[3][3];
  quat_rotmat(u, R);
  const float tv[3] = {t[3 * i], t[3 * i + 1], t[3 * i + 2]};
  float gR[3][3], gt[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int r = 0; r < 3; ++r) {
#pragma unroll
    for (int c = 0; c < 3; ++c) gR[r][c] = dE != nullptr ? dE[16 * i + 4 * r + c] : 0.0f;
    if (dE != nullptr) gt[r] = dE[16 * i + 4 * r + 3];
  }
  if (dcenter != nullptr) {
    // centre_c = -sum_r R[r][c] t_r
    const float dc[3] = {dcenter[3 * i], dcenter[3 * i + 1], dcenter[3 * i + 2]};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) gR[r][c] -= tv[r] * dc[c];
      gt[r] -= (R[r][0] * dc[0] + R[r][1] * dc[1]) + R[r][2] * dc[2];
    }
  }
  float gu[4];
  rotmat_vjp(u, gR, gu);
  // through q / max(|q|, 1e-12): (g - u (u.g)) / |q|, or g / 1e-12 when clamped
  const float dot = ((u.w * gu[0] + u.x * gu[1]) + u.y * gu[2]) + u.z * gu[3];
  const float uq[4] = {u.w, u.x, u.y, u.z};
#pragma unroll
  for (int k = 0; k < 4; ++k)
    dq[4 * i + k] = n > 1e-12f ? (gu[k] - uq[k] * dot) / n : gu[k] / 1e-12f;
#pragma unroll
  for (int r = 0; r < 3; ++r) dt[3 * i + r] = gt[r];
}

}  // namespace

extern "C" {

int gsr_camera_pose_fwd(const float *q, const float *t, int C, float *E, float *center,
                        void *stream) {
  GSR_REQUIRE(C >= 0, "camera_pose: bad size");
  if (C == 0) return GSR_OK;
  pose_fwd_kernel<<<gsr_blocks(C, 64), 64, 0, (hipStream_t)stream>>>(q, t, C, E, center);
  return gsr_launch_status("camera_pose_fwd");
}

int gsr_camera_pose_bwd(const float *q, const float *t, int C, const float *dE,
                        const float *dcenter, float *dq, float *dt, void *stream) {
  GSR_REQUIRE(C >= 0, "camera_pose_bwd: bad size");
  if (C == 0) return GSR_OK;
  pose_bwd_kernel<<<gsr_blocks(C, 64), 64, 0, (hipStream_t)stream>>>(q, t, C, dE, dcenter, dq,
                                                                    dt);
  return gsr_launch_status("camera_pose_bwd");
}

}  // extern "C"
