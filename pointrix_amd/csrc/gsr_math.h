// Per-splat math of the preprocessing stage: camera transform and pinhole
// projection, 3D covariance from scale + quaternion, EWA 2D covariance /
// conic / radius, real spherical harmonics -- forward (exact FP32 order) and
// vector-Jacobian products for the backward pass.
#pragma once
#include "gsr_common.h"

namespace gsr {

// Camera values, read by every thread from the caller's device arrays
// intr = (fx, fy, cx, cy) and extr = 3x4 row-major [R|t] (scalar-cached
// broadcast loads; no host round trip, so launches stay graph-capturable).
struct CamV {
  float fx, fy, cx, cy;
  float R[9];
  float t[3];
};

__device__ __forceinline__ CamV cam_load(const float *__restrict__ intr,
                                         const float *__restrict__ extr) {
  CamV c;
  c.fx = intr[0];
  c.fy = intr[1];
  c.cx = intr[2];
  c.cy = intr[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) c.R[3 * i + j] = extr[4 * i + j];
    c.t[i] = extr[4 * i + 3];
  }
  return c;
}

__device__ __forceinline__ float3 cam_point(const CamV &c, float3 x) {
  float3 p;
  p.x = c.R[0] * x.x + c.R[1] * x.y + c.R[2] * x.z + c.t[0];
  p.y = c.R[3] * x.x + c.R[4] * x.y + c.R[5] * x.z + c.t[1];
  p.z = c.R[6] * x.x + c.R[7] * x.y + c.R[8] * x.z + c.t[2];
  return p;
}

// ---------------------------------------------------------------------------
// project_point
// ---------------------------------------------------------------------------
__device__ __forceinline__ void project(const CamV &c, float3 x, float nearest, float extent,
                                        int W, int H, float &u, float &v, float &depth) {
  const float3 p = cam_point(c, x);
  u = 0.0f;
  v = 0.0f;
  depth = 0.0f;
  if (!(p.z > nearest)) return;
  const float pu = c.fx * p.x / p.z + c.cx;
  const float pv = c.fy * p.y / p.z + c.cy;
  const float lo_x = 0.5f * (float)W * (1.0f - extent);
  const float hi_x = 0.5f * (float)W * (1.0f + extent);
  const float lo_y = 0.5f * (float)H * (1.0f - extent);
  const float hi_y = 0.5f * (float)H * (1.0f + extent);
  if (pu < lo_x || pu > hi_x || pv < lo_y || pv > hi_y) return;
  u = pu;
  v = pv;
  depth = p.z;
}

// gcam layout: [0..3] = d(fx, fy, cx, cy), [4 + 4*i + j] = d extr[i][j].
__device__ __forceinline__ void project_vjp(const CamV &c, float3 x, float gu, float gv,
                                            float gd, float3 &gx, float (&gcam)[16]) {
  const float3 p = cam_point(c, x);
  const float iz = 1.0f / p.z;
  const float gpx = gu * c.fx * iz;
  const float gpy = gv * c.fy * iz;
  const float gpz = -(gu * c.fx * p.x + gv * c.fy * p.y) * iz * iz + gd;
  gx.x += c.R[0] * gpx + c.R[3] * gpy + c.R[6] * gpz;
  gx.y += c.R[1] * gpx + c.R[4] * gpy + c.R[7] * gpz;
  gx.z += c.R[2] * gpx + c.R[5] * gpy + c.R[8] * gpz;
  gcam[0] += gu * p.x * iz;
  gcam[1] += gv * p.y * iz;
  gcam[2] += gu;
  gcam[3] += gv;
  const float gp[3] = {gpx, gpy, gpz};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    gcam[4 + 4 * i + 0] += gp[i] * x.x;
    gcam[4 + 4 * i + 1] += gp[i] * x.y;
    gcam[4 + 4 * i + 2] += gp[i] * x.z;
    gcam[4 + 4 * i + 3] += gp[i];
  }
}

// ---------------------------------------------------------------------------
// compute_cov3d: R(q) from a wxyz quaternion with the homogeneous formula of
// pointrix/utils/pose.py:40-84, Sigma = (R S)(R S)^T.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void quat_rotmat(float4 q, float (&R)[9]) {
  const float w = q.x, x = q.y, y = q.z, z = q.w;
  const float x2 = x * x, y2 = y * y, z2 = z * z, w2 = w * w;
  const float xy = x * y, zw = z * w, xz = x * z, yw = y * w, yz = y * z, xw = x * w;
  R[0] = x2 - y2 - z2 + w2;
  R[1] = 2.0f * (xy - zw);
  R[2] = 2.0f * (xz + yw);
  R[3] = 2.0f * (xy + zw);
  R[4] = -x2 + y2 - z2 + w2;
  R[5] = 2.0f * (yz - xw);
  R[6] = 2.0f * (xz - yw);
  R[7] = 2.0f * (yz + xw);
  R[8] = -x2 - y2 + z2 + w2;
}

__device__ __forceinline__ void covariance3d(float3 s, float4 q, float (&cov)[6]) {
  float R[9];
  quat_rotmat(q, R);
  float M[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    M[3 * i + 0] = R[3 * i + 0] * s.x;
    M[3 * i + 1] = R[3 * i + 1] * s.y;
    M[3 * i + 2] = R[3 * i + 2] * s.z;
  }
  // (0,0) (0,1) (0,2) (1,1) (1,2) (2,2)
  const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int i = ii[k], j = jj[k];
    cov[k] = M[3 * i] * M[3 * j] + M[3 * i + 1] * M[3 * j + 1] + M[3 * i + 2] * M[3 * j + 2];
  }
}

__device__ __forceinline__ void covariance3d_vjp(float3 s, float4 q, const float (&g)[6],
                                                 float (&gs)[3], float (&gq)[4]) {
  float R[9];
  quat_rotmat(q, R);
  const float sv[3] = {s.x, s.y, s.z};
  float M[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) M[3 * i + j] = R[3 * i + j] * sv[j];
  // symmetric gradient of the packed output
  const float G[9] = {g[0], 0.5f * g[1], 0.5f * g[2], 0.5f * g[1], g[3],
                      0.5f * g[4], 0.5f * g[2], 0.5f * g[4], g[5]};
  float dR[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float dM = 2.0f * (G[3 * i] * M[k] + G[3 * i + 1] * M[3 + k] + G[3 * i + 2] * M[6 + k]);
      dR[3 * i + k] = dM * sv[k];
    }
  // gs_k = sum_i dM_ik R_ik
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float acc = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float dM = 2.0f * (G[3 * i] * M[k] + G[3 * i + 1] * M[3 + k] + G[3 * i + 2] * M[6 + k]);
      acc += dM * R[3 * i + k];
    }
    gs[k] = acc;
  }
  const float w = q.x, x = q.y, y = q.z, z = q.w;
  // d R / d (w, x, y, z), row-major entries
  gq[0] = 2.0f * (w * dR[0] - z * dR[1] + y * dR[2] + z * dR[3] + w * dR[4] - x * dR[5] -
                  y * dR[6] + x * dR[7] + w * dR[8]);
  gq[1] = 2.0f * (x * dR[0] + y * dR[1] + z * dR[2] + y * dR[3] - x * dR[4] - w * dR[5] +
                  z * dR[6] + w * dR[7] - x * dR[8]);
  gq[2] = 2.0f * (-y * dR[0] + x * dR[1] + w * dR[2] + x * dR[3] + y * dR[4] + z * dR[5] -
                  w * dR[6] + z * dR[7] - y * dR[8]);
  gq[3] = 2.0f * (-z * dR[0] - w * dR[1] + x * dR[2] + w * dR[3] - z * dR[4] + y * dR[5] +
                  x * dR[6] + y * dR[7] + z * dR[8]);
}

// ---------------------------------------------------------------------------
// ewa_project: 2D covariance (with the 0.3 px low-pass), conic, radius.
// ---------------------------------------------------------------------------
struct EwaFwd {
  float3 p;
  float z, limx, limy, txz, tyz, tx, ty;
  float J00, J02, J11, J12;
  float T[6];  // T = J R, 2x3 row-major
  float U[6];  // U = T Sigma
  float a, b, c;
};

__device__ __forceinline__ void ewa_forward(const CamV &cam, float3 x, const float (&S)[6], int W,
                                            int H, EwaFwd &e) {
  e.p = cam_point(cam, x);
  e.z = e.p.z;
  e.limx = kJClamp * (0.5f * (float)W / cam.fx);
  e.limy = kJClamp * (0.5f * (float)H / cam.fy);
  e.txz = fminf(fmaxf(e.p.x / e.z, -e.limx), e.limx);
  e.tyz = fminf(fmaxf(e.p.y / e.z, -e.limy), e.limy);
  e.tx = e.txz * e.z;
  e.ty = e.tyz * e.z;
  e.J00 = cam.fx / e.z;
  e.J02 = -(cam.fx * e.tx) / (e.z * e.z);
  e.J11 = cam.fy / e.z;
  e.J12 = -(cam.fy * e.ty) / (e.z * e.z);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    e.T[j] = e.J00 * cam.R[j] + e.J02 * cam.R[6 + j];
    e.T[3 + j] = e.J11 * cam.R[3 + j] + e.J12 * cam.R[6 + j];
  }
  const float Sf[9] = {S[0], S[1], S[2], S[1], S[3], S[4], S[2], S[4], S[5]};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      e.U[3 * i + j] = e.T[3 * i] * Sf[j] + e.T[3 * i + 1] * Sf[3 + j] + e.T[3 * i + 2] * Sf[6 + j];
  e.a = e.U[0] * e.T[0] + e.U[1] * e.T[1] + e.U[2] * e.T[2];
  e.b = e.U[0] * e.T[3] + e.U[1] * e.T[4] + e.U[2] * e.T[5];
  e.c = e.U[3] * e.T[3] + e.U[4] * e.T[4] + e.U[5] * e.T[5];
  e.a = e.a + kLowPass;
  e.c = e.c + kLowPass;
}

// conic = inverse 2D covariance; radius = ceil(3 sqrt(lambda_max)).
__device__ __forceinline__ bool conic_radius(float a, float b, float c, float (&con)[3], int &r) {
  const float det = a * c - b * b;
  if (!(det > 0.0f)) return false;
  const float inv = 1.0f / det;
  con[0] = c * inv;
  con[1] = -b * inv;
  con[2] = a * inv;
  const float mid = 0.5f * (a + c);
  const float disc = fmaxf(mid * mid - det, 0.1f);
  const float lam = mid + sqrtf(disc);
  r = (int)ceilf(3.0f * sqrtf(lam));
  return true;
}

__device__ __forceinline__ void ewa_vjp(const CamV &cam, float3 x, const float (&S)[6], int W,
                                        int H, const float (&gcon)[3], float3 &gx,
                                        float (&gS)[6], float (&gcam)[16]) {
  EwaFwd e;
  ewa_forward(cam, x, S, W, H, e);
  const float a = e.a, b = e.b, c = e.c;
  const float det = a * c - b * b;
  const float id2 = 1.0f / (det * det);
  const float gA = gcon[0], gB = gcon[1], gC = gcon[2];
  const float ga = (-c * c * gA + b * c * gB + (det - a * c) * gC) * id2;
  const float gb = (2.0f * b * c * gA + (-det - 2.0f * b * b) * gB + 2.0f * a * b * gC) * id2;
  const float gc = ((det - a * c) * gA + a * b * gB - a * a * gC) * id2;
  // symmetric 2x2 gradient of (a, b, c)
  const float G00 = ga, G01 = 0.5f * gb, G11 = gc;
  // dSigma = T^T G T (packed, off-diagonals doubled)
  float dSf[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const float t0i = e.T[i], t1i = e.T[3 + i], t0j = e.T[j], t1j = e.T[3 + j];
      dSf[3 * i + j] = t0i * (G00 * t0j + G01 * t1j) + t1i * (G01 * t0j + G11 * t1j);
    }
  gS[0] = dSf[0];
  gS[1] = 2.0f * dSf[1];
  gS[2] = 2.0f * dSf[2];
  gS[3] = dSf[4];
  gS[4] = 2.0f * dSf[5];
  gS[5] = dSf[8];
  // dT = 2 G U
  float dT[6];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    dT[j] = 2.0f * (G00 * e.U[j] + G01 * e.U[3 + j]);
    dT[3 + j] = 2.0f * (G01 * e.U[j] + G11 * e.U[3 + j]);
  }
  // T = J R
  float dR[9];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    dR[j] = e.J00 * dT[j];
    dR[3 + j] = e.J11 * dT[3 + j];
    dR[6 + j] = e.J02 * dT[j] + e.J12 * dT[3 + j];
  }
  const float dJ00 = dT[0] * cam.R[0] + dT[1] * cam.R[1] + dT[2] * cam.R[2];
  const float dJ02 = dT[0] * cam.R[6] + dT[1] * cam.R[7] + dT[2] * cam.R[8];
  const float dJ11 = dT[3] * cam.R[3] + dT[4] * cam.R[4] + dT[5] * cam.R[5];
  const float dJ12 = dT[3] * cam.R[6] + dT[4] * cam.R[7] + dT[5] * cam.R[8];
  const float z = e.z, iz = 1.0f / z, iz2 = iz * iz, iz3 = iz2 * iz;
  float gfx = dJ00 * iz - dJ02 * e.tx * iz2;
  float gfy = dJ11 * iz - dJ12 * e.ty * iz2;
  float gz = -dJ00 * cam.fx * iz2 + 2.0f * dJ02 * cam.fx * e.tx * iz3 - dJ11 * cam.fy * iz2 +
             2.0f * dJ12 * cam.fy * e.ty * iz3;
  const float gtx = -dJ02 * cam.fx * iz2;
  const float gty = -dJ12 * cam.fy * iz2;
  float gpx = 0.0f, gpy = 0.0f;
  const float rx = e.p.x / e.z, ry = e.p.y / e.z;
  if (rx >= -e.limx && rx <= e.limx) {
    gpx = gtx;  // tx = p.x
  } else {
    const float sg = rx > 0.0f ? 1.0f : -1.0f;  // tx = sg * limx * z
    gz += gtx * sg * e.limx;
    gfx += gtx * sg * z * (-e.limx / cam.fx);
  }
  if (ry >= -e.limy && ry <= e.limy) {
    gpy = gty;
  } else {
    const float sg = ry > 0.0f ? 1.0f : -1.0f;
    gz += gty * sg * e.limy;
    gfy += gty * sg * z * (-e.limy / cam.fy);
  }
  const float gp[3] = {gpx, gpy, gz};
  gx.x += cam.R[0] * gp[0] + cam.R[3] * gp[1] + cam.R[6] * gp[2];
  gx.y += cam.R[1] * gp[0] + cam.R[4] * gp[1] + cam.R[7] * gp[2];
  gx.z += cam.R[2] * gp[0] + cam.R[5] * gp[1] + cam.R[8] * gp[2];
  gcam[0] += gfx;
  gcam[1] += gfy;
  const float xv[3] = {x.x, x.y, x.z};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) gcam[4 + 4 * i + j] += dR[3 * i + j] + gp[i] * xv[j];
    gcam[4 + 4 * i + 3] += gp[i];
  }
}

// ---------------------------------------------------------------------------
// Real spherical harmonics, degrees 0..3 (3DGS sign convention).  Constants
// are the closed-form normalisations, e.g. C0 = 1/(2 sqrt(pi)),
// C1 = sqrt(3/(4 pi)), C2_0 = sqrt(15/pi)/2, C3_0 = sqrt(35/(2 pi))/4.
// ---------------------------------------------------------------------------
constexpr float kC0 = 0.28209479177387814f;
constexpr float kC1 = 0.48860251190291992f;
constexpr float kC2_0 = 1.0925484305920792f;
constexpr float kC2_2 = 0.31539156525252005f;
constexpr float kC2_4 = 0.54627421529603959f;
constexpr float kC3_0 = 0.59004358992664352f;
constexpr float kC3_1 = 2.8906114426405538f;
constexpr float kC3_2 = 0.45704579946446572f;
constexpr float kC3_3 = 0.37317633259011540f;
constexpr float kC3_5 = 1.4453057213202769f;

__device__ __forceinline__ void sh_basis(float3 d, int K, float (&Y)[16]) {
  const float x = d.x, y = d.y, z = d.z;
  Y[0] = kC0;
  if (K > 1) {
    Y[1] = -kC1 * y;
    Y[2] = kC1 * z;
    Y[3] = -kC1 * x;
  }
  if (K > 4) {
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, yz = y * z, xz = x * z;
    Y[4] = kC2_0 * xy;
    Y[5] = -kC2_0 * yz;
    Y[6] = kC2_2 * (2.0f * zz - xx - yy);
    Y[7] = -kC2_0 * xz;
    Y[8] = kC2_4 * (xx - yy);
    if (K > 9) {
      Y[9] = -kC3_0 * y * (3.0f * xx - yy);
      Y[10] = kC3_1 * xy * z;
      Y[11] = -kC3_2 * y * (4.0f * zz - xx - yy);
      Y[12] = kC3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
      Y[13] = -kC3_2 * x * (4.0f * zz - xx - yy);
      Y[14] = kC3_5 * z * (xx - yy);
      Y[15] = -kC3_0 * x * (xx - 3.0f * yy);
    }
  }
}

// gd += sum_k coef[k] * dY_k/d(x, y, z)
__device__ __forceinline__ void sh_basis_vjp(float3 d, int K, const float (&coef)[16], float3 &gd) {
  const float x = d.x, y = d.y, z = d.z;
  float gx = 0.0f, gy = 0.0f, gz = 0.0f;
  if (K > 1) {
    gy += -kC1 * coef[1];
    gz += kC1 * coef[2];
    gx += -kC1 * coef[3];
  }
  if (K > 4) {
    const float xx = x * x, yy = y * y, zz = z * z;
    gx += kC2_0 * y * coef[4];
    gy += kC2_0 * x * coef[4];
    gy += -kC2_0 * z * coef[5];
    gz += -kC2_0 * y * coef[5];
    gx += -2.0f * kC2_2 * x * coef[6];
    gy += -2.0f * kC2_2 * y * coef[6];
    gz += 4.0f * kC2_2 * z * coef[6];
    gx += -kC2_0 * z * coef[7];
    gz += -kC2_0 * x * coef[7];
    gx += 2.0f * kC2_4 * x * coef[8];
    gy += -2.0f * kC2_4 * y * coef[8];
    if (K > 9) {
      // Y9 = -C3_0 (3 x^2 y - y^3)
      gx += -kC3_0 * 6.0f * x * y * coef[9];
      gy += -kC3_0 * (3.0f * xx - 3.0f * yy) * coef[9];
      // Y10 = C3_1 x y z
      gx += kC3_1 * y * z * coef[10];
      gy += kC3_1 * x * z * coef[10];
      gz += kC3_1 * x * y * coef[10];
      // Y11 = -C3_2 (4 y z^2 - x^2 y - y^3)
      gx += -kC3_2 * (-2.0f * x * y) * coef[11];
      gy += -kC3_2 * (4.0f * zz - xx - 3.0f * yy) * coef[11];
      gz += -kC3_2 * (8.0f * y * z) * coef[11];
      // Y12 = C3_3 (2 z^3 - 3 x^2 z - 3 y^2 z)
      gx += kC3_3 * (-6.0f * x * z) * coef[12];
      gy += kC3_3 * (-6.0f * y * z) * coef[12];
      gz += kC3_3 * (6.0f * zz - 3.0f * xx - 3.0f * yy) * coef[12];
      // Y13 = -C3_2 (4 x z^2 - x^3 - x y^2)
      gx += -kC3_2 * (4.0f * zz - 3.0f * xx - yy) * coef[13];
      gy += -kC3_2 * (-2.0f * x * y) * coef[13];
      gz += -kC3_2 * (8.0f * x * z) * coef[13];
      // Y14 = C3_5 (x^2 z - y^2 z)
      gx += kC3_5 * (2.0f * x * z) * coef[14];
      gy += kC3_5 * (-2.0f * y * z) * coef[14];
      gz += kC3_5 * (xx - yy) * coef[14];
      // Y15 = -C3_0 (x^3 - 3 x y^2)
      gx += -kC3_0 * (3.0f * xx - 3.0f * yy) * coef[15];
      gy += -kC3_0 * (-6.0f * x * y) * coef[15];
    }
  }
  gd.x += gx;
  gd.y += gy;
  gd.z += gz;
}

// ---------------------------------------------------------------------------
// Per-splat normal channel (examples/supervise/model.py:50-74): the axis of
// the smallest scale, m = R(q) e_k, normalised, turned towards the camera,
// rotated into camera space by the extrinsic rotation.
// ---------------------------------------------------------------------------
struct NormalFwd {
  int k;        // index of the smallest scale (first on ties)
  float sgn;    // +1 or -1 (faces the camera)
  float m[3];   // R(q) column k
  float inv;    // 1 / |m|
  float nh[3];  // sgn * m / |m| (world space)
};

__device__ __forceinline__ void splat_normal(float3 s, float4 q, float3 x, const float (&center)[3],
                                             const CamV &cam, NormalFwd &nf, float (&ncam)[3]) {
  float R[9];
  quat_rotmat(q, R);
  int k = 0;
  float smin = s.x;
  if (s.y < smin) { k = 1; smin = s.y; }
  if (s.z < smin) { k = 2; }
  nf.k = k;
  nf.m[0] = R[k];
  nf.m[1] = R[3 + k];
  nf.m[2] = R[6 + k];
  const float len = sqrtf(nf.m[0] * nf.m[0] + nf.m[1] * nf.m[1] + nf.m[2] * nf.m[2]);
  nf.inv = 1.0f / fmaxf(len, 1e-12f);
  const float dot = (center[0] - x.x) * nf.m[0] + (center[1] - x.y) * nf.m[1] +
                    (center[2] - x.z) * nf.m[2];
  nf.sgn = dot < 0.0f ? -1.0f : 1.0f;
#pragma unroll
  for (int i = 0; i < 3; ++i) nf.nh[i] = nf.sgn * (nf.m[i] * nf.inv);
#pragma unroll
  for (int i = 0; i < 3; ++i)
    ncam[i] = cam.R[3 * i] * nf.nh[0] + cam.R[3 * i + 1] * nf.nh[1] + cam.R[3 * i + 2] * nf.nh[2];
}

// g (camera-space normal gradient) -> quaternion gradient (+=) and extrinsic
// rotation gradient gcam[4 + 4i + j] (+=).
__device__ __forceinline__ void splat_normal_vjp(float4 q, const CamV &cam, const NormalFwd &nf,
                                                 const float (&g)[3], float (&gq)[4],
                                                 float (&gcam)[16]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) gcam[4 + 4 * i + j] += g[i] * nf.nh[j];
  float gn[3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
    gn[j] = nf.sgn * (cam.R[j] * g[0] + cam.R[3 + j] * g[1] + cam.R[6 + j] * g[2]);
  // through m / |m|
  const float u0 = nf.m[0] * nf.inv, u1 = nf.m[1] * nf.inv, u2 = nf.m[2] * nf.inv;
  const float dp = u0 * gn[0] + u1 * gn[1] + u2 * gn[2];
  float dR[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  dR[nf.k] = (gn[0] - u0 * dp) * nf.inv;
  dR[3 + nf.k] = (gn[1] - u1 * dp) * nf.inv;
  dR[6 + nf.k] = (gn[2] - u2 * dp) * nf.inv;
  const float w = q.x, x = q.y, y = q.z, z = q.w;
  gq[0] += 2.0f * (w * dR[0] - z * dR[1] + y * dR[2] + z * dR[3] + w * dR[4] - x * dR[5] -
                   y * dR[6] + x * dR[7] + w * dR[8]);
  gq[1] += 2.0f * (x * dR[0] + y * dR[1] + z * dR[2] + y * dR[3] - x * dR[4] - w * dR[5] +
                   z * dR[6] + w * dR[7] - x * dR[8]);
  gq[2] += 2.0f * (-y * dR[0] + x * dR[1] + w * dR[2] + x * dR[3] + y * dR[4] + z * dR[5] -
                   w * dR[6] + z * dR[7] - y * dR[8]);
  gq[3] += 2.0f * (-z * dR[0] - w * dR[1] + x * dR[2] + w * dR[3] - z * dR[4] + y * dR[5] +
                   x * dR[6] + y * dR[7] + z * dR[8]);
}

// Camera-gradient block reduction: gcam[0..3] -> d_intr, gcam[4..15] -> d_extr.
// Must be reached by every thread of the block.
__device__ __forceinline__ void block_cam_grad(float (&gcam)[16], float *d_intr, float *d_extr) {
  __shared__ float s_red[16][16];
  const int wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const float s = wave_sum(gcam[k]);
    if (lane_id() == 0) s_red[wid][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < 16) {
    float s = 0.0f;
    for (int w = 0; w < nw; ++w) s += s_red[w][threadIdx.x];
    if (s != 0.0f) {
      if (threadIdx.x < 4) {
        if (d_intr) atomicAdd(d_intr + threadIdx.x, s);
      } else if (d_extr) {
        atomicAdd(d_extr + threadIdx.x - 4, s);
      }
    }
  }
}

}  // namespace gsr
