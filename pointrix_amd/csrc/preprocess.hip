// Per-splat preprocessing kernels: project_point, compute_cov3d, ewa_project,
// compute_sh -- forward and backward.  One thread per splat; inputs and outputs
// are the dense row-major arrays msplat's call sites pass
// (pointrix/model/renderer/msplat.py:104-130).  These kernels stream their
// inputs once and are HBM-bound; the forward arithmetic order is part of the
// bit-reproducibility contract described in gsr_common.h.
#include "gsr_internal.h"
#include "gsr_math.h"

using namespace gsr;

namespace {

constexpr int kThreads = 256;

// --------------------------------------------------------------------------
// project_point
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void project_fwd_kernel(
    const float *__restrict__ xyz, int64_t n, const float *__restrict__ intr, const float *__restrict__ extr, float nearest, float extent, int W,
    int H, float *__restrict__ uv, float *__restrict__ depth) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float3 x = make_float3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  const CamV cam = cam_load(intr, extr);
  float u, v, d;
  project(cam, x, nearest, extent, W, H, u, v, d);
  uv[2 * i] = u;
  uv[2 * i + 1] = v;
  depth[i] = d;
}

__global__ __launch_bounds__(kThreads) void project_bwd_kernel(
    const float *__restrict__ xyz, int64_t n, const float *__restrict__ intr, const float *__restrict__ extr, const float *__restrict__ depth,
    const float *__restrict__ d_uv, const float *__restrict__ d_depth,
    float *__restrict__ d_xyz, float *__restrict__ d_intr, float *__restrict__ d_extr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float gcam[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) gcam[k] = 0.0f;
  if (i < n) {
    float3 gx = make_float3(0.f, 0.f, 0.f);
    if (depth[i] != 0.0f) {
      const float3 x = make_float3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
      const CamV cam = cam_load(intr, extr);
      project_vjp(cam, x, d_uv[2 * i], d_uv[2 * i + 1], d_depth ? d_depth[i] : 0.0f, gx,
                  gcam);
    }
    d_xyz[3 * i] = gx.x;
    d_xyz[3 * i + 1] = gx.y;
    d_xyz[3 * i + 2] = gx.z;
  }
  if (d_intr != nullptr || d_extr != nullptr) block_cam_grad(gcam, d_intr, d_extr);
}

// --------------------------------------------------------------------------
// compute_cov3d
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void cov3d_fwd_kernel(
    const float *__restrict__ scale, const float *__restrict__ rot,
    const uint8_t *__restrict__ visible, int64_t n, float *__restrict__ cov3d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float c[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (visible[i]) {
    const float3 s = make_float3(scale[3 * i], scale[3 * i + 1], scale[3 * i + 2]);
    const float4 q = make_float4(rot[4 * i], rot[4 * i + 1], rot[4 * i + 2], rot[4 * i + 3]);
    covariance3d(s, q, c);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) cov3d[6 * i + k] = c[k];
}

__global__ __launch_bounds__(kThreads) void cov3d_bwd_kernel(
    const float *__restrict__ scale, const float *__restrict__ rot,
    const uint8_t *__restrict__ visible, int64_t n, const float *__restrict__ d_cov3d,
    float *__restrict__ d_scale, float *__restrict__ d_rot) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gs[3] = {0.f, 0.f, 0.f}, gq[4] = {0.f, 0.f, 0.f, 0.f};
  if (visible[i]) {
    const float3 s = make_float3(scale[3 * i], scale[3 * i + 1], scale[3 * i + 2]);
    const float4 q = make_float4(rot[4 * i], rot[4 * i + 1], rot[4 * i + 2], rot[4 * i + 3]);
    float g[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) g[k] = d_cov3d[6 * i + k];
    covariance3d_vjp(s, q, g, gs, gq);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) d_scale[3 * i + k] = gs[k];
#pragma unroll
  for (int k = 0; k < 4; ++k) d_rot[4 * i + k] = gq[k];
}

// --------------------------------------------------------------------------
// ewa_project
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void ewa_fwd_kernel(
    const float *__restrict__ xyz, const float *__restrict__ cov3d, int64_t n, const float *__restrict__ intr, const float *__restrict__ extr,
    const float *__restrict__ uv, int W, int H, int gx, int gy,
    const uint8_t *__restrict__ visible, float *__restrict__ conic,
    int32_t *__restrict__ radius, int32_t *__restrict__ tiles) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float con[3] = {0.f, 0.f, 0.f};
  int r = 0, t = 0;
  if (visible[i]) {
    const float3 x = make_float3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    float S[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) S[k] = cov3d[6 * i + k];
    const CamV cam = cam_load(intr, extr);
    EwaFwd e;
    ewa_forward(cam, x, S, W, H, e);
    int rr;
    float cc[3];
    if (conic_radius(e.a, e.b, e.c, cc, rr)) {
      const Rect q = tile_rect(uv[2 * i], uv[2 * i + 1], rr, gx, gy, 16, 16);
      const int area = (q.x1 - q.x0) * (q.y1 - q.y0);
      if (area > 0) {
        r = rr;
        t = area;
        con[0] = cc[0];
        con[1] = cc[1];
        con[2] = cc[2];
      }
    }
  }
  conic[3 * i] = con[0];
  conic[3 * i + 1] = con[1];
  conic[3 * i + 2] = con[2];
  radius[i] = r;
  tiles[i] = t;
}

__global__ __launch_bounds__(kThreads) void ewa_bwd_kernel(
    const float *__restrict__ xyz, const float *__restrict__ cov3d, int64_t n, const float *__restrict__ intr, const float *__restrict__ extr, int W,
    int H, const int32_t *__restrict__ radius, const float *__restrict__ d_conic,
    float *__restrict__ d_xyz, float *__restrict__ d_cov3d, float *__restrict__ d_intr,
    float *__restrict__ d_extr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float gcam[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) gcam[k] = 0.0f;
  if (i < n) {
    float3 gx = make_float3(0.f, 0.f, 0.f);
    float gS[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (radius[i] > 0) {
      const float3 x = make_float3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
      float S[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) S[k] = cov3d[6 * i + k];
      const float g3[3] = {d_conic[3 * i], d_conic[3 * i + 1], d_conic[3 * i + 2]};
      const CamV cam = cam_load(intr, extr);
      ewa_vjp(cam, x, S, W, H, g3, gx, gS, gcam);
    }
    d_xyz[3 * i] = gx.x;
    d_xyz[3 * i + 1] = gx.y;
    d_xyz[3 * i + 2] = gx.z;
#pragma unroll
    for (int k = 0; k < 6; ++k) d_cov3d[6 * i + k] = gS[k];
  }
  if (d_intr != nullptr || d_extr != nullptr) block_cam_grad(gcam, d_intr, d_extr);
}

// --------------------------------------------------------------------------
// compute_sh
// --------------------------------------------------------------------------
template <int KE>
__global__ __launch_bounds__(kThreads) void sh_fwd_kernel(
    const float *__restrict__ shs, int64_t s_n, int64_t s_c, int64_t s_k,
    const float *__restrict__ dirs, int64_t n, int C, float *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float Y[16];
  sh_basis(make_float3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), KE, Y);
  const float *base = shs + i * s_n;
  for (int c = 0; c < C; ++c) {
    const float *sc = base + c * s_c;
    float acc = sc[0] * Y[0];
#pragma unroll
    for (int k = 1; k < KE; ++k) acc = acc + sc[k * s_k] * Y[k];
    out[i * C + c] = acc;
  }
}

template <int KE>
__global__ __launch_bounds__(kThreads) void sh_bwd_kernel(
    const float *__restrict__ shs, int64_t s_n, int64_t s_c, int64_t s_k,
    const float *__restrict__ dirs, int64_t n, int C, int Kfull,
    const float *__restrict__ d_out, float *__restrict__ d_shs, float *__restrict__ d_dirs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float3 d = make_float3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
  float Y[16];
  sh_basis(d, KE, Y);
  float3 gd = make_float3(0.f, 0.f, 0.f);
  const float *base = shs + i * s_n;
  for (int c = 0; c < C; ++c) {
    const float g = d_out[i * C + c];
    float *dst = d_shs + (i * C + c) * Kfull;
#pragma unroll
    for (int k = 0; k < KE; ++k) dst[k] = g * Y[k];
    for (int k = KE; k < Kfull; ++k) dst[k] = 0.0f;
    if (KE > 1 && d_dirs != nullptr) {
      float coef[16];
      const float *sc = base + c * s_c;
#pragma unroll
      for (int k = 0; k < 16; ++k) coef[k] = k < KE ? g * sc[(k < KE ? k : 0) * s_k] : 0.0f;
      sh_basis_vjp(d, KE, coef, gd);
    }
  }
  if (d_dirs != nullptr) {
    d_dirs[3 * i] = gd.x;
    d_dirs[3 * i + 1] = gd.y;
    d_dirs[3 * i + 2] = gd.z;
  }
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

int gsr_project_point_fwd(const float *xyz, int64_t n, const float *intr, const float *extr,
                          int W, int H, float nearest, float extent, float *uv, float *depth,
                          void *stream) {
  GSR_REQUIRE(n >= 0 && W > 0 && H > 0, "project_point: bad sizes");
  if (n == 0) return GSR_OK;
  project_fwd_kernel<<<gsr_blocks(n, kThreads), kThreads, 0, (hipStream_t)stream>>>(
      xyz, n, intr, extr, nearest, extent, W, H, uv, depth);
  return gsr_launch_status("project_point_fwd");
}

int gsr_project_point_bwd(const float *xyz, int64_t n, const float *intr, const float *extr,
                          const float *depth, const float *d_uv, const float *d_depth,
                          float *d_xyz, float *d_intr, float *d_extr, void *stream) {
  GSR_REQUIRE(n >= 0, "project_point_bwd: bad sizes");
  if (n == 0) return GSR_OK;
  project_bwd_kernel<<<gsr_blocks(n, kThreads), kThreads, 0, (hipStream_t)stream>>>(
      xyz, n, intr, extr, depth, d_uv, d_depth, d_xyz, d_intr, d_extr);
  return gsr_launch_status("project_point_bwd");
}

int gsr_compute_cov3d_fwd(const float *scale, const float *rot, const uint8_t *visible,
                          int64_t n, float *cov3d, void *stream) {
  GSR_REQUIRE(n >= 0, "compute_cov3d: bad sizes");
  if (n == 0) return GSR_OK;
  cov3d_fwd_kernel<<<gsr_blocks(n, kThreads), kThreads, 0, (hipStream_t)stream>>>(
      scale, rot, visible, n, cov3d);
  return gsr_launch_status("compute_cov3d_fwd");
}

int gsr_compute_cov3d_bwd(const float *scale, const float *rot, const uint8_t *visible,
                          int64_t n, const float *d_cov3d, float *d_scale, float *d_rot,
                          void *stream) {
  GSR_REQUIRE(n >= 0, "compute_cov3d_bwd: bad sizes");
  if (n == 0) return GSR_OK;
  cov3d_bwd_kernel<<<gsr_blocks(n, kThreads), kThreads, 0, (hipStream_t)stream>>>(
      scale, rot, visible, n, d_cov3d, d_scale, d_rot);
  return gsr_launch_status("compute_cov3d_bwd");
}

int gsr_ewa_project_fwd(const float *xyz, const float *cov3d, int64_t n, const float *intr,
                        const float *extr, const float *uv, int W, int H,
                        const uint8_t *visible, float *conic, int32_t *radius,
                        int32_t *tiles, void *stream) {
  GSR_REQUIRE(n >= 0 && W > 0 && H > 0, "ewa_project: bad sizes");
  if (n == 0) return GSR_OK;
  const int gx = (W + 15) / 16, gy = (H + 15) / 16;
  ewa_fwd_kernel<<<gsr_blocks(n, kThreads), kThreads, 0, (hipStream_t)stream>>>(
      xyz, cov3d, n, intr, extr, uv, W, H, gx, gy, visible, conic, radius, tiles);
  return gsr_launch_status("ewa_project_fwd");
}

int gsr_ewa_project_bwd(const float *xyz, const float *cov3d, int64_t n, const float *intr,
                        const float *extr, int W, int H, const int32_t *radius,
                        const float *d_conic, float *d_xyz, float *d_cov3d, float *d_intr,
                        float *d_extr, void *stream) {
  GSR_REQUIRE(n >= 0 && W > 0 && H > 0, "ewa_project_bwd: bad sizes");
  if (n == 0) return GSR_OK;
  ewa_bwd_kernel<<<gsr_blocks(n, kThreads), kThreads, 0, (hipStream_t)stream>>>(
      xyz, cov3d, n, intr, extr, W, H, radius, d_conic, d_xyz, d_cov3d, d_intr, d_extr);
  return gsr_launch_status("ewa_project_bwd");
}

int gsr_compute_sh_fwd(const float *shs, int64_t s_n, int64_t s_c, int64_t s_k,
                       const float *dirs, int64_t n, int C, int K, int degree, float *out,
                       void *stream) {
  GSR_REQUIRE(n >= 0 && C > 0 && (K == 1 || K == 4 || K == 9 || K == 16),
              "compute_sh: K must be 1, 4, 9 or 16 (got %d)", K);
  if (n == 0) return GSR_OK;
  const int Ke = degree < 0 ? K : (K < (degree + 1) * (degree + 1) ? K : (degree + 1) * (degree + 1));
  const unsigned nb = gsr_blocks(n, kThreads);
  hipStream_t st = (hipStream_t)stream;
  switch (Ke) {
    case 1: sh_fwd_kernel<1><<<nb, kThreads, 0, st>>>(shs, s_n, s_c, s_k, dirs, n, C, out); break;
    case 4: sh_fwd_kernel<4><<<nb, kThreads, 0, st>>>(shs, s_n, s_c, s_k, dirs, n, C, out); break;
    case 9: sh_fwd_kernel<9><<<nb, kThreads, 0, st>>>(shs, s_n, s_c, s_k, dirs, n, C, out); break;
    default: sh_fwd_kernel<16><<<nb, kThreads, 0, st>>>(shs, s_n, s_c, s_k, dirs, n, C, out);
  }
  return gsr_launch_status("compute_sh_fwd");
}

int gsr_compute_sh_bwd(const float *shs, int64_t s_n, int64_t s_c, int64_t s_k,
                       const float *dirs, int64_t n, int C, int K, int degree,
                       const float *d_out, float *d_shs, float *d_dirs, void *stream) {
  GSR_REQUIRE(n >= 0 && C > 0 && (K == 1 || K == 4 || K == 9 || K == 16),
              "compute_sh_bwd: K must be 1, 4, 9 or 16 (got %d)", K);
  if (n == 0) return GSR_OK;
  const int Ke = degree < 0 ? K : (K < (degree + 1) * (degree + 1) ? K : (degree + 1) * (degree + 1));
  const unsigned nb = gsr_blocks(n, kThreads);
  hipStream_t st = (hipStream_t)stream;
#define GSR_SHB(KE_) \
  sh_bwd_kernel<KE_><<<nb, kThreads, 0, st>>>(shs, s_n, s_c, s_k, dirs, n, C, K, d_out, d_shs, d_dirs)
  switch (Ke) {
    case 1: GSR_SHB(1); break;
    case 4: GSR_SHB(4); break;
    case 9: GSR_SHB(9); break;
    default: GSR_SHB(16);
  }
#undef GSR_SHB
  return gsr_launch_status("compute_sh_bwd");
}

}  // extern "C"
