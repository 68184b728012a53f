// Fused per-splat preprocessing for MsplatRender.render_iter
// (pointrix/model/renderer/msplat.py:94-139): view direction, SH colour
// (+0.5, clamped at 0), project_point, compute_cov3d, ewa_project and the
// blend feature pack (rgb | depth | camera-space normal) in one pass over the
// splats, and the matching backward that chains every vector-Jacobian
// product in registers.  Same per-op arithmetic (gsr_math.h) as the separate
// msplat ops, so uv / depth / cov3d / conic / radius are identical to theirs.
//
// Layouts: shs is the [N, KS, 3] array render_iter receives (coefficient
// major, channel minor), or the pair DC [N, 1, 3] + rest [N, KS-1, 3]; only
// the first KE = min(KS, (degree+1)^2) coefficients are read.  feat is [N, C] with C = 3 (+1 depth) (+3 normal).
#include "gsr_internal.h"
#include "gsr_math.h"

using namespace gsr;

namespace {

constexpr int kThreads = 256;
constexpr int kFlagDepth = 1;
constexpr int kFlagNormal = 2;
// scale, rotation and opacity arrive as the model's raw parameters (log
// scale, unnormalised quaternion, opacity logit): the kernels apply exp,
// normalize and sigmoid (GaussianPointCloud's activations,
// gaussian_points.py:30-35,70-81) and return gradients w.r.t. the raw values
constexpr int kFlagRaw = 4;
// colour = the SH sum itself: no +0.5 offset, no clamp (the normal-supervision
// renderer of examples/supervise/renderer.py:41-42)
constexpr int kFlagPlain = 8;

__device__ __forceinline__ float3 act_scale(float3 s) {
  return make_float3(expf(s.x), expf(s.y), expf(s.z));
}

// torch.nn.functional.normalize: q / max(|q|, 1e-12)
__device__ __forceinline__ float4 act_rot(float4 q, float &norm) {
  norm = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  const float d = fmaxf(norm, 1e-12f);
  return make_float4(q.x / d, q.y / d, q.z / d, q.w / d);
}

__device__ __forceinline__ float act_opacity(float o) { return 1.0f / (1.0f + expf(-o)); }

// Load the first KE coefficients (x3 channels) of splat i, either from one
// [N, KS, 3] array or, when rest != nullptr, from the DC array [N, 1, 3] and
// the higher-order array [N, KS-1, 3] (pointrix keeps them as two
// parameters, gaussian_points.py:86-90; reading both avoids concatenating).
template <int KS, int KE>
__device__ __forceinline__ void load_sh(const float *__restrict__ shs,
                                        const float *__restrict__ rest, int64_t i,
                                        float (&sh)[KE * 3]) {
  if (rest != nullptr) {
    const float3 dc = ld_f3(shs + 3 * i);
    sh[0] = dc.x;
    sh[1] = dc.y;
    sh[2] = dc.z;
    // the rest row (3(K-1) floats) as 16-byte accesses plus the tail
    const float *rb = rest + i * ((KS - 1) * 3);
    constexpr int nr = KE * 3 - 3;
#pragma unroll
    for (int j = 0; j + 4 <= nr; j += 4) {
      const float4 t = ld_f4u(rb + j);
      sh[3 + j] = t.x;
      sh[3 + j + 1] = t.y;
      sh[3 + j + 2] = t.z;
      sh[3 + j + 3] = t.w;
    }
#pragma unroll
    for (int j = nr & ~3; j < nr; ++j) sh[3 + j] = rb[j];
    return;
  }
  const float *base = shs + i * (KS * 3);
  if constexpr ((KS * 3) % 4 == 0) {
    constexpr int n4 = (KE * 3 + 3) / 4;
    const float4 *b4 = reinterpret_cast<const float4 *>(base);
#pragma unroll
    for (int j = 0; j < n4; ++j) {
      const float4 t = b4[j];
      if (4 * j + 0 < KE * 3) sh[4 * j + 0] = t.x;
      if (4 * j + 1 < KE * 3) sh[4 * j + 1] = t.y;
      if (4 * j + 2 < KE * 3) sh[4 * j + 2] = t.z;
      if (4 * j + 3 < KE * 3) sh[4 * j + 3] = t.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < KE * 3; ++j) sh[j] = base[j];
  }
}

__device__ __forceinline__ float3 view_dir(float3 x, const float (&c)[3], float &len) {
  const float dx = x.x - c[0], dy = x.y - c[1], dz = x.z - c[2];
  len = sqrtf(dx * dx + dy * dy + dz * dz);
  return make_float3(dx / len, dy / len, dz / len);
}

template <int KS, int KE>
__global__ __launch_bounds__(kThreads) void preprocess_fwd_kernel(
    const float *__restrict__ xyz, const float *__restrict__ scale,
    const float *__restrict__ rot, const float *__restrict__ shs,
    const float *__restrict__ shs_rest, int64_t n, const float *__restrict__ intr,
    const float *__restrict__ extr, const float *__restrict__ center, int W, int H, int gx, int gy,
    float nearest,
    float extent, int flags, int C, float *__restrict__ uv, float *__restrict__ depth,
    float *__restrict__ cov_out, float *__restrict__ conic, int32_t *__restrict__ radius,
    int32_t *__restrict__ tiles, float *__restrict__ feat, const float *__restrict__ opac_raw,
    float *__restrict__ opac_act) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const CamV cam = cam_load(intr, extr);
  const float3 x = ld_f3(xyz + 3 * i);
  // requested with the position; stored once both have arrived
  const float o_raw = opac_act != nullptr ? opac_raw[i] : 0.0f;
  float u, v, d;
  project(cam, x, nearest, extent, W, H, u, v, d);
  st_f2(uv + 2 * i, u, v);
  depth[i] = d;
  if (opac_act != nullptr) opac_act[i] = act_opacity(o_raw);
  float con[3] = {0.f, 0.f, 0.f};
  float cov[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float f[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int r = 0, t = 0;
  if (d != 0.0f) {
    float3 s = ld_f3(scale + 3 * i);
    float4 q = ld_f4u(rot + 4 * i);
    // the SH coefficients are requested together with scale and rotation
    // (one memory round trip instead of two); almost every splat inside the
    // frustum is also rasterised
    float sh[KE * 3];
    load_sh<KS, KE>(shs, shs_rest, i, sh);
    if (flags & kFlagRaw) {
      float qn;
      s = act_scale(s);
      q = act_rot(q, qn);
    }
    covariance3d(s, q, cov);
    EwaFwd e;
    ewa_forward(cam, x, cov, W, H, e);
    int rr;
    float cc[3];
    if (conic_radius(e.a, e.b, e.c, cc, rr)) {
      const Rect qr = tile_rect(u, v, rr, gx, gy, 16, 16);
      const int area = (qr.x1 - qr.x0) * (qr.y1 - qr.y0);
      if (area > 0) {
        r = rr;
        t = area;
        con[0] = cc[0];
        con[1] = cc[1];
        con[2] = cc[2];
      }
    }
    if (r > 0) {
      const float cen[3] = {center[0], center[1], center[2]};
      float len;
      const float3 dir = view_dir(x, cen, len);
      float Y[16];
      sh_basis(dir, KE, Y);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float acc = sh[c] * Y[0];
#pragma unroll
        for (int k = 1; k < KE; ++k) acc = acc + sh[3 * k + c] * Y[k];
        f[c] = (flags & kFlagPlain) ? acc : fmaxf(acc + 0.5f, 0.0f);
      }
      int o = 3;
      if (flags & kFlagDepth) f[o++] = d;
      if (flags & kFlagNormal) {
        NormalFwd nf;
        float nc[3];
        splat_normal(s, q, x, cen, cam, nf, nc);
        f[o] = nc[0];
        f[o + 1] = nc[1];
        f[o + 2] = nc[2];
      }
    }
  }
  if (cov_out != nullptr) {
#pragma unroll
    for (int k = 0; k < 6; ++k) cov_out[6 * i + k] = cov[k];
  }
  st_f3(conic + 3 * i, con[0], con[1], con[2]);
  radius[i] = r;
  tiles[i] = t;
  if (C == 7) {  // rgb | depth | normal: one 16-byte and one 12-byte store
    st_f4u(feat + 7 * i, f[0], f[1], f[2], f[3]);
    st_f3(feat + 7 * i + 4, f[4], f[5], f[6]);
  } else {
    for (int c = 0; c < C; ++c) feat[i * C + c] = f[c];
  }
}

template <int KS, int KE>
__global__ __launch_bounds__(kThreads) void preprocess_bwd_kernel(
    const float *__restrict__ xyz, const float *__restrict__ scale,
    const float *__restrict__ rot, const float *__restrict__ shs,
    const float *__restrict__ shs_rest, int64_t n, const float *__restrict__ intr,
    const float *__restrict__ extr, const float *__restrict__ center, int W, int H,
    float nearest, float extent, int flags,
    int C, const int32_t *__restrict__ radius, const float *__restrict__ d_uv, int ld_uv,
    const float *__restrict__ d_depth, const float *__restrict__ d_conic, int ld_conic,
    const float *__restrict__ d_feat, int ld_feat, float *__restrict__ d_xyz, float *__restrict__ d_scale,
    float *__restrict__ d_rot, float *__restrict__ d_shs, float *__restrict__ d_shs_rest,
    float *__restrict__ d_intr, float *__restrict__ d_extr, float *__restrict__ d_center,
    const float *__restrict__ opac_raw, const float *__restrict__ d_opac_act, int ld_oact,
    float *__restrict__ d_opac_raw) {
  constexpr int kRest = (KS - 1) * 3;
  __shared__ __align__(16) float s_rest[KS > 1 ? kThreads * kRest : 4];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float gcam[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) gcam[k] = 0.0f;
  float gcen[3] = {0.f, 0.f, 0.f};
  if (i < n) {
    const CamV cam = cam_load(intr, extr);
    float3 gx = make_float3(0.f, 0.f, 0.f);
    float gs[3] = {0.f, 0.f, 0.f}, gq[4] = {0.f, 0.f, 0.f, 0.f};
    float dsh[KE * 3];
#pragma unroll
    for (int j = 0; j < KE * 3; ++j) dsh[j] = 0.0f;
    const float3 x = ld_f3(xyz + 3 * i);
    // inputs that do not depend on visibility are requested with the position
    const float g_depth_in = d_depth != nullptr ? d_depth[i] : 0.0f;
    const float o_raw = d_opac_raw != nullptr ? opac_raw[i] : 0.0f;
    const float g_oact = d_opac_raw != nullptr ? d_opac_act[i * ld_oact] : 0.0f;
    float3 s_act = make_float3(1.f, 1.f, 1.f);
    float4 q_act = make_float4(1.f, 0.f, 0.f, 0.f);
    float q_norm = 1.0f;
    if (radius[i] > 0) {
      // every input of the splat is requested up front (one memory round trip)
      float3 s = ld_f3(scale + 3 * i);
      float4 q = ld_f4u(rot + 4 * i);
      float sh[KE * 3];
      load_sh<KS, KE>(shs, shs_rest, i, sh);
      float gf[7];
      if (C == 7) {  // rgb | depth | normal: one 16-byte and one 12-byte access
        const float4 a = ld_f4u(d_feat + i * ld_feat);
        const float3 b = ld_f3(d_feat + i * ld_feat + 4);
        gf[0] = a.x;
        gf[1] = a.y;
        gf[2] = a.z;
        gf[3] = a.w;
        gf[4] = b.x;
        gf[5] = b.y;
        gf[6] = b.z;
      } else {
#pragma unroll
        for (int c = 0; c < 7; ++c) gf[c] = c < C ? d_feat[i * ld_feat + c] : 0.0f;
      }
      const float2 guv2 = ld_f2(d_uv + i * ld_uv);
      const float guv[2] = {guv2.x, guv2.y};
      const float3 gc3 = ld_f3(d_conic + i * ld_conic);
      const float g3[3] = {gc3.x, gc3.y, gc3.z};
      if (flags & kFlagRaw) {
        s = act_scale(s);
        q = act_rot(q, q_norm);
        s_act = s;
        q_act = q;
      }
      const float cen[3] = {center[0], center[1], center[2]};
      // colour: SH -> +0.5 -> clamp(min=0) (gradient passes where x >= 0)
      float len;
      const float3 dir = view_dir(x, cen, len);
      float Y[16];
      sh_basis(dir, KE, Y);
      float coef[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) coef[k] = 0.0f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float acc = sh[c] * Y[0];
#pragma unroll
        for (int k = 1; k < KE; ++k) acc = acc + sh[3 * k + c] * Y[k];
        const float g = ((flags & kFlagPlain) || acc + 0.5f >= 0.0f) ? gf[c] : 0.0f;
#pragma unroll
        for (int k = 0; k < KE; ++k) {
          dsh[3 * k + c] = g * Y[k];
          coef[k] += g * sh[3 * k + c];
        }
      }
      if (KE > 1) {
        float3 gdir = make_float3(0.f, 0.f, 0.f);
        sh_basis_vjp(dir, KE, coef, gdir);
        const float dp = dir.x * gdir.x + dir.y * gdir.y + dir.z * gdir.z;
        const float il = 1.0f / len;
        const float3 gd = make_float3((gdir.x - dir.x * dp) * il, (gdir.y - dir.y * dp) * il,
                                      (gdir.z - dir.z * dp) * il);
        gx.x += gd.x;
        gx.y += gd.y;
        gx.z += gd.z;
        gcen[0] -= gd.x;
        gcen[1] -= gd.y;
        gcen[2] -= gd.z;
      }
      // channel layout rgb | depth (flag 1) | normal (flag 2)
      const bool has_d = (flags & kFlagDepth) != 0;
      float gdepth = g_depth_in;
      if (has_d) gdepth += gf[3];
      if (flags & kFlagNormal) {
        NormalFwd nf;
        float nc[3];
        splat_normal(s, q, x, cen, cam, nf, nc);
        const float gn[3] = {has_d ? gf[4] : gf[3], has_d ? gf[5] : gf[4],
                             has_d ? gf[6] : gf[5]};
        splat_normal_vjp(q, cam, nf, gn, gq, gcam);
      }
      project_vjp(cam, x, guv[0], guv[1], gdepth, gx, gcam);
      float cov[6];
      covariance3d(s, q, cov);
      float gS[6];
      ewa_vjp(cam, x, cov, W, H, g3, gx, gS, gcam);
      float gs2[3], gq2[4];
      covariance3d_vjp(s, q, gS, gs2, gq2);
#pragma unroll
      for (int k = 0; k < 3; ++k) gs[k] += gs2[k];
#pragma unroll
      for (int k = 0; k < 4; ++k) gq[k] += gq2[k];
    } else if (d_depth != nullptr) {
      // projected but not rasterised: only a direct depth gradient applies
      float u, v, dd;
      project(cam, x, nearest, extent, W, H, u, v, dd);
      if (dd != 0.0f) project_vjp(cam, x, 0.0f, 0.0f, g_depth_in, gx, gcam);
    }
    st_f3(d_xyz + 3 * i, gx.x, gx.y, gx.z);
    if (flags & kFlagRaw) {
      // through exp: ds_raw = ds * s; through normalize: (g - q (q.g)) / |q|
      // (1 / 1e-12 where the norm is clamped); through sigmoid: s (1 - s)
      gs[0] *= s_act.x;
      gs[1] *= s_act.y;
      gs[2] *= s_act.z;
      const float qa[4] = {q_act.x, q_act.y, q_act.z, q_act.w};
      const float dot = qa[0] * gq[0] + qa[1] * gq[1] + qa[2] * gq[2] + qa[3] * gq[3];
      if (q_norm > 1e-12f) {
#pragma unroll
        for (int k = 0; k < 4; ++k) gq[k] = (gq[k] - qa[k] * dot) / q_norm;
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) gq[k] = gq[k] / 1e-12f;
      }
      if (d_opac_raw != nullptr) {
        const float a = act_opacity(o_raw);
        d_opac_raw[i] = g_oact * (a * (1.0f - a));
      }
    }
    st_f3(d_scale + 3 * i, gs[0], gs[1], gs[2]);
    st_f4u(d_rot + 4 * i, gq[0], gq[1], gq[2], gq[3]);
    if (d_shs_rest != nullptr) {
      st_f3(d_shs + 3 * i, dsh[0], dsh[1], dsh[2]);
      if constexpr (KS > 1) {
        // this splat's row of the rest gradient, staged for the block's
        // coalesced store below
#pragma unroll
        for (int j = 3; j < KS * 3; ++j)
          s_rest[threadIdx.x * kRest + (j - 3)] = (j < KE * 3) ? dsh[j % (KE * 3)] : 0.0f;
      }
    } else {
      float *ds = d_shs + i * (KS * 3);
      if constexpr ((KS * 3) % 4 == 0) {
        float4 *d4 = reinterpret_cast<float4 *>(ds);
#pragma unroll
        for (int j = 0; j < KS * 3 / 4; ++j) {
          float4 t;
          t.x = (4 * j + 0 < KE * 3) ? dsh[(4 * j + 0) % (KE * 3)] : 0.0f;
          t.y = (4 * j + 1 < KE * 3) ? dsh[(4 * j + 1) % (KE * 3)] : 0.0f;
          t.z = (4 * j + 2 < KE * 3) ? dsh[(4 * j + 2) % (KE * 3)] : 0.0f;
          t.w = (4 * j + 3 < KE * 3) ? dsh[(4 * j + 3) % (KE * 3)] : 0.0f;
          d4[j] = t;
        }
      } else {
#pragma unroll
        for (int j = 0; j < KS * 3; ++j) ds[j] = (j < KE * 3) ? dsh[j % (KE * 3)] : 0.0f;
      }
    }
  }
  if constexpr (KS > 1) {
    // The block's rows of d_shs_rest are one contiguous range: write it with
    // consecutive lanes on consecutive 16-byte words (the rows are 12(K-1)
    // bytes, so per-lane row stores would touch a new cache line per lane).
    if (d_shs_rest != nullptr) {
      __syncthreads();
      const int64_t first = (int64_t)blockIdx.x * kThreads;
      const int rows = (int)min((int64_t)kThreads, n - first);
      const int total = rows * kRest;
      float *dst = d_shs_rest + first * kRest;
      if ((reinterpret_cast<uintptr_t>(dst) & 15) == 0) {
        float4 *d4 = reinterpret_cast<float4 *>(dst);
        const float4 *s4 = reinterpret_cast<const float4 *>(s_rest);
        for (int k = threadIdx.x; k < total / 4; k += kThreads) d4[k] = s4[k];
        for (int k = (total / 4) * 4 + threadIdx.x; k < total; k += kThreads) dst[k] = s_rest[k];
      } else {
        for (int k = threadIdx.x; k < total; k += kThreads) dst[k] = s_rest[k];
      }
    }
  }
  if (d_intr != nullptr || d_extr != nullptr) block_cam_grad(gcam, d_intr, d_extr);
  if (d_center != nullptr) {
    __shared__ float s_c[16][3];
    __syncthreads();
    const int wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float sum = wave_sum(gcen[k]);
      if (lane_id() == 0) s_c[wid][k] = sum;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      float sum = 0.0f;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) sum += s_c[w][threadIdx.x];
      if (sum != 0.0f) atomicAdd(d_center + threadIdx.x, sum);
    }
  }
}

template <int KS>
int launch_fwd_ks(int KE, unsigned nb, hipStream_t st, const float *xyz, const float *scale,
                  const float *rot, const float *shs, const float *shs_rest, int64_t n, const float *intr,
                  const float *extr, const float *center, int W, int H, int gx, int gy,
                  float nearest, float extent, int flags, int C, float *uv, float *depth,
                  float *cov_out, float *conic, int32_t *radius, int32_t *tiles, float *feat,
                  const float *opac_raw, float *opac_act) {
#define GSR_FWD_ARGS                                                                          \
  xyz, scale, rot, shs, shs_rest, n, intr, extr, center, W, H, gx, gy, nearest, extent, flags, \
      C, uv, depth, cov_out, conic, radius, tiles, feat, opac_raw, opac_act
  switch (KE) {
    case 1: preprocess_fwd_kernel<KS, 1><<<nb, kThreads, 0, st>>>(GSR_FWD_ARGS); break;
    case 4: if constexpr (KS >= 4) { preprocess_fwd_kernel<KS, 4><<<nb, kThreads, 0, st>>>(GSR_FWD_ARGS); break; } return GSR_ERR_ARG;
    case 9: if constexpr (KS >= 9) { preprocess_fwd_kernel<KS, 9><<<nb, kThreads, 0, st>>>(GSR_FWD_ARGS); break; } return GSR_ERR_ARG;
    case 16: if constexpr (KS >= 16) { preprocess_fwd_kernel<KS, 16><<<nb, kThreads, 0, st>>>(GSR_FWD_ARGS); break; } return GSR_ERR_ARG;
    default: return GSR_ERR_ARG;
  }
#undef GSR_FWD_ARGS
  return GSR_OK;
}

template <int KS>
int launch_bwd_ks(int KE, unsigned nb, hipStream_t st, const float *xyz, const float *scale,
                  const float *rot, const float *shs, const float *shs_rest, int64_t n, const float *intr,
                  const float *extr, const float *center, int W, int H, float nearest,
                  float extent, int flags, int C, const int32_t *radius, const float *d_uv,
                  int ld_uv, const float *d_depth, const float *d_conic, int ld_conic,
                  const float *d_feat, int ld_feat, float *d_xyz,
                  float *d_scale, float *d_rot, float *d_shs, float *d_shs_rest, float *d_intr,
                  float *d_extr, float *d_center,
                  const float *opac_raw, const float *d_opac_act, int ld_oact,
                  float *d_opac_raw) {
#define GSR_BWD_ARGS                                                                          \
  xyz, scale, rot, shs, shs_rest, n, intr, extr, center, W, H, nearest, extent, flags, C,      \
      radius, d_uv, ld_uv, d_depth, d_conic, ld_conic, d_feat, ld_feat, d_xyz, d_scale, d_rot,   \
      d_shs, d_shs_rest, d_intr, d_extr, d_center, opac_raw, d_opac_act, ld_oact, d_opac_raw
  switch (KE) {
    case 1: preprocess_bwd_kernel<KS, 1><<<nb, kThreads, 0, st>>>(GSR_BWD_ARGS); break;
    case 4: if constexpr (KS >= 4) { preprocess_bwd_kernel<KS, 4><<<nb, kThreads, 0, st>>>(GSR_BWD_ARGS); break; } return GSR_ERR_ARG;
    case 9: if constexpr (KS >= 9) { preprocess_bwd_kernel<KS, 9><<<nb, kThreads, 0, st>>>(GSR_BWD_ARGS); break; } return GSR_ERR_ARG;
    case 16: if constexpr (KS >= 16) { preprocess_bwd_kernel<KS, 16><<<nb, kThreads, 0, st>>>(GSR_BWD_ARGS); break; } return GSR_ERR_ARG;
    default: return GSR_ERR_ARG;
  }
#undef GSR_BWD_ARGS
  return GSR_OK;
}

inline int effective_k(int K, int degree) {
  if (degree < 0) return K;
  const int k = (degree + 1) * (degree + 1);
  return k < K ? k : K;
}

inline int feat_channels(int flags) {
  return 3 + ((flags & kFlagDepth) ? 1 : 0) + ((flags & kFlagNormal) ? 3 : 0);
}

}  // namespace

extern "C" {

int gsr_preprocess_fwd(const float *xyz, const float *scale, const float *rot, const float *shs,
                       const float *shs_rest, int64_t n, int K, int degree, const float *intr, const float *extr,
                       const float *cam_center, int W, int H, float nearest, float extent,
                       int flags, float *uv, float *depth, float *cov3d, float *conic,
                       int32_t *radius, int32_t *tiles, float *feat, const float *opacity_raw,
                       float *opacity_act, void *stream) {
  GSR_REQUIRE(n >= 0 && W > 0 && H > 0 && (K == 1 || K == 4 || K == 9 || K == 16) &&
                  (flags & ~15) == 0 && ((opacity_raw == nullptr) == (opacity_act == nullptr)) &&
                  (opacity_raw == nullptr || (flags & kFlagRaw)),
              "preprocess_fwd: bad arguments (K=%d flags=%d)", K, flags);
  if (n == 0) return GSR_OK;
  const int gx = (W + 15) / 16, gy = (H + 15) / 16;
  const int KE = effective_k(K, degree), C = feat_channels(flags);
  const unsigned nb = gsr_blocks(n, kThreads);
  hipStream_t st = (hipStream_t)stream;
  int rc;
#define GSR_ARGS                                                                              \
  KE, nb, st, xyz, scale, rot, shs, shs_rest, n, intr, extr, cam_center, W, H, gx, gy, nearest, \
      extent, flags, C, uv, depth, cov3d, conic, radius, tiles, feat, opacity_raw, opacity_act
  switch (K) {
    case 1: rc = launch_fwd_ks<1>(GSR_ARGS); break;
    case 4: rc = launch_fwd_ks<4>(GSR_ARGS); break;
    case 9: rc = launch_fwd_ks<9>(GSR_ARGS); break;
    default: rc = launch_fwd_ks<16>(GSR_ARGS); break;
  }
#undef GSR_ARGS
  GSR_REQUIRE(rc == GSR_OK, "preprocess_fwd: unsupported SH size");
  return gsr_launch_status("preprocess_fwd");
}

int gsr_preprocess_bwd(const float *xyz, const float *scale, const float *rot, const float *shs,
                       const float *shs_rest, int64_t n, int K, int degree, const float *intr, const float *extr,
                       const float *cam_center, int W, int H, float nearest, float extent,
                       int flags, const int32_t *radius, const float *d_uv, int ld_uv,
                       const float *d_depth, const float *d_conic, int ld_conic,
                       const float *d_feat, int ld_feat, float *d_xyz, float *d_scale,
                       float *d_rot, float *d_shs, float *d_shs_rest, float *d_intr,
                       float *d_extr, float *d_center, const float *opacity_raw,
                       const float *d_opacity_act, int ld_opacity_act, float *d_opacity_raw,
                       void *stream) {
  GSR_REQUIRE(n >= 0 && W > 0 && H > 0 && (K == 1 || K == 4 || K == 9 || K == 16) &&
                  (flags & ~15) == 0 &&
                  ((opacity_raw == nullptr) == (d_opacity_raw == nullptr)) &&
                  ((d_opacity_raw == nullptr) || (d_opacity_act != nullptr && ld_opacity_act >= 1)) &&
                  (opacity_raw == nullptr || (flags & kFlagRaw)),
              "preprocess_bwd: bad arguments (K=%d flags=%d)", K, flags);
  if (n == 0) return GSR_OK;
  const int KE = effective_k(K, degree), C = feat_channels(flags);
  const unsigned nb = gsr_blocks(n, kThreads);
  hipStream_t st = (hipStream_t)stream;
  int rc;
#define GSR_ARGS                                                                               \
  KE, nb, st, xyz, scale, rot, shs, shs_rest, n, intr, extr, cam_center, W, H, nearest, extent, \
      flags, C, radius, d_uv, ld_uv, d_depth, d_conic, ld_conic, d_feat, ld_feat, d_xyz, d_scale, \
      d_rot, d_shs, d_shs_rest, d_intr, d_extr, d_center, opacity_raw, d_opacity_act,             \
      ld_opacity_act, d_opacity_raw
  switch (K) {
    case 1: rc = launch_bwd_ks<1>(GSR_ARGS); break;
    case 4: rc = launch_bwd_ks<4>(GSR_ARGS); break;
    case 9: rc = launch_bwd_ks<9>(GSR_ARGS); break;
    default: rc = launch_bwd_ks<16>(GSR_ARGS); break;
  }
#undef GSR_ARGS
  GSR_REQUIRE(rc == GSR_OK, "preprocess_bwd: unsupported SH size");
  return gsr_launch_status("preprocess_bwd");
}

}  // extern "C"
