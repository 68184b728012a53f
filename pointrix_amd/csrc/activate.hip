// Activations of the Gaussian point cloud (GaussianPointCloud's get_scaling /
// get_rotation / get_opacity, pointrix/model/point_cloud/gaussian_points.py:
// 30-35,70-81): scale = exp(log_scale), rotation = q / max(|q|, 1e-12)
// (F.normalize over the last dim), opacity = sigmoid(logit).  One elementwise
// pass per splat forward and one backward, instead of torch's separate
// exp / norm / divide / sigmoid kernels and their backward nodes.
#include "gsr_internal.h"

using namespace gsr;

namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void activate_fwd_kernel(
    const float *__restrict__ ls, const float *__restrict__ q, const float *__restrict__ lo,
    int64_t n, float *__restrict__ s, float *__restrict__ r, float *__restrict__ o) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float3 l = ld_f3(ls + 3 * i);
  st_f3(s + 3 * i, expf(l.x), expf(l.y), expf(l.z));
  const float4 v = reinterpret_cast<const float4 *>(q)[i];
  const float d = fmaxf(sqrtf(((v.x * v.x + v.y * v.y) + v.z * v.z) + v.w * v.w), 1e-12f);
  reinterpret_cast<float4 *>(r)[i] = make_float4(v.x / d, v.y / d, v.z / d, v.w / d);
  o[i] = 1.0f / (1.0f + expf(-lo[i]));
}

// d log_scale = g_s * s;  d q = (g_r - r (r . g_r)) / |q|  (g_r / 1e-12 where
// the norm is clamped);  d logit = g_o * o (1 - o).  A NULL gradient is zero.
// r is recomputed from q with the forward's operations (bit-identical), which
// saves reading it back.
__global__ __launch_bounds__(kThreads) void activate_bwd_kernel(
    const float *__restrict__ q, const float *__restrict__ s,
    const float *__restrict__ o, const float *__restrict__ gs, const float *__restrict__ gr,
    const float *__restrict__ go, int64_t n, float *__restrict__ dls, float *__restrict__ dq,
    float *__restrict__ dlo) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  if (dls != nullptr) {
    if (gs != nullptr) {
      const float3 g = ld_f3(gs + 3 * i), sv = ld_f3(s + 3 * i);
      st_f3(dls + 3 * i, g.x * sv.x, g.y * sv.y, g.z * sv.z);
    } else {
      st_f3(dls + 3 * i, 0.0f, 0.0f, 0.0f);
    }
  }
  if (dq != nullptr) {
    float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
    if (gr != nullptr) g = reinterpret_cast<const float4 *>(gr)[i];
    const float4 v = reinterpret_cast<const float4 *>(q)[i];
    const float nq = sqrtf(((v.x * v.x + v.y * v.y) + v.z * v.z) + v.w * v.w);
    const float dn = fmaxf(nq, 1e-12f);
    const float4 u = make_float4(v.x / dn, v.y / dn, v.z / dn, v.w / dn);
    float4 d;
    if (nq > 1e-12f) {
      const float dot = ((u.x * g.x + u.y * g.y) + u.z * g.z) + u.w * g.w;
      d = make_float4((g.x - u.x * dot) / nq, (g.y - u.y * dot) / nq, (g.z - u.z * dot) / nq,
                      (g.w - u.w * dot) / nq);
    } else {
      d = make_float4(g.x / 1e-12f, g.y / 1e-12f, g.z / 1e-12f, g.w / 1e-12f);
    }
    reinterpret_cast<float4 *>(dq)[i] = d;
  }
  if (dlo != nullptr) dlo[i] = go != nullptr ? go[i] * (o[i] * (1.0f - o[i])) : 0.0f;
}

}  // namespace

extern "C" {

int gsr_activate_fwd(const float *log_scale, const float *quat, const float *logit, int64_t n,
                     float *scale, float *rot, float *opacity, void *stream) {
  GSR_REQUIRE(n >= 0, "activate: bad size");
  GSR_REQUIRE(((reinterpret_cast<uintptr_t>(quat) | reinterpret_cast<uintptr_t>(rot)) & 15) == 0,
              "activate: quaternion arrays must be 16-byte aligned");
  if (n == 0) return GSR_OK;
  activate_fwd_kernel<<<gsr_blocks(n, kThreads), kThreads, 0, (hipStream_t)stream>>>(
      log_scale, quat, logit, n, scale, rot, opacity);
  return gsr_launch_status("activate_fwd");
}

int gsr_activate_bwd(const float *quat, const float *scale, const float *rot,
                     const float *opacity, const float *g_scale, const float *g_rot,
                     const float *g_opacity, int64_t n, float *d_log_scale, float *d_quat,
                     float *d_logit, void *stream) {
  GSR_REQUIRE(n >= 0, "activate_bwd: bad size");
  GSR_REQUIRE(((reinterpret_cast<uintptr_t>(quat) | reinterpret_cast<uintptr_t>(rot) |
                reinterpret_cast<uintptr_t>(g_rot) | reinterpret_cast<uintptr_t>(d_quat)) &
               15) == 0,
              "activate_bwd: quaternion arrays must be 16-byte aligned");
  if (n == 0) return GSR_OK;
  activate_bwd_kernel<<<gsr_blocks(n, kThreads), kThreads, 0, (hipStream_t)stream>>>(
      quat, scale, opacity, g_scale, g_rot, g_opacity, n, d_log_scale, d_quat, d_logit);
  return gsr_launch_status("activate_bwd");
}

}  // extern "C"
