// Densification statistics (DensificationController.preprocess,
// pointrix/controller/gs.py:318-333, with accumulate_viewspace_grad,
// gs.py:259-284) in one pass over the splats, and the viewspace-gradient
// scaling of the blend backward's records.  Elementwise HBM streams: one
// thread per splat, no atomics, no host synchronisation.
#include "gsr_internal.h"

using namespace gsr;

namespace {

constexpr int kThreads = 256;

// g = (sum over views of grad[v][i]) * (sx, sy), the views summed in order
// and then scaled, as the reference does; for visible splats
// grad_accum += |g|, acc_steps += 1, max_radii = max(max_radii, radii).
__global__ __launch_bounds__(kThreads) void densify_stats_kernel(
    const float *__restrict__ grad, int64_t vstride, int64_t rstride, int V, int64_t n,
    const uint8_t *__restrict__ vis, const int32_t *__restrict__ radii, float sx, float sy,
    float *__restrict__ grad_accum, float *__restrict__ acc_steps,
    float *__restrict__ max_radii) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n || vis[i] == 0) return;
  float gx = 0.0f, gy = 0.0f;
  for (int v = 0; v < V; ++v) {
    const float *p = grad + v * vstride + i * rstride;
    gx = gx + p[0];
    gy = gy + p[1];
  }
  gx = gx * sx;
  gy = gy * sy;
  grad_accum[i] = grad_accum[i] + sqrtf(gx * gx + gy * gy);
  acc_steps[i] = acc_steps[i] + 1.0f;
  max_radii[i] = fmaxf(max_radii[i], (float)radii[i]);
}

// out[i] = (rec[i*ld] * sx, rec[i*ld + 1] * sy): uv_points.grad from the
// per-splat gradient records of the blend backward.
__global__ __launch_bounds__(kThreads) void uv_grad_scale_kernel(const float *__restrict__ rec,
                                                                 int ld, int64_t n, float sx,
                                                                 float sy,
                                                                 float2 *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float *r = rec + i * ld;
  out[i] = make_float2(r[0] * sx, r[1] * sy);
}

}  // namespace

extern "C" {

int gsr_densify_stats(const float *grad, int64_t view_stride, int64_t row_stride, int V,
                      int64_t n, const uint8_t *visibility, const int32_t *radii, float sx,
                      float sy, float *grad_accum, float *acc_steps, float *max_radii,
                      void *stream) {
  GSR_REQUIRE(n >= 0 && V >= 1 && row_stride >= 2, "densify_stats: bad sizes");
  if (n == 0) return GSR_OK;
  densify_stats_kernel<<<gsr_blocks(n, kThreads), kThreads, 0, (hipStream_t)stream>>>(
      grad, view_stride, row_stride, V, n, visibility, radii, sx, sy, grad_accum, acc_steps,
      max_radii);
  return gsr_launch_status("densify_stats");
}

int gsr_uv_grad_scale(const float *rec, int ld, int64_t n, float sx, float sy, float *out,
                      void *stream) {
  GSR_REQUIRE(n >= 0 && ld >= 2, "uv_grad_scale: bad sizes");
  if (n == 0) return GSR_OK;
  uv_grad_scale_kernel<<<gsr_blocks(n, kThreads), kThreads, 0, (hipStream_t)stream>>>(
      rec, ld, n, sx, sy, (float2 *)out);
  return gsr_launch_status("uv_grad_scale");
}

}  // extern "C"
