// Output planes of the renderer.  MsplatRender returns the blended [C,H,W]
// image split along dim 0 per feature in RenderFeatures' order
// (renderer_utils.py:57-71, msplat.py:152), and maps a normal group n of
// three planes to (n / |n| + 1) / 2 with |n| over the channel dimension
// (examples/supervise/renderer.py:89-96).  Forward: one pass writes the split
// planes, the normal group normalised.  Backward: one pass assembles the
// gradient of the blended image from the per-group gradients, instead of a
// zero-filled full-size tensor per slice plus the separate norm / divide
// kernels of the unfused formulation.
#include "gsr_internal.h"

using namespace gsr;

namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void planes_fwd_kernel(const float *__restrict__ in,
                                                              int C, int64_t P, int noff,
                                                              float *__restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p >= P) return;
  for (int c = 0; c < C; ++c) {
    if (c == noff) {
      const float a = in[c * P + p], b = in[(c + 1) * P + p], d = in[(c + 2) * P + p];
      const float r = sqrtf((a * a + b * b) + d * d);
      out[c * P + p] = (a / r + 1.0f) / 2.0f;
      out[(c + 1) * P + p] = (b / r + 1.0f) / 2.0f;
      out[(c + 2) * P + p] = (d / r + 1.0f) / 2.0f;
      c += 2;
    } else {
      out[c * P + p] = in[c * P + p];
    }
  }
}

// d in[c] from the gradient of the output plane c (NULL: zero); for the
// normal group o = (n / r + 1) / 2:  d n = (g - n (n.g) / r^2) / (2 r).
__global__ __launch_bounds__(kThreads) void planes_bwd_kernel(const float *__restrict__ in,
                                                              int C, int64_t P, int noff,
                                                              GsrPlaneGrads g,
                                                              float *__restrict__ d_in) {
  const int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p >= P) return;
  for (int k = 0; k < g.ngroups; ++k) {
    const int c0 = g.first[k], cn = g.count[k];
    const float *gp = g.grad[k];
    if (c0 == noff) {
      const float n[3] = {in[c0 * P + p], in[(c0 + 1) * P + p], in[(c0 + 2) * P + p]};
      float gv[3] = {0.0f, 0.0f, 0.0f};
      if (gp != nullptr) {
        gv[0] = gp[p];
        gv[1] = gp[P + p];
        gv[2] = gp[2 * P + p];
      }
      const float r2 = (n[0] * n[0] + n[1] * n[1]) + n[2] * n[2];
      const float r = sqrtf(r2);
      const float ng = (n[0] * gv[0] + n[1] * gv[1]) + n[2] * gv[2];
#pragma unroll
      for (int j = 0; j < 3; ++j) d_in[(c0 + j) * P + p] = (gv[j] - n[j] * (ng / r2)) / (2.0f * r);
    } else {
      for (int j = 0; j < cn; ++j) d_in[(c0 + j) * P + p] = gp != nullptr ? gp[j * P + p] : 0.0f;
    }
  }
}

}  // namespace

extern "C" {

int gsr_planes_fwd(const float *in, int C, int64_t P, int normal_first, float *out,
                   void *stream) {
  GSR_REQUIRE(C >= 1 && P >= 0 && (normal_first < 0 || normal_first + 3 <= C),
              "planes_fwd: bad sizes (C=%d, normal at %d)", C, normal_first);
  if (P == 0) return GSR_OK;
  planes_fwd_kernel<<<gsr_blocks(P, kThreads), kThreads, 0, (hipStream_t)stream>>>(
      in, C, P, normal_first, out);
  return gsr_launch_status("planes_fwd");
}

int gsr_planes_bwd(const float *in, int C, int64_t P, int normal_first,
                   const GsrPlaneGrads *grads, float *d_in, void *stream) {
  GSR_REQUIRE(grads != nullptr && grads->ngroups >= 1 && grads->ngroups <= GSR_MAX_PLANE_GROUPS,
              "planes_bwd: bad group table");
  int covered = 0;
  for (int k = 0; k < grads->ngroups; ++k) {
    GSR_REQUIRE(grads->first[k] == covered && grads->count[k] >= 1,
                "planes_bwd: groups must tile the channels in order");
    GSR_REQUIRE(grads->first[k] != normal_first || grads->count[k] == 3,
                "planes_bwd: the normal group has three planes");
    covered += grads->count[k];
  }
  GSR_REQUIRE(covered == C && P >= 0 && (normal_first < 0 || normal_first + 3 <= C),
              "planes_bwd: bad sizes (C=%d, normal at %d)", C, normal_first);
  if (P == 0) return GSR_OK;
  planes_bwd_kernel<<<gsr_blocks(P, kThreads), kThreads, 0, (hipStream_t)stream>>>(
      in, C, P, normal_first, *grads, d_in);
  return gsr_launch_status("planes_bwd");
}

int gsr_plane_grads_bytes(void) { return (int)sizeof(GsrPlaneGrads); }

}  // extern "C"
