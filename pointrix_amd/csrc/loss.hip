// Fused L1 loss (pointrix/model/loss.py:27 l1_loss = mean |x - y|, used by
// BaseModel.get_loss_dict, pointrix/model/base_model.py:112-124), forward
// and backward, as kernels that are safe inside a captured HIP graph: the
// forward is a two-stage deterministic reduction (per-block partial sums,
// then one block sums the partials in a fixed order), with no atomics and no
// memset; the backward is elementwise and reads the upstream gradient from
// device memory.
#include "gsr_internal.h"

using namespace gsr;

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;

__global__ __launch_bounds__(kThreads) void l1_partial_kernel(const float *__restrict__ x,
                                                              const float *__restrict__ y,
                                                              int64_t n,
                                                              float *__restrict__ partial) {
  __shared__ float s_w[kThreads / 64];
  float acc = 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    acc += fabsf(x[i] - y[i]);
  acc = wave_sum(acc);
  if (lane_id() == 0) s_w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.0f;
    for (int w = 0; w < kThreads / 64; ++w) s += s_w[w];
    partial[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(kThreads) void l1_final_kernel(const float *__restrict__ partial,
                                                            int nb, float scale,
                                                            float *__restrict__ out) {
  __shared__ float s_w[kThreads / 64];
  float acc = 0.0f;
  for (int i = threadIdx.x; i < nb; i += kThreads) acc += partial[i];
  acc = wave_sum(acc);
  if (lane_id() == 0) s_w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.0f;
    for (int w = 0; w < kThreads / 64; ++w) s += s_w[w];
    out[0] = s * scale;
  }
}

__global__ __launch_bounds__(kThreads) void l1_bwd_kernel(const float *__restrict__ x,
                                                          const float *__restrict__ y, int64_t n,
                                                          const float *__restrict__ g,
                                                          float scale, float *__restrict__ dx) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float d = x[i] - y[i];
  const float s = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
  dx[i] = s * g[0] * scale;
}

}  // namespace

extern "C" {

int gsr_l1_loss_partials(int64_t n) {
  const int64_t nb = (n + kThreads * 8 - 1) / (kThreads * 8);
  return (int)(nb < 1 ? 1 : (nb > kMaxBlocks ? kMaxBlocks : nb));
}

int gsr_l1_loss_fwd(const float *x, const float *y, int64_t n, float *partial, float *out,
                    void *stream) {
  GSR_REQUIRE(n > 0, "l1_loss: empty input");
  hipStream_t st = (hipStream_t)stream;
  const int nb = gsr_l1_loss_partials(n);
  l1_partial_kernel<<<nb, kThreads, 0, st>>>(x, y, n, partial);
  l1_final_kernel<<<1, kThreads, 0, st>>>(partial, nb, 1.0f / (float)n, out);
  return gsr_launch_status("l1_loss_fwd");
}

int gsr_l1_loss_bwd(const float *x, const float *y, int64_t n, const float *g_out, float *dx,
                    void *stream) {
  GSR_REQUIRE(n > 0, "l1_loss_bwd: empty input");
  l1_bwd_kernel<<<gsr_blocks(n, kThreads), kThreads, 0, (hipStream_t)stream>>>(
      x, y, n, g_out, 1.0f / (float)n, dx);
  return gsr_launch_status("l1_loss_bwd");
}

}  // extern "C"
