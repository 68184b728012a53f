// Fused Adam over all parameter tensors of a step in one launch (the update
// torch.optim.Adam performs for pointrix's optimizer, pointrix/optimizer/
// optimizer.py:128-140 with eps from the config, weight decay 0, no amsgrad):
//   m = b1 m + (1 - b1) g,  v = b2 v + (1 - b2) g^2,
//   p -= (lr / (1 - b1^t)) * m / (sqrt(v) / sqrt(1 - b2^t) + eps)
// A table of tensors (pointers, length, step size, bias correction) lives in
// device memory; each workgroup takes kChunk consecutive elements of one
// tensor (a prefix of per-tensor chunk counts maps workgroups to tensors),
// streams p, g, m, v once (coalesced), writes p, m, v once, and optionally
// zeroes g (zero_grad without a separate pass): 28-32 bytes per element,
// against several passes for an unfused update.
#include "gsr_internal.h"

using namespace gsr;

namespace {

constexpr int kThreads = 256;
constexpr int kChunk = kThreads * 16;  // elements per workgroup

__global__ __launch_bounds__(kThreads) void adam_kernel(const GsrAdamTensor *__restrict__ tab,
                                                        int ntab, float b1, float b2, float a1,
                                                        float a2, float eps, int zero_grad) {
  // tensor of this workgroup: the last entry whose first chunk <= blockIdx.x
  __shared__ int s_t;
  if (threadIdx.x == 0) {
    int lo = 0, hi = ntab - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tab[mid].first_chunk <= (int64_t)blockIdx.x) lo = mid; else hi = mid - 1;
    }
    s_t = lo;
  }
  __syncthreads();
  const GsrAdamTensor t = tab[s_t];
  const int64_t begin = ((int64_t)blockIdx.x - t.first_chunk) * kChunk;
  const int64_t end = min(begin + (int64_t)kChunk, t.n);
  for (int64_t i = begin + threadIdx.x; i < end; i += kThreads) {
    const float g = t.g[i];
    const float m = fmaf(b1, t.m[i], a1 * g);
    const float v = fmaf(b2, t.v[i], a2 * g * g);
    t.m[i] = m;
    t.v[i] = v;
    t.p[i] = t.p[i] - t.step_size * (m / (sqrtf(v) / t.bc2_sqrt + eps));
    if (zero_grad) t.g[i] = 0.0f;
  }
}

}  // namespace

extern "C" {

int gsr_adam_step(const GsrAdamTensor *table, int ntab, int64_t nchunks, double beta1,
                  double beta2, double eps, int zero_grad, void *stream) {
  GSR_REQUIRE(ntab >= 0 && nchunks >= 0 && nchunks < (int64_t)INT32_MAX,
              "adam_step: bad sizes");
  if (ntab == 0 || nchunks == 0) return GSR_OK;
  // 1 - beta in double: from a float beta2 = 0.999 it would carry a 1e-5
  // relative error
  adam_kernel<<<(unsigned)nchunks, kThreads, 0, (hipStream_t)stream>>>(
      table, ntab, (float)beta1, (float)beta2, (float)(1.0 - beta1), (float)(1.0 - beta2),
      (float)eps, zero_grad);
  return gsr_launch_status("adam_step");
}

int gsr_adam_chunk(void) { return kChunk; }

int gsr_adam_entry_bytes(void) { return (int)sizeof(GsrAdamTensor); }

}  // extern "C"
