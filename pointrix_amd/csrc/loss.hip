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
constexpr int kBwdItems = 4 * kThreads;  // elements per block of the multi-term backward

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

// A weighted sum of L1 terms, sum_k w_k mean|x_k - y_k| (a loss such as
// l1(rgb) + 0.1 l1(depth) + 0.1 l1(normal)), in the same two-stage order:
// each term owns a contiguous range of partial blocks, each partial already
// scaled by w_k / n_k, and one block sums them all.  The backward is one
// elementwise launch over every term.
struct L1Table {
  const float *x[GSR_MAX_L1_TERMS];
  const float *y[GSR_MAX_L1_TERMS];
  float *dx[GSR_MAX_L1_TERMS];
  int64_t n[GSR_MAX_L1_TERMS];
  float scale[GSR_MAX_L1_TERMS];
  int first_block[GSR_MAX_L1_TERMS + 1];
  int nterms;
};

__device__ __forceinline__ int term_of_block(const L1Table &t, int b) {
  int k = 0;
  while (k + 1 < t.nterms && b >= t.first_block[k + 1]) ++k;
  return k;
}

__global__ __launch_bounds__(kThreads) void l1_sum_partial_kernel(L1Table t,
                                                                  float *__restrict__ partial) {
  __shared__ float s_w[kThreads / 64];
  const int k = term_of_block(t, blockIdx.x);
  const int nbk = t.first_block[k + 1] - t.first_block[k];
  const float *__restrict__ x = t.x[k];
  const float *__restrict__ y = t.y[k];
  const int64_t n = t.n[k];
  const int64_t stride = (int64_t)nbk * kThreads;
  int64_t i = (int64_t)(blockIdx.x - t.first_block[k]) * kThreads + threadIdx.x;
  float acc = 0.0f;
  for (; i + 3 * stride < n; i += 4 * stride) {  // four independent loads in flight
    const float a0 = x[i], a1 = x[i + stride], a2 = x[i + 2 * stride], a3 = x[i + 3 * stride];
    const float b0 = y[i], b1 = y[i + stride], b2 = y[i + 2 * stride], b3 = y[i + 3 * stride];
    acc += (fabsf(a0 - b0) + fabsf(a1 - b1)) + (fabsf(a2 - b2) + fabsf(a3 - b3));
  }
  for (; i < n; i += stride) acc += fabsf(x[i] - y[i]);
  acc = wave_sum(acc);
  if (lane_id() == 0) s_w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.0f;
    for (int w = 0; w < kThreads / 64; ++w) s += s_w[w];
    partial[blockIdx.x] = s * t.scale[k];
  }
}

__global__ __launch_bounds__(kThreads) void l1_sum_bwd_kernel(L1Table t,
                                                              const float *__restrict__ g) {
  const int k = term_of_block(t, blockIdx.x);
  float *__restrict__ dx = t.dx[k];
  if (dx == nullptr) return;
  const float gs = g[0] * t.scale[k];
  const int64_t base = (int64_t)(blockIdx.x - t.first_block[k]) * kBwdItems + threadIdx.x;
#pragma unroll
  for (int r = 0; r < kBwdItems / kThreads; ++r) {
    const int64_t i = base + r * kThreads;
    if (i < t.n[k]) {
      const float d = t.x[k][i] - t.y[k][i];
      dx[i] = (d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f)) * gs;
    }
  }
}

// Builds the launch table; blocks_of(n) = the blocks a term of n elements
// gets.  Returns the total block count, or -1 on a bad table.
template <typename F>
int l1_table(const GsrL1Terms &in, F blocks_of, L1Table &t) {
  if (in.nterms < 1 || in.nterms > GSR_MAX_L1_TERMS) return -1;
  t.nterms = in.nterms;
  int64_t total = 0;
  for (int k = 0; k < in.nterms; ++k) {
    if (in.n[k] <= 0 || in.x[k] == nullptr || in.y[k] == nullptr) return -1;
    t.x[k] = in.x[k];
    t.y[k] = in.y[k];
    t.dx[k] = in.dx[k];
    t.n[k] = in.n[k];
    t.scale[k] = in.weight[k] / (float)in.n[k];
    t.first_block[k] = (int)total;
    total += blocks_of(in.n[k]);
    if (total > INT32_MAX) return -1;
  }
  t.first_block[in.nterms] = (int)total;
  return (int)total;
}

}  // namespace

extern "C" {

int gsr_l1_terms_bytes(void) { return (int)sizeof(GsrL1Terms); }

int gsr_l1_sum_partials(const GsrL1Terms *terms) {
  if (terms == nullptr) return -1;
  L1Table t;
  return l1_table(*terms, gsr_l1_loss_partials, t);
}

int gsr_l1_sum_fwd(const GsrL1Terms *terms, float *partial, float *out, void *stream) {
  GSR_REQUIRE(terms != nullptr, "l1_sum: no term table");
  L1Table t;
  const int nb = l1_table(*terms, gsr_l1_loss_partials, t);
  GSR_REQUIRE(nb > 0, "l1_sum: bad term table (1..%d non-empty terms)", GSR_MAX_L1_TERMS);
  hipStream_t st = (hipStream_t)stream;
  l1_sum_partial_kernel<<<nb, kThreads, 0, st>>>(t, partial);
  l1_final_kernel<<<1, kThreads, 0, st>>>(partial, nb, 1.0f, out);
  return gsr_launch_status("l1_sum_fwd");
}

int gsr_l1_sum_bwd(const GsrL1Terms *terms, const float *g_out, void *stream) {
  GSR_REQUIRE(terms != nullptr && g_out != nullptr, "l1_sum_bwd: no term table / gradient");
  L1Table t;
  const int nb = l1_table(
      *terms, [](int64_t n) -> int64_t { return (n + kBwdItems - 1) / kBwdItems; }, t);
  GSR_REQUIRE(nb > 0, "l1_sum_bwd: bad term table (1..%d non-empty terms)", GSR_MAX_L1_TERMS);
  l1_sum_bwd_kernel<<<nb, kThreads, 0, (hipStream_t)stream>>>(t, g_out);
  return gsr_launch_status("l1_sum_bwd");
}

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
