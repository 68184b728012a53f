// Host-side helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>

#include "gsr.h"
#include "gsr_common.h"

void gsr_set_error(const char *fmt, ...);

// Map the launch status of the last kernel to the C-ABI return code.
static inline int gsr_launch_status(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    gsr_set_error("%s: %s", what, hipGetErrorString(e));
    return -(int)e;
  }
  return GSR_OK;
}

#define GSR_REQUIRE(cond, ...)        \
  do {                                \
    if (!(cond)) {                    \
      gsr_set_error(__VA_ARGS__);     \
      return GSR_ERR_ARG;             \
    }                                 \
  } while (0)

static inline unsigned gsr_blocks(int64_t n, int threads) {
  return (unsigned)((n + threads - 1) / threads);
}

namespace gsr {

// Block-wide sum of K per-thread values, then one atomicAdd per value per
// block into dst[0..K).  Used for the camera gradients, which reduce over all
// splats.  blockDim.x must be a multiple of 64 and at most 1024.
template <int K>
__device__ __forceinline__ void block_atomic_sum(float (&v)[K], float *dst) {
  __shared__ float s_red[16][K];
  const int wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    float s = wave_sum(v[k]);
    if (lane_id() == 0) s_red[wid][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    float s = 0.0f;
    for (int w = 0; w < nw; ++w) s += s_red[w][threadIdx.x];
    if (s != 0.0f) atomicAdd(dst + threadIdx.x, s);
  }
}

}  // namespace gsr
