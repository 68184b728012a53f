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

// ---- row gather for densify / prune with optimizer-state remap --------------
// dst row r of every table entry = src row index[r] (index < 0: zeros), for
// all entries (parameters and their Adam moments) in one launch: the
// compaction / extension pointrix does per tensor with boolean-mask indexing
// and torch.cat (model/point_cloud/utils/point_utils.py:51-108,
// controller/gs.py:182-247).
namespace {

__global__ __launch_bounds__(256) void gather_rows_kernel(const GsrRowCopy *__restrict__ tab,
                                                          const int32_t *__restrict__ index,
                                                          int64_t rows) {
  const GsrRowCopy t = tab[blockIdx.y];
  const int64_t total = rows * t.width;
  const uint32_t w = (uint32_t)t.width;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    // rows * width < 2^32 (required of the caller): 32-bit division
    const uint32_t e32 = (uint32_t)e;
    const uint32_t r = e32 / w, c = e32 - r * w;
    const int32_t s = index[r];
    t.dst[e] = s >= 0 ? t.src[(int64_t)s * t.width + c] : 0.0f;
  }
}

}  // namespace

extern "C" {

int gsr_gather_rows(const GsrRowCopy *table, int ntab, const int32_t *index, int64_t rows,
                    void *stream) {
  GSR_REQUIRE(ntab >= 0 && ntab <= 65535 && rows >= 0 && rows < (int64_t)INT32_MAX,
              "gather_rows: bad sizes");
  if (ntab == 0 || rows == 0) return GSR_OK;
  const int64_t want = (rows * 16 + 255) / 256;
  const unsigned gx = (unsigned)(want < 65536 ? want : 65536);
  gather_rows_kernel<<<dim3(gx, (unsigned)ntab), 256, 0, (hipStream_t)stream>>>(table, index,
                                                                              rows);
  return gsr_launch_status("gather_rows");
}

int gsr_row_copy_bytes(void) { return (int)sizeof(GsrRowCopy); }

}  // extern "C"
