// Fused L1 + SSIM loss (pointrix/model/loss.py:27 l1_loss and :73-116 ssim,
// combined as in BaseModel.get_loss_dict, pointrix/model/base_model.py:
// 112-124), forward and backward, over P planes of H x W (a [B, C, H, W]
// batch is P = B*C planes).
//
// SSIM uses the reference's 11-tap Gaussian window (sigma 1.5) with zero
// padding, per plane; the map is averaged over every plane and pixel.  The
// five window statistics (mu_x, mu_y, E[x^2], E[y^2], E[xy]) are separable
// blurs: each workgroup stages its output tile plus halo in LDS, blurs rows,
// then columns.  The backward recomputes the statistics (no [5, P, H, W]
// intermediate in HBM) and applies the adjoint blur:
//   dx(q) = (G*A)(q) + 2 x(q) (G*B)(q) + y(q) (G*C)(q)
// with A, B, C the derivatives of the map w.r.t. mu_x, E[x^2], E[xy].
// Means are two-stage fixed-order reductions (graph-safe, no atomics).
#include "gsr_internal.h"

using namespace gsr;

namespace {

constexpr int kR = 5;            // window radius (11 taps)
constexpr int kTX = 32, kTY = 16;  // output tile
constexpr int kThreads = 256;
constexpr float kC1 = 0.01f * 0.01f, kC2 = 0.03f * 0.03f;

struct Window {
  float g[2 * kR + 1];
};

__device__ __forceinline__ float at(const float *__restrict__ p, int W, int H, int x, int y) {
  return (x >= 0 && x < W && y >= 0 && y < H) ? p[(int64_t)y * W + x] : 0.0f;
}

// ssim map value and its partials from the window statistics
struct SsimTerms {
  float s, dmu1, ds11, ds12;
};

__device__ __forceinline__ SsimTerms ssim_terms(float mu1, float mu2, float s11, float s22,
                                                float s12) {
  const float mu1_sq = mu1 * mu1, mu2_sq = mu2 * mu2, mu12 = mu1 * mu2;
  const float sig1 = s11 - mu1_sq, sig2 = s22 - mu2_sq, sig12 = s12 - mu12;
  const float n1 = 2.0f * mu12 + kC1, n2 = 2.0f * sig12 + kC2;
  const float d1 = mu1_sq + mu2_sq + kC1, d2 = sig1 + sig2 + kC2;
  const float inv = 1.0f / (d1 * d2);
  SsimTerms t;
  t.s = (n1 * n2) * inv;
  t.dmu1 = 2.0f * mu2 * (n2 - n1) * inv - 2.0f * mu1 * t.s * (1.0f / d1 - 1.0f / d2);
  t.ds11 = -t.s / d2;
  t.ds12 = 2.0f * n1 * inv;
  return t;
}

// Forward: per workgroup one kTX x kTY tile of one plane; partial sums of
// the ssim map and of |x - y| into part[2 * block].
__global__ __launch_bounds__(kThreads) void ssim_fwd_kernel(const float *__restrict__ x,
                                                            const float *__restrict__ y, int W,
                                                            int H, int tiles_x, int tiles_y,
                                                            Window win,
                                                            float *__restrict__ part) {
  constexpr int RX = kTX + 2 * kR, RY = kTY + 2 * kR;
  __shared__ float sx[RY][RX], sy[RY][RX];
  __shared__ float h[5][RY][kTX];
  __shared__ float s_w[2][kThreads / 64];
  const int tile = blockIdx.x % (tiles_x * tiles_y);
  const int plane = blockIdx.x / (tiles_x * tiles_y);
  const int ox = (tile % tiles_x) * kTX, oy = (tile / tiles_x) * kTY;
  const float *px = x + (int64_t)plane * W * H, *py = y + (int64_t)plane * W * H;
  for (int k = threadIdx.x; k < RX * RY; k += kThreads) {
    const int r = k / RX, c = k % RX;
    sx[r][c] = at(px, W, H, ox + c - kR, oy + r - kR);
    sy[r][c] = at(py, W, H, ox + c - kR, oy + r - kR);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < RY * kTX; k += kThreads) {
    const int r = k / kTX, c = k % kTX;
    float a = 0.f, b = 0.f, p = 0.f, q = 0.f, s = 0.f;
#pragma unroll
    for (int t = 0; t <= 2 * kR; ++t) {
      const float g = win.g[t], xv = sx[r][c + t], yv = sy[r][c + t];
      a = fmaf(g, xv, a);
      b = fmaf(g, yv, b);
      p = fmaf(g, xv * xv, p);
      q = fmaf(g, yv * yv, q);
      s = fmaf(g, xv * yv, s);
    }
    h[0][r][c] = a;
    h[1][r][c] = b;
    h[2][r][c] = p;
    h[3][r][c] = q;
    h[4][r][c] = s;
  }
  __syncthreads();
  float acc_s = 0.f, acc_l = 0.f;
  for (int k = threadIdx.x; k < kTY * kTX; k += kThreads) {
    const int r = k / kTX, c = k % kTX;
    if (ox + c >= W || oy + r >= H) continue;
    float m[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t <= 2 * kR; ++t) {
      const float g = win.g[t];
#pragma unroll
      for (int j = 0; j < 5; ++j) m[j] = fmaf(g, h[j][r + t][c], m[j]);
    }
    acc_s += ssim_terms(m[0], m[1], m[2], m[3], m[4]).s;
    acc_l += fabsf(sx[r + kR][c + kR] - sy[r + kR][c + kR]);
  }
  acc_s = wave_sum(acc_s);
  acc_l = wave_sum(acc_l);
  if (lane_id() == 0) {
    s_w[0][threadIdx.x >> 6] = acc_s;
    s_w[1][threadIdx.x >> 6] = acc_l;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float a = 0.f, b = 0.f;
    for (int w = 0; w < kThreads / 64; ++w) {
      a += s_w[0][w];
      b += s_w[1][w];
    }
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// out[0] = mean ssim, out[1] = mean |x - y| (fixed-order sums of part)
__global__ __launch_bounds__(kThreads) void ssim_final_kernel(const float *__restrict__ part,
                                                              int nb, float inv_n,
                                                              float *__restrict__ out) {
  __shared__ float s_w[2][kThreads / 64];
  float a = 0.f, b = 0.f;
  for (int i = threadIdx.x; i < nb; i += kThreads) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  a = wave_sum(a);
  b = wave_sum(b);
  if (lane_id() == 0) {
    s_w[0][threadIdx.x >> 6] = a;
    s_w[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float sa = 0.f, sb = 0.f;
    for (int w = 0; w < kThreads / 64; ++w) {
      sa += s_w[0][w];
      sb += s_w[1][w];
    }
    out[0] = sa * inv_n;
    out[1] = sb * inv_n;
  }
}

// Backward: dx for one kTX x kTY tile.  Needs the map's partials on the tile
// plus a kR halo, hence the statistics on that ring, hence inputs with a 2kR
// halo.  g[0] = dL/d(mean ssim), g[1] = dL/d(mean |x - y|).
__global__ __launch_bounds__(kThreads) void ssim_bwd_kernel(const float *__restrict__ x,
                                                            const float *__restrict__ y, int W,
                                                            int H, int tiles_x, int tiles_y,
                                                            Window win,
                                                            const float *__restrict__ g,
                                                            float inv_n,
                                                            float *__restrict__ dx) {
  constexpr int IX = kTX + 4 * kR, IY = kTY + 4 * kR;  // inputs
  constexpr int MX = kTX + 2 * kR, MY = kTY + 2 * kR;  // map partials
  __shared__ float sx[IY][IX], sy[IY][IX];
  __shared__ float h[5][IY][MX];  // row blurs of the statistics (reused for A, B, C)
  __shared__ float abc[3][MY][MX];
  const int tile = blockIdx.x % (tiles_x * tiles_y);
  const int plane = blockIdx.x / (tiles_x * tiles_y);
  const int ox = (tile % tiles_x) * kTX, oy = (tile / tiles_x) * kTY;
  const float *px = x + (int64_t)plane * W * H, *py = y + (int64_t)plane * W * H;
  const float gs = g[0] * inv_n, gl = g[1] * inv_n;
  for (int k = threadIdx.x; k < IX * IY; k += kThreads) {
    const int r = k / IX, c = k % IX;
    sx[r][c] = at(px, W, H, ox + c - 2 * kR, oy + r - 2 * kR);
    sy[r][c] = at(py, W, H, ox + c - 2 * kR, oy + r - 2 * kR);
  }
  __syncthreads();
  // row blurs: rows 0..IY-1, map columns 0..MX-1 (input columns c..c+2kR)
  for (int k = threadIdx.x; k < IY * MX; k += kThreads) {
    const int r = k / MX, c = k % MX;
    float a = 0.f, b = 0.f, p = 0.f, q = 0.f, s = 0.f;
#pragma unroll
    for (int t = 0; t <= 2 * kR; ++t) {
      const float w = win.g[t], xv = sx[r][c + t], yv = sy[r][c + t];
      a = fmaf(w, xv, a);
      b = fmaf(w, yv, b);
      p = fmaf(w, xv * xv, p);
      q = fmaf(w, yv * yv, q);
      s = fmaf(w, xv * yv, s);
    }
    h[0][r][c] = a;
    h[1][r][c] = b;
    h[2][r][c] = p;
    h[3][r][c] = q;
    h[4][r][c] = s;
  }
  __syncthreads();
  // column blurs -> map partials at positions (oy - kR + r, ox - kR + c);
  // zero outside the image (the map exists only there)
  for (int k = threadIdx.x; k < MY * MX; k += kThreads) {
    const int r = k / MX, c = k % MX;
    const int X = ox - kR + c, Y = oy - kR + r;
    float A = 0.f, B = 0.f, Cc = 0.f;
    if (X >= 0 && X < W && Y >= 0 && Y < H) {
      float m[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int t = 0; t <= 2 * kR; ++t) {
        const float w = win.g[t];
#pragma unroll
        for (int j = 0; j < 5; ++j) m[j] = fmaf(w, h[j][r + t][c], m[j]);
      }
      const SsimTerms st = ssim_terms(m[0], m[1], m[2], m[3], m[4]);
      A = gs * st.dmu1;
      B = gs * st.ds11;
      Cc = gs * st.ds12;
    }
    abc[0][r][c] = A;
    abc[1][r][c] = B;
    abc[2][r][c] = Cc;
  }
  __syncthreads();
  // adjoint blur of A, B, C: rows first (reuse h[0..2] as [MY][kTX])
  for (int k = threadIdx.x; k < MY * kTX; k += kThreads) {
    const int r = k / kTX, c = k % kTX;
    float a = 0.f, b = 0.f, s = 0.f;
#pragma unroll
    for (int t = 0; t <= 2 * kR; ++t) {
      const float w = win.g[t];
      a = fmaf(w, abc[0][r][c + t], a);
      b = fmaf(w, abc[1][r][c + t], b);
      s = fmaf(w, abc[2][r][c + t], s);
    }
    h[0][r][c] = a;
    h[1][r][c] = b;
    h[2][r][c] = s;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kTY * kTX; k += kThreads) {
    const int r = k / kTX, c = k % kTX;
    const int X = ox + c, Y = oy + r;
    if (X >= W || Y >= H) continue;
    float a = 0.f, b = 0.f, s = 0.f;
#pragma unroll
    for (int t = 0; t <= 2 * kR; ++t) {
      const float w = win.g[t];
      a = fmaf(w, h[0][r + t][c], a);
      b = fmaf(w, h[1][r + t][c], b);
      s = fmaf(w, h[2][r + t][c], s);
    }
    const float xv = sx[r + 2 * kR][c + 2 * kR], yv = sy[r + 2 * kR][c + 2 * kR];
    const float d = xv - yv;
    const float sg = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
    dx[(int64_t)plane * W * H + (int64_t)Y * W + X] = a + 2.0f * xv * b + yv * s + gl * sg;
  }
}

Window make_window() {
  Window w;
  double g[2 * kR + 1], sum = 0.0;
  for (int i = 0; i <= 2 * kR; ++i) {
    const double d = i - kR;
    g[i] = exp(-d * d / (2.0 * 1.5 * 1.5));
    sum += g[i];
  }
  for (int i = 0; i <= 2 * kR; ++i) w.g[i] = (float)(g[i] / sum);
  return w;
}

}  // namespace

extern "C" {

int gsr_ssim_partials(int planes, int W, int H) {
  const int tx = (W + kTX - 1) / kTX, ty = (H + kTY - 1) / kTY;
  return 2 * planes * tx * ty;
}

int gsr_l1_ssim_fwd(const float *x, const float *y, int planes, int W, int H, float *partial,
                    float *out, void *stream) {
  GSR_REQUIRE(planes > 0 && W > 0 && H > 0, "l1_ssim: bad sizes");
  const int tx = (W + kTX - 1) / kTX, ty = (H + kTY - 1) / kTY;
  const int nb = planes * tx * ty;
  hipStream_t st = (hipStream_t)stream;
  ssim_fwd_kernel<<<nb, kThreads, 0, st>>>(x, y, W, H, tx, ty, make_window(), partial);
  ssim_final_kernel<<<1, kThreads, 0, st>>>(partial, nb,
                                           (float)(1.0 / ((double)planes * W * H)), out);
  return gsr_launch_status("l1_ssim_fwd");
}

int gsr_l1_ssim_bwd(const float *x, const float *y, int planes, int W, int H, const float *g,
                    float *dx, void *stream) {
  GSR_REQUIRE(planes > 0 && W > 0 && H > 0, "l1_ssim_bwd: bad sizes");
  const int tx = (W + kTX - 1) / kTX, ty = (H + kTY - 1) / kTY;
  ssim_bwd_kernel<<<planes * tx * ty, kThreads, 0, (hipStream_t)stream>>>(
      x, y, W, H, tx, ty, make_window(), g, (float)(1.0 / ((double)planes * W * H)), dx);
  return gsr_launch_status("l1_ssim_bwd");
}

}  // extern "C"
