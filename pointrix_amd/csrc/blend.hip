// Per-tile front-to-back alpha blending (msplat.alpha_blending,
// pointrix/model/renderer/msplat.py:148-151) and its reverse-order backward.
//
// One workgroup per tile, one thread per pixel.  Splats of the tile's
// depth-sorted range are staged through LDS in batches of one splat per
// thread (position, conic pre-scaled to the exponent's coefficients, opacity,
// C feature channels); every pixel of the tile then walks the batch from LDS.
// A block-wide count of finished pixels ends the tile early.
//
// Backward: each pixel starts from its saved T_final and contributor count and
// walks its splats back to front, recovering T_i = T_{i+1} / (1 - alpha_i)
// instead of storing it.  The 6 + C per-splat gradients of the wave's 64
// pixels are summed with a butterfly transpose reduction, after which one
// atomic instruction (one lane per value) adds them into the splat's
// contiguous gradient record.
#include "gsr_internal.h"

using namespace gsr;

namespace {

struct SplatLds {
  float2 xy;
  float4 co;  // (-a/2, -b, -c/2, opacity)
};

// exponent of the splat at pixel offset (dx, dy) = (u - px, v - py):
// -(a dx^2 + 2 b dx dy + c dy^2) / 2 with the conic (a, b, c).
__device__ __forceinline__ float splat_power(float4 co, float dx, float dy) {
  return fmaf(co.x * dx, dx, fmaf(co.y * dx, dy, (co.z * dy) * dy));
}

// alpha of a splat, or a negative value when the splat is skipped (exponent
// positive, or alpha below 1/255).  Shared by forward and backward so the
// backward replays exactly the forward's decisions.
__device__ __forceinline__ float splat_alpha(float4 co, float power, float &G) {
  G = 0.0f;
  if (power > 0.0f) return -1.0f;
  if (power < kPowerCut && co.w <= 1.0f) return -1.0f;
  G = exp_neg(power);
  const float alpha = fminf(kAlphaMax, co.w * G);
  if (alpha < kAlphaMin) return -1.0f;
  return alpha;
}

template <int C, int TW, int TH>
__global__ __launch_bounds__(TW *TH) void blend_fwd_kernel(
    const float *__restrict__ uv, const float *__restrict__ conic,
    const float *__restrict__ opacity, const float *__restrict__ feat,
    const int32_t *__restrict__ ids, const int32_t *__restrict__ ranges, float bg, int W, int H,
    int gx, float *__restrict__ out, float *__restrict__ Tfin, int32_t *__restrict__ ncon) {
  constexpr int B = TW * TH;
  __shared__ SplatLds s_sp[B];
  __shared__ float s_f[C * B];
  const int tile = blockIdx.x;
  const int tx = tile % gx, ty = tile / gx;
  const int px = tx * TW + (int)threadIdx.x % TW;
  const int py = ty * TH + (int)threadIdx.x / TW;
  const bool inside = px < W && py < H;
  const float fpx = (float)px, fpy = (float)py;
  const int start = ranges[2 * tile], end = ranges[2 * tile + 1];
  float T = 1.0f;
  float acc[C];
#pragma unroll
  for (int c = 0; c < C; ++c) acc[c] = 0.0f;
  int last = 0;
  bool done = !inside;
  for (int base = start; base < end; base += B) {
    if (__syncthreads_count(done) == B) break;
    const int j = base + (int)threadIdx.x;
    if (j < end) {
      const int g = ids[j];
      SplatLds sp;
      sp.xy = make_float2(uv[2 * g], uv[2 * g + 1]);
      sp.co = make_float4(-0.5f * conic[3 * g], -conic[3 * g + 1], -0.5f * conic[3 * g + 2],
                          opacity[g]);
      s_sp[threadIdx.x] = sp;
#pragma unroll
      for (int c = 0; c < C; ++c) s_f[c * B + threadIdx.x] = feat[(int64_t)g * C + c];
    }
    __syncthreads();
    if (!done) {
      const int cnt = min(B, end - base);
      for (int k = 0; k < cnt; ++k) {
        const SplatLds sp = s_sp[k];
        const float dx = sp.xy.x - fpx, dy = sp.xy.y - fpy;
        float G;
        const float alpha = splat_alpha(sp.co, splat_power(sp.co, dx, dy), G);
        if (alpha < 0.0f) continue;
        const float test_T = T * (1.0f - alpha);
        if (test_T < kTMin) {
          done = true;
          break;
        }
        const float w = alpha * T;
#pragma unroll
        for (int c = 0; c < C; ++c) acc[c] = fmaf(s_f[c * B + k], w, acc[c]);
        T = test_T;
        last = base - start + k + 1;
      }
    }
  }
  if (inside) {
    const int64_t pix = (int64_t)py * W + px;
    const int64_t HW = (int64_t)H * W;
#pragma unroll
    for (int c = 0; c < C; ++c) out[c * HW + pix] = acc[c] + bg * T;
    Tfin[pix] = T;
    ncon[pix] = last;
  }
}

// Per-splat gradient record written by the backward: [du, dv, da, db, dc,
// dopacity, dfeature_0 .. dfeature_{C-1}] at d_splat[g * ld + ...].
template <int C>
struct GradLayout {
  static constexpr int Q = 6 + C;
  static constexpr int QP = Q <= 8 ? 8 : (Q <= 16 ? 16 : 32);  // butterfly width
};

template <int C, int TW, int TH>
__global__ __launch_bounds__(TW *TH) void blend_bwd_kernel(
    const float *__restrict__ uv, const float *__restrict__ conic,
    const float *__restrict__ opacity, const float *__restrict__ feat,
    const int32_t *__restrict__ ids, const int32_t *__restrict__ ranges, float bg, int W, int H,
    int gx, const float *__restrict__ Tfin, const int32_t *__restrict__ ncon,
    const float *__restrict__ dout, const float *__restrict__ dTfin, float *__restrict__ d_splat,
    int ld) {
  constexpr int B = TW * TH;
  constexpr int QP = GradLayout<C>::QP;
  constexpr int Q = GradLayout<C>::Q;
  __shared__ SplatLds s_sp[B];
  __shared__ float s_f[C * B];
  __shared__ int s_id[B];
  __shared__ int s_max;
  const int tile = blockIdx.x;
  const int tx = tile % gx, ty = tile / gx;
  const int px = tx * TW + (int)threadIdx.x % TW;
  const int py = ty * TH + (int)threadIdx.x / TW;
  const bool inside = px < W && py < H;
  const float fpx = (float)px, fpy = (float)py;
  const int start = ranges[2 * tile], end = ranges[2 * tile + 1];
  const int64_t pix = (int64_t)py * W + px;
  const int64_t HW = (int64_t)H * W;
  const int lane = lane_id();
  const int my_q = bfly_index<QP>(lane);
  const bool writer = (lane & (64 / QP - 1)) == 0 && my_q < Q;

  float T = 1.0f, gpix[C];
  int last = 0;
  float gT = 0.0f;
  if (inside) {
    T = Tfin[pix];
    last = ncon[pix];
#pragma unroll
    for (int c = 0; c < C; ++c) gpix[c] = dout[c * HW + pix];
    if (dTfin != nullptr) gT = dTfin[pix];
  } else {
#pragma unroll
    for (int c = 0; c < C; ++c) gpix[c] = 0.0f;
  }
  if (threadIdx.x == 0) s_max = 0;
  __syncthreads();
  if (last > 0) atomicMax(&s_max, last);
  __syncthreads();
  const int stop = start + s_max;
  if (stop <= start || end <= start) return;  // uniform
  const float T_final = T;
  float gsum = 0.0f;
#pragma unroll
  for (int c = 0; c < C; ++c) gsum += gpix[c];
  const float bg_term = bg * gsum + gT;  // d out / d T_final summed with d T_final
  float accum[C], last_f[C];
#pragma unroll
  for (int c = 0; c < C; ++c) accum[c] = last_f[c] = 0.0f;
  float last_alpha = 0.0f;

  for (int bend = stop; bend > start; bend -= B) {
    const int bstart = max(start, bend - B);
    __syncthreads();
    const int j = bstart + (int)threadIdx.x;
    if (j < bend) {
      const int g = ids[j];
      SplatLds sp;
      sp.xy = make_float2(uv[2 * g], uv[2 * g + 1]);
      sp.co = make_float4(-0.5f * conic[3 * g], -conic[3 * g + 1], -0.5f * conic[3 * g + 2],
                          opacity[g]);
      s_sp[threadIdx.x] = sp;
      s_id[threadIdx.x] = g;
#pragma unroll
      for (int c = 0; c < C; ++c) s_f[c * B + threadIdx.x] = feat[(int64_t)g * C + c];
    }
    __syncthreads();
    for (int k = bend - bstart - 1; k >= 0; --k) {
      const int local = bstart + k - start;
      float v[QP];
#pragma unroll
      for (int q = 0; q < QP; ++q) v[q] = 0.0f;
      bool contrib = local < last;
      if (contrib) {
        const SplatLds sp = s_sp[k];
        const float dx = sp.xy.x - fpx, dy = sp.xy.y - fpy;
        float G;
        const float alpha = splat_alpha(sp.co, splat_power(sp.co, dx, dy), G);
        contrib = alpha >= 0.0f;
        if (contrib) {
          const float inv_om = __builtin_amdgcn_rcpf(1.0f - alpha);
          T = T * inv_om;
          const float w = alpha * T;
          float dL_dalpha = 0.0f;
#pragma unroll
          for (int c = 0; c < C; ++c) {
            const float f = s_f[c * B + k];
            v[6 + c] = w * gpix[c];
            accum[c] = last_alpha * last_f[c] + (1.0f - last_alpha) * accum[c];
            last_f[c] = f;
            dL_dalpha += (f - accum[c]) * gpix[c];
          }
          dL_dalpha *= T;
          dL_dalpha -= T_final * inv_om * bg_term;
          last_alpha = alpha;
          if (sp.co.w * G < kAlphaMax) {  // alpha not clamped
            v[5] = G * dL_dalpha;
            const float dpow = alpha * dL_dalpha;
            v[0] = dpow * (2.0f * sp.co.x * dx + sp.co.y * dy);
            v[1] = dpow * (sp.co.y * dx + 2.0f * sp.co.z * dy);
            v[2] = dpow * (-0.5f * dx * dx);
            v[3] = dpow * (-dx * dy);
            v[4] = dpow * (-0.5f * dy * dy);
          }
        }
      }
      if (__ballot(contrib) == 0ull) continue;  // wave-uniform
      Bfly<QP, 32>::run(v, lane);
      if (writer) atomicAdd(d_splat + (int64_t)s_id[k] * ld + my_q, v[0]);
    }
  }
}

template <int C>
int launch_fwd(const float *uv, const float *conic, const float *opacity, const float *feature,
               const int32_t *ids, const int32_t *tile_range, float bg, int W, int H, int tw,
               int th, float *out, float *T_final, int32_t *n_contrib, hipStream_t s) {
  const int gx = (W + tw - 1) / tw, gy = (H + th - 1) / th;
  const unsigned nt = (unsigned)(gx * gy);
  if (tw == 16 && th == 16)
    blend_fwd_kernel<C, 16, 16><<<nt, 256, 0, s>>>(uv, conic, opacity, feature, ids, tile_range,
                                                   bg, W, H, gx, out, T_final, n_contrib);
  else if (tw == 8 && th == 32)
    blend_fwd_kernel<C, 8, 32><<<nt, 256, 0, s>>>(uv, conic, opacity, feature, ids, tile_range,
                                                  bg, W, H, gx, out, T_final, n_contrib);
  else if (tw == 32 && th == 8)
    blend_fwd_kernel<C, 32, 8><<<nt, 256, 0, s>>>(uv, conic, opacity, feature, ids, tile_range,
                                                  bg, W, H, gx, out, T_final, n_contrib);
  else
    return GSR_ERR_ARG;
  return GSR_OK;
}

template <int C>
int launch_bwd(const float *uv, const float *conic, const float *opacity, const float *feature,
               const int32_t *ids, const int32_t *tile_range, float bg, int W, int H, int tw,
               int th, const float *T_final, const int32_t *n_contrib, const float *d_out,
               const float *d_T_final, float *d_splat, int ld, hipStream_t s) {
  const int gx = (W + tw - 1) / tw, gy = (H + th - 1) / th;
  const unsigned nt = (unsigned)(gx * gy);
#define GSR_BWD_ARGS                                                                          \
  uv, conic, opacity, feature, ids, tile_range, bg, W, H, gx, T_final, n_contrib, d_out, \
      d_T_final, d_splat, ld
  if (tw == 16 && th == 16)
    blend_bwd_kernel<C, 16, 16><<<nt, 256, 0, s>>>(GSR_BWD_ARGS);
  else if (tw == 8 && th == 32)
    blend_bwd_kernel<C, 8, 32><<<nt, 256, 0, s>>>(GSR_BWD_ARGS);
  else if (tw == 32 && th == 8)
    blend_bwd_kernel<C, 32, 8><<<nt, 256, 0, s>>>(GSR_BWD_ARGS);
  else
    return GSR_ERR_ARG;
#undef GSR_BWD_ARGS
  return GSR_OK;
}

#define GSR_DISPATCH_C(C, EXPR)        \
  switch (C) {                         \
    case 1: { constexpr int CC = 1; EXPR; } break;   \
    case 2: { constexpr int CC = 2; EXPR; } break;   \
    case 3: { constexpr int CC = 3; EXPR; } break;   \
    case 4: { constexpr int CC = 4; EXPR; } break;   \
    case 5: { constexpr int CC = 5; EXPR; } break;   \
    case 6: { constexpr int CC = 6; EXPR; } break;   \
    case 7: { constexpr int CC = 7; EXPR; } break;   \
    case 8: { constexpr int CC = 8; EXPR; } break;   \
    case 9: { constexpr int CC = 9; EXPR; } break;   \
    case 10: { constexpr int CC = 10; EXPR; } break; \
    case 11: { constexpr int CC = 11; EXPR; } break; \
    case 12: { constexpr int CC = 12; EXPR; } break; \
    case 13: { constexpr int CC = 13; EXPR; } break; \
    case 14: { constexpr int CC = 14; EXPR; } break; \
    case 15: { constexpr int CC = 15; EXPR; } break; \
    case 16: { constexpr int CC = 16; EXPR; } break; \
    default: rc = GSR_ERR_ARG;         \
  }

}  // namespace

extern "C" {

int gsr_alpha_blending_fwd(const float *uv, const float *conic, const float *opacity,
                           const float *feature, int C, const int32_t *ids,
                           const int32_t *tile_range, float bg, int W, int H, int tile_w,
                           int tile_h, float *out, float *T_final, int32_t *n_contrib,
                           void *stream) {
  GSR_REQUIRE(W > 0 && H > 0 && C >= 1 && C <= 16, "alpha_blending: bad sizes (C=%d)", C);
  int rc = GSR_OK;
  GSR_DISPATCH_C(C, rc = launch_fwd<CC>(uv, conic, opacity, feature, ids, tile_range, bg, W, H,
                                        tile_w, tile_h, out, T_final, n_contrib,
                                        (hipStream_t)stream));
  GSR_REQUIRE(rc == GSR_OK, "alpha_blending: unsupported tile shape %dx%d", tile_w, tile_h);
  return gsr_launch_status("alpha_blending_fwd");
}

int gsr_alpha_blending_bwd(const float *uv, const float *conic, const float *opacity,
                           const float *feature, int C, const int32_t *ids,
                           const int32_t *tile_range, float bg, int W, int H, int tile_w,
                           int tile_h, const float *T_final, const int32_t *n_contrib,
                           const float *d_out, const float *d_T_final, float *d_splat, int ld,
                           void *stream) {
  GSR_REQUIRE(W > 0 && H > 0 && C >= 1 && C <= 16 && ld >= 6 + C,
              "alpha_blending_bwd: bad sizes (C=%d, ld=%d)", C, ld);
  int rc = GSR_OK;
  GSR_DISPATCH_C(C, rc = launch_bwd<CC>(uv, conic, opacity, feature, ids, tile_range, bg, W, H,
                                        tile_w, tile_h, T_final, n_contrib, d_out, d_T_final,
                                        d_splat, ld, (hipStream_t)stream));
  GSR_REQUIRE(rc == GSR_OK, "alpha_blending_bwd: unsupported tile shape %dx%d", tile_w, tile_h);
  return gsr_launch_status("alpha_blending_bwd");
}

}  // extern "C"
