// Shared device helpers for the gfx950 Gaussian rasteriser.
//
// Every forward kernel is compiled with -ffp-contract=off: a float expression
// is evaluated exactly as written (left to right, one rounding per operator)
// and fused multiply-adds appear only where fmaf() is written.  Together with
// correctly rounded '/' and sqrtf (HIP's default) and the exp below, built only
// from IEEE-exact operations, the forward path is bit-reproducible by a plain
// C program on the host that performs the same operations in the same order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gsr {

constexpr float kAlphaMax = 0.99f;
constexpr float kAlphaMin = 1.0f / 255.0f;
constexpr float kTMin = 1e-4f;
constexpr float kLowPass = 0.3f;
constexpr float kJClamp = 1.3f;
// exp(-5.6) < 1/255 with a 6% margin: for opacity <= 1 a splat with
// power < -5.6 cannot reach the alpha threshold, so exp can be skipped without
// changing any decision.
constexpr float kPowerCut = -5.6f;

// Per-splat rows ([N,2], [N,3], [N,4], [N,7] float arrays) read and written
// as one vector access per row instead of one per float: each access costs the
// address unit one instruction for the wave, and the preprocess kernels are
// bound by that unit when every lane issues a dword access per value.  The
// types carry 4-byte alignment (a row of 3 floats is 12-byte strided); gfx950
// performs unaligned global vector accesses.
typedef float gsr_f2u __attribute__((ext_vector_type(2), aligned(4)));
typedef float gsr_f3u __attribute__((ext_vector_type(3), aligned(4)));
typedef float gsr_f4u __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ float3 ld_f3(const float *p) {
  const gsr_f3u v = *reinterpret_cast<const gsr_f3u *>(p);
  return make_float3(v.x, v.y, v.z);
}
__device__ __forceinline__ float2 ld_f2(const float *p) {
  const gsr_f2u v = *reinterpret_cast<const gsr_f2u *>(p);
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ float4 ld_f4u(const float *p) {
  const gsr_f4u v = *reinterpret_cast<const gsr_f4u *>(p);
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_f2(float *p, float a, float b) {
  gsr_f2u v;
  v.x = a;
  v.y = b;
  *reinterpret_cast<gsr_f2u *>(p) = v;
}
__device__ __forceinline__ void st_f3(float *p, float a, float b, float c) {
  gsr_f3u v;
  v.x = a;
  v.y = b;
  v.z = c;
  *reinterpret_cast<gsr_f3u *>(p) = v;
}
__device__ __forceinline__ void st_f4u(float *p, float a, float b, float c, float d) {
  gsr_f4u v;
  v.x = a;
  v.y = b;
  v.z = c;
  v.w = d;
  *reinterpret_cast<gsr_f4u *>(p) = v;
}

// exp(x) for x <= 0 (inputs below -80 are clamped; exp(-80) ~ 1.8e-35 is far
// below any alpha threshold).  Cody-Waite reduction x = k*ln2 + r with ln2
// split into a 15-bit head (k*head is exact for |k| < 2^8) and a tail, then
// the degree-7 Taylor polynomial of e^r on |r| <= ln2/2 (truncation error
// < 1e-8 relative), then an exact scaling by 2^k.  Uses only mul, fma, rint,
// ldexp: every step is correctly rounded on the GPU and on the host alike.
__device__ __forceinline__ float exp_neg(float x) {
  x = fmaxf(x, -80.0f);
  const float kf = rintf(x * 1.44269504088896341f);
  float r = fmaf(kf, -0.693145751953125f, x);
  r = fmaf(kf, -1.42860682030941723212e-6f, r);
  float p = (float)(1.0 / 5040.0);
  p = fmaf(p, r, (float)(1.0 / 720.0));
  p = fmaf(p, r, (float)(1.0 / 120.0));
  p = fmaf(p, r, (float)(1.0 / 24.0));
  p = fmaf(p, r, (float)(1.0 / 6.0));
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)kf);
}

// Tile rectangle [x0,x1) x [y0,y1) covered by a splat of integer radius r at
// (u, v): x0 = trunc((u - r)/tw), x1 = trunc((u + r + tw - 1)/tw), clamped to
// the tile grid (clamping to >= 0 first makes the truncation a floor).
struct Rect {
  int x0, x1, y0, y1;
};

__device__ __forceinline__ Rect tile_rect(float u, float v, int r, int gx, int gy, int tw,
                                          int th) {
  const float rf = (float)r;
  const float ftw = (float)tw, fth = (float)th;
  Rect q;
  q.x0 = (int)fminf(fmaxf((u - rf) / ftw, 0.0f), (float)gx);
  q.x1 = (int)fminf(fmaxf(((u + rf) + (ftw - 1.0f)) / ftw, 0.0f), (float)gx);
  q.y0 = (int)fminf(fmaxf((v - rf) / fth, 0.0f), (float)gy);
  q.y1 = (int)fminf(fmaxf(((v + rf) + (fth - 1.0f)) / fth, 0.0f), (float)gy);
  return q;
}

// ln(x) for x > 0 in normal range: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(s), s = (m-1)/(m+1), |s| < 0.172, by its odd series to s^9
// (truncation < 4e-10).  frexp, one correctly rounded division, mul and fma:
// bit-identical on the GPU and in a host C build.
__device__ __forceinline__ float log_pos(float x) {
  int e;
  float m = frexpf(x, &e);
  if (m < 0.70710678f) {
    m = m * 2.0f;
    e = e - 1;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float s2 = s * s;
  float p = 1.0f / 9.0f;
  p = fmaf(p, s2, 1.0f / 7.0f);
  p = fmaf(p, s2, 1.0f / 5.0f);
  p = fmaf(p, s2, 1.0f / 3.0f);
  p = fmaf(p, s2, 1.0f);
  return fmaf(2.0f * s, p, (float)e * 0.69314718f);
}

// Opacity-aware tile test.  A splat centred at (u, v) with conic (a, b, c)
// and opacity o reaches alpha >= kAlphaMin at pixel p only if
// q = a dx^2 + 2 b dx dy + c dy^2 <= tau = 2 ln(o / kAlphaMin), d = (u,v) - p
// (alpha = min(0.99, o exp(-q/2)); o < kAlphaMin reaches it nowhere).  The
// test takes the minimum of the convex q over the tile's continuous
// pixel-centre box -- zero with the centre inside, else the least of the four
// edge minima (each a 1-D quadratic, minimiser -(b/c) X or -(b/a) Y clamped
// to the edge) -- and keeps the tile when that minimum, lowered by the
// margin 1e-3 + 1e-4 * S (S bounds |a| dx^2 + 2|b||dx dy| + |c| dy^2 on the
// box: far above the rounding of the blend's FP32 q, of exp and of log_pos),
// is <= tau.  Never false where a pixel of the box reaches the threshold, so
// dropping the tiles it rejects leaves every blended pixel unchanged.
struct TightSplat {
  float u, v, a, b, c, rbc, rba, tau;
};

// reach == false: the splat reaches the threshold nowhere
__device__ __forceinline__ TightSplat tight_setup(float u, float v, float a, float b, float c,
                                                  float o, bool &reach) {
  TightSplat t;
  t.u = u;
  t.v = v;
  t.a = a;
  t.b = b;
  t.c = c;
  reach = o >= kAlphaMin;
  t.tau = reach ? 2.0f * log_pos(o / kAlphaMin) : 0.0f;
  const bool pd = a > 0.0f && c > 0.0f;
  t.rbc = pd ? b / c : 0.0f;
  t.rba = pd ? b / a : 0.0f;
  return t;
}

__device__ __forceinline__ bool tile_reachable(const TightSplat &t, float x0, float x1, float y0,
                                               float y1) {
  if (!(t.a > 0.0f) || !(t.c > 0.0f)) return true;
  const float dx0 = t.u - x1, dx1 = t.u - x0, dy0 = t.v - y1, dy1 = t.v - y0;
  if (dx0 <= 0.0f && dx1 >= 0.0f && dy0 <= 0.0f && dy1 >= 0.0f) return true;
  float m;
  {
    const float ya = fminf(fmaxf(-(t.rbc * dx0), dy0), dy1);
    const float yb = fminf(fmaxf(-(t.rbc * dx1), dy0), dy1);
    const float xa = fminf(fmaxf(-(t.rba * dy0), dx0), dx1);
    const float xb = fminf(fmaxf(-(t.rba * dy1), dx0), dx1);
    m = t.a * dx0 * dx0 + 2.0f * t.b * dx0 * ya + t.c * ya * ya;
    m = fminf(m, t.a * dx1 * dx1 + 2.0f * t.b * dx1 * yb + t.c * yb * yb);
    m = fminf(m, t.c * dy0 * dy0 + 2.0f * t.b * dy0 * xa + t.a * xa * xa);
    m = fminf(m, t.c * dy1 * dy1 + 2.0f * t.b * dy1 * xb + t.a * xb * xb);
  }
  const float DX = fmaxf(fabsf(dx0), fabsf(dx1)), DY = fmaxf(fabsf(dy0), fabsf(dy1));
  const float S = fabsf(t.a) * DX * DX + 2.0f * fabsf(t.b) * DX * DY + fabsf(t.c) * DY * DY;
  return m - (1e-3f + 1e-4f * S) <= t.tau;
}

// ---------------------------------------------------------------------------
// wave64 helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// Cross-lane exchange for the butterfly below, without the LDS crossbar
// where the ISA allows: lanes L and L^M pair up.  M = 32 and 16 use gfx950's
// v_permlane32_swap / v_permlane16_swap (swap the upper half-wave / odd rows
// of the first operand with the lower half-wave / even rows of the second),
// M = 8, 2, 1 use DPP (row rotate by 8, quad permutes), M = 4 a ds_swizzle
// xor pattern.
//
// xpair<M>(lo, hi): a lane with bit M clear returns lo + (partner's lo); a
// lane with bit M set returns (partner's hi) + hi.
template <int M>
__device__ __forceinline__ float xpair(float lo, float hi, int lane);

template <int M>
__device__ __forceinline__ float xor_get(float v) {
  const int x = __float_as_int(v);
  if constexpr (M == 1) return __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false));
  else if constexpr (M == 2) return __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false));
  else if constexpr (M == 8) return __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, false));
  else if constexpr (M == 4) return __int_as_float(__builtin_amdgcn_ds_swizzle(x, 0x101F));
  else return __shfl_xor(v, M, 64);
}

template <int M>
__device__ __forceinline__ float xpair(float lo, float hi, int lane) {
  if constexpr (M == 32 || M == 16) {
    const unsigned a = __float_as_uint(lo), b = __float_as_uint(hi);
    const auto r = M == 32 ? __builtin_amdgcn_permlane32_swap(a, b, false, false)
                           : __builtin_amdgcn_permlane16_swap(a, b, false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
  } else {
    const bool up = (lane & M) != 0;
    const float keep = up ? hi : lo, send = up ? lo : hi;
    const float got = xor_get<M>(send);
    return up ? got + keep : keep + got;
  }
}

// Butterfly transpose reduction of N values per lane across the wave64.
// Each xor stage M exchanges half of the live values with the partner lane
// and adds, halving the live count; after log2(N) stages every lane holds one
// fully or partially reduced value, and the remaining stages are plain
// xor-adds.  Costs ~2N exchanges instead of 6N for N independent reductions.
// On return v[0] of lane L is the wave sum of value bfly_index<N>(L).
template <int N, int M>
struct Bfly {
  __device__ static __forceinline__ void run(float *v, int lane) {
    if constexpr (N > 1) {
      constexpr int H = N / 2;
#pragma unroll
      for (int i = 0; i < H; ++i) v[i] = xpair<M>(v[i], v[i + H], lane);
      Bfly<H, M / 2>::run(v, lane);
    } else if constexpr (M >= 1) {
      v[0] = xpair<M>(v[0], v[0], lane);
      Bfly<1, M / 2>::run(v, lane);
    }
  }
};

template <int N>
__device__ __forceinline__ int bfly_index(int lane) {
  int idx = 0, h = N / 2;
#pragma unroll
  for (int m = 32; m >= 1 && h >= 1; m >>= 1, h >>= 1) idx += (lane & m) ? h : 0;
  return idx;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

}  // namespace gsr

// Status codes returned through the C ABI.
#define GSR_OK 0
#define GSR_ERR_ARG -1000
#define GSR_ERR_CAPACITY 1
