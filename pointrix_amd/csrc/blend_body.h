// Per-tile bodies of the alpha-blending forward and backward, shared by the
// one-view kernels (blend.hip) and the multi-view persistent kernels
// (multiview.hip).  A tile is processed by one workgroup of TW*TH threads,
// one thread per pixel; the tile's depth-sorted splat range is staged through
// LDS in batches of TW*TH splats.
#pragma once
#include "gsr_common.h"

namespace gsr {

struct SplatLds {
  float2 xy;
  float4 co;  // (-a/2, -b, -c/2, opacity)
};

// exponent of the splat at pixel offset (dx, dy) = (u - px, v - py):
// -(a dx^2 + 2 b dx dy + c dy^2) / 2 with the conic (a, b, c).
__device__ __forceinline__ float splat_power(float4 co, float dx, float dy) {
  return fmaf(co.x * dx, dx, fmaf(co.y * dx, dy, (co.z * dy) * dy));
}

// A splat is skipped at a pixel when its exponent is positive, when the
// exponent is below kPowerCut with opacity <= 1 (alpha is then certainly below
// 1/255), or when alpha = min(0.99, opacity * exp(power)) is below 1/255.  The
// forward and the backward evaluate exactly these tests, so the backward
// replays the forward's decisions.

// Where a tile lives: its view's image size and tile grid, the tile's index
// within that grid, the offset of the view's pixels in the (packed) outputs,
// the channel stride of the outputs, and the offset subtracted from a splat
// id to index per-splat opacity (multi-view ids are view * N + splat).
struct TileView {
  int W, H, gx, tile;
  int64_t pix_base, chan_stride;
  int64_t op_offset;
};

constexpr int kMaxTileWaves = 4;  // 256-pixel tiles, one wave per 64 pixels

// A splat's feature row [C] into batch column dst[c * BD]; rows of 7 (rgb |
// depth | normal) as one 16-byte and one 12-byte access.
template <int C, int BD>
__device__ __forceinline__ void load_feat_row(const float *__restrict__ row, float *dst) {
  if constexpr (C == 7) {
    const float4 a = ld_f4u(row);
    const float3 b = ld_f3(row + 4);
    dst[0] = a.x;
    dst[BD] = a.y;
    dst[2 * BD] = a.z;
    dst[3 * BD] = a.w;
    dst[4 * BD] = b.x;
    dst[5 * BD] = b.y;
    dst[6 * BD] = b.z;
  } else {
#pragma unroll
    for (int c = 0; c < C; ++c) dst[c * BD] = row[c];
  }
}

// LDS of one tile: a batch of BD splats (BD <= TW*TH; the "batch depth"), and
// for the forward the per-wave lists of batch slots that can reach the wave's
// pixels (slot indices in depth order) with their per-loading-wave counts.
template <int C, int BD>
struct BlendLds {
  SplatLds sp[BD];
  float f[C * BD];
  int id[BD];
  int wcnt[kMaxTileWaves][kMaxTileWaves];  // [strip][loading wave]
  uint8_t list[kMaxTileWaves][BD];
  int misc;
};

// Forward of one tile.  Each wave owns a strip of 64 / TW pixel rows across
// the tile.  When a batch is staged, the thread loading a splat also tests it
// against every strip with the opacity-aware reachability test of the
// binning (tile_reachable: never false where a pixel of the box reaches
// alpha >= kAlphaMin), and the batch slots are compacted into one depth-ordered
// list per strip.  A wave then walks only its list: a splat left out could not
// have changed any of its pixels (it is neither a contributor nor the splat at
// which a pixel stops, both of which need alpha >= kAlphaMin), so every
// pixel's colour, T and contributor count are exactly those of the walk over
// the whole range.
template <int C, int TW, int TH, int BD>
__device__ __forceinline__ void blend_fwd_tile(
    const TileView &tv, int start, int end, const float *__restrict__ uv,
    const float *__restrict__ conic, const float *__restrict__ opacity,
    const float *__restrict__ feat, const int32_t *__restrict__ ids, float bg,
    float *__restrict__ out, float *__restrict__ Tfin, int32_t *__restrict__ ncon,
    BlendLds<C, BD> &sm) {
  constexpr int B = TW * TH;
  constexpr int NW = B / 64;      // waves (strips) per tile
  constexpr int RW = 64 / TW;     // pixel rows per strip
  static_assert(BD <= B, "batch depth exceeds the workgroup");
  static_assert(BD <= 256, "batch slots are stored as bytes");
  static_assert(64 % TW == 0 && B % 64 == 0 && NW <= kMaxTileWaves, "tile shape");
  const int tx = tv.tile % tv.gx, ty = tv.tile / tv.gx;
  const int px = tx * TW + (int)threadIdx.x % TW;
  const int py = ty * TH + (int)threadIdx.x / TW;
  const bool inside = px < tv.W && py < tv.H;
  float fpx = (float)px, fpy = (float)py;
  // kept in registers through the walk (the compiler would otherwise convert
  // them again on every splat)
  asm volatile("" : "+v"(fpx), "+v"(fpy));
  const int wave = (int)threadIdx.x >> 6;
  const int lane = lane_id();
  const uint64_t lt = lanemask_lt();
  // pixel-centre box of the tile (clipped to the image) and of each strip
  const float bx0 = (float)(tx * TW), bx1 = (float)(min(tx * TW + TW, tv.W) - 1);
  const int by0 = ty * TH, byl = min(ty * TH + TH, tv.H) - 1;
  float T = 1.0f;
  float acc[C];
#pragma unroll
  for (int c = 0; c < C; ++c) acc[c] = 0.0f;
  int last = 0;
  // the pixel is done (outside the image, or stopped) when nonzero; an int
  // rather than a bool, so the wave-uniform tests on it compile to scalar
  // compares of a lane mask
  int done = inside ? 0 : 1;
  for (int base = start; base < end; base += BD) {
    if (__syncthreads_count(done != 0) == B) break;
    const int j = base + (int)threadIdx.x;
    unsigned mask = 0u;
    if ((int)threadIdx.x < BD && j < end) {
      const int g = ids[j];
      const float2 uv2 = ld_f2(uv + 2 * (int64_t)g);
      const float3 co3 = ld_f3(conic + 3 * (int64_t)g);
      const float u = uv2.x, v = uv2.y, ca = co3.x, cb = co3.y, cc = co3.z;
      const float o = opacity[g - tv.op_offset];
      SplatLds sp;
      sp.xy = make_float2(u, v);
      sp.co = make_float4(-0.5f * ca, -cb, -0.5f * cc, o);
      sm.sp[threadIdx.x] = sp;
      load_feat_row<C, BD>(feat + (int64_t)g * C, sm.f + threadIdx.x);
      bool reach;
      const TightSplat ts = tight_setup(u, v, ca, cb, cc, o, reach);
      if (reach) {
#pragma unroll
        for (int w = 0; w < NW; ++w) {
          const int y0 = by0 + w * RW, y1 = min(y0 + RW - 1, byl);
          if (y0 <= y1 && tile_reachable(ts, bx0, bx1, (float)y0, (float)y1)) mask |= 1u << w;
        }
      }
    }
    // compaction: slot t goes to position (slots before t in strip w) of list w
    int pos[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint64_t bal = __ballot((mask >> w) & 1u);
      pos[w] = (int)__popcll(bal & lt);
      if (lane == 0) sm.wcnt[w][wave] = (int)__popcll(bal);
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      if ((mask >> w) & 1u) {
        int off = pos[w];
        for (int q = 0; q < wave; ++q) off += sm.wcnt[w][q];
        sm.list[w][off] = (uint8_t)threadIdx.x;
      }
    }
    int n_mine = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) n_mine += sm.wcnt[wave][q];
    __syncthreads();
    for (int i = 0; i < n_mine; ++i) {
      // Branch-light walk: lanes keep their own state through selects; the
      // wave only branches (uniformly) to skip a splat no active lane needs
      // or to leave once every lane is done.  Decisions are exactly those of
      // the straightforward loop (a lane that skips or stops changes nothing).
      if (__ballot(done == 0) == 0ull) break;
      const int k = __builtin_amdgcn_readfirstlane((int)sm.list[wave][i]);
      const SplatLds sp = sm.sp[k];
      const float dx = sp.xy.x - fpx, dy = sp.xy.y - fpy;
      const float power = splat_power(sp.co, dx, dy);
      const bool near = done == 0 && !(power > 0.0f) && !(power < kPowerCut && sp.co.w <= 1.0f);
      if (__ballot(near) == 0ull) continue;
      const float alpha = fminf(kAlphaMax, sp.co.w * exp_neg(power));
      const bool valid = near && !(alpha < kAlphaMin);
      const float test_T = T * (1.0f - alpha);
      const bool stop = valid && test_T < kTMin;
      const bool contrib = valid && !stop;
      done = stop ? 1 : done;
      const float w = contrib ? alpha * T : 0.0f;
#pragma unroll
      for (int c = 0; c < C; ++c) acc[c] = fmaf(sm.f[c * BD + k], w, acc[c]);
      T = contrib ? test_T : T;
      last = contrib ? base - start + k + 1 : last;
    }
  }
  if (inside) {
    const int64_t pix = tv.pix_base + (int64_t)py * tv.W + px;
#pragma unroll
    for (int c = 0; c < C; ++c) out[c * tv.chan_stride + pix] = acc[c] + bg * T;
    Tfin[pix] = T;
    ncon[pix] = last;
  }
}

// Per-splat gradient record written by the backward: [du, dv, da, db, dc,
// dopacity, dfeature_0 .. dfeature_{C-1}] at d_splat[g * ld + ...]; (du, dv)
// and dopacity instead into dense [N,2] / [N] arrays when those are given
// (their consumers then read 12 bytes per splat, not a record row), and with
// both given the record row is compact, [da, db, dc, dfeature...] (3 + C).
template <int C>
struct GradLayout {
  static constexpr int Q = 6 + C;
  static constexpr int QP = Q <= 8 ? 8 : (Q <= 16 ? 16 : 32);  // butterfly width
};

// Backward of one tile with PP vertically adjacent pixels per lane
// (TW*TH/PP threads; a wave covers a TW x 64 PP/TW strip).  Each pixel replays
// the forward's decisions in reverse from its n_contrib, recovering T_i from
// T_final.  dL/dalpha_i = T_i sum_c (f_ic - A_ic) g_c - ..., where A_c (the
// colour blended behind splat i) obeys A <- a' f' + (1 - a') A over the
// splats behind it (a', f' the next splat back); the recurrence is linear, so
// a pixel keeps only the scalars S = sum_c A_c g_c and F = sum_c f'_c g_c.
// Each lane adds its pixels' contributions before the butterfly, so the
// cross-lane reduction, the atomic and the LDS reads of the splat are paid
// once per PP pixels.  PP = 2 is used: at full3m four pixels per lane (one
// wave per 16x16 tile) measured 358 against 333 us.
template <int C>
struct PixBwd {
  float T, S, F, last_alpha, T_final, bg_term;
  float g[C];
  int last;
  float fx, fy;
};

// The pixel's offset to the splat, the exponent and whether the splat can
// reach alpha >= 1/255 there within the pixel's contributors (the skip tests
// of the forward).
struct PixNear {
  float dx, dy, power;
  bool near;
};

template <int C>
__device__ __forceinline__ PixNear pix_near(const PixBwd<C> &p, const SplatLds &sp, int local) {
  PixNear r;
  r.dx = sp.xy.x - p.fx;
  r.dy = sp.xy.y - p.fy;
  r.power = splat_power(sp.co, r.dx, r.dy);
  r.near = local < p.last && !(r.power > 0.0f) && !(r.power < kPowerCut && sp.co.w <= 1.0f);
  return r;
}

// BG: the pixels have a background / T_final gradient term (bg != 0 or a
// T_final gradient); without one the term is exactly zero and is skipped.
template <int C, int QP, bool BG>
__device__ __forceinline__ bool pix_bwd_step(PixBwd<C> &p, const SplatLds &sp, const PixNear &pn,
                                             const float *f, int fstride, float (&v)[QP]) {
  const float dx = pn.dx, dy = pn.dy, power = pn.power;
  const bool near = pn.near;
  // pixels that do not use this splat evaluate exp at 0 (power may be large
  // and positive there, and G feeds a product that must stay finite)
  const float G = exp_neg(near ? power : 0.0f);
  const float alpha = fminf(kAlphaMax, sp.co.w * G);
  const bool contrib = near && !(alpha < kAlphaMin);
  const float inv_om = __builtin_amdgcn_rcpf(1.0f - alpha);
  const float Tn = p.T * inv_om;
  p.T = contrib ? Tn : p.T;
  const float w = contrib ? alpha * Tn : 0.0f;
  float fg = f[0] * p.g[0];
#pragma unroll
  for (int c = 1; c < C; ++c) fg = fmaf(f[c * fstride], p.g[c], fg);
#pragma unroll
  for (int c = 0; c < C; ++c) v[6 + c] = fmaf(w, p.g[c], v[6 + c]);
  const float S_n = fmaf(p.last_alpha, p.F, (1.0f - p.last_alpha) * p.S);
  p.S = contrib ? S_n : p.S;
  p.F = contrib ? fg : p.F;
  float dL_dalpha = (fg - S_n) * Tn;
  if constexpr (BG) dL_dalpha -= p.T_final * inv_om * p.bg_term;
  p.last_alpha = contrib ? alpha : p.last_alpha;
  // alpha = min(0.99, o G): no gradient where clamped
  const float gsel = (contrib && sp.co.w * G < kAlphaMax) ? dL_dalpha : 0.0f;
  v[5] = fmaf(G, gsel, v[5]);
  const float dpow = alpha * gsel;
  v[0] = fmaf(dpow, 2.0f * sp.co.x * dx + sp.co.y * dy, v[0]);
  v[1] = fmaf(dpow, sp.co.y * dx + 2.0f * sp.co.z * dy, v[1]);
  v[2] = fmaf(dpow, -0.5f * dx * dx, v[2]);
  v[3] = fmaf(dpow, -dx * dy, v[3]);
  v[4] = fmaf(dpow, -0.5f * dy * dy, v[4]);
  return contrib;
}

template <int C, int TW, int TH, int BD, int PP = 2, bool BG = true>
__device__ __forceinline__ void blend_bwd_tile2(
    const TileView &tv, int start, int end, const float *__restrict__ uv,
    const float *__restrict__ conic, const float *__restrict__ opacity,
    const float *__restrict__ feat, const int32_t *__restrict__ ids, float bg,
    const float *__restrict__ Tfin, const int32_t *__restrict__ ncon,
    const float *__restrict__ dout, const float *__restrict__ dTfin, float *__restrict__ d_splat,
    int ld, BlendLds<C, BD> &sm, float *__restrict__ d_uv_sep, float *__restrict__ d_op_sep) {
  constexpr int NT = TW * TH / PP;
  constexpr int NW = NT / 64;       // waves (strips) per tile
  constexpr int RW = 64 * PP / TW;  // pixel rows per strip (PP per lane)
  static_assert(BD <= NT, "batch depth exceeds the workgroup");
  static_assert(BD <= 256, "batch slots are stored as bytes");
  static_assert(TH % PP == 0 && NT % 64 == 0 && 64 % TW == 0 && NW <= kMaxTileWaves,
                "tile shape");
  constexpr int QP = GradLayout<C>::QP;
  constexpr int Q = GradLayout<C>::Q;
  const int tx = tv.tile % tv.gx, ty = tv.tile / tv.gx;
  const int px = tx * TW + (int)threadIdx.x % TW;
  const int py0 = ty * TH + PP * ((int)threadIdx.x / TW);
  const int lane = lane_id();
  const int wave = (int)threadIdx.x >> 6;
  const uint64_t lt = lanemask_lt();
  // pixel-centre box of the tile (clipped to the image); strip w = rows
  // [by0 + w RW, by0 + (w + 1) RW)
  const float bx0 = (float)(tx * TW), bx1 = (float)(min(tx * TW + TW, tv.W) - 1);
  const int by0 = ty * TH, byl = min(ty * TH + TH, tv.H) - 1;
  const int my_q = bfly_index<QP>(lane);
  const bool writer = (lane & (64 / QP - 1)) == 0 && my_q < Q;

  // per batch slot, the tile's gradient record summed over its strips (LDS
  // float atomics from each strip's butterfly), then one global atomic per
  // non-zero quantity for the whole tile instead of one per strip.
  // Destination of quantity q of splat g: (du, dv) and dopacity go to their
  // own dense arrays when given (then the record rows are compact, [da, db,
  // dc, dfeature...]), else to the full row
  __shared__ float s_acc[BD * Q];
  auto grad_addr = [&](int q, int g) -> float * {
    if (q < 2 && d_uv_sep != nullptr) return d_uv_sep + 2 * (int64_t)g + q;
    if (q == 5 && d_op_sep != nullptr) return d_op_sep + ((int64_t)g - tv.op_offset);
    const bool compact = d_uv_sep != nullptr && d_op_sep != nullptr;
    return d_splat + (int64_t)g * ld + (compact ? q - (q < 5 ? 2 : 3) : q);
  };

  PixBwd<C> pp[PP];
  int tile_last = 0;
#pragma unroll
  for (int h = 0; h < PP; ++h) {
    PixBwd<C> &p = pp[h];
    const int py = py0 + h;
    const bool inside = px < tv.W && py < tv.H;
    const int64_t pix = tv.pix_base + (int64_t)py * tv.W + px;
    p.T = 1.0f;
    p.last = 0;
    float gT = 0.0f;
    if (inside) {
      p.T = Tfin[pix];
      p.last = ncon[pix];
#pragma unroll
      for (int c = 0; c < C; ++c) p.g[c] = dout[c * tv.chan_stride + pix];
      if (dTfin != nullptr) gT = dTfin[pix];
    } else {
#pragma unroll
      for (int c = 0; c < C; ++c) p.g[c] = 0.0f;
    }
    p.T_final = p.T;
    float gsum = 0.0f;
#pragma unroll
    for (int c = 0; c < C; ++c) gsum += p.g[c];
    p.bg_term = bg * gsum + gT;  // d out / d T_final summed with d T_final
    p.S = p.F = p.last_alpha = 0.0f;
    p.fx = (float)px;
    p.fy = (float)py;
    tile_last = max(tile_last, p.last);
  }
  __syncthreads();  // sm.misc may still be read by the previous tile
  if (threadIdx.x == 0) sm.misc = 0;
  __syncthreads();
  if (tile_last > 0) atomicMax(&sm.misc, tile_last);
  __syncthreads();
  const int stop = start + sm.misc;
  if (stop <= start || end <= start) return;  // uniform

  for (int bend = stop; bend > start; bend -= BD) {
    const int bstart = max(start, bend - BD);
    __syncthreads();
    for (int i = (int)threadIdx.x; i < BD * Q; i += NT) s_acc[i] = 0.0f;
    const int j = bstart + (int)threadIdx.x;
    unsigned mask = 0u;
    if ((int)threadIdx.x < BD && j < bend) {
      const int g = ids[j];
      const float2 uv2 = ld_f2(uv + 2 * (int64_t)g);
      const float3 co3 = ld_f3(conic + 3 * (int64_t)g);
      const float u = uv2.x, v = uv2.y, ca = co3.x, cb = co3.y, cc = co3.z;
      const float o = opacity[g - tv.op_offset];
      SplatLds sp;
      sp.xy = make_float2(u, v);
      sp.co = make_float4(-0.5f * ca, -cb, -0.5f * cc, o);
      sm.sp[threadIdx.x] = sp;
      sm.id[threadIdx.x] = g;
      load_feat_row<C, BD>(feat + (int64_t)g * C, sm.f + threadIdx.x);
      // the forward's reachability test per strip: a splat no pixel of the
      // strip can see at alpha >= kAlphaMin was neither a contributor nor a
      // stopping splat there, so the strip's walk leaves it out
      bool reach;
      const TightSplat ts = tight_setup(u, v, ca, cb, cc, o, reach);
      if (reach) {
#pragma unroll
        for (int w = 0; w < NW; ++w) {
          const int y0 = by0 + w * RW, y1 = min(y0 + RW - 1, byl);
          if (y0 <= y1 && tile_reachable(ts, bx0, bx1, (float)y0, (float)y1)) mask |= 1u << w;
        }
      }
    }
    int pos[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint64_t bal = __ballot((mask >> w) & 1u);
      pos[w] = (int)__popcll(bal & lt);
      if (lane == 0) sm.wcnt[w][wave] = (int)__popcll(bal);
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      if ((mask >> w) & 1u) {
        int off = pos[w];
        for (int q = 0; q < wave; ++q) off += sm.wcnt[w][q];
        sm.list[w][off] = (uint8_t)threadIdx.x;
      }
    }
    int n_mine = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) n_mine += sm.wcnt[wave][q];
    __syncthreads();
    for (int i = n_mine - 1; i >= 0; --i) {  // the strip's list, back to front
      const int k = __builtin_amdgcn_readfirstlane((int)sm.list[wave][i]);
      const int local = bstart + k - start;
      const SplatLds sp = sm.sp[k];
      // wave-uniform skip when no lane's pixels can use the splat (every
      // pixel's test evaluated, no short circuit: the tests feed the step)
      PixNear pn[PP];
      bool any = false;
#pragma unroll
      for (int h = 0; h < PP; ++h) {
        pn[h] = pix_near(pp[h], sp, local);
        any = any | pn[h].near;
      }
      if (__ballot(any) == 0ull) continue;
      float v[QP];
#pragma unroll
      for (int q = 0; q < QP; ++q) v[q] = 0.0f;
      bool contrib = false;
#pragma unroll
      for (int h = 0; h < PP; ++h)
        contrib = pix_bwd_step<C, QP, BG>(pp[h], sp, pn[h], sm.f + k, BD, v) | contrib;
      if (__ballot(contrib) == 0ull) continue;  // wave-uniform
      Bfly<QP, 32>::run(v, lane);
      if (writer) atomicAdd(&s_acc[k * Q + my_q], v[0]);
    }
    __syncthreads();
    // the batch's records to memory: lanes over (slot, quantity), adding 0
    // would change nothing
    const int nq = (bend - bstart) * Q;
    for (int i = (int)threadIdx.x; i < nq; i += NT) {
      const float val = s_acc[i];
      if (val != 0.0f) {
        const int k = i / Q;
        atomicAdd(grad_addr(i - k * Q, sm.id[k]), val);
      }
    }
  }
}

}  // namespace gsr
