// Multi-view batched rendering: one scene, V cameras of uneven resolution,
// one launch per stage.  Replaces the per-view loop of
// MsplatRender.render_batch (pointrix/model/renderer/msplat.py:160-213).
//
// * Preprocess: one thread per splat loops over the views.  Position, scale,
//   rotation and SH are read from HBM once (SH only if some view keeps the
//   splat) and the 3D covariance is formed once; projection, EWA, SH colour
//   and the feature pack are per view.  Outputs are indexed by the virtual
//   id j = view * N + splat.
// * Binning: the one-view pipeline over the V*N virtual splats, with each
//   view's tile grid offset into one global tile numbering (binning.hip).
// * Blending: a persistent grid pulls global tiles from a device work queue
//   (uneven views, no per-view launches); each tile looks up its view.
// * Backward: blending writes per-virtual-splat gradient records; one thread
//   per splat sums its views' contributions and writes each gradient once.
//
// View tables (device): vf float[V][20] = intr(4) | extr 3x4 (12) | centre(3)
// | pad; vi int32[V][8] = W, H, gx, gy, tile_off, pix_off, -, -.
#include "blend_body.h"
#include "gsr_internal.h"
#include "gsr_math.h"

using namespace gsr;

namespace {

constexpr int kThreads = 256;
constexpr int kMaxViews = 64;
constexpr int kVF = 20;
constexpr int kVI = 8;
constexpr int kFlagDepth = 1;
constexpr int kFlagNormal = 2;
// forward only, intermediates not inspected: for culled (view, splat) pairs
// write only radius = tiles = 0 (nothing downstream reads their uv, depth,
// conic or features), saving most of the per-pair output traffic
constexpr int kFlagSparse = 4;

__device__ __forceinline__ CamV cam_from(const float *p) {
  CamV c;
  c.fx = p[0];
  c.fy = p[1];
  c.cx = p[2];
  c.cy = p[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) c.R[3 * i + j] = p[4 + 4 * i + j];
    c.t[i] = p[4 + 4 * i + 3];
  }
  return c;
}

template <int KS, int KE>
__device__ __forceinline__ void mv_load_sh(const float *__restrict__ shs,
                                           const float *__restrict__ rest, int64_t i,
                                           float (&sh)[KE * 3]) {
  if (rest != nullptr) {
    const float3 dc = ld_f3(shs + 3 * i);
    sh[0] = dc.x;
    sh[1] = dc.y;
    sh[2] = dc.z;
    // the rest row as 16-byte accesses plus the tail
    const float *rb = rest + i * ((KS - 1) * 3);
    constexpr int nr = KE * 3 - 3;
#pragma unroll
    for (int j = 0; j + 4 <= nr; j += 4) {
      const float4 t = ld_f4u(rb + j);
      sh[3 + j] = t.x;
      sh[3 + j + 1] = t.y;
      sh[3 + j + 2] = t.z;
      sh[3 + j + 3] = t.w;
    }
#pragma unroll
    for (int j = nr & ~3; j < nr; ++j) sh[3 + j] = rb[j];
  } else {
    const float *b = shs + i * (KS * 3);
#pragma unroll
    for (int j = 0; j < KE * 3; ++j) sh[j] = b[j];
  }
}

__device__ __forceinline__ float3 mv_view_dir(float3 x, const float *c, float &len) {
  const float dx = x.x - c[0], dy = x.y - c[1], dz = x.z - c[2];
  len = sqrtf(dx * dx + dy * dy + dz * dz);
  return make_float3(dx / len, dy / len, dz / len);
}

template <int KS, int KE>
__global__ __launch_bounds__(kThreads) void mv_preprocess_fwd_kernel(
    const float *__restrict__ xyz, const float *__restrict__ scale,
    const float *__restrict__ rot, const float *__restrict__ shs,
    const float *__restrict__ shs_rest, int64_t n, int V, const float *__restrict__ vf,
    const int32_t *__restrict__ vi, int tw, int th, float nearest, float extent, int flags,
    int C, float *__restrict__ uv, float *__restrict__ depth, float *__restrict__ conic,
    int32_t *__restrict__ radius, int32_t *__restrict__ tiles, float *__restrict__ feat) {
  __shared__ float s_vf[kMaxViews * kVF];
  __shared__ int s_vi[kMaxViews * kVI];
  for (int k = threadIdx.x; k < V * kVF; k += blockDim.x) s_vf[k] = vf[k];
  for (int k = threadIdx.x; k < V * kVI; k += blockDim.x) s_vi[k] = vi[k];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float3 x = ld_f3(xyz + 3 * i);
  const float3 s = ld_f3(scale + 3 * i);
  const float4 q = ld_f4u(rot + 4 * i);
  float cov[6];
  covariance3d(s, q, cov);
  float sh[KE * 3];
  bool have_sh = false;
  for (int v = 0; v < V; ++v) {
    const float *pv = s_vf + v * kVF;
    const int *iv = s_vi + v * kVI;
    const CamV cam = cam_from(pv);
    const int W = iv[0], H = iv[1];
    const int64_t j = (int64_t)v * n + i;
    float u, vv, d;
    project(cam, x, nearest, extent, W, H, u, vv, d);
    float con[3] = {0.f, 0.f, 0.f};
    float f[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int r = 0, t = 0;
    if (d != 0.0f) {
      EwaFwd e;
      ewa_forward(cam, x, cov, W, H, e);
      int rr;
      float cc[3];
      if (conic_radius(e.a, e.b, e.c, cc, rr)) {
        const Rect qr = tile_rect(u, vv, rr, iv[2], iv[3], tw, th);
        const int area = (qr.x1 - qr.x0) * (qr.y1 - qr.y0);
        if (area > 0) {
          r = rr;
          t = area;
          con[0] = cc[0];
          con[1] = cc[1];
          con[2] = cc[2];
        }
      }
      if (r > 0) {
        if (!have_sh) {
          mv_load_sh<KS, KE>(shs, shs_rest, i, sh);
          have_sh = true;
        }
        float len;
        const float3 dir = mv_view_dir(x, pv + 16, len);
        float Y[16];
        sh_basis(dir, KE, Y);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          float acc = sh[c] * Y[0];
#pragma unroll
          for (int k = 1; k < KE; ++k) acc = acc + sh[3 * k + c] * Y[k];
          f[c] = fmaxf(acc + 0.5f, 0.0f);
        }
        int o = 3;
        if (flags & kFlagDepth) f[o++] = d;
        if (flags & kFlagNormal) {
          NormalFwd nf;
          float nc[3];
          const float cen[3] = {pv[16], pv[17], pv[18]};
          splat_normal(s, q, x, cen, cam, nf, nc);
          f[o] = nc[0];
          f[o + 1] = nc[1];
          f[o + 2] = nc[2];
        }
      }
    }
    radius[j] = r;
    tiles[j] = t;
    if (r == 0 && (flags & kFlagSparse)) continue;
    uv[2 * j] = u;
    uv[2 * j + 1] = vv;
    depth[j] = d;
    st_f3(conic + 3 * j, con[0], con[1], con[2]);
    for (int c = 0; c < C; ++c) feat[j * C + c] = f[c];
  }
}

template <int KS, int KE, bool CAMGRAD>
__global__ __launch_bounds__(kThreads) void mv_preprocess_bwd_kernel(
    const float *__restrict__ xyz, const float *__restrict__ scale,
    const float *__restrict__ rot, const float *__restrict__ shs,
    const float *__restrict__ shs_rest, int64_t n, int V, const float *__restrict__ vf,
    const int32_t *__restrict__ vi, float nearest, float extent, int flags, int C,
    const int32_t *__restrict__ radius, const float *__restrict__ rec, int ld,
    float *__restrict__ d_xyz, float *__restrict__ d_scale, float *__restrict__ d_rot,
    float *__restrict__ d_opacity, float *__restrict__ d_shs, float *__restrict__ d_shs_rest,
    float *__restrict__ d_vf) {
  constexpr int kRest = (KS - 1) * 3;
  __shared__ float s_vf[kMaxViews * kVF];
  __shared__ int s_vi[kMaxViews * kVI];
  __shared__ float s_cam[kThreads / 64][32];
  __shared__ __align__(16) float s_rest[KS > 1 ? kThreads * kRest : 4];
  for (int k = threadIdx.x; k < V * kVF; k += blockDim.x) s_vf[k] = vf[k];
  for (int k = threadIdx.x; k < V * kVI; k += blockDim.x) s_vi[k] = vi[k];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < n;
  const int lane = lane_id();
  float3 x = make_float3(0.f, 0.f, 0.f), s = make_float3(1.f, 1.f, 1.f);
  float4 q = make_float4(1.f, 0.f, 0.f, 0.f);
  if (active) {
    x = ld_f3(xyz + 3 * i);
    s = ld_f3(scale + 3 * i);
    q = ld_f4u(rot + 4 * i);
  }
  float cov[6];
  covariance3d(s, q, cov);
  float sh[KE * 3], dsh[KE * 3];
#pragma unroll
  for (int k = 0; k < KE * 3; ++k) dsh[k] = 0.0f;
  bool have_sh = false;
  float3 gx = make_float3(0.f, 0.f, 0.f);
  float gS[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float gq[4] = {0.f, 0.f, 0.f, 0.f};
  float gop = 0.0f;
  // the splat's radius in every view, requested together (one memory round
  // trip instead of one per view), as a bit mask of the views that raster it
  uint64_t vis = 0;
  if (active) {
    for (int v0 = 0; v0 < V; v0 += 16) {
      int rv[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) rv[k] = v0 + k < V ? radius[(int64_t)(v0 + k) * n + i] : 0;
#pragma unroll
      for (int k = 0; k < 16; ++k) vis |= (uint64_t)(rv[k] > 0) << (v0 + k);
    }
  }
  for (int v = 0; v < V; ++v) {
    const float *pv = s_vf + v * kVF;
    const int *iv = s_vi + v * kVI;
    float gcam[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) gcam[k] = 0.0f;
    float gcen[3] = {0.f, 0.f, 0.f};
    const int64_t j = (int64_t)v * n + i;
    if ((vis >> v) & 1u) {
      const CamV cam = cam_from(pv);
      const int W = iv[0], H = iv[1];
      const float *g = rec + j * ld;
      gop += g[5];
      if (!have_sh) {
        mv_load_sh<KS, KE>(shs, shs_rest, i, sh);
        have_sh = true;
      }
      float len;
      const float3 dir = mv_view_dir(x, pv + 16, len);
      float Y[16];
      sh_basis(dir, KE, Y);
      float coef[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) coef[k] = 0.0f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float acc = sh[c] * Y[0];
#pragma unroll
        for (int k = 1; k < KE; ++k) acc = acc + sh[3 * k + c] * Y[k];
        const float gc = (acc + 0.5f >= 0.0f) ? g[6 + c] : 0.0f;
#pragma unroll
        for (int k = 0; k < KE; ++k) {
          dsh[3 * k + c] += gc * Y[k];
          coef[k] += gc * sh[3 * k + c];
        }
      }
      if (KE > 1) {
        float3 gdir = make_float3(0.f, 0.f, 0.f);
        sh_basis_vjp(dir, KE, coef, gdir);
        const float dp = dir.x * gdir.x + dir.y * gdir.y + dir.z * gdir.z;
        const float il = 1.0f / len;
        const float3 gd = make_float3((gdir.x - dir.x * dp) * il, (gdir.y - dir.y * dp) * il,
                                      (gdir.z - dir.z * dp) * il);
        gx.x += gd.x;
        gx.y += gd.y;
        gx.z += gd.z;
        gcen[0] -= gd.x;
        gcen[1] -= gd.y;
        gcen[2] -= gd.z;
      }
      int o = 9;
      float gdepth = 0.0f;
      if (flags & kFlagDepth) gdepth = g[o++];
      if (flags & kFlagNormal) {
        NormalFwd nf;
        float nc[3];
        const float cen[3] = {pv[16], pv[17], pv[18]};
        splat_normal(s, q, x, cen, cam, nf, nc);
        const float gn[3] = {g[o], g[o + 1], g[o + 2]};
        splat_normal_vjp(q, cam, nf, gn, gq, gcam);
      }
      project_vjp(cam, x, g[0], g[1], gdepth, gx, gcam);
      const float g3[3] = {g[2], g[3], g[4]};
      float gSv[6];
      ewa_vjp(cam, x, cov, W, H, g3, gx, gSv, gcam);
#pragma unroll
      for (int k = 0; k < 6; ++k) gS[k] += gSv[k];
    }
    if constexpr (CAMGRAD) {  // per-view camera gradients, block-reduced
      float gv[32];
#pragma unroll
      for (int k = 0; k < 16; ++k) gv[k] = gcam[k];
      gv[16] = gcen[0];
      gv[17] = gcen[1];
      gv[18] = gcen[2];
#pragma unroll
      for (int k = 19; k < 32; ++k) gv[k] = 0.0f;
      Bfly<32, 32>::run(gv, lane);
      if ((lane & 1) == 0) s_cam[threadIdx.x >> 6][bfly_index<32>(lane)] = gv[0];
      __syncthreads();
      if (threadIdx.x < 19) {
        float sum = 0.0f;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) sum += s_cam[w][threadIdx.x];
        if (sum != 0.0f) atomicAdd(d_vf + v * kVF + threadIdx.x, sum);
      }
      __syncthreads();
    }
  }
  if (active) {
    float gs[3], gq2[4];
    covariance3d_vjp(s, q, gS, gs, gq2);
    st_f3(d_xyz + 3 * i, gx.x, gx.y, gx.z);
    d_opacity[i] = gop;
#pragma unroll
    for (int k = 0; k < 3; ++k) d_scale[3 * i + k] = gs[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) d_rot[4 * i + k] = gq[k] + gq2[k];
  }
  if (d_shs_rest != nullptr) {
    if (active) {
      st_f3(d_shs + 3 * i, dsh[0], dsh[1], dsh[2]);
    }
    if constexpr (KS > 1) {
      // stage the rows, then store the block's contiguous range of
      // d_shs_rest with consecutive lanes on consecutive 16-byte words
      if (active) {
#pragma unroll
        for (int j = 3; j < KS * 3; ++j)
          s_rest[threadIdx.x * kRest + (j - 3)] = (j < KE * 3) ? dsh[j % (KE * 3)] : 0.0f;
      }
      __syncthreads();
      const int64_t first = (int64_t)blockIdx.x * kThreads;
      const int rows = (int)min((int64_t)kThreads, n - first);
      const int total = rows * kRest;
      float *dst = d_shs_rest + first * kRest;
      if ((reinterpret_cast<uintptr_t>(dst) & 15) == 0) {
        float4 *d4 = reinterpret_cast<float4 *>(dst);
        const float4 *s4 = reinterpret_cast<const float4 *>(s_rest);
        for (int k = threadIdx.x; k < total / 4; k += kThreads) d4[k] = s4[k];
        for (int k = (total / 4) * 4 + threadIdx.x; k < total; k += kThreads) dst[k] = s_rest[k];
      } else {
        for (int k = threadIdx.x; k < total; k += kThreads) dst[k] = s_rest[k];
      }
    }
  } else if (active) {
    float *ds = d_shs + i * (KS * 3);
#pragma unroll
    for (int j = 0; j < KS * 3; ++j) ds[j] = (j < KE * 3) ? dsh[j % (KE * 3)] : 0.0f;
  }
}

// ---------------------------------------------------------------------------
// Persistent multi-view blending over a device work queue of global tiles.
// ---------------------------------------------------------------------------
__device__ __forceinline__ TileView mv_tile_view(const int *s_vi, int V, int tile, int64_t npv,
                                                 int64_t chan_stride) {
  int v = 0;
  for (int k = 1; k < V; ++k)
    if (s_vi[k * kVI + 4] <= tile) v = k;
  const int *iv = s_vi + v * kVI;
  return TileView{iv[0], iv[1], iv[2], tile - iv[4], (int64_t)iv[5], chan_stride,
                  (int64_t)v * npv};
}

template <int C, int TW, int TH, int BD>
__global__ __launch_bounds__(TW *TH) void mv_blend_fwd_kernel(
    const float *__restrict__ uv, const float *__restrict__ conic,
    const float *__restrict__ opacity, const float *__restrict__ feat,
    const int32_t *__restrict__ ids, const int32_t *__restrict__ ranges, float bg, int V,
    const int32_t *__restrict__ vi, int64_t npv, int T_total, int64_t P_total,
    int *__restrict__ queue, float *__restrict__ out, float *__restrict__ Tfin,
    int32_t *__restrict__ ncon) {
  __shared__ BlendLds<C, BD> sm;
  __shared__ int s_vi[kMaxViews * kVI];
  __shared__ int s_tile;
  for (int k = threadIdx.x; k < V * kVI; k += blockDim.x) s_vi[k] = vi[k];
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_tile = atomicAdd(queue, 1);
    __syncthreads();
    const int tile = s_tile;
    if (tile >= T_total) break;  // uniform
    const TileView tv = mv_tile_view(s_vi, V, tile, npv, P_total);
    blend_fwd_tile<C, TW, TH, BD>(tv, ranges[2 * tile], ranges[2 * tile + 1], uv, conic,
                                  opacity, feat, ids, bg, out, Tfin, ncon, sm);
  }
}

// Two pixels per lane (TW*TH/2 threads), as the one-view 16x16 backward; the
// batch depth is capped at the thread count.
template <int C, int TW, int TH, int BD, bool BG>
__global__ __launch_bounds__(TW *TH / 2) void mv_blend_bwd_kernel(
    const float *__restrict__ uv, const float *__restrict__ conic,
    const float *__restrict__ opacity, const float *__restrict__ feat,
    const int32_t *__restrict__ ids, const int32_t *__restrict__ ranges, float bg, int V,
    const int32_t *__restrict__ vi, int64_t npv, int T_total, int64_t P_total,
    int *__restrict__ queue, const float *__restrict__ Tfin, const int32_t *__restrict__ ncon,
    const float *__restrict__ dout, const float *__restrict__ dTfin, float *__restrict__ d_splat,
    int ld) {
  constexpr int BDX = BD < TW * TH / 2 ? BD : TW * TH / 2;
  __shared__ BlendLds<C, BDX> sm;
  __shared__ int s_vi[kMaxViews * kVI];
  __shared__ int s_tile;
  for (int k = threadIdx.x; k < V * kVI; k += blockDim.x) s_vi[k] = vi[k];
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_tile = atomicAdd(queue, 1);
    __syncthreads();
    const int tile = s_tile;
    if (tile >= T_total) break;  // uniform
    const TileView tv = mv_tile_view(s_vi, V, tile, npv, P_total);
    blend_bwd_tile2<C, TW, TH, BDX, 2, BG>(tv, ranges[2 * tile], ranges[2 * tile + 1], uv, conic,
                                       opacity, feat, ids, bg, Tfin, ncon, dout, dTfin, d_splat,
                                       ld, sm, nullptr, nullptr);
  }
}

__global__ void zero_int_kernel(int *p) { *p = 0; }

int persistent_blocks() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
  }
  return cus * 4;
}

inline int effective_k(int K, int degree) {
  if (degree < 0) return K;
  const int k = (degree + 1) * (degree + 1);
  return k < K ? k : K;
}

template <template <int, int> class Launch, class... A>
int dispatch_sh(int K, int KE, A... args) {
#define GSR_CASE(KS_, KE_) \
  if (K == KS_ && KE == KE_) return Launch<KS_, KE_>::run(args...);
  GSR_CASE(1, 1)
  GSR_CASE(4, 1) GSR_CASE(4, 4)
  GSR_CASE(9, 1) GSR_CASE(9, 4) GSR_CASE(9, 9)
  GSR_CASE(16, 1) GSR_CASE(16, 4) GSR_CASE(16, 9) GSR_CASE(16, 16)
#undef GSR_CASE
  return GSR_ERR_ARG;
}

template <int KS, int KE>
struct LaunchPreFwd {
  template <class... A>
  static int run(unsigned nb, hipStream_t st, A... a) {
    mv_preprocess_fwd_kernel<KS, KE><<<nb, kThreads, 0, st>>>(a...);
    return GSR_OK;
  }
};
template <int KS, int KE>
struct LaunchPreBwd {
  template <class... A>
  static int run(unsigned nb, hipStream_t st, bool camgrad, A... a) {
    if (camgrad)
      mv_preprocess_bwd_kernel<KS, KE, true><<<nb, kThreads, 0, st>>>(a...);
    else
      mv_preprocess_bwd_kernel<KS, KE, false><<<nb, kThreads, 0, st>>>(a...);
    return GSR_OK;
  }
};

// blend dispatch over channel count (3, 4, 6, 7), tile shape and batch depth
// bg_term: the backward keeps the background term (bg != 0 or a T_final gradient); the
// variant without it exists for 16x16 tiles only
template <bool FWD, class... A>
int dispatch_blend(int C, int tw, int th, int bd, bool bg_term, unsigned nb, hipStream_t st,
                   A... a) {
#define GSR_K(C_, TW_, TH_, BD_)                                                        \
  if (C == C_ && tw == TW_ && th == TH_ && bd == BD_) {                                 \
    if constexpr (FWD) {                                                                \
      mv_blend_fwd_kernel<C_, TW_, TH_, BD_><<<nb, TW_ * TH_, 0, st>>>(a...);           \
    } else if constexpr (TW_ == 16 && TH_ == 16) {                                      \
      if (bg_term)                                                                      \
        mv_blend_bwd_kernel<C_, TW_, TH_, BD_, true><<<nb, TW_ * TH_ / 2, 0, st>>>(a...); \
      else                                                                              \
        mv_blend_bwd_kernel<C_, TW_, TH_, BD_, false><<<nb, TW_ * TH_ / 2, 0, st>>>(a...); \
    } else {                                                                            \
      mv_blend_bwd_kernel<C_, TW_, TH_, BD_, true><<<nb, TW_ * TH_ / 2, 0, st>>>(a...); \
    }                                                                                   \
    return GSR_OK;                                                                      \
  }
#define GSR_SHAPES(C_)                                                                  \
  GSR_K(C_, 16, 16, 256) GSR_K(C_, 16, 16, 128) GSR_K(C_, 16, 16, 64)                   \
  GSR_K(C_, 8, 32, 256) GSR_K(C_, 8, 32, 128) GSR_K(C_, 8, 32, 64)                      \
  GSR_K(C_, 32, 8, 256) GSR_K(C_, 32, 8, 128) GSR_K(C_, 32, 8, 64)
  GSR_SHAPES(3) GSR_SHAPES(4) GSR_SHAPES(6) GSR_SHAPES(7)
#undef GSR_SHAPES
#undef GSR_K
  return GSR_ERR_ARG;
}

}  // namespace

extern "C" {

int gsr_mv_preprocess_fwd(const float *xyz, const float *scale, const float *rot,
                          const float *shs, const float *shs_rest, int64_t n, int K, int degree,
                          int V, const float *view_f, const int32_t *view_i, int tile_w,
                          int tile_h, float nearest, float extent, int flags, float *uv,
                          float *depth, float *conic, int32_t *radius, int32_t *tiles,
                          float *feat, void *stream) {
  GSR_REQUIRE(n >= 0 && V >= 1 && V <= kMaxViews && (flags & ~7) == 0 && tile_w > 0 &&
                  tile_h > 0,
              "mv_preprocess_fwd: bad arguments (V=%d)", V);
  if (n == 0) return GSR_OK;
  const int C = 3 + ((flags & kFlagDepth) ? 1 : 0) + ((flags & kFlagNormal) ? 3 : 0);
  const int rc = dispatch_sh<LaunchPreFwd>(
      K, effective_k(K, degree), gsr_blocks(n, kThreads), (hipStream_t)stream, xyz, scale, rot,
      shs, shs_rest, n, V, view_f, view_i, tile_w, tile_h, nearest, extent, flags, C, uv, depth,
      conic, radius, tiles, feat);
  GSR_REQUIRE(rc == GSR_OK, "mv_preprocess_fwd: unsupported SH size K=%d", K);
  return gsr_launch_status("mv_preprocess_fwd");
}

int gsr_mv_preprocess_bwd(const float *xyz, const float *scale, const float *rot,
                          const float *shs, const float *shs_rest, int64_t n, int K, int degree,
                          int V, const float *view_f, const int32_t *view_i, float nearest,
                          float extent, int flags, const int32_t *radius, const float *d_splat,
                          int ld, float *d_xyz, float *d_scale, float *d_rot, float *d_opacity,
                          float *d_shs, float *d_shs_rest, float *d_view_f, void *stream) {
  const int C = 3 + ((flags & kFlagDepth) ? 1 : 0) + ((flags & kFlagNormal) ? 3 : 0);
  GSR_REQUIRE(n >= 0 && V >= 1 && V <= kMaxViews && (flags & ~3) == 0 && ld >= 6 + C,
              "mv_preprocess_bwd: bad arguments (V=%d, ld=%d)", V, ld);
  if (n == 0) return GSR_OK;
  const int rc = dispatch_sh<LaunchPreBwd>(
      K, effective_k(K, degree), gsr_blocks(n, kThreads), (hipStream_t)stream,
      d_view_f != nullptr, xyz, scale, rot,
      shs, shs_rest, n, V, view_f, view_i, nearest, extent, flags, C, radius, d_splat, ld,
      d_xyz, d_scale, d_rot, d_opacity, d_shs, d_shs_rest, d_view_f);
  GSR_REQUIRE(rc == GSR_OK, "mv_preprocess_bwd: unsupported SH size K=%d", K);
  return gsr_launch_status("mv_preprocess_bwd");
}

int gsr_mv_alpha_blending_fwd(const float *uv, const float *conic, const float *opacity,
                              const float *feature, int C, const int32_t *ids,
                              const int32_t *tile_range, float bg, int64_t n_per_view, int V,
                              const int32_t *view_i, int total_tiles, int64_t total_pixels,
                              int tile_w, int tile_h, int batch_depth, int32_t *queue,
                              float *out, float *T_final, int32_t *n_contrib, void *stream) {
  GSR_REQUIRE(V >= 1 && V <= kMaxViews && total_tiles >= 0 && total_pixels >= 0,
              "mv_alpha_blending_fwd: bad sizes");
  if (total_tiles == 0) return GSR_OK;
  hipStream_t st = (hipStream_t)stream;
  zero_int_kernel<<<1, 1, 0, st>>>(queue);
  const unsigned nb = (unsigned)min(total_tiles, persistent_blocks());
  const int rc = dispatch_blend<true>(C, tile_w, tile_h, batch_depth, true, nb, st, uv, conic, opacity,
                                      feature, ids, tile_range, bg, V, view_i, n_per_view,
                                      total_tiles, total_pixels, queue, out, T_final, n_contrib);
  GSR_REQUIRE(rc == GSR_OK, "mv_alpha_blending_fwd: unsupported C=%d / tile %dx%d / depth %d", C,
              tile_w, tile_h, batch_depth);
  return gsr_launch_status("mv_alpha_blending_fwd");
}

int gsr_mv_alpha_blending_bwd(const float *uv, const float *conic, const float *opacity,
                              const float *feature, int C, const int32_t *ids,
                              const int32_t *tile_range, float bg, int64_t n_per_view, int V,
                              const int32_t *view_i, int total_tiles, int64_t total_pixels,
                              int tile_w, int tile_h, int batch_depth, int32_t *queue,
                              const float *T_final, const int32_t *n_contrib, const float *d_out,
                              const float *d_T_final, float *d_splat, int ld, void *stream) {
  GSR_REQUIRE(V >= 1 && V <= kMaxViews && total_tiles >= 0 && ld >= 6 + C,
              "mv_alpha_blending_bwd: bad sizes");
  if (total_tiles == 0) return GSR_OK;
  hipStream_t st = (hipStream_t)stream;
  zero_int_kernel<<<1, 1, 0, st>>>(queue);
  // half-size workgroups: twice as many keep the same waves per CU
  const unsigned nb = (unsigned)min(total_tiles, 2 * persistent_blocks());
  const int rc = dispatch_blend<false>(C, tile_w, tile_h, batch_depth,
                                       bg != 0.0f || d_T_final != nullptr, nb, st, uv, conic,
                                       opacity, feature, ids, tile_range, bg, V, view_i,
                                       n_per_view, total_tiles, total_pixels, queue, T_final,
                                       n_contrib, d_out, d_T_final, d_splat, ld);
  GSR_REQUIRE(rc == GSR_OK, "mv_alpha_blending_bwd: unsupported C=%d / tile %dx%d / depth %d",
              C, tile_w, tile_h, batch_depth);
  return gsr_launch_status("mv_alpha_blending_bwd");
}

}  // extern "C"
