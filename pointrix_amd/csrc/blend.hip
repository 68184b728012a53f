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
#include "blend_body.h"
#include "gsr_internal.h"

using namespace gsr;

namespace {

template <int C, int TW, int TH>
__global__ __launch_bounds__(TW *TH) void blend_fwd_kernel(
    const float *__restrict__ uv, const float *__restrict__ conic,
    const float *__restrict__ opacity, const float *__restrict__ feat,
    const int32_t *__restrict__ ids, const int32_t *__restrict__ ranges, float bg, int W, int H,
    int gx, float *__restrict__ out, float *__restrict__ Tfin, int32_t *__restrict__ ncon) {
  __shared__ BlendLds<C, TW * TH> sm;
  const int tile = blockIdx.x;
  const TileView tv{W, H, gx, tile, 0, (int64_t)H * W, 0};
  blend_fwd_tile<C, TW, TH, TW * TH>(tv, ranges[2 * tile], ranges[2 * tile + 1], uv, conic, opacity,
                            feat, ids, bg, out, Tfin, ncon, sm);
}

// PP pixels per lane: TW*TH/PP threads, batches of TW*TH/PP splats.  BG:
// whether the pixels carry a background / T_final gradient term.
template <int C, int TW, int TH, int PP, bool BG>
__global__ __launch_bounds__(TW *TH / PP) void blend_bwd2_kernel(
    const float *__restrict__ uv, const float *__restrict__ conic,
    const float *__restrict__ opacity, const float *__restrict__ feat,
    const int32_t *__restrict__ ids, const int32_t *__restrict__ ranges, float bg, int W, int H,
    int gx, const float *__restrict__ Tfin, const int32_t *__restrict__ ncon,
    const float *__restrict__ dout, const float *__restrict__ dTfin, float *__restrict__ d_splat,
    int ld, float *__restrict__ d_uv, float *__restrict__ d_op) {
  __shared__ BlendLds<C, TW * TH / PP> sm;
  const int tile = blockIdx.x;
  const TileView tv{W, H, gx, tile, 0, (int64_t)H * W, 0};
  blend_bwd_tile2<C, TW, TH, TW * TH / PP, PP, BG>(tv, ranges[2 * tile], ranges[2 * tile + 1], uv,
                                               conic, opacity, feat, ids, bg, Tfin, ncon, dout,
                                               dTfin, d_splat, ld, sm, d_uv, d_op);
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
               const float *d_T_final, float *d_splat, int ld, float *d_uv, float *d_op,
               hipStream_t s) {
  const int gx = (W + tw - 1) / tw, gy = (H + th - 1) / th;
  const unsigned nt = (unsigned)(gx * gy);
#define GSR_BWD_ARGS                                                                          \
  uv, conic, opacity, feature, ids, tile_range, bg, W, H, gx, T_final, n_contrib, d_out, \
      d_T_final, d_splat, ld, d_uv, d_op
  // without a background and a T_final gradient the pixels' background term is
  // exactly zero: the kernel without it gives the same results
  const bool bgt = bg != 0.0f || d_T_final != nullptr;
  if (tw == 16 && th == 16) {
    if (bgt)
      blend_bwd2_kernel<C, 16, 16, 2, true><<<nt, 128, 0, s>>>(GSR_BWD_ARGS);
    else
      blend_bwd2_kernel<C, 16, 16, 2, false><<<nt, 128, 0, s>>>(GSR_BWD_ARGS);
  } else if (tw == 8 && th == 32)
    blend_bwd2_kernel<C, 8, 32, 2, true><<<nt, 128, 0, s>>>(GSR_BWD_ARGS);
  else if (tw == 32 && th == 8)
    blend_bwd2_kernel<C, 32, 8, 2, true><<<nt, 128, 0, s>>>(GSR_BWD_ARGS);
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

int gsr_alpha_blending_bwd_split(const float *uv, const float *conic, const float *opacity,
                                 const float *feature, int C, const int32_t *ids,
                                 const int32_t *tile_range, float bg, int W, int H, int tile_w,
                                 int tile_h, const float *T_final, const int32_t *n_contrib,
                                 const float *d_out, const float *d_T_final, float *d_splat,
                                 int ld, float *d_uv, float *d_opacity, void *stream) {
  const bool compact = d_uv != nullptr && d_opacity != nullptr;
  GSR_REQUIRE(W > 0 && H > 0 && C >= 1 && C <= 16 && ld >= (compact ? 3 : 6) + C,
              "alpha_blending_bwd: bad sizes (C=%d, ld=%d)", C, ld);
  int rc = GSR_OK;
  GSR_DISPATCH_C(C, rc = launch_bwd<CC>(uv, conic, opacity, feature, ids, tile_range, bg, W, H,
                                        tile_w, tile_h, T_final, n_contrib, d_out, d_T_final,
                                        d_splat, ld, d_uv, d_opacity, (hipStream_t)stream));
  GSR_REQUIRE(rc == GSR_OK, "alpha_blending_bwd: unsupported tile shape %dx%d", tile_w, tile_h);
  return gsr_launch_status("alpha_blending_bwd");
}

int gsr_alpha_blending_bwd(const float *uv, const float *conic, const float *opacity,
                           const float *feature, int C, const int32_t *ids,
                           const int32_t *tile_range, float bg, int W, int H, int tile_w,
                           int tile_h, const float *T_final, const int32_t *n_contrib,
                           const float *d_out, const float *d_T_final, float *d_splat, int ld,
                           void *stream) {
  return gsr_alpha_blending_bwd_split(uv, conic, opacity, feature, C, ids, tile_range, bg, W, H,
                                      tile_w, tile_h, T_final, n_contrib, d_out, d_T_final,
                                      d_splat, ld, nullptr, nullptr, stream);
}

}  // extern "C"
