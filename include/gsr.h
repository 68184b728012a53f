/*
 * gsr.h -- C ABI of libgsr, the gfx950 (MI355X) differentiable Gaussian
 * rasteriser.
 *
 * Every entry point replaces one msplat call made by the reference renderer,
 * pointrix/model/renderer/msplat.py (cited per function), forward and
 * backward.  All pointers are device pointers unless noted; tensors are dense
 * row-major fp32 / int32 / uint8 arrays owned by the caller.  Work is enqueued
 * on `stream` (a hipStream_t) and never synchronises the host; the calls are
 * capturable into a HIP graph.  Return value: 0 on success, a negative
 * hipError_t-derived code on a launch failure, GSR_ERR_ARG (-1000) on a bad
 * argument; gsr_last_error() describes the last failure of the calling thread.
 *
 * Conventions (fixed by this library, following 3DGS practice): intrinsics
 * intr = (fx, fy, cx, cy); extrinsics extr = 3x4 row-major world-to-camera
 * [R|t]; pixel centres at integer coordinates; 16x16 tiles unless a tile shape
 * argument says otherwise; cov3d packed (xx, xy, xz, yy, yz, zz); quaternions
 * wxyz.
 */
#ifndef GSR_H_
#define GSR_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int gsr_version(void);
const char *gsr_last_error(void);

/* ---- msplat.project_point (msplat.py:109-113) ---------------------------
 * uv[N,2], depth[N] (depth 0 and uv 0 for culled points: z <= nearest or
 * |2u/W - 1| > extent or |2v/H - 1| > extent). */
int gsr_project_point_fwd(const float *xyz, int64_t n, const float *intr, const float *extr,
                          int W, int H, float nearest, float extent, float *uv, float *depth,
                          void *stream);
/* Backward.  depth is the forward output (0 marks culled points).  d_intr[4]
 * and d_extr[12] are accumulated (+=) and may be NULL. */
int gsr_project_point_bwd(const float *xyz, int64_t n, const float *intr, const float *extr,
                          const float *depth, const float *d_uv, const float *d_depth,
                          float *d_xyz, float *d_intr, float *d_extr, void *stream);

/* ---- msplat.compute_cov3d (msplat.py:118) -------------------------------
 * cov3d[N,6] = R(q) diag(s)^2 R(q)^T packed; 0 where visible[i] == 0. */
int gsr_compute_cov3d_fwd(const float *scale, const float *rot, const uint8_t *visible,
                          int64_t n, float *cov3d, void *stream);
int gsr_compute_cov3d_bwd(const float *scale, const float *rot, const uint8_t *visible,
                          int64_t n, const float *d_cov3d, float *d_scale, float *d_rot,
                          void *stream);

/* ---- msplat.ewa_project (msplat.py:121-130) ------------------------------
 * conic[N,3] (inverse EWA 2D covariance), radius[N] (3-sigma, pixels) and
 * tiles[N] (tiles touched; 16x16 tiles).  All 0 for invisible points. */
int gsr_ewa_project_fwd(const float *xyz, const float *cov3d, int64_t n, const float *intr,
                        const float *extr, const float *uv, int W, int H,
                        const uint8_t *visible, float *conic, int32_t *radius,
                        int32_t *tiles, void *stream);
/* Backward.  radius is the forward output (0 marks skipped points).  d_intr
 * and d_extr are accumulated and may be NULL. */
int gsr_ewa_project_bwd(const float *xyz, const float *cov3d, int64_t n, const float *intr,
                        const float *extr, int W, int H, const int32_t *radius,
                        const float *d_conic, float *d_xyz, float *d_cov3d, float *d_intr,
                        float *d_extr, void *stream);

/* ---- msplat.compute_sh (msplat.py:104) -----------------------------------
 * out[n,c] = sum_{k < K'} shs[n,c,k] * Y_k(dirs[n]) with K' = min(K,
 * (degree+1)^2) (degree < 0: K' = K).  shs is addressed through element
 * strides (s_n, s_c, s_k) so the permuted [N,16,3] -> [N,3,16] view the
 * reference passes needs no copy. */
int gsr_compute_sh_fwd(const float *shs, int64_t s_n, int64_t s_c, int64_t s_k,
                       const float *dirs, int64_t n, int C, int K, int degree, float *out,
                       void *stream);
/* d_shs is dense [N,C,K] (overwritten, zero for k >= K'); d_dirs[N,3]
 * overwritten, may be NULL. */
int gsr_compute_sh_bwd(const float *shs, int64_t s_n, int64_t s_c, int64_t s_k,
                       const float *dirs, int64_t n, int C, int K, int degree,
                       const float *d_out, float *d_shs, float *d_dirs, void *stream);

/* ---- fused preprocess of MsplatRender.render_iter (msplat.py:94-139) -----
 * One pass per splat replacing the view-direction normalisation (:94-96),
 * the SH warm-up mask + compute_sh + "+0.5, clamp" (:98-105), project_point,
 * compute_cov3d, ewa_project (:107-130) and RenderFeatures.combine (:137-139).
 * shs: [N,K,3] (K in 1,4,9,16), or -- when shs_rest != NULL -- the DC
 * coefficients shs[N,1,3] and shs_rest[N,K-1,3] (pointrix's two SH
 * parameters, gaussian_points.py:86-90, read without concatenation); summed
 * over the first min(K,(degree+1)^2) coefficients (degree < 0: all).  flags: 1 = depth channel, 2 = camera-space
 * normal channel (shortest splat axis facing the camera,
 * examples/supervise/model.py:50-74).  feat[N,C], C = 3 (+1) (+3):
 * rgb | depth | normal.  uv/depth/conic/radius/tiles equal those of the
 * separate ops; cov3d[N,6] may be NULL.  flags 4 = raw parameters: scale is
 * log scale, rot an unnormalised quaternion, and the kernel applies exp and
 * F.normalize (GaussianPointCloud's activations, gaussian_points.py:30-35,
 * 70-81); with flag 4, opacity_raw[N] (logits) -> opacity_act[N] = sigmoid
 * (both NULL, or both set).  flags 8 = plain colour: rgb is the SH sum
 * itself, without the +0.5 and the clamp (MsplatNormalRender,
 * examples/supervise/renderer.py:41-42). */
int gsr_preprocess_fwd(const float *xyz, const float *scale, const float *rot, const float *shs,
                       const float *shs_rest, int64_t n, int K, int degree, const float *intr, const float *extr,
                       const float *cam_center, int W, int H, float nearest, float extent,
                       int flags, float *uv, float *depth, float *cov3d, float *conic,
                       int32_t *radius, int32_t *tiles, float *feat, const float *opacity_raw,
                       float *opacity_act, void *stream);
/* Backward: d_uv[N,2], d_depth[N] (may be NULL), d_conic[N,3], d_feat[N,C],
 * each addressed with its row stride (ld_*; e.g. views into the d_splat
 * records of gsr_alpha_blending_bwd) -> d_xyz[N,3], d_scale[N,3], d_rot[N,4],
 * d_shs[N,K,3] (or d_shs[N,1,3] + d_shs_rest[N,K-1,3] when split; all
 * overwritten), d_intr[4], d_extr[12], d_center[3] (accumulated; each may be
 * NULL).  With flag 4 the scale and rotation gradients are w.r.t. the raw
 * parameters, and d_opacity_raw[N] = d_opacity_act[i * ld_opacity_act] *
 * s (1 - s) when opacity_raw / d_opacity_act / d_opacity_raw are set (all
 * NULL otherwise; d_opacity_act may be a column of the blend's records). */
int gsr_preprocess_bwd(const float *xyz, const float *scale, const float *rot, const float *shs,
                       const float *shs_rest, int64_t n, int K, int degree, const float *intr, const float *extr,
                       const float *cam_center, int W, int H, float nearest, float extent,
                       int flags, const int32_t *radius, const float *d_uv, int ld_uv,
                       const float *d_depth, const float *d_conic, int ld_conic,
                       const float *d_feat, int ld_feat, float *d_xyz, float *d_scale,
                       float *d_rot, float *d_shs, float *d_shs_rest, float *d_intr,
                       float *d_extr, float *d_center, const float *opacity_raw,
                       const float *d_opacity_act, int ld_opacity_act, float *d_opacity_raw,
                       void *stream);

/* ---- msplat.sort_gaussian (msplat.py:133-135) ----------------------------
 * Two phases around one host read of M (the instance count):
 *   gsr_sort_count: per-splat tile counts from (uv, radius), depth-sort of
 *     the splats that touch a tile, exclusive scan -> *m_out (device int32).
 *   gsr_sort_emit: one (tile, id) pair per touched tile, stable tile sort,
 *     tile ranges.  ids_out[M] is ordered by (tile, f32 depth bits, splat id);
 *     tile_range[T,2] = [first, last+1), (0, 0) for empty tiles.
 * ws_gauss must hold gsr_sort_ws_gauss_bytes(n) bytes; ws_inst
 * gsr_sort_ws_inst_bytes(M) bytes; both 256-byte aligned. */
size_t gsr_sort_ws_gauss_bytes(int64_t n);
size_t gsr_sort_ws_inst_bytes(int64_t m);
int gsr_sort_count(const float *uv, const float *depth, const int32_t *radius, int64_t n,
                   int W, int H, int tile_w, int tile_h, void *ws_gauss, int32_t *m_out,
                   void *stream);
int gsr_sort_emit(const float *uv, const int32_t *radius, int64_t n, int W, int H, int tile_w,
                  int tile_h, void *ws_gauss, int64_t m, void *ws_inst, int32_t *ids_out,
                  int32_t *tile_range, void *stream);
/* m in gsr_sort_emit is the instance capacity of ids_out / ws_inst; the
 * kernels use min(device count, m), so a fixed capacity (no host read, e.g.
 * inside a captured graph) is safe: if the device count exceeds it the
 * emission is truncated in bounds and the caller detects it from *m_out. */

/* Exact binning of one view has three implementations, all giving the same
 * ids_out and tile_range (process-wide switch, for comparisons; returns the
 * previous mode, ignores values outside 0..2; no msplat counterpart):
 *   2 (default): grids of at most 256 x 256 tiles are binned by rows -- one
 *     segment per splat and rectangle row sorted by the row, then each row's
 *     segments expanded and sorted by the column, the last pass writing
 *     ids_out and tile_range; larger grids as 1;
 *   1: the first tile radix pass generates its items from the depth-ordered
 *     splats where it uses them, the other passes sort lists;
 *   0: the instance list is written, then sorted by list passes.
 * gsr_sort_count prepares mode 2 (one more scan) for every grid it can take. */
int gsr_set_binning_mode(int mode);

/* gsr_sort_count + gsr_sort_emit in one call that allocates nothing: no
 * counterpart in msplat or in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The workspaces are carved from a
 * file-scope device array of 64 MiB that each GPU the library runs on holds
 * until the process ends (its address is taken once per device, on the first
 * call there).  n splats and the instance capacity m_cap of ids_out must fit
 * it together: gsr_sort_ws_gauss_bytes(n) + gsr_sort_ws_inst_bytes(m_cap) <=
 * 64 MiB (about 500k splats with 2.5M instances), else GSR_ERR_ARG before
 * anything is queued.  Capacity semantics as gsr_sort_emit with a fixed m:
 * *m_out receives the device instance count, and if it exceeds m_cap the
 * sort covers the first m_cap instances in bounds.  Every call on one GPU
 * uses that GPU's one array: calls queued on one stream are ordered, calls
 * that overlap on one GPU (other streams, other host threads) race. */
int gsr_sort_gaussian_noalloc(const float *uv, const float *depth, const int32_t *radius,
                              int64_t n, int W, int H, int tile_w, int tile_h, int64_t m_cap,
                              int32_t *m_out, int32_t *ids_out, int32_t *tile_range,
                              void *stream);

/* Opacity-aware ("tight") binning, no counterpart in msplat: as
 * gsr_sort_count / gsr_sort_emit, but a splat is emitted only into the tiles
 * of its rectangle where it can reach alpha >= 1/255 at some pixel centre
 * (a conservative test on the minimum of its conic's quadratic form over the
 * tile; conic [N,3] and opacity [N] as the blend reads them).  Blending the
 * result gives the same images bit for bit with fewer instances; ids_out and
 * tile_range then differ from the msplat contract (each tile's list is a
 * subsequence of the exact one).  conic == opacity == NULL is the exact
 * contract. */
int gsr_sort_count_tight(const float *uv, const float *depth, const int32_t *radius,
                         const float *conic, const float *opacity, int64_t n, int W, int H,
                         int tile_w, int tile_h, void *ws_gauss, int32_t *m_out, void *stream);
/* As gsr_sort_count_tight, and M also reaches m_host (pinned host memory) as
 * soon as it is known -- before the depth sort.  When m_host has a device
 * address (hipHostGetDevicePointer) the kernel that sums M stores it there;
 * otherwise a copy is queued.  `event` (a hipEvent_t, may be NULL) is recorded
 * after that.  A host reads M while the device still sorts, instead of
 * draining the stream to read m_out: by waiting on the event, or, with event
 * NULL, by polling a value it set negative before the call (no event: no
 * barrier in the stream ahead of the depth sort).  No msplat counterpart
 * (msplat.sort_gaussian, msplat.py:133-135, sizes its outputs internally). */
int gsr_sort_count_async(const float *uv, const float *depth, const int32_t *radius,
                         const float *conic, const float *opacity, int64_t n, int W, int H,
                         int tile_w, int tile_h, void *ws_gauss, int32_t *m_out,
                         int32_t *m_host, void *event, void *stream);
/* As gsr_sort_count_tight, and *m_high (device int32) receives max(*m_high,
 * M): the largest count since the caller last reset it, stored by the kernel
 * that sums M (no launch of its own).  For a caller that replays a captured
 * sort with a fixed instance capacity and checks once, after many replays,
 * whether any of them overflowed it.  No msplat counterpart. */
int gsr_sort_count_high(const float *uv, const float *depth, const int32_t *radius,
                        const float *conic, const float *opacity, int64_t n, int W, int H,
                        int tile_w, int tile_h, void *ws_gauss, int32_t *m_out, int32_t *m_high,
                        void *stream);
int gsr_sort_emit_tight(const float *uv, const int32_t *radius, const float *conic,
                        const float *opacity, int64_t n, int W, int H, int tile_w, int tile_h,
                        void *ws_gauss, int64_t m, void *ws_inst, int32_t *ids_out,
                        int32_t *tile_range, void *stream);

/* ---- multi-view batches (MsplatRender.render_batch, msplat.py:160-213) ---
 * V cameras of any resolution over one scene, one launch per stage.  View
 * tables: view_f float[V][20] = intr(4) | extr 3x4 row-major (12) | centre(3)
 * | pad; view_i int32[V][8] = W, H, gx, gy, first global tile, first pixel,
 * 0, 0 (tiles of tile_w x tile_h; views' tiles and pixels numbered
 * consecutively).  Per-view outputs are indexed by j = view * N + splat;
 * images are packed [C, total_pixels].  V <= 64. */
/* flags: 1 depth channel, 2 normal channel, 4 "sparse": culled (view, splat)
 * pairs get radius = tiles = 0 and nothing else written (no later stage,
 * backward included, reads anything else of them). */
int gsr_mv_preprocess_fwd(const float *xyz, const float *scale, const float *rot,
                          const float *shs, const float *shs_rest, int64_t n, int K, int degree,
                          int V, const float *view_f, const int32_t *view_i, int tile_w,
                          int tile_h, float nearest, float extent, int flags, float *uv,
                          float *depth, float *conic, int32_t *radius, int32_t *tiles,
                          float *feat, void *stream);
/* d_splat[V*N, ld]: the records of gsr_mv_alpha_blending_bwd.  Gradients
 * (position, scale, rotation, opacity, SH) are summed over views and written
 * once; d_view_f[V][20] (may be NULL) accumulates camera gradients. */
int gsr_mv_preprocess_bwd(const float *xyz, const float *scale, const float *rot,
                          const float *shs, const float *shs_rest, int64_t n, int K, int degree,
                          int V, const float *view_f, const int32_t *view_i, float nearest,
                          float extent, int flags, const int32_t *radius, const float *d_splat,
                          int ld, float *d_xyz, float *d_scale, float *d_rot, float *d_opacity,
                          float *d_shs, float *d_shs_rest, float *d_view_f, void *stream);
/* conic [V*N,3] / opacity [N]: both NULL for the exact tile contract, both
 * set for tight binning (see gsr_sort_count_tight).  Every view's tile grid
 * (view_i columns 2, 3) must be under 65536 tiles per side. */
int gsr_mv_sort_count(const float *uv, const float *depth, const int32_t *radius,
                      const float *conic, const float *opacity, int64_t n_per_view, int V,
                      const int32_t *view_i, int tile_w, int tile_h, void *ws_gauss,
                      int32_t *m_out, void *stream);
/* As gsr_mv_sort_count, with M also in m_host (pinned) ahead of the depth
 * sort and `event` recorded after it (see gsr_sort_count_async). */
int gsr_mv_sort_count_async(const float *uv, const float *depth, const int32_t *radius,
                            const float *conic, const float *opacity, int64_t n_per_view, int V,
                            const int32_t *view_i, int tile_w, int tile_h, void *ws_gauss,
                            int32_t *m_out, int32_t *m_host, void *event, void *stream);
int gsr_mv_sort_emit(const float *uv, const int32_t *radius, const float *conic,
                     const float *opacity, int64_t n_per_view, int V, const int32_t *view_i,
                     int total_tiles, int tile_w, int tile_h, void *ws_gauss, int64_t m,
                     void *ws_inst, int32_t *ids_out, int32_t *tile_range, void *stream);
/* Persistent blending: min(total_tiles, 4 x CUs) workgroups pull global
 * tiles from *queue (one device int32 of caller scratch, reset in-stream).
 * C in {3,4,6,7}; tiles 16x16, 8x32, 32x8; batch_depth 64, 128 or 256 splats
 * staged per LDS batch. */
int gsr_mv_alpha_blending_fwd(const float *uv, const float *conic, const float *opacity,
                              const float *feature, int C, const int32_t *ids,
                              const int32_t *tile_range, float bg, int64_t n_per_view, int V,
                              const int32_t *view_i, int total_tiles, int64_t total_pixels,
                              int tile_w, int tile_h, int batch_depth, int32_t *queue,
                              float *out, float *T_final, int32_t *n_contrib, void *stream);
int gsr_mv_alpha_blending_bwd(const float *uv, const float *conic, const float *opacity,
                              const float *feature, int C, const int32_t *ids,
                              const int32_t *tile_range, float bg, int64_t n_per_view, int V,
                              const int32_t *view_i, int total_tiles, int64_t total_pixels,
                              int tile_w, int tile_h, int batch_depth, int32_t *queue,
                              const float *T_final, const int32_t *n_contrib, const float *d_out,
                              const float *d_T_final, float *d_splat, int ld, void *stream);

/* ---- msplat.alpha_blending (msplat.py:148-151) ---------------------------
 * out[C,H,W] = sum_i f_i alpha_i T_i + bg * T_final, front to back over the
 * tile's depth-sorted splats.  Also writes T_final[H,W] and
 * n_contrib[H,W] (1 + index within the tile range of the last contributing
 * splat), which the backward pass needs.  C in 1..16. */
int gsr_alpha_blending_fwd(const float *uv, const float *conic, const float *opacity,
                           const float *feature, int C, const int32_t *ids,
                           const int32_t *tile_range, float bg, int W, int H, int tile_w,
                           int tile_h, float *out, float *T_final, int32_t *n_contrib,
                           void *stream);
/* Backward: d_out[C,H,W], d_T_final[H,W] (may be NULL) -> per-splat gradient
 * records accumulated (+=) into d_splat[N, ld] (ld >= 6 + C; caller
 * zero-fills): [d_u, d_v, d_conic_a, d_conic_b, d_conic_c, d_opacity,
 * d_feature_0 .. d_feature_{C-1}]. */
int gsr_alpha_blending_bwd(const float *uv, const float *conic, const float *opacity,
                           const float *feature, int C, const int32_t *ids,
                           const int32_t *tile_range, float bg, int W, int H, int tile_w,
                           int tile_h, const float *T_final, const int32_t *n_contrib,
                           const float *d_out, const float *d_T_final, float *d_splat, int ld,
                           void *stream);
/* As gsr_alpha_blending_bwd, with (d_u, d_v) accumulated into d_uv[N,2] and
 * d_opacity into d_opacity[N] (dense, caller zero-fills) instead of record
 * slots 0, 1 and 5, when those pointers are not NULL.  With both given the
 * records are compact: [d_conic_a, d_conic_b, d_conic_c, d_feature_0 ..
 * d_feature_{C-1}] (ld >= 3 + C). */
int gsr_alpha_blending_bwd_split(const float *uv, const float *conic, const float *opacity,
                                 const float *feature, int C, const int32_t *ids,
                                 const int32_t *tile_range, float bg, int W, int H, int tile_w,
                                 int tile_h, const float *T_final, const int32_t *n_contrib,
                                 const float *d_out, const float *d_T_final, float *d_splat,
                                 int ld, float *d_uv, float *d_opacity, void *stream);

/* ---- fused L1 loss (pointrix/model/loss.py:27, mean |x - y|) -----------
 * Graph-safe: two-stage fixed-order reduction into *out (device scalar),
 * using partial[gsr_l1_loss_partials(n)] caller scratch; no atomics, no
 * memset.  Backward: dx = sign(x - y) * g_out[0] / n. */
int gsr_l1_loss_partials(int64_t n);
int gsr_l1_loss_fwd(const float *x, const float *y, int64_t n, float *partial, float *out,
                    void *stream);
int gsr_l1_loss_bwd(const float *x, const float *y, int64_t n, const float *g_out, float *dx,
                    void *stream);

/* ---- weighted sum of L1 terms, sum_k weight[k] * mean |x[k] - y[k]| (the
 * per-feature L1 terms of a training loss, pointrix/model/loss.py:27 each,
 * e.g. l1(rgb) + 0.1 l1(depth) + 0.1 l1(normal)).  Same reduction as
 * gsr_l1_loss_fwd over all terms at once: partial[gsr_l1_sum_partials(terms)]
 * caller scratch, *out a device scalar; fixed order, no atomics, no memset.
 * Backward: dx[k] = sign(x[k] - y[k]) * g_out[0] * weight[k] / n[k] for each
 * term whose dx[k] is not NULL (dx is ignored by the forward). */
#define GSR_MAX_L1_TERMS 8
typedef struct {
  const float *x[GSR_MAX_L1_TERMS];
  const float *y[GSR_MAX_L1_TERMS];
  float *dx[GSR_MAX_L1_TERMS];
  int64_t n[GSR_MAX_L1_TERMS];
  float weight[GSR_MAX_L1_TERMS];
  int nterms;
} GsrL1Terms;
int gsr_l1_terms_bytes(void); /* sizeof(GsrL1Terms), for bindings */
int gsr_l1_sum_partials(const GsrL1Terms *terms); /* -1: bad table */
int gsr_l1_sum_fwd(const GsrL1Terms *terms, float *partial, float *out, void *stream);
int gsr_l1_sum_bwd(const GsrL1Terms *terms, const float *g_out, void *stream);

/* ---- fused L1 + SSIM (pointrix/model/loss.py:27, :73-116; combined as in
 * base_model.py:112-124) over `planes` planes of W x H ([B,C,H,W]: B*C planes).
 * out[0] = mean SSIM (11-tap Gaussian window, sigma 1.5, zero padding),
 * out[1] = mean |x - y|; partial[gsr_ssim_partials(planes, W, H)] caller
 * scratch; fixed-order sums, no atomics, no memset (graph-safe).  Backward:
 * dx = d(out)/dx weighted by g[0] (for out[0]) and g[1] (for out[1]), both
 * device scalars. */
int gsr_ssim_partials(int planes, int W, int H);
int gsr_l1_ssim_fwd(const float *x, const float *y, int planes, int W, int H, float *partial,
                    float *out, void *stream);
int gsr_l1_ssim_bwd(const float *x, const float *y, int planes, int W, int H, const float *g,
                    float *dx, void *stream);

/* ---- fused Adam (torch.optim.Adam as pointrix configures it,
 * pointrix/optimizer/__init__.py:61-63, optimizer.py:128-140; weight decay 0,
 * no amsgrad) over every tensor of a step in one launch.  table[ntab] lives in
 * device memory; entry k covers workgroups [first_chunk, first_chunk +
 * ceil(n / gsr_adam_chunk())), nchunks is the total.  Per element:
 *   m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
 *   p -= step_size * m / (sqrt(v) / bc2_sqrt + eps)
 * with step_size = lr / (1 - b1^t), bc2_sqrt = sqrt(1 - b2^t); zero_grad != 0
 * also writes g = 0. */
typedef struct {
  float *p, *g, *m, *v;
  int64_t n;
  int64_t first_chunk;
  float step_size, bc2_sqrt;
} GsrAdamTensor;
int gsr_adam_chunk(void);
int gsr_adam_entry_bytes(void); /* sizeof(GsrAdamTensor), for bindings */
int gsr_adam_step(const GsrAdamTensor *table, int ntab, int64_t nchunks, double beta1,
                  double beta2, double eps, int zero_grad, void *stream);

/* ---- densification statistics (controller/gs.py:318-333, 259-284) -------
 * For every splat with visibility[i] != 0: g = (sum over the V views of
 * grad[v*view_stride + i*row_stride + {0,1}]) * (sx, sy) (the views summed in
 * order, then scaled: normalize_grad's (W/2, H/2), or (1, 1));
 * grad_accum[i] += |g|; acc_steps[i] += 1;
 * max_radii[i] = max(max_radii[i], radii[i]).  Other splats are untouched. */
int gsr_densify_stats(const float *grad, int64_t view_stride, int64_t row_stride, int V,
                      int64_t n, const uint8_t *visibility, const int32_t *radii, float sx,
                      float sy, float *grad_accum, float *acc_steps, float *max_radii,
                      void *stream);
/* Densify / prune with optimizer-state remap (point_utils.py:51-108,
 * controller/gs.py:182-247) as one row gather: for every entry of table
 * (device memory; src [*, width], dst [rows, width] float32) dst row r = src
 * row index[r], or zeros where index[r] < 0 (new rows' Adam moments);
 * rows * width < 2^32 for every entry. */
typedef struct {
  const float *src;
  float *dst;
  int64_t width;
} GsrRowCopy;
int gsr_row_copy_bytes(void); /* sizeof(GsrRowCopy), for bindings */
int gsr_gather_rows(const GsrRowCopy *table, int ntab, const int32_t *index, int64_t rows,
                    void *stream);

/* uv_points.grad [n,2] = the (du, dv) of the blend backward's records
 * (d_splat[i*ld + 0..1]) scaled by (sx, sy) = (W/2, H/2) (msplat.py:141-145). */
int gsr_uv_grad_scale(const float *d_splat, int ld, int64_t n, float sx, float sy, float *out,
                      void *stream);

/* ---- point-cloud activations (gaussian_points.py:30-35,70-81) -----------
 * scale[N,3] = exp(log_scale), rot[N,4] = quat / max(|quat|, 1e-12),
 * opacity[N] = sigmoid(logit).  Backward: gradients of the three outputs
 * (each may be NULL = zero) -> d_log_scale, d_quat, d_logit (each may be NULL
 * = not wanted), from the forward's inputs and outputs; rot is not read (the
 * unit quaternion is recomputed from quat) and may be NULL.  quat, rot, g_rot
 * and d_quat are 16-byte aligned. */
int gsr_activate_fwd(const float *log_scale, const float *quat, const float *logit, int64_t n,
                     float *scale, float *rot, float *opacity, void *stream);
int gsr_activate_bwd(const float *quat, const float *scale, const float *rot,
                     const float *opacity, const float *g_scale, const float *g_rot,
                     const float *g_opacity, int64_t n, float *d_log_scale, float *d_quat,
                     float *d_logit, void *stream);

/* ---- camera pose (camera_model.py:92-175, utils/pose.py:40-84) -----------
 * q[C,4] wxyz (normalised inside as q / max(|q|, 1e-12)), t[C,3] ->
 * E[C,4,4] = [R | t; 0 0 0 1] and center[C,3] = -R^T t (may be NULL).
 * Backward: dE[C,4,4] and dcenter[C,3] (either may be NULL) -> dq[C,4],
 * dt[C,3] (overwritten). */
int gsr_camera_pose_fwd(const float *q, const float *t, int C, float *E, float *center,
                        void *stream);
int gsr_camera_pose_bwd(const float *q, const float *t, int C, const float *dE,
                        const float *dcenter, float *dq, float *dt, void *stream);

/* ---- output planes (RenderFeatures.split, renderer_utils.py:57-71, and the
 * normal mapping of examples/supervise/renderer.py:89-96) -------------------
 * in[C,P] (P = H*W) -> out[C,P]: planes copied, except the three planes from
 * normal_first (< 0: none), which become (n / |n| + 1) / 2 with |n| over the
 * three.  Backward: the gradient of each output group (channels first ..
 * first+count-1, groups in channel order covering all C; grad NULL = zero)
 * -> d_in[C,P] (overwritten). */
#define GSR_MAX_PLANE_GROUPS 8
typedef struct {
  const float *grad[GSR_MAX_PLANE_GROUPS];
  int first[GSR_MAX_PLANE_GROUPS];
  int count[GSR_MAX_PLANE_GROUPS];
  int ngroups;
} GsrPlaneGrads;
int gsr_plane_grads_bytes(void); /* sizeof(GsrPlaneGrads), for bindings */
int gsr_planes_fwd(const float *in, int C, int64_t P, int normal_first, float *out,
                   void *stream);
int gsr_planes_bwd(const float *in, int C, int64_t P, int normal_first,
                   const GsrPlaneGrads *grads, float *d_in, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* GSR_H_ */
