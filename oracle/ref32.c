/*
 * ref32.c -- FP32 C restatement of the forward rasterisation path (test
 * infrastructure; built by oracle/build_ref32.py, loaded by oracle/ref32.py).
 *
 * Restates, in plain sequential C, the forward of the six msplat ops the
 * reference renderer calls (pointrix/model/renderer/msplat.py:104-151) under
 * the conventions of oracle/torch64.py, with every float operation in the
 * order the HIP kernels perform it (compiled with -ffp-contract=off: one
 * rounding per written operator, fmaf only where written).  The HIP results
 * must equal these bit for bit: radius, tiles touched, sorted ids, tile
 * ranges and per-pixel contributor counts (the integer outputs), and also the
 * float outputs.  The exp used by the blend is defined here from its method
 * (Cody-Waite reduction + degree-7 Taylor + exact 2^k scaling), not copied.
 *
 * It also restates the backward of the path (ref_blend_bwd, ref_preprocess_bwd:
 * reverse-mode chain rule of the forward above, FP32), which the FP64
 * autograd of oracle/torch64.py pins; float sums there are not in the
 * kernels' order, so the backward is compared within a tolerance.  Loops
 * over pixels, tiles and splats run on OpenMP threads; every output element
 * is written by one thread in a fixed order, so the results do not depend on
 * the thread count (the camera gradients are summed per thread in double).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define TILE_DEFAULT 16

static float ln2_hi(void) { return 0.693145751953125f; } /* 15-bit head of ln 2 */
static float ln2_lo(void) { return (float)(log(2.0) - 0.693145751953125); }

static float exp_neg(float x) {
  const float log2e = (float)(1.0 / log(2.0));
  float kf, r, p;
  x = fmaxf(x, -80.0f);
  kf = rintf(x * log2e);
  r = fmaf(kf, -ln2_hi(), x);
  r = fmaf(kf, -ln2_lo(), r);
  p = (float)(1.0 / 5040.0);
  p = fmaf(p, r, (float)(1.0 / 720.0));
  p = fmaf(p, r, (float)(1.0 / 120.0));
  p = fmaf(p, r, (float)(1.0 / 24.0));
  p = fmaf(p, r, (float)(1.0 / 6.0));
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)kf);
}

float ref_exp_neg(float x) { return exp_neg(x); }

/* camera: intr = fx fy cx cy; extr 3x4 row-major */
static void cam_point(const float *E, const float *x, float *p) {
  p[0] = E[0] * x[0] + E[1] * x[1] + E[2] * x[2] + E[3];
  p[1] = E[4] * x[0] + E[5] * x[1] + E[6] * x[2] + E[7];
  p[2] = E[8] * x[0] + E[9] * x[1] + E[10] * x[2] + E[11];
}

void ref_project(const float *xyz, int64_t n, const float *K, const float *E, int W, int H,
                 float nearest, float extent, float *uv, float *depth) {
  for (int64_t i = 0; i < n; ++i) {
    float p[3], u, v, lo_x, hi_x, lo_y, hi_y;
    cam_point(E, xyz + 3 * i, p);
    uv[2 * i] = 0.0f;
    uv[2 * i + 1] = 0.0f;
    depth[i] = 0.0f;
    if (!(p[2] > nearest)) continue;
    u = K[0] * p[0] / p[2] + K[2];
    v = K[1] * p[1] / p[2] + K[3];
    lo_x = 0.5f * (float)W * (1.0f - extent);
    hi_x = 0.5f * (float)W * (1.0f + extent);
    lo_y = 0.5f * (float)H * (1.0f - extent);
    hi_y = 0.5f * (float)H * (1.0f + extent);
    if (u < lo_x || u > hi_x || v < lo_y || v > hi_y) continue;
    uv[2 * i] = u;
    uv[2 * i + 1] = v;
    depth[i] = p[2];
  }
}

static void rotmat(const float *q, float *R) {
  const float w = q[0], x = q[1], y = q[2], z = q[3];
  const float x2 = x * x, y2 = y * y, z2 = z * z, w2 = w * w;
  const float xy = x * y, zw = z * w, xz = x * z, yw = y * w, yz = y * z, xw = x * w;
  R[0] = x2 - y2 - z2 + w2;
  R[1] = 2.0f * (xy - zw);
  R[2] = 2.0f * (xz + yw);
  R[3] = 2.0f * (xy + zw);
  R[4] = -x2 + y2 - z2 + w2;
  R[5] = 2.0f * (yz - xw);
  R[6] = 2.0f * (xz - yw);
  R[7] = 2.0f * (yz + xw);
  R[8] = -x2 - y2 + z2 + w2;
}

void ref_cov3d(const float *scale, const float *rot, const uint8_t *vis, int64_t n, float *cov) {
  static const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
  for (int64_t i = 0; i < n; ++i) {
    float R[9], M[9];
    int k, r;
    for (k = 0; k < 6; ++k) cov[6 * i + k] = 0.0f;
    if (!vis[i]) continue;
    rotmat(rot + 4 * i, R);
    for (r = 0; r < 3; ++r) {
      M[3 * r + 0] = R[3 * r + 0] * scale[3 * i + 0];
      M[3 * r + 1] = R[3 * r + 1] * scale[3 * i + 1];
      M[3 * r + 2] = R[3 * r + 2] * scale[3 * i + 2];
    }
    for (k = 0; k < 6; ++k) {
      const int a = ii[k], b = jj[k];
      cov[6 * i + k] = M[3 * a] * M[3 * b] + M[3 * a + 1] * M[3 * b + 1] + M[3 * a + 2] * M[3 * b + 2];
    }
  }
}

static void rect_of(float u, float v, int r, int gx, int gy, int tw, int th, int *x0, int *x1,
                    int *y0, int *y1) {
  const float rf = (float)r, ftw = (float)tw, fth = (float)th;
  *x0 = (int)fminf(fmaxf((u - rf) / ftw, 0.0f), (float)gx);
  *x1 = (int)fminf(fmaxf(((u + rf) + (ftw - 1.0f)) / ftw, 0.0f), (float)gx);
  *y0 = (int)fminf(fmaxf((v - rf) / fth, 0.0f), (float)gy);
  *y1 = (int)fminf(fmaxf(((v + rf) + (fth - 1.0f)) / fth, 0.0f), (float)gy);
}

void ref_ewa(const float *xyz, const float *cov, int64_t n, const float *K, const float *E,
             const float *uv, int W, int H, const uint8_t *vis, float *conic, int32_t *radius,
             int32_t *tiles) {
  const int gx = (W + TILE_DEFAULT - 1) / TILE_DEFAULT, gy = (H + TILE_DEFAULT - 1) / TILE_DEFAULT;
  const float lowpass = 0.3f, jclamp = 1.3f;
  for (int64_t i = 0; i < n; ++i) {
    float p[3], z, limx, limy, txz, tyz, tx, ty, J00, J02, J11, J12, T[6], U[6], Sf[9];
    float a, b, c, det, inv, mid, disc, lam;
    const float *S = cov + 6 * i;
    int r, x0, x1, y0, y1, j, q;
    conic[3 * i] = conic[3 * i + 1] = conic[3 * i + 2] = 0.0f;
    radius[i] = 0;
    tiles[i] = 0;
    if (!vis[i]) continue;
    cam_point(E, xyz + 3 * i, p);
    z = p[2];
    limx = jclamp * (0.5f * (float)W / K[0]);
    limy = jclamp * (0.5f * (float)H / K[1]);
    txz = fminf(fmaxf(p[0] / z, -limx), limx);
    tyz = fminf(fmaxf(p[1] / z, -limy), limy);
    tx = txz * z;
    ty = tyz * z;
    J00 = K[0] / z;
    J02 = -(K[0] * tx) / (z * z);
    J11 = K[1] / z;
    J12 = -(K[1] * ty) / (z * z);
    for (j = 0; j < 3; ++j) {
      T[j] = J00 * E[j] + J02 * E[8 + j];
      T[3 + j] = J11 * E[4 + j] + J12 * E[8 + j];
    }
    Sf[0] = S[0]; Sf[1] = S[1]; Sf[2] = S[2];
    Sf[3] = S[1]; Sf[4] = S[3]; Sf[5] = S[4];
    Sf[6] = S[2]; Sf[7] = S[4]; Sf[8] = S[5];
    for (q = 0; q < 2; ++q)
      for (j = 0; j < 3; ++j)
        U[3 * q + j] = T[3 * q] * Sf[j] + T[3 * q + 1] * Sf[3 + j] + T[3 * q + 2] * Sf[6 + j];
    a = U[0] * T[0] + U[1] * T[1] + U[2] * T[2];
    b = U[0] * T[3] + U[1] * T[4] + U[2] * T[5];
    c = U[3] * T[3] + U[4] * T[4] + U[5] * T[5];
    a = a + lowpass;
    c = c + lowpass;
    det = a * c - b * b;
    if (!(det > 0.0f)) continue;
    inv = 1.0f / det;
    mid = 0.5f * (a + c);
    disc = fmaxf(mid * mid - det, 0.1f);
    lam = mid + sqrtf(disc);
    r = (int)ceilf(3.0f * sqrtf(lam));
    rect_of(uv[2 * i], uv[2 * i + 1], r, gx, gy, TILE_DEFAULT, TILE_DEFAULT, &x0, &x1, &y0, &y1);
    if ((x1 - x0) * (y1 - y0) <= 0) continue;
    conic[3 * i] = c * inv;
    conic[3 * i + 1] = -b * inv;
    conic[3 * i + 2] = a * inv;
    radius[i] = r;
    tiles[i] = (x1 - x0) * (y1 - y0);
  }
}

/* real SH, 3DGS sign convention; constants from their closed forms */
void ref_sh(const float *shs, const float *dirs, int64_t n, int C, int K, int degree, float *out) {
  const double pi = 3.14159265358979323846;
  const float C0 = (float)(0.5 / sqrt(pi));
  const float C1 = (float)sqrt(3.0 / (4.0 * pi));
  const float C20 = (float)(0.5 * sqrt(15.0 / pi));
  const float C22 = (float)(0.25 * sqrt(5.0 / pi));
  const float C24 = (float)(0.25 * sqrt(15.0 / pi));
  const float C30 = (float)(0.25 * sqrt(35.0 / (2.0 * pi)));
  const float C31 = (float)(0.5 * sqrt(105.0 / pi));
  const float C32 = (float)(0.25 * sqrt(21.0 / (2.0 * pi)));
  const float C33 = (float)(0.25 * sqrt(7.0 / pi));
  const float C35 = (float)(0.25 * sqrt(105.0 / pi));
  int Ke = K;
  if (degree >= 0 && (degree + 1) * (degree + 1) < K) Ke = (degree + 1) * (degree + 1);
  for (int64_t i = 0; i < n; ++i) {
    const float x = dirs[3 * i], y = dirs[3 * i + 1], z = dirs[3 * i + 2];
    float Y[16];
    int c, k;
    Y[0] = C0;
    if (Ke > 1) {
      Y[1] = -C1 * y;
      Y[2] = C1 * z;
      Y[3] = -C1 * x;
    }
    if (Ke > 4) {
      const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
      Y[4] = C20 * xy;
      Y[5] = -C20 * yz;
      Y[6] = C22 * (2.0f * zz - xx - yy);
      Y[7] = -C20 * xz;
      Y[8] = C24 * (xx - yy);
      if (Ke > 9) {
        Y[9] = -C30 * y * (3.0f * xx - yy);
        Y[10] = C31 * xy * z;
        Y[11] = -C32 * y * (4.0f * zz - xx - yy);
        Y[12] = C33 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
        Y[13] = -C32 * x * (4.0f * zz - xx - yy);
        Y[14] = C35 * z * (xx - yy);
        Y[15] = -C30 * x * (xx - 3.0f * yy);
      }
    }
    for (c = 0; c < C; ++c) {
      const float *s = shs + (i * C + c) * K;
      float acc = s[0] * Y[0];
      for (k = 1; k < Ke; ++k) acc = acc + s[k] * Y[k];
      out[i * C + c] = acc;
    }
  }
}

/* ---- binning ------------------------------------------------------------ */
typedef struct {
  uint64_t key;
  uint32_t seq;
  int32_t id;
} Inst;

static int inst_cmp(const void *pa, const void *pb) {
  const Inst *a = (const Inst *)pa, *b = (const Inst *)pb;
  if (a->key != b->key) return a->key < b->key ? -1 : 1;
  return a->seq < b->seq ? -1 : (a->seq > b->seq ? 1 : 0);
}

int64_t ref_sort_count(const float *uv, const int32_t *radius, int64_t n, int W, int H, int tw,
                       int th) {
  const int gx = (W + tw - 1) / tw, gy = (H + th - 1) / th;
  int64_t m = 0;
  for (int64_t i = 0; i < n; ++i) {
    int x0, x1, y0, y1;
    if (radius[i] <= 0) continue;
    rect_of(uv[2 * i], uv[2 * i + 1], radius[i], gx, gy, tw, th, &x0, &x1, &y0, &y1);
    m += (int64_t)(x1 - x0) * (y1 - y0);
  }
  return m;
}

/* stable LSD radix sort of 32-bit keys (4 passes of 8 bits) carrying a
 * 32-bit value; in/out are swapped through tmp */
static void radix32(uint32_t *key, int32_t *val, uint32_t *tk, int32_t *tv, int64_t n) {
  for (int shift = 0; shift < 32; shift += 8) {
    int64_t cnt[257];
    memset(cnt, 0, sizeof(cnt));
    for (int64_t i = 0; i < n; ++i) ++cnt[((key[i] >> shift) & 255u) + 1];
    for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
    for (int64_t i = 0; i < n; ++i) {
      const int64_t o = cnt[(key[i] >> shift) & 255u]++;
      tk[o] = key[i];
      tv[o] = val[i];
    }
    memcpy(key, tk, sizeof(uint32_t) * (size_t)n);
    memcpy(val, tv, sizeof(int32_t) * (size_t)n);
  }
}

/* ids[M] sorted by (tile, f32 depth bits, splat id); ranges[T][2]; returns M.
 * The splats that touch a tile are sorted by (depth bits, id) -- a stable
 * radix sort of the index-ordered list -- and dealt into per-tile buckets in
 * that order, which is the order of the (tile, depth bits, id) sort; the
 * brute-force sort of the keys pins it (tests/test_oracle.py). */
int64_t ref_sort(const float *uv, const float *depth, const int32_t *radius, int64_t n, int W,
                 int H, int tw, int th, int32_t *ids, int32_t *ranges) {
  const int gx = (W + tw - 1) / tw, gy = (H + th - 1) / th;
  const int64_t T = (int64_t)gx * gy;
  int64_t nv = 0, m = 0;
  uint32_t *key, *tk;
  int32_t *val, *tv;
  int64_t *cur;
  for (int64_t i = 0; i < n; ++i) {
    int x0, x1, y0, y1;
    if (radius[i] <= 0) continue;
    rect_of(uv[2 * i], uv[2 * i + 1], radius[i], gx, gy, tw, th, &x0, &x1, &y0, &y1);
    if ((x1 - x0) * (y1 - y0) > 0) ++nv;
  }
  key = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(nv + 1));
  tk = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(nv + 1));
  val = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv + 1));
  tv = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv + 1));
  cur = (int64_t *)calloc((size_t)T + 1, sizeof(int64_t));
  nv = 0;
  for (int64_t i = 0; i < n; ++i) {
    int x0, x1, y0, y1, tx, ty;
    if (radius[i] <= 0) continue;
    rect_of(uv[2 * i], uv[2 * i + 1], radius[i], gx, gy, tw, th, &x0, &x1, &y0, &y1);
    if ((x1 - x0) * (y1 - y0) <= 0) continue;
    memcpy(key + nv, depth + i, 4);
    val[nv++] = (int32_t)i;
    for (ty = y0; ty < y1; ++ty)
      for (tx = x0; tx < x1; ++tx) ++cur[ty * gx + tx + 1];
  }
  radix32(key, val, tk, tv, nv);
  for (int64_t t = 0; t < T; ++t) cur[t + 1] += cur[t];
  m = cur[T];
  for (int64_t t = 0; t < T; ++t) {
    const int64_t a = cur[t], b = cur[t + 1];
    ranges[2 * t] = b > a ? (int32_t)a : 0;
    ranges[2 * t + 1] = b > a ? (int32_t)b : 0;
  }
  for (int64_t k = 0; k < nv; ++k) {
    const int64_t i = val[k];
    int x0, x1, y0, y1, tx, ty;
    rect_of(uv[2 * i], uv[2 * i + 1], radius[i], gx, gy, tw, th, &x0, &x1, &y0, &y1);
    for (ty = y0; ty < y1; ++ty)
      for (tx = x0; tx < x1; ++tx) ids[cur[ty * gx + tx]++] = (int32_t)i;
  }
  free(key);
  free(tk);
  free(val);
  free(tv);
  free(cur);
  return m;
}

/* ---- opacity-aware ("tight") binning -----------------------------------
 * Extension of the build with no msplat counterpart.  A splat reaches
 * alpha >= 1/255 at pixel p only if q = a dx^2 + 2 b dx dy + c dy^2 <=
 * tau = 2 ln(o * 255), d = uv - p (none at all when o < 1/255).  A tile of
 * its rectangle is kept when q's minimum over the tile's pixel-centre box (0
 * with the centre inside, else the least edge minimum, each edge's
 * minimiser -(b/c) X or -(b/a) Y clamped to the edge), lowered by the margin
 * 1e-3 + 1e-4 * S (S bounds the terms' magnitudes on the box), is <= tau.
 * ln is a frexp + atanh-series evaluation.  Same operations, in the same
 * order, as the kernels; the blend is checked to give identical images. */
static float log_pos(float x) {
  int e;
  float m = frexpf(x, &e), s, s2, p;
  if (m < 0.70710678f) {
    m = m * 2.0f;
    e = e - 1;
  }
  s = (m - 1.0f) / (m + 1.0f);
  s2 = s * s;
  p = 1.0f / 9.0f;
  p = fmaf(p, s2, 1.0f / 7.0f);
  p = fmaf(p, s2, 1.0f / 5.0f);
  p = fmaf(p, s2, 1.0f / 3.0f);
  p = fmaf(p, s2, 1.0f);
  return fmaf(2.0f * s, p, (float)e * 0.69314718f);
}

float ref_log_pos(float x) { return log_pos(x); }

typedef struct {
  float u, v, a, b, c, rbc, rba, tau;
  int reach;
} Tight;

static Tight tight_setup(float u, float v, float a, float b, float c, float o) {
  const float alpha_min = 1.0f / 255.0f;
  Tight t;
  const int pd = a > 0.0f && c > 0.0f;
  t.u = u;
  t.v = v;
  t.a = a;
  t.b = b;
  t.c = c;
  t.reach = o >= alpha_min;
  t.tau = t.reach ? 2.0f * log_pos(o / alpha_min) : 0.0f;
  t.rbc = pd ? b / c : 0.0f;
  t.rba = pd ? b / a : 0.0f;
  return t;
}

static int tile_reachable(const Tight *t, float x0, float x1, float y0, float y1) {
  float dx0, dx1, dy0, dy1, m, ya, yb, xa, xb, DX, DY, S;
  if (!(t->a > 0.0f) || !(t->c > 0.0f)) return 1;
  dx0 = t->u - x1;
  dx1 = t->u - x0;
  dy0 = t->v - y1;
  dy1 = t->v - y0;
  if (dx0 <= 0.0f && dx1 >= 0.0f && dy0 <= 0.0f && dy1 >= 0.0f) return 1;
  ya = fminf(fmaxf(-(t->rbc * dx0), dy0), dy1);
  yb = fminf(fmaxf(-(t->rbc * dx1), dy0), dy1);
  xa = fminf(fmaxf(-(t->rba * dy0), dx0), dx1);
  xb = fminf(fmaxf(-(t->rba * dy1), dx0), dx1);
  m = t->a * dx0 * dx0 + 2.0f * t->b * dx0 * ya + t->c * ya * ya;
  m = fminf(m, t->a * dx1 * dx1 + 2.0f * t->b * dx1 * yb + t->c * yb * yb);
  m = fminf(m, t->c * dy0 * dy0 + 2.0f * t->b * dy0 * xa + t->a * xa * xa);
  m = fminf(m, t->c * dy1 * dy1 + 2.0f * t->b * dy1 * xb + t->a * xb * xb);
  DX = fmaxf(fabsf(dx0), fabsf(dx1));
  DY = fmaxf(fabsf(dy0), fabsf(dy1));
  S = fabsf(t->a) * DX * DX + 2.0f * fabsf(t->b) * DX * DY + fabsf(t->c) * DY * DY;
  return m - (1e-3f + 1e-4f * S) <= t->tau;
}

int ref_tile_reachable(float u, float v, float a, float b, float c, float o, float x0, float x1,
                       float y0, float y1) {
  const Tight t = tight_setup(u, v, a, b, c, o);
  return t.reach && tile_reachable(&t, x0, x1, y0, y1);
}

static int tile_keep(const float *uv, const float *conic, const float *opacity, int64_t i,
                     int tx, int ty, int tw, int th, int W, int H) {
  const int px0 = tx * tw, py0 = ty * th;
  const int px1 = (px0 + tw < W ? px0 + tw : W) - 1, py1 = (py0 + th < H ? py0 + th : H) - 1;
  const Tight t = tight_setup(uv[2 * i], uv[2 * i + 1], conic[3 * i], conic[3 * i + 1],
                              conic[3 * i + 2], opacity[i]);
  return t.reach && tile_reachable(&t, (float)px0, (float)px1, (float)py0, (float)py1);
}

/* as ref_sort, keeping only the tiles tile_keep accepts; returns M */
int64_t ref_sort_tight(const float *uv, const float *depth, const int32_t *radius,
                       const float *conic, const float *opacity, int64_t n, int W, int H, int tw,
                       int th, int32_t *ids, int32_t *ranges) {
  const int gx = (W + tw - 1) / tw, gy = (H + th - 1) / th;
  int64_t m = 0, o = 0, j;
  Inst *v;
  for (int64_t i = 0; i < n; ++i) {
    int x0, x1, y0, y1, tx, ty;
    if (radius[i] <= 0) continue;
    rect_of(uv[2 * i], uv[2 * i + 1], radius[i], gx, gy, tw, th, &x0, &x1, &y0, &y1);
    for (ty = y0; ty < y1; ++ty)
      for (tx = x0; tx < x1; ++tx) m += tile_keep(uv, conic, opacity, i, tx, ty, tw, th, W, H);
  }
  v = (Inst *)malloc(sizeof(Inst) * (size_t)(m > 0 ? m : 1));
  for (int64_t i = 0; i < n; ++i) {
    int x0, x1, y0, y1, tx, ty;
    uint32_t bits;
    if (radius[i] <= 0) continue;
    memcpy(&bits, depth + i, 4);
    rect_of(uv[2 * i], uv[2 * i + 1], radius[i], gx, gy, tw, th, &x0, &x1, &y0, &y1);
    for (ty = y0; ty < y1; ++ty)
      for (tx = x0; tx < x1; ++tx) {
        if (!tile_keep(uv, conic, opacity, i, tx, ty, tw, th, W, H)) continue;
        v[o].key = ((uint64_t)(uint32_t)(ty * gx + tx) << 32) | bits;
        v[o].seq = (uint32_t)o;
        v[o].id = (int32_t)i;
        ++o;
      }
  }
  qsort(v, (size_t)m, sizeof(Inst), inst_cmp);
  memset(ranges, 0, sizeof(int32_t) * 2 * (size_t)gx * gy);
  for (j = 0; j < m; ++j) {
    const uint32_t t = (uint32_t)(v[j].key >> 32);
    ids[j] = v[j].id;
    if (j == 0 || (uint32_t)(v[j - 1].key >> 32) != t) ranges[2 * t] = (int32_t)j;
    if (j == m - 1 || (uint32_t)(v[j + 1].key >> 32) != t) ranges[2 * t + 1] = (int32_t)(j + 1);
  }
  free(v);
  return m;
}

/* ---- blending ----------------------------------------------------------- */
void ref_blend(const float *uv, const float *conic, const float *opacity, const float *feat,
               int C, const int32_t *ids, const int32_t *ranges, float bg, int W, int H, int tw,
               int th, float *out, float *Tfin, int32_t *ncon) {
  const int gx = (W + tw - 1) / tw;
  const float power_cut = -5.6f, alpha_max = 0.99f, alpha_min = 1.0f / 255.0f, t_min = 1e-4f;
#pragma omp parallel for schedule(dynamic, 1)
  for (int py = 0; py < H; ++py)
    for (int px = 0; px < W; ++px) {
      float acc[64];
      const int t = (py / th) * gx + (px / tw);
      const int s = ranges[2 * t], e = ranges[2 * t + 1];
      const float fpx = (float)px, fpy = (float)py;
      float T = 1.0f;
      int last = 0, c, j;
      for (c = 0; c < C; ++c) acc[c] = 0.0f;
      for (j = s; j < e; ++j) {
        const int g = ids[j];
        const float ha = -0.5f * conic[3 * g], nb = -conic[3 * g + 1],
                    hc = -0.5f * conic[3 * g + 2], o = opacity[g];
        const float dx = uv[2 * g] - fpx, dy = uv[2 * g + 1] - fpy;
        const float power = fmaf(ha * dx, dx, fmaf(nb * dx, dy, (hc * dy) * dy));
        float alpha, test_T, w;
        if (power > 0.0f) continue;
        if (power < power_cut && o <= 1.0f) continue;
        alpha = fminf(alpha_max, o * exp_neg(power));
        if (alpha < alpha_min) continue;
        test_T = T * (1.0f - alpha);
        if (test_T < t_min) break;
        w = alpha * T;
        for (c = 0; c < C; ++c) acc[c] = fmaf(feat[(int64_t)g * C + c], w, acc[c]);
        T = test_T;
        last = j - s + 1;
      }
      for (c = 0; c < C; ++c) out[(int64_t)c * H * W + (int64_t)py * W + px] = acc[c] + bg * T;
      Tfin[(int64_t)py * W + px] = T;
      ncon[(int64_t)py * W + px] = last;
    }
}

/* Work counts of ref_blend's walk over the same inputs (nothing is written
 * but the counts): pairs[0] = (pixel, instance) pairs examined before each
 * pixel stops (the early-out at T < 1e-4 or the end of its tile's range),
 * pairs[1] = pairs that pass the alpha >= 1/255 test and are blended (the
 * forward's contributors; the backward replays exactly these), pairs[2] =
 * sum of n_contrib (the prefix the backward walks). */
void ref_blend_pairs(const float *uv, const float *conic, const float *opacity,
                     const int32_t *ids, const int32_t *ranges, int W, int H, int tw, int th,
                     int64_t *pairs) {
  const int gx = (W + tw - 1) / tw;
  const float power_cut = -5.6f, alpha_max = 0.99f, alpha_min = 1.0f / 255.0f, t_min = 1e-4f;
  int64_t examined = 0, blended = 0, prefix = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : examined, blended, prefix)
  for (int py = 0; py < H; ++py)
    for (int px = 0; px < W; ++px) {
      const int t = (py / th) * gx + (px / tw);
      const int s = ranges[2 * t], e = ranges[2 * t + 1];
      const float fpx = (float)px, fpy = (float)py;
      float T = 1.0f;
      int last = 0, j;
      for (j = s; j < e; ++j) {
        const int g = ids[j];
        const float ha = -0.5f * conic[3 * g], nb = -conic[3 * g + 1],
                    hc = -0.5f * conic[3 * g + 2], o = opacity[g];
        const float dx = uv[2 * g] - fpx, dy = uv[2 * g + 1] - fpy;
        const float power = fmaf(ha * dx, dx, fmaf(nb * dx, dy, (hc * dy) * dy));
        float alpha, test_T;
        ++examined;
        if (power > 0.0f) continue;
        if (power < power_cut && o <= 1.0f) continue;
        alpha = fminf(alpha_max, o * exp_neg(power));
        if (alpha < alpha_min) continue;
        test_T = T * (1.0f - alpha);
        if (test_T < t_min) break;
        T = test_T;
        ++blended;
        last = j - s + 1;
      }
      prefix += last;
    }
  pairs[0] = examined;
  pairs[1] = blended;
  pairs[2] = prefix;
}

/* OpenMP threads the parallel loops above use. */
int ref_num_threads(void) {
  int n = 1;
#pragma omp parallel
  {
#pragma omp single
    n = omp_get_num_threads();
  }
  return n;
}

/* ---- fused preprocess (direction, SH colour, project, cov3d, EWA, feature
 * pack) in the fused kernel's order; feat = rgb | depth (flag 1) | normal
 * (flag 2) ---------------------------------------------------------------- */
void ref_preprocess(const float *xyz, const float *scale, const float *rot, const float *shs,
                    int64_t n, int K, int degree, const float *Kin, const float *E,
                    const float *center, int W, int H, float nearest, float extent, int flags,
                    float *uv, float *depth, float *conic, int32_t *radius, int32_t *tiles,
                    float *feat) {
  const int C = 3 + ((flags & 1) ? 1 : 0) + ((flags & 2) ? 3 : 0);
  int Ke = K;
  uint8_t one = 1;
  if (degree >= 0 && (degree + 1) * (degree + 1) < K) Ke = (degree + 1) * (degree + 1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    float cov[6], f[7] = {0, 0, 0, 0, 0, 0, 0};
    const float *x = xyz + 3 * i;
    int c;
    ref_project(x, 1, Kin, E, W, H, nearest, extent, uv + 2 * i, depth + i);
    conic[3 * i] = conic[3 * i + 1] = conic[3 * i + 2] = 0.0f;
    radius[i] = 0;
    tiles[i] = 0;
    if (depth[i] != 0.0f) {
      ref_cov3d(scale + 3 * i, rot + 4 * i, &one, 1, cov);
      ref_ewa(x, cov, 1, Kin, E, uv + 2 * i, W, H, &one, conic + 3 * i, radius + i, tiles + i);
    }
    if (radius[i] > 0) {
      const float dx = x[0] - center[0], dy = x[1] - center[1], dz = x[2] - center[2];
      const float len = sqrtf(dx * dx + dy * dy + dz * dz);
      float dir[3], out3[3];
      int o = 3;
      dir[0] = dx / len;
      dir[1] = dy / len;
      dir[2] = dz / len;
      {
        /* per-channel SH sum over the first Ke coefficients of [K,3] */
        float tmp[16 * 3];
        int k;
        for (c = 0; c < 3; ++c)
          for (k = 0; k < K; ++k) tmp[c * K + k] = shs[(i * K + k) * 3 + c];
        ref_sh(tmp, dir, 1, 3, K, Ke == K ? -1 : degree, out3);
      }
      for (c = 0; c < 3; ++c) f[c] = (flags & 8) ? out3[c] : fmaxf(out3[c] + 0.5f, 0.0f);
      if (flags & 1) f[o++] = depth[i];
      if (flags & 2) {
        float R[9], m[3], nh[3], len2, inv, dot, sgn, smin;
        const float *sc = scale + 3 * i;
        int k = 0, j;
        rotmat(rot + 4 * i, R);
        smin = sc[0];
        if (sc[1] < smin) { k = 1; smin = sc[1]; }
        if (sc[2] < smin) { k = 2; }
        m[0] = R[k];
        m[1] = R[3 + k];
        m[2] = R[6 + k];
        len2 = sqrtf(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        inv = 1.0f / fmaxf(len2, 1e-12f);
        dot = (center[0] - x[0]) * m[0] + (center[1] - x[1]) * m[1] + (center[2] - x[2]) * m[2];
        sgn = dot < 0.0f ? -1.0f : 1.0f;
        for (j = 0; j < 3; ++j) nh[j] = sgn * (m[j] * inv);
        for (j = 0; j < 3; ++j) f[o + j] = E[4 * j] * nh[0] + E[4 * j + 1] * nh[1] + E[4 * j + 2] * nh[2];
      }
    }
    for (c = 0; c < C; ++c) feat[i * C + c] = f[c];
  }
}

/* ======================================================================
 * Backward (FP32 restatement of the chain rule; OpenMP)
 * ====================================================================== */

/* ---- alpha_blending backward ------------------------------------------
 * Per pixel, the contributors of the forward are replayed back to front from
 * n_contrib, T_i recovered from T_final (T_i = T_{i+1} / (1 - alpha_i)):
 *   dL/df_ic   = alpha_i T_i g_c
 *   dL/dalpha_i = T_i sum_c (f_ic - A_ic) g_c - T_final / (1 - alpha_i) (bg sum_c g_c + gT)
 * with A_c the colour composited behind splat i (A <- alpha_i f_i + (1 -
 * alpha_i) A).  alpha = min(0.99, o G), G = exp(power): no gradient where
 * clamped; else dL/do = G dL/dalpha and dL/dpower = alpha dL/dalpha, and
 * power = -(a dx^2 + c dy^2)/2 - b dx dy with d = uv - pixel.
 * Each tile is one task; its pixels add into rec[j] (one row of 6 + C floats
 * per instance j of the tile's range: du, dv, da, db, dc, dopacity, dfeat),
 * which no other tile touches.  The rows are then summed per splat in
 * instance order: d_uv[n,2], d_conic[n,3], d_op[n], d_feat[n,C] (overwritten). */
void ref_blend_bwd(const float *uv, const float *conic, const float *opacity, const float *feat,
                   int C, const int32_t *ids, const int32_t *ranges, float bg, int W, int H,
                   int tw, int th, const float *Tfin, const int32_t *ncon, const float *dout,
                   const float *dT, int64_t n, int64_t m, float *rec, float *d_uv,
                   float *d_conic, float *d_op, float *d_feat) {
  const int gx = (W + tw - 1) / tw, gy = (H + th - 1) / th;
  const int Q = 6 + C;
  const float power_cut = -5.6f, alpha_max = 0.99f, alpha_min = 1.0f / 255.0f;
  const int64_t P = (int64_t)W * H;
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < gx * gy; ++t) {
    const int s = ranges[2 * t], e = ranges[2 * t + 1];
    const int tx = t % gx, ty = t / gx;
    if (e > s) memset(rec + (int64_t)s * Q, 0, sizeof(float) * (size_t)(e - s) * Q);
    for (int py = ty * th; py < ty * th + th && py < H; ++py)
      for (int px = tx * tw; px < tx * tw + tw && px < W; ++px) {
        const int64_t pix = (int64_t)py * W + px;
        const int last = ncon[pix];
        float g[64], A[64], gsum = 0.0f, T = Tfin[pix], T_final, bg_term;
        int c, j;
        for (c = 0; c < C; ++c) {
          g[c] = dout[c * P + pix];
          A[c] = 0.0f;
          gsum += g[c];
        }
        T_final = T;
        bg_term = bg * gsum + (dT != NULL ? dT[pix] : 0.0f);
        for (j = s + last - 1; j >= s; --j) {
          const int gi = ids[j];
          const float ca = conic[3 * gi], cb = conic[3 * gi + 1], cc = conic[3 * gi + 2];
          const float ha = -0.5f * ca, nb = -cb, hc = -0.5f * cc, o = opacity[gi];
          const float dx = uv[2 * gi] - (float)px, dy = uv[2 * gi + 1] - (float)py;
          const float power = fmaf(ha * dx, dx, fmaf(nb * dx, dy, (hc * dy) * dy));
          float G, alpha, dLda, fg = 0.0f, *r;
          if (power > 0.0f) continue;
          if (power < power_cut && o <= 1.0f) continue;
          G = exp_neg(power);
          alpha = fminf(alpha_max, o * G);
          if (alpha < alpha_min) continue;
          T = T / (1.0f - alpha);
          r = rec + (int64_t)j * Q;
          for (c = 0; c < C; ++c) {
            const float f = feat[(int64_t)gi * C + c];
            r[6 + c] += alpha * T * g[c];
            fg += (f - A[c]) * g[c];
            A[c] = alpha * f + (1.0f - alpha) * A[c];
          }
          dLda = T * fg - T_final / (1.0f - alpha) * bg_term;
          if (o * G < alpha_max) {
            const float dpow = alpha * dLda;
            r[0] += dpow * (-ca * dx - cb * dy);
            r[1] += dpow * (-cb * dx - cc * dy);
            r[2] += dpow * (-0.5f * dx * dx);
            r[3] += dpow * (-dx * dy);
            r[4] += dpow * (-0.5f * dy * dy);
            r[5] += G * dLda;
          }
        }
      }
  }
  memset(d_uv, 0, sizeof(float) * 2 * (size_t)n);
  memset(d_conic, 0, sizeof(float) * 3 * (size_t)n);
  memset(d_op, 0, sizeof(float) * (size_t)n);
  memset(d_feat, 0, sizeof(float) * (size_t)C * (size_t)n);
  for (int64_t j = 0; j < m; ++j) {
    const int64_t gi = ids[j];
    const float *r = rec + j * Q;
    d_uv[2 * gi] += r[0];
    d_uv[2 * gi + 1] += r[1];
    d_conic[3 * gi] += r[2];
    d_conic[3 * gi + 1] += r[3];
    d_conic[3 * gi + 2] += r[4];
    d_op[gi] += r[5];
    for (int c = 0; c < C; ++c) d_feat[gi * C + c] += r[6 + c];
  }
}

/* dR/dq of the rotation of rotmat(), q = (w, x, y, z): gq += sum_i gR_i dR_i/dq */
static void rotmat_vjp(const float *q, const float *gR, float *gq) {
  const float w = q[0], x = q[1], y = q[2], z = q[3];
  const float dw[9] = {2 * w, -2 * z, 2 * y, 2 * z, 2 * w, -2 * x, -2 * y, 2 * x, 2 * w};
  const float dxv[9] = {2 * x, 2 * y, 2 * z, 2 * y, -2 * x, -2 * w, 2 * z, 2 * w, -2 * x};
  const float dy[9] = {-2 * y, 2 * x, 2 * w, 2 * x, 2 * y, 2 * z, -2 * w, 2 * z, -2 * y};
  const float dz[9] = {-2 * z, -2 * w, 2 * x, 2 * w, -2 * z, 2 * y, 2 * x, 2 * y, 2 * z};
  for (int i = 0; i < 9; ++i) {
    gq[0] += gR[i] * dw[i];
    gq[1] += gR[i] * dxv[i];
    gq[2] += gR[i] * dy[i];
    gq[3] += gR[i] * dz[i];
  }
}

/* gradient of the real-SH basis (the polynomials of ref_sh) at dir,
 * contracted with gY: gdir += sum_k gY_k dY_k/ddir, k < Ke */
static void sh_basis_vjp(const float *d, int Ke, const float *gY, float *gd) {
  const double pi = 3.14159265358979323846;
  const float C1 = (float)sqrt(3.0 / (4.0 * pi));
  const float C20 = (float)(0.5 * sqrt(15.0 / pi));
  const float C22 = (float)(0.25 * sqrt(5.0 / pi));
  const float C24 = (float)(0.25 * sqrt(15.0 / pi));
  const float C30 = (float)(0.25 * sqrt(35.0 / (2.0 * pi)));
  const float C31 = (float)(0.5 * sqrt(105.0 / pi));
  const float C32 = (float)(0.25 * sqrt(21.0 / (2.0 * pi)));
  const float C33 = (float)(0.25 * sqrt(7.0 / pi));
  const float C35 = (float)(0.25 * sqrt(105.0 / pi));
  const float x = d[0], y = d[1], z = d[2];
  float J[16][3];
  int k;
  memset(J, 0, sizeof(J));
  J[1][1] = -C1;
  J[2][2] = C1;
  J[3][0] = -C1;
  J[4][0] = C20 * y; J[4][1] = C20 * x;
  J[5][1] = -C20 * z; J[5][2] = -C20 * y;
  J[6][0] = -2 * C22 * x; J[6][1] = -2 * C22 * y; J[6][2] = 4 * C22 * z;
  J[7][0] = -C20 * z; J[7][2] = -C20 * x;
  J[8][0] = 2 * C24 * x; J[8][1] = -2 * C24 * y;
  J[9][0] = -6 * C30 * x * y; J[9][1] = -C30 * (3 * x * x - 3 * y * y);
  J[10][0] = C31 * y * z; J[10][1] = C31 * x * z; J[10][2] = C31 * x * y;
  J[11][0] = 2 * C32 * x * y; J[11][1] = -C32 * (4 * z * z - x * x - 3 * y * y);
  J[11][2] = -8 * C32 * y * z;
  J[12][0] = -6 * C33 * x * z; J[12][1] = -6 * C33 * y * z;
  J[12][2] = C33 * (6 * z * z - 3 * x * x - 3 * y * y);
  J[13][0] = -C32 * (4 * z * z - 3 * x * x - y * y); J[13][1] = 2 * C32 * x * y;
  J[13][2] = -8 * C32 * x * z;
  J[14][0] = 2 * C35 * x * z; J[14][1] = -2 * C35 * y * z; J[14][2] = C35 * (x * x - y * y);
  J[15][0] = -C30 * (3 * x * x - 3 * y * y); J[15][1] = 6 * C30 * x * y;
  for (k = 1; k < Ke; ++k) {
    gd[0] += gY[k] * J[k][0];
    gd[1] += gY[k] * J[k][1];
    gd[2] += gY[k] * J[k][2];
  }
}

/* ---- preprocess backward ------------------------------------------------
 * Reverse of ref_preprocess for the splats with radius > 0 (every gradient
 * of the others is zero): from d_uv[n,2], d_conic[n,3] and d_feat[n,C]
 * (rgb | depth (flag 1) | normal (flag 2); flag 8: rgb is the SH sum without
 * +0.5 and clamp) to d_xyz[n,3], d_scale[n,3], d_rot[n,4] (w.r.t. the unit
 * quaternion the forward used), d_shs[n,K,3] (overwritten), and the camera:
 * d_intr[4] (fx, fy, cx, cy), d_extr[12] (3x4 row-major), d_center[3]
 * (overwritten; summed per thread in double).
 *   uv = (fx p_x / p_z + cx, fy p_y / p_z + cy), p = E [x; 1];
 *   conic = inverse of S2 = T S3 T^T + 0.3 I, T = J W, J the clamped EWA
 *   Jacobian (its limits 1.3 * 0.5 W / fx depend on fx), S3 = R diag(s)^2 R^T;
 *   colour = SH(dir) [+0.5, clamp], dir = (x - centre) / |x - centre|;
 *   depth = p_z; normal = W (+-m / |m|), m the column of R of the smallest
 *   scale (the sign is fixed: no gradient). */
void ref_preprocess_bwd(const float *xyz, const float *scale, const float *rot,
                        const float *shs, int64_t n, int K, int degree, const float *Kin,
                        const float *E, const float *center, int W, int H, float nearest,
                        float extent, int flags, const int32_t *radius, const float *d_uv,
                        const float *d_conic, const float *d_feat, float *d_xyz,
                        float *d_scale, float *d_rot, float *d_shs, float *d_intr,
                        float *d_extr, float *d_center) {
  const int C = 3 + ((flags & 1) ? 1 : 0) + ((flags & 2) ? 3 : 0);
  const float lowpass = 0.3f, jclamp = 1.3f;
  int Ke = K;
  double cam_acc[19];
  (void)nearest;
  (void)extent;
  if (degree >= 0 && (degree + 1) * (degree + 1) < K) Ke = (degree + 1) * (degree + 1);
  memset(cam_acc, 0, sizeof(cam_acc));
#pragma omp parallel
  {
    double loc[19];
    memset(loc, 0, sizeof(loc));
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      const float *x = xyz + 3 * i, *sc = scale + 3 * i, *q = rot + 4 * i;
      float p[3], gp[3] = {0, 0, 0}, gK[4] = {0, 0, 0, 0}, gE[12], gx[3] = {0, 0, 0};
      float R[9], M[9], S[6], Sf[9], Jm[4], T[6], U[6], gT[6], gR[9], gq[4] = {0, 0, 0, 0};
      float z, limx, limy, txz, tyz, tx, ty, a, b, c, det, id2, ga, gb, gcc, gA, gB, gC;
      float G[9], gM[9], gs[3], gsh[48];
      int k, r, j;
      memset(d_shs + i * K * 3, 0, sizeof(float) * 3 * (size_t)K);
      for (k = 0; k < 3; ++k) d_xyz[3 * i + k] = d_scale[3 * i + k] = 0.0f;
      for (k = 0; k < 4; ++k) d_rot[4 * i + k] = 0.0f;
      if (radius[i] <= 0) continue;
      memset(gE, 0, sizeof(gE));
      memset(gR, 0, sizeof(gR));
      cam_point(E, x, p);
      z = p[2];
      /* uv */
      {
        const float gu = d_uv[2 * i], gv = d_uv[2 * i + 1];
        gK[0] += gu * (p[0] / z);
        gK[2] += gu;
        gK[1] += gv * (p[1] / z);
        gK[3] += gv;
        gp[0] += gu * (Kin[0] / z);
        gp[1] += gv * (Kin[1] / z);
        gp[2] += -gu * Kin[0] * p[0] / (z * z) - gv * Kin[1] * p[1] / (z * z);
      }
      /* depth channel */
      if (flags & 1) gp[2] += d_feat[i * C + 3];
      /* covariance forward */
      rotmat(q, R);
      for (r = 0; r < 3; ++r)
        for (k = 0; k < 3; ++k) M[3 * r + k] = R[3 * r + k] * sc[k];
      {
        static const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
        for (k = 0; k < 6; ++k)
          S[k] = M[3 * ii[k]] * M[3 * jj[k]] + M[3 * ii[k] + 1] * M[3 * jj[k] + 1] +
                 M[3 * ii[k] + 2] * M[3 * jj[k] + 2];
      }
      Sf[0] = S[0]; Sf[1] = S[1]; Sf[2] = S[2];
      Sf[3] = S[1]; Sf[4] = S[3]; Sf[5] = S[4];
      Sf[6] = S[2]; Sf[7] = S[4]; Sf[8] = S[5];
      /* EWA forward */
      limx = jclamp * (0.5f * (float)W / Kin[0]);
      limy = jclamp * (0.5f * (float)H / Kin[1]);
      txz = fminf(fmaxf(p[0] / z, -limx), limx);
      tyz = fminf(fmaxf(p[1] / z, -limy), limy);
      tx = txz * z;
      ty = tyz * z;
      Jm[0] = Kin[0] / z;
      Jm[1] = -(Kin[0] * tx) / (z * z);
      Jm[2] = Kin[1] / z;
      Jm[3] = -(Kin[1] * ty) / (z * z);
      for (j = 0; j < 3; ++j) {
        T[j] = Jm[0] * E[j] + Jm[1] * E[8 + j];
        T[3 + j] = Jm[2] * E[4 + j] + Jm[3] * E[8 + j];
      }
      for (r = 0; r < 2; ++r)
        for (j = 0; j < 3; ++j)
          U[3 * r + j] = T[3 * r] * Sf[j] + T[3 * r + 1] * Sf[3 + j] + T[3 * r + 2] * Sf[6 + j];
      a = U[0] * T[0] + U[1] * T[1] + U[2] * T[2] + lowpass;
      b = U[0] * T[3] + U[1] * T[4] + U[2] * T[5];
      c = U[3] * T[3] + U[4] * T[4] + U[5] * T[5] + lowpass;
      det = a * c - b * b;
      id2 = 1.0f / (det * det);
      /* conic = (c, -b, a) / det */
      gA = d_conic[3 * i];
      gB = d_conic[3 * i + 1];
      gC = d_conic[3 * i + 2];
      ga = (-gA * c * c + gB * b * c - gC * b * b) * id2;
      gb = (2.0f * gA * b * c - gB * (det + 2.0f * b * b) + 2.0f * gC * a * b) * id2;
      gcc = (-gA * b * b + gB * a * b - gC * a * a) * id2;
      /* S2 = T S3 T^T: dS3 = ga T0'T0 + gb T0'T1 + gc T1'T1 (symmetrised
       * into the packed entries); dT0 = 2 ga U0 + gb U1, dT1 = gb U0 + 2 gc U1 */
      {
        float D[9];
        for (r = 0; r < 3; ++r)
          for (j = 0; j < 3; ++j)
            D[3 * r + j] = ga * T[r] * T[j] + gb * T[r] * T[3 + j] + gcc * T[3 + r] * T[3 + j];
        /* G: full symmetric gradient of S3 with the packed entries' weight
         * split over both off-diagonal positions */
        G[0] = D[0];
        G[4] = D[4];
        G[8] = D[8];
        G[1] = G[3] = 0.5f * (D[1] + D[3]);
        G[2] = G[6] = 0.5f * (D[2] + D[6]);
        G[5] = G[7] = 0.5f * (D[5] + D[7]);
      }
      for (j = 0; j < 3; ++j) {
        gT[j] = 2.0f * ga * U[j] + gb * U[3 + j];
        gT[3 + j] = gb * U[j] + 2.0f * gcc * U[3 + j];
      }
      /* T = J W */
      {
        float gJ[4] = {0, 0, 0, 0}, gtx, gty, gtxz, gtyz;
        for (j = 0; j < 3; ++j) {
          gJ[0] += gT[j] * E[j];
          gJ[1] += gT[j] * E[8 + j];
          gJ[2] += gT[3 + j] * E[4 + j];
          gJ[3] += gT[3 + j] * E[8 + j];
          gE[j] += gT[j] * Jm[0];
          gE[8 + j] += gT[j] * Jm[1] + gT[3 + j] * Jm[3];
          gE[4 + j] += gT[3 + j] * Jm[2];
        }
        /* J00 = fx / z, J02 = -(fx tx) / z^2, tx = clamp(p_x / z) z */
        gK[0] += gJ[0] / z - gJ[1] * tx / (z * z);
        gp[2] += -gJ[0] * Kin[0] / (z * z) + 2.0f * gJ[1] * Kin[0] * tx / (z * z * z);
        gtx = -gJ[1] * Kin[0] / (z * z);
        gtxz = gtx * z;
        gp[2] += gtx * txz;
        if (p[0] / z >= -limx && p[0] / z <= limx) {
          gp[0] += gtxz / z;
          gp[2] += -gtxz * p[0] / (z * z);
        } else {
          const float glim = p[0] / z > limx ? gtxz : -gtxz;
          gK[0] += glim * (-jclamp * 0.5f * (float)W / (Kin[0] * Kin[0]));
        }
        gK[1] += gJ[2] / z - gJ[3] * ty / (z * z);
        gp[2] += -gJ[2] * Kin[1] / (z * z) + 2.0f * gJ[3] * Kin[1] * ty / (z * z * z);
        gty = -gJ[3] * Kin[1] / (z * z);
        gtyz = gty * z;
        gp[2] += gty * tyz;
        if (p[1] / z >= -limy && p[1] / z <= limy) {
          gp[1] += gtyz / z;
          gp[2] += -gtyz * p[1] / (z * z);
        } else {
          const float glim = p[1] / z > limy ? gtyz : -gtyz;
          gK[1] += glim * (-jclamp * 0.5f * (float)H / (Kin[1] * Kin[1]));
        }
      }
      /* S3 = M M^T, M = R diag(s): dM = 2 G M */
      for (r = 0; r < 3; ++r)
        for (k = 0; k < 3; ++k)
          gM[3 * r + k] = 2.0f * (G[3 * r] * M[k] + G[3 * r + 1] * M[3 + k] + G[3 * r + 2] * M[6 + k]);
      for (k = 0; k < 3; ++k) {
        gs[k] = 0.0f;
        for (r = 0; r < 3; ++r) {
          gR[3 * r + k] += gM[3 * r + k] * sc[k];
          gs[k] += gM[3 * r + k] * R[3 * r + k];
        }
      }
      /* colour */
      {
        const float dx = x[0] - center[0], dy = x[1] - center[1], dz = x[2] - center[2];
        const float len = sqrtf(dx * dx + dy * dy + dz * dz);
        float dir[3], Y[16], gY[16], out3[3], gdir[3] = {0, 0, 0}, tmp[48], dp;
        dir[0] = dx / len;
        dir[1] = dy / len;
        dir[2] = dz / len;
        for (int cc2 = 0; cc2 < 3; ++cc2)
          for (k = 0; k < K; ++k) tmp[cc2 * K + k] = shs[(i * K + k) * 3 + cc2];
        ref_sh(tmp, dir, 1, 3, K, Ke == K ? -1 : degree, out3);
        {
          /* the basis values: ref_sh of unit coefficient rows */
          float e[16 * 16], yy[16];
          memset(e, 0, sizeof(float) * 16 * (size_t)K);
          for (k = 0; k < K; ++k) e[k * K + k] = 1.0f;
          ref_sh(e, dir, 1, K, K, Ke == K ? -1 : degree, yy);
          for (k = 0; k < 16; ++k) Y[k] = k < K ? yy[k] : 0.0f;
        }
        memset(gY, 0, sizeof(gY));
        memset(gsh, 0, sizeof(gsh));
        for (int cc2 = 0; cc2 < 3; ++cc2) {
          const float gcol = ((flags & 8) || out3[cc2] + 0.5f >= 0.0f) ? d_feat[i * C + cc2] : 0.0f;
          for (k = 0; k < Ke; ++k) {
            gsh[3 * k + cc2] = gcol * Y[k];
            gY[k] += gcol * shs[(i * K + k) * 3 + cc2];
          }
        }
        sh_basis_vjp(dir, Ke, gY, gdir);
        dp = dir[0] * gdir[0] + dir[1] * gdir[1] + dir[2] * gdir[2];
        for (k = 0; k < 3; ++k) {
          const float gd = (gdir[k] - dir[k] * dp) / len;
          gx[k] += gd;
          loc[16 + k] -= gd;
        }
        for (k = 0; k < 3 * K; ++k) d_shs[i * K * 3 + k] = gsh[k];
      }
      /* normal channel */
      if (flags & 2) {
        const int o = (flags & 1) ? 4 : 3;
        float m[3], len2, inv, dot, sgn, nh[3], gnh[3] = {0, 0, 0}, gm[3], mg;
        int kk = 0;
        float smin = sc[0];
        if (sc[1] < smin) { kk = 1; smin = sc[1]; }
        if (sc[2] < smin) kk = 2;
        m[0] = R[kk];
        m[1] = R[3 + kk];
        m[2] = R[6 + kk];
        len2 = sqrtf(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        inv = 1.0f / fmaxf(len2, 1e-12f);
        dot = (center[0] - x[0]) * m[0] + (center[1] - x[1]) * m[1] + (center[2] - x[2]) * m[2];
        sgn = dot < 0.0f ? -1.0f : 1.0f;
        for (j = 0; j < 3; ++j) nh[j] = sgn * (m[j] * inv);
        for (j = 0; j < 3; ++j) {
          const float gn = d_feat[i * C + o + j];
          for (k = 0; k < 3; ++k) {
            gnh[k] += gn * E[4 * j + k];
            gE[4 * j + k] += gn * nh[k];
          }
        }
        mg = m[0] * gnh[0] + m[1] * gnh[1] + m[2] * gnh[2];
        for (k = 0; k < 3; ++k)
          gm[k] = len2 > 1e-12f ? sgn * (gnh[k] * inv - m[k] * mg * inv * inv * inv)
                                : sgn * gnh[k] * inv;
        gR[kk] += gm[0];
        gR[3 + kk] += gm[1];
        gR[6 + kk] += gm[2];
      }
      /* p = E [x; 1] */
      for (r = 0; r < 3; ++r) {
        for (k = 0; k < 3; ++k) {
          gx[k] += E[4 * r + k] * gp[r];
          gE[4 * r + k] += gp[r] * x[k];
        }
        gE[4 * r + 3] += gp[r];
      }
      rotmat_vjp(q, gR, gq);
      for (k = 0; k < 3; ++k) {
        d_xyz[3 * i + k] = gx[k];
        d_scale[3 * i + k] = gs[k];
      }
      for (k = 0; k < 4; ++k) d_rot[4 * i + k] = gq[k];
      for (k = 0; k < 4; ++k) loc[k] += gK[k];
      for (k = 0; k < 12; ++k) loc[4 + k] += gE[k];
    }
#pragma omp critical
    for (int k2 = 0; k2 < 19; ++k2) cam_acc[k2] += loc[k2];
  }
  for (int k = 0; k < 4; ++k) d_intr[k] = (float)cam_acc[k];
  for (int k = 0; k < 12; ++k) d_extr[k] = (float)cam_acc[4 + k];
  for (int k = 0; k < 3; ++k) d_center[k] = (float)cam_acc[16 + k];
}
