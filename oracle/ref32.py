"""ctypes wrapper of libref32.so, the FP32 C restatement (test infrastructure).

numpy in, numpy out; all arrays float32 / int32 / uint8, C-contiguous.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libref32.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from . import build_ref32
            build_ref32.build(verbose=False)
        L = ctypes.CDLL(LIB)
        P, I64, I, F = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_float
        L.ref_exp_neg.restype = F
        L.ref_exp_neg.argtypes = [F]
        L.ref_project.argtypes = [P, I64, P, P, I, I, F, F, P, P]
        L.ref_cov3d.argtypes = [P, P, P, I64, P]
        L.ref_ewa.argtypes = [P, P, I64, P, P, P, I, I, P, P, P, P]
        L.ref_sh.argtypes = [P, P, I64, I, I, I, P]
        L.ref_sort_count.restype = I64
        L.ref_sort_count.argtypes = [P, P, I64, I, I, I, I]
        L.ref_sort.restype = I64
        L.ref_sort.argtypes = [P, P, P, I64, I, I, I, I, P, P]
        L.ref_log_pos.restype = F
        L.ref_log_pos.argtypes = [F]
        L.ref_tile_reachable.restype = I
        L.ref_tile_reachable.argtypes = [F] * 10
        L.ref_sort_tight.restype = I64
        L.ref_sort_tight.argtypes = [P, P, P, P, P, I64, I, I, I, I, P, P]
        L.ref_blend.argtypes = [P, P, P, P, I, P, P, F, I, I, I, I, P, P, P]
        L.ref_blend_pairs.argtypes = [P, P, P, P, P, I, I, I, I, P]
        L.ref_num_threads.restype = I
        L.ref_preprocess.argtypes = [P, P, P, P, I64, I, I, P, P, P, I, I, F, F, I, P, P, P, P, P,
                                     P]
        L.ref_blend_bwd.argtypes = [P, P, P, P, I, P, P, F, I, I, I, I, P, P, P, P, I64, I64, P,
                                    P, P, P, P]
        L.ref_preprocess_bwd.argtypes = [P, P, P, P, I64, I, I, P, P, P, I, I, F, F, I] + [P] * 11
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(np.asarray(a), dtype=dt)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def exp_neg(x: float) -> float:
    return float(lib().ref_exp_neg(float(x)))


def project_point(xyz, intr, extr, W, H, nearest=0.2, extent=1.3):
    x = _c(xyz, np.float32)
    K = _c(np.asarray(intr).reshape(-1)[:4], np.float32)
    E = _c(np.asarray(extr)[:3, :4], np.float32)
    n = x.shape[0]
    uv = np.zeros((n, 2), np.float32)
    d = np.zeros((n, 1), np.float32)
    lib().ref_project(_p(x), n, _p(K), _p(E), W, H, nearest, extent, _p(uv), _p(d))
    return uv, d


def compute_cov3d(scale, rot, visible):
    s, q = _c(scale, np.float32), _c(rot, np.float32)
    v = _c(np.asarray(visible).reshape(-1), np.uint8)
    n = s.shape[0]
    cov = np.zeros((n, 6), np.float32)
    lib().ref_cov3d(_p(s), _p(q), _p(v), n, _p(cov))
    return cov


def ewa_project(xyz, cov3d, intr, extr, uv, W, H, visible):
    x, c, u = _c(xyz, np.float32), _c(cov3d, np.float32), _c(uv, np.float32)
    K = _c(np.asarray(intr).reshape(-1)[:4], np.float32)
    E = _c(np.asarray(extr)[:3, :4], np.float32)
    v = _c(np.asarray(visible).reshape(-1), np.uint8)
    n = x.shape[0]
    conic = np.zeros((n, 3), np.float32)
    radius = np.zeros(n, np.int32)
    tiles = np.zeros(n, np.int32)
    lib().ref_ewa(_p(x), _p(c), n, _p(K), _p(E), _p(u), W, H, _p(v), _p(conic), _p(radius),
                  _p(tiles))
    return conic, radius, tiles


def compute_sh(shs, dirs, degree=-1):
    """shs [N,C,K] (made dense), dirs [N,3] -> [N,C]."""
    s, d = _c(shs, np.float32), _c(dirs, np.float32)
    n, C, K = s.shape
    out = np.zeros((n, C), np.float32)
    lib().ref_sh(_p(s), _p(d), n, C, K, -1 if degree is None else degree, _p(out))
    return out


def sort_gaussian(uv, depth, W, H, radius, tile_w=16, tile_h=16):
    u = _c(uv, np.float32)
    d = _c(np.asarray(depth).reshape(-1), np.float32)
    r = _c(np.asarray(radius).reshape(-1), np.int32)
    n = u.shape[0]
    m = lib().ref_sort_count(_p(u), _p(r), n, W, H, tile_w, tile_h)
    gx, gy = (W + tile_w - 1) // tile_w, (H + tile_h - 1) // tile_h
    ids = np.zeros(max(m, 1), np.int32)
    rng = np.zeros((gx * gy, 2), np.int32)
    lib().ref_sort(_p(u), _p(d), _p(r), n, W, H, tile_w, tile_h, _p(ids), _p(rng))
    return ids[:m], rng


def log_pos(x: float) -> float:
    return float(lib().ref_log_pos(float(x)))


def tile_reachable(u, v, a, b, c, o, x0, x1, y0, y1) -> bool:
    """The tight-binning test of one splat against one pixel-centre box."""
    return bool(lib().ref_tile_reachable(*[float(t) for t in (u, v, a, b, c, o, x0, x1, y0,
                                                                y1)]))


def sort_gaussian_tight(uv, depth, W, H, radius, conic, opacity, tile_w=16, tile_h=16):
    """Opacity-aware binning -> (ids [M] int32, tile_range [T,2] int32)."""
    u, d, r = _c(uv, np.float32), _c(np.asarray(depth).reshape(-1), np.float32), \
        _c(np.asarray(radius).reshape(-1), np.int32)
    cn, op = _c(conic, np.float32), _c(np.asarray(opacity).reshape(-1), np.float32)
    n = u.shape[0]
    gx, gy = (W + tile_w - 1) // tile_w, (H + tile_h - 1) // tile_h
    # the count pass is repeated inside; size ids by the exact (rect) count
    cap = int(lib().ref_sort_count(_p(u), _p(r), n, W, H, tile_w, tile_h))
    ids = np.zeros(max(cap, 1), np.int32)
    rng = np.zeros((gx * gy, 2), np.int32)
    m = lib().ref_sort_tight(_p(u), _p(d), _p(r), _p(cn), _p(op), n, W, H, tile_w, tile_h,
                             _p(ids), _p(rng))
    return ids[:m].copy(), rng


def alpha_blending(uv, conic, opacity, feature, ids, tile_range, bg, W, H, tile_w=16, tile_h=16):
    """-> (out [C,H,W], T_final [H,W], n_contrib [H,W])."""
    u, c = _c(uv, np.float32), _c(conic, np.float32)
    o = _c(np.asarray(opacity).reshape(-1), np.float32)
    f = _c(feature, np.float32)
    i = _c(ids, np.int32)
    r = _c(tile_range, np.int32)
    C = f.shape[1]
    out = np.zeros((C, H, W), np.float32)
    T = np.zeros((H, W), np.float32)
    nc = np.zeros((H, W), np.int32)
    lib().ref_blend(_p(u), _p(c), _p(o), _p(f), C, _p(i), _p(r), bg, W, H, tile_w, tile_h, _p(out),
                    _p(T), _p(nc))
    return out, T, nc


def blend_pairs(uv, conic, opacity, ids, tile_range, W, H, tile_w=16, tile_h=16):
    """Work counts of the blend walk -> (examined, blended, sum of n_contrib):
    (pixel, instance) pairs examined before each pixel stops, pairs blended
    (alpha >= 1/255, the forward's contributors), and the prefix lengths the
    backward walks."""
    u, c = _c(uv, np.float32), _c(conic, np.float32)
    o = _c(np.asarray(opacity).reshape(-1), np.float32)
    i = _c(ids, np.int32)
    r = _c(tile_range, np.int32)
    out = np.zeros(3, np.int64)
    lib().ref_blend_pairs(_p(u), _p(c), _p(o), _p(i), _p(r), W, H, tile_w, tile_h, _p(out))
    return int(out[0]), int(out[1]), int(out[2])


def num_threads() -> int:
    """OpenMP threads of the restatement's parallel loops."""
    return int(lib().ref_num_threads())


def preprocess(xyz, scale, rot, shs, intr, extr, center, W, H, degree=-1, flags=0, nearest=0.2,
               extent=1.3):
    """Fused preprocess restatement: shs [N,K,3] -> (uv, depth [N,1], conic, radius, tiles,
    feat [N,C]) with C = 3 (+1 if flags & 1) (+3 if flags & 2)."""
    x, s, q = _c(xyz, np.float32), _c(scale, np.float32), _c(rot, np.float32)
    sh = _c(shs, np.float32)
    K = _c(np.asarray(intr).reshape(-1)[:4], np.float32)
    E = _c(np.asarray(extr)[:3, :4], np.float32)
    cen = _c(np.asarray(center).reshape(-1)[:3], np.float32)
    n, Kc = x.shape[0], sh.shape[1]
    C = 3 + (1 if flags & 1 else 0) + (3 if flags & 2 else 0)
    uv = np.zeros((n, 2), np.float32)
    d = np.zeros((n, 1), np.float32)
    conic = np.zeros((n, 3), np.float32)
    radius = np.zeros(n, np.int32)
    tiles = np.zeros(n, np.int32)
    feat = np.zeros((n, C), np.float32)
    lib().ref_preprocess(_p(x), _p(s), _p(q), _p(sh), n, Kc, -1 if degree is None else degree,
                         _p(K), _p(E), _p(cen), W, H, nearest, extent, flags, _p(uv), _p(d),
                         _p(conic), _p(radius), _p(tiles), _p(feat))
    return uv, d, conic, radius, tiles, feat


def alpha_blending_bwd(uv, conic, opacity, feature, ids, tile_range, bg, W, H, T_final,
                       n_contrib, dout, dT=None, tile_w=16, tile_h=16):
    """Backward of alpha_blending from dout [C,H,W] (and dT [H,W]) ->
    (d_uv [n,2], d_conic [n,3], d_opacity [n], d_feature [n,C])."""
    u, c = _c(uv, np.float32), _c(conic, np.float32)
    o = _c(np.asarray(opacity).reshape(-1), np.float32)
    f = _c(feature, np.float32)
    i = _c(ids, np.int32)
    r = _c(tile_range, np.int32)
    T, nc, g = _c(T_final, np.float32), _c(n_contrib, np.int32), _c(dout, np.float32)
    gT = None if dT is None else _c(dT, np.float32)
    n, C = f.shape
    m = i.shape[0]
    rec = np.empty((max(m, 1), 6 + C), np.float32)
    d_uv = np.zeros((n, 2), np.float32)
    d_c = np.zeros((n, 3), np.float32)
    d_o = np.zeros(n, np.float32)
    d_f = np.zeros((n, C), np.float32)
    lib().ref_blend_bwd(_p(u), _p(c), _p(o), _p(f), C, _p(i), _p(r), bg, W, H, tile_w, tile_h,
                        _p(T), _p(nc), _p(g), None if gT is None else _p(gT), n, m, _p(rec),
                        _p(d_uv), _p(d_c), _p(d_o), _p(d_f))
    return d_uv, d_c, d_o, d_f


def preprocess_bwd(xyz, scale, rot, shs, intr, extr, center, W, H, radius, d_uv, d_conic,
                   d_feat, degree=-1, flags=0, nearest=0.2, extent=1.3):
    """Backward of preprocess -> (d_xyz [n,3], d_scale [n,3], d_rot [n,4],
    d_shs [n,K,3], d_intr [4], d_extr [3,4], d_center [3])."""
    x, sc, q = _c(xyz, np.float32), _c(scale, np.float32), _c(rot, np.float32)
    sh = _c(shs, np.float32)
    K = _c(np.asarray(intr).reshape(-1)[:4], np.float32)
    E = _c(np.asarray(extr)[:3, :4], np.float32)
    cen = _c(np.asarray(center).reshape(-1)[:3], np.float32)
    rad = _c(np.asarray(radius).reshape(-1), np.int32)
    n, Kc = x.shape[0], sh.shape[1]
    gu, gc, gf = _c(d_uv, np.float32), _c(d_conic, np.float32), _c(d_feat, np.float32)
    d_x = np.zeros((n, 3), np.float32)
    d_s = np.zeros((n, 3), np.float32)
    d_q = np.zeros((n, 4), np.float32)
    d_sh = np.zeros((n, Kc, 3), np.float32)
    d_k = np.zeros(4, np.float32)
    d_e = np.zeros((3, 4), np.float32)
    d_c = np.zeros(3, np.float32)
    lib().ref_preprocess_bwd(_p(x), _p(sc), _p(q), _p(sh), n, Kc,
                             -1 if degree is None else degree, _p(K), _p(E), _p(cen), W, H,
                             nearest, extent, flags, _p(rad), _p(gu), _p(gc), _p(gf), _p(d_x),
                             _p(d_s), _p(d_q), _p(d_sh), _p(d_k), _p(d_e), _p(d_c))
    return d_x, d_s, d_q, d_sh, d_k, d_e, d_c


def normal_planes(raw):
    """(n / |n| + 1) / 2 of a [3,H,W] blended normal (examples/supervise/
    renderer.py:89-96), in float32."""
    n = np.asarray(raw, np.float32)
    r = np.sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2])
    return ((n / r + np.float32(1.0)) / np.float32(2.0)).astype(np.float32)


def normal_planes_bwd(raw, g):
    """Gradient of normal_planes: d n = (g - n (n.g) / |n|^2) / (2 |n|)."""
    n = np.asarray(raw, np.float32)
    g = np.asarray(g, np.float32)
    r2 = (n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]
    r = np.sqrt(r2)
    ng = (n[0] * g[0] + n[1] * g[1]) + n[2] * g[2]
    return ((g - n * (ng / r2)) / (np.float32(2.0) * r)).astype(np.float32)


def render_step(inputs, camera, degree=3, flags=0, bg=0.0, image_grad=None, keep=False):
    """One render_iter forward and backward on the host, FP32 throughout:
    preprocess, binning, blending, then the blend backward and the
    preprocess backward.

    inputs: activated splat tensors as numpy arrays -- position [N,3],
    scaling [N,3], rotation [N,4] (unit wxyz), opacity [N] or [N,1], shs
    [N,K,3].  camera: (W, H, intr [4], extr [3,4] or [4,4], centre [3]).
    flags: 1 depth channel, 2 camera-space normal channel, 8 plain colour.
    image_grad(planes) -> (loss, grads): planes {"rgb", "depth"?, "normal"?}
    as the renderer returns them (the normal group mapped to [0,1]), grads
    the same keys.  Returns {"loss", "planes", "grads": {position, scaling,
    rotation, opacity, shs, intr, extr, center}, and with ``keep`` the
    forward intermediates}."""
    W, H, K, E, cen = camera
    x = _c(inputs["position"], np.float32)
    s = _c(inputs["scaling"], np.float32)
    q = _c(inputs["rotation"], np.float32)
    op = _c(np.asarray(inputs["opacity"]).reshape(-1), np.float32)
    sh = _c(inputs["shs"], np.float32)
    uv, d, con, rad, tl, feat = preprocess(x, s, q, sh, K, E, cen, W, H, degree, flags)
    ids, rng = sort_gaussian(uv, d, W, H, rad)
    out, T, nc = alpha_blending(uv, con, op, feat, ids, rng, bg, W, H)
    names = [("rgb", 3)] + ([("depth", 1)] if flags & 1 else []) + \
        ([("normal", 3)] if flags & 2 else [])
    planes, start = {}, 0
    for name, c in names:
        planes[name] = out[start:start + c]
        start += c
    if "normal" in planes:
        planes["normal"] = normal_planes(planes["normal"])
    res = {"planes": planes, "T_final": T, "n_contrib": nc}
    if keep:
        res.update(uv=uv, depth=d, conic=con, radius=rad, tiles=tl, features=feat, ids=ids,
                   tile_range=rng, rendered=out)
    if image_grad is None:
        return res
    loss, gp = image_grad(planes)
    dout = np.zeros_like(out)
    start = 0
    for name, c in names:
        g = np.asarray(gp[name], np.float32)
        if name == "normal":
            g = normal_planes_bwd(out[start:start + c], g)
        dout[start:start + c] = g
        start += c
    du, dc, do, df = alpha_blending_bwd(uv, con, op, feat, ids, rng, bg, W, H, T, nc, dout)
    gx, gs, gq, gsh, gk, ge, gcen = preprocess_bwd(x, s, q, sh, K, E, cen, W, H, rad, du, dc, df,
                                                  degree, flags)
    res["loss"] = loss
    res["grads"] = {"position": gx, "scaling": gs, "rotation": gq, "opacity": do, "shs": gsh,
                    "intr": gk, "extr": ge, "center": gcen, "uv": du}
    return res
