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
        L.ref_preprocess.argtypes = [P, P, P, P, I64, I, I, P, P, P, I, I, F, F, I, P, P, P, P, P,
                                     P]
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
