"""Opacity-aware ("tight") binning in the FP32 C restatement (host only).

The tile test must never reject a tile where the blend would let the splat
contribute: for single-splat renders of one 16x16 tile, every rejected case
leaves all 256 pixels untouched (T_final exactly 1), with the splat swept
densely across the acceptance boundary.  The tight sort keeps, per tile, a
subsequence of the exact sort's list, and blending either gives the same
image bit for bit.
"""
import math

import numpy as np

from oracle import ref32 as r32


def random_conic(rng, thin=False):
    s1, s2 = np.exp(rng.uniform(-1.0, 3.5, 2))
    if thin:
        s1, s2 = max(s1, s2) * 6.0, min(s1, s2) * 0.2
    th = rng.uniform(0, math.pi)
    R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
    cov = R @ np.diag([s1 * s1, s2 * s2]) @ R.T + 0.3 * np.eye(2)
    inv = np.linalg.inv(cov)
    return np.float32(inv[0, 0]), np.float32(inv[0, 1]), np.float32(inv[1, 1])


def one_tile_render(u, v, a, b, c, o):
    uv = np.array([[u, v]], np.float32)
    conic = np.array([[a, b, c]], np.float32)
    op = np.array([o], np.float32)
    feat = np.ones((1, 1), np.float32)
    ids = np.array([0], np.int32)
    rng = np.array([[0, 1]], np.int32)
    out, T, nc = r32.alpha_blending(uv, conic, op, feat, ids, rng, 0.0, 16, 16)
    return T


def check_case(u, v, a, b, c, o, stats):
    keep = r32.tile_reachable(u, v, a, b, c, o, 0.0, 15.0, 0.0, 15.0)
    T = one_tile_render(u, v, a, b, c, o)
    touched = bool((T != 1.0).any())
    if touched:
        assert keep, (u, v, a, b, c, o)
    stats["keep"] += keep
    stats["touched"] += touched
    stats["n"] += 1


def test_tile_test_is_conservative_random():
    rng = np.random.default_rng(5)
    stats = {"keep": 0, "touched": 0, "n": 0}
    for k in range(6000):
        a, b, c = random_conic(rng, thin=(k % 3 == 0))
        o = np.float32(np.exp(rng.uniform(math.log(1e-3), 0.0)))
        u = np.float32(rng.uniform(-80, 96))
        v = np.float32(rng.uniform(-80, 96))
        check_case(u, v, a, b, c, o, stats)
    # the test rejects a good share of the cases and keeps few untouched ones
    assert stats["keep"] < 0.7 * stats["n"]
    assert stats["keep"] - stats["touched"] < 0.25 * stats["n"]


def test_tile_test_is_conservative_at_the_boundary():
    rng = np.random.default_rng(6)
    stats = {"keep": 0, "touched": 0, "n": 0}
    for k in range(40):
        a, b, c = random_conic(rng, thin=(k % 2 == 0))
        o = np.float32(np.exp(rng.uniform(math.log(5e-3), 0.0)))
        v = np.float32(rng.uniform(-10, 26))
        # sweep u across the acceptance boundary in 0.01 px steps
        lo = None
        for u in np.arange(16.0, 120.0, 0.5, dtype=np.float32):
            if not r32.tile_reachable(u, v, a, b, c, o, 0, 15, 0, 15):
                lo = u
                break
        if lo is None:
            continue
        for u in np.arange(lo - 1.0, lo + 1.0, 0.01, dtype=np.float32):
            check_case(np.float32(u), v, a, b, c, o, stats)
    assert stats["n"] > 1000


def test_tight_sort_is_a_subsequence_and_blends_identically():
    rng = np.random.default_rng(7)
    n, W, H = 3000, 96, 80
    uv = np.stack([rng.uniform(-10, W + 10, n), rng.uniform(-10, H + 10, n)], 1).astype(np.float32)
    depth = rng.uniform(0.5, 9.0, n).astype(np.float32)
    conic = np.array([random_conic(rng, thin=(i % 4 == 0)) for i in range(n)], np.float32)
    # radius as the EWA rule would give for these conics (3 sigma of the major axis)
    radius = np.zeros(n, np.int32)
    for i in range(n):
        a, b, c = conic[i]
        det = a * c - b * b
        cov = np.array([[c, -b], [-b, a]]) / det
        lam = np.linalg.eigvalsh(cov)[-1]
        radius[i] = math.ceil(3.0 * math.sqrt(lam))
    op = np.exp(rng.uniform(math.log(2e-3), 0.0, n)).astype(np.float32)
    feat = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    ids_e, rng_e = r32.sort_gaussian(uv, depth, W, H, radius)
    ids_t, rng_t = r32.sort_gaussian_tight(uv, depth, W, H, radius, conic, op)
    assert 0 < ids_t.shape[0] < ids_e.shape[0]
    for t in range(rng_e.shape[0]):
        le = list(ids_e[rng_e[t, 0]:rng_e[t, 1]])
        lt = list(ids_t[rng_t[t, 0]:rng_t[t, 1]])
        it = iter(le)
        assert all(x in it for x in lt), t          # subsequence, order kept
    out_e, T_e, _ = r32.alpha_blending(uv, conic, op, feat, ids_e, rng_e, 0.0, W, H)
    out_t, T_t, _ = r32.alpha_blending(uv, conic, op, feat, ids_t, rng_t, 0.0, W, H)
    assert np.array_equal(out_e, out_t)
    assert np.array_equal(T_e, T_t)


def test_log_pos_accuracy():
    xs = np.concatenate([np.exp(np.linspace(-20, 20, 4001)), [1.0, 2.0, 0.5, 255.0, 1e-3]])
    worst = 0.0
    for x in xs.astype(np.float32):
        got = r32.log_pos(float(x))
        ref = math.log(float(x))
        worst = max(worst, abs(got - ref) / max(abs(ref), 1.0))
    assert worst < 3e-7, worst
    assert r32.log_pos(1.0) == 0.0
