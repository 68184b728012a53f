"""GPU parity of the HIP ops (through the C ABI) against the oracles.

* forward: bit-exact against the FP32 C restatement (oracle/ref32.c) given the
  same inputs -- uv, depth, cov3d, conic, radius, tiles, sorted ids, tile
  ranges, images, T_final, contributor counts;
* forward vs the FP64 transcription (oracle/torch64.py): images within 1e-5
  absolute (the FP32 error of the transmittance product is stated per test);
* backward vs FP64 autograd: max |g - g64| <= 1e-4 * max |g64| per tensor
  (atomic accumulation order is not fixed).
"""
import numpy as np
import pytest
import torch

from oracle import ref32 as r32
from oracle import torch64 as o64
from pointrix_amd import ops
from pointrix_amd import scene as S

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-5
GRAD_TOL = 1e-4
# FP32 vs FP64: the transmittance product and the colour sums round once per
# contributing splat, so a pixel with ~100 contributors can sit a few 1e-6
# above 1e-5; the image bound is met by >= 99.9% of the pixels and the worst
# pixel is reported (the FP32 restatement itself is matched bit for bit).
PIX_FRAC = 0.999


def rel_err(a, b):
    """max |a - b| / max |b| (b None means an all-zero gradient)."""
    a = a.detach().double().cpu()
    b = torch.zeros_like(a) if b is None else b.detach().double().cpu()
    den = b.abs().max().item()
    return (a - b).abs().max().item() / (den if den > 0 else 1.0)


@pytest.fixture(scope="module")
def small():
    scene, cams = S.cpu10k()
    inp = {k: v.detach() for k, v in scene.render_inputs().items()}
    return inp, cams[0]


@pytest.fixture(scope="module")
def pre32(small):
    """FP32 restatement outputs of the preprocess on the 10k scene."""
    inp, cam = small
    K = cam.intrinsic_params.numpy()
    E = cam.extrinsic_matrix.numpy()
    uv, d = r32.project_point(inp["position"].numpy(), K, E, cam.width, cam.height, 0.2)
    vis = d != 0
    cov = r32.compute_cov3d(inp["scaling"].numpy(), inp["rotation"].numpy(), vis)
    con, rad, tl = r32.ewa_project(inp["position"].numpy(), cov, K, E, uv, cam.width, cam.height,
                                   vis)
    ids, rng = r32.sort_gaussian(uv, d, cam.width, cam.height, rad)
    return dict(uv=uv, depth=d, vis=vis, cov=cov, conic=con, radius=rad, tiles=tl, ids=ids,
                rng=rng)


@pytest.mark.parametrize("a", [2.0 ** -4, 2.0 ** -7, 2.0 ** -10])
def test_exp_matches_host(device, a):
    """The GPU's exp_neg, read out of one-splat blends: a splat at (0, 8)
    with conic (a, 0, a) and opacity 0.5 over a 128x16 image.  On the row y
    = 8 the power is -(a/2) px^2, exact in FP32 for these a (powers of two)
    and integer px, and the blended value is 0.5 exp_neg(power) exactly (one
    contributor, T = 1, feature 1, background 0), so twice the pixel must
    equal the host restatement's exp_neg(power) bit for bit; the whole image
    must equal the restatement's blend."""
    W, H = 128, 16
    uv = torch.tensor([[0.0, 8.0]])
    conic = torch.tensor([[a, 0.0, a]])
    op = torch.tensor([[0.5]])
    feat = torch.ones(1, 1)
    T = ((W + 15) // 16) * ((H + 15) // 16)
    ids = torch.zeros(T, dtype=torch.int32)
    rng = torch.stack([torch.arange(T), torch.arange(T) + 1], 1).to(torch.int32)
    t = lambda x: x.to(device)
    out, Tf, nc = ops.alpha_blending_ex(t(uv), t(conic), t(op), t(feat), t(ids), t(rng), 0.0, W,
                                        H)
    row = out[0, 8].cpu().numpy()
    checked = 0
    for px in range(W):
        power = np.float32(-(a / 2.0) * px * px)
        e = r32.exp_neg(float(power))
        alpha = np.float32(0.5) * np.float32(e)
        if power < np.float32(-5.6) or alpha < np.float32(1.0) / np.float32(255.0):
            assert row[px] == 0.0 and int(nc[8, px]) == 0
            continue
        assert np.float32(2.0) * row[px] == np.float32(e), (px, float(power), row[px], e)
        assert abs(e - np.exp(np.float64(power))) <= 2.0 * np.spacing(np.float32(e))
        checked += 1
    assert checked >= 3
    ref = r32.alpha_blending(uv.numpy(), conic.numpy(), op.numpy(), feat.numpy(), ids.numpy(),
                             rng.numpy(), 0.0, W, H)
    assert np.array_equal(out.cpu().numpy(), ref[0])
    assert np.array_equal(nc.cpu().numpy(), ref[2])


def test_project_point(device, small, pre32):
    inp, cam = small
    uv, d = ops.project_point(inp["position"].to(device), cam.intrinsic_params.to(device),
                              cam.extrinsic_matrix[:3].to(device), cam.width, cam.height, 0.2)
    assert np.array_equal(uv.cpu().numpy(), pre32["uv"])
    assert np.array_equal(d.cpu().numpy(), pre32["depth"])


def test_cov3d_ewa(device, small, pre32):
    inp, cam = small
    vis = torch.from_numpy(pre32["vis"]).to(device)
    cov = ops.compute_cov3d(inp["scaling"].to(device), inp["rotation"].to(device), vis)
    assert np.array_equal(cov.cpu().numpy(), pre32["cov"])
    con, rad, tl = ops.ewa_project(inp["position"].to(device), cov, cam.intrinsic_params.to(device),
                                   cam.extrinsic_matrix[:3].to(device),
                                   torch.from_numpy(pre32["uv"]).to(device), cam.width,
                                   cam.height, vis)
    assert np.array_equal(rad.cpu().numpy(), pre32["radius"])
    assert np.array_equal(tl.cpu().numpy(), pre32["tiles"])
    assert np.array_equal(con.cpu().numpy(), pre32["conic"])


@pytest.mark.parametrize("K,degree", [(1, None), (4, None), (9, None), (16, None), (16, 1),
                                      (16, 2)])
def test_compute_sh(device, K, degree):
    g = torch.Generator().manual_seed(K)
    n, C = 5000, 3
    shs = torch.randn(n, K, C, generator=g)
    dirs = torch.nn.functional.normalize(torch.randn(n, 3, generator=g), dim=1)
    out = ops.compute_sh(shs.to(device).permute(0, 2, 1), dirs.to(device), degree=degree)
    ref = r32.compute_sh(shs.permute(0, 2, 1).numpy(), dirs.numpy(), -1 if degree is None else degree)
    assert np.array_equal(out.cpu().numpy(), ref)
    # backward vs FP64
    sd = shs.to(device).requires_grad_(True)
    dd = dirs.to(device).requires_grad_(True)
    gout = torch.randn(n, C, generator=g)
    (ops.compute_sh(sd.permute(0, 2, 1), dd, degree=degree) * gout.to(device)).sum().backward()
    s64 = shs.double().requires_grad_(True)
    d64 = dirs.double().requires_grad_(True)
    (o64.compute_sh(s64.permute(0, 2, 1), d64, degree=degree) * gout.double()).sum().backward()
    assert rel_err(sd.grad, s64.grad) < GRAD_TOL
    assert rel_err(dd.grad, d64.grad) < GRAD_TOL


def test_sort_bitexact(device, small, pre32):
    inp, cam = small
    ids, rng = ops.sort_gaussian(torch.from_numpy(pre32["uv"]).to(device),
                                 torch.from_numpy(pre32["depth"]).to(device), cam.width,
                                 cam.height, torch.from_numpy(pre32["radius"]).to(device), None)
    assert np.array_equal(ids.cpu().numpy(), pre32["ids"])
    assert np.array_equal(rng.cpu().numpy(), pre32["rng"])


@pytest.mark.parametrize("n,W,H,seed,ties", [(200_000, 1920, 1080, 1, False),
                                             (50_000, 979, 546, 2, True),
                                             (3_000, 33, 17, 3, True),
                                             (1, 16, 16, 4, False)])
def test_sort_random_bitexact(device, n, W, H, seed, ties):
    """Random splats incl. off-screen rects, radius 0, equal depths (ties by id)."""
    g = torch.Generator().manual_seed(seed)
    uv = (torch.rand(n, 2, generator=g) * 1.4 - 0.2) * torch.tensor([W, H])
    radius = torch.randint(0, 40, (n,), generator=g, dtype=torch.int32)
    radius[torch.rand(n, generator=g) < 0.01] = 300
    depth = torch.rand(n, 1, generator=g) * 50 + 0.2
    if ties:
        depth = torch.round(depth)
    ids, rng = ops.sort_gaussian(uv.to(device), depth.to(device), W, H, radius.to(device), None)
    ids_r, rng_r = r32.sort_gaussian(uv.numpy(), depth.numpy(), W, H, radius.numpy())
    assert np.array_equal(ids.cpu().numpy(), ids_r)
    assert np.array_equal(rng.cpu().numpy(), rng_r)


def _queued_emission_case(seed=9, n_big=192, n_small=3000, W=1920, H=1080):
    """192 large splats (radius 700-1200 px: most of the 120x68 tile grid)
    consecutive in depth order, at random splat indices among 3000 small far
    ones.  Each 64-splat wave of large splats owns ~500k instances: far past
    the 8192 a wave writes itself, so it queues more than 64 items of 4096
    for the second emission kernel."""
    g = torch.Generator().manual_seed(seed)
    n = n_big + n_small
    uv = torch.rand(n, 2, generator=g) * torch.tensor([W, H])
    radius = torch.randint(1, 20, (n,), generator=g, dtype=torch.int32)
    depth = 1.0 + 10.0 * torch.rand(n, 1, generator=g)
    big = torch.randperm(n, generator=g)[:n_big]
    radius[big] = torch.randint(700, 1200, (n_big,), generator=g, dtype=torch.int32)
    depth[big] = 0.5 + 0.1 * torch.rand(n_big, 1, generator=g)
    return uv, depth, radius, W, H


def _emission_order_reference(uv, depth, radius, W, H, cap):
    """The first ``cap`` instances of the depth-first emission (splats in
    (depth bits, index) order, each its tile rectangle row by row), then
    stably sorted by tile: what a sort whose instance buffers hold only
    ``cap`` instances produces.  Rectangles in FP32 as the kernels compute
    them."""
    u, v = uv[:, 0].astype(np.float32), uv[:, 1].astype(np.float32)
    r = radius.astype(np.float32)
    gx, gy = (W + 15) // 16, (H + 15) // 16
    f16, f15 = np.float32(16.0), np.float32(15.0)
    x0 = np.minimum(np.maximum((u - r) / f16, 0), gx).astype(np.int64)
    x1 = np.minimum(np.maximum(((u + r) + f15) / f16, 0), gx).astype(np.int64)
    y0 = np.minimum(np.maximum((v - r) / f16, 0), gy).astype(np.int64)
    y1 = np.minimum(np.maximum(((v + r) + f15) / f16, 0), gy).astype(np.int64)
    area = (x1 - x0) * (y1 - y0)
    vis = np.nonzero((radius > 0) & (area > 0))[0]
    bits = depth.reshape(-1).astype(np.float32).view(np.uint32)
    order = vis[np.argsort(bits[vis], kind="stable")]
    tiles, owners, total = [], [], 0
    for i in order:
        if total >= cap:
            break
        ty, tx = np.meshgrid(np.arange(y0[i], y1[i]), np.arange(x0[i], x1[i]), indexing="ij")
        t = (ty * gx + tx).reshape(-1)
        tiles.append(t)
        owners.append(np.full(t.shape, i, np.int32))
        total += t.size
    tiles = np.concatenate(tiles)[:cap]
    owners = np.concatenate(owners)[:cap]
    perm = np.argsort(tiles, kind="stable")
    ts, ids = tiles[perm], owners[perm]
    rng = np.zeros((gx * gy, 2), np.int32)
    first = np.searchsorted(ts, np.arange(gx * gy), side="left")
    last = np.searchsorted(ts, np.arange(gx * gy), side="right")
    nz = last > first
    rng[nz, 0], rng[nz, 1] = first[nz], last[nz]
    # instances per 64-splat emission wave (splats in depth order)
    a = area[order]
    waves = np.add.reduceat(a, np.arange(0, len(a), 64)) if len(a) else a
    return ids, rng, int(area[vis].sum()), waves


def test_sort_queued_emission_bitexact(device):
    """Waves of large splats past the per-wave emission span (the queued
    second emission kernel): ids and tile ranges bit-exact."""
    uv, depth, radius, W, H = _queued_emission_case()
    ids, rng = ops.sort_gaussian(uv.to(device), depth.to(device), W, H, radius.to(device), None)
    ids_r, rng_r = r32.sort_gaussian(uv.numpy(), depth.numpy(), W, H, radius.numpy())
    waves = _emission_order_reference(uv.numpy(), depth.numpy(), radius.numpy(), W, H, 1)[3]
    print(f"queued emission: M = {len(ids_r)}, largest wave {waves.max()} instances")
    # a wave writes 8192 itself and queues the rest as 4096-instance items:
    # more than 64 items means the item loop's second pass runs
    assert waves.max() > 8192 + 64 * 4096
    assert np.array_equal(ids.cpu().numpy(), ids_r)
    assert np.array_equal(rng.cpu().numpy(), rng_r)


@pytest.mark.parametrize("frac", [0.4, 0.97])
def test_sort_capacity_below_count(device, frac):
    """A fixed instance capacity below the instance count (what a captured
    graph uses): the device count still reports the true M, the emission
    stops at the capacity -- the ids are the first ``cap`` instances of the
    depth-first emission, tile-sorted, and the ranges index them -- and
    nothing is written past it."""
    uv, depth, radius, W, H = _queued_emission_case(seed=10)
    ids_full, _ = r32.sort_gaussian(uv.numpy(), depth.numpy(), W, H, radius.numpy())
    m = len(ids_full)
    cap = int(m * frac)
    ws = ops.SortWorkspace()
    ws.capacity = cap
    ids, rng = ops.sort_gaussian(uv.to(device), depth.to(device), W, H, radius.to(device), None,
                                 workspace=ws)
    assert ids.shape[0] == cap and ws.instances() == m
    ids_e, rng_e, m_e, _ = _emission_order_reference(uv.numpy(), depth.numpy(),
                                                     radius.numpy(), W, H, cap)
    assert m_e == m
    assert np.array_equal(ids.cpu().numpy(), ids_e)
    assert np.array_equal(rng.cpu().numpy(), rng_e)
    # the same workspace with an exact sort afterwards is unaffected
    ws.capacity = None
    ids2, rng2 = ops.sort_gaussian(uv.to(device), depth.to(device), W, H, radius.to(device),
                                   None, workspace=ws)
    assert np.array_equal(ids2.cpu().numpy(), ids_full)


def test_sort_empty(device):
    uv = torch.zeros(0, 2, device=device)
    d = torch.zeros(0, 1, device=device)
    r = torch.zeros(0, dtype=torch.int32, device=device)
    ids, rng = ops.sort_gaussian(uv, d, 64, 48, r, None)
    assert ids.numel() == 0 and rng.shape == (12, 2) and int(rng.abs().sum()) == 0
    r1 = torch.zeros(5, dtype=torch.int32, device=device)
    ids, rng = ops.sort_gaussian(torch.ones(5, 2, device=device), torch.ones(5, 1, device=device),
                                 64, 48, r1, None)
    assert ids.numel() == 0 and int(rng.abs().sum()) == 0


@pytest.mark.parametrize("C,bg", [(3, 0.0), (4, 1.0), (7, 0.0), (16, 1.0)])
def test_blend_forward(device, small, pre32, C, bg):
    inp, cam = small
    W, H = cam.width, cam.height
    g = torch.Generator().manual_seed(C)
    feat = torch.rand(len(pre32["uv"]), C, generator=g)
    t = lambda a: torch.from_numpy(np.asarray(a)).to(device)
    out, T, nc = ops.alpha_blending_ex(t(pre32["uv"]), t(pre32["conic"]), inp["opacity"].to(device),
                                       feat.to(device), t(pre32["ids"]), t(pre32["rng"]), bg, W, H)
    out_r, T_r, nc_r = r32.alpha_blending(pre32["uv"], pre32["conic"], inp["opacity"].numpy(),
                                          feat.numpy(), pre32["ids"], pre32["rng"], bg, W, H)
    assert np.array_equal(nc.cpu().numpy(), nc_r)
    assert np.array_equal(T.cpu().numpy(), T_r)
    assert np.array_equal(out.cpu().numpy(), out_r)
    out64, T64, nc64 = o64.alpha_blending(torch.from_numpy(pre32["uv"]),
                                          torch.from_numpy(pre32["conic"]), inp["opacity"], feat,
                                          torch.from_numpy(pre32["ids"]),
                                          torch.from_numpy(pre32["rng"]), bg, W, H,
                                          return_aux=True)
    same = nc.cpu() == nc64
    assert same.float().mean().item() > 0.999
    perr = (out.cpu().double() - out64).abs().amax(0)[same]
    frac = (perr <= IMG_TOL).double().mean().item()
    print(f"blend C={C}: max |img - fp64| = {perr.max().item():.3g}, "
          f"{100 * frac:.4f}% of pixels within {IMG_TOL}")
    assert frac >= PIX_FRAC and perr.max().item() < 5 * IMG_TOL


@pytest.mark.parametrize("C,bg,with_T", [(3, 0.0, False), (7, 1.0, True)])
def test_blend_backward(device, small, pre32, C, bg, with_T):
    inp, cam = small
    W, H = cam.width, cam.height
    g = torch.Generator().manual_seed(100 + C)
    n = len(pre32["uv"])
    feat = torch.rand(n, C, generator=g)
    gout = torch.randn(C, H, W, generator=g)
    gT = torch.randn(H, W, generator=g)
    leaves = [torch.from_numpy(pre32["uv"]), torch.from_numpy(pre32["conic"]),
              inp["opacity"].clone(), feat]
    dl = [x.to(device).requires_grad_(True) for x in leaves]
    ndc = torch.zeros(n, 2, device=device, requires_grad=True)
    ids = torch.from_numpy(pre32["ids"])
    rng = torch.from_numpy(pre32["rng"])
    out, T, _ = ops.alpha_blending_ex(*dl, ids.to(device), rng.to(device), bg, W, H, ndc)
    loss = (out * gout.to(device)).sum()
    if with_T:
        loss = loss + (T * gT.to(device)).sum()
    loss.backward()
    hl = [x.double().requires_grad_(True) for x in leaves]
    out64, T64, _ = o64.alpha_blending(*hl, ids, rng, bg, W, H, return_aux=True)
    loss64 = (out64 * gout.double()).sum()
    if with_T:
        loss64 = loss64 + (T64 * gT.double()).sum()
    loss64.backward()
    for name, a, b in zip(["uv", "conic", "opacity", "feature"], dl, hl):
        e = rel_err(a.grad, b.grad)
        print(f"blend bwd C={C} d{name}: rel err {e:.3g}")
        assert e < GRAD_TOL, name
    scale = torch.tensor([0.5 * W, 0.5 * H], dtype=torch.float64)
    assert rel_err(ndc.grad, hl[0].grad * scale) < GRAD_TOL


@pytest.mark.parametrize("tile", [(8, 32), (32, 8)])
def test_blend_tile_shapes(device, small, pre32, tile):
    """Non-square tiles: the binning and forward bit-exact against the FP32
    restatement with the same tile shape, the backward against FP64 autograd
    of the same tiling (within the gradient tolerance)."""
    inp, cam = small
    W, H = cam.width, cam.height
    tw, th = tile
    ids_r, rng_r = r32.sort_gaussian(pre32["uv"], pre32["depth"], W, H, pre32["radius"], tw, th)
    t = lambda a: torch.from_numpy(np.asarray(a)).to(device)
    ids, rng = ops.sort_gaussian(t(pre32["uv"]), t(pre32["depth"]), W, H, t(pre32["radius"]),
                                 None, tile_w=tw, tile_h=th)
    assert np.array_equal(ids.cpu().numpy(), ids_r)
    assert np.array_equal(rng.cpu().numpy(), rng_r)
    g = torch.Generator().manual_seed(tw)
    n, C = len(pre32["uv"]), 4
    feat = torch.rand(n, C, generator=g)
    gout = torch.randn(C, H, W, generator=g)
    leaves = [torch.from_numpy(pre32["uv"]), torch.from_numpy(pre32["conic"]),
              inp["opacity"].clone(), feat]
    dl = [x.to(device).requires_grad_(True) for x in leaves]
    out, T, nc = ops.alpha_blending_ex(*dl, ids, rng, 0.0, W, H, None, tw, th)
    out_r, T_r, nc_r = r32.alpha_blending(pre32["uv"], pre32["conic"], inp["opacity"].numpy(),
                                          feat.numpy(), ids_r, rng_r, 0.0, W, H, tw, th)
    assert np.array_equal(out.detach().cpu().numpy(), out_r)
    assert np.array_equal(nc.cpu().numpy(), nc_r)
    (out * gout.to(device)).sum().backward()
    hl = [x.double().requires_grad_(True) for x in leaves]
    out64 = o64.alpha_blending(*hl, torch.from_numpy(ids_r), torch.from_numpy(rng_r), 0.0, W, H,
                               tw=tw, th=th)
    (out64 * gout.double()).sum().backward()
    for name, a, b in zip(["uv", "conic", "opacity", "feature"], dl, hl):
        e = rel_err(a.grad, b.grad)
        print(f"tile {tw}x{th} blend bwd d{name}: rel err {e:.3g}")
        assert e < GRAD_TOL, name


def test_preprocess_backward(device, small, pre32):
    inp, cam = small
    W, H = cam.width, cam.height
    g = torch.Generator().manual_seed(7)
    n = inp["position"].shape[0]
    K = cam.intrinsic_params.clone()
    E = cam.extrinsic_matrix[:3].clone()
    vis = torch.from_numpy(pre32["vis"])
    # project_point
    g_uv, g_d = torch.randn(n, 2, generator=g), torch.randn(n, 1, generator=g)
    xs = [inp["position"].to(device).requires_grad_(True), K.to(device).requires_grad_(True),
          E.to(device).requires_grad_(True)]
    uv, d = ops.project_point(*xs, W, H, 0.2)
    ((uv * g_uv.to(device)).sum() + (d * g_d.to(device)).sum()).backward()
    hs = [inp["position"].double().requires_grad_(True), K.double().requires_grad_(True),
          E.double().requires_grad_(True)]
    uv64, d64 = o64.project_point(*hs, W, H, 0.2)
    ((uv64 * g_uv.double()).sum() + (d64 * g_d.double()).sum()).backward()
    for name, a, b in zip(["xyz", "intr", "extr"], xs, hs):
        assert rel_err(a.grad, b.grad) < GRAD_TOL, f"project d{name}"
    # compute_cov3d
    g_c = torch.randn(n, 6, generator=g)
    xs = [inp["scaling"].to(device).requires_grad_(True),
          inp["rotation"].to(device).requires_grad_(True)]
    (ops.compute_cov3d(*xs, vis.to(device)) * g_c.to(device)).sum().backward()
    hs = [inp["scaling"].double().requires_grad_(True),
          inp["rotation"].double().requires_grad_(True)]
    (o64.compute_cov3d(*hs, vis) * g_c.double()).sum().backward()
    for name, a, b in zip(["scale", "rot"], xs, hs):
        assert rel_err(a.grad, b.grad) < GRAD_TOL, f"cov3d d{name}"
    # ewa_project
    g_con = torch.randn(n, 3, generator=g)
    cov = torch.from_numpy(pre32["cov"])
    uvt = torch.from_numpy(pre32["uv"])
    xs = [inp["position"].to(device).requires_grad_(True), cov.to(device).requires_grad_(True),
          K.to(device).requires_grad_(True), E.to(device).requires_grad_(True)]
    con, _, _ = ops.ewa_project(xs[0], xs[1], xs[2], xs[3], uvt.to(device), W, H, vis.to(device))
    (con * g_con.to(device)).sum().backward()
    hs = [inp["position"].double().requires_grad_(True), cov.double().requires_grad_(True),
          K.double().requires_grad_(True), E.double().requires_grad_(True)]
    con64, _, _ = o64.ewa_project(hs[0], hs[1], hs[2], hs[3], uvt, W, H, vis)
    (con64 * g_con.double()).sum().backward()
    for name, a, b in zip(["xyz", "cov3d", "intr", "extr"], xs, hs):
        e = rel_err(a.grad, b.grad)
        print(f"ewa d{name}: rel err {e:.3g}")
        assert e < GRAD_TOL, f"ewa d{name}"


@pytest.mark.parametrize("n", [1, 1000, 3 * 1920 * 1080])
def test_l1_loss(device, n):
    g = torch.Generator().manual_seed(n)
    x = torch.rand(n, generator=g).to(device).requires_grad_(True)
    y = torch.rand(n, generator=g).to(device)
    loss = ops.l1_loss(x, y)
    loss.backward()
    x64 = x.detach().double().cpu().requires_grad_(True)
    ref = (x64 - y.double().cpu()).abs().mean()
    ref.backward()
    assert abs(loss.item() - ref.item()) <= 1e-6 * max(1.0, abs(ref.item()))
    assert torch.allclose(x.grad.double().cpu(), x64.grad, rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("sizes", [[1], [3, 1, 3], [5, 2, 1, 7, 3, 1, 2, 4]])
@pytest.mark.parametrize("P", [17, 1237 * 822])
def test_l1_loss_sum(device, sizes, P):
    """sum_k w_k mean|x_k - y_k| against the FP64 sum of separate means; the
    x_k are planes of one image split along dim 0 (as the renderer's output
    planes are), one target also takes a gradient and one term's input
    none."""
    g = torch.Generator().manual_seed(len(sizes) * 7 + P)
    img = torch.rand(sum(sizes), P, generator=g)
    tg = [torch.rand(c, P, generator=g) for c in sizes]
    w = [1.0 if k == 0 else 0.1 * (k + 1) for k in range(len(sizes))]
    xd = img.to(device).requires_grad_(True)
    xs = list(torch.split(xd, sizes, dim=0))
    ys = [t.to(device) for t in tg]
    ys[-1].requires_grad_(True)
    frozen = len(sizes) > 2
    if frozen:
        xs[1] = xs[1].detach()
    loss = ops.l1_loss_sum([(x, y, wk) for x, y, wk in zip(xs, ys, w)])
    loss.backward()
    x64 = img.double().requires_grad_(True)
    y64 = tg[-1].double().requires_grad_(True)
    xs64 = list(torch.split(x64, sizes, dim=0))
    if frozen:
        xs64[1] = xs64[1].detach()
    ys64 = [t.double() for t in tg[:-1]] + [y64]
    ref = sum(wk * (x - y).abs().mean() for x, y, wk in zip(xs64, ys64, w))
    ref.backward()
    assert abs(loss.item() - ref.item()) <= 1e-6 * max(1.0, abs(ref.item()))
    assert torch.allclose(xd.grad.double().cpu(), x64.grad, rtol=1e-6, atol=1e-12)
    assert torch.allclose(ys[-1].grad.double().cpu(), y64.grad, rtol=1e-6, atol=1e-12)
