"""Pins the oracles (oracle/torch64.py FP64, oracle/ref32.c FP32) on the CPU.

Each op is checked against something that does not restate its formula:
closed forms and invariants (SH addition theorem and orthonormality by
quadrature, pinhole back-projection round trip, scipy's rotation matrices,
covariance eigenvalues, EWA covariance = J Sigma J^T with J from autograd),
brute force (Python-sorted binning keys, a literal per-pixel blending loop),
special cases (one splat, zero opacity) and central finite differences.
The reference holds no golden vectors for this path; its one constant on the
path, C0 = 0.28209479177387814 (pointrix/utils/common.py:2), is checked.
"""
import math

import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from oracle import ref32 as r32
from oracle import torch64 as o
from pointrix_amd import scene as S

F64 = torch.float64


def unit(v):
    return v / v.norm(dim=-1, keepdim=True)


# ---------------------------------------------------------------------------
# spherical harmonics
# ---------------------------------------------------------------------------
def test_sh_c0_matches_reference_constant():
    assert o.sh_constants()[0] == pytest.approx(0.28209479177387814, abs=1e-17)


def test_sh_addition_theorem():
    """sum_m Y_lm(d)^2 = (2l+1)/(4 pi) for every unit d and band l."""
    d = unit(torch.randn(500, 3, dtype=F64, generator=torch.Generator().manual_seed(0)))
    Y = o.sh_basis(d, 16)
    for l, (a, b) in enumerate([(0, 1), (1, 4), (4, 9), (9, 16)]):
        s = (Y[:, a:b] ** 2).sum(1)
        assert torch.allclose(s, torch.full_like(s, (2 * l + 1) / (4 * math.pi)), atol=1e-12)


def test_sh_orthonormal_by_quadrature():
    """Gauss-Legendre in cos(theta) x uniform phi is exact for these degrees."""
    x, w = np.polynomial.legendre.leggauss(12)
    phi = np.arange(24) * 2 * np.pi / 24
    ct, ph = np.meshgrid(x, phi, indexing="ij")
    st = np.sqrt(1 - ct ** 2)
    d = torch.tensor(np.stack([st * np.cos(ph), st * np.sin(ph), ct], -1).reshape(-1, 3))
    wt = torch.tensor((w[:, None] * np.full_like(ph, 2 * np.pi / 24)).reshape(-1))
    Y = o.sh_basis(d, 16)
    G = (Y * wt[:, None]).T @ Y
    assert torch.allclose(G, torch.eye(16, dtype=F64), atol=1e-12)


def test_sh_degree0_and_degree_mask():
    g = torch.Generator().manual_seed(1)
    shs = torch.randn(50, 3, 16, dtype=F64, generator=g)
    d = unit(torch.randn(50, 3, dtype=F64, generator=g))
    assert torch.allclose(o.compute_sh(shs, d, degree=0), shs[:, :, 0] * 0.28209479177387814)
    masked = shs.clone()
    masked[:, :, 4:] = 0
    assert torch.allclose(o.compute_sh(shs, d, degree=1), o.compute_sh(masked, d))


def test_sh_fp32_restatement_agrees():
    g = torch.Generator().manual_seed(2)
    shs = torch.randn(2000, 3, 16, generator=g)
    d = unit(torch.randn(2000, 3, generator=g))
    a = r32.compute_sh(shs.numpy(), d.numpy())
    b = o.compute_sh(shs, d).numpy()
    assert np.abs(a - b).max() < 2e-6


# ---------------------------------------------------------------------------
# project_point
# ---------------------------------------------------------------------------
def _camera():
    return S.make_camera(320, 200, eye=(1.0, -0.5, -4.0), fov_deg=55.0, target=(0.1, 0.2, 0.3))


def test_project_backprojection_round_trip():
    cam = _camera()
    g = torch.Generator().manual_seed(3)
    pts = torch.randn(400, 3, dtype=F64, generator=g) * 0.8
    uv, d = o.project_point(pts, cam.intrinsic_params, cam.extrinsic_matrix, 320, 200)
    vis = d[:, 0] != 0
    assert vis.sum() > 300
    fx, fy, cx, cy = cam.intrinsic_params.double().tolist()
    z = d[vis, 0]
    xc = (uv[vis, 0] - cx) / fx * z
    yc = (uv[vis, 1] - cy) / fy * z
    pc = torch.stack([xc, yc, z, torch.ones_like(z)], 1)
    c2w = torch.linalg.inv(cam.extrinsic_matrix.double())
    back = (pc @ c2w.T)[:, :3]
    assert torch.allclose(back, pts[vis], atol=1e-10)


def test_project_axis_and_culling():
    cam = _camera()
    E = cam.extrinsic_matrix.double()
    c2w = torch.linalg.inv(E)
    on_axis = (c2w @ torch.tensor([0.0, 0.0, 2.0, 1.0], dtype=F64))[:3]
    behind = (c2w @ torch.tensor([0.0, 0.0, 0.1, 1.0], dtype=F64))[:3]   # z < nearest 0.2
    outside = (c2w @ torch.tensor([10.0, 0.0, 2.0, 1.0], dtype=F64))[:3]  # |2u/W-1| > 1.3
    uv, d = o.project_point(torch.stack([on_axis, behind, outside]), cam.intrinsic_params, E,
                            320, 200)
    assert torch.allclose(uv[0], torch.tensor([160.0, 100.0], dtype=F64), atol=1e-9)
    assert d[0, 0] == pytest.approx(2.0)
    assert torch.all(uv[1:] == 0) and torch.all(d[1:] == 0)


# ---------------------------------------------------------------------------
# compute_cov3d
# ---------------------------------------------------------------------------
def test_rotation_matches_scipy():
    q = unit(torch.randn(100, 4, dtype=F64, generator=torch.Generator().manual_seed(4)))
    R = o.quat_to_rotmat(q)
    ref = Rotation.from_quat(q[:, [1, 2, 3, 0]].numpy()).as_matrix()   # scipy is xyzw
    assert np.allclose(R.numpy(), ref, atol=1e-12)


def test_cov3d_identity_and_eigenvalues():
    g = torch.Generator().manual_seed(5)
    s = torch.rand(200, 3, dtype=F64, generator=g) + 0.1
    q = unit(torch.randn(200, 4, dtype=F64, generator=g))
    vis = torch.ones(200, dtype=torch.bool)
    cov = o.compute_cov3d(s, q, vis)
    ident = torch.tensor([[1.0, 0, 0, 0]], dtype=F64).expand(200, 4)
    ci = o.compute_cov3d(s, ident, vis)
    assert torch.allclose(ci, torch.stack([s[:, 0] ** 2, 0 * s[:, 0], 0 * s[:, 0],
                                           s[:, 1] ** 2, 0 * s[:, 0], s[:, 2] ** 2], 1))
    ev = torch.linalg.eigvalsh(o.unpack_cov3d(cov))
    assert torch.allclose(ev, torch.sort(s ** 2, dim=1).values, atol=1e-12)
    assert torch.all(o.compute_cov3d(s, q, ~vis) == 0)


# ---------------------------------------------------------------------------
# ewa_project
# ---------------------------------------------------------------------------
def test_ewa_isotropic_on_axis_closed_form():
    W = H = 256
    f = 300.0
    z = 3.0
    s = 0.05
    intr = torch.tensor([f, f, W / 2, H / 2], dtype=F64)
    E = torch.eye(4, dtype=F64)[:3]
    xyz = torch.tensor([[0.0, 0.0, z]], dtype=F64)
    cov = torch.tensor([[s * s, 0, 0, s * s, 0, s * s]], dtype=F64)
    uv = torch.tensor([[W / 2, H / 2]], dtype=F64)
    conic, radius, tiles = o.ewa_project(xyz, cov, intr, E, uv, W, H, torch.tensor([True]))
    var = f * f * s * s / (z * z) + 0.3
    assert torch.allclose(conic[0], torch.tensor([1 / var, 0.0, 1 / var], dtype=F64))
    assert int(radius[0]) == math.ceil(3 * math.sqrt(var + math.sqrt(0.1)))


def test_ewa_covariance_is_projection_jacobian():
    """cov2d - 0.3 I = J Sigma J^T with J the autograd Jacobian of the pinhole
    map (points inside the 1.3 tan(fov/2) clamp)."""
    cam = _camera()
    g = torch.Generator().manual_seed(6)
    pts = torch.randn(40, 3, dtype=F64, generator=g) * 0.5
    s = torch.rand(40, 3, dtype=F64, generator=g) * 0.1 + 0.01
    q = unit(torch.randn(40, 4, dtype=F64, generator=g))
    cov = o.compute_cov3d(s, q, torch.ones(40, dtype=torch.bool))
    uv, d = o.project_point(pts, cam.intrinsic_params, cam.extrinsic_matrix, 320, 200)
    conic, radius, _ = o.ewa_project(pts, cov, cam.intrinsic_params, cam.extrinsic_matrix, uv,
                                     320, 200, d[:, 0] != 0)
    for i in range(40):
        if radius[i] == 0:
            continue

        def proj(x):
            return o.project_point(x[None], cam.intrinsic_params, cam.extrinsic_matrix,
                                   320, 200)[0][0]

        J = torch.autograd.functional.jacobian(proj, pts[i])
        c2 = J @ o.unpack_cov3d(cov)[i] @ J.T + 0.3 * torch.eye(2, dtype=F64)
        inv = torch.linalg.inv(c2)
        assert torch.allclose(conic[i], torch.stack([inv[0, 0], inv[0, 1], inv[1, 1]]),
                              rtol=1e-9)
        lam = torch.linalg.eigvalsh(c2).max()
        mid, det = c2.trace() / 2, torch.linalg.det(c2)
        if mid * mid - det > 0.1:
            assert int(radius[i]) == math.ceil(3 * math.sqrt(float(lam)) - 1e-12)


def test_tiles_touched_brute_force():
    g = torch.Generator().manual_seed(7)
    W, H = 100, 70
    uv = torch.rand(300, 2, generator=g, dtype=F64) * 140 - 20
    r = torch.randint(0, 30, (300,), generator=g)
    x0, x1, y0, y1 = o.tile_rect(uv, r, W, H)
    for i in range(300):
        u, v, rr = float(uv[i, 0]), float(uv[i, 1]), int(r[i])
        # definition: tiles t with trunc((u-r)/16) <= t < trunc((u+r+15)/16), clamped
        lo = max(0, int(max(0.0, (u - rr) / 16)))
        hi = min(7, int(max(0.0, (u + rr + 15) / 16)))
        assert (int(x0[i]), int(x1[i])) == (min(lo, 7), hi)
        lo = max(0, int(max(0.0, (v - rr) / 16)))
        hi = min(5, int(max(0.0, (v + rr + 15) / 16)))
        assert (int(y0[i]), int(y1[i])) == (min(lo, 5), hi)


# ---------------------------------------------------------------------------
# sort_gaussian
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ties", [False, True])
def test_sort_brute_force(ties):
    g = torch.Generator().manual_seed(8)
    W, H, n = 90, 60, 120
    uv = torch.rand(n, 2, generator=g) * 110 - 10
    r = torch.randint(0, 25, (n,), generator=g, dtype=torch.int32)
    depth = torch.rand(n, 1, generator=g) * 10 + 0.3
    if ties:
        depth = torch.round(depth)
    ids, rng = o.sort_gaussian(uv, depth, W, H, r, None)
    gx, gy = 6, 4
    items = []
    for i in range(n):
        if r[i] <= 0:
            continue
        u, v, rr = float(uv[i, 0]), float(uv[i, 1]), int(r[i])
        xa = min(gx, int(max(0.0, (u - rr) / 16)))
        xb = min(gx, int(max(0.0, (u + rr + 15) / 16)))
        ya = min(gy, int(max(0.0, (v - rr) / 16)))
        yb = min(gy, int(max(0.0, (v + rr + 15) / 16)))
        dbits = int(np.float32(depth[i, 0]).view(np.uint32))
        for ty in range(ya, yb):
            for tx in range(xa, xb):
                items.append((ty * gx + tx, dbits, i))
    items.sort()
    assert ids.tolist() == [it[2] for it in items]
    for t in range(gx * gy):
        members = [k for k, it in enumerate(items) if it[0] == t]
        exp = (members[0], members[-1] + 1) if members else (0, 0)
        assert tuple(rng[t].tolist()) == exp
    # the FP32 restatement sorts identically
    ids32, rng32 = r32.sort_gaussian(uv.numpy(), depth.numpy(), W, H, r.numpy())
    assert np.array_equal(ids32, ids.numpy()) and np.array_equal(rng32, rng.numpy())


# ---------------------------------------------------------------------------
# alpha_blending
# ---------------------------------------------------------------------------
def _tiny_blend_inputs(seed, n=40, W=32, H=24):
    g = torch.Generator().manual_seed(seed)
    uv = torch.rand(n, 2, generator=g, dtype=F64) * torch.tensor([W, H], dtype=F64)
    a = torch.rand(n, generator=g, dtype=F64) * 0.4 + 0.05
    c = torch.rand(n, generator=g, dtype=F64) * 0.4 + 0.05
    b = (torch.rand(n, generator=g, dtype=F64) - 0.5) * 0.8 * torch.sqrt(a * c)
    conic = torch.stack([a, b, c], 1)
    op = torch.rand(n, 1, generator=g, dtype=F64) * 0.98 + 0.01
    feat = torch.rand(n, 3, generator=g, dtype=F64)
    depth = torch.rand(n, 1, generator=g, dtype=F64) + 1
    radius = torch.full((n,), 12, dtype=torch.int32)
    ids, rng = o.sort_gaussian(uv, depth, W, H, radius, None)
    return uv, conic, op, feat, ids, rng, W, H


@pytest.mark.parametrize("seed,bg", [(0, 0.0), (1, 1.0), (2, 0.3)])
def test_blend_matches_bruteforce_loop(seed, bg):
    uv, conic, op, feat, ids, rng, W, H = _tiny_blend_inputs(seed)
    a, T, nc = o.alpha_blending(uv, conic, op, feat, ids, rng, bg, W, H, return_aux=True)
    b, T2, nc2 = o.alpha_blending_bruteforce(uv, conic, op, feat, ids, rng, bg, W, H)
    assert torch.allclose(a, b, atol=1e-13) and torch.allclose(T, T2, atol=1e-13)
    assert torch.equal(nc, nc2)


def test_blend_single_splat_closed_form():
    W = H = 16
    uv = torch.tensor([[7.3, 8.1]], dtype=F64)
    conic = torch.tensor([[0.2, 0.05, 0.3]], dtype=F64)
    op = torch.tensor([[0.7]], dtype=F64)
    f = torch.tensor([[0.25, 0.5, 1.0]], dtype=F64)
    ids = torch.tensor([0], dtype=torch.int32)
    rng = torch.tensor([[0, 1]], dtype=torch.int32)
    out = o.alpha_blending(uv, conic, op, f, ids, rng, 0.4, W, H)
    ys, xs = torch.meshgrid(torch.arange(H, dtype=F64), torch.arange(W, dtype=F64), indexing="ij")
    dx, dy = 7.3 - xs, 8.1 - ys
    alpha = torch.clamp(0.7 * torch.exp(-0.5 * (0.2 * dx * dx + 0.3 * dy * dy) - 0.05 * dx * dy),
                        max=0.99)
    alpha = torch.where(alpha >= 1 / 255, alpha, torch.zeros_like(alpha))
    exp = f[0][:, None, None] * alpha + 0.4 * (1 - alpha)
    assert torch.allclose(out, exp, atol=1e-14)
    zero = o.alpha_blending(uv, conic, 0 * op, f, ids, rng, 0.4, W, H)
    assert torch.allclose(zero, torch.full_like(zero, 0.4))


def test_blend_fp32_restatement_agrees():
    uv, conic, op, feat, ids, rng, W, H = _tiny_blend_inputs(3, n=200, W=64, H=48)
    a, T, nc = o.alpha_blending(uv, conic, op, feat, ids, rng, 0.0, W, H, return_aux=True)
    b, T2, nc2 = r32.alpha_blending(uv.float().numpy(), conic.float().numpy(),
                                    op.float().numpy(), feat.float().numpy(), ids.numpy(),
                                    rng.numpy(), 0.0, W, H)
    assert np.array_equal(nc.numpy(), nc2)
    assert np.abs(a.numpy() - b).max() < 1e-5


def test_exp_restatement_accuracy():
    xs = np.float32(-np.linspace(0.0, 80.0, 20001))
    rel = max(abs(r32.exp_neg(float(x)) / math.exp(float(x)) - 1.0) for x in xs)
    assert rel < 2e-7


# ---------------------------------------------------------------------------
# gradients: central finite differences in FP64
# ---------------------------------------------------------------------------
def test_op_gradcheck():
    g = torch.Generator().manual_seed(9)
    cam = _camera()
    K = cam.intrinsic_params.double().requires_grad_(True)
    E = cam.extrinsic_matrix[:3].double().requires_grad_(True)
    x = (torch.randn(6, 3, generator=g, dtype=F64) * 0.3).requires_grad_(True)
    assert torch.autograd.gradcheck(
        lambda a, b, c: o.project_point(a, b, c, 320, 200), (x, K, E))
    s = (torch.rand(6, 3, generator=g, dtype=F64) * 0.1 + 0.02).requires_grad_(True)
    q = unit(torch.randn(6, 4, generator=g, dtype=F64)).requires_grad_(True)
    vis = torch.ones(6, dtype=torch.bool)
    assert torch.autograd.gradcheck(lambda a, b: o.compute_cov3d(a, b, vis), (s, q))
    cov = o.compute_cov3d(s, q, vis).detach().requires_grad_(True)
    uv = o.project_point(x, K, E, 320, 200)[0].detach()
    assert torch.autograd.gradcheck(
        lambda a, c, k, e: o.ewa_project(a, c, k, e, uv, 320, 200, vis)[0], (x, cov, K, E))
    shs = torch.randn(6, 3, 16, generator=g, dtype=F64, requires_grad=True)
    d = unit(torch.randn(6, 3, generator=g, dtype=F64)).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda a, b: o.compute_sh(a, b), (shs, d))


def test_blend_gradcheck():
    uv, conic, op, feat, ids, rng, W, H = _tiny_blend_inputs(4, n=12, W=16, H=16)
    leaves = [t.clone().requires_grad_(True) for t in (uv, conic, op, feat)]
    assert torch.autograd.gradcheck(
        lambda a, b, c, d: o.alpha_blending(a, b, c, d, ids, rng, 0.5, W, H), leaves,
        eps=1e-7, atol=1e-6)


def test_render_iter_finite_differences():
    """2,000 splats, 64x64, SH degree 3, rgb+depth: autograd of the FP64
    render_iter vs central differences on randomly chosen coordinates of every
    input, including the camera extrinsics and intrinsics."""
    scene = S.random_scene(2000, 21, sh_degree=3, means="normal", means_std=(0.6, 0.6, 0.6),
                           log_scale=(-3.0, 0.5), opacity_std=1.5)
    cam = S.make_camera(64, 64, eye=(0.3, 0.2, -2.5), fov_deg=55.0)
    inp = {k: v.detach().double() for k, v in scene.render_inputs().items()}
    inp["extrinsic_matrix"] = cam.extrinsic_matrix.double()
    inp["intrinsic_params"] = cam.intrinsic_params.double()
    g = torch.Generator().manual_seed(22)
    w_rgb = torch.randn(3, 64, 64, generator=g, dtype=F64)
    w_d = torch.randn(1, 64, 64, generator=g, dtype=F64)

    def loss_of(t):
        res = o.render_iter(64, 64, t["extrinsic_matrix"], t["intrinsic_params"],
                            cam.camera_center.double(), t["position"], t["opacity"],
                            t["scaling"], t["rotation"], t["shs"], sh_degree=3,
                            render_depth=True)
        f = res["rendered_features_split"]
        return (f["rgb"] * w_rgb).sum() + (f["depth"] * w_d).sum()

    leaves = {k: v.clone().requires_grad_(True) for k, v in inp.items()}
    loss_of(leaves).backward()
    eps = 1e-6
    checked = 0
    for k, v in inp.items():
        grad = leaves[k].grad
        flat = grad.reshape(-1)
        # coordinates with the largest gradients plus random ones
        idx = torch.cat([flat.abs().topk(4).indices,
                         torch.randint(0, flat.numel(), (4,), generator=g)])
        for j in idx.tolist():
            plus = {kk: vv.clone() for kk, vv in inp.items()}
            minus = {kk: vv.clone() for kk, vv in inp.items()}
            plus[k].view(-1)[j] += eps
            minus[k].view(-1)[j] -= eps
            with torch.no_grad():
                fd = (loss_of(plus) - loss_of(minus)).item() / (2 * eps)
            an = flat[j].item()
            scale = max(1.0, flat.abs().max().item())
            assert abs(fd - an) <= 1e-4 * scale, (k, j, fd, an)
            checked += 1
    assert checked == 8 * len(inp)


# ---------------------------------------------------------------------------
# splat normals (examples/supervise/model.py:50-74)
# ---------------------------------------------------------------------------
def _normal_case(n, seed):
    g = torch.Generator().manual_seed(seed)
    pos = torch.randn(n, 3, dtype=F64, generator=g)
    scale = torch.rand(n, 3, dtype=F64, generator=g) + 0.05
    q = unit(torch.randn(n, 4, dtype=F64, generator=g))
    cam = S.make_camera(64, 48, eye=(0.4, -0.3, -4.0), fov_deg=50.0)
    return pos, scale, q, cam.camera_center.double(), cam.extrinsic_matrix.double()


def test_splat_normals_identity_quaternion():
    """Identity rotation: the normal is +-e_k for the smallest scale axis k,
    the sign facing the camera, then rotated into camera space by E[:3,:3]."""
    pos, scale, _, c, E = _normal_case(300, 40)
    q = torch.tensor([[1.0, 0.0, 0.0, 0.0]], dtype=F64).expand(300, 4)
    n_cam = o.splat_normals(pos, scale, q, c, E)
    k = scale.argmin(1)
    for i in range(300):
        e = torch.zeros(3, dtype=F64)
        e[k[i]] = 1.0
        if torch.dot(c - pos[i], e) < 0:
            e = -e
        assert torch.allclose(n_cam[i], E[:3, :3] @ e, atol=1e-14)


def test_splat_normals_match_scipy_rotation():
    """Random rotations: the world normal is the scipy rotation matrix's
    column of the smallest scale axis, unit length, turned to the camera."""
    pos, scale, q, c, E = _normal_case(500, 41)
    n_cam = o.splat_normals(pos, scale, q, c, E)
    Rs = torch.from_numpy(Rotation.from_quat(q[:, [1, 2, 3, 0]].numpy()).as_matrix())
    k = scale.argmin(1)
    world = Rs[torch.arange(500), :, k]
    sign = torch.where(((c[None] - pos) * world).sum(1) < 0, -1.0, 1.0).to(F64)
    want = (sign[:, None] * world) @ E[:3, :3].T
    assert torch.allclose(n_cam, want, atol=1e-12)
    assert torch.allclose(n_cam.norm(dim=1), torch.ones(500, dtype=F64), atol=1e-12)
    # facing the camera: in camera space the normal points against the
    # camera-to-splat direction
    p_cam = pos @ E[:3, :3].T + E[:3, 3]
    assert torch.all((n_cam * p_cam).sum(1) <= 1e-12)


def test_splat_normals_gradcheck():
    pos, scale, q, c, E = _normal_case(20, 42)
    leaves = [t.clone().requires_grad_(True) for t in (q, E)]
    assert torch.autograd.gradcheck(lambda qq, EE: o.splat_normals(pos, scale, qq, c, EE),
                                    leaves, eps=1e-7, atol=1e-7)


def test_render_iter_normal_finite_differences():
    """1,000 splats, 48x40, SH degree 3, rgb + depth + normal (the normal
    channel from splat_normals, mapped to [0,1] after blending): autograd vs
    central differences on every input, the camera centre included."""
    scene = S.random_scene(1000, 23, sh_degree=3, means="normal", means_std=(0.6, 0.6, 0.6),
                           log_scale=(-2.8, 0.5), opacity_std=1.5)
    cam = S.make_camera(48, 40, eye=(0.3, 0.2, -2.5), fov_deg=55.0)
    inp = {k: v.detach().double() for k, v in scene.render_inputs().items()}
    inp["extrinsic_matrix"] = cam.extrinsic_matrix.double()
    inp["intrinsic_params"] = cam.intrinsic_params.double()
    inp["camera_center"] = cam.camera_center.double()
    g = torch.Generator().manual_seed(24)
    w = {k: torch.randn(c, 40, 48, generator=g, dtype=F64)
         for k, c in (("rgb", 3), ("depth", 1), ("normal", 3))}

    def loss_of(t):
        nrm = o.splat_normals(t["position"], t["scaling"], t["rotation"], t["camera_center"],
                              t["extrinsic_matrix"])
        res = o.render_iter(40, 48, t["extrinsic_matrix"], t["intrinsic_params"],
                            t["camera_center"], t["position"], t["opacity"], t["scaling"],
                            t["rotation"], t["shs"], sh_degree=3, render_depth=True,
                            normals=nrm)
        f = res["rendered_features_split"]
        return sum((f[k] * w[k]).sum() for k in w)

    leaves = {k: v.clone().requires_grad_(True) for k, v in inp.items()}
    loss_of(leaves).backward()
    eps = 1e-6
    for k, v in inp.items():
        flat = leaves[k].grad.reshape(-1)
        idx = torch.cat([flat.abs().topk(min(4, flat.numel())).indices,
                         torch.randint(0, flat.numel(), (3,), generator=g)])
        for j in idx.tolist():
            plus = {kk: vv.clone() for kk, vv in inp.items()}
            minus = {kk: vv.clone() for kk, vv in inp.items()}
            plus[k].view(-1)[j] += eps
            minus[k].view(-1)[j] -= eps
            with torch.no_grad():
                fd = (loss_of(plus) - loss_of(minus)).item() / (2 * eps)
            an = flat[j].item()
            assert abs(fd - an) <= 1e-4 * max(1.0, flat.abs().max().item()), (k, j, fd, an)


# ---------------------------------------------------------------------------
# the FP32 restatement's backward (ref_blend_bwd, ref_preprocess_bwd)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("flags,degree,bg", [(0, 3, 0.0), (1, 1, 1.0), (3, 3, 0.0), (8, 3, 0.0),
                                             (11, 2, 1.0)])
def test_ref32_backward_matches_fp64_autograd(flags, degree, bg):
    """ref32.render_step's gradients (FP32 chain rule of the restatement) vs
    autograd of the FP64 transcription at the restatement's forward point
    (uv and conic snapped): every splat and camera gradient within 1e-5 of
    its largest entry.  flags: 1 depth, 2 normal, 8 plain colour."""
    W, H = 64, 56
    scene = S.random_scene(2000, 11 + flags, sh_degree=3, means="normal",
                           means_std=(0.6, 0.6, 0.6), log_scale=(-3.0, 0.5), opacity_std=1.5)
    cam = S.make_camera(W, H, eye=(0.4, -0.3, -2.5), fov_deg=55.0)
    inp = {k: v.detach() for k, v in scene.render_inputs().items()}
    names = [("rgb", 3)] + ([("depth", 1)] if flags & 1 else []) + \
        ([("normal", 3)] if flags & 2 else [])
    g = torch.Generator().manual_seed(3)
    w = {k: torch.randn(c, H, W, generator=g, dtype=F64) for k, c in names}
    ref = r32.render_step({k: v.numpy() for k, v in inp.items()},
                          (W, H, cam.intrinsic_params.numpy(), cam.extrinsic_matrix.numpy(),
                           cam.camera_center.numpy()), degree=degree, flags=flags, bg=bg,
                          keep=True,
                          image_grad=lambda p: (0.0, {k: w[k].numpy() for k in w}))
    hl = {k: v.double().requires_grad_(True) for k, v in inp.items()}
    E = cam.extrinsic_matrix.double().requires_grad_(True)
    K = cam.intrinsic_params.double().requires_grad_(True)
    c = cam.camera_center.double().requires_grad_(True)
    normals = o.splat_normals(hl["position"], hl["scaling"], hl["rotation"], c, E) \
        if flags & 2 else None
    snap = {"uv": torch.from_numpy(ref["uv"]), "conic": torch.from_numpy(ref["conic"])}
    res = o.render_iter(H, W, E, K, c, hl["position"], hl["opacity"], hl["scaling"],
                        hl["rotation"], hl["shs"], sh_degree=degree, bg=bg,
                        render_depth=bool(flags & 1), normals=normals,
                        add_half=not flags & 8, snap=snap)
    f = res["rendered_features_split"]
    for k in f:
        assert np.abs(ref["planes"][k] - f[k].detach().numpy()).max() < 1e-5, k
    sum((f[k] * w[k]).sum() for k in w).backward()
    pairs = [("position", hl["position"].grad), ("scaling", hl["scaling"].grad),
             ("rotation", hl["rotation"].grad), ("shs", hl["shs"].grad),
             ("opacity", hl["opacity"].grad.reshape(-1)), ("intr", K.grad),
             ("extr", E.grad[:3]), ("center", c.grad)]
    for k, want in pairs:
        got = np.asarray(ref["grads"][k], np.float64)
        e = np.abs(got - want.numpy()).max() / max(np.abs(want.numpy()).max(), 1e-30)
        assert e < 1e-5, (k, e)


def test_ref32_sort_matches_brute_force_at_size():
    """The restatement's bucket sort against the definition, at size: every
    (splat, tile) pair of the FP32 tile rectangles as a 64-bit key tile << 32
    | f32 depth bits, in splat order, stably sorted (numpy) -- 50k random
    splats with ties and off-screen rectangles."""
    g = torch.Generator().manual_seed(12)
    n, W, H = 50_000, 700, 500
    uv = ((torch.rand(n, 2, generator=g) * 1.4 - 0.2) * torch.tensor([W, H])).numpy()
    r = torch.randint(0, 60, (n,), generator=g, dtype=torch.int32).numpy()
    depth = torch.round(torch.rand(n, 1, generator=g) * 40 + 1).numpy()
    gx, gy = (W + 15) // 16, (H + 15) // 16
    u, v, rf = uv[:, 0].astype(np.float32), uv[:, 1].astype(np.float32), r.astype(np.float32)
    f16, f15 = np.float32(16.0), np.float32(15.0)
    x0 = np.minimum(np.maximum((u - rf) / f16, 0), gx).astype(np.int64)
    x1 = np.minimum(np.maximum(((u + rf) + f15) / f16, 0), gx).astype(np.int64)
    y0 = np.minimum(np.maximum((v - rf) / f16, 0), gy).astype(np.int64)
    y1 = np.minimum(np.maximum(((v + rf) + f15) / f16, 0), gy).astype(np.int64)
    bits = depth.reshape(-1).astype(np.float32).view(np.uint32).astype(np.uint64)
    keys, owner = [], []
    for i in np.nonzero(r > 0)[0]:
        ty, tx = np.meshgrid(np.arange(y0[i], y1[i]), np.arange(x0[i], x1[i]), indexing="ij")
        t = (ty * gx + tx).reshape(-1).astype(np.uint64)
        keys.append((t << np.uint64(32)) | bits[i])
        owner.append(np.full(t.shape, i, np.int32))
    keys, owner = np.concatenate(keys), np.concatenate(owner)
    perm = np.argsort(keys, kind="stable")
    ids32, rng32 = r32.sort_gaussian(uv, depth, W, H, r)
    assert np.array_equal(ids32, owner[perm])
    tiles = (keys[perm] >> np.uint64(32)).astype(np.int64)
    first = np.searchsorted(tiles, np.arange(gx * gy), side="left")
    last = np.searchsorted(tiles, np.arange(gx * gy), side="right")
    want = np.where((last > first)[:, None], np.stack([first, last], 1), 0)
    assert np.array_equal(rng32, want.astype(np.int32))
