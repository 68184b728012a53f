"""Camera pose and output-plane kernels against their torch formulations.

camera_pose: pointrix's CameraModel composition (camera_model.py:111-175 with
unitquat_to_rotmat, utils/pose.py:40-84) -- E = [R(q/|q|) | t; 0 0 0 1],
centre = -R^T t -- forward within 1e-6, gradients for q and t against FP64
autograd of the same formulation within 1e-5 relative.

output_planes: the renderer's split of the blended image with the normal
group mapped to (n/|n| + 1)/2 (examples/supervise/renderer.py:89-96) --
forward within 1e-6 of the FP32 torch formulation, image gradient against
FP64 autograd within 1e-5 relative, with some groups' gradients absent.
"""
import pytest
import torch

from oracle.torch64 import quat_to_rotmat
from pointrix_amd import ops

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    den = b.abs().max().item()
    return (a - b).abs().max().item() / (den if den > 0 else 1.0)


def pose_ref(q, t):
    """The oracle's rotation (pinned against scipy in test_oracle.py) composed
    as CameraModel does."""
    R = quat_to_rotmat(torch.nn.functional.normalize(q, dim=-1).reshape(-1, 4))
    R = R.reshape(*q.shape[:-1], 3, 3)
    Rt = torch.cat([R, t[..., None]], -1)
    bottom = torch.zeros(*q.shape[:-1], 1, 4, dtype=q.dtype)
    bottom[..., 0, 3] = 1.0
    E = torch.cat([Rt, bottom], -2)
    center = torch.matmul(-R.transpose(-1, -2), t[..., None])[..., 0]
    return E, center


@pytest.mark.parametrize("C", [1, 5])
def test_camera_pose(C):
    g = torch.Generator().manual_seed(C)
    q = torch.randn(C, 4, generator=g) * (0.5 + torch.rand(C, 1, generator=g))
    t = torch.randn(C, 3, generator=g) * 3
    dE = torch.randn(C, 4, 4, generator=g)
    dc = torch.randn(C, 3, generator=g)
    qd, td = q.cuda().requires_grad_(True), t.cuda().requires_grad_(True)
    E, c = ops.camera_pose(qd, td)
    q64, t64 = q.double().requires_grad_(True), t.double().requires_grad_(True)
    E64, c64 = pose_ref(q64, t64)
    assert (E.cpu().double() - E64).abs().max().item() <= 1e-6
    assert (c.cpu().double() - c64).abs().max().item() <= 1e-5 * max(1.0, c64.abs().max().item())
    ((E * dE.cuda()).sum() + (c * dc.cuda()).sum()).backward()
    ((E64 * dE.double()).sum() + (c64 * dc.double()).sum()).backward()
    assert rel_err(qd.grad, q64.grad) <= 1e-5
    assert rel_err(td.grad, t64.grad) <= 1e-5
    # one camera as [4] / [3] (the shapes of a single pose)
    E1, c1 = ops.camera_pose(q[0].cuda(), t[0].cuda())
    assert E1.shape == (4, 4) and c1.shape == (3,)
    assert torch.equal(E1.cpu(), E[0].detach().cpu())


def test_camera_pose_only_centre_used():
    g = torch.Generator().manual_seed(9)
    q, t = torch.randn(4, generator=g), torch.randn(3, generator=g)
    qd, td = q.cuda().requires_grad_(True), t.cuda().requires_grad_(True)
    _, c = ops.camera_pose(qd, td)
    c.sum().backward()
    q64, t64 = q.double().requires_grad_(True), t.double().requires_grad_(True)
    pose_ref(q64[None], t64[None])[1].sum().backward()
    assert rel_err(qd.grad, q64.grad) <= 1e-5 and rel_err(td.grad, t64.grad) <= 1e-5


def planes_ref(x, sizes, nf):
    out, start = [], 0
    for c in sizes:
        p = x[start:start + c]
        if start == nf:
            p = (p / p.norm(dim=0, keepdim=True) + 1) / 2
        out.append(p)
        start += c
    return out


@pytest.mark.parametrize("sizes,nf,used", [([3, 1, 3], 4, [1, 1, 1]), ([3, 3, 1], 3, [1, 0, 1]),
                                           ([3, 1], -1, [1, 1]), ([3, 1, 3], 4, [0, 1, 0])])
def test_output_planes(sizes, nf, used):
    H, W = 37, 53
    g = torch.Generator().manual_seed(sum(sizes) + nf)
    x = torch.randn(sum(sizes), H, W, generator=g)
    xd = x.cuda().requires_grad_(True)
    outs = ops.output_planes(xd, sizes, nf)
    x64 = x.double().requires_grad_(True)
    ref = planes_ref(x64, sizes, nf)
    ref32 = planes_ref(x, sizes, nf)
    loss, loss64 = 0.0, 0.0
    for o, r, r32, c, u in zip(outs, ref, ref32, sizes, used):
        assert o.shape == (c, H, W)
        assert (o.detach().cpu() - r32).abs().max().item() <= 1e-6
        if u:
            w = torch.randn(c, H, W, generator=g)
            loss = loss + (o * w.cuda()).sum()
            loss64 = loss64 + (r * w.double()).sum()
    loss.backward()
    loss64.backward()
    assert rel_err(xd.grad, x64.grad) <= 1e-5


@pytest.mark.parametrize("n", [1, 1000, 70001])
def test_activate_matches_torch(n):
    """ops.activate against the point cloud's torch activations
    (gaussian_points.py:30-35,70-81), forward within 1e-6 relative and the
    gradients of each raw parameter within 1e-5 relative; includes one
    quaternion of zero norm (the clamped branch of F.normalize)."""
    g = torch.Generator().manual_seed(n)
    ls = torch.randn(n, 3, generator=g) - 2.0
    q = torch.randn(n, 4, generator=g) * (0.2 + torch.rand(n, 1, generator=g))
    lo = torch.randn(n, 1, generator=g) * 2
    if n > 1:
        q[n // 2] = 0.0
    gs, gq, go = (torch.randn(n, 3, generator=g), torch.randn(n, 4, generator=g),
                  torch.randn(n, 1, generator=g))
    dev = [t.cuda().requires_grad_(True) for t in (ls, q, lo)]
    s, r, o = ops.activate(*dev)
    ((s * gs.cuda()).sum() + (r * gq.cuda()).sum() + (o * go.cuda()).sum()).backward()
    ref = [t.clone().cuda().requires_grad_(True) for t in (ls, q, lo)]
    s2, r2, o2 = (torch.exp(ref[0]), torch.nn.functional.normalize(ref[1], dim=-1),
                  torch.sigmoid(ref[2]))
    ((s2 * gs.cuda()).sum() + (r2 * gq.cuda()).sum() + (o2 * go.cuda()).sum()).backward()
    for a, b in ((s, s2), (r, r2), (o, o2)):
        assert rel_err(a, b) <= 1e-6
    for a, b in zip(dev, ref):
        assert rel_err(a.grad, b.grad) <= 1e-5
