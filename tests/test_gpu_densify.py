"""Fused densification statistics (pointrix_amd.densify) against the
restatement of controller/gs.py (oracle/densify_ref.py), on the side outputs
of real renders: one view (render_iter), several views (render_batch's list
of uv_points) and the batched [V,N,2] leaf; three consecutive steps."""
import pytest
import torch

from oracle import densify_ref as D
from pointrix_amd import scene as S
from pointrix_amd.batched import BatchedRender
from pointrix_amd.densify import DensifyStats
from pointrix_amd.renderer import MsplatRender

pytestmark = pytest.mark.gpu


def close(a, b):
    assert a.shape == b.shape
    err = (a.cpu() - b).abs().max().item()
    assert err <= 1e-6 * max(b.abs().max().item(), 1.0), err


def make(device, n=20000, seed=81, views=3):
    scene = S.random_scene(n, seed, sh_degree=3, means="normal", means_std=(0.8, 0.8, 0.8),
                           log_scale=(-3.3, 0.5), opacity_std=1.5).to(device, requires_grad=True)
    g = torch.Generator().manual_seed(seed)
    cams = [S.make_camera(128, 96, eye=S.random_sphere_eye(g, 2.5, 3.5), fov_deg=60.0).to(device)
            for _ in range(views)]
    return scene, cams


@pytest.mark.parametrize("normalize", [True, False])
def test_stats_single_and_multi_view(device, normalize):
    scene, cams = make(device)
    n = len(scene)
    r = MsplatRender({"max_sh_degree": 3}, device="cuda")
    r.sh_degree = 3
    fused = DensifyStats(n, device, 128, 96, normalize_grad=normalize)
    ga, st, mr = torch.zeros(n, 1), torch.zeros(n, 1), torch.zeros(n)
    for step in range(3):
        views = cams[: 1 + step % 3]
        outs = [r.render_iter(**c.render_kwargs(), **scene.render_inputs()) for c in views]
        loss = sum((o["rendered_features_split"]["rgb"] * (k + 1)).sum()
                   for k, o in enumerate(outs))
        loss.backward()
        uv = [o["uv_points"] for o in outs]
        vis = torch.stack([o["visibility"] for o in outs]).any(0)
        rad = torch.stack([o["radii"] for o in outs]).max(0).values
        fused.update(uv, vis, rad)
        D.update(ga, st, mr, [u.grad.cpu() for u in uv], vis.cpu(), rad.cpu(), 128, 96,
                 normalize_grad=normalize)
        for p in scene.params():
            p.grad = None
    torch.cuda.synchronize()
    close(fused.grad_accum, ga)
    assert torch.equal(fused.acc_steps.cpu(), st)
    assert torch.equal(fused.max_radii.cpu(), mr)
    assert (st > 0).any() and (st == 0).any()
    avg = fused.average_grad()
    assert not avg.isnan().any()


def test_stats_batched_leaf(device):
    scene, cams = make(device, seed=82, views=4)
    n = len(scene)
    br = BatchedRender(sh_degree=3)
    res = br.render([c.render_kwargs() for c in cams], **scene.render_inputs(split_sh=True))
    sum(d["rgb"].sum() for d in res["views"]).backward()
    fused = DensifyStats(n, device, 128, 96)
    fused.update(res["uv_points"], res["visibility"], res["radii"])
    ga, st, mr = torch.zeros(n, 1), torch.zeros(n, 1), torch.zeros(n)
    D.update(ga, st, mr, list(res["uv_points"].grad.cpu()), res["visibility"].cpu(),
             res["radii"].cpu(), 128, 96)
    torch.cuda.synchronize()
    close(fused.grad_accum, ga)
    assert torch.equal(fused.acc_steps.cpu(), st)
    assert torch.equal(fused.max_radii.cpu(), mr)
