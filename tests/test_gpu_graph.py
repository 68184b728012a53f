"""One-view training step captured as one HIP graph (pointrix_amd.graph).

* every replay equals an eager step (loss; gradients of the splats and of the
  camera tensors), with eager steps that replace ``.grad`` run between the
  replays;
* a capacity below the instance count is detected and the step re-captured,
  so the replayed result is still complete;
* a fixed-capacity sort returns the same ids and tile ranges as the exact one,
  and its high-water word keeps the largest instance count.
"""
import pytest
import torch

from pointrix_amd import ops
from pointrix_amd import scene as S
from pointrix_amd.graph import CapturedStep
from pointrix_amd.renderer import MsplatRender

pytestmark = pytest.mark.gpu


def setup(device, n=15000, seed=61, w=160, h=120):
    scene = S.random_scene(n, seed, sh_degree=3, means="normal", means_std=(0.7, 0.7, 0.7),
                           log_scale=(-3.3, 0.5), opacity_std=1.5).to(device, requires_grad=True)
    g = torch.Generator().manual_seed(seed)
    cam = S.make_camera(w, h, eye=S.random_sphere_eye(g, 2.5, 3.0), fov_deg=60.0).to(device)
    kw = {k: (v.clone().requires_grad_(True) if torch.is_tensor(v) else v)
          for k, v in cam.render_kwargs().items()}
    r = MsplatRender({"render_depth": True, "render_normal": True, "max_sh_degree": 3},
                     white_bg=False, device="cuda")
    r.sh_degree = 3
    tg = torch.Generator().manual_seed(seed + 1)
    targets = {"rgb": torch.rand(3, h, w, generator=tg).to(device),
               "depth": (torch.rand(1, h, w, generator=tg) * 4).to(device),
               "normal": torch.rand(3, h, w, generator=tg).to(device)}
    leaves = dict(scene.named_params())
    leaves.update({k: v for k, v in kw.items() if torch.is_tensor(v) and v.requires_grad})

    def loss_fn():
        res = r.render_iter(**kw, **scene.render_inputs(split_sh=True))
        f = res["rendered_features_split"]
        loss = ops.l1_loss(f["rgb"], targets["rgb"]) + \
            0.1 * ops.l1_loss(f["depth"], targets["depth"]) + \
            0.1 * ops.l1_loss(f["normal"], targets["normal"])
        loss.backward()
        return loss

    return r, leaves, loss_fn


def eager(r, leaves, loss_fn):
    cap = r.workspace.capacity
    r.workspace.capacity = None
    for p in leaves.values():
        p.grad = None
    loss = float(loss_fn())
    grads = {k: p.grad.clone() for k, p in leaves.items()}
    r.workspace.capacity = cap
    return loss, grads


def check(loss, grads, l0, g0, tag):
    assert abs(loss - l0) <= 1e-6 * abs(l0), (tag, loss, l0)
    for k in g0:
        e = (grads[k] - g0[k]).abs().max().item() / max(g0[k].abs().max().item(), 1e-30)
        assert e < 1e-5, (tag, k, e)


def test_captured_step_matches_eager(device):
    r, leaves, loss_fn = setup(device)
    l0, g0 = eager(r, leaves, loss_fn)
    step = CapturedStep(loss_fn, leaves, r.workspace).capture()
    assert r.workspace.capacity >= r.workspace.instances()
    for rep in range(3):
        assert step.replay()
        torch.cuda.synchronize()
        check(float(step.loss), {k: p.grad.clone() for k, p in leaves.items()}, l0, g0, rep)
        l1, g1 = eager(r, leaves, loss_fn)       # replaces every .grad
        check(l1, g1, l0, g0, f"eager {rep}")
    r.workspace.capacity = None


def test_captured_step_recaptures_on_overflow(device):
    r, leaves, loss_fn = setup(device, seed=62)
    l0, g0 = eager(r, leaves, loss_fn)
    # a capacity of half the instance count: the captured graph truncates
    step = CapturedStep(loss_fn, leaves, r.workspace, headroom=0.5, min_capacity=1).capture()
    m = r.workspace.instances()
    assert r.workspace.capacity < m
    step.headroom = 1.25
    assert not step.replay()                     # overflow seen, re-captured and replayed
    torch.cuda.synchronize()
    assert r.workspace.capacity >= m
    check(float(step.loss), {k: p.grad.clone() for k, p in leaves.items()}, l0, g0, "recap")
    assert step.replay()
    r.workspace.capacity = None


def test_fixed_capacity_sort_equals_exact(device):
    r, leaves, loss_fn = setup(device, seed=63)
    r.keep_intermediates = True
    r.workspace.capacity = None
    loss_fn()
    exact = {k: r.intermediates[k].clone() for k in ("ids", "tile_range")}
    m = r.workspace.instances()
    r.workspace.capacity = m + 777
    loss_fn()
    capped = r.intermediates
    assert capped["ids"].shape[0] == m + 777
    assert torch.equal(capped["tile_range"], exact["tile_range"])
    assert torch.equal(capped["ids"][:m], exact["ids"])
    r.workspace.capacity = None
    r.keep_intermediates = False


def test_high_water_keeps_largest_count(device):
    """Fixed-capacity sorts (gsr_sort_count_high): the workspace's high-water
    word holds the largest instance count since it was zeroed, stored by the
    count kernel -- a later, smaller count leaves it, zeroing restarts it."""
    ws = ops.SortWorkspace()
    cases = []
    for seed, n, r_max in ((71, 4000, 30), (72, 4000, 60), (73, 4000, 10)):
        g = torch.Generator().manual_seed(seed)
        uv = (torch.rand(n, 2, generator=g) * torch.tensor([320.0, 200.0])).to(device)
        depth = (torch.rand(n, 1, generator=g) * 10 + 0.5).to(device)
        radius = torch.randint(0, r_max, (n,), generator=g, dtype=torch.int32).to(device)
        ws.capacity = None
        ops.sort_gaussian(uv, depth, 320, 200, radius, None, workspace=ws)
        cases.append((uv, depth, radius, ws.instances()))
    counts = [c[3] for c in cases]
    assert counts[1] > counts[0] > counts[2] > 0
    ws.capacity = max(counts) + 100
    seen = []
    for uv, depth, radius, m in cases:
        ops.sort_gaussian(uv, depth, 320, 200, radius, None, workspace=ws)
        assert ws.instances() == m
        seen.append(int(ws.high_water_tensor().item()))
    assert seen == [counts[0], counts[1], counts[1]]
    ws.high_water_tensor().zero_()
    uv, depth, radius, m = cases[2]
    ops.sort_gaussian(uv, depth, 320, 200, radius, None, workspace=ws)
    assert int(ws.high_water_tensor().item()) == counts[2]
    ws.capacity = None
