"""Opacity-aware ("tight") binning on the GPU.

* the tight sort equals the FP32 C restatement's bit for bit (one view);
* rendering with it gives the same images as the exact tile contract, bit
  for bit, and the same gradients up to float-atomic ordering -- one view
  (MsplatRender tight_tiles) and many views (BatchedRender, every tile shape).
"""
import numpy as np
import pytest
import torch

from oracle import ref32 as r32
from pointrix_amd import ops
from pointrix_amd import scene as S
from pointrix_amd.batched import BatchedRender
from pointrix_amd.renderer import MsplatRender

pytestmark = pytest.mark.gpu


def scene_cam(device, n, seed, w, h, opacity_std=2.0):
    scene = S.random_scene(n, seed, sh_degree=3, means="normal", means_std=(1.5, 1.0, 1.5),
                           log_scale=(-3.8, 0.8),
                           opacity_std=opacity_std).to(device, requires_grad=True)
    g = torch.Generator().manual_seed(seed)
    cam = S.make_camera(w, h, eye=S.random_sphere_eye(g, 2.5, 4.0), focal=1.1 * w).to(device)
    return scene, cam


@pytest.mark.parametrize("n,w,h,seed", [(30000, 320, 240, 1), (120000, 979, 546, 2)])
def test_tight_sort_bitexact(device, n, w, h, seed):
    scene, cam = scene_cam(device, n, seed, w, h)
    r = MsplatRender({"max_sh_degree": 3}, device="cuda")
    r.sh_degree = 3
    r.keep_intermediates = True
    with torch.no_grad():
        r.render_iter(**cam.render_kwargs(), **scene.render_inputs())
    it = r.intermediates
    op = scene.render_inputs()["opacity"].detach()
    ids, rng = ops.sort_gaussian(it["uv"], it["depth"], w, h, it["radius"], None,
                                 conic=it["conic"], opacity=op)
    ids_r, rng_r = r32.sort_gaussian_tight(it["uv"].cpu().numpy(), it["depth"].cpu().numpy(), w,
                                           h, it["radius"].cpu().numpy(),
                                           it["conic"].cpu().numpy(), op.cpu().numpy())
    assert np.array_equal(ids.cpu().numpy(), ids_r)
    assert np.array_equal(rng.cpu().numpy(), rng_r)
    assert ids_r.shape[0] < it["ids"].shape[0]


def rel(ga, gb):
    return {k: (ga[k] - gb[k]).abs().max().item() / max(gb[k].abs().max().item(), 1e-30)
            for k in gb}


def grads_close(g_tight, g_exact, *g_more):
    """The two binnings add the same per-(wave, splat) partial sums into the
    per-splat records with float atomics, in another order; repeated exact
    runs differ the same way.  Bounds, per parameter: the L2 norm of the
    difference within 1e-5 of the gradient's norm, and the largest element
    difference within the north-star 1e-4 of max |g| -- each or a multiple
    (4 for L2, 8 for the largest element) of the largest exact-vs-exact spread
    over the repeated runs, where summation cancellation makes the spread
    larger (the scale gradients of a few strongly anisotropic splats sum many
    terms of both signs and move by percents between identical runs)."""
    def l2(a, b):
        return {k: ((a[k] - b[k]).norm() / b[k].norm().clamp(min=1e-30)).item() for k in b}
    e, e2 = rel(g_tight, g_exact), l2(g_tight, g_exact)
    spread = {k: 0.0 for k in e}
    spread2 = {k: 0.0 for k in e}
    for g in g_more:
        for k, v in rel(g, g_exact).items():
            spread[k] = max(spread[k], v)
        for k, v in l2(g, g_exact).items():
            spread2[k] = max(spread2[k], v)
    for k in e:
        assert e2[k] < max(1e-5, 4.0 * spread2[k]), (k, e2[k], spread2[k])
        assert e[k] < max(1e-4, 8.0 * spread[k]), (k, e[k], spread[k])


def test_tight_single_view_same_images(device):
    scene, cam = scene_cam(device, 20000, 3, 640, 360, opacity_std=1.0)
    out = {}
    for key, tight in (("exact", False), ("tight", True), ("exact2", False), ("exact3", False),
                       ("exact4", False)):
        r = MsplatRender({"render_depth": True, "render_normal": True, "max_sh_degree": 3,
                          "tight_tiles": tight}, device="cuda")
        r.sh_degree = 3
        for p in scene.params():
            p.grad = None
        res = r.render_iter(**cam.render_kwargs(), **scene.render_inputs(split_sh=True))
        f = res["rendered_features_split"]
        g = torch.Generator().manual_seed(9)
        loss = sum((v * torch.randn(v.shape, generator=g).to(device)).sum() for v in f.values())
        loss.backward()
        out[key] = ({k: v.detach().clone() for k, v in f.items()},
                    {k: p.grad.clone() for k, p in scene.named_params().items()},
                    ops.last_instance_count)
    for k in out["exact"][0]:
        assert torch.equal(out["exact"][0][k], out["tight"][0][k]), k
    grads_close(out["tight"][1], out["exact"][1], out["exact2"][1], out["exact3"][1],
                out["exact4"][1])
    assert out["tight"][2] < out["exact"][2]


@pytest.mark.parametrize("tile", [(16, 16), (8, 32), (32, 8)])
def test_tight_batched_same_images(device, tile):
    scene, _ = scene_cam(device, 40000, 4, 64, 64)
    g = torch.Generator().manual_seed(4)
    cams = []
    for w, h in [(320, 180), (256, 256), (480, 270), (97, 61)]:
        eye = S.random_sphere_eye(g, 2.5, 4.0)
        cams.append(S.make_camera(w, h, eye=eye, focal=1.1 * w).to(device).render_kwargs())
    res = {}
    for tight in (False, True, "exact2", "exact3"):
        r = BatchedRender(render_depth=True, sh_degree=3, tile=tile)
        r.tight_tiles = tight is True
        for p in scene.params():
            p.grad = None
        out = r.render(cams, **scene.render_inputs(split_sh=True))
        loss = sum((d["rgb"] * 0.7 - d["depth"] * 0.1).sum() for d in out["views"])
        loss.backward()
        res[tight] = (out["packed"].detach().clone(),
                      {k: p.grad.clone() for k, p in scene.named_params().items()},
                      r.instances())
    assert torch.equal(res[False][0], res[True][0])
    grads_close(res[True][1], res[False][1], res["exact2"][1], res["exact3"][1])
    assert res[True][2] < res[False][2]
