"""Raw-parameter mode of the fused preprocess (activations inside the kernel).

The fused path with ``raw=True`` takes the point cloud's raw opacity logits,
log scales and unnormalised quaternions and applies sigmoid / exp /
F.normalize itself (gaussian_points.py:30-35,70-81).  It must render what the
drop-in path renders from torch-activated tensors (BaseModel.forward,
base_model.py:66-71), and its gradients w.r.t. the raw parameters must equal
torch's chain rule through those activations applied to the drop-in path's
gradients:
  images within 1e-5 (the activations are evaluated by different exp
  implementations, one ulp apart at most), gradients within
  max|g - g_ref| <= 1e-4 * max|g_ref| per parameter.
The drop-in path itself is pinned against the oracle in test_gpu_renderer.
"""
import pytest
import torch

from pointrix_amd import ops
from pointrix_amd import scene as S
from pointrix_amd.renderer import MsplatRender

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-5
GRAD_TOL = 1e-4


def rel_err(a, b):
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    den = b.abs().max().item()
    return (a - b).abs().max().item() / (den if den > 0 else 1.0)


def _setup(sh_degree, n, seed):
    scene = S.random_scene(n, seed, sh_degree=sh_degree, means="normal",
                           means_std=(0.6, 0.6, 0.6), log_scale=(-3.0, 0.5), opacity_std=1.5)
    # unnormalised quaternions with spread-out norms
    g = torch.Generator().manual_seed(seed + 1)
    scene.rotation = scene.rotation * (0.3 + 2.0 * torch.rand(n, 1, generator=g))
    cam = S.make_camera(96, 80, eye=(0.4, -0.3, -2.5), fov_deg=55.0)
    return scene, cam


def _run(scene, cam, cfg, raw, split_sh, seed, sh_degree):
    dev = "cuda:0"
    sc = scene.to(dev, requires_grad=True)
    r = MsplatRender(cfg)
    r.sh_degree = sh_degree
    inp = sc.render_inputs(split_sh=split_sh, raw=raw)
    out = r.render_iter(**cam.to(dev).render_kwargs(), **inp)
    imgs = out["rendered_features_split"]
    gen = torch.Generator(device=dev).manual_seed(seed)
    loss = sum((img * torch.randn(img.shape, generator=gen, device=dev)).sum()
               for img in imgs.values())
    loss.backward()
    return ({k: v.detach() for k, v in imgs.items()},
            {k: (None if p is None or p.grad is None else p.grad.detach())
             for k, p in sc.named_params().items()},
            out["uv_points"].grad.detach(), out["radii"].detach())


@pytest.mark.parametrize("case", ["sh0", "sh3_depth_normal_split"])
def test_raw_matches_torch_activations(case):
    if case == "sh0":
        sh, cfg, split = 0, {}, False
    else:
        sh, cfg, split = 3, {"render_depth": True, "render_normal": True,
                             "render_opacity": True}, True
    scene, cam = _setup(sh, 3000, 5)
    print(f"[raw] case={case}", flush=True)
    ref_img, ref_g, ref_ndc, ref_r = _run(scene, cam, cfg, False, split, 3, sh)
    img, g, ndc, r = _run(scene, cam, cfg, True, split, 3, sh)
    assert torch.equal(r, ref_r)
    assert int((r > 0).sum()) > 500
    for k in ref_img:
        e = (img[k] - ref_img[k]).abs().max().item()
        print(f"  image {k}: max abs {e:.3e}")
        assert e <= IMG_TOL, k
    for k, gr in ref_g.items():
        if gr is None:
            continue
        e = rel_err(g[k], gr)
        print(f"  grad {k}: rel {e:.3e}")
        assert e <= GRAD_TOL, k
    assert rel_err(ndc, ref_ndc) <= GRAD_TOL


def test_raw_render_batch_and_unfused():
    """render_batch forwards ``raw`` to every view; the unfused path applies
    the activations with torch and renders the same images."""
    scene, cam = _setup(1, 1500, 9)
    dev = "cuda:0"
    sc = scene.to(dev)
    cams = [cam, S.make_camera(96, 80, eye=(-0.5, 0.2, -2.4), fov_deg=55.0)]
    rd = {"height": 80, "width": 96,
          "extrinsic_matrix": torch.stack([c.extrinsic_matrix for c in cams]).to(dev),
          "intrinsic_params": cams[0].intrinsic_params.to(dev),
          "camera_center": torch.stack([c.camera_center for c in cams]).to(dev)}
    outs = []
    for fused, raw in ((True, False), (True, True), (False, True)):
        r = MsplatRender({"fused": fused})
        r.sh_degree = 1
        with torch.no_grad():
            outs.append(r.render_batch({**rd, **sc.render_inputs(raw=raw)})["rgb"])
    assert (outs[1] - outs[0]).abs().max().item() <= IMG_TOL
    assert (outs[2] - outs[0]).abs().max().item() <= IMG_TOL


def test_raw_preprocess_opacity_gradient():
    """The activated opacity output and its gradient alone: sigmoid and
    s (1 - s) against torch."""
    dev = "cuda:0"
    scene, cam = _setup(0, 1000, 21)
    sc = scene.to(dev, requires_grad=True)
    kw = cam.to(dev).render_kwargs()
    out = ops.preprocess(sc.position, sc.scaling, sc.rotation, sc.features, kw["intrinsic_params"],
                         kw["extrinsic_matrix"], kw["camera_center"], cam.width, cam.height,
                         raw=True, opacity=sc.opacity)
    assert len(out) == 8
    o = out[7]
    ref = torch.sigmoid(sc.opacity.detach())
    assert (o - ref).abs().max().item() <= 1e-6
    w = torch.randn(o.shape, device=dev, generator=torch.Generator(device=dev).manual_seed(0))
    (o * w).sum().backward()
    g_ref = w * ref * (1 - ref)
    assert rel_err(sc.opacity.grad, g_ref) <= 1e-6
    with pytest.raises(ValueError):
        ops.preprocess(sc.position, sc.scaling, sc.rotation, sc.features, kw["intrinsic_params"],
                       kw["extrinsic_matrix"], kw["camera_center"], cam.width, cam.height,
                       opacity=sc.opacity)
