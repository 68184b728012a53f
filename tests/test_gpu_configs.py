"""The benchmark workloads at their full sizes, on the GPU, against the FP32 C
restatement (oracle/ref32.c) run on the host with the same inputs.

* train500k: 500,000 splats, SH degree 3, one 979x546 view (rgb).
* full3m: 3,000,000 splats, one 1237x822 view, rgb + depth + normal,
  gradients into the splats and the camera (extrinsics, intrinsics, centre).
* batched1m: 1,000,000 splats, 16 views of uneven resolution in one batched
  render, each view checked on its own.
* serve2m: 2,000,000 resident splats, a request mix at 540p / 720p / 1080p /
  2160p served batched and one view at a time.

Forward (the contract of msplat.py:133-151): uv, depth, conic, radius, the
sorted ids, tile ranges, per-pixel contributor counts, T_final and the images
equal the restatement bit for bit.  Backward: every gradient within
max|g - g_ref| <= 1e-4 max|g_ref| of the restatement's FP32 chain rule
(float sums in another order; the restatement's backward is pinned against
FP64 autograd in tests/test_oracle.py).  The restatement is multi-threaded
(OpenMP) so the whole module runs in about a minute.
"""
import numpy as np
import pytest
import torch

from oracle import ref32 as r32
from pointrix_amd import scene as S
from pointrix_amd.batched import BatchedRender
from pointrix_amd.renderer import MsplatRender
from pointrix_amd.serve import BatchServer, PlainServer, make_requests

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-4


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def np_inputs(scene):
    inp = scene.render_inputs()
    return {k: v.detach().numpy() for k, v in inp.items()}


def cam_tuple(cam):
    return (cam.width, cam.height, cam.intrinsic_params.numpy(), cam.extrinsic_matrix.numpy(),
            cam.camera_center.numpy())


def weights_for(names, W, H, seed):
    """Seeded image weights {plane: [c,H,W]} for names [(plane, c)]."""
    g = np.random.default_rng(seed)
    return {k: g.standard_normal((c, H, W)).astype(np.float32) for k, c in names}


RGB = [("rgb", 3)]
RGB_D_N = [("rgb", 3), ("depth", 1), ("normal", 3)]


def gpu_render(device, inp, cam, depth, normal, weights=None, camera_grad=False):
    """MsplatRender.render_iter with the intermediates kept; with weights,
    backward of sum_k w_k * plane_k.  Returns (renderer, output planes,
    splat leaves, camera leaves)."""
    r = MsplatRender({"render_depth": depth, "render_normal": normal, "max_sh_degree": 3},
                     device="cuda")
    r.sh_degree = 3
    r.keep_intermediates = True
    leaves = {k: torch.from_numpy(v).to(device).requires_grad_(weights is not None)
              for k, v in inp.items()}
    ck = {"extrinsic_matrix": cam.extrinsic_matrix, "intrinsic_params": cam.intrinsic_params,
          "camera_center": cam.camera_center}
    cl = {k: v.to(device).requires_grad_(camera_grad) for k, v in ck.items()}
    res = r.render_iter(cam.height, cam.width, **cl, **leaves)
    planes = res["rendered_features_split"]
    if weights is not None:
        sum((planes[k] * torch.from_numpy(weights[k]).to(device)).sum() for k in planes).backward()
    return r, planes, leaves, cl


def assert_forward_equal(im, ref, tag):
    for k in ("uv", "depth", "conic", "radius", "ids", "tile_range", "n_contrib", "T_final",
              "rendered"):
        a = im[k].detach().cpu().numpy() if torch.is_tensor(im[k]) else im[k]
        assert np.array_equal(a.reshape(ref[k].shape), ref[k]), (tag, k)


def test_train500k_full_size(device):
    """979x546, 500k splats, SH 3: forward bit-exact, gradients vs the
    restatement's backward."""
    scene, cams = S.train500k(0)
    cam = cams[0]
    inp = np_inputs(scene)
    w = weights_for(RGB, cam.width, cam.height, 1)
    ref = r32.render_step(inp, cam_tuple(cam), degree=3, flags=0, keep=True,
                          image_grad=lambda p: (0.0, w))
    r, planes, leaves, _ = gpu_render(device, inp, cam, False, False, weights=w)
    im = r.intermediates
    print(f"train500k: M = {len(ref['ids'])}, visible = {(ref['radius'] > 0).sum()}")
    assert_forward_equal(im, ref, "train500k")
    assert np.array_equal(planes["rgb"].detach().cpu().numpy(), ref["planes"]["rgb"])
    for k in ("position", "scaling", "rotation", "opacity", "shs"):
        e = rel_err(leaves[k].grad.cpu().numpy().reshape(ref["grads"][k].shape), ref["grads"][k])
        print(f"train500k d{k}: rel err {e:.3g}")
        assert e < GRAD_TOL, k


def test_full3m_full_size(device):
    """1237x822, 3M splats, rgb + depth + normal, camera gradients."""
    scene, cams = S.full3m(0)
    cam = cams[0]
    inp = np_inputs(scene)
    w = weights_for(RGB_D_N, cam.width, cam.height, 2)
    ref = r32.render_step(inp, cam_tuple(cam), degree=3, flags=3, keep=True,
                          image_grad=lambda p: (0.0, w))
    r, planes, leaves, cl = gpu_render(device, inp, cam, True, True, weights=w, camera_grad=True)
    im = r.intermediates
    print(f"full3m: M = {len(ref['ids'])}, visible = {(ref['radius'] > 0).sum()}")
    assert len(ref["ids"]) > 20_000_000
    assert_forward_equal(im, ref, "full3m")
    for k in ("rgb", "depth", "normal"):
        assert np.array_equal(planes[k].detach().cpu().numpy(), ref["planes"][k]), k
    for k in ("position", "scaling", "rotation", "opacity", "shs"):
        e = rel_err(leaves[k].grad.cpu().numpy().reshape(ref["grads"][k].shape), ref["grads"][k])
        print(f"full3m d{k}: rel err {e:.3g}")
        assert e < GRAD_TOL, k
    cam_pairs = (("intrinsic_params", "intr", lambda g: g.reshape(-1)[:4]),
                 ("extrinsic_matrix", "extr", lambda g: g[:3]),
                 ("camera_center", "center", lambda g: g.reshape(-1)))
    for k, rk, view in cam_pairs:
        e = rel_err(view(cl[k].grad.cpu().numpy()), ref["grads"][rk])
        print(f"full3m d{k}: rel err {e:.3g}")
        assert e < GRAD_TOL, k


def test_batched1m_full_size(device):
    """16 views of uneven resolution over 1M splats in one batched render:
    each view's binning, contributor counts, T_final and image equal the
    restatement run on that view alone; the splat gradients of the summed
    loss match the sum of the restatement's per-view backward."""
    scene, cams = S.batched1m(0)
    inp = np_inputs(scene)
    n = len(scene)
    refs, ws = [], []
    for i, cam in enumerate(cams):
        w = weights_for(RGB, cam.width, cam.height, 100 + i)
        refs.append(r32.render_step(inp, cam_tuple(cam), degree=3, flags=0, keep=True,
                                    image_grad=lambda p, w=w: (0.0, w)))
        ws.append(w)
    br = BatchedRender(sh_degree=3, tile=(16, 16), batch_depth=256)
    br.keep_intermediates = True
    leaves = {k: torch.from_numpy(v).to(device).requires_grad_(True) for k, v in inp.items()}
    cds = [c.to(device).render_kwargs() for c in cams]
    res = br.render(cds, **leaves)
    lay = res["layout"]
    sum((d["rgb"] * torch.from_numpy(ws[v]["rgb"]).to(device)).sum()
        for v, d in enumerate(res["views"])).backward()
    im = {k: v.detach().cpu().numpy() for k, v in br.intermediates.items()}
    total = 0
    for v, ref in enumerate(refs):
        sl = slice(v * n, (v + 1) * n)
        W, H = lay.widths[v], lay.heights[v]
        assert np.array_equal(im["uv"][sl], ref["uv"]), v
        assert np.array_equal(im["radius"][sl], ref["radius"]), v
        assert np.array_equal(im["conic"][sl], ref["conic"]), v
        t0 = lay.tile_off[v]
        grng = im["tile_range"][t0:t0 + len(ref["tile_range"])]
        nz = grng[:, 1] > grng[:, 0]
        base = grng[nz, 0].min() if nz.any() else 0
        assert np.array_equal(np.where(nz[:, None], grng - base, 0), ref["tile_range"]), v
        m = len(ref["ids"])
        total += m
        assert np.array_equal(im["ids"][base:base + m] - v * n, ref["ids"]), v
        p0 = lay.pix_off[v]
        assert np.array_equal(im["n_contrib"][p0:p0 + W * H].reshape(H, W), ref["n_contrib"]), v
        assert np.array_equal(im["T_final"][p0:p0 + W * H].reshape(H, W), ref["T_final"]), v
        img = res["views"][v]["rgb"].detach().cpu().numpy()
        assert np.array_equal(img, ref["planes"]["rgb"]), v
    print(f"batched1m: {len(cams)} views, {total} instances")
    for k in ("position", "scaling", "rotation", "opacity", "shs"):
        want = sum(ref["grads"][k] for ref in refs)
        e = rel_err(leaves[k].grad.cpu().numpy().reshape(want.shape), want)
        print(f"batched1m d{k}: rel err {e:.3g}")
        assert e < GRAD_TOL, k


def test_serve2m_mix_batched_equals_plain(device):
    """2M resident splats; 12 requests at the serving resolution mix (one
    2160p among them), 20% with depth: the batched server's images equal the
    one-view-at-a-time server's bit for bit, and the 2160p view equals the
    restatement."""
    scene, _ = S.serve2m(0)
    sd = scene.to(device)
    with torch.no_grad():
        inputs = sd.render_inputs(split_sh=True)
    reqs = make_requests(200.0, 12, seed=7)
    big = make_requests(200.0, 1, seed=8, resolutions=[((3840, 2160), 1.0)])[0]
    big.rid = len(reqs)
    reqs.append(big)
    got_b = BatchServer(inputs, key_budget_bytes=16 << 30).render(reqs)
    got_p = PlainServer(inputs).render(reqs)
    sizes = sorted({(r.width, r.height) for r in reqs})
    print(f"serve2m: {len(reqs)} requests, sizes {sizes}")
    for r, a, b in zip(reqs, got_b, got_p):
        assert set(a) == set(b) == ({"rgb", "depth"} if r.depth else {"rgb"})
        for k in a:
            assert torch.equal(a[k], b[k]), (r.rid, k)
    # the server's inputs were activated on the GPU (ops.activate): the
    # restatement takes those same tensors
    inp = {k: v.detach().cpu().numpy() for k, v in inputs.items() if k != "shs"}
    inp["shs"] = torch.cat([t.detach().cpu() for t in inputs["shs"]], 1).numpy()
    cam = (big.width, big.height, big.intrinsic_params.numpy(), big.extrinsic_matrix.numpy(),
           big.camera_center.numpy())
    ref = r32.render_step(inp, cam, degree=3, flags=1 if big.depth else 0)
    assert np.array_equal(got_b[-1]["rgb"].cpu().numpy(), ref["planes"]["rgb"])
