"""Benchmark: fwd+bwd steps/s of the Gaussian rasterisation path on one MI355X.

Workload (default ``full3m``): 3,000,000 splats (SH degree 3, anisotropic
means + 10% far shell), one 1237x822 camera, rendering rgb + depth + normal,
L1 losses against seeded random targets, backward into every Gaussian
parameter and into the camera's 6-DoF pose (quaternion + translation) and
intrinsics (fx, fy, cx, cy).  One step = one pass of MsplatRender.render_iter
+ loss + backward over that input, inputs resident in HBM.  The step is timed
eagerly and captured as one HIP graph; ``value`` is the faster of the two,
the graph only when every checked replay matches an eager step.

Prints one JSON line (see the keys below), plus per-kernel times measured
with device events on the launch stream, the HBM-roofline fraction of the
dominant kernel, and the FP64 PyTorch-CPU oracle timed on a bounded sample on
this host.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on the
# full3m workload, each scaled by the counter's reading for a copy of known
# size; summary and commands in PMC_SOURCE
PMC_SOURCE = "profiles/blend_full3m_pmc.md"
PMC_TRAFFIC_BYTES = {"alpha_blending_bwd": 64996 * 1024 * 2 + 46788 * 1024,
                     "alpha_blending_fwd": 22057 * 1024 * 2 + 51523 * 1024,
                     "preprocess_bwd": 296245 * 1024 * 2 + 680019 * 1024,
                     "preprocess_fwd": 255529 * 1024 * 2 + 175784 * 1024}
# the blend kernels' instruction counters (rocprofv3 --pmc SQ_*) on the
# full3m workload, from the same summary: VALU instructions per dispatch and
# their ratio to the dispatch's CU-cycles (256 CUs x GRBM_GUI_ACTIVE / 8 XCDs),
# against a bound of one wave64 VALU instruction per CU-cycle
VALU_PMC = {"alpha_blending_bwd": {"pmc_valu_insts": 1.449e8, "pmc_valu_issue_frac": 0.85},
            "alpha_blending_fwd": {"pmc_valu_insts": 8.298e7, "pmc_valu_issue_frac": 1.09}}


def rotmat_to_quat(R):
    """wxyz unit quaternion of a rotation matrix (Shepperd's method)."""
    R = R.double()
    tr = R.trace()
    if tr > 0:
        s = torch.sqrt(tr + 1.0) * 2
        q = [0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s]
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = torch.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2
        q = [(R[2, 1] - R[1, 2]) / s, 0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s]
    elif R[1, 1] > R[2, 2]:
        s = torch.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2
        q = [(R[0, 2] - R[2, 0]) / s, (R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s]
    else:
        s = torch.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2
        q = [(R[1, 0] - R[0, 1]) / s, (R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s]
    return torch.stack([torch.as_tensor(v) for v in q]).float()


class TrainableCamera:
    """Camera pose as (wxyz quaternion, translation) + intrinsics, all leaves,
    composed into the [4,4] extrinsic and the centre the way pointrix's
    CameraModel does (pointrix/model/camera/camera_model.py:111-175), by the
    fused pointrix_amd.ops.camera_pose."""

    def __init__(self, cam, dev):
        self.width, self.height = cam.width, cam.height
        self.q = rotmat_to_quat(cam.extrinsic_matrix[:3, :3]).to(dev).requires_grad_(True)
        self.t = cam.extrinsic_matrix[:3, 3].clone().to(dev).requires_grad_(True)
        self.intr = cam.intrinsic_params.clone().to(dev).requires_grad_(True)

    def params(self):
        return [self.q, self.t, self.intr]

    def render_kwargs(self):
        from pointrix_amd.ops import camera_pose
        E, center = camera_pose(self.q, self.t)
        return {"height": self.height, "width": self.width, "extrinsic_matrix": E,
                "intrinsic_params": self.intr, "camera_center": center}


def l1(a, b):
    from pointrix_amd.ops import l1_loss
    return l1_loss(a, b)


def l1_sum(terms):
    from pointrix_amd.ops import l1_loss_sum
    return l1_loss_sum(terms)


SETTLE_S = 2.0


def settle_steps(step, t_start, min_s=SETTLE_S):
    """Untimed steps until ``min_s`` seconds have passed since ``t_start``
    (synchronised every step); returns how many were run."""
    n = 0
    while time.perf_counter() - t_start < min_s:
        step()
        torch.cuda.synchronize()
        n += 1
    return n


def make_step(renderer, scene, cam, targets, raw=False):
    """-> (step, loss_fn): ``step(stats)`` = one eager fwd + loss + bwd with
    the gradients dropped afterwards; ``loss_fn()`` = fwd + loss + bwd leaving
    the gradients in place (what a captured step replays).  ``raw``: the
    renderer takes the raw parameters and applies the activations in its
    preprocess kernel."""
    params = scene.params() + cam.params()

    def loss_fn(stats=None):
        inp = scene.render_inputs(split_sh=True, raw=raw)
        res = renderer.render_iter(**cam.render_kwargs(), **inp)
        f = res["rendered_features_split"]
        # l1(rgb) + 0.1 l1(depth) + 0.1 l1(normal) as one fused reduction
        loss = l1_sum([(f[k], targets[k], 1.0 if k == "rgb" else 0.1)
                       for k in ("rgb", "depth", "normal") if k in f])
        loss.backward()
        if stats is not None:
            stats["visible"] = int(res["visibility"].sum().item())
        return loss

    def step(stats=None):
        loss = loss_fn(stats)
        for p in params:
            p.grad = None
        return loss

    return step, loss_fn


def graphed_headline(renderer, scene, cam, loss_fn, steps, warmup, log):
    """The same step captured as one HIP graph (pointrix_amd.graph): timed
    over ``steps`` replays, then three more replays each checked against an
    eager step (loss, and every gradient within four times the largest eager
    run-to-run spread of float-atomic accumulation over three eager steps, at
    least 1e-5 of the gradient's max)."""
    from pointrix_amd.graph import CapturedStep
    leaves = {f"p{i}": p for i, p in enumerate(scene.params() + cam.params())}
    params = list(leaves.values())

    def eager():
        for p in params:
            p.grad = None
        loss = float(loss_fn().detach())
        return loss, [p.grad.clone() for p in params]

    def rel(a, b):
        return max(float((x - y).abs().max() / y.abs().max().clamp(min=1e-30))
                   for x, y in zip(a, b))

    renderer.workspace.capacity = None
    loss_e, ref = eager()
    spread = max(rel(eager()[1], ref) for _ in range(3))
    tol = max(1e-5, 4 * spread)
    cs = CapturedStep(loss_fn, leaves, renderer.workspace).capture()
    for _ in range(warmup):
        cs.replay()
    torch.cuda.synchronize()
    # replays back to back (no host wait between them); the instance count is
    # checked once afterwards, and an overflow (the step then re-captured)
    # repeats the timing
    for _ in range(2):
        t0 = time.perf_counter()
        for _ in range(steps):
            cs.replay(check=False)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1000 / steps
        if not cs.overflowed():
            break
    ok, worst = True, 0.0
    for _ in range(3):
        cs.replay()
        torch.cuda.synchronize()
        dl = abs(float(cs.loss) - loss_e) / abs(loss_e)
        dg = rel([p.grad for p in params], ref)
        worst = max(worst, dg)
        if dl > 1e-6 or dg > tol:
            ok = False
            log(f"graph replay differs from eager: loss rel {dl:.3g}, grad rel {dg:.3g} "
                f"(eager spread {spread:.3g})")
    capacity = renderer.workspace.capacity
    renderer.workspace.capacity = None
    for p in params:
        p.grad = None
    log(f"graphed step {ms:.3f} ms (replays match eager: {ok}, grad rel {worst:.3g}, "
        f"eager spread {spread:.3g}, capacity {capacity})")
    return {"ms_per_step": round(ms, 4), "replays_match_eager": ok,
            "grad_rel": float(f"{worst:.3g}"), "eager_grad_spread": float(f"{spread:.3g}"),
            "instance_capacity": capacity, "recaptures": cs.captures - 1}


def algorithmic_bytes(N, K, C, P, T, Nv, M):
    """Compulsory HBM bytes per stage: every stage reads its inputs once and
    writes its outputs once; the sort is one read + one write of key + value."""
    return {
        "preprocess_fwd": N * (4 * (11 + 3 * K) + 44),
        "binning": 8 * N + 20 * Nv + 44 * M + 8 * T,
        "alpha_blending_fwd": 8 * T + 4 * M + 4 * Nv * (6 + C) + 4 * P * (C + 2),
        # reads as the forward; writes (du, dv), dopacity and the compact
        # [dconic, dfeature] record: 6 + C floats per visible splat
        "alpha_blending_bwd": 4 * P * (C + 2) + 8 * T + 4 * M + 4 * Nv * (6 + C) + 4 * Nv * (6 + C),
        "preprocess_bwd": N * (8 * (11 + 3 * K) + 80),
    }


# VALU issue ceiling: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz, one FP32
# operation per lane per cycle (packed FP32 is not generated)
VALU_PEAK_OPS = 256 * 64 * 2.4e9
# FP32 VALU operations per (pixel, splat) pair of the blend recurrences as the
# oracle writes them (oracle/ref32.c ref_blend / ref_blend_bwd): one per
# arithmetic operator, compare, min / max or select; exp_neg expanded into its
# 14 operations (clamp, scale, round, 2 + 7 fma, exponent convert and scale);
# a division counted once.
#   forward, blended pair: dx, dy 2; power 6; two cut tests 2; exp 14; alpha
#     (mul, min) 2; alpha test 1; T test (sub, mul, cmp) 3; weight 1; C
#     feature fma; state selects 2 -> 33 + C
#   backward, replayed pair: the forward's alpha 27; T recovery (sub, div) 2;
#     dfeature (C mul-add) and the colour-behind dot product (C fma) 2C; its
#     recurrence 4; dL/dalpha 4; clamp test 1; dpower 1; du, dv, da, db, dc,
#     dopacity record terms 24 -> 63 + 2C
#   rejected pair (either walk): dx, dy, power, one test -> 9
def blend_ops(C, blended, examined):
    rej = max(examined - blended, 0)
    return {"alpha_blending_fwd": (33 + C) * blended + 9 * rej,
            "alpha_blending_bwd": (63 + 2 * C) * blended + 9 * rej}


def copy_bandwidth(nbytes=2 << 30, reps=10):
    """Device-to-device copy of ``nbytes`` (torch copy_), timed with events on
    the current stream: bytes read + written per second (GB/s), median of
    ``reps`` after 3 warm-up copies."""
    dev = torch.device("cuda:0")
    src = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    times = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b) * 1e-3)
    times.sort()
    t = times[len(times) // 2]
    del src, dst
    torch.cuda.empty_cache()
    return round(2 * nbytes / t / 1e9, 1)


def _median(xs):
    xs = sorted(xs)
    return xs[len(xs) // 2]


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.abs(b).max() if b.size else 0.0
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0)) if a.size else 0.0


L1_WEIGHTS = {"rgb": 1.0, "depth": 0.1, "normal": 0.1}


def gpu_reference_step(renderer, scene, cam, targets):
    """One eager step of the benchmark's render + loss with the render_iter
    inputs (activated splat tensors, the composed extrinsic, intrinsics and
    centre) taken as leaves: -> numpy inputs, camera, planes, intermediates
    and gradients, for the comparison with the host restatement."""
    with torch.no_grad():
        inp = scene.render_inputs(split_sh=True)
        ck = cam.render_kwargs()
    leaves = {k: (v.detach().clone().requires_grad_(True) if torch.is_tensor(v) else
                  [t.detach().clone().requires_grad_(True) for t in v]) for k, v in inp.items()}
    cl = {k: ck[k].detach().clone().requires_grad_(True)
          for k in ("extrinsic_matrix", "intrinsic_params", "camera_center")}
    renderer.keep_intermediates = True
    res = renderer.render_iter(ck["height"], ck["width"], **cl, **leaves)
    f = res["rendered_features_split"]
    loss = l1_sum([(f[k], targets[k], L1_WEIGHTS[k]) for k in ("rgb", "depth", "normal")])
    loss.backward()
    im = {k: v.detach().cpu().numpy() for k, v in renderer.intermediates.items()}
    renderer.keep_intermediates = False
    renderer.intermediates = None
    host = {k: v.detach().cpu().numpy() for k, v in leaves.items() if k != "shs"}
    host["shs"] = torch.cat([t.detach().cpu() for t in leaves["shs"]], 1).numpy()
    grads = {k: leaves[k].grad.cpu().numpy() for k in ("position", "scaling", "rotation",
                                                        "opacity")}
    grads["shs"] = torch.cat([t.grad.cpu() for t in leaves["shs"]], 1).numpy()
    grads["extr"] = cl["extrinsic_matrix"].grad.cpu().numpy()[:3]
    grads["intr"] = cl["intrinsic_params"].grad.cpu().numpy().reshape(-1)[:4]
    grads["center"] = cl["camera_center"].grad.cpu().numpy().reshape(-1)
    camera = (ck["width"], ck["height"], cl["intrinsic_params"].detach().cpu().numpy(),
              cl["extrinsic_matrix"].detach().cpu().numpy(),
              cl["camera_center"].detach().cpu().numpy())
    planes = {k: v.detach().cpu().numpy() for k, v in f.items()}
    return host, camera, planes, im, grads, float(loss.detach())


def restatement_leg(renderer, scene, cam, targets, reps=5):
    """The whole step (render, the same weighted L1 losses, backward into the
    splats and the camera) by the FP32 C restatement (oracle/ref32.c, OpenMP on
    this host's cores) on the GPU step's own inputs: timed in full (median of
    ``reps`` after one warm-up) as the CPU baseline, compared with the GPU step
    (forward bit for bit, gradients relative to the largest), and its exact
    blend work counts (the restatement's walk is the kernels' walk)."""
    from oracle import ref32 as r32
    host, camera, planes, im, grads, _ = gpu_reference_step(renderer, scene, cam, targets)
    tg = {k: v.detach().cpu().numpy() for k, v in targets.items()}

    def image_grad(p):
        out, loss = {}, 0.0
        for k, x in p.items():
            w = np.float32(L1_WEIGHTS[k] / x.size)
            out[k] = (np.sign(x - tg[k]).astype(np.float32) * w).astype(np.float32)
            loss += L1_WEIGHTS[k] * float(np.abs(x - tg[k]).mean())
        return loss, out

    def run(keep=False):
        return r32.render_step(host, camera, degree=3, flags=3, keep=keep,
                               image_grad=image_grad)

    ref = run(keep=True)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    t = _median(times)
    W, H = camera[0], camera[1]
    exact = {k: bool(np.array_equal(im[k].reshape(np.shape(ref[k])), ref[k]))
             for k in ("uv", "depth", "conic", "radius", "ids", "tile_range", "n_contrib",
                       "T_final")}
    exact.update({f"image_{k}": bool(np.array_equal(planes[k], ref["planes"][k]))
                  for k in planes})
    gerr = {k: float(f"{_rel(grads[k].reshape(np.shape(ref['grads'][k])), ref['grads'][k]):.3g}")
            for k in ("position", "scaling", "rotation", "opacity", "shs", "extr", "intr",
                      "center")}
    examined, blended, prefix = r32.blend_pairs(ref["uv"], ref["conic"], host["opacity"],
                                                ref["ids"], ref["tile_range"], W, H)
    threads = r32.num_threads()
    cpu = {"value": round(1.0 / t, 4), "unit": "steps/s", "cores": threads, "kind": "port",
           "mpixel_per_s": round(W * H / 1e6 / t, 3),
           "sample": f"the full step at its full size: the FP32 C restatement of the path "
                     f"(oracle/ref32.c: preprocess, binning, blending, their backward, camera "
                     f"gradients) with the same weighted L1 losses, on the GPU step's own "
                     f"inputs; median of {reps} steps after one warm-up "
                     f"({', '.join(f'{x:.2f}' for x in times)} s); OpenMP threads={threads}, "
                     f"os.cpu_count()={os.cpu_count()}"}
    parity = {"reference": "FP32 C restatement at full size, same inputs",
              "bit_exact": exact, "grad_rel_err": gerr}
    pairs = {"examined": examined, "blended": blended, "n_contrib_sum": prefix}
    return cpu, parity, pairs


def fp64_cpu10k(reps=5):
    """The FP64 PyTorch-CPU transcription (oracle/torch64.py) of the 10k-splat
    256x256 SH-0 L1 step, timed in full: median of ``reps`` after one warm-up,
    torch threads as set on this host."""
    from oracle import torch64 as o64
    from pointrix_amd import scene as S
    scene, cams = S.cpu10k()
    cam = cams[0]
    target = torch.rand(3, cam.height, cam.width,
                        generator=torch.Generator().manual_seed(5)).double()
    times = []
    for i in range(reps + 1):
        inp = {k: v.detach().double().requires_grad_(True)
               for k, v in scene.render_inputs().items()}
        t0 = time.perf_counter()
        ref = o64.render_iter(cam.height, cam.width, cam.extrinsic_matrix.double(),
                              cam.intrinsic_params.double(), cam.camera_center.double(),
                              inp["position"], inp["opacity"], inp["scaling"], inp["rotation"],
                              inp["shs"], sh_degree=0)
        (ref["rendered_features_split"]["rgb"] - target).abs().mean().backward()
        if i:
            times.append(time.perf_counter() - t0)
    t = _median(times)
    return {"workload": "cpu10k: 10k splats SH0, 256x256, L1", "value": round(1.0 / t, 3),
            "unit": "steps/s", "mpixel_per_s": round(256 * 256 / 1e6 / t, 4),
            "threads": torch.get_num_threads(), "cpu_count": os.cpu_count(),
            "steps_s": [round(x, 3) for x in times]}


def _fp64_errors(inp, cam, flags, sh_degree, bg, dev, snapped):
    """GPU render_iter vs the FP64 transcription on the same leaves: max |image
    error| per output plane, max relative error per gradient (backward of a
    seeded random weighting of every plane)."""
    from oracle import torch64 as o64
    from pointrix_amd.renderer import MsplatRender
    r = MsplatRender({"render_depth": bool(flags & 1), "render_normal": bool(flags & 2),
                      "render_opacity": bool(flags & 4)}, white_bg=bool(bg), device="cuda")
    r.sh_degree = sh_degree
    r.keep_intermediates = True
    camk = {"extrinsic_matrix": cam.extrinsic_matrix, "intrinsic_params": cam.intrinsic_params,
            "camera_center": cam.camera_center}
    dl = {k: v.to(dev).requires_grad_(True) for k, v in {**inp, **camk}.items()}
    res = r.render_iter(cam.height, cam.width, **dl)
    imgs = res["rendered_features_split"]
    g = torch.Generator().manual_seed(1)
    wts = {k: torch.randn(v.shape, generator=g) for k, v in imgs.items()}
    sum((imgs[k] * wts[k].to(dev)).sum() for k in imgs).backward()
    snap = None
    if snapped:
        snap = {"uv": r.intermediates["uv"].detach().cpu(),
                "conic": r.intermediates["conic"].detach().cpu()}
    hl = {k: v.double().requires_grad_(True) for k, v in {**inp, **camk}.items()}
    normals = None
    if flags & 2:
        normals = o64.splat_normals(hl["position"], hl["scaling"], hl["rotation"],
                                    hl["camera_center"], hl["extrinsic_matrix"])
    ref = o64.render_iter(cam.height, cam.width, hl["extrinsic_matrix"], hl["intrinsic_params"],
                          hl["camera_center"], hl["position"], hl["opacity"], hl["scaling"],
                          hl["rotation"], hl["shs"], sh_degree=sh_degree, bg=bg,
                          render_depth=bool(flags & 1), normals=normals,
                          render_opacity=bool(flags & 4), snap=snap)
    rimgs = ref["rendered_features_split"]
    sum((rimgs[k] * wts[k].double()).sum() for k in rimgs).backward()
    err = {k: (imgs[k].detach().cpu().double() - rimgs[k].detach()).abs().amax(0) for k in imgs}
    img = {k: float(f"{e.max():.3g}") for k, e in err.items()}
    # pixels within 1e-5 (the FP32 rounding of pixel coordinates moves a splat
    # tail by ~1e-4 relative, and can flip an alpha >= 1/255 decision)
    within = {k: float(f"{(e <= 1e-5).double().mean():.6f}") for k, e in err.items()}
    names = {"extrinsic_matrix": "extr", "intrinsic_params": "intr", "camera_center": "center"}
    def g_of(t):  # a leaf the output does not depend on (the centre at SH 0) has none
        return np.zeros(tuple(t.shape)) if t.grad is None else t.grad.detach().cpu().double().numpy()

    grad = {names.get(k, k): float(f"{_rel(g_of(dl[k]), g_of(hl[k])):.3g}") for k in dl}
    return {"image_max_abs_err": img, "pixels_within_1e-5": within, "grad_rel_err": grad}


def crop2000(scene_cpu, cam, n=2000, size=64, seed=0):
    """2,000 splats of the scene whose footprint meets a size x size window at
    the centre of the camera's image, and the camera cropped to that window
    (principal point shifted): the FP64-checkable crop of the full step."""
    from oracle import ref32 as r32
    from pointrix_amd import scene as S
    inp = {k: v.detach() for k, v in scene_cpu.render_inputs().items()}
    K = cam.intrinsic_params.clone().reshape(-1)
    E = cam.extrinsic_matrix
    W, H = cam.width, cam.height
    uv, d = r32.project_point(inp["position"].numpy(), K.numpy(), E.numpy(), W, H, 0.2)
    vis = d != 0
    cov = r32.compute_cov3d(inp["scaling"].numpy(), inp["rotation"].numpy(), vis)
    _, rad, _ = r32.ewa_project(inp["position"].numpy(), cov, K.numpy(), E.numpy(), uv, W, H, vis)
    x0, y0 = (W - size) // 2, (H - size) // 2
    hit = ((rad > 0) & (uv[:, 0] + rad >= x0) & (uv[:, 0] - rad < x0 + size) &
           (uv[:, 1] + rad >= y0) & (uv[:, 1] - rad < y0 + size))
    idx = np.flatnonzero(hit)
    g = np.random.default_rng(seed)
    idx = np.sort(g.choice(idx, size=min(n, len(idx)), replace=False))
    sub = {k: v[torch.from_numpy(idx)].contiguous() for k, v in inp.items()}
    Kc = K.clone()
    Kc[2] -= x0
    Kc[3] -= y0
    crop = S.Camera(size, size, E.clone(), Kc.reshape(cam.intrinsic_params.shape),
                    cam.camera_center.clone())
    return sub, crop


def fp64_errors(scene_cpu, cam, dev):
    """Error of every output against the FP64 transcription: the 2,000-splat
    64x64 crop of the full step (rgb, depth, normal, opacity; SH 3; white
    background; gradients into the splats and the camera), unsnapped (both
    sides from the same FP32 leaves, which includes FP32 pixel-coordinate
    rounding) and snapped (FP64 evaluated at the GPU's uv and conic), and the
    10k-splat SH-0 256x256 workload unsnapped."""
    from pointrix_amd import scene as S
    sub, crop = crop2000(scene_cpu, cam)
    out = {"crop2000_64x64": {"unsnapped": _fp64_errors(sub, crop, 7, 3, 1.0, dev, False),
                              "snapped": _fp64_errors(sub, crop, 7, 3, 1.0, dev, True)}}
    sc, cams = S.cpu10k()
    inp = {k: v.detach() for k, v in sc.render_inputs().items()}
    out["cpu10k"] = {"unsnapped": _fp64_errors(inp, cams[0], 0, 0, 0.0, dev, False),
                     "snapped": _fp64_errors(inp, cams[0], 0, 0, 0.0, dev, True)}
    return out


def time_tight(renderer, step, steps):
    """The step with MsplatRender's tight_tiles on (opacity-aware binning)."""
    from pointrix_amd import ops
    renderer.cfg.tight_tiles = True
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000 / steps
    m = ops.last_instance_count
    renderer.cfg.tight_tiles = False
    print(f"[bench] tight tiles: {ms:.3f} ms/step, {m} instances", file=sys.stderr, flush=True)
    return {"ms_per_step": round(ms, 4), "instances": m}


def time_raw(renderer, scene, cam, targets, steps):
    """The step with the activations (sigmoid, exp, normalize) applied inside
    the preprocess kernel instead of by torch before render_iter."""
    step, _ = make_step(renderer, scene, cam, targets, raw=True)
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000 / steps
    print(f"[bench] raw-parameter step: {ms:.3f} ms/step", file=sys.stderr, flush=True)
    return {"ms_per_step": round(ms, 4)}


def bench_batched(steps, warmup, log, n=1_000_000, tune=None):
    """Batched workload: 1M splats (SH 3), 16 cameras of uneven resolution,
    summed L1 over views, fwd+bwd.  Plain = one view per launch
    (MsplatRender.render_iter per view, one backward); batched = one launch per
    stage (autotuned tile shape / batch depth); graph = the batched step
    captured as one HIP graph, every replay checked against eager."""
    from pointrix_amd import scene as S
    from pointrix_amd.autotune import autotune, time_step
    from pointrix_amd.batched import BatchedRender, GraphedStep
    from pointrix_amd.renderer import MsplatRender

    dev = torch.device("cuda:0")
    scene_cpu, cams = S.batched1m(0, n=n)
    scene = scene_cpu.to(dev, requires_grad=True)
    params = scene.params()
    cd = [c.to(dev).render_kwargs() for c in cams]
    g = torch.Generator().manual_seed(321)
    targets = [torch.rand(3, c.height, c.width, generator=g).to(dev) for c in cams]
    V = len(cams)

    def inputs():
        return scene.render_inputs(split_sh=True)

    def loss_b(res):
        return sum(l1(d["rgb"], t) for d, t in zip(res["views"], targets))

    plain = MsplatRender({"max_sh_degree": 3}, device="cuda")
    plain.sh_degree = 3

    def plain_step():
        inp = inputs()
        loss = sum(l1(plain.render_iter(**c, **inp)["rendered_features_split"]["rgb"], t)
                   for c, t in zip(cd, targets))
        loss.backward()
        for p in params:
            p.grad = None

    def timed(fn, n):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1000 / n

    ms_plain = timed(plain_step, steps)
    log(f"plain {ms_plain:.2f} ms / {V} views")
    if tune is None:
        (tile, bd), tune = autotune(cd, inputs, loss_b, params, len(scene), sh_degree=3, reps=2)
        log(f"autotune -> tile {tile[0]}x{tile[1]} depth {bd}")
    else:
        (tile, bd), tune = tune
    br = BatchedRender(sh_degree=3, tile=tile, batch_depth=bd)

    def batched_step():
        loss = loss_b(br.render(cd, **inputs()))
        loss.backward()
        for p in params:
            p.grad = None

    ms_batched = timed(batched_step, steps)
    log(f"batched {ms_batched:.2f} ms")
    # eager reference for the graph check.  Gradients are accumulated with
    # float atomics, so eager itself varies run to run in the last bits; the
    # replay tolerance is four times the largest spread measured over three
    # more eager steps (at least 1e-5).
    def eager_grads():
        loss = loss_b(br.render(cd, **inputs()))
        loss.backward()
        grads = [p.grad.clone() for p in params]
        for p in params:
            p.grad = None
        return float(loss.detach()), grads

    def grad_rel(a, b):
        return max(float((x - y).abs().max() / y.abs().max().clamp(min=1e-30))
                   for x, y in zip(a, b))

    loss_e, ref_grads = eager_grads()
    # the run-to-run spread of three more eager steps (the largest of them)
    noise = max(grad_rel(eager_grads()[1], ref_grads) for _ in range(3))
    tol = max(1e-5, 4 * noise)
    gs = GraphedStep(br, cd, scene.named_params(), inputs, loss_b).capture()
    graph_ok = True
    for _ in range(warmup):
        gs.replay()
    torch.cuda.synchronize()
    for _ in range(2):  # back to back, the count checked once afterwards
        t0 = time.perf_counter()
        for _ in range(steps):
            gs.replay(check=False)
        torch.cuda.synchronize()
        ms_graph = (time.perf_counter() - t0) * 1000 / steps
        if not gs.overflowed():
            break
    worst = 0.0
    for _ in range(3):   # every replay against eager
        gs.replay()
        torch.cuda.synchronize()
        dl = abs(float(gs.loss) - loss_e) / abs(loss_e)
        dg = grad_rel([p.grad for p in params], ref_grads)
        worst = max(worst, dg)
        if dl > 1e-5 or dg > tol:
            graph_ok = False
            log(f"graph replay differs from eager: loss rel {dl:.3g}, grad rel {dg:.3g} "
                f"(eager spread {noise:.3g})")
    log(f"graph {ms_graph:.2f} ms (replays match eager: {graph_ok})")
    best = min(ms_batched, ms_graph if graph_ok else float("inf"))
    return {"workload": f"batched: {n // 1_000_000}M splats SH3, 16 views of {{256x256, "
                        "640x360, 979x546, 1237x822, 1920x1080}, summed L1, fwd+bwd",
            "views": V, "pixels": sum(c.width * c.height for c in cams),
            "plain_ms": round(ms_plain, 3), "batched_ms": round(ms_batched, 3),
            "graph_ms": round(ms_graph, 3), "graph_replays_match_eager": graph_ok,
            "graph_grad_rel": float(f"{worst:.3g}"), "eager_grad_spread": float(f"{noise:.3g}"),
            "plain_views_per_s": round(V * 1000 / ms_plain, 1),
            "batched_views_per_s": round(V * 1000 / best, 1),
            "speedup_vs_plain": round(ms_plain / best, 3),
            "autotune": {"chosen": f"{tile[0]}x{tile[1]}/{bd}",
                         "ms": {f"{t[0]}x{t[1]}/{d}": round(v, 3) for (t, d), v in tune.items()}},
            "_tune": ((tile, bd), tune)}


def bench_train500k(steps, warmup, log):
    """Full training step: 500k splats (SH 3), one 979x546 view, loss
    0.8 L1 + 0.2 (1 - SSIM), densification statistics, Adam (pointrix's
    per-attribute learning rates, eps 1e-15): the full training step.
    fused = ops.l1_ssim_loss + DensifyStats + FusedAdam; fused_raw = the same
    with the activations applied inside the preprocess kernel; torch = the
    step with the reference's torch formulations (model/loss.py SSIM with
    conv2d, the controller's masked updates, torch.optim.Adam)."""
    import torch.nn.functional as F

    from pointrix_amd import scene as S
    from pointrix_amd.densify import DensifyStats
    from pointrix_amd.ops import l1_ssim_loss
    from pointrix_amd.optim import FusedAdam
    from pointrix_amd.renderer import MsplatRender

    dev = torch.device("cuda:0")
    scene_cpu, cams = S.train500k(0)
    cam = cams[0].to(dev)
    W, H = cam.width, cam.height
    lrs = {"position": 1.6e-4, "features": 2.5e-3, "features_rest": 1.25e-4, "scaling": 5e-3,
           "rotation": 1e-3, "opacity": 5e-2}
    target = torch.rand(3, H, W, generator=torch.Generator().manual_seed(5)).to(dev)
    # the unfused comparison: the reference's torch formulations, on the GPU
    # (model/loss.py:27,67-116 SSIM by conv2d; controller/gs.py:318-333)
    g1 = torch.tensor([math.exp(-(x - 5) ** 2 / 4.5) for x in range(11)])
    g1 = g1 / g1.sum()
    win = (g1[:, None] @ g1[None, :]).expand(3, 1, 11, 11).contiguous().to(dev)

    def torch_loss(x, y):
        x, y = x[None], y[None]
        mu1, mu2 = F.conv2d(x, win, padding=5, groups=3), F.conv2d(y, win, padding=5, groups=3)
        s11 = F.conv2d(x * x, win, padding=5, groups=3) - mu1 * mu1
        s22 = F.conv2d(y * y, win, padding=5, groups=3) - mu2 * mu2
        s12 = F.conv2d(x * y, win, padding=5, groups=3) - mu1 * mu2
        smap = ((2 * mu1 * mu2 + 1e-4) * (2 * s12 + 9e-4)) / \
            ((mu1 * mu1 + mu2 * mu2 + 1e-4) * (s11 + s22 + 9e-4))
        return 0.8 * (x - y).abs().mean() + 0.2 * (1.0 - smap.mean())

    def torch_stats(ga, st, mr, uv, vis, radii):
        g = uv.grad.clone()
        g[..., 0] *= W / 2.0
        g[..., 1] *= H / 2.0
        mr[vis] = torch.max(mr[vis], radii[vis].float())
        ga[vis] += torch.norm(g[vis, :2], dim=-1, keepdim=True)
        st[vis] += 1

    out = {}
    for mode in ("torch", "fused", "fused_raw"):
        scene = scene_cpu.to(dev, requires_grad=True)
        named = scene.named_params()
        groups = [{"params": [named[k]], "lr": lrs[k], "name": k} for k in named]
        fused = mode != "torch"
        opt = FusedAdam(groups, eps=1e-15) if fused else \
            torch.optim.Adam(groups, eps=1e-15)
        r = MsplatRender({"max_sh_degree": 3}, device="cuda")
        r.sh_degree = 3
        n = len(scene)
        stats = DensifyStats(n, dev, W, H)
        ga, st, mr = (torch.zeros(n, 1, device=dev), torch.zeros(n, 1, device=dev),
                      torch.zeros(n, device=dev))

        def step():
            res = r.render_iter(**cam.render_kwargs(),
                                **scene.render_inputs(split_sh=True, raw=mode == "fused_raw"))
            rgb = res["rendered_features_split"]["rgb"]
            if fused:
                loss = l1_ssim_loss(rgb, target, 0.2)["loss"]
            else:
                loss = torch_loss(rgb, target)
            loss.backward()
            if fused:
                stats.update(res["uv_points"], res["visibility"], res["radii"])
            else:
                torch_stats(ga, st, mr, res["uv_points"], res["visibility"], res["radii"])
            opt.step()
            opt.zero_grad(set_to_none=True)

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        out[mode] = (time.perf_counter() - t0) * 1000 / steps
        log(f"train500k {mode}: {out[mode]:.3f} ms/step")
    return {"workload": "train500k: 500k splats SH3, 979x546, 0.8 L1 + 0.2 (1-SSIM), "
                        "densification statistics, Adam",
            "fused_ms_per_step": round(out["fused"], 3),
            "fused_raw_ms_per_step": round(out["fused_raw"], 3),
            "torch_ms_per_step": round(out["torch"], 3),
            "speedup": round(out["torch"] / out["fused"], 3),
            "speedup_raw": round(out["torch"] / out["fused_raw"], 3)}


def bench_serve(log, n_requests=300, rates=(50.0, 200.0, 800.0)):
    """Serving: 2M resident splats (SH 3), Poisson request arrivals at each
    rate, resolutions {540p 40%, 720p 30%, 1080p 25%, 2160p 5%}, eye at
    distance U[2,8], 20% of requests also want depth.  The same trace is
    replayed in real time against the one-view-per-call server (plain) and
    the continuous-batching server; latency is arrival to images complete."""
    from pointrix_amd import scene as S
    from pointrix_amd.serve import BatchServer, PlainServer, make_requests, run

    dev = torch.device("cuda:0")
    scene_cpu, _ = S.serve2m(0)
    scene = scene_cpu.to(dev)
    with torch.no_grad():
        inputs = scene.render_inputs(split_sh=True)
    servers = {"plain": PlainServer(inputs), "batched": BatchServer(inputs)}
    warm = make_requests(1000.0, 24, seed=99)
    for srv in servers.values():               # allocator / code warm-up at every size
        run(srv, warm)
    out = {}
    for rate in rates:
        reqs = make_requests(rate, n_requests, seed=int(rate))
        row = {}
        for name, srv in servers.items():
            st = run(srv, reqs)
            row[name] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()}
            log(f"serve {rate:.0f} req/s {name}: p50 {st['p50_ms']:.2f} ms p99 "
                f"{st['p99_ms']:.2f} ms, {st['views_per_s']:.1f} views/s, mean batch "
                f"{st['mean_batch']:.2f}")
        row["p99_speedup"] = round(row["plain"]["p99_ms"] / row["batched"]["p99_ms"], 3)
        out[f"{rate:.0f}"] = row
    b = servers["batched"]
    return {"workload": "serve2m: 2M splats SH3, Poisson arrivals, {540p .40, 720p .30, "
                        "1080p .25, 2160p .05}, eye r U[2,8], 20% depth",
            "requests_per_rate": n_requests,
            "key_budget_instances": b.max_instances, "budget_splits": b.splits,
            "rates": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="full3m")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--batched", type=int, default=1, help="also run the 16-view batched case")
    ap.add_argument("--serve", type=int, default=1, help="also run the serving trace")
    ap.add_argument("--train500k", type=int, default=1,
                    help="also time the full config-2 training step (loss, stats, Adam)")
    ap.add_argument("--tight", type=int, default=1,
                    help="also time the step with opacity-aware binning")
    ap.add_argument("--raw", type=int, default=1,
                    help="also time the step with activations fused into preprocess")
    ap.add_argument("--graph", type=int, default=1,
                    help="also time the step captured as one HIP graph (value = the faster "
                         "of eager and graph, the graph only if its replays match eager)")
    ap.add_argument("--only", default="",
                    help="run only this section: batched, serve, or copy (HBM counter "
                         "calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the "
              f"renderer runs on a single GPU (one process, one device)", file=sys.stderr)
        sys.exit(2)

    from pointrix_amd import ops
    from pointrix_amd import scene as S
    from pointrix_amd.renderer import MsplatRender

    if args.only == "copy":
        # calibration for HBM counters: 4 x a 256 MiB device-to-device copy
        dev = torch.device("cuda:0")
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": src.numel() * 4, "copies": 4}), flush=True)
        return
    if args.only == "serve":
        res = bench_serve(lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))
        print(json.dumps({"serving": res}), flush=True)
        return
    if args.only == "batched":
        res = bench_batched(args.steps, args.warmup,
                            lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))
        res.pop("_tune")
        print(json.dumps({"batched": res}), flush=True)
        return

    dev = torch.device("cuda:0")
    wl = S.WORKLOADS[args.workload]
    scene_cpu, cams = wl(0)
    scene = scene_cpu.to(dev, requires_grad=True)
    cam = TrainableCamera(cams[0], dev)
    W, H = cam.width, cam.height
    g = torch.Generator().manual_seed(123)
    targets = {"rgb": torch.rand(3, H, W, generator=g).to(dev),
               "depth": (torch.rand(1, H, W, generator=g) * 8).to(dev),
               "normal": torch.rand(3, H, W, generator=g).to(dev)}
    renderer = MsplatRender({"render_depth": True, "render_normal": True, "max_sh_degree": 3},
                            white_bg=False, device="cuda")
    renderer.sh_degree = 3
    step, loss_fn = make_step(renderer, scene, cam, targets)
    print(f"[bench] workload {args.workload}: N={len(scene)} {W}x{H}", file=sys.stderr, flush=True)

    t_w = time.perf_counter()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # the GPU reaches its sustained clocks only after some hundreds of ms of
    # load: keep stepping (untimed) until the warmup has lasted SETTLE_S
    settle = settle_steps(step, t_w)
    print(f"[bench] warmup done ({args.warmup} steps + {settle} settle steps)", file=sys.stderr,
          flush=True)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt * 1000 / args.steps
    print("[bench] timed steps done", file=sys.stderr, flush=True)
    # per-kernel device times: the same steps again, with events around
    # every C-ABI call on its launch stream (not part of the timed region)
    ops.timer = ops.KernelTimer()
    for i in range(args.steps):
        step()
    torch.cuda.synchronize()
    kt = ops.timer.summary()
    ops.timer = None
    print("[bench] instrumented steps done", file=sys.stderr, flush=True)

    # sizes for the algorithmic-byte model
    with torch.no_grad():
        res = renderer.render_iter(**cam.render_kwargs(), **scene.render_inputs(split_sh=True))
    Nv = int(res["visibility"].sum().item())
    M = ops.last_instance_count
    N = len(scene)
    K, C, P = 16, 7, W * H
    T = ((W + 15) // 16) * ((H + 15) // 16)
    ab = algorithmic_bytes(N, K, C, P, T, Nv, M)
    per_kernel = {k: {"launches": n, "avg_ms": tot / max(n, 1)} for k, (n, tot) in kt.items()}
    # the roofline line is for one kernel: the sort spans cover several each
    # (count, scans, radix passes, emission), so the dominant kernel is taken
    # among the single-kernel spans
    single = [k for k in per_kernel if not k.startswith("sort")]
    dom = max(single, key=lambda k: per_kernel[k]["avg_ms"] * per_kernel[k]["launches"])
    dom_bytes = ab[dom]
    achieved = dom_bytes / (per_kernel[dom]["avg_ms"] * 1e-3) / 1e9
    traffic = PMC_TRAFFIC_BYTES.get(dom) if args.workload == "full3m" else None
    copy_gbs = copy_bandwidth()
    print(f"[bench] device copy {copy_gbs} GB/s", file=sys.stderr, flush=True)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "algorithmic_bytes": int(dom_bytes),
                "traffic_source": PMC_SOURCE if traffic else None,
                "copy_measured_GBs": copy_gbs,
                "frac_of_measured_copy": round(achieved / copy_gbs, 4)}
    if "sort_count" in per_kernel and "sort_emit" in per_kernel:
        t_bin = per_kernel["sort_count"]["avg_ms"] + per_kernel["sort_emit"]["avg_ms"]
        roofline["binning"] = {"ms": round(t_bin, 4), "algorithmic_bytes": int(ab["binning"]),
                               "achieved": round(ab["binning"] / (t_bin * 1e-3) / 1e9, 1),
                               "frac": round(ab["binning"] / (t_bin * 1e-3) / 1e9 /
                                             HBM_PEAK_GBS, 4)}
    print(f"[bench] per-kernel: {json.dumps(per_kernel)}", file=sys.stderr, flush=True)

    # the FP32 C restatement of the whole step on the same inputs: CPU
    # baseline (timed in full), parity at full size, exact blend work counts
    cpu = parity = None
    if args.cpu_baseline:
        print("[bench] FP32 restatement of the full step on the host ...", file=sys.stderr,
              flush=True)
        cpu, parity, pairs = restatement_leg(renderer, scene, cam, targets)
        ops_ = blend_ops(C, pairs["blended"], pairs["examined"])
        valu = {}
        for k in ("alpha_blending_fwd", "alpha_blending_bwd"):
            if k not in per_kernel:
                continue
            t_k = per_kernel[k]["avg_ms"] * 1e-3
            walked = pairs["examined"] if k.endswith("fwd") else pairs["n_contrib_sum"]
            valu[k] = {"pairs_blended": pairs["blended"], "pairs_walked": walked,
                       "gpairs_per_s": round(pairs["blended"] / t_k / 1e9, 1),
                       "algorithmic_ops": int(ops_[k]),
                       "achieved_tops": round(ops_[k] / t_k / 1e12, 2),
                       "peak_tops": round(VALU_PEAK_OPS / 1e12, 2),
                       "frac": round(ops_[k] / t_k / VALU_PEAK_OPS, 4)}
            valu[k].update(VALU_PMC.get(k, {}))
        roofline["valu"] = valu
        print(f"[bench] cpu baseline {cpu['value']} steps/s; parity {json.dumps(parity)}; "
              f"valu {json.dumps(valu)}", file=sys.stderr, flush=True)
        print("[bench] FP64 transcription: cpu10k timing and per-output errors ...",
              file=sys.stderr, flush=True)
        cpu["fp64_cpu10k"] = fp64_cpu10k()
        parity["vs_fp64"] = fp64_errors(scene_cpu, cams[0], dev)
        print(f"[bench] fp64: {json.dumps(parity['vs_fp64'])}", file=sys.stderr, flush=True)

    # the same step with opacity-aware binning (same images, fewer instances;
    # not the msplat tile contract, so reported beside the headline only)
    tight = None
    if args.tight:
        tight = time_tight(renderer, step, args.steps)

    # the same step fed the raw parameters (activations fused into preprocess)
    raw = None
    if args.raw:
        raw = time_raw(renderer, scene, cam, targets, args.steps)

    graph = None
    if args.graph:
        print("[bench] graphed step ...", file=sys.stderr, flush=True)
        graph = graphed_headline(renderer, scene, cam, loss_fn, args.steps, args.warmup,
                                 lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))
    ms_eager = ms
    if graph is not None and graph["replays_match_eager"]:
        ms = min(ms, graph["ms_per_step"])

    batched = None
    if args.batched:
        print("[bench] batched multi-view ...", file=sys.stderr, flush=True)
        blog = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
        b1 = bench_batched(max(3, args.steps // 4), 2, blog)
        print("[bench] batched multi-view at 3M splats ...", file=sys.stderr, flush=True)
        b3 = bench_batched(3, 1, blog, n=3_000_000, tune=b1.pop("_tune"))
        b3.pop("_tune")
        b3["autotune"] = {"chosen": b1["autotune"]["chosen"], "note": "choice of the 1M run"}
        batched = {"1M": b1, "3M": b3}

    train500k = None
    if args.train500k:
        print("[bench] train500k full step ...", file=sys.stderr, flush=True)
        train500k = bench_train500k(max(5, args.steps // 2), 3,
                                    lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))

    serving = None
    if args.serve:
        print("[bench] serving ...", file=sys.stderr, flush=True)
        serving = bench_serve(lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))

    line = {
        "metric": "fwd+bwd steps/s @3M Gaussians 1237x822",
        "value": round(1000.0 / ms, 3), "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded random scene and targets)",
        "config": {"workload": f"{args.workload}: {N} splats SH3, {W}x{H}, rgb+depth+normal, "
                               "grads to splats + camera pose/intrinsics",
                   "splats": N, "width": W, "height": H, "visible": Nv, "instances": M,
                   "mpixel_per_s": round(P / 1e6 * 1000.0 / ms, 2),
                   "eager_ms_per_step": round(ms_eager, 4), "graph": graph,
                   "warmup_settle": {"seconds": SETTLE_S, "extra_untimed_steps": settle},
                   "tight_tiles": tight, "raw_activations": raw,
                   "kernels_ms": {k: round(v["avg_ms"], 4) for k, v in per_kernel.items()}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
        "batched": batched,
        "serving": serving,
        "train500k": train500k,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
