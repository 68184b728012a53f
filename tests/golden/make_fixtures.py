"""Write the golden fixtures of this directory from the FP64 oracle.

The reference ships no vectors for this path (its msplat back-end is an
un-vendored submodule), so the fixtures are outputs of oracle/torch64.py --
itself pinned by tests/test_oracle.py -- on seeded synthetic scenes:

* sh3_1200_56.npz: 1,200 splats, SH degree 3, 56x56, white background: the
  exact float32 inputs; the FP64 oracle evaluated at the FP32 forward point
  (uv and conic from the FP32 C restatement, which the GPU matches bit for
  bit, FP64 derivatives kept): rgb + depth + normal + opacity images and the
  gradients of a fixed random weighting of them into every input (splats and
  camera); and the plain end-to-end FP64 images;
* cpu10k.npz: the 10,000-splat 256x256 SH-0 scene, rgb image and per-pixel
  contributor counts.

The cpu10k inputs are regenerated from the seed (pointrix_amd.scene); a
checksum of them is stored so that a change of generator is detected.
Run: python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ref32 as r32  # noqa: E402
from oracle import torch64 as o64  # noqa: E402
from pointrix_amd import scene as S  # noqa: E402


def sh3_case():
    scene = S.random_scene(1200, 31, sh_degree=3, means="normal", means_std=(0.6, 0.6, 0.6),
                           log_scale=(-3.0, 0.5), opacity_std=1.5)
    cam = S.make_camera(56, 56, eye=(0.5, -0.4, -2.4), fov_deg=60.0)
    inp = {k: v.detach().float().contiguous() for k, v in scene.render_inputs().items()}
    return inp, cam


def input_checksum(inp, cam):
    parts = [v.detach().double().reshape(-1) for v in inp.values()]
    parts += [cam.extrinsic_matrix.double().reshape(-1), cam.intrinsic_params.double()]
    return float(torch.cat(parts).abs().sum())


def weights_for(shapes):
    g = torch.Generator().manual_seed(77)
    return {k: torch.randn(s, generator=g, dtype=torch.float64) for k, s in shapes.items()}


def fp32_point(inp, cam):
    """uv and conic of the FP32 forward (C restatement == GPU, bit for bit)."""
    K = cam.intrinsic_params.numpy()
    E = cam.extrinsic_matrix.numpy()
    uv, d = r32.project_point(inp["position"].numpy(), K, E, cam.width, cam.height, 0.2)
    vis = d != 0
    cov = r32.compute_cov3d(inp["scaling"].numpy(), inp["rotation"].numpy(), vis)
    con, _, _ = r32.ewa_project(inp["position"].numpy(), cov, K, E, uv, cam.width, cam.height,
                                vis)
    return torch.from_numpy(uv), torch.from_numpy(con)


def oracle_sh3(inp, cam, snap):
    camk = {"extrinsic_matrix": cam.extrinsic_matrix, "intrinsic_params": cam.intrinsic_params,
            "camera_center": cam.camera_center}
    hl = {k: v.double().requires_grad_(True) for k, v in {**inp, **camk}.items()}
    normals = o64.splat_normals(hl["position"], hl["scaling"], hl["rotation"],
                                hl["camera_center"], hl["extrinsic_matrix"])
    res = o64.render_iter(cam.height, cam.width, hl["extrinsic_matrix"],
                          hl["intrinsic_params"], hl["camera_center"], hl["position"],
                          hl["opacity"], hl["scaling"], hl["rotation"], hl["shs"], sh_degree=3,
                          bg=1.0, render_depth=True, normals=normals, render_opacity=True,
                          snap=snap)
    imgs = res["rendered_features_split"]
    w = weights_for({k: tuple(v.shape) for k, v in imgs.items()})
    sum((imgs[k] * w[k]).sum() for k in imgs).backward()
    return res, imgs, hl


def run_sh3():
    inp, cam = sh3_case()
    out = {"in_" + k: v.numpy() for k, v in inp.items()}
    uv, con = fp32_point(inp, cam)
    res, imgs, hl = oracle_sh3(inp, cam, {"uv": uv, "conic": con})
    for k, v in imgs.items():
        out["img_" + k] = v.detach().float().numpy()
    for k, v in hl.items():
        out["grad_" + k] = v.grad.float().numpy()
    out["radii"] = res["radii"].numpy().astype(np.int16)
    _, imgs64, _ = oracle_sh3(inp, cam, None)
    for k, v in imgs64.items():
        out["e2e_img_" + k] = v.detach().float().numpy()
    return out


def run_cpu10k():
    scene, cams = S.cpu10k()
    cam = cams[0]
    inp = {k: v.detach() for k, v in scene.render_inputs().items()}
    res = o64.render_iter(256, 256, cam.extrinsic_matrix, cam.intrinsic_params,
                          cam.camera_center, **inp, sh_degree=0, return_aux=True)
    return {"checksum": input_checksum(inp, cam),
            "img_rgb": res["rendered_features_split"]["rgb"].detach().float().numpy(),
            "n_contrib": res["aux"]["n_contrib"].numpy().astype(np.int16)}


def main():
    np.savez_compressed(os.path.join(HERE, "sh3_1200_56.npz"), **run_sh3())
    np.savez_compressed(os.path.join(HERE, "cpu10k.npz"), **run_cpu10k())
    for f in ("sh3_1200_56.npz", "cpu10k.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
