"""Seeded synthetic Gaussian scenes and pinhole cameras.

Scenes hold the raw (pre-activation) parameters of a Gaussian point cloud
the way pointrix's GaussianPointCloud does (pointrix/model/point_cloud/
gaussian_points.py:31-35,70-90): position, features (SH DC) [N,1,3],
features_rest [N,K-1,3], log-scale, raw quaternion (wxyz) and opacity logit;
``render_inputs()`` applies exp / normalize / sigmoid and concatenates the SH.
Cameras use the OpenCV convention (x right, y down, z forward): a 4x4
world-to-camera matrix, intrinsics (fx, fy, cx, cy) with cx = W/2, cy = H/2
(the convention of pointrix/webgui/gui.py:179), and centre = -R^T t.

Workloads are named by content; their distributions are listed in WORKLOADS.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import torch


def look_at(eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)) -> torch.Tensor:
    """4x4 world-to-camera matrix of a camera at ``eye`` looking at ``target``."""
    eye = torch.as_tensor(eye, dtype=torch.float64)
    target = torch.as_tensor(target, dtype=torch.float64)
    up = torch.as_tensor(up, dtype=torch.float64)
    z = target - eye
    z = z / z.norm()
    if torch.abs(torch.dot(z, up / up.norm())) > 0.999:
        up = torch.tensor([0.0, 0.0, 1.0], dtype=torch.float64)
    x = torch.linalg.cross(-up, z)
    x = x / x.norm()
    y = torch.linalg.cross(z, x)
    R = torch.stack([x, y, z], 0)
    t = -R @ eye
    E = torch.eye(4, dtype=torch.float64)
    E[:3, :3] = R
    E[:3, 3] = t
    return E.float()


@dataclass
class Camera:
    width: int
    height: int
    extrinsic_matrix: torch.Tensor  # [4,4] world-to-camera
    intrinsic_params: torch.Tensor  # [4] (fx, fy, cx, cy)
    camera_center: torch.Tensor     # [3]

    def to(self, device):
        return Camera(self.width, self.height, self.extrinsic_matrix.to(device),
                      self.intrinsic_params.to(device), self.camera_center.to(device))

    def render_kwargs(self) -> dict:
        return {"height": self.height, "width": self.width,
                "extrinsic_matrix": self.extrinsic_matrix,
                "intrinsic_params": self.intrinsic_params,
                "camera_center": self.camera_center}


def make_camera(width, height, eye, focal=None, fov_deg=None, target=(0.0, 0.0, 0.0)) -> Camera:
    """Pinhole camera at ``eye`` looking at ``target``; focal in pixels or horizontal FOV."""
    if focal is None:
        focal = 0.5 * width / math.tan(math.radians(fov_deg) / 2)
    E = look_at(eye, target)
    intr = torch.tensor([focal, focal, width / 2.0, height / 2.0], dtype=torch.float32)
    R = E[:3, :3].double()
    center = (-R.T @ E[:3, 3].double()).float()
    return Camera(int(width), int(height), E, intr, center)


def random_sphere_eye(gen: torch.Generator, r_lo: float, r_hi: float | None = None):
    d = torch.randn(3, generator=gen, dtype=torch.float64)
    d = d / d.norm()
    r = r_lo if r_hi is None else r_lo + (r_hi - r_lo) * torch.rand(1, generator=gen).item()
    return (d * r).tolist()


class GaussianScene:
    """Raw Gaussian parameters (leaf tensors) + the activations of GaussianPointCloud."""

    def __init__(self, position, features, features_rest, scaling, rotation, opacity):
        self.position = position
        self.features = features
        self.features_rest = features_rest
        self.scaling = scaling
        self.rotation = rotation
        self.opacity = opacity

    def params(self):
        return [self.position, self.features, self.features_rest, self.scaling, self.rotation,
                self.opacity]

    def named_params(self):
        return {"position": self.position, "features": self.features,
                "features_rest": self.features_rest, "scaling": self.scaling,
                "rotation": self.rotation, "opacity": self.opacity}

    def __len__(self):
        return self.position.shape[0]

    def to(self, device, requires_grad=False):
        ps = [p.detach().to(device).requires_grad_(requires_grad) for p in self.params()]
        return GaussianScene(*ps)

    def render_inputs(self, split_sh: bool = False, raw: bool = False) -> dict:
        """Activated tensors, as BaseModel.forward passes them (base_model.py:66-71);
        on the GPU the three activations run as one fused kernel (ops.activate).

        split_sh: pass SH as the pair (features, features_rest) instead of
        their concatenation (get_shs, gaussian_points.py:86-90).
        raw: pass opacity, scaling and rotation unactivated with ``raw=True``,
        for a renderer that applies the activations in its own kernel."""
        if self.features_rest is None or self.features_rest.shape[1] == 0:
            shs = self.features
        elif split_sh:
            shs = (self.features, self.features_rest)
        else:
            shs = torch.cat([self.features, self.features_rest], dim=1)
        if raw:
            return {"position": self.position, "opacity": self.opacity,
                    "scaling": self.scaling, "rotation": self.rotation, "shs": shs, "raw": True}
        if self.scaling.is_cuda:
            from .ops import activate
            scaling, rotation, opacity = activate(self.scaling, self.rotation, self.opacity)
            return {"position": self.position, "opacity": opacity, "scaling": scaling,
                    "rotation": rotation, "shs": shs}
        return {"position": self.position,
                "opacity": torch.sigmoid(self.opacity),
                "scaling": torch.exp(self.scaling),
                "rotation": torch.nn.functional.normalize(self.rotation, dim=-1),
                "shs": shs}


def random_scene(n, seed, sh_degree=0, means="uniform", log_scale=(-4.0, 0.5),
                 opacity_std=1.0, dc_std=0.3, rest_std=0.05, shell_frac=0.0,
                 shell_r=(20.0, 50.0), means_std=(1.0, 1.0, 1.0)) -> GaussianScene:
    """Seeded scene on the CPU (float32)."""
    g = torch.Generator().manual_seed(seed)
    if means == "uniform":
        pos = torch.rand(n, 3, generator=g) * 2 - 1
    else:
        pos = torch.randn(n, 3, generator=g) * torch.tensor(means_std)
    if shell_frac > 0:
        k = int(round(n * shell_frac))
        d = torch.randn(k, 3, generator=g)
        d = d / d.norm(dim=1, keepdim=True)
        r = shell_r[0] + (shell_r[1] - shell_r[0]) * torch.rand(k, 1, generator=g)
        pos[n - k:] = d * r
    scaling = torch.randn(n, 3, generator=g) * log_scale[1] + log_scale[0]
    rot = torch.randn(n, 4, generator=g)
    rot = rot / rot.norm(dim=1, keepdim=True)
    opacity = torch.randn(n, 1, generator=g) * opacity_std
    K = (sh_degree + 1) ** 2
    features = torch.randn(n, 1, 3, generator=g) * dc_std
    features_rest = torch.randn(n, K - 1, 3, generator=g) * rest_std
    return GaussianScene(pos, features, features_rest, scaling, rot, opacity)


# ---------------------------------------------------------------------------
# Named workloads (the five synthetic configurations of the benchmark suite)
# ---------------------------------------------------------------------------
def cpu10k(seed=0):
    """10k splats, means U[-1,1]^3, log-scale N(-4,0.5), opacity sigmoid(N(0,1)),
    SH degree 0 (DC N(0,0.3)); one 256x256 camera, 60 deg FOV, 3 units away."""
    scene = random_scene(10_000, seed, sh_degree=0, means="uniform", log_scale=(-4.0, 0.5),
                         opacity_std=1.0, dc_std=0.3)
    cam = make_camera(256, 256, eye=(0.0, 0.0, -3.0), fov_deg=60.0)
    return scene, [cam]


def train500k(seed=0):
    """500k splats, means N(0,1.5^2 I), log-scale N(-4.5,0.7), opacity
    sigmoid(N(0,1.5)), SH degree 3 (DC N(0,0.3), rest N(0,0.05)); one 979x546
    camera, focal 1.1*W, radius 4 looking at the origin."""
    scene = random_scene(500_000, seed, sh_degree=3, means="normal", means_std=(1.5, 1.5, 1.5),
                         log_scale=(-4.5, 0.7), opacity_std=1.5)
    g = torch.Generator().manual_seed(seed + 1)
    cam = make_camera(979, 546, eye=random_sphere_eye(g, 4.0), focal=1.1 * 979)
    return scene, [cam]


def full3m(seed=0, n=3_000_000):
    """3M splats with the cpu10k parameter distributions (log-scale N(-4,0.5),
    opacity sigmoid(N(0,1)), unit quaternions), SH degree 3 (DC N(0,0.3),
    rest N(0,0.05)), anisotropic means N(0, diag(4,1,4)) plus a 10% far shell
    at radius 20-50; one 1237x822 camera (focal 1.1*W, radius 4 looking at
    the origin, the train500k camera convention)."""
    scene = random_scene(n, seed, sh_degree=3, means="normal", means_std=(2.0, 1.0, 2.0),
                         log_scale=(-4.0, 0.5), opacity_std=1.0, shell_frac=0.1)
    g = torch.Generator().manual_seed(seed + 1)
    cam = make_camera(1237, 822, eye=random_sphere_eye(g, 4.0), focal=1.1 * 1237)
    return scene, [cam]


def full3m_light(seed=0, n=3_000_000):
    """full3m's means and camera with the train500k parameter distributions
    (log-scale N(-4.5,0.7), opacity sigmoid(N(0,1.5))): smaller splats, about
    30% fewer (splat, tile) instances."""
    scene = random_scene(n, seed, sh_degree=3, means="normal", means_std=(2.0, 1.0, 2.0),
                         log_scale=(-4.5, 0.7), opacity_std=1.5, shell_frac=0.1)
    g = torch.Generator().manual_seed(seed + 1)
    cam = make_camera(1237, 822, eye=random_sphere_eye(g, 4.0), focal=1.1 * 1237)
    return scene, [cam]


BATCH_RESOLUTIONS = [(256, 256), (640, 360), (979, 546), (1237, 822), (1920, 1080)]


def batched1m(seed=0, views=16, n=1_000_000):
    """1M splats with the cpu10k distributions (means U[-1,1]^3, log-scale
    N(-4,0.5), opacity sigmoid(N(0,1))), SH degree 3; 16 cameras: resolution
    uniform over BATCH_RESOLUTIONS, eye on a radius 3-6 shell, horizontal FOV
    U[40,80] deg."""
    scene = random_scene(n, seed, sh_degree=3, means="uniform", log_scale=(-4.0, 0.5),
                         opacity_std=1.0)
    g = torch.Generator().manual_seed(seed + 1)
    cams = []
    for _ in range(views):
        w, h = BATCH_RESOLUTIONS[int(torch.randint(len(BATCH_RESOLUTIONS), (1,), generator=g))]
        fov = 40.0 + 40.0 * torch.rand(1, generator=g).item()
        cams.append(make_camera(w, h, eye=random_sphere_eye(g, 3.0, 6.0), fov_deg=fov))
    return scene, cams


SERVE_RESOLUTIONS = [((960, 540), 0.40), ((1280, 720), 0.30), ((1920, 1080), 0.25),
                     ((3840, 2160), 0.05)]


def serve2m(seed=0, n=2_000_000):
    """2M splats with the full3m distributions; requests are generated by the server."""
    scene = random_scene(n, seed, sh_degree=3, means="normal", means_std=(2.0, 1.0, 2.0),
                         log_scale=(-4.0, 0.5), opacity_std=1.0, shell_frac=0.1)
    return scene, []


WORKLOADS = {"cpu10k": cpu10k, "train500k": train500k, "full3m": full3m,
             "full3m_light": full3m_light, "batched1m": batched1m, "serve2m": serve2m}
