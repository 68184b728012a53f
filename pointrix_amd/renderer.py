"""MsplatRender-compatible renderer plugin running on libgsr.

Mirrors pointrix/model/renderer/msplat.py (class MsplatRender, its Config,
setup, render_iter, render_batch, update_sh_degree, state_dict) and
pointrix/model/renderer/utils/renderer_utils.py (RenderFeatures), with the
msplat calls replaced by ``pointrix_amd.ops``.  Same inputs, same output dict
("rendered_features_split", "uv_points", "visibility", "radii"; render_batch
stacks features per name).  Extensions, all off by default so the default
output keys equal the reference's: ``render_normal`` (per-splat normal channel,
as examples/supervise renders it) and ``render_opacity`` (1 - T_final).
"""
from __future__ import annotations

from dataclasses import dataclass, fields
from collections.abc import Mapping
from typing import Dict

import torch

from . import ops


class Registry:
    """Name -> class registry with the semantics of pointrix/utils/registry.py:
    an object registers under its ``__name__`` (as a decorator or a plain
    call), a name can be registered once only, ``get`` raises KeyError for an
    unknown name."""

    def __init__(self, name):
        self.name = name
        self._map = {}

    def _add(self, name, obj):
        if name in self._map:
            raise AssertionError(f"An object named '{name}' was already registered in "
                                 f"'{self.name}' registry!")
        self._map[name] = obj

    def register(self, obj=None):
        if obj is None:
            def deco(o):
                self._add(o.__name__, o)
                return o
            return deco
        self._add(obj.__name__, obj)
        return obj

    def get(self, name):
        if name not in self._map:
            raise KeyError(f"No object named '{name}' found in '{self.name}' registry!")
        return self._map[name]

    def __contains__(self, name):
        return name in self._map

    def __iter__(self):
        return iter(self._map.items())


RENDERER_REGISTRY = Registry("RENDERER")


class RenderFeatures:
    """Per-splat feature channels concatenated for one blend, split afterwards
    (renderer_utils.py:4-71); insertion order defines the channel order."""

    def __init__(self, **kwargs) -> None:
        for k, v in kwargs.items():
            setattr(self, k, v)

    def to(self, device: str) -> None:
        for k, v in self.__dict__.items():
            if isinstance(v, torch.Tensor):
                setattr(self, k, v.to(device))

    def combine(self) -> torch.Tensor:
        return torch.cat([v for v in self.__dict__.values() if isinstance(v, torch.Tensor)], -1)

    def split(self, feature: torch.Tensor) -> Dict[str, torch.Tensor]:
        out, start = {}, 0
        for k, v in self.__dict__.items():
            if isinstance(v, torch.Tensor):
                end = start + v.shape[-1]
                out[k] = feature[start:end, ...]
                start = end
        return out


def quat_to_rotmat(q: torch.Tensor) -> torch.Tensor:
    """wxyz quaternion -> rotation matrix (pointrix/utils/pose.py:40-84)."""
    w, x, y, z = q.unbind(-1)
    x2, y2, z2, w2 = x * x, y * y, z * z, w * w
    xy, zw, xz, yw, yz, xw = x * y, z * w, x * z, y * w, y * z, x * w
    return torch.stack([
        torch.stack([x2 - y2 - z2 + w2, 2 * (xy - zw), 2 * (xz + yw)], -1),
        torch.stack([2 * (xy + zw), -x2 + y2 - z2 + w2, 2 * (yz - xw)], -1),
        torch.stack([2 * (xz - yw), 2 * (yz + xw), -x2 - y2 + z2 + w2], -1)], -2)


def splat_normals(position, scaling, rotation, camera_center, extrinsic_matrix):
    """Camera-space normal of each splat: its shortest axis, turned towards the
    camera (examples/supervise/model.py:50-74)."""
    R = quat_to_rotmat(rotation)
    idx = torch.argmin(scaling.detach(), dim=-1)
    n = torch.gather(R, 2, idx[:, None, None].expand(-1, 3, 1)).squeeze(-1)
    n = torch.nn.functional.normalize(n, dim=-1)
    d = (camera_center.reshape(1, 3) - position).detach()
    flip = (d * n.detach()).sum(-1) < 0
    n = torch.where(flip[:, None], -n, n)
    return n @ extrinsic_matrix[:3, :3].T


@RENDERER_REGISTRY.register()
class MsplatRender:
    """Differentiable Gaussian rasteriser plugin (msplat.py:12-248)."""

    @dataclass
    class Config:
        update_sh_iter: int = 1000
        max_sh_degree: int = 3
        render_depth: bool = False
        render_normal: bool = False
        render_opacity: bool = False
        # one fused HIP kernel for direction/SH/projection/covariance/EWA
        # (False: the six msplat-style ops one after another)
        fused: bool = True
        # opacity-aware binning in the fused path: same images, fewer
        # (splat, tile) instances; off by default because the sorted ids and
        # tile ranges then differ from the msplat tile contract
        tight_tiles: bool = False

    def __init__(self, cfg=None, white_bg=False, device="cuda", **kwargs):
        self.cfg = self.parse_cfg(cfg)
        self.setup(white_bg, device, **kwargs)

    @classmethod
    def parse_cfg(cls, cfg):
        """Config from what BaseObject receives (pointrix/utils/base.py:32-39,
        parse_structured at utils/config.py:102-107): None (defaults), a
        Config, any Mapping (dict, OmegaConf DictConfig ...) or an object with
        the fields as attributes.  As ``Config(**cfg)`` does there, a key that
        is not a Config field raises TypeError."""
        if cfg is None:
            return cls.Config()
        if isinstance(cfg, cls.Config):
            return cfg
        names = [f.name for f in fields(cls.Config)]
        if isinstance(cfg, Mapping) or (hasattr(cfg, "keys") and hasattr(cfg, "__getitem__")):
            items = {str(k): cfg[k] for k in cfg.keys()}
            unknown = sorted(set(items) - set(names))
            if unknown:
                raise TypeError(f"{cls.__name__}.Config got unexpected keys {unknown}")
            return cls.Config(**items)
        return cls.Config(**{k: getattr(cfg, k) for k in names if hasattr(cfg, k)})

    def setup(self, white_bg, device, **kwargs):
        self.sh_degree = 0
        self.device = device
        self.bg_color = 1.0 if white_bg else 0.0
        self.workspace = ops.SortWorkspace()
        # when True, render_iter keeps its per-splat / per-pixel intermediates
        # (uv, depth, conic, radius, ids, tile_range, features, T_final,
        # n_contrib) in self.intermediates, for parity checks and debugging
        self.keep_intermediates = False
        self.intermediates = None

    def render_iter(self, height, width, extrinsic_matrix, intrinsic_params, camera_center,
                    position, opacity, scaling, rotation, shs, normals=None, raw=False,
                    **kwargs) -> dict:
        """One view (msplat.py:51-158).  ``shs`` is [N,K,3], or the pair
        (features [N,1,3], features_rest [N,K-1,3]) of pointrix's point cloud,
        which the fused path reads without concatenating.  ``raw=True``:
        opacity, scaling and rotation are the point cloud's raw parameters
        (logit, log scale, unnormalised wxyz); the fused path applies sigmoid,
        exp and normalize inside its preprocess kernel (gaussian_points.py:
        30-35,70-81), the unfused path with torch."""
        if isinstance(shs, (tuple, list)) and not self.cfg.fused:
            shs = torch.cat(list(shs), dim=1)
        if self.cfg.fused:
            return self._render_iter_fused(height, width, extrinsic_matrix, intrinsic_params,
                                           camera_center, position, opacity, scaling, rotation,
                                           shs, normals, raw)
        if raw:
            opacity, scaling = torch.sigmoid(opacity), torch.exp(scaling)
            rotation = torch.nn.functional.normalize(rotation, dim=-1)
        direction = position - camera_center.reshape(1, 3)
        direction = direction / direction.norm(dim=1, keepdim=True)
        # SH warm-up: the first (sh_degree+1)^2 coefficients (msplat.py:98-101)
        rgb = ops.compute_sh(shs.permute(0, 2, 1), direction, degree=self.sh_degree)
        rgb = (rgb + 0.5).clamp(min=0.0)
        extr = extrinsic_matrix[:3, :]
        uv, depth = ops.project_point(position, intrinsic_params, extr, width, height,
                                      nearest=0.2)
        visible = depth != 0
        cov3d = ops.compute_cov3d(scaling, rotation, visible)
        conic, radius, tiles = ops.ewa_project(position, cov3d, intrinsic_params, extr, uv, width,
                                               height, visible)
        ids, tile_range = ops.sort_gaussian(uv, depth, width, height, radius, tiles,
                                            workspace=self.workspace)
        feats = {"rgb": rgb}
        if self.cfg.render_depth:
            feats["depth"] = depth
        if normals is None and self.cfg.render_normal:
            normals = splat_normals(position, scaling, rotation, camera_center, extrinsic_matrix)
        if normals is not None:
            feats["normal"] = normals
        Render_Features = RenderFeatures(**feats)
        render_features = Render_Features.combine()
        ndc = torch.zeros_like(uv, requires_grad=True)
        ndc.retain_grad()
        rendered, T_final, n_contrib = ops.alpha_blending_ex(uv, conic, opacity, render_features,
                                                             ids, tile_range, self.bg_color,
                                                             width, height, ndc)
        if self.keep_intermediates:
            self.intermediates = {"uv": uv, "depth": depth, "cov3d": cov3d, "conic": conic,
                                  "radius": radius, "ids": ids, "tile_range": tile_range,
                                  "features": render_features, "rendered": rendered,
                                  "T_final": T_final, "n_contrib": n_contrib}
        split = Render_Features.split(rendered)
        if "normal" in split:
            # [-1,1] normals to [0,1] (examples/supervise/renderer.py:89-96)
            nrm = split["normal"]
            split["normal"] = (nrm / nrm.norm(dim=0, keepdim=True) + 1) / 2
        if self.cfg.render_opacity:
            split["opacity"] = (1.0 - T_final)[None]
        return {"rendered_features_split": split, "uv_points": ndc,
                "visibility": radius > 0, "radii": radius}

    def _render_iter_fused(self, height, width, extrinsic_matrix, intrinsic_params,
                           camera_center, position, opacity, scaling, rotation, shs, normals,
                           raw=False, plain_sh=False):
        """The fused path.  ``plain_sh``: colour = the SH sum over every
        coefficient, without the +0.5 offset, the clamp and the warm-up degree
        (MsplatNormalRender, examples/supervise/renderer.py:41-42), and no
        depth channel."""
        normal_in_kernel = normals is None and self.cfg.render_normal and not plain_sh
        depth_channel = self.cfg.render_depth and not plain_sh
        out = ops.preprocess(
            position, scaling, rotation, shs, intrinsic_params, extrinsic_matrix, camera_center,
            width, height, degree=None if plain_sh else self.sh_degree,
            depth_channel=depth_channel, normal_channel=normal_in_kernel,
            keep_cov=self.keep_intermediates, raw=raw, opacity=opacity if raw else None,
            plain_sh=plain_sh)
        uv, depth, conic, radius, tiles, feat, cov3d = out[:7]
        if raw:
            opacity = out[7]
        names = [("rgb", 3)] + ([("depth", 1)] if depth_channel else []) + \
            ([("normal", 3)] if normal_in_kernel else [])
        if normals is not None:
            feat = torch.cat([feat, normals.to(feat.dtype)], -1)
            names.append(("normal", normals.shape[-1]))
        tight = (conic.detach(), opacity.detach()) if self.cfg.tight_tiles else (None, None)
        ids, tile_range = ops.sort_gaussian(uv, depth, width, height, radius, tiles,
                                            workspace=self.workspace, conic=tight[0],
                                            opacity=tight[1])
        ndc = torch.zeros_like(uv, requires_grad=True)
        ndc.retain_grad()
        rendered, T_final, n_contrib = ops.alpha_blending_ex(uv, conic, opacity, feat, ids,
                                                             tile_range, self.bg_color, width,
                                                             height, ndc)
        if self.keep_intermediates:
            self.intermediates = {"uv": uv, "depth": depth, "cov3d": cov3d, "conic": conic,
                                  "radius": radius, "ids": ids, "tile_range": tile_range,
                                  "features": feat, "rendered": rendered, "T_final": T_final,
                                  "n_contrib": n_contrib}
        split = _split_planes(rendered, names)
        if self.cfg.render_opacity:
            split["opacity"] = (1.0 - T_final)[None]
        return {"rendered_features_split": split, "uv_points": ndc,
                "visibility": radius > 0, "radii": radius}

    def render_batch(self, render_dict: dict, batch=None) -> dict:
        """Per-view loop of msplat.py:160-213 (extrinsics and centres per view;
        intrinsics and image size shared)."""
        rendered, uv_points, vis, radii = {}, [], [], []
        per_view = ["extrinsic_matrix", "camera_center"]
        for i in range(render_dict["extrinsic_matrix"].shape[0]):
            kw = {k: (v[i, ...] if k in per_view else v) for k, v in render_dict.items()}
            res = self.render_iter(**kw)
            for name, img in res["rendered_features_split"].items():
                rendered.setdefault(name, []).append(img)
            uv_points.append(res["uv_points"])
            vis.append(res["visibility"].unsqueeze(0))
            radii.append(res["radii"].unsqueeze(0))
        for name in rendered:
            rendered[name] = torch.stack(rendered[name], dim=0)
        return {**rendered, "uv_points": uv_points,
                "visibility": torch.cat(vis).any(dim=0),
                "radii": torch.cat(radii, 0).max(dim=0).values}

    def update_sh_degree(self, step):
        if step % self.cfg.update_sh_iter == 0 and self.sh_degree < self.cfg.max_sh_degree:
            self.sh_degree += 1

    def load_state_dict(self, state_dict):
        self.sh_degree = state_dict["sh_degree"]

    def state_dict(self):
        return {"sh_degree": self.sh_degree}


def _split_planes(rendered, names):
    """{name: planes} of the blended image in channel order, the "normal"
    group mapped to [0,1] as (n / |n| + 1) / 2 (examples/supervise/
    renderer.py:89-96): one fused kernel each way when there is more than one
    group, a plain view for a single group."""
    if len(names) == 1 and names[0][0] != "normal":
        return {names[0][0]: rendered}
    first, nf = 0, -1
    for name, c in names:
        if name == "normal":
            if c != 3:
                nf = -2
                break
            nf = first
        first += c
    if nf == -2 or len(names) > 8:
        split, start = {}, 0
        for name, c in names:
            split[name] = rendered[start:start + c]
            start += c
        nrm = split["normal"]
        split["normal"] = (nrm / nrm.norm(dim=0, keepdim=True) + 1) / 2
        return split
    planes = ops.output_planes(rendered, [c for _, c in names], nf)
    return {name: p for (name, _), p in zip(names, planes)}


@RENDERER_REGISTRY.register()
class MsplatNormalRender(MsplatRender):
    """Normal-supervision variant (examples/supervise/renderer.py:8-102).
    Unlike MsplatRender it takes the colour as the raw SH sum over every
    coefficient (no +0.5, no clamp, no warm-up degree; :41-42), blends only
    rgb and the ``normals`` kwarg (never depth; :73), and maps the blended
    normal to [0,1] (:86-96).  ``render_opacity`` still adds 1 - T_final."""

    def render_iter(self, height, width, extrinsic_matrix, intrinsic_params, camera_center,
                    position, opacity, scaling, rotation, shs, normals=None, raw=False,
                    **kwargs) -> dict:
        """One view.  ``shs`` is [N,K,3], or the pair (features [N,1,3],
        features_rest [N,K-1,3]), which the fused path reads without
        concatenating; ``raw`` as for MsplatRender.render_iter."""
        if isinstance(shs, (tuple, list)) and not self.cfg.fused:
            shs = torch.cat(list(shs), dim=1)
        if self.cfg.fused:
            return self._render_iter_fused(height, width, extrinsic_matrix, intrinsic_params,
                                           camera_center, position, opacity, scaling, rotation,
                                           shs, normals, raw, plain_sh=True)
        if raw:
            opacity, scaling = torch.sigmoid(opacity), torch.exp(scaling)
            rotation = torch.nn.functional.normalize(rotation, dim=-1)
        direction = position - camera_center.reshape(1, 3)
        direction = direction / direction.norm(dim=1, keepdim=True)
        rgb = ops.compute_sh(shs.permute(0, 2, 1), direction)
        extr = extrinsic_matrix[:3, :]
        uv, depth = ops.project_point(position, intrinsic_params, extr, width, height)
        visible = depth != 0
        cov3d = ops.compute_cov3d(scaling, rotation, visible)
        conic, radius, tiles = ops.ewa_project(position, cov3d, intrinsic_params, extr, uv, width,
                                               height, visible)
        ids, tile_range = ops.sort_gaussian(uv, depth, width, height, radius, tiles,
                                            workspace=self.workspace)
        Render_Features = RenderFeatures(rgb=rgb, normal=normals) if normals is not None \
            else RenderFeatures(rgb=rgb)
        ndc = torch.zeros_like(uv, requires_grad=True)
        ndc.retain_grad()
        rendered, T_final, n_contrib = ops.alpha_blending_ex(uv, conic, opacity,
                                                             Render_Features.combine(), ids,
                                                             tile_range, self.bg_color, width,
                                                             height, ndc)
        if self.keep_intermediates:
            self.intermediates = {"uv": uv, "depth": depth, "cov3d": cov3d, "conic": conic,
                                  "radius": radius, "ids": ids, "tile_range": tile_range,
                                  "features": Render_Features.combine(), "rendered": rendered,
                                  "T_final": T_final, "n_contrib": n_contrib}
        split = Render_Features.split(rendered)
        if normals is not None:
            nrm = split["normal"]
            split["normal"] = (nrm / nrm.norm(dim=0, keepdim=True) + 1) / 2
        if self.cfg.render_opacity:
            split["opacity"] = (1.0 - T_final)[None]
        return {"rendered_features_split": split, "uv_points": ndc,
                "visibility": radius > 0, "radii": radius}


def parse_renderer(cfg, **kwargs):
    """Build a renderer from a config carrying ``name``
    (pointrix/model/renderer/__init__.py:4-19): ``name`` is taken out and the
    rest of the mapping becomes the renderer's Config."""
    rest = {k: cfg[k] for k in cfg.keys() if k != "name"}
    name = cfg["name"] if "name" in cfg else "MsplatRender"
    return RENDERER_REGISTRY.get(name)(rest, **kwargs)
