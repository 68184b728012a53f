"""Densification statistics, fused: the per-step bookkeeping pointrix's
DensificationController does on the renderer's side outputs
(pointrix/controller/gs.py:318-333 ``preprocess`` with
``accumulate_viewspace_grad`` gs.py:259-284; buffers as
``reset_controller_state`` gs.py:249-257) in one HIP kernel per step: no boolean-mask
indexing (which needs a host sync), no [N,2] temporaries.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import call, ptr, stream_of


class DensifyStats:
    """grad_accum [N,1], acc_steps [N,1], max_radii [N] (float32, device).

    ``update(uv_points, visibility, radii)`` is ``preprocess(**render_out)``:
    uv_points is render_iter's leaf (its ``.grad`` [N,2]), render_batch's list
    of them, or a batched [V,N,2] leaf; visibility [N] bool; radii [N].
    normalize_grad scales the summed gradient by (width/2, height/2)."""

    def __init__(self, num_points: int, device, width: int, height: int,
                 normalize_grad: bool = True):
        self.width, self.height = width, height
        self.normalize_grad = normalize_grad
        self.device = torch.device(device)
        self.reset(num_points)

    def reset(self, num_points: int):
        self.grad_accum = torch.zeros(num_points, 1, device=self.device)
        self.acc_steps = torch.zeros(num_points, 1, device=self.device)
        self.max_radii = torch.zeros(num_points, device=self.device)

    def update(self, uv_points, visibility, radii):
        if isinstance(uv_points, (list, tuple)):
            grads = [u.grad.reshape(-1, 2) for u in uv_points]
            g = grads[0] if len(grads) == 1 else torch.stack(grads)
        else:
            g = uv_points.grad if isinstance(uv_points, torch.Tensor) and \
                uv_points.grad is not None else uv_points
        g = g.float()
        if g.dim() == 2:
            g = g[None]
        if g.stride(-1) != 1:
            g = g.contiguous()
        V, n = g.shape[0], g.shape[1]
        if n != self.grad_accum.shape[0]:
            raise ValueError(f"densify stats hold {self.grad_accum.shape[0]} splats, got {n}")
        vis = visibility.reshape(-1).to(torch.uint8).contiguous()
        rad = radii.reshape(-1).to(torch.int32).contiguous()
        sx, sy = (self.width / 2.0, self.height / 2.0) if self.normalize_grad else (1.0, 1.0)
        call("gsr_densify_stats", ptr(g), g.stride(0) if V > 1 else 0, g.stride(1), V, n,
             ptr(vis), ptr(rad), sx, sy, ptr(self.grad_accum), ptr(self.acc_steps),
             ptr(self.max_radii), stream_of(g))

    def average_grad(self) -> torch.Tensor:
        """grad_accum / acc_steps with NaN (never visible) set to 0
        (pointrix/controller/base.py:91-92)."""
        avg = self.grad_accum / self.acc_steps
        avg[avg.isnan()] = 0.0
        return avg


# ---------------------------------------------------------------------------
# densify / prune with optimizer-state remap, as one row gather per call
# ---------------------------------------------------------------------------
class _RowCopy(ctypes.Structure):
    _fields_ = [("src", ctypes.c_void_p), ("dst", ctypes.c_void_p), ("width", ctypes.c_int64)]


def _gather(pairs, index: torch.Tensor, rows: int):
    """[(src [*, ...], dst [rows, ...])]: dst row r = src row index[r] or 0."""
    if int(_lib.load().gsr_row_copy_bytes()) != ctypes.sizeof(_RowCopy):
        raise RuntimeError("gather_rows: table layout differs from the library's")
    if not pairs or rows == 0:
        return
    table = (_RowCopy * len(pairs))()
    for k, (src, dst) in enumerate(pairs):
        width = dst[0].numel() if dst.dim() > 1 else 1
        if rows * width >= 2 ** 32:
            raise ValueError("gather_rows: too many elements")
        table[k] = _RowCopy(ptr(src), ptr(dst), width)
    raw = torch.frombuffer(bytearray(table), dtype=torch.uint8).to(index.device)
    call("gsr_gather_rows", ptr(raw), len(pairs), ptr(index), rows, stream_of(index))
    _gather.keep = raw


def _swap_param(optimizer, old, new, state):
    """Put ``new`` in ``old``'s place in the param groups with ``state``."""
    for group in optimizer.param_groups:
        for k, p in enumerate(group["params"]):
            if p is old:
                group["params"][k] = new
    if old in optimizer.state:
        del optimizer.state[old]
    if state is not None:
        optimizer.state[new] = state


def remap_points(params: dict, optimizer, index: torch.Tensor, overrides: dict | None = None,
                 fresh: int = 0) -> dict:
    """New parameter set whose row r is old row index[r]; the last ``fresh``
    rows are new points: their Adam moments start at zero (pointrix's
    extend_opt_by_tensor), while every other row keeps its moments
    (reduce_opt_by_mask).  ``overrides`` {name: [fresh, ...]} replaces the
    values of the new rows of a parameter (split positions / scalings).
    One gather launch moves every parameter and moment.  Returns the new
    leaves {name: nn.Parameter}, already swapped into ``optimizer``."""
    overrides = overrides or {}
    idx = index.to(torch.int32).contiguous()
    rows = idx.numel()
    sidx = idx.clone()
    if fresh:
        sidx[rows - fresh:] = -1
    pairs_p, pairs_s, out, states = [], [], {}, {}
    for name, p in params.items():
        new = torch.empty((rows,) + tuple(p.shape[1:]), device=p.device, dtype=p.dtype)
        pairs_p.append((p.detach().contiguous(), new))
        out[name] = new
        st = optimizer.state.get(p) if optimizer is not None else None
        if st is not None and "exp_avg" in st:
            nst = dict(st)
            for key in ("exp_avg", "exp_avg_sq"):
                nst[key] = torch.empty_like(new)
                pairs_s.append((st[key].contiguous(), nst[key]))
            states[name] = nst
        else:
            states[name] = st
    _gather(pairs_p, idx, rows)
    _gather(pairs_s, sidx, rows)
    leaves = {}
    for name, p in params.items():
        if name in overrides and fresh:
            out[name][rows - fresh:] = overrides[name].to(out[name].dtype)
        leaf = torch.nn.Parameter(out[name].requires_grad_(True))
        if optimizer is not None:
            _swap_param(optimizer, p, leaf, states[name])
        leaves[name] = leaf
    return leaves


def densify(params: dict, optimizer, grads: torch.Tensor, scaling_act: torch.Tensor,
            threshold: float, percent_dense: float, cameras_extent: float, split_num: int = 2,
            samples: torch.Tensor | None = None, generator=None):
    """DensificationController.densify_clone + densify_split
    (controller/gs.py:91-241) fused: clone small high-gradient splats, split
    large ones into ``split_num`` samples (position drawn from the splat's
    Gaussian, scale / (0.8 split_num)), drop the split originals.  Final row
    order is pointrix's: [unsplit originals] + [clones] + [split samples,
    copy-major].  ``grads`` [N,1] are the averaged viewspace gradients,
    ``scaling_act`` the activated scales; ``samples`` [split_num * Ns, 3] may
    be given (standard normals) to reproduce a run.  Returns (leaves, masks)."""
    g = grads.reshape(-1)
    big = scaling_act.max(dim=1).values > percent_dense * cameras_extent
    clone_mask = (g.abs() >= threshold) & ~big
    split_mask = (g >= threshold) & big
    keep_idx = torch.nonzero(~split_mask).reshape(-1)
    clone_idx = torch.nonzero(clone_mask).reshape(-1)
    split_idx = torch.nonzero(split_mask).reshape(-1)
    ns = split_idx.numel()
    rep = split_idx.repeat(split_num)
    index = torch.cat([keep_idx, clone_idx, rep])
    fresh = clone_idx.numel() + rep.numel()
    overrides = {}
    if ns:
        from .renderer import quat_to_rotmat
        stds = scaling_act[split_idx].repeat(split_num, 1)
        if samples is None:
            samples = torch.randn(stds.shape, device=stds.device, generator=generator)
        r = params["rotation"].detach()[split_idx]
        r = r / torch.sqrt(r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1] + r[:, 2] * r[:, 2] +
                           r[:, 3] * r[:, 3])[:, None]   # pose.py:86-117 normalises
        rots = quat_to_rotmat(r).repeat(split_num, 1, 1)
        new_pos = torch.bmm(rots, (samples * stds).unsqueeze(-1)).squeeze(-1) + \
            params["position"].detach()[split_idx].repeat(split_num, 1)
        new_scaling = torch.log(stds / (0.8 * split_num))
        pos_over = torch.cat([params["position"].detach()[clone_idx], new_pos])
        sc_over = torch.cat([params["scaling"].detach()[clone_idx], new_scaling])
        overrides = {"position": pos_over, "scaling": sc_over}
    leaves = remap_points(params, optimizer, index, overrides, fresh)
    return leaves, {"clone": clone_mask, "split": split_mask}


def prune(params: dict, optimizer, keep_mask: torch.Tensor, stats: DensifyStats | None = None):
    """PointCloud.remove_points + prune_postprocess (controller/gs.py:196-247):
    keep the rows where keep_mask holds, parameters and Adam moments in one
    gather, and the densification statistics with them."""
    index = torch.nonzero(keep_mask.reshape(-1)).reshape(-1)
    leaves = remap_points(params, optimizer, index)
    if stats is not None:
        stats.grad_accum = stats.grad_accum[index]
        stats.acc_steps = stats.acc_steps[index]
        stats.max_radii = stats.max_radii[index]
    return leaves
