"""Densification statistics, fused: the per-step bookkeeping pointrix's
DensificationController does on the renderer's side outputs
(pointrix/controller/gs.py:318-333 ``preprocess`` with
``accumulate_viewspace_grad`` gs.py:259-284; buffers as
``reset_controller_state`` gs.py:249-257) in one HIP kernel per step: no boolean-mask
indexing (which needs a host sync), no [N,2] temporaries.
"""
from __future__ import annotations

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
