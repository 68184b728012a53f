"""Restatement of pointrix's densification bookkeeping (test infrastructure):
accumulate_viewspace_grad (pointrix/controller/gs.py:259-284) and the stat
updates of DensificationController.preprocess (gs.py:318-333), in plain torch
on the CPU, statement for statement."""
import torch


def viewspace_grad(view_grads, width, height, normalize_grad=True):
    """gs.py:271-284: sum of the views' uv_points grads, then scaled."""
    g = torch.zeros_like(view_grads[0])
    for vp in view_grads:
        g += vp.clone()
    if normalize_grad:
        g[..., 0] *= width / 2.0
        g[..., 1] *= height / 2.0
    return g


def update(grad_accum, acc_steps, max_radii, view_grads, visibility, radii, width, height,
           normalize_grad=True):
    """gs.py:318-333 on the given buffers (modified in place)."""
    point_grad = viewspace_grad(view_grads, width, height, normalize_grad)
    sel = visibility
    max_radii[sel] = torch.max(max_radii[sel], radii[sel].to(max_radii.dtype))
    grad_accum[sel] += torch.norm(point_grad[sel, :2], dim=-1, keepdim=True)
    acc_steps[sel] += 1
