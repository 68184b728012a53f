"""Fused Adam: the update of torch.optim.Adam as pointrix configures it
(pointrix/optimizer/__init__.py:61-63 builds ``torch.optim.Adam`` with
per-attribute groups and ``eps: 1e-15``; BaseOptimizer.update_model,
pointrix/optimizer/optimizer.py:128-140, calls ``step`` then
``zero_grad(set_to_none=True)``), with every parameter tensor of the step
updated by one HIP launch.

``FusedAdam`` is a ``torch.optim.Optimizer``: same constructor arguments,
param groups, ``state[p]`` keys (``step``, ``exp_avg``, ``exp_avg_sq``) and
state_dict as torch's Adam, so pointrix's optimizer-state surgery on
densify/prune (point_utils.py ``reduce_opt_by_mask`` / ``extend_opt_by_tensor``
/ ``replace_opt_tensor``) works on it unchanged.
"""
from __future__ import annotations

import ctypes
import math

import torch

from . import _lib
from ._lib import call, ptr, stream_of


class _Entry(ctypes.Structure):
    _fields_ = [("p", ctypes.c_void_p), ("g", ctypes.c_void_p), ("m", ctypes.c_void_p),
                ("v", ctypes.c_void_p), ("n", ctypes.c_int64), ("first_chunk", ctypes.c_int64),
                ("step_size", ctypes.c_float), ("bc2_sqrt", ctypes.c_float)]


class FusedAdam(torch.optim.Optimizer):
    """Adam (weight decay 0, no amsgrad) with one kernel launch per step for
    all parameters sharing (betas, eps).  ``zero_grad_in_step``: the kernel
    also writes the gradients to zero (saves the separate pass; leave False
    where ``zero_grad(set_to_none=True)`` follows anyway)."""

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.0,
                 zero_grad_in_step=False):
        if weight_decay != 0.0:
            raise ValueError("FusedAdam: weight decay is not supported")
        super().__init__(params, dict(lr=lr, betas=tuple(betas), eps=eps))
        self.zero_grad_in_step = zero_grad_in_step
        if int(_lib.load().gsr_adam_entry_bytes()) != ctypes.sizeof(_Entry):
            raise RuntimeError("FusedAdam: table layout differs from the library's")
        self._chunk = int(_lib.load().gsr_adam_chunk())

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        launches = {}
        for group in self.param_groups:
            b1, b2 = group["betas"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                if p.dtype != torch.float32 or not p.is_cuda:
                    raise RuntimeError("FusedAdam: float32 GPU parameters only")
                if not (p.is_contiguous() and p.grad.is_contiguous()):
                    raise RuntimeError("FusedAdam: contiguous parameters and gradients only")
                st = self.state[p]
                if len(st) == 0:
                    st["step"] = torch.tensor(0.0)
                    st["exp_avg"] = torch.zeros_like(p, memory_format=torch.contiguous_format)
                    st["exp_avg_sq"] = torch.zeros_like(p,
                                                        memory_format=torch.contiguous_format)
                st["step"] += 1
                t = float(st["step"])
                key = (float(b1), float(b2), float(group["eps"]))
                launches.setdefault(key, []).append(
                    (p, p.grad, st["exp_avg"], st["exp_avg_sq"],
                     group["lr"] / (1.0 - b1 ** t), math.sqrt(1.0 - b2 ** t)))
        for (b1, b2, eps), items in launches.items():
            self._launch(items, b1, b2, eps)
        return loss

    def _launch(self, items, b1, b2, eps):
        dev = items[0][0].device
        table = (_Entry * len(items))()
        first = 0
        for k, (p, g, m, v, step_size, bc2) in enumerate(items):
            n = p.numel()
            table[k] = _Entry(ptr(p), ptr(g), ptr(m), ptr(v), n, first, step_size, bc2)
            first += (n + self._chunk - 1) // self._chunk
        raw = torch.frombuffer(bytearray(table), dtype=torch.uint8).to(dev, non_blocking=False)
        call("gsr_adam_step", ptr(raw), len(items), first, b1, b2, eps,
             1 if self.zero_grad_in_step else 0, stream_of(items[0][0]))
        self._last_table = raw  # keep alive until the launch has run
