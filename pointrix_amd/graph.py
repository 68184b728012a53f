"""Whole training steps (forward + loss + backward) captured as one HIP graph.

A Gaussian-splatting step launches a few hundred kernels (the renderer's
stages plus the small tensor ops of activations, camera composition and
losses).  Replaying them as one captured graph removes the per-launch host
cost and the host read of the instance count that sizes the binning buffers
in eager mode.

The instance count is data-dependent, so a captured step sizes the binning
buffers for a fixed capacity (the count seen at capture times a headroom);
the emission kernels never write past it, and ``replay()`` reads the true
count afterwards: when it outgrew the capacity, the step is captured again
at the new size and replayed, so every result the caller sees is complete.

The object that sizes the buffers (the "sizer": a renderer's sort workspace,
or the batched renderer) exposes ``capacity`` (None = exact, with a host
read), ``instances()`` (host read of the device count of the last forward)
and ``count_tensor()`` (that device count, an int32[1] tensor).  A sizer may
also expose ``high_water_tensor()``: the largest count of its fixed-capacity
sorts since it was last zeroed, kept by the library's count kernel, which
then replaces the step's own tracking kernel.
"""
from __future__ import annotations

import math

import torch


class CapturedStep:
    """``step_fn() -> loss`` runs forward + loss + ``backward()`` into the
    ``.grad`` of ``leaves`` ({name: tensor}).

    The gradients are not zeroed and accumulated: every leaf's ``.grad`` is
    None when the captured step starts, so the backward writes fresh gradient
    buffers (from the graph's memory pool) instead of reading, adding and
    writing back zero-filled ones.  They are held here and bound back to the
    leaves' ``.grad`` on every replay, so eager steps in between (which may
    replace ``.grad``) never leave the graph writing into freed memory.

    The graph also keeps the largest instance count it has seen on the device
    (the sizer's high-water word when it has one, else one small kernel).  ``replay(check=True)`` reads it after the replay (a
    host wait) and re-captures when it outgrew the capacity;
    ``replay(check=False)`` does not wait, and ``overflowed()`` tells later
    whether any unchecked replay since the last check overflowed (its
    results are then incomplete and the step has been captured again)."""

    def __init__(self, step_fn, leaves: dict, sizer, headroom=1.25, min_capacity=1024):
        self.step_fn, self.leaves, self.sizer = step_fn, leaves, sizer
        self.headroom, self.min_capacity = headroom, min_capacity
        self.graph = None
        self.loss = None
        self.grads = {}
        self.captures = 0
        self.max_count = None
        self.own_max = True  # False: the sizer's high-water word is read instead

    def _step(self):
        for p in self.leaves.values():
            p.grad = None
        loss = self.step_fn()
        if self.own_max:
            m = self.sizer.count_tensor()
            if m is not None:
                torch.maximum(self.max_count, m, out=self.max_count)
        return loss

    def _high(self):
        hw = getattr(self.sizer, "high_water_tensor", None)
        return hw() if hw is not None else None

    def capture(self):
        sz = self.sizer
        sz.capacity = None
        dev = next(iter(self.leaves.values())).device
        self.max_count = torch.zeros(1, dtype=torch.int32, device=dev)
        self.own_max = True
        self._step()                                   # eager: exact instance count
        torch.cuda.synchronize()
        m = sz.instances()
        sz.capacity = max(self.min_capacity, int(math.ceil(m * self.headroom)))
        self._step()                                   # fixed capacity: creates the high-water word
        self.own_max = self._high() is None
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(2):                         # warm the allocator on the side stream
                self._step()
        torch.cuda.current_stream().wait_stream(s)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.loss = self._step()
        self.grads = {k: p.grad for k, p in self.leaves.items()}
        self._reset_max()
        self.captures += 1
        return self

    def _max_tensor(self):
        return self.max_count if self.own_max else self._high()

    def _reset_max(self):
        self._max_tensor().zero_()

    def overflowed(self) -> bool:
        """Host read of the largest instance count seen since the last check:
        True (and the step captured again at the new size) when it exceeded
        the capacity."""
        m = int(self._max_tensor().item())
        self._reset_max()
        if m > self.sizer.capacity:
            self.capture()
            return True
        return False

    def replay(self, check=True) -> bool:
        """Run the captured step.  With ``check`` (a host wait), returns False
        when the instance count had outgrown the capacity (the step was then
        captured again and replayed, so the results are complete either way);
        without, returns True at once (see ``overflowed``)."""
        for k, p in self.leaves.items():
            if p.grad is not self.grads[k]:
                p.grad = self.grads[k]
        self.graph.replay()
        if check and self.overflowed():
            for k, p in self.leaves.items():
                p.grad = self.grads[k]
            self.graph.replay()
            return False
        return True
