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
read) and ``instances()`` (the device count of the last forward).
"""
from __future__ import annotations

import math

import torch


class CapturedStep:
    """``step_fn() -> loss`` runs forward + loss + ``backward()`` into the
    ``.grad`` of ``leaves`` ({name: tensor}).  Gradients are zeroed inside the
    graph.  The graph writes into the gradient tensors that existed at
    capture; they are held here and bound back to the leaves' ``.grad`` on
    every replay, so eager steps in between (which may replace ``.grad``)
    never leave the graph writing into freed memory."""

    def __init__(self, step_fn, leaves: dict, sizer, headroom=1.25, min_capacity=1024):
        self.step_fn, self.leaves, self.sizer = step_fn, leaves, sizer
        self.headroom, self.min_capacity = headroom, min_capacity
        self.graph = None
        self.loss = None
        self.grads = {}
        self.captures = 0

    def _step(self):
        for p in self.leaves.values():
            if p.grad is not None:
                p.grad.zero_()
        return self.step_fn()

    def capture(self):
        sz = self.sizer
        sz.capacity = None
        self._step()                                   # eager: exact instance count
        torch.cuda.synchronize()
        m = sz.instances()
        sz.capacity = max(self.min_capacity, int(math.ceil(m * self.headroom)))
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
        self.captures += 1
        return self

    def replay(self) -> bool:
        """Run the captured step.  Returns False when the instance count had
        outgrown the capacity (the step was then captured again and replayed,
        so the results are complete either way)."""
        for k, p in self.leaves.items():
            if p.grad is not self.grads[k]:
                p.grad = self.grads[k]
        self.graph.replay()
        if self.sizer.instances() > self.sizer.capacity:
            self.capture()
            self.graph.replay()
            return False
        return True
