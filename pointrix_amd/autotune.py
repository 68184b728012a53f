"""Autotuner for the batched renderer's launch shape.

Picks the tile shape (16x16, 8x32, 32x8) and the Gaussian batch depth (64,
128 or 256 splats staged per LDS batch) that give the fastest forward +
backward on the actual scene and camera set, by timing each candidate with
device events.  Tile shape changes the binning (instances per splat, tiles
per view) as well as the blend kernels, so whole steps are timed.  Results
are cached per (splat count, resolutions, channels) for the process.
"""
from __future__ import annotations

import itertools
import statistics

import torch

from .batched import BATCH_DEPTHS, TILE_SHAPES, BatchedRender

_cache: dict = {}


def time_step(renderer: BatchedRender, cameras, inputs_fn, loss_fn, params, reps=3, warmup=1):
    """Median milliseconds of forward + loss + backward (gradients of
    ``params`` are reset between steps)."""
    times = []
    for i in range(warmup + reps):
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record()
        loss = loss_fn(renderer.render(cameras, **inputs_fn()))
        loss.backward()
        stop.record()
        torch.cuda.synchronize()
        if i >= warmup:
            times.append(start.elapsed_time(stop))
        for t in params:
            t.grad = None
    return statistics.median(times)


def autotune(cameras, inputs_fn, loss_fn, params, n_splats, render_depth=False,
             render_normal=False, sh_degree=3, tiles=TILE_SHAPES, depths=BATCH_DEPTHS, reps=3,
             verbose=False):
    """-> (best (tile, batch_depth), {(tile, depth): ms})."""
    key = (n_splats, tuple((int(c["width"]), int(c["height"])) for c in cameras), render_depth,
           render_normal, sh_degree)
    if key in _cache:
        return _cache[key]
    timings = {}
    for tile, bd in itertools.product(tiles, depths):
        r = BatchedRender(render_depth=render_depth, render_normal=render_normal,
                          sh_degree=sh_degree, tile=tile, batch_depth=bd)
        timings[(tile, bd)] = time_step(r, cameras, inputs_fn, loss_fn, params, reps=reps)
        if verbose:
            print(f"[autotune] tile {tile[0]}x{tile[1]} depth {bd}: "
                  f"{timings[(tile, bd)]:.3f} ms", flush=True)
    best = min(timings, key=timings.get)
    _cache[key] = (best, timings)
    return best, timings
