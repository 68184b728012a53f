"""Serving views of a resident scene: continuous batching under a
duplicate-key memory budget, with batches overlapping on the device.

A request asks for one view (its own resolution and camera) of the scene
held in HBM, rgb and optionally depth.  ``BatchServer`` keeps a few slots,
each a HIP stream with its own renderers and sort workspace.  Whenever a
slot is free, the requests that have arrived are admitted in arrival order
while the batch's estimated (splat, tile) instance count -- the duplicated
sort keys, the memory that grows with the views -- fits the budget left over
by the batches still in flight, and the batch is launched forward-only
through BatchedRender (one launch per stage for all its views) on the slot's
stream.  The next batch is admitted and launched while earlier ones are
still on the GPU, so their work overlaps on the device.  The estimate is
pixels times the running instances-per-pixel of past batches; the renderer
counts the exact instances before it allocates them, and a batch that still
exceeds the budget returns its second half to the head of the queue, so the
budget holds.

``PlainServer`` is the one-view-per-call baseline: each request is one
MsplatRender.render_iter call (pointrix/model/renderer/msplat.py:51-158), in
arrival order, one at a time.

``run`` replays a request trace in real time (arrival offsets from the
start) against either server and reports latency from arrival to the moment
the request's images are complete on the device (a device event, placed on
the host clock through an event recorded at the start of the replay).
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass

import torch

from . import _lib, ops
from . import scene as S
from .batched import BatchedRender, InstanceBudgetExceeded, ViewLayout, view_floats
from .renderer import MsplatRender

MAX_VIEWS = 64  # views per launch supported by the multi-view kernels


@dataclass
class Request:
    rid: int
    arrival: float              # seconds after the start of the trace
    width: int
    height: int
    extrinsic_matrix: torch.Tensor   # [4,4] world-to-camera, host
    intrinsic_params: torch.Tensor   # [4] fx, fy, cx, cy, host
    camera_center: torch.Tensor      # [3], host
    depth: bool

    @property
    def pixels(self) -> int:
        return self.width * self.height

    def camera(self) -> dict:
        return {"width": self.width, "height": self.height,
                "extrinsic_matrix": self.extrinsic_matrix,
                "intrinsic_params": self.intrinsic_params, "camera_center": self.camera_center}


def make_requests(rate: float, n: int, seed: int = 0, resolutions=None, depth_frac: float = 0.2,
                  radius=(2.0, 8.0), focal_per_width: float = 1.1):
    """``n`` requests with Poisson arrivals at ``rate`` per second; resolution
    drawn from ``resolutions`` [((W, H), probability)], eye uniform in
    direction at a distance U[radius] looking at the origin, focal =
    focal_per_width * W, depth requested with probability ``depth_frac``."""
    resolutions = resolutions or S.SERVE_RESOLUTIONS
    g = torch.Generator().manual_seed(seed)
    probs = torch.tensor([p for _, p in resolutions], dtype=torch.float64)
    cdf = torch.cumsum(probs / probs.sum(), 0)
    t, out = 0.0, []
    for rid in range(n):
        u = float(torch.rand(1, generator=g, dtype=torch.float64))
        t += -math.log(max(u, 1e-300)) / rate
        pick = float(torch.rand(1, generator=g, dtype=torch.float64))
        k = min(int(torch.searchsorted(cdf, torch.tensor([pick], dtype=torch.float64))),
                len(resolutions) - 1)
        w, h = resolutions[k][0]
        eye = S.random_sphere_eye(g, radius[0], radius[1])
        cam = S.make_camera(w, h, eye=eye, focal=focal_per_width * w)
        depth = float(torch.rand(1, generator=g)) < depth_frac
        out.append(Request(rid, t, w, h, cam.extrinsic_matrix, cam.intrinsic_params,
                           cam.camera_center, depth))
    return out


def instance_bytes() -> int:
    """Device bytes per (splat, tile) instance of a render: the sort's key and
    value double buffers plus the output id."""
    m = 1 << 24
    return int(_lib.load().gsr_sort_ws_inst_bytes(m)) // m + 4


class _Inflight:
    """A launched batch: its requests, results (device views), end event,
    instance count and (for BatchServer) the event recorded on the slot's
    stream before its first kernel."""

    def __init__(self, reqs, results, event, instances, start=None):
        self.reqs, self.results, self.event, self.instances = reqs, results, event, instances
        self.start = start


class BatchServer:
    """Continuous batching (see the module docstring).  ``inputs``: the
    resident scene as render_iter takes it (activated tensors on the device;
    ``shs`` may be the (dc, rest) pair).  ``slots`` batches may be in flight
    at once, each on its own stream.  Batches of at least ``tight_min_views``
    views use opacity-aware binning (the same images with fewer duplicated
    keys; it pays off on large batches, its extra tests cost more than they
    save on a few views)."""

    def __init__(self, inputs: dict, sh_degree=3, key_budget_bytes=4 << 30, max_views=MAX_VIEWS,
                 slots=2, tile=(16, 16), batch_depth=256, white_bg=False, tight_min_views=4):
        self.inputs = inputs
        self.dev = inputs["position"].device
        self.slots = []
        for _ in range(slots):
            rs = {depth: BatchedRender(render_depth=depth, white_bg=white_bg,
                                       sh_degree=sh_degree, tile=tile, batch_depth=batch_depth)
                  for depth in (False, True)}

            # a batch of one view runs through the one-view renderer (the same
            # images: a view of a batch equals that view rendered alone), which
            # skips the multi-view tables and queue
            single = {}
            for depth in (False, True):
                r = MsplatRender({"render_depth": depth, "max_sh_degree": sh_degree},
                                 white_bg=white_bg, device="cuda")
                r.sh_degree = sh_degree
                single[depth] = r
            self.slots.append({"stream": torch.cuda.Stream(device=self.dev), "renderers": rs,
                               "single": single, "busy": None})
        self.max_views = min(max_views, MAX_VIEWS)
        self.tight_min_views = tight_min_views
        self.max_instances = max(1, key_budget_bytes // instance_bytes())
        self.per_pixel = None      # running instances per pixel (estimate)
        self.splits = 0
        self.batch_sizes = []

    # -- admission -----------------------------------------------------------
    def estimate(self, req: Request) -> float:
        return req.pixels * (self.per_pixel if self.per_pixel is not None else 0.0)

    def in_flight_instances(self) -> int:
        return sum(sl["busy"].instances for sl in self.slots if sl["busy"] is not None)

    def free_slot(self):
        for sl in self.slots:
            if sl["busy"] is None:
                return sl
        return None

    def take(self, queue: collections.deque) -> list:
        """Admit queued requests in arrival order while the batch fits the
        memory left by the batches in flight (at least one request)."""
        room = self.max_instances - self.in_flight_instances()
        batch = [queue.popleft()]
        est = self.estimate(batch[0])
        while queue and len(batch) < self.max_views:
            e = self.estimate(queue[0])
            if est + e > room:
                break
            est += e
            batch.append(queue.popleft())
        return batch

    # -- launch / completion -------------------------------------------------
    def submit(self, batch: list, queue: collections.deque, slot=None) -> None:
        """Launch ``batch`` on a free slot.  If its exact instance count does
        not fit the memory left, its second half goes back to the head of
        ``queue`` and the first half is tried (a single view always runs)."""
        sl = slot or self.free_slot()
        assert sl is not None and sl["busy"] is None, "no free slot"
        if len(batch) == 1:
            self._submit_one(batch[0], sl)
            return
        while True:
            need_depth = any(r.depth for r in batch)
            rd = sl["renderers"][need_depth]
            rd.tight_tiles = len(batch) >= self.tight_min_views
            room = self.max_instances - self.in_flight_instances()
            rd.max_instances = max(room, 1) if len(batch) > 1 else None
            lay = ViewLayout([r.width for r in batch], [r.height for r in batch], rd.tile[0],
                             rd.tile[1])
            try:
                with torch.cuda.stream(sl["stream"]), torch.no_grad():
                    start = torch.cuda.Event(enable_timing=True)
                    start.record()
                    vf = view_floats([r.camera() for r in batch]).to(self.dev, non_blocking=True)
                    out, _, _, _, _ = rd.render_packed(None, **self.inputs, layout=lay, vf=vf)
                    ev = torch.cuda.Event(enable_timing=True)
                    ev.record()
                    # the count the renderer read (ahead of its depth sort) to
                    # size the instance buffers: no further host wait, so the
                    # next batch is admitted while this one is on the GPU
                    m = rd.last_instances
            except InstanceBudgetExceeded as e:
                self.splits += 1
                self._learn(e.needed, lay.total_pixels)
                half = len(batch) // 2
                queue.extendleft(reversed(batch[half:]))
                batch = batch[:half]
                continue
            break
        self._learn(m, lay.total_pixels)
        res = []
        for v, r in enumerate(batch):
            img = lay.unpack(out, v)
            d = {"rgb": img[0:3]}
            if r.depth:
                d["depth"] = img[3:4]
            res.append(d)
        sl["busy"] = _Inflight(batch, res, ev, m, start)
        self.batch_sizes.append(len(batch))

    def _submit_one(self, r: Request, sl) -> None:
        with torch.cuda.stream(sl["stream"]), torch.no_grad():
            start = torch.cuda.Event(enable_timing=True)
            start.record()
            o = sl["single"][r.depth].render_iter(
                r.height, r.width, r.extrinsic_matrix.to(self.dev, non_blocking=True),
                r.intrinsic_params.to(self.dev, non_blocking=True),
                r.camera_center.to(self.dev, non_blocking=True), **self.inputs)
            ev = torch.cuda.Event(enable_timing=True)
            ev.record()
        # the count sort_gaussian read ahead of its depth sort
        m = ops.last_instance_count
        self._learn(m, r.pixels)
        f = o["rendered_features_split"]
        res = {k: f[k] for k in (("rgb", "depth") if r.depth else ("rgb",))}
        sl["busy"] = _Inflight([r], [res], ev, m, start)
        self.batch_sizes.append(1)

    def poll(self, wait=False) -> list:
        """-> [(request, result, end event)] of the batches that completed
        (with ``wait``, of every batch in flight)."""
        done = []
        for sl in self.slots:
            b = sl["busy"]
            if b is not None and (wait or b.event.query()):
                b.event.synchronize()
                done.extend((r, o, b.event) for r, o in zip(b.reqs, b.results))
                sl["busy"] = None
        return done

    def render(self, batch: list) -> list:
        """Synchronous helper: render ``batch`` (splitting as the budget
        requires) and return the results in request order."""
        queue = collections.deque(batch)
        got = {}
        while queue:
            if self.free_slot() is None:
                for r, o, _ in self.poll(wait=True):
                    got[r.rid] = o
            sub = [queue.popleft() for _ in range(min(len(queue), self.max_views))]
            self.submit(sub, queue)
        for r, o, _ in self.poll(wait=True):
            got[r.rid] = o
        return [got[r.rid] for r in batch]

    def _learn(self, instances: int, pixels: int):
        rate = instances / max(pixels, 1)
        self.per_pixel = rate if self.per_pixel is None else 0.7 * self.per_pixel + 0.3 * rate


class PlainServer:
    """One render_iter call per request, in arrival order, one at a time."""

    def __init__(self, inputs: dict, sh_degree=3, white_bg=False):
        self.inputs = inputs
        self.dev = inputs["position"].device
        self.renderers = {}
        for depth in (False, True):
            r = MsplatRender({"render_depth": depth, "max_sh_degree": sh_degree},
                             white_bg=white_bg, device="cuda")
            r.sh_degree = sh_degree
            self.renderers[depth] = r
        self.busy = None
        self.batch_sizes = []

    def free_slot(self):
        return self if self.busy is None else None

    def take(self, queue: collections.deque) -> list:
        return [queue.popleft()]

    def submit(self, batch: list, queue=None, slot=None) -> None:
        res = []
        for r in batch:
            with torch.no_grad():
                o = self.renderers[r.depth].render_iter(
                    r.height, r.width, r.extrinsic_matrix.to(self.dev),
                    r.intrinsic_params.to(self.dev), r.camera_center.to(self.dev), **self.inputs)
            f = o["rendered_features_split"]
            res.append({k: f[k] for k in (("rgb", "depth") if r.depth else ("rgb",))})
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        self.busy = _Inflight(batch, res, ev, 0)
        self.batch_sizes.append(len(batch))

    def poll(self, wait=False) -> list:
        b = self.busy
        if b is None or not (wait or b.event.query()):
            return []
        b.event.synchronize()
        self.busy = None
        return [(r, o, b.event) for r, o in zip(b.reqs, b.results)]

    def render(self, batch: list) -> list:
        out = []
        for r in batch:
            self.submit([r])
            out.extend(o for _, o, _ in self.poll(wait=True))
        return out


def percentile(xs, q):
    s = sorted(xs)
    if not s:
        return float("nan")
    k = (len(s) - 1) * q
    lo, hi = math.floor(k), math.ceil(k)
    return s[lo] + (s[hi] - s[lo]) * (k - lo)


def run(server, requests, keep_results=False):
    """Replay ``requests`` in real time against ``server``.  Returns latency
    statistics (ms; overall and per resolution) and throughput; with
    keep_results, also {rid: images}."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    queue = collections.deque()
    lat, by_res, results = [], collections.defaultdict(list), {}
    first_batch = len(server.batch_sizes)
    i, pending = 0, len(reqs)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    start.record()
    start.synchronize()
    t0 = time.perf_counter()

    def finish(done):
        nonlocal pending
        for r, o, ev in done:
            ms = start.elapsed_time(ev) - r.arrival * 1000.0
            lat.append(ms)
            by_res[f"{r.width}x{r.height}"].append(ms)
            if keep_results:
                results[r.rid] = o
            pending -= 1

    while pending:
        now = time.perf_counter() - t0
        while i < len(reqs) and reqs[i].arrival <= now:
            queue.append(reqs[i])
            i += 1
        finish(server.poll())
        if queue and server.free_slot() is not None:
            server.submit(server.take(queue), queue)
            continue
        # nothing to launch: wait for the next arrival or completion
        if i < len(reqs) and not queue:
            wait = reqs[i].arrival - (time.perf_counter() - t0)
            if wait > 1e-3:
                time.sleep(min(wait - 5e-4, 2e-3))
        else:
            time.sleep(5e-5)
    end = time.perf_counter() - t0
    n = len(reqs)
    sizes = server.batch_sizes[first_batch:]
    stats = {"requests": n, "p50_ms": percentile(lat, 0.5), "p99_ms": percentile(lat, 0.99),
             "mean_ms": sum(lat) / max(n, 1), "views_per_s": n / end,
             "batches": len(sizes), "mean_batch": sum(sizes) / max(len(sizes), 1),
             "max_batch": max(sizes) if sizes else 0, "wall_s": end,
             "by_resolution": {k: {"n": len(v), "p50_ms": percentile(v, 0.5),
                                   "p99_ms": percentile(v, 0.99)}
                               for k, v in sorted(by_res.items())}}
    return (stats, results) if keep_results else stats
