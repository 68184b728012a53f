"""Serving loop on the GPU (pointrix_amd.serve).

* a request's images are the same rendered alone and inside a mixed batch
  (other resolutions, depth requested by some), bit for bit;
* a duplicate-key budget too small for the batch splits it, with the same
  images;
* the plain one-view server returns the same images as the batched server;
* a short real-time replay serves every request.
"""
import collections

import pytest
import torch

from pointrix_amd import scene as S
from pointrix_amd.serve import BatchServer, PlainServer, make_requests, run

pytestmark = pytest.mark.gpu

RES = [((320, 180), 0.4), ((256, 144), 0.3), ((480, 270), 0.25), ((640, 360), 0.05)]


def setup(device, n=30000, seed=71):
    scene = S.random_scene(n, seed, sh_degree=3, means="normal", means_std=(2.0, 1.0, 2.0),
                           log_scale=(-3.5, 0.6), opacity_std=1.5, shell_frac=0.1).to(device)
    inputs = scene.render_inputs(split_sh=True)
    reqs = make_requests(100.0, 9, seed=seed, resolutions=RES, depth_frac=0.4)
    return inputs, reqs


def equal(a, b):
    return set(a) == set(b) and all(torch.equal(a[k], b[k]) for k in a)


def test_alone_equals_in_mix(device):
    inputs, reqs = setup(device)
    assert any(r.depth for r in reqs) and not all(r.depth for r in reqs)
    srv = BatchServer(inputs)
    mixed = srv.render(reqs)
    for r, m in zip(reqs, mixed):
        alone = srv.render([r])[0]
        assert equal(alone, m), r.rid


def test_budget_split(device):
    inputs, reqs = setup(device, seed=72)
    full = BatchServer(inputs).render(reqs)
    small = BatchServer(inputs, key_budget_bytes=1)   # every multi-view batch must split
    assert small.max_instances == 1
    out = small.render(reqs)
    assert small.splits >= len(reqs) - 1
    for a, b in zip(out, full):
        assert equal(a, b)


def test_plain_equals_batched(device):
    inputs, reqs = setup(device, seed=73)
    b = BatchServer(inputs).render(reqs)
    p = PlainServer(inputs)
    for r, x in zip(reqs, b):
        y = p.render([r])[0]
        assert equal({k: v.contiguous() for k, v in x.items()},
                     {k: v.contiguous() for k, v in y.items()}), r.rid


def test_realtime_replay(device):
    inputs, _ = setup(device, seed=74)
    reqs = make_requests(400.0, 40, seed=74, resolutions=RES)
    srv = BatchServer(inputs)
    stats, res = run(srv, reqs, keep_results=True)
    assert stats["requests"] == 40 and len(res) == 40
    assert 0 < stats["p50_ms"] <= stats["p99_ms"]
    assert stats["mean_batch"] >= 1.0
    q = collections.Counter(tuple(v["rgb"].shape) for v in res.values())
    assert sum(q.values()) == 40


def test_overlapping_batches(device):
    inputs, reqs = setup(device, seed=75)
    ref = BatchServer(inputs, slots=1).render(reqs)
    srv = BatchServer(inputs, slots=2)
    q = collections.deque()
    srv.submit(reqs[:4], q)
    srv.submit(reqs[4:], q)          # both batches in flight on two streams
    assert srv.free_slot() is None and not q
    got = {r.rid: o for r, o, _ in srv.poll(wait=True)}
    for r, o in zip(reqs, ref):
        assert equal(got[r.rid], o), r.rid


def test_batches_overlap_on_device(device):
    """Continuous batching overlaps on the device: a second batch admitted
    while a long first one is in flight starts on the GPU before the first
    one ends (device events recorded on the two slot streams), and both
    return the images they give rendered alone."""
    scene = S.random_scene(1_000_000, 76, sh_degree=3, means="normal",
                           means_std=(2.0, 1.0, 2.0), log_scale=(-4.0, 0.5), opacity_std=1.0,
                           shell_frac=0.1).to(device)
    inputs = scene.render_inputs(split_sh=True)
    long = make_requests(100.0, 12, seed=76, resolutions=[((1920, 1080), 1.0)],
                         radius=(2.0, 3.0))
    short = make_requests(100.0, 3, seed=77, resolutions=RES)
    for k, r in enumerate(short):
        r.rid = 100 + k
    srv = BatchServer(inputs, slots=2, key_budget_bytes=64 << 30)
    alone = {r.rid: o for r, o in zip(long + short, srv.render(long) + srv.render(short))}
    q = collections.deque()
    srv.submit(long, q)
    a = srv.slots[0]["busy"]
    srv.submit(short, q)
    b = srv.slots[1]["busy"]
    assert a is not None and b is not None and not q
    got = {r.rid: o for r, o, _ in srv.poll(wait=True)}
    a_len = a.start.elapsed_time(a.event)
    b_start = a.start.elapsed_time(b.start)
    print(f"first batch {a_len:.3f} ms on the device; second started {b_start:.3f} ms after "
          f"the first started")
    assert b_start < a_len
    for rid, o in alone.items():
        assert equal(got[rid], o), rid
