"""Request traces for the serving loop (host only): Poisson arrivals at the
requested rate, the resolution mix, the depth fraction, seeded."""
import math

from pointrix_amd import scene as S
from pointrix_amd.serve import make_requests, percentile


def test_trace_statistics():
    n, rate = 4000, 200.0
    reqs = make_requests(rate, n, seed=3)
    assert len(reqs) == n
    arr = [r.arrival for r in reqs]
    assert all(b > a for a, b in zip(arr, arr[1:]))
    # mean inter-arrival 1/rate (Poisson); std error of the mean ~ 1/(rate sqrt n)
    mean_gap = arr[-1] / n
    assert abs(mean_gap * rate - 1.0) < 5.0 / math.sqrt(n)
    counts = {}
    for r in reqs:
        counts[(r.width, r.height)] = counts.get((r.width, r.height), 0) + 1
    for (wh, p) in S.SERVE_RESOLUTIONS:
        f = counts.get(wh, 0) / n
        assert abs(f - p) < 5 * math.sqrt(p * (1 - p) / n), (wh, f, p)
    fd = sum(r.depth for r in reqs) / n
    assert abs(fd - 0.2) < 5 * math.sqrt(0.16 / n)
    for r in reqs[:50]:
        d = float(r.camera_center.norm())
        assert 2.0 - 1e-4 <= d <= 8.0 + 1e-4
        assert abs(float(r.intrinsic_params[0]) - 1.1 * r.width) < 1e-3


def test_trace_seeded():
    a = make_requests(50.0, 20, seed=9)
    b = make_requests(50.0, 20, seed=9)
    assert [(r.arrival, r.width, r.depth) for r in a] == [(r.arrival, r.width, r.depth) for r in b]


def test_percentile():
    xs = list(range(101))
    assert percentile(xs, 0.5) == 50
    assert percentile(xs, 0.99) == 99
    assert percentile([3.0], 0.99) == 3.0
