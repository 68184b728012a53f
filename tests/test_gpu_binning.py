"""Exact binning on the GPU in its three implementations, each bit-exact
against the FP32 C restatement (oracle/ref32.c; msplat.sort_gaussian's
contract of msplat.py:133-135: ids ordered by (tile, f32 depth bits, splat
index), tile_range = [first, last+1), (0, 0) for empty tiles) and so equal to
each other: mode 2 (default) bins grids of at most 256 x 256 tiles by rows
(segments sorted by row, then expanded and sorted by column inside each
row), mode 1 generates the first tile radix pass's items from the
depth-ordered splats, mode 0 sorts the emitted instance list.

Cases: random splats with off-screen rectangles, radius 0 and depth ties;
the full 64x64 = 4096-tile grid, a 1080p grid and a grid 263 tiles wide
(beyond row binning: mode 2 runs as mode 1); grids one tile wide and one
tile high; large near splats whose instance runs cross many 1024-instance
wave ranges and 4096-instance radix blocks; one splat; a one-tile image (no
tile pass); an instance capacity below the count (the in-bounds prefix a
captured graph gets)."""
import numpy as np
import pytest
import torch

from oracle import ref32 as r32
from pointrix_amd import ops
from tests.test_gpu_parity import _emission_order_reference

pytestmark = pytest.mark.gpu


def _random_case(n, W, H, seed, ties, big=0):
    g = torch.Generator().manual_seed(seed)
    uv = (torch.rand(n, 2, generator=g) * 1.4 - 0.2) * torch.tensor([W, H])
    radius = torch.randint(0, 40, (n,), generator=g, dtype=torch.int32)
    radius[torch.rand(n, generator=g) < 0.01] = 300
    depth = torch.rand(n, 1, generator=g) * 50 + 0.2
    if ties:
        depth = torch.round(depth)
    if big:
        # near splats covering most of the grid, consecutive in depth order
        idx = torch.randperm(n, generator=g)[:big]
        radius[idx] = torch.randint(max(W, H) // 3, max(W, H), (big,), generator=g,
                                    dtype=torch.int32)
        depth[idx] = 0.05 + 0.1 * torch.rand(big, 1, generator=g)
    return uv, depth, radius


MODES = (2, 1, 0)


def _sort(device, uv, depth, W, H, radius, mode, ws=None):
    prev = ops.set_binning_mode(mode)
    try:
        ids, rng = ops.sort_gaussian(uv.to(device), depth.to(device), W, H, radius.to(device),
                                     None, workspace=ws)
        torch.cuda.synchronize()
    finally:
        ops.set_binning_mode(prev)
    return ids.cpu().numpy(), rng.cpu().numpy()


CASES = [(200_000, 1024, 1024, 1, False, 0),   # 64 x 64 = 4096 tiles
         (500_000, 1237, 822, 2, False, 0),    # 78 x 52 = 4056 tiles
         (300_000, 1920, 1080, 8, False, 0),   # 120 x 68 = 8160 tiles
         (50_000, 979, 546, 3, True, 0),
         (3_000, 33, 17, 4, True, 0),
         (1, 16, 16, 5, False, 0),
         (2_000, 16, 16, 9, True, 0),          # one tile: no tile pass
         (20_000, 1024, 1024, 6, False, 300),  # long single-splat runs
         (5_000, 640, 360, 7, True, 64),
         (30_000, 4200, 700, 12, False, 0),    # 263 x 44 tiles: wider than row binning
         (60_000, 3000, 2200, 15, True, 0),    # 188 x 138 tiles: 8-bit rows and columns
         (2_000, 16, 700, 13, True, 0),        # one tile wide
         (2_000, 700, 16, 14, True, 0)]        # one tile high


@pytest.mark.parametrize("n,W,H,seed,ties,big", CASES)
def test_binning_bitexact_all_modes(device, n, W, H, seed, ties, big):
    uv, depth, radius = _random_case(n, W, H, seed, ties, big)
    ids_r, rng_r = r32.sort_gaussian(uv.numpy(), depth.numpy(), W, H, radius.numpy())
    print(f"{n} splats {W}x{H}: M = {len(ids_r)}", flush=True)
    for mode in MODES:
        ids, rng = _sort(device, uv, depth, W, H, radius, mode)
        assert np.array_equal(ids, ids_r), mode
        assert np.array_equal(rng, rng_r), mode


@pytest.mark.parametrize("frac", [0.3, 0.97])
def test_binning_capacity_below_count(device, frac):
    """Instance capacity below the count: every mode sorts the first ``cap``
    instances of the depth-first sequence (an in-bounds prefix), the device
    count still reports M, and ranges index only the prefix."""
    W, H = 1024, 1024
    uv, depth, radius = _random_case(20_000, W, H, 11, False, 200)
    m = len(r32.sort_gaussian(uv.numpy(), depth.numpy(), W, H, radius.numpy())[0])
    cap = int(m * frac)
    ids_e, rng_e, m_e, _ = _emission_order_reference(uv.numpy(), depth.numpy(), radius.numpy(),
                                                     W, H, cap)
    assert m_e == m
    for mode in MODES:
        ws = ops.SortWorkspace()
        ws.capacity = cap
        ids, rng = _sort(device, uv, depth, W, H, radius, mode, ws)
        assert ws.instances() == m
        assert np.array_equal(ids, ids_e), mode
        assert np.array_equal(rng, rng_e), mode


@pytest.mark.parametrize("mode", MODES)
def test_binning_empty(device, mode):
    ids, rng = _sort(device, torch.zeros(0, 2), torch.zeros(0, 1), 64, 48,
                     torch.zeros(0, dtype=torch.int32), mode)
    assert ids.size == 0 and rng.shape == (12, 2) and not rng.any()
    ids, rng = _sort(device, torch.ones(5, 2), torch.ones(5, 1), 64, 48,
                     torch.zeros(5, dtype=torch.int32), mode)
    assert ids.size == 0 and not rng.any()
