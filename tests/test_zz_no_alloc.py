"""gsr_sort_gaussian_noalloc (the binning on the library's static scratch,
no allocation) through the C ABI against the allocating entry points
(gsr_sort_count + gsr_sort_emit via ops.sort_gaussian): ids, tile ranges and
the device instance count bit-exact, on two calls in a row (the second with a
smaller input), with a truncated capacity, and the size limit rejected before
anything is queued."""
import numpy as np
import pytest
import torch

from pointrix_amd import _lib, ops
from pointrix_amd._lib import ptr, stream_of

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _case(n, W, H, seed):
    g = torch.Generator().manual_seed(seed)
    uv = (torch.rand(n, 2, generator=g) * 1.4 - 0.2) * torch.tensor([W, H])
    radius = torch.randint(0, 40, (n,), generator=g, dtype=torch.int32)
    radius[torch.rand(n, generator=g) < 0.01] = 300
    depth = torch.round(torch.rand(n, 1, generator=g) * 50 + 0.2)
    return uv, depth, radius


def _noalloc(device, uv, depth, radius, W, H, m_cap):
    u = uv.to(device).contiguous()
    d = depth.to(device).reshape(-1).contiguous()
    r = radius.to(device).contiguous()
    n = u.shape[0]
    gx, gy = (W + 15) // 16, (H + 15) // 16
    m_out = torch.full((1,), -1, dtype=torch.int32, device=device)
    ids = torch.full((max(m_cap, 1),), -7, dtype=torch.int32, device=device)
    rng = torch.full((gx * gy, 2), -7, dtype=torch.int32, device=device)
    rc = _lib.load().gsr_sort_gaussian_noalloc(ptr(u), ptr(d), ptr(r), n, W, H, 16, 16, m_cap,
                                               ptr(m_out), ptr(ids), ptr(rng), stream_of(u))
    assert rc == 0, _lib.load().gsr_last_error()
    torch.cuda.synchronize()
    return int(m_out.item()), ids[:m_cap].cpu().numpy(), rng.cpu().numpy()


def _reference(device, uv, depth, radius, W, H, cap=None):
    ws = ops.SortWorkspace()
    ws.capacity = cap
    ids, rng = ops.sort_gaussian(uv.to(device), depth.to(device), W, H, radius.to(device), None,
                                 workspace=ws)
    torch.cuda.synchronize()
    return ws.instances(), ids.cpu().numpy(), rng.cpu().numpy()


def test_noalloc_matches_allocating_sort_twice(device):
    lib = _lib.load()
    for n, W, H, seed in ((200_000, 1237, 822, 1), (40_000, 640, 360, 2)):
        uv, depth, radius = _case(n, W, H, seed)
        m, ids_r, rng_r = _reference(device, uv, depth, radius, W, H)
        need = lib.gsr_sort_ws_gauss_bytes(n) + lib.gsr_sort_ws_inst_bytes(m)
        assert need <= SCRATCH_BYTES, need
        m_n, ids_n, rng_n = _noalloc(device, uv, depth, radius, W, H, m)
        print(f"{n} splats {W}x{H}: M = {m}")
        assert m_n == m
        assert np.array_equal(ids_n, ids_r)
        assert np.array_equal(rng_n, rng_r)


def test_noalloc_truncated_capacity(device):
    W, H = 979, 546
    uv, depth, radius = _case(60_000, W, H, 3)
    m_full = _reference(device, uv, depth, radius, W, H)[0]
    cap = m_full // 2
    m, ids_r, rng_r = _reference(device, uv, depth, radius, W, H, cap)
    m_n, ids_n, rng_n = _noalloc(device, uv, depth, radius, W, H, cap)
    assert m_n == m == m_full
    assert np.array_equal(ids_n, ids_r)
    assert np.array_equal(rng_n, rng_r)


def test_noalloc_rejects_oversize(device):
    lib = _lib.load()
    n = 4_000_000  # the splat workspace alone exceeds the scratch
    assert lib.gsr_sort_ws_gauss_bytes(n) > SCRATCH_BYTES
    u = torch.zeros(1, 2, device=device)
    m_out = torch.full((1,), -1, dtype=torch.int32, device=device)
    rc = lib.gsr_sort_gaussian_noalloc(ptr(u), ptr(u), ptr(u), n, 64, 64, 16, 16, 10,
                                       ptr(m_out), ptr(u), ptr(u), stream_of(u))
    assert rc == -1000
    assert b"scratch" in lib.gsr_last_error()
    torch.cuda.synchronize()
    assert int(m_out.item()) == -1  # nothing was queued
