"""msplat-compatible differentiable ops backed by libgsr (HIP, gfx950).

The six functions below have the names, positional arguments and return
values of the msplat calls made by the reference renderer
(pointrix/model/renderer/msplat.py:104-151, examples/supervise/renderer.py:42-86),
so a renderer written against msplat ports by changing its import.  Each is a
``torch.autograd.Function`` whose forward and backward are hand-written HIP
kernels reached through the C ABI in include/gsr.h; tensors must live on the
GPU -- there is no CPU path.
"""
from __future__ import annotations

import ctypes

import time

import torch

from . import _lib
from ._lib import call, ptr, stream_of

TILE = 16  # msplat's tile edge (pixels); tile_range has ceil(W/16)*ceil(H/16) rows


class KernelTimer:
    """Optional per-op device timing: torch events recorded on the stream the
    kernels are launched on, around each C-ABI call.  Enabled by the
    benchmark; off (None) by default."""

    def __init__(self):
        self.events = {}

    def span(self, name, dev):
        return _Span(self, name, dev)

    def summary(self):
        """name -> (launches, total ms) after a device synchronisation."""
        out = {}
        for name, evs in self.events.items():
            out[name] = (len(evs), sum(a.elapsed_time(b) for a, b in evs))
        return out


class _Span:
    def __init__(self, timer, name, dev):
        self.t, self.name, self.dev = timer, name, dev

    def __enter__(self):
        self.a = torch.cuda.Event(enable_timing=True)
        self.a.record(torch.cuda.current_stream(self.dev))

    def __exit__(self, *exc):
        b = torch.cuda.Event(enable_timing=True)
        b.record(torch.cuda.current_stream(self.dev))
        self.t.events.setdefault(self.name, []).append((self.a, b))


timer: KernelTimer | None = None


class _NoSpan:
    def __enter__(self):
        return None

    def __exit__(self, *exc):
        return False


def _span(name, dev):
    return timer.span(name, dev) if timer is not None else _NoSpan()


def _dev_check(*ts):
    for t in ts:
        if t is not None and not t.is_cuda:
            raise RuntimeError("pointrix_amd ops run on the GPU only; got a CPU tensor")


def _f32(t: torch.Tensor) -> torch.Tensor:
    return t.detach().to(torch.float32).contiguous()


def _intr4(intr: torch.Tensor) -> torch.Tensor:
    t = _f32(intr).reshape(-1)
    if t.numel() < 4:
        raise ValueError("intrinsics must hold (fx, fy, cx, cy)")
    return t[:4].contiguous()


def _extr34(extr: torch.Tensor) -> torch.Tensor:
    t = _f32(extr)
    if t.shape == (4, 4):
        t = t[:3].contiguous()
    if t.shape != (3, 4):
        raise ValueError(f"extrinsics must be [3,4] or [4,4], got {tuple(extr.shape)}")
    return t


def _intr_grad(g: torch.Tensor, like: torch.Tensor) -> torch.Tensor:
    out = torch.zeros(like.numel(), dtype=like.dtype, device=like.device)
    out[:4] = g.to(like.dtype)
    return out.reshape(like.shape)


def _extr_grad(g: torch.Tensor, like: torch.Tensor) -> torch.Tensor:
    g = g.reshape(3, 4).to(like.dtype)
    if like.shape == (4, 4):
        g = torch.cat([g, torch.zeros(1, 4, dtype=g.dtype, device=g.device)], 0)
    return g


# ---------------------------------------------------------------------------
# project_point  (msplat.py:109-113)
# ---------------------------------------------------------------------------
class _ProjectPoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, intr, extr, W, H, nearest, extent):
        _dev_check(xyz, intr, extr)
        x, k, e = _f32(xyz), _intr4(intr), _extr34(extr)
        n = x.shape[0]
        uv = torch.empty(n, 2, dtype=torch.float32, device=x.device)
        depth = torch.empty(n, 1, dtype=torch.float32, device=x.device)
        call("gsr_project_point_fwd", ptr(x), n, ptr(k), ptr(e), int(W), int(H), float(nearest),
             float(extent), ptr(uv), ptr(depth), stream_of(x))
        ctx.save_for_backward(x, k, e, depth)
        ctx.shapes = (intr, extr)
        return uv, depth

    @staticmethod
    def backward(ctx, g_uv, g_depth):
        x, k, e, depth = ctx.saved_tensors
        intr, extr = ctx.shapes
        n = x.shape[0]
        g_uv = torch.zeros(n, 2, device=x.device) if g_uv is None else _f32(g_uv)
        g_d = None if g_depth is None else _f32(g_depth)
        d_xyz = torch.empty_like(x)
        need_k, need_e = ctx.needs_input_grad[1], ctx.needs_input_grad[2]
        d_k = torch.zeros(4, device=x.device) if need_k else None
        d_e = torch.zeros(12, device=x.device) if need_e else None
        call("gsr_project_point_bwd", ptr(x), n, ptr(k), ptr(e), ptr(depth), ptr(g_uv), ptr(g_d),
             ptr(d_xyz), ptr(d_k), ptr(d_e), stream_of(x))
        return (d_xyz, _intr_grad(d_k, intr) if need_k else None,
                _extr_grad(d_e, extr) if need_e else None, None, None, None, None)


def project_point(xyz, intr, extr, W, H, nearest=0.2, extent=1.3):
    """World points [N,3] -> (uv [N,2] pixels, depth [N,1]); culled points get 0."""
    return _ProjectPoint.apply(xyz, intr, extr, W, H, nearest, extent)


# ---------------------------------------------------------------------------
# compute_cov3d  (msplat.py:118)
# ---------------------------------------------------------------------------
class _ComputeCov3d(torch.autograd.Function):
    @staticmethod
    def forward(ctx, scale, rot, visible):
        _dev_check(scale, rot, visible)
        s, q = _f32(scale), _f32(rot)
        vis = visible.reshape(-1).to(torch.uint8).contiguous()
        n = s.shape[0]
        cov = torch.empty(n, 6, dtype=torch.float32, device=s.device)
        call("gsr_compute_cov3d_fwd", ptr(s), ptr(q), ptr(vis), n, ptr(cov), stream_of(s))
        ctx.save_for_backward(s, q, vis)
        return cov

    @staticmethod
    def backward(ctx, g):
        s, q, vis = ctx.saved_tensors
        n = s.shape[0]
        d_s = torch.empty_like(s)
        d_q = torch.empty_like(q)
        call("gsr_compute_cov3d_bwd", ptr(s), ptr(q), ptr(vis), n, ptr(_f32(g)), ptr(d_s),
             ptr(d_q), stream_of(s))
        return d_s, d_q, None


def compute_cov3d(scale, rot, visible):
    """(scale [N,3], unit quaternion wxyz [N,4], visible [N] or [N,1]) -> cov3d [N,6]."""
    return _ComputeCov3d.apply(scale, rot, visible)


# ---------------------------------------------------------------------------
# ewa_project  (msplat.py:121-130)
# ---------------------------------------------------------------------------
class _EwaProject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, cov3d, intr, extr, uv, W, H, visible):
        _dev_check(xyz, cov3d, intr, extr, uv, visible)
        x, c, k, e = _f32(xyz), _f32(cov3d), _intr4(intr), _extr34(extr)
        u = _f32(uv)
        vis = visible.reshape(-1).to(torch.uint8).contiguous()
        n = x.shape[0]
        conic = torch.empty(n, 3, dtype=torch.float32, device=x.device)
        radius = torch.empty(n, dtype=torch.int32, device=x.device)
        tiles = torch.empty(n, dtype=torch.int32, device=x.device)
        call("gsr_ewa_project_fwd", ptr(x), ptr(c), n, ptr(k), ptr(e), ptr(u), int(W), int(H),
             ptr(vis), ptr(conic), ptr(radius), ptr(tiles), stream_of(x))
        ctx.save_for_backward(x, c, k, e, radius)
        ctx.shapes = (intr, extr, int(W), int(H))
        ctx.mark_non_differentiable(radius, tiles)
        return conic, radius, tiles

    @staticmethod
    def backward(ctx, g_conic, _gr, _gt):
        x, c, k, e, radius = ctx.saved_tensors
        intr, extr, W, H = ctx.shapes
        n = x.shape[0]
        g = torch.zeros(n, 3, device=x.device) if g_conic is None else _f32(g_conic)
        d_x = torch.empty_like(x)
        d_c = torch.empty_like(c)
        need_k, need_e = ctx.needs_input_grad[2], ctx.needs_input_grad[3]
        d_k = torch.zeros(4, device=x.device) if need_k else None
        d_e = torch.zeros(12, device=x.device) if need_e else None
        call("gsr_ewa_project_bwd", ptr(x), ptr(c), n, ptr(k), ptr(e), W, H, ptr(radius), ptr(g),
             ptr(d_x), ptr(d_c), ptr(d_k), ptr(d_e), stream_of(x))
        return (d_x, d_c, _intr_grad(d_k, intr) if need_k else None,
                _extr_grad(d_e, extr) if need_e else None, None, None, None, None)


def ewa_project(xyz, cov3d, intr, extr, uv, W, H, visible):
    """-> (conic [N,3], radius [N] int32, tiles touched [N] int32)."""
    return _EwaProject.apply(xyz, cov3d, intr, extr, uv, W, H, visible)


# ---------------------------------------------------------------------------
# sort_gaussian  (msplat.py:133-135) -- not differentiable
# ---------------------------------------------------------------------------
class SortWorkspace:
    """Reusable device workspace for the binning stage (grows, never shrinks).

    ``capacity`` = None sizes the instance buffers exactly, with one host read
    of the device instance count per sort.  An integer capacity skips that
    read (as a captured graph requires): the instance list is allocated at the
    capacity, emission never writes past it, and ``instances()`` returns the
    true count of the last sort so the caller can detect an overflow;
    ``high_water_tensor()`` holds the largest count of the fixed-capacity
    sorts since it was last zeroed (the library's count kernel updates it)."""

    def __init__(self):
        self.gauss = None
        self.inst = None
        self.m_dev = None
        self.high_dev = None
        self.capacity = None

    def gauss_buf(self, n, device):
        need = int(_lib.load().gsr_sort_ws_gauss_bytes(n))
        if self.gauss is None or self.gauss.numel() < need or self.gauss.device != device:
            self.gauss = torch.empty(need, dtype=torch.uint8, device=device)
        return self.gauss

    def inst_buf(self, m, device):
        need = int(_lib.load().gsr_sort_ws_inst_bytes(m))
        if self.inst is None or self.inst.numel() < need or self.inst.device != device:
            self.inst = torch.empty(need, dtype=torch.uint8, device=device)
        return self.inst

    def count_buf(self, device):
        if self.m_dev is None or self.m_dev.device != device:
            self.m_dev = torch.zeros(1, dtype=torch.int32, device=device)
        return self.m_dev

    def high_buf(self, device):
        if self.high_dev is None or self.high_dev.device != device:
            self.high_dev = torch.zeros(1, dtype=torch.int32, device=device)
        return self.high_dev

    def host_count(self, device):
        """Pinned host int32[1] the instance count reaches ahead of the depth
        sort (the library's count kernel stores it there).  It is set to -1
        here; ``wait_count`` polls it."""
        if getattr(self, "m_host", None) is None or self.m_host_dev != device:
            self.m_host = torch.empty(1, dtype=torch.int32, pin_memory=True)
            self.m_host_dev = device
        self.m_host[0] = -1
        return self.m_host

    def wait_count(self, device, timeout_s=30.0):
        """The count of the sort just queued, read by polling the pinned word
        (no event, so no barrier in the stream ahead of the depth sort).  If
        it has not arrived by the deadline the device is synchronised -- a
        failed launch then raises there -- and the word is read."""
        m_host = self.m_host
        m = int(m_host[0])
        if m >= 0:
            return m
        deadline = time.perf_counter() + timeout_s
        while m < 0 and time.perf_counter() < deadline:
            m = int(m_host[0])
        if m < 0:
            torch.cuda.synchronize(device)
            m = int(m_host[0])
            if m < 0:
                raise _lib.GsrError("sort_count: the instance count never reached the host")
        return m

    def instances(self) -> int:
        """Device instance count of the last sort (host read)."""
        return int(self.m_dev.item()) if self.m_dev is not None else 0

    def count_tensor(self):
        """The device int32[1] the instance count of each sort is written to."""
        return self.m_dev

    def high_water_tensor(self):
        """The device int32[1] holding the largest instance count of the
        fixed-capacity sorts since it was last zeroed (None before the first)."""
        return self.high_dev


def set_binning_mode(mode: int) -> int:
    """Implementation of exact one-view binning, process-wide: 2 (default)
    bins grids of at most 256 x 256 tiles by rows (segments sorted by row,
    then expanded and sorted by column); 1 generates the first tile radix
    pass's items from the depth-ordered splats; 0 writes the instance list
    and sorts it.  All give the same ids and tile ranges.  Returns the
    previous mode."""
    if mode not in (0, 1, 2):
        raise ValueError(f"binning mode must be 0, 1 or 2, not {mode!r}")
    return int(_lib.load().gsr_set_binning_mode(int(mode)))


_default_ws = SortWorkspace()
last_instance_count = 0  # M of the most recent sort_gaussian call (diagnostics)


def sort_gaussian(uv, depth, W, H, radius, tiles, tile_w=TILE, tile_h=TILE, workspace=None,
                  conic=None, opacity=None):
    """-> (gaussian_ids_sorted [M] int32, tile_range [T,2] int32).

    Order: (tile, float32 bits of depth, splat index).  One host read of the
    instance count M sizes the outputs, unless the workspace carries a fixed
    capacity (then ids has that length and only its first M entries, the
    ones tile_range points at, are meaningful).  ``tiles`` is accepted for msplat
    signature compatibility; counts are recomputed from (uv, radius), which is
    what ewa_project's ``tiles`` holds.

    With ``conic`` [N,3] and ``opacity`` [N] (an extension, no msplat
    counterpart) the binning is opacity-aware: a splat goes only to the tiles
    of its rectangle where it can reach alpha >= 1/255, which blends to the
    same images with fewer instances (each tile's list is then a subsequence
    of the exact one).
    """
    _dev_check(uv, depth, radius)
    ws = workspace or _default_ws
    u = _f32(uv)
    d = _f32(depth).reshape(-1)
    r = radius.reshape(-1).to(torch.int32).contiguous()
    n = u.shape[0]
    dev = u.device
    st = stream_of(u)
    gx, gy = (W + tile_w - 1) // tile_w, (H + tile_h - 1) // tile_h
    m_dev = ws.count_buf(dev)
    g = ws.gauss_buf(n, dev)
    if (conic is None) != (opacity is None):
        raise ValueError("sort_gaussian: pass conic and opacity together")
    if conic is not None:
        conic = _f32(conic)
        opacity = _f32(opacity).reshape(-1)
    tc, to = (None, None) if conic is None else (ptr(conic), ptr(opacity))
    if ws.capacity is None:
        # M reaches the host ahead of the depth sort: waiting for it does not
        # drain the stream, so the sort keeps the device busy meanwhile
        m_host = ws.host_count(dev)
        with _span("sort_count", dev):
            call("gsr_sort_count_async", ptr(u), ptr(d), ptr(r), tc, to, n, int(W), int(H),
                 tile_w, tile_h, ptr(g), ptr(m_dev), ctypes.c_void_p(m_host.data_ptr()), None,
                 st)
        m = ws.wait_count(dev)
        global last_instance_count
        last_instance_count = m
    else:
        with _span("sort_count", dev):
            call("gsr_sort_count_high", ptr(u), ptr(d), ptr(r), tc, to, n, int(W), int(H),
                 tile_w, tile_h, ptr(g), ptr(m_dev), ptr(ws.high_buf(dev)), st)
        m = int(ws.capacity)
    ids = torch.empty(m, dtype=torch.int32, device=dev)
    tile_range = torch.empty(gx * gy, 2, dtype=torch.int32, device=dev)
    inst = ws.inst_buf(m, dev)
    with _span("sort_emit", dev):
        call("gsr_sort_emit_tight", ptr(u), ptr(r), tc, to, n, int(W), int(H), tile_w, tile_h,
             ptr(g), m, ptr(inst), ptr(ids), ptr(tile_range), st)
    return ids, tile_range


# ---------------------------------------------------------------------------
# alpha_blending  (msplat.py:148-151)
# ---------------------------------------------------------------------------
class _AlphaBlending(torch.autograd.Function):
    @staticmethod
    def forward(ctx, uv, conic, opacity, feature, ids, tile_range, bg, W, H, ndc, tile_w, tile_h):
        _dev_check(uv, conic, opacity, feature, ids, tile_range)
        u, c, o, f = _f32(uv), _f32(conic), _f32(opacity).reshape(-1), _f32(feature)
        ids_c = ids.to(torch.int32).contiguous()
        tr = tile_range.to(torch.int32).contiguous()
        C = f.shape[1]
        if not 1 <= C <= 16:
            raise ValueError(f"alpha_blending supports 1..16 channels, got {C}")
        dev = u.device
        out = torch.empty(C, H, W, dtype=torch.float32, device=dev)
        T = torch.empty(H, W, dtype=torch.float32, device=dev)
        nc = torch.empty(H, W, dtype=torch.int32, device=dev)
        with _span("alpha_blending_fwd", dev):
            call("gsr_alpha_blending_fwd", ptr(u), ptr(c), ptr(o), ptr(f), C, ptr(ids_c),
                 ptr(tr), float(bg), int(W), int(H), tile_w, tile_h, ptr(out), ptr(T), ptr(nc),
                 stream_of(u))
        ctx.save_for_backward(u, c, o, f, ids_c, tr, T, nc)
        ctx.cfg = (float(bg), int(W), int(H), tile_w, tile_h, opacity.shape)
        ctx.mark_non_differentiable(nc)
        ctx.set_materialize_grads(False)
        return out, T, nc

    @staticmethod
    def backward(ctx, g_out, g_T, _g_nc):
        u, c, o, f, ids, tr, T, nc = ctx.saved_tensors
        bg, W, H, tw, th, op_shape = ctx.cfg
        n, C = f.shape
        dev = u.device
        g_out = torch.zeros(C, H, W, device=dev) if g_out is None else _f32(g_out)
        g_T = None if g_T is None else _f32(g_T)
        # one zero-filled buffer: compact records [da, db, dc, dfeature...]
        # (3 + C per splat), then (du, dv) [n,2] and dopacity [n] dense, so
        # their consumers (the uv / ndc and opacity gradients) read 12 bytes
        # per splat, not a row
        ld = 3 + C
        a64 = lambda v: (v + 63) // 64 * 64  # 256-byte aligned sections
        off_u = a64(n * ld)
        off_o = off_u + a64(2 * n)
        buf = torch.zeros(off_o + n, device=dev)
        rec = buf[:n * ld].view(n, ld)
        d_u = buf[off_u:off_u + 2 * n].view(n, 2)
        d_o = buf[off_o:off_o + n]
        with _span("alpha_blending_bwd", dev):
            call("gsr_alpha_blending_bwd_split", ptr(u), ptr(c), ptr(o), ptr(f), C, ptr(ids),
                 ptr(tr), bg, W, H, tw, th, ptr(T), ptr(nc), ptr(g_out), ptr(g_T), ptr(rec), ld,
                 ptr(d_u), ptr(d_o), stream_of(u))
        d_c, d_f = rec[:, 0:3], rec[:, 3:3 + C]
        d_ndc = None
        if ctx.needs_input_grad[9]:
            d_ndc = torch.empty(n, 2, device=dev)
            call("gsr_uv_grad_scale", ptr(d_u), 2, n, 0.5 * W, 0.5 * H, ptr(d_ndc),
                 stream_of(u))
        return (d_u, d_c, d_o.reshape(op_shape), d_f, None, None, None, None, None, d_ndc, None,
                None)


def grad_record_width(C: int) -> int:
    """Row width of the per-splat gradient records of the blend backward:
    [du, dv, da, db, dc, dopacity, dfeature...] padded to 8, 16 or 32."""
    q = 6 + C
    return 8 if q <= 8 else (16 if q <= 16 else 32)


def alpha_blending_ex(uv, conic, opacity, feature, ids, tile_range, bg, W, H, ndc=None,
                      tile_w=TILE, tile_h=TILE):
    """alpha_blending that also returns T_final [H,W] (differentiable) and
    n_contrib [H,W] int32; 1 - T_final is the rendered opacity."""
    return _AlphaBlending.apply(uv, conic, opacity, feature, ids, tile_range, bg, W, H, ndc,
                                tile_w, tile_h)


def alpha_blending(uv, conic, opacity, feature, ids, tile_range, bg, W, H, ndc=None):
    """Front-to-back compositing of feature [N,C] -> [C,H,W].

    ``ndc`` (optional, [N,2], a zero leaf with requires_grad) receives
    dL/duv * (W/2, H/2), the viewspace-gradient the densification controller
    reads (pointrix/controller/gs.py:259-284); its value is not used.
    """
    out, _, _ = alpha_blending_ex(uv, conic, opacity, feature, ids, tile_range, bg, W, H, ndc)
    return out


# ---------------------------------------------------------------------------
# compute_sh  (msplat.py:104)
# ---------------------------------------------------------------------------
class _ComputeSH(torch.autograd.Function):
    @staticmethod
    def forward(ctx, shs, dirs, degree):
        _dev_check(shs, dirs)
        s = shs.detach()
        if s.dtype != torch.float32:
            s = s.float()
        d = _f32(dirs)
        n, C, K = s.shape
        out = torch.empty(n, C, dtype=torch.float32, device=d.device)
        sn, sc, sk = s.stride()
        call("gsr_compute_sh_fwd", ptr(s), sn, sc, sk, ptr(d), n, C, K, int(degree), ptr(out),
             stream_of(d))
        ctx.save_for_backward(s, d)
        ctx.degree = int(degree)
        return out

    @staticmethod
    def backward(ctx, g):
        s, d = ctx.saved_tensors
        n, C, K = s.shape
        d_s = torch.empty(n, C, K, dtype=torch.float32, device=d.device)
        d_d = torch.empty_like(d) if ctx.needs_input_grad[1] else None
        sn, sc, sk = s.stride()
        call("gsr_compute_sh_bwd", ptr(s), sn, sc, sk, ptr(d), n, C, K, ctx.degree, ptr(_f32(g)),
             ptr(d_s), ptr(d_d), stream_of(d))
        return d_s, d_d, None


def compute_sh(shs, dirs, degree=None):
    """Real-SH evaluation: shs [N,C,K] (any strides), unit dirs [N,3] -> [N,C].

    ``degree`` restricts the sum to the first (degree+1)^2 coefficients (the
    SH warm-up of msplat.py:98-101 without materialising a mask).
    """
    return _ComputeSH.apply(shs, dirs, -1 if degree is None else degree)


# ---------------------------------------------------------------------------
# Fused preprocess for MsplatRender.render_iter (msplat.py:94-139)
# ---------------------------------------------------------------------------
FLAG_DEPTH = 1
FLAG_NORMAL = 2
FLAG_RAW = 4
FLAG_PLAIN = 8   # colour = the SH sum, no +0.5 / clamp


def _rows(t: torch.Tensor) -> torch.Tensor:
    """fp32 2-D tensor whose rows are contiguous (row stride may be larger
    than the row, e.g. a column slice of the blend's gradient records)."""
    t = t.detach()
    if t.dtype != torch.float32 or t.dim() != 2 or t.stride(1) != 1:
        t = t.to(torch.float32).contiguous()
    return t


def _aligned16(t: torch.Tensor) -> torch.Tensor:
    return t if t.data_ptr() % 16 == 0 else t.clone()


class _Preprocess(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, scale, rot, shs, shs_rest, intr, extr, center, W, H, degree, flags,
                nearest, extent, keep_cov, opacity_raw):
        _dev_check(xyz, scale, rot, shs, shs_rest, intr, extr, center, opacity_raw)
        x, s, q = _f32(xyz), _f32(scale), _f32(rot)
        o_raw = None if opacity_raw is None else _f32(opacity_raw).reshape(-1)
        if shs_rest is None:
            sh, rest = _aligned16(_f32(shs)), None
            K = sh.shape[1]
        else:
            sh, rest = _f32(shs), _f32(shs_rest)
            if sh.shape[1] != 1:
                raise ValueError("with shs_rest, shs holds the DC coefficients [N,1,3]")
            K = 1 + rest.shape[1]
        k, e = _intr4(intr), _extr34(extr)
        c = _f32(center).reshape(-1)[:3].contiguous()
        n = x.shape[0]
        if sh.shape[-1] != 3 or K not in (1, 4, 9, 16):
            raise ValueError(f"SH must be [N,K,3] with K in 1,4,9,16; got K={K}")
        C = 3 + (1 if flags & FLAG_DEPTH else 0) + (3 if flags & FLAG_NORMAL else 0)
        dev = x.device
        uv = torch.empty(n, 2, device=dev)
        depth = torch.empty(n, 1, device=dev)
        conic = torch.empty(n, 3, device=dev)
        radius = torch.empty(n, dtype=torch.int32, device=dev)
        tiles = torch.empty(n, dtype=torch.int32, device=dev)
        feat = torch.empty(n, C, device=dev)
        cov = torch.empty(n, 6, device=dev) if keep_cov else torch.empty(0, device=dev)
        if o_raw is not None and o_raw.shape[0] != n:
            raise ValueError("opacity must hold one logit per splat")
        o_act = torch.empty(n, 1, device=dev) if o_raw is not None else torch.empty(0, device=dev)
        with _span("preprocess_fwd", dev):
            call("gsr_preprocess_fwd", ptr(x), ptr(s), ptr(q), ptr(sh), ptr(rest), n, K,
                 int(degree), ptr(k), ptr(e), ptr(c), int(W), int(H), float(nearest),
                 float(extent), int(flags), ptr(uv), ptr(depth), ptr(cov) if keep_cov else None,
                 ptr(conic), ptr(radius), ptr(tiles), ptr(feat), ptr(o_raw),
                 ptr(o_act) if o_raw is not None else None, stream_of(x))
        ctx.save_for_backward(x, s, q, sh, rest, k, e, c, radius, o_raw)
        ctx.cfg = (int(W), int(H), int(degree), int(flags), float(nearest), float(extent), intr,
                   extr, center, C, K)
        ctx.o_shape = None if opacity_raw is None else opacity_raw.shape
        ctx.mark_non_differentiable(radius, tiles, cov)
        ctx.set_materialize_grads(False)
        return uv, depth, conic, radius, tiles, feat, cov, o_act

    @staticmethod
    def backward(ctx, g_uv, g_depth, g_conic, _gr, _gt, g_feat, _gc, g_oact):
        x, s, q, sh, rest, k, e, c, radius, o_raw = ctx.saved_tensors
        W, H, degree, flags, nearest, extent, intr, extr, center, C, K = ctx.cfg
        n = x.shape[0]
        dev = x.device
        z = lambda *shape: torch.zeros(*shape, device=dev)
        g_uv = z(n, 2) if g_uv is None else _rows(g_uv)
        g_conic = z(n, 3) if g_conic is None else _rows(g_conic)
        g_feat = z(n, C) if g_feat is None else _rows(g_feat)
        g_depth = None if g_depth is None else _f32(g_depth).reshape(-1)
        d_x, d_s, d_q = torch.empty_like(x), torch.empty_like(s), torch.empty_like(q)
        d_sh = torch.empty_like(sh)
        d_rest = None if rest is None else torch.empty_like(rest)
        need_k, need_e, need_c = ctx.needs_input_grad[5], ctx.needs_input_grad[6], \
            ctx.needs_input_grad[7]
        # the camera gradients share one zeroed buffer (one fill): intrinsics
        # [0:4], extrinsics [4:20] as a 4x4 whose bottom row stays zero, centre
        # [20:23]
        cam = z(24) if (need_k or need_e or need_c) else None
        d_k = cam[0:4] if need_k else None
        d_e = cam[4:16] if need_e else None
        d_c = cam[20:23] if need_c else None
        d_o, ld_o = None, 1
        if o_raw is not None and ctx.needs_input_grad[15]:
            if g_oact is None:
                g_oact = z(n)
            else:
                # a column of the blend's gradient records is read in place
                g_oact = g_oact.detach().reshape(-1)
                if g_oact.dtype != torch.float32 or g_oact.stride(0) < 1:
                    g_oact = _f32(g_oact)
                ld_o = g_oact.stride(0)
            d_o = torch.empty_like(o_raw)
        with _span("preprocess_bwd", dev):
            call("gsr_preprocess_bwd", ptr(x), ptr(s), ptr(q), ptr(sh), ptr(rest), n, K, degree,
                 ptr(k), ptr(e), ptr(c), W, H, nearest, extent, flags, ptr(radius), ptr(g_uv),
                 g_uv.stride(0), ptr(g_depth), ptr(g_conic), g_conic.stride(0), ptr(g_feat),
                 g_feat.stride(0), ptr(d_x), ptr(d_s), ptr(d_q), ptr(d_sh), ptr(d_rest),
                 ptr(d_k), ptr(d_e), ptr(d_c), ptr(o_raw) if d_o is not None else None,
                 ptr(g_oact) if d_o is not None else None, ld_o, ptr(d_o), stream_of(x))
        g_k = g_e = None
        if need_k:
            g_k = (cam[0:4].view(intr.shape).to(intr.dtype) if intr.numel() == 4
                   else _intr_grad(d_k, intr))
        if need_e:
            g_e = (cam[4:20].view(4, 4).to(extr.dtype) if tuple(extr.shape) == (4, 4)
                   else _extr_grad(d_e, extr))
        return (d_x, d_s, d_q, d_sh, d_rest, g_k, g_e,
                d_c.reshape(center.shape).to(center.dtype) if need_c else None,
                None, None, None, None, None, None, None,
                None if d_o is None else d_o.view(ctx.o_shape))


def preprocess(position, scaling, rotation, shs, intrinsic_params, extrinsic_matrix,
               camera_center, W, H, degree=None, depth_channel=False, normal_channel=False,
               nearest=0.2, extent=1.3, keep_cov=False, raw=False, opacity=None,
               plain_sh=False):
    """Everything render_iter does per splat before sorting, in one kernel.

    Inputs as MsplatRender.render_iter receives them (activated scaling,
    unit wxyz rotation, shs [N,K,3] -- or the pair (DC [N,1,3], rest
    [N,K-1,3]), which is read without concatenating).  Returns (uv [N,2], depth [N,1], conic
    [N,3], radius [N], tiles [N], feat [N,C], cov3d [N,6] or empty) where
    feat = clamp(SH colour + 0.5, 0) | depth | camera-space normal.  uv,
    depth, conic, radius and tiles equal those of project_point +
    compute_cov3d + ewa_project.

    ``raw=True``: scaling, rotation (and ``opacity``, if given) are the
    model's raw parameters -- log scale, unnormalised quaternion, opacity
    logit -- and the kernel applies GaussianPointCloud's activations (exp,
    F.normalize, sigmoid; gaussian_points.py:30-35,70-81) itself, so the
    gradients land on the raw parameters without separate elementwise
    passes.  With ``opacity`` the activated opacity [N,1] is appended to the
    returned tuple.

    ``plain_sh=True``: the rgb channels are the SH sum itself, without the
    +0.5 offset and the clamp (MsplatNormalRender's colour,
    examples/supervise/renderer.py:41-42).
    """
    flags = (FLAG_DEPTH if depth_channel else 0) | (FLAG_NORMAL if normal_channel else 0)
    if plain_sh:
        flags |= FLAG_PLAIN
    if raw:
        flags |= FLAG_RAW
    elif opacity is not None:
        raise ValueError("preprocess: opacity logits are taken only with raw=True")
    shs_dc, shs_rest = (shs[0], shs[1]) if isinstance(shs, (tuple, list)) else (shs, None)
    out = _Preprocess.apply(position, scaling, rotation, shs_dc, shs_rest, intrinsic_params,
                            extrinsic_matrix, camera_center, W, H,
                            -1 if degree is None else degree, flags, nearest, extent, keep_cov,
                            opacity)
    return out if opacity is not None else out[:7]


# ---------------------------------------------------------------------------
# Fused L1 loss (pointrix/model/loss.py:27)
# ---------------------------------------------------------------------------
class _L1Loss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y):
        _dev_check(x, y)
        if x.shape != y.shape:
            raise ValueError(f"l1_loss shapes differ: {tuple(x.shape)} vs {tuple(y.shape)}")
        xc, yc = _f32(x), _f32(y)
        n = xc.numel()
        lib = _lib.load()
        partial = torch.empty(int(lib.gsr_l1_loss_partials(n)), device=xc.device)
        out = torch.empty((), device=xc.device)
        call("gsr_l1_loss_fwd", ptr(xc), ptr(yc), n, ptr(partial), ptr(out), stream_of(xc))
        ctx.save_for_backward(xc, yc)
        ctx.y_needs = ctx.needs_input_grad[1]
        return out

    @staticmethod
    def backward(ctx, g):
        xc, yc = ctx.saved_tensors
        gc = _f32(g)
        dx = torch.empty_like(xc)
        call("gsr_l1_loss_bwd", ptr(xc), ptr(yc), xc.numel(), ptr(gc), ptr(dx), stream_of(xc))
        return dx, (-dx if ctx.y_needs else None)


def l1_loss(x, y):
    """mean |x - y| as one deterministic, graph-capturable reduction."""
    return _L1Loss.apply(x, y)


class _L1Terms(ctypes.Structure):
    _fields_ = [("x", ctypes.c_void_p * 8), ("y", ctypes.c_void_p * 8),
                ("dx", ctypes.c_void_p * 8), ("n", ctypes.c_int64 * 8),
                ("weight", ctypes.c_float * 8), ("nterms", ctypes.c_int)]


def _l1_table(xs, ys, weights, dxs=None):
    tab = _L1Terms()
    for k, (x, y, w) in enumerate(zip(xs, ys, weights)):
        tab.x[k], tab.y[k] = x.data_ptr(), y.data_ptr()
        tab.dx[k] = None if dxs is None or dxs[k] is None else dxs[k].data_ptr()
        tab.n[k], tab.weight[k] = x.numel(), w
    tab.nterms = len(xs)
    return tab


class _L1Sum(torch.autograd.Function):
    @staticmethod
    def forward(ctx, weights, *xy):
        xs, ys = [_f32(t) for t in xy[0::2]], [_f32(t) for t in xy[1::2]]
        _dev_check(*xs, *ys)
        tab = _l1_table(xs, ys, weights)
        lib = _lib.load()
        nb = int(lib.gsr_l1_sum_partials(ctypes.byref(tab)))
        if nb <= 0:
            raise ValueError("l1_loss_sum: every term must be non-empty")
        partial = torch.empty(nb, device=xs[0].device)
        out = torch.empty((), device=xs[0].device)
        call("gsr_l1_sum_fwd", ctypes.byref(tab), ptr(partial), ptr(out), stream_of(xs[0]))
        ctx.save_for_backward(*xs, *ys)
        ctx.weights = weights
        return out

    @staticmethod
    def backward(ctx, g):
        saved = ctx.saved_tensors
        k = len(saved) // 2
        xs, ys = saved[:k], saved[k:]
        need = [ctx.needs_input_grad[1 + 2 * i] or ctx.needs_input_grad[2 + 2 * i]
                for i in range(k)]
        dxs = [torch.empty_like(x) if nd else None for x, nd in zip(xs, need)]
        call("gsr_l1_sum_bwd", ctypes.byref(_l1_table(xs, ys, ctx.weights, dxs)),
             ptr(_f32(g)), stream_of(xs[0]))
        grads = [None]
        for i in range(k):
            grads.append(dxs[i] if ctx.needs_input_grad[1 + 2 * i] else None)
            grads.append(-dxs[i] if ctx.needs_input_grad[2 + 2 * i] else None)
        return tuple(grads)


def l1_loss_sum(terms):
    """sum of w * mean |x - y| over ``terms`` = [(x, y, w), ...] (at most 8),
    e.g. l1(rgb) + 0.1 l1(depth) + 0.1 l1(normal) as one two-kernel reduction
    and a one-kernel backward, instead of a reduction per term plus the
    scalar multiply / add kernels (and their backward) that combine them."""
    terms = list(terms)
    if not 1 <= len(terms) <= 8:
        raise ValueError("l1_loss_sum takes 1..8 terms")
    flat, weights = [], []
    for x, y, w in terms:
        if x.shape != y.shape:
            raise ValueError(f"l1_loss_sum shapes differ: {tuple(x.shape)} vs {tuple(y.shape)}")
        flat += [x, y]
        weights.append(float(w))
    return _L1Sum.apply(tuple(weights), *flat)


# ---------------------------------------------------------------------------
# fused L1 + SSIM loss (pointrix/model/loss.py:27, :73-116; base_model.py:112-124)
# ---------------------------------------------------------------------------
class _L1Ssim(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y):
        xc, yc = _f32(x), _f32(y)
        H, W = xc.shape[-2], xc.shape[-1]
        planes = xc.numel() // (H * W)
        dev = xc.device
        part = torch.empty(int(_lib.load().gsr_ssim_partials(planes, W, H)), device=dev)
        out = torch.empty(2, device=dev)
        call("gsr_l1_ssim_fwd", ptr(xc), ptr(yc), planes, W, H, ptr(part), ptr(out),
             stream_of(xc))
        ctx.save_for_backward(xc, yc)
        ctx.dims = (planes, W, H, x.shape)
        return out

    @staticmethod
    def backward(ctx, g):
        xc, yc = ctx.saved_tensors
        planes, W, H, shape = ctx.dims
        dx = torch.empty_like(xc)
        call("gsr_l1_ssim_bwd", ptr(xc), ptr(yc), planes, W, H, ptr(_f32(g)), ptr(dx),
             stream_of(xc))
        return dx.view(shape), None


def l1_ssim(pred, gt):
    """-> tensor [2] = (mean SSIM, mean |pred - gt|), differentiable in pred;
    pred and gt [..., H, W] (e.g. [B,3,H,W]) on the GPU."""
    if pred.shape != gt.shape:
        raise ValueError("l1_ssim: shapes differ")
    _dev_check(pred, gt)
    return _L1Ssim.apply(pred, gt)


def l1_ssim_loss(pred, gt, lambda_ssim=0.2):
    """BaseModel.get_loss_dict (base_model.py:112-124): loss =
    (1 - lambda) * L1 + lambda * (1 - SSIM); returns the same dict."""
    s = l1_ssim(pred, gt)
    ssim_loss = 1.0 - s[0]
    loss = (1.0 - lambda_ssim) * s[1] + lambda_ssim * ssim_loss
    return {"loss": loss, "L1_loss": s[1], "ssim_loss": ssim_loss}


# ---------------------------------------------------------------------------
# Point-cloud activations (gaussian_points.py:30-35,70-81)
# ---------------------------------------------------------------------------
def _aligned_f32(t):
    t = _f32(t)
    return t if t.data_ptr() % 16 == 0 else t.clone()


class _Activate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_scale, quat, logit):
        _dev_check(log_scale, quat, logit)
        ls, q, lo = _f32(log_scale), _aligned_f32(quat), _f32(logit)
        n = ls.shape[0]
        if q.shape[0] != n or lo.numel() != n or ls.numel() != 3 * n or q.numel() != 4 * n:
            raise ValueError("activate: log_scale [N,3], quat [N,4], logit [N,1] expected")
        s, r, o = torch.empty_like(ls), torch.empty_like(q), torch.empty_like(lo)
        call("gsr_activate_fwd", ptr(ls), ptr(q), ptr(lo), n, ptr(s), ptr(r), ptr(o),
             stream_of(ls))
        ctx.save_for_backward(q, s, o)  # the backward recomputes r from q
        ctx.shapes = (log_scale.shape, quat.shape, logit.shape)
        ctx.set_materialize_grads(False)
        return s.view(log_scale.shape), r.view(quat.shape), o.view(logit.shape)

    @staticmethod
    def backward(ctx, gs, gr, go):
        q, s, o = ctx.saved_tensors
        n = q.shape[0]
        need = ctx.needs_input_grad
        d_ls = torch.empty_like(s) if need[0] else None
        d_q = torch.empty_like(q) if need[1] else None
        d_lo = torch.empty_like(o) if need[2] else None
        gs = None if gs is None else _f32(gs)
        gr = None if gr is None else _aligned_f32(gr)
        go = None if go is None else _f32(go)
        call("gsr_activate_bwd", ptr(q), ptr(s), None, ptr(o), ptr(gs), ptr(gr), ptr(go), n,
             ptr(d_ls), ptr(d_q), ptr(d_lo), stream_of(q))
        shp = ctx.shapes
        return (None if d_ls is None else d_ls.view(shp[0]),
                None if d_q is None else d_q.view(shp[1]),
                None if d_lo is None else d_lo.view(shp[2]))


def activate(log_scale, quat, logit):
    """The point cloud's activations in one kernel (and one backward):
    (exp(log_scale) [N,3], F.normalize(quat) [N,4], sigmoid(logit) [N,1])."""
    return _Activate.apply(log_scale, quat, logit)


# ---------------------------------------------------------------------------
# Camera pose (camera_model.py:92-175 with utils/pose.py:40-84)
# ---------------------------------------------------------------------------
class _CameraPose(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, t):
        _dev_check(q, t)
        qc, tc = _f32(q).reshape(-1, 4), _f32(t).reshape(-1, 3)
        C = qc.shape[0]
        if tc.shape[0] != C:
            raise ValueError("camera_pose: one translation per quaternion")
        E = torch.empty(C, 4, 4, device=qc.device)
        center = torch.empty(C, 3, device=qc.device)
        call("gsr_camera_pose_fwd", ptr(qc), ptr(tc), C, ptr(E), ptr(center), stream_of(qc))
        ctx.save_for_backward(qc, tc)
        ctx.shapes = (q.shape, t.shape)
        ctx.set_materialize_grads(False)
        return E, center

    @staticmethod
    def backward(ctx, gE, gc):
        qc, tc = ctx.saved_tensors
        dq, dt = torch.empty_like(qc), torch.empty_like(tc)
        gE = None if gE is None else _f32(gE)
        gc = None if gc is None else _f32(gc)
        call("gsr_camera_pose_bwd", ptr(qc), ptr(tc), qc.shape[0], ptr(gE), ptr(gc), ptr(dq),
             ptr(dt), stream_of(qc))
        return dq.view(ctx.shapes[0]), dt.view(ctx.shapes[1])


def camera_pose(q, t):
    """pointrix's CameraModel composition in one kernel (and one backward):
    wxyz quaternion(s) q [...,4] (normalised as F.normalize does) and
    translation(s) t [...,3] -> (world-to-camera E [...,4,4] = [R | t; 0 0 0 1],
    camera centre [...,3] = -R^T t), differentiable in q and t."""
    lead = q.shape[:-1]
    E, c = _CameraPose.apply(q, t)
    return E.view(*lead, 4, 4), c.view(*lead, 3)


# ---------------------------------------------------------------------------
# Output planes (RenderFeatures.split, renderer_utils.py:57-71; the normal
# mapping of examples/supervise/renderer.py:89-96)
# ---------------------------------------------------------------------------
class _PlaneGrads(ctypes.Structure):
    _fields_ = [("grad", ctypes.c_void_p * 8), ("first", ctypes.c_int * 8),
                ("count", ctypes.c_int * 8), ("ngroups", ctypes.c_int)]


class _Planes(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rendered, sizes, normal_first):
        _dev_check(rendered)
        x = _f32(rendered)
        C, P = x.shape[0], x[0].numel()
        out = torch.empty_like(x)
        call("gsr_planes_fwd", ptr(x), C, P, normal_first, ptr(out), stream_of(x))
        ctx.save_for_backward(x)
        ctx.cfg = (tuple(sizes), normal_first)
        ctx.set_materialize_grads(False)
        return tuple(torch.split(out, list(sizes), dim=0))

    @staticmethod
    def backward(ctx, *grads):
        x, = ctx.saved_tensors
        sizes, nf = ctx.cfg
        C, P = x.shape[0], x[0].numel()
        tab, keep, first = _PlaneGrads(), [], 0
        for k, (c, g) in enumerate(zip(sizes, grads)):
            if g is not None:
                g = _f32(g)
                keep.append(g)
            tab.grad[k] = None if g is None else g.data_ptr()
            tab.first[k] = first
            tab.count[k] = c
            first += c
        tab.ngroups = len(sizes)
        d = torch.empty_like(x)
        call("gsr_planes_bwd", ptr(x), C, P, nf, ctypes.byref(tab), ptr(d), stream_of(x))
        return d, None, None


def output_planes(rendered, sizes, normal_first=-1):
    """Split the blended image [C,H,W] into planes of ``sizes`` (channel
    order), the group starting at channel ``normal_first`` (three planes;
    -1 for none) mapped to (n / |n| + 1) / 2 -- as MsplatRender's
    post-processing does -- in one kernel, with a one-kernel backward that
    assembles the image gradient from the groups' gradients."""
    sizes = [int(c) for c in sizes]
    if sum(sizes) != rendered.shape[0] or len(sizes) > 8:
        raise ValueError(f"output_planes: sizes {sizes} do not tile {rendered.shape[0]} channels")
    return _Planes.apply(rendered, sizes, int(normal_first))
