"""Multi-view batched rendering: many cameras of uneven resolution over one
scene, one kernel launch per stage.

The reference renders a batch view by view (MsplatRender.render_batch,
pointrix/model/renderer/msplat.py:160-213, sharing one image size).  Here a
``BatchedRender`` takes any list of cameras (each with its own size and
intrinsics), runs the multi-view preprocess, one binning pass over all
(view, splat) pairs, and persistent blending kernels that pull tiles of every
view from one work queue; the backward mirrors it.  Gradients flow into the
splat parameters, the cameras (when their tensors require grad) and a
per-view viewspace ``uv_points`` leaf, as in render_batch.

``GraphedStep`` records forward + loss + backward of a fixed camera set into
a HIP graph (torch.cuda.CUDAGraph) with an instance capacity instead of the
host read of the instance count; ``replay()`` reruns it and re-captures with a
larger capacity if the device count outgrew it.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib, ops
from ._lib import call, ptr, stream_of
from .graph import CapturedStep
from .renderer import _split_planes

TILE_SHAPES = ((16, 16), (8, 32), (32, 8))
BATCH_DEPTHS = (256, 128, 64)


@dataclass
class ViewLayout:
    """Packing of V views: tile offsets, pixel offsets, sizes (host side)."""
    widths: list
    heights: list
    tile_w: int
    tile_h: int

    def __post_init__(self):
        self.V = len(self.widths)
        self.grid = [((w + self.tile_w - 1) // self.tile_w, (h + self.tile_h - 1) // self.tile_h)
                     for w, h in zip(self.widths, self.heights)]
        if any(gx >= 65536 or gy >= 65536 for gx, gy in self.grid):
            raise ValueError("a view's tile grid must be under 65536 tiles per side")
        self.tile_off, self.pix_off = [], []
        t = p = 0
        for (gx, gy), w, h in zip(self.grid, self.widths, self.heights):
            self.tile_off.append(t)
            self.pix_off.append(p)
            t += gx * gy
            p += w * h
        self.total_tiles, self.total_pixels = t, p

    def view_int(self, device) -> torch.Tensor:
        rows = [[w, h, gx, gy, to, po, 0, 0] for w, h, (gx, gy), to, po in
                zip(self.widths, self.heights, self.grid, self.tile_off, self.pix_off)]
        return torch.tensor(rows, dtype=torch.int32, device=device)

    def unpack(self, packed: torch.Tensor, v: int) -> torch.Tensor:
        """[C, total_pixels] -> view v's [C, H, W] (a view, no copy)."""
        o, w, h = self.pix_off[v], self.widths[v], self.heights[v]
        return packed[:, o:o + w * h].view(packed.shape[0], h, w)


def view_floats(cameras) -> torch.Tensor:
    """[V, 20] = intr(4) | extr[:3,:4] (12) | centre(3) | 0, differentiable."""
    rows = []
    for c in cameras:
        intr = c["intrinsic_params"].reshape(-1)[:4].float()
        extr = c["extrinsic_matrix"][:3, :4].reshape(-1).float()
        cen = c["camera_center"].reshape(-1)[:3].float()
        rows.append(torch.cat([intr, extr, cen, torch.zeros(1, device=intr.device)]))
    return torch.stack(rows)


class _Unpack(torch.autograd.Function):
    """Packed [C, total_pixels] -> the V views' [C, H, W] (views, no copy).
    The backward writes the views' gradients into one packed buffer -- one
    copy per view -- instead of autograd's full-size zero tensor per sliced
    view."""

    @staticmethod
    def forward(ctx, packed, layout):
        ctx.layout, ctx.shape = layout, packed.shape
        return tuple(layout.unpack(packed, v) for v in range(layout.V))

    @staticmethod
    def backward(ctx, *grads):
        lay = ctx.layout
        g = torch.empty(ctx.shape, device=next(x for x in grads if x is not None).device)
        for v, gv in enumerate(grads):
            dst = lay.unpack(g, v)
            if gv is None:
                dst.zero_()
            else:
                dst.copy_(gv)
        return g, None


class InstanceBudgetExceeded(RuntimeError):
    """The batch needs more (splat, tile) instances than the renderer's
    ``max_instances`` allows; raised after counting, before allocating."""

    def __init__(self, needed: int, budget: int):
        super().__init__(f"batch needs {needed} instances, budget is {budget}")
        self.needed, self.budget = needed, budget


class _Workspace:
    def __init__(self):
        self.sort = ops.SortWorkspace()
        self.queue = None

    def queue_buf(self, device):
        if self.queue is None or self.queue.device != device:
            self.queue = torch.zeros(1, dtype=torch.int32, device=device)
        return self.queue


class _BatchedFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, scale, rot, shs_dc, shs_rest, opacity, vf, ndc, meta):
        lay, vi, flags, degree, bg, bd, ws, capacity = (meta["layout"], meta["vi"],
                                                         meta["flags"], meta["degree"],
                                                         meta["bg"], meta["batch_depth"],
                                                         meta["ws"], meta["capacity"])
        x, s, q, o = ops._f32(xyz), ops._f32(scale), ops._f32(rot), ops._f32(opacity).reshape(-1)
        dc = ops._f32(shs_dc)
        rest = None if shs_rest is None else ops._f32(shs_rest)
        K = dc.shape[1] if rest is None else 1 + rest.shape[1]
        vfc = ops._f32(vf)
        n, V = x.shape[0], lay.V
        C = 3 + (1 if flags & ops.FLAG_DEPTH else 0) + (3 if flags & ops.FLAG_NORMAL else 0)
        dev = x.device
        st = stream_of(x)
        nv = V * n
        uv = torch.empty(nv, 2, device=dev)
        depth = torch.empty(nv, device=dev)
        conic = torch.empty(nv, 3, device=dev)
        radius = torch.empty(nv, dtype=torch.int32, device=dev)
        tiles = torch.empty(nv, dtype=torch.int32, device=dev)
        feat = torch.empty(nv, C, device=dev)
        with ops._span("mv_preprocess_fwd", dev):
            # unless the intermediates are kept for inspection, culled (view,
            # splat) pairs get only radius = tiles = 0 (the backward and every
            # later stage read nothing else of them)
            sparse = 4 if meta.get("keep") is None else 0
            call("gsr_mv_preprocess_fwd", ptr(x), ptr(s), ptr(q), ptr(dc), ptr(rest), n, K,
                 int(degree), V, ptr(vfc), ptr(vi), lay.tile_w, lay.tile_h, 0.2, 1.3,
                 int(flags) | sparse,
                 ptr(uv), ptr(depth), ptr(conic), ptr(radius), ptr(tiles), ptr(feat), st)
        m_dev = meta["m_dev"]
        g = ws.sort.gauss_buf(nv, dev)
        with ops._span("mv_sort", dev):
            tc, to = (ptr(conic), ptr(o)) if meta["tight"] else (None, None)
            if capacity is None:
                # M reaches the host ahead of the depth sort (no stream drain)
                m_host = ws.sort.host_count(dev)
                call("gsr_mv_sort_count_async", ptr(uv), ptr(depth), ptr(radius), tc, to, n, V,
                     ptr(vi), lay.tile_w, lay.tile_h, ptr(g), ptr(m_dev),
                     ctypes.c_void_p(m_host.data_ptr()), None, st)
                m = ws.sort.wait_count(dev)
                meta["count"] = m
            else:
                call("gsr_mv_sort_count", ptr(uv), ptr(depth), ptr(radius), tc, to, n, V,
                     ptr(vi), lay.tile_w, lay.tile_h, ptr(g), ptr(m_dev), st)
                m = int(capacity)
            if meta["max_instances"] is not None and m > meta["max_instances"]:
                raise InstanceBudgetExceeded(m, meta["max_instances"])
            ids = torch.empty(max(m, 1), dtype=torch.int32, device=dev)
            rng = torch.empty(lay.total_tiles, 2, dtype=torch.int32, device=dev)
            inst = ws.sort.inst_buf(m, dev)
            call("gsr_mv_sort_emit", ptr(uv), ptr(radius), tc, to, n, V, ptr(vi),
                 lay.total_tiles, lay.tile_w, lay.tile_h, ptr(g), m, ptr(inst), ptr(ids),
                 ptr(rng), st)
        out = torch.empty(C, lay.total_pixels, device=dev)
        T = torch.empty(lay.total_pixels, device=dev)
        nc = torch.empty(lay.total_pixels, dtype=torch.int32, device=dev)
        queue = ws.queue_buf(dev)
        with ops._span("mv_alpha_blending_fwd", dev):
            call("gsr_mv_alpha_blending_fwd", ptr(uv), ptr(conic), ptr(o), ptr(feat), C, ptr(ids),
                 ptr(rng), float(bg), n, V, ptr(vi), lay.total_tiles, lay.total_pixels,
                 lay.tile_w, lay.tile_h, int(bd), ptr(queue), ptr(out), ptr(T), ptr(nc), st)
        ctx.save_for_backward(x, s, q, dc, rest, o, vfc, uv, conic, feat, radius, ids, rng, T, nc)
        if meta.get("keep") is not None:
            meta["keep"].update(uv=uv, depth=depth, conic=conic, radius=radius, ids=ids,
                                tile_range=rng, features=feat, T_final=T, n_contrib=nc)
        ctx.meta = meta
        ctx.sizes = (n, V, C, K)
        ctx.op_shape = opacity.shape
        ctx.mark_non_differentiable(radius, nc)
        # an unused T output gets no zero-filled gradient (the backward then
        # runs its kernel variant without the background / T_final term)
        ctx.set_materialize_grads(False)
        return out, T, radius, nc

    @staticmethod
    def backward(ctx, g_out, g_T, _gr, _gn):
        (x, s, q, dc, rest, o, vfc, uv, conic, feat, radius, ids, rng, T,
         nc) = ctx.saved_tensors
        meta = ctx.meta
        lay, vi, flags, degree, bg, bd, ws = (meta["layout"], meta["vi"], meta["flags"],
                                              meta["degree"], meta["bg"], meta["batch_depth"],
                                              meta["ws"])
        n, V, C, K = ctx.sizes
        dev = x.device
        st = stream_of(x)
        g_out = torch.zeros(C, lay.total_pixels, device=dev) if g_out is None else ops._f32(g_out)
        g_T = None if g_T is None else ops._f32(g_T)
        ld = ops.grad_record_width(C)
        rec = torch.zeros(V * n, ld, device=dev)
        queue = ws.queue_buf(dev)
        with ops._span("mv_alpha_blending_bwd", dev):
            call("gsr_mv_alpha_blending_bwd", ptr(uv), ptr(conic), ptr(o), ptr(feat), C,
                 ptr(ids), ptr(rng), float(bg), n, V, ptr(vi), lay.total_tiles,
                 lay.total_pixels, lay.tile_w, lay.tile_h, int(bd), ptr(queue), ptr(T), ptr(nc),
                 ptr(g_out), ptr(g_T), ptr(rec), ld, st)
        d_x, d_s, d_q = torch.empty_like(x), torch.empty_like(s), torch.empty_like(q)
        d_o = torch.empty_like(o)
        d_dc = torch.empty_like(dc)
        d_rest = None if rest is None else torch.empty_like(rest)
        need_vf = ctx.needs_input_grad[6]
        d_vf = torch.zeros_like(vfc) if need_vf else None
        with ops._span("mv_preprocess_bwd", dev):
            call("gsr_mv_preprocess_bwd", ptr(x), ptr(s), ptr(q), ptr(dc), ptr(rest), n, K,
                 int(degree), V, ptr(vfc), ptr(vi), 0.2, 1.3, int(flags), ptr(radius), ptr(rec),
                 ld, ptr(d_x), ptr(d_s), ptr(d_q), ptr(d_o), ptr(d_dc), ptr(d_rest), ptr(d_vf),
                 st)
        d_ndc = None
        if ctx.needs_input_grad[7]:
            d_ndc = rec.view(V, n, ld)[:, :, 0:2] * meta["half"][:, None, :]
        return d_x, d_s, d_q, d_dc, d_rest, d_o.reshape(ctx.op_shape), d_vf, d_ndc, None


class BatchedRender:
    """Render V cameras of uneven resolution over one scene (fwd + bwd).

    cameras: list of dicts with height, width, extrinsic_matrix [4,4],
    intrinsic_params [4], camera_center [3] (tensors on the GPU; they may
    require grad).  Splat inputs as MsplatRender.render_iter takes them;
    ``shs`` may be [N,K,3] or (features [N,1,3], features_rest [N,K-1,3]).
    """

    def __init__(self, render_depth=False, render_normal=False, white_bg=False, sh_degree=3,
                 tile=(16, 16), batch_depth=256, render_opacity=False):
        if tuple(tile) not in TILE_SHAPES or batch_depth not in BATCH_DEPTHS:
            raise ValueError(f"tile must be one of {TILE_SHAPES}, batch_depth one of "
                             f"{BATCH_DEPTHS}")
        self.render_depth = render_depth
        self.render_normal = render_normal
        self.render_opacity = render_opacity   # per view "opacity" = 1 - T_final
        self.bg = 1.0 if white_bg else 0.0
        self.sh_degree = sh_degree
        self.tile = tuple(tile)
        self.batch_depth = batch_depth
        self.ws = _Workspace()
        self._m_dev = None
        self.capacity = None      # instance capacity (graph mode); None = exact
        # upper bound on the instances (duplicated keys) of one render; a
        # render that needs more raises InstanceBudgetExceeded before any
        # instance buffer is allocated (the serving admission control)
        self.max_instances = None
        # opacity-aware binning: each splat goes only to the tiles where it
        # can reach alpha >= 1/255 (same images, fewer instances; ids and
        # tile ranges then differ from the msplat tile contract)
        self.tight_tiles = False
        # when True, render() keeps the packed intermediates (uv, depth,
        # conic, radius, ids, tile_range, features, T_final, n_contrib) of
        # all views in self.intermediates
        self.keep_intermediates = False
        self.intermediates = None
        # exact instance count of the last forward, as read on the host to
        # size its buffers (None in capacity mode, where no read is made)
        self.last_instances = None

    @property
    def flags(self):
        return (ops.FLAG_DEPTH if self.render_depth else 0) | \
            (ops.FLAG_NORMAL if self.render_normal else 0)

    def channel_names(self):
        return [("rgb", 3)] + ([("depth", 1)] if self.render_depth else []) + \
            ([("normal", 3)] if self.render_normal else [])

    def layout(self, cameras) -> ViewLayout:
        return ViewLayout([int(c["width"]) for c in cameras], [int(c["height"]) for c in cameras],
                          self.tile[0], self.tile[1])

    def render(self, cameras, position, opacity, scaling, rotation, shs, layout=None,
               vi=None, half=None, vf=None):
        """layout / vi / half (the view table and the (W/2, H/2) per view) may
        be passed precomputed, as a captured graph requires (no host data
        copies inside the step); vf, the [V,20] camera table of view_floats,
        may be passed built on the host and copied once (serving).  Without
        autograd (torch.no_grad) no ``uv_points`` leaf is allocated."""
        out, T, radius, ndc, lay = self.render_packed(cameras, position, opacity, scaling,
                                                      rotation, shs, layout, vi, half, vf)
        imgs = _Unpack.apply(out, lay)
        Ts = _Unpack.apply(T[None], lay) if self.render_opacity else None
        views = []
        for v in range(lay.V):
            d = _split_planes(imgs[v], self.channel_names())
            if Ts is not None:
                d["opacity"] = 1.0 - Ts[v]
            views.append(d)
        rad = radius.view(lay.V, -1)
        return {"views": views, "uv_points": ndc, "visibility": (rad > 0).any(0),
                "radii": rad.max(0).values, "packed": out, "layout": lay}

    def render_packed(self, cameras, position, opacity, scaling, rotation, shs, layout=None,
                      vi=None, half=None, vf=None):
        """-> (out [C, total_pixels], T_final [total_pixels], radius [V*N],
        uv_points leaf or None, layout): the packed images without the
        per-view split (layout.unpack(out, v) is view v's [C,H,W])."""
        lay = layout or self.layout(cameras)
        dev = position.device
        if vi is None:
            vi = lay.view_int(dev)
        if self._m_dev is None or self._m_dev.device != dev:
            self._m_dev = torch.zeros(1, dtype=torch.int32, device=dev)
        dc, rest = (shs[0], shs[1]) if isinstance(shs, (tuple, list)) else (shs, None)
        if vf is None:
            vf = view_floats(cameras)
        if torch.is_grad_enabled():
            ndc = torch.zeros(lay.V, position.shape[0], 2, device=dev, requires_grad=True)
            half = torch.tensor([[0.5 * w, 0.5 * h] for w, h in zip(lay.widths, lay.heights)],
                                device=dev) if half is None else half
        else:
            ndc = None
        meta = {"layout": lay, "vi": vi, "flags": self.flags, "degree": self.sh_degree,
                "bg": self.bg, "batch_depth": self.batch_depth, "ws": self.ws,
                "capacity": self.capacity, "m_dev": self._m_dev, "half": half,
                "max_instances": self.max_instances, "tight": self.tight_tiles,
                "keep": {} if self.keep_intermediates else None}
        self.intermediates = meta["keep"]
        out, T, radius, nc = _BatchedFn.apply(position, scaling, rotation, dc, rest, opacity, vf,
                                              ndc, meta)
        self.last_instances = meta.get("count")
        return out, T, radius, ndc, lay

    def instances(self) -> int:
        """Device instance count of the last forward (host read)."""
        return int(self._m_dev.item()) if self._m_dev is not None else 0

    def count_tensor(self):
        """The device int32[1] the instance count of each forward is written to."""
        return self._m_dev


class GraphedStep(CapturedStep):
    """A fixed camera set's forward + loss + backward through a BatchedRender,
    captured as one HIP graph (pointrix_amd.graph.CapturedStep).
    ``loss_fn(result) -> scalar``; gradients land in the ``.grad`` of the
    splat leaves."""

    def __init__(self, renderer: BatchedRender, cameras, leaves: dict, inputs_fn, loss_fn,
                 headroom=1.25):
        self.r, self.cameras = renderer, cameras
        self.inputs_fn, self.loss_fn = inputs_fn, loss_fn
        self.layout = renderer.layout(cameras)
        dev = next(iter(leaves.values())).device
        self.vi = self.layout.view_int(dev)
        self.half = torch.tensor([[0.5 * w, 0.5 * h] for w, h in
                                  zip(self.layout.widths, self.layout.heights)], device=dev)
        super().__init__(self._render_loss, leaves, renderer, headroom=headroom)

    def _render_loss(self):
        res = self.r.render(self.cameras, **self.inputs_fn(), layout=self.layout, vi=self.vi,
                            half=self.half)
        loss = self.loss_fn(res)
        loss.backward()
        return loss
