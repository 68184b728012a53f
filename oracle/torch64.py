"""FP64 PyTorch-CPU transcription of the msplat rasterisation path (test oracle).

Test infrastructure: imported only by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg.  It is the tolerance target for the HIP
path and, being differentiable through autograd, the gradient target.

Each function mirrors one msplat call site of the reference renderer
(pointrix/model/renderer/msplat.py) with the same positional arguments.  msplat
itself is an un-vendored submodule (.gitmodules:1-3), so its numeric
conventions are not in the reference tree; the ones fixed here are these
3DGS-compatible choices:

* pinhole uv = (fx*x/z + cx, fy*y/z + cy), pixel centres at integer coords;
* cull: z <= nearest (0.2 at msplat.py:113) or |2u/W-1| > extent (1.3);
* cov3d packing (xx, xy, xz, yy, yz, zz); quaternion wxyz, R as in
  pointrix/utils/pose.py:40-84;
* EWA: J clamped at 1.3*tan(fov/2), +0.3 low-pass, radius ceil(3*sqrt(l_max)),
  l_max = mid + sqrt(max(0.1, mid^2 - det));
* 16x16 tiles; rect = [trunc((u-r)/16), trunc((u+r+15)/16)) clamped;
* key = tile << 32 | f32bits(depth), stable, ties by Gaussian index;
* blend: alpha = min(0.99, o*exp(power)), skip power > 0 or alpha < 1/255,
  stop before a Gaussian that would take T below 1e-4; scalar bg on every
  channel (msplat.py:49); ndc.grad = dL/duv * (W/2, H/2).
"""
from __future__ import annotations

import math

import numpy as np
import torch

F64 = torch.float64
TILE_W = 16
TILE_H = 16
ALPHA_MAX = 0.99
ALPHA_MIN = 1.0 / 255.0
T_MIN = 1e-4
LOWPASS = 0.3
J_CLAMP = 1.3


# ----------------------------------------------------------------------------
# Spherical harmonics (msplat.compute_sh, called at msplat.py:104)
# ----------------------------------------------------------------------------
def sh_constants():
    """Real-SH normalisation constants, derived from their closed forms.

    Degree 0: 1/(2 sqrt(pi)) -- the same C0 as pointrix/utils/common.py:2.
    Signs follow the 3DGS real-SH convention (l=1 band is (-y, z, -x)).
    """
    pi = math.pi
    c0 = 0.5 / math.sqrt(pi)
    c1 = math.sqrt(3.0 / (4.0 * pi))
    a = 0.5 * math.sqrt(15.0 / pi)
    c2 = [a, -a, 0.25 * math.sqrt(5.0 / pi), -a, 0.25 * math.sqrt(15.0 / pi)]
    b = 0.25 * math.sqrt(35.0 / (2.0 * pi))
    d = 0.25 * math.sqrt(21.0 / (2.0 * pi))
    c3 = [-b, 0.5 * math.sqrt(105.0 / pi), -d, 0.25 * math.sqrt(7.0 / pi), -d,
          0.25 * math.sqrt(105.0 / pi), -b]
    return c0, c1, c2, c3


def sh_basis(dirs: torch.Tensor, k: int) -> torch.Tensor:
    """Real SH basis Y_0..Y_{k-1} at unit directions ``dirs`` [N,3] -> [N,k]."""
    c0, c1, c2, c3 = sh_constants()
    x, y, z = dirs[:, 0], dirs[:, 1], dirs[:, 2]
    ys = [torch.full_like(x, c0)]
    if k > 1:
        ys += [-c1 * y, c1 * z, -c1 * x]
    if k > 4:
        xx, yy, zz = x * x, y * y, z * z
        ys += [c2[0] * x * y, c2[1] * y * z, c2[2] * (2 * zz - xx - yy),
               c2[3] * x * z, c2[4] * (xx - yy)]
    if k > 9:
        ys += [c3[0] * y * (3 * xx - yy), c3[1] * x * y * z,
               c3[2] * y * (4 * zz - xx - yy), c3[3] * z * (2 * zz - 3 * xx - 3 * yy),
               c3[4] * x * (4 * zz - xx - yy), c3[5] * z * (xx - yy),
               c3[6] * x * (xx - 3 * yy)]
    return torch.stack(ys[:k], dim=1)


def compute_sh(shs: torch.Tensor, dirs: torch.Tensor, degree: int | None = None) -> torch.Tensor:
    """sum_k shs[:, c, k] * Y_k(dirs): [N,C,K], [N,3] -> [N,C].

    K must be 1, 4, 9 or 16.  ``degree`` (optional) drops coefficients
    k >= (degree+1)^2, the effect of the SH warm-up mask at msplat.py:98-101.
    """
    shs = shs.to(F64)
    dirs = dirs.to(F64)
    k = shs.shape[-1]
    assert k in (1, 4, 9, 16)
    if degree is not None:
        k = min(k, (degree + 1) ** 2)
    basis = sh_basis(dirs, k)                      # [N,k]
    return (shs[:, :, :k] * basis[:, None, :]).sum(-1)


# ----------------------------------------------------------------------------
# Camera helpers
# ----------------------------------------------------------------------------
def _intr(intr):
    intr = intr.to(F64).reshape(-1)
    return intr[0], intr[1], intr[2], intr[3]


def _extr(extr):
    extr = extr.to(F64)
    if extr.shape[-2] == 4:
        extr = extr[..., :3, :]
    return extr[:, :3], extr[:, 3]


def _cam_points(xyz, extr):
    R, t = _extr(extr)
    return xyz.to(F64) @ R.T + t


# ----------------------------------------------------------------------------
# project_point (msplat.py:109-113)
# ----------------------------------------------------------------------------
def project_point(xyz, intr, extr, W, H, nearest=0.2, extent=1.3):
    """World points -> (uv [N,2], depth [N,1]); culled points get 0 for both."""
    fx, fy, cx, cy = _intr(intr)
    p = _cam_points(xyz, extr)
    z = p[:, 2]
    front = z > nearest
    zs = torch.where(front, z, torch.ones_like(z))
    u = fx * p[:, 0] / zs + cx
    v = fy * p[:, 1] / zs + cy
    lo_x, hi_x = 0.5 * W * (1 - extent), 0.5 * W * (1 + extent)
    lo_y, hi_y = 0.5 * H * (1 - extent), 0.5 * H * (1 + extent)
    with torch.no_grad():
        vis = front & (u >= lo_x) & (u <= hi_x) & (v >= lo_y) & (v <= hi_y)
    zero = torch.zeros_like(u)
    uv = torch.stack([torch.where(vis, u, zero), torch.where(vis, v, zero)], dim=1)
    depth = torch.where(vis, z, zero)[:, None]
    return uv, depth


# ----------------------------------------------------------------------------
# compute_cov3d (msplat.py:118)
# ----------------------------------------------------------------------------
def quat_to_rotmat(q):
    """wxyz quaternion -> rotation matrix, the formula of pointrix/utils/pose.py:40-84."""
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    x2, y2, z2, w2 = x * x, y * y, z * z, w * w
    xy, zw, xz, yw, yz, xw = x * y, z * w, x * z, y * w, y * z, x * w
    r0 = torch.stack([x2 - y2 - z2 + w2, 2 * (xy - zw), 2 * (xz + yw)], dim=1)
    r1 = torch.stack([2 * (xy + zw), -x2 + y2 - z2 + w2, 2 * (yz - xw)], dim=1)
    r2 = torch.stack([2 * (xz - yw), 2 * (yz + xw), -x2 - y2 + z2 + w2], dim=1)
    return torch.stack([r0, r1, r2], dim=1)


def compute_cov3d(scale, rot, visible):
    """Sigma = R diag(s)^2 R^T, packed (xx, xy, xz, yy, yz, zz); 0 where not visible."""
    R = quat_to_rotmat(rot.to(F64))
    M = R * scale.to(F64)[:, None, :]
    S = M @ M.transpose(1, 2)
    cov = torch.stack([S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]], 1)
    vis = visible.reshape(-1).to(torch.bool)
    return torch.where(vis[:, None], cov, torch.zeros_like(cov))


def unpack_cov3d(cov):
    a, b, c, d, e, f = cov.unbind(1)
    return torch.stack([torch.stack([a, b, c], 1), torch.stack([b, d, e], 1),
                        torch.stack([c, e, f], 1)], 1)


# ----------------------------------------------------------------------------
# Tile rectangle shared by ewa_project (tiles touched) and sort_gaussian
# ----------------------------------------------------------------------------
def tile_grid(W, H, tw=TILE_W, th=TILE_H):
    return (W + tw - 1) // tw, (H + th - 1) // th


def tile_rect(uv, radius, W, H, tw=TILE_W, th=TILE_H):
    """[x0, x1) x [y0, y1) tile rectangle of each splat, as int64 tensors."""
    gx, gy = tile_grid(W, H, tw, th)
    u = uv[:, 0].to(F64)
    v = uv[:, 1].to(F64)
    r = radius.to(F64)
    x0 = torch.clamp((u - r) / tw, 0, gx).floor().long()
    x1 = torch.clamp((u + r + tw - 1) / tw, 0, gx).floor().long()
    y0 = torch.clamp((v - r) / th, 0, gy).floor().long()
    y1 = torch.clamp((v + r + th - 1) / th, 0, gy).floor().long()
    return x0, x1, y0, y1


# ----------------------------------------------------------------------------
# ewa_project (msplat.py:121-130)
# ----------------------------------------------------------------------------
def ewa_project(xyz, cov3d, intr, extr, uv, W, H, visible):
    """EWA splat: -> conic [N,3] (inverse 2D covariance), radius [N] int32, tiles [N] int32."""
    fx, fy, cx, cy = _intr(intr)
    R, _ = _extr(extr)
    p = _cam_points(xyz, extr)
    vis = visible.reshape(-1).to(torch.bool)
    z = torch.where(vis, p[:, 2], torch.ones_like(p[:, 2]))
    limx = J_CLAMP * 0.5 * W / fx
    limy = J_CLAMP * 0.5 * H / fy
    tx = torch.clamp(p[:, 0] / z, -limx, limx) * z
    ty = torch.clamp(p[:, 1] / z, -limy, limy) * z
    zero = torch.zeros_like(z)
    J = torch.stack([torch.stack([fx / z, zero, -fx * tx / (z * z)], 1),
                     torch.stack([zero, fy / z, -fy * ty / (z * z)], 1)], 1)   # [N,2,3]
    Tm = J @ R                                                                  # [N,2,3]
    S = unpack_cov3d(cov3d.to(F64))
    c2 = Tm @ S @ Tm.transpose(1, 2)
    a = c2[:, 0, 0] + LOWPASS
    b = c2[:, 0, 1]
    c = c2[:, 1, 1] + LOWPASS
    det = a * c - b * b
    ok = vis & (det > 0)
    dets = torch.where(ok, det, torch.ones_like(det))
    conic = torch.stack([c / dets, -b / dets, a / dets], 1)
    with torch.no_grad():
        mid = 0.5 * (a + c)
        lam = mid + torch.sqrt(torch.clamp(mid * mid - det, min=0.1))
        radius = torch.ceil(3.0 * torch.sqrt(torch.clamp(lam, min=0))).long()
        radius = torch.where(ok, radius, torch.zeros_like(radius))
        x0, x1, y0, y1 = tile_rect(uv, radius, W, H)
        tiles = (x1 - x0) * (y1 - y0)
        ok = ok & (tiles > 0)
        radius = torch.where(ok, radius, torch.zeros_like(radius))
        tiles = torch.where(ok, tiles, torch.zeros_like(tiles))
    conic = torch.where(ok[:, None], conic, torch.zeros_like(conic))
    return conic, radius.to(torch.int32), tiles.to(torch.int32)


# ----------------------------------------------------------------------------
# sort_gaussian (msplat.py:133-135) -- integer work, numpy
# ----------------------------------------------------------------------------
def emit_keys(uv, depth, W, H, radius, tw=TILE_W, th=TILE_H):
    """Emit (key, gaussian id) pairs in Gaussian-index order, tiles row-major."""
    gx, _ = tile_grid(W, H, tw, th)
    radius = radius.reshape(-1)
    x0, x1, y0, y1 = tile_rect(uv, radius, W, H, tw, th)
    w = (x1 - x0).clamp(min=0)
    cnt = torch.where(radius > 0, w * (y1 - y0).clamp(min=0), torch.zeros_like(w))
    n = cnt.numel()
    gid = torch.repeat_interleave(torch.arange(n), cnt)
    start = torch.cumsum(cnt, 0) - cnt
    local = torch.arange(int(cnt.sum())) - start[gid]
    wg = w[gid].clamp(min=1)
    tx = x0[gid] + local % wg
    ty = y0[gid] + local // wg
    tile = (ty * gx + tx).numpy().astype(np.uint64)
    dbits = depth.reshape(-1).to(torch.float32).numpy().view(np.uint32).astype(np.uint64)
    keys = (tile << np.uint64(32)) | dbits[gid.numpy()]
    return keys, gid.numpy().astype(np.int32)


def sort_gaussian(uv, depth, W, H, radius, tiles, tw=TILE_W, th=TILE_H):
    """-> (gaussian_ids_sorted [M] int32, tile_range [T,2] int32, [first, last+1), (0,0) if empty)."""
    gx, gy = tile_grid(W, H, tw, th)
    keys, vals = emit_keys(uv, depth, W, H, radius, tw, th)
    order = np.argsort(keys, kind="stable")
    ids = vals[order]
    tiles_sorted = (keys[order] >> np.uint64(32)).astype(np.int64)
    T = gx * gy
    rng = np.zeros((T, 2), np.int32)
    if ids.size:
        first = np.searchsorted(tiles_sorted, np.arange(T), side="left")
        last = np.searchsorted(tiles_sorted, np.arange(T), side="right")
        nz = last > first
        rng[nz, 0] = first[nz]
        rng[nz, 1] = last[nz]
    return torch.from_numpy(ids), torch.from_numpy(rng)


# ----------------------------------------------------------------------------
# alpha_blending (msplat.py:148-151)
# ----------------------------------------------------------------------------
def alpha_blending(uv, conic, opacity, feature, ids, tile_range, bg, W, H, ndc=None,
                   return_aux=False, tw=TILE_W, th=TILE_H):
    """Front-to-back compositing -> [C,H,W] (+ T_final [H,W], n_contrib [H,W]).

    Vectorised per tile: alpha matrix [pixels, splats], transmittance by
    cumulative product.  Skipped splats act as alpha = 0; the T < 1e-4 stop is
    a prefix mask because the inclusive product never increases.
    """
    gx, gy = tile_grid(W, H, tw, th)
    uv = uv.to(F64)
    if ndc is not None:
        uv = uv + ndc.to(F64) * torch.tensor([0.5 * W, 0.5 * H], dtype=F64)
    conic = conic.to(F64)
    op = opacity.to(F64).reshape(-1)
    feat = feature.to(F64)
    C = feat.shape[1]
    out = torch.zeros(C, H, W, dtype=F64)
    Tfin = torch.ones(H, W, dtype=F64)
    ncon = torch.zeros(H, W, dtype=torch.int32)
    rows_out, rows_T = [], []
    tr = tile_range.to(torch.int64)
    ids = ids.to(torch.int64)
    out_tiles = {}
    for ty in range(gy):
        for tx in range(gx):
            t = ty * gx + tx
            s, e = int(tr[t, 0]), int(tr[t, 1])
            xs = torch.arange(tx * tw, min((tx + 1) * tw, W))
            ys = torch.arange(ty * th, min((ty + 1) * th, H))
            py, px = torch.meshgrid(ys, xs, indexing="ij")
            px = px.reshape(-1).to(F64)
            py = py.reshape(-1).to(F64)
            if e <= s:
                out_tiles[t] = (ys, xs, None, None, None)
                continue
            g = ids[s:e]
            dx = uv[g, 0][None, :] - px[:, None]
            dy = uv[g, 1][None, :] - py[:, None]
            a, b, c = conic[g, 0][None, :], conic[g, 1][None, :], conic[g, 2][None, :]
            power = -0.5 * (a * dx * dx + c * dy * dy) - b * dx * dy
            alpha = torch.clamp(op[g][None, :] * torch.exp(power), max=ALPHA_MAX)
            with torch.no_grad():
                keep = (power <= 0) & (alpha >= ALPHA_MIN)
            ae = torch.where(keep, alpha, torch.zeros_like(alpha))
            with torch.no_grad():
                incl = torch.cumprod(1 - ae, dim=1)
                live = incl >= T_MIN
                contrib = keep & live
            ae = torch.where(contrib, ae, torch.zeros_like(ae))
            one_m = 1 - ae
            incl = torch.cumprod(one_m, dim=1)
            excl = torch.cat([torch.ones_like(incl[:, :1]), incl[:, :-1]], 1)
            w = ae * excl
            col = w @ feat[g]                                  # [P, C]
            tfin = incl[:, -1]
            with torch.no_grad():
                idx = torch.arange(1, e - s + 1)[None, :].expand_as(contrib)
                nc = torch.where(contrib, idx, torch.zeros_like(idx)).max(1).values
            out_tiles[t] = (ys, xs, col, tfin, nc)
    # assemble (keeps autograd graph)
    pieces_c = []
    for ty in range(gy):
        row_c, row_t = [], []
        for tx in range(gx):
            ys, xs, col, tfin, nc = out_tiles[ty * gx + tx]
            h, w = len(ys), len(xs)
            if col is None:
                row_c.append(torch.full((C, h, w), float(bg), dtype=F64))
                row_t.append(torch.ones(h, w, dtype=F64))
            else:
                tf = tfin.reshape(h, w)
                row_c.append(col.T.reshape(C, h, w) + float(bg) * tf[None])
                row_t.append(tf)
                ncon[ys[0]:ys[-1] + 1, xs[0]:xs[-1] + 1] = nc.reshape(h, w).to(torch.int32)
        pieces_c.append((torch.cat(row_c, 2), torch.cat(row_t, 1)))
    out = torch.cat([p[0] for p in pieces_c], 1)
    Tfin = torch.cat([p[1] for p in pieces_c], 0)
    if return_aux:
        return out, Tfin, ncon
    return out


def alpha_blending_bruteforce(uv, conic, opacity, feature, ids, tile_range, bg, W, H,
                              tw=TILE_W, th=TILE_H):
    """Literal per-pixel loop in Python floats (FP64); for tiny inputs only.

    Pins the vectorised ``alpha_blending``: the 3DGS compositing loop with
    early stop, written the obvious way.
    """
    gx, _ = tile_grid(W, H, tw, th)
    uvl = uv.double().tolist()
    con = conic.double().tolist()
    op = opacity.double().reshape(-1).tolist()
    ft = feature.double().tolist()
    idl = ids.tolist()
    tr = tile_range.tolist()
    C = len(ft[0]) if ft else feature.shape[1]
    out = [[[0.0] * W for _ in range(H)] for _ in range(C)]
    Tf = [[1.0] * W for _ in range(H)]
    nc = [[0] * W for _ in range(H)]
    for y in range(H):
        for x in range(W):
            t = (y // th) * gx + (x // tw)
            s, e = tr[t]
            T = 1.0
            acc = [0.0] * C
            last = 0
            for j in range(s, e):
                g = idl[j]
                dx = uvl[g][0] - x
                dy = uvl[g][1] - y
                a, b, c = con[g]
                power = -0.5 * (a * dx * dx + c * dy * dy) - b * dx * dy
                if power > 0:
                    continue
                alpha = min(ALPHA_MAX, op[g] * math.exp(power))
                if alpha < ALPHA_MIN:
                    continue
                test_T = T * (1 - alpha)
                if test_T < T_MIN:
                    break
                for ch in range(C):
                    acc[ch] += ft[g][ch] * alpha * T
                T = test_T
                last = j - s + 1
            for ch in range(C):
                out[ch][y][x] = acc[ch] + bg * T
            Tf[y][x] = T
            nc[y][x] = last
    return (torch.tensor(out, dtype=F64), torch.tensor(Tf, dtype=F64),
            torch.tensor(nc, dtype=torch.int32))


def splat_normals(position, scaling, rotation, camera_center, extrinsic_matrix):
    """Per-splat camera-space normal: shortest scale axis of R(q), flipped to
    face the camera, rotated by the world-to-camera R
    (examples/supervise/model.py:50-74)."""
    pos = position.to(F64)
    R = quat_to_rotmat(rotation.to(F64))
    idx = torch.argmin(scaling.detach(), dim=-1)
    onehot = torch.nn.functional.one_hot(idx, 3).to(F64)
    n = torch.nn.functional.normalize((R @ onehot[:, :, None])[:, :, 0], dim=-1)
    d = camera_center.to(F64).reshape(1, 3) - pos.detach()
    d = d / d.norm(dim=1, keepdim=True)
    flip = (d * n.detach()).sum(-1) < 0
    n = torch.where(flip[:, None], -n, n)
    return n @ extrinsic_matrix.to(F64)[:3, :3].T


# ----------------------------------------------------------------------------
# MsplatRender.render_iter host glue (msplat.py:51-158)
# ----------------------------------------------------------------------------
def render_iter(height, width, extrinsic_matrix, intrinsic_params, camera_center,
                position, opacity, scaling, rotation, shs, sh_degree=3, bg=0.0,
                render_depth=False, normals=None, add_half=True, render_opacity=False,
                return_aux=False, snap=None):
    """FP64 transcription of MsplatRender.render_iter.

    Returns the same dict as the reference (rendered_features_split, uv_points,
    visibility, radii), plus ``opacity`` (1 - T_final) under ``aux``.

    ``snap`` (optional dict with "uv" and/or "conic") moves the forward value
    of those intermediates onto the given tensors while keeping the FP64
    derivative (x + (given - x).detach()): the oracle is then differentiated
    at the same forward point as an FP32 run, which separates the kernels'
    arithmetic from the rounding of their FP32 inputs.
    """
    pos = position.to(F64)
    cc = camera_center.to(F64).reshape(1, 3)
    direction = pos - cc
    direction = direction / direction.norm(dim=1, keepdim=True)
    sh_coeff = shs.to(F64).permute(0, 2, 1)
    rgb = compute_sh(sh_coeff, direction, degree=sh_degree)
    if add_half:
        rgb = (rgb + 0.5).clamp(min=0.0)
    extr = extrinsic_matrix[:3, :]
    uv, depth = project_point(pos, intrinsic_params, extr, width, height, nearest=0.2)
    if snap is not None and "uv" in snap:
        uv = uv + (snap["uv"].to(F64) - uv).detach()
    visible = depth != 0
    cov3d = compute_cov3d(scaling, rotation, visible)
    conic, radius, tiles = ewa_project(pos, cov3d, intrinsic_params, extr, uv, width, height, visible)
    if snap is not None and "conic" in snap:
        conic = conic + (snap["conic"].to(F64) - conic).detach()
    ids, tr = sort_gaussian(uv.detach(), depth.detach(), width, height, radius, tiles)
    feats = {"rgb": rgb}
    if render_depth:
        feats["depth"] = depth
    if normals is not None:
        feats["normal"] = normals.to(F64)
    feat = torch.cat(list(feats.values()), dim=-1)
    ndc = torch.zeros_like(uv, requires_grad=True)
    out, tfin, ncon = alpha_blending(uv, conic, opacity, feat, ids, tr, bg, width, height,
                                     ndc, return_aux=True)
    split, start = {}, 0
    for k, v in feats.items():
        split[k] = out[start:start + v.shape[-1]]
        start += v.shape[-1]
    if "normal" in split:
        # [-1,1] -> [0,1] as examples/supervise/renderer.py:89-96
        nrm = split["normal"]
        split["normal"] = (nrm / nrm.norm(dim=0, keepdim=True) + 1) / 2
    if render_opacity:
        split["opacity"] = (1.0 - tfin)[None]
    res = {"rendered_features_split": split, "uv_points": ndc,
           "visibility": radius > 0, "radii": radius}
    if return_aux:
        res["aux"] = {"opacity": 1.0 - tfin, "T_final": tfin, "n_contrib": ncon,
                      "ids": ids, "tile_range": tr, "uv": uv, "depth": depth,
                      "conic": conic, "cov3d": cov3d, "rgb": rgb}
    return res
