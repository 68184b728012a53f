"""Restatement of pointrix's densification bookkeeping (test infrastructure):
accumulate_viewspace_grad (pointrix/controller/gs.py:259-284) and the stat
updates of DensificationController.preprocess (gs.py:318-333), in plain torch
on the CPU, statement for statement."""
import torch


def viewspace_grad(view_grads, width, height, normalize_grad=True):
    """gs.py:271-284: sum of the views' uv_points grads, then scaled."""
    g = torch.zeros_like(view_grads[0])
    for vp in view_grads:
        g += vp.clone()
    if normalize_grad:
        g[..., 0] *= width / 2.0
        g[..., 1] *= height / 2.0
    return g


def update(grad_accum, acc_steps, max_radii, view_grads, visibility, radii, width, height,
           normalize_grad=True):
    """gs.py:318-333 on the given buffers (modified in place)."""
    point_grad = viewspace_grad(view_grads, width, height, normalize_grad)
    sel = visibility
    max_radii[sel] = torch.max(max_radii[sel], radii[sel].to(max_radii.dtype))
    grad_accum[sel] += torch.norm(point_grad[sel, :2], dim=-1, keepdim=True)
    acc_steps[sel] += 1


# --- densify / prune (controller/gs.py:91-247; point_utils.py:51-108) -------
def _quat_to_rotmat(r):
    """pose.py:40-117: normalise, then the unit-quaternion matrix (wxyz)."""
    q = r / torch.sqrt(r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1] + r[:, 2] * r[:, 2] +
                       r[:, 3] * r[:, 3])[:, None]
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    m = torch.empty(q.shape[0], 3, 3, dtype=q.dtype)
    m[:, 0, 0] = x * x - y * y - z * z + w * w
    m[:, 1, 0] = 2 * (x * y + z * w)
    m[:, 2, 0] = 2 * (x * z - y * w)
    m[:, 0, 1] = 2 * (x * y - z * w)
    m[:, 1, 1] = -x * x + y * y - z * z + w * w
    m[:, 2, 1] = 2 * (y * z + x * w)
    m[:, 0, 2] = 2 * (x * z + y * w)
    m[:, 1, 2] = 2 * (y * z - x * w)
    m[:, 2, 2] = -x * x - y * y + z * z + w * w
    return m


def densify_and_split(params, states, grads, threshold, percent_dense, extent, split_num,
                      samples):
    """densify_clone then densify_split on plain tensors: params {name: [N,...]},
    states {name: (exp_avg, exp_avg_sq)}; samples are the standard normals
    torch.normal would draw (std applied here).  Returns (params, states)."""
    params = dict(params)
    states = {k: list(v) for k, v in states.items()}
    scaling = torch.exp(params["scaling"])
    # densify_clone (gs.py:186-199): select, extend (zero moments for new rows)
    mask = torch.where(torch.norm(grads, dim=-1) >= threshold, True, False)
    mask = torch.logical_and(mask, torch.max(scaling, dim=1).values <= percent_dense * extent)
    for k in params:
        sel = params[k][mask]
        params[k] = torch.cat([params[k], sel])
        states[k] = [torch.cat([s, torch.zeros_like(sel)]) for s in states[k]]
    # densify_split (gs.py:201-241)
    n = params["position"].shape[0]
    scaling = torch.exp(params["scaling"])
    padded = torch.zeros(n, dtype=grads.dtype)
    padded[:grads.shape[0]] = grads.squeeze()
    mask = torch.where(padded >= threshold, True, False)
    mask = torch.logical_and(mask, torch.max(scaling, dim=1).values > percent_dense * extent)
    stds = scaling[mask].repeat(split_num, 1)
    s = samples * stds
    rots = _quat_to_rotmat(params["rotation"][mask]).repeat(split_num, 1, 1)
    new_pos = torch.bmm(rots, s.unsqueeze(-1)).squeeze(-1) + \
        params["position"][mask].repeat(split_num, 1)
    new_scaling = torch.log(scaling[mask].repeat(split_num, 1) / (0.8 * split_num))
    new = {k: params[k][mask].repeat(*([split_num] + [1] * (params[k].dim() - 1)))
           for k in params}
    new["position"], new["scaling"] = new_pos, new_scaling
    for k in params:
        params[k] = torch.cat([params[k], new[k]])
        states[k] = [torch.cat([st, torch.zeros_like(new[k])]) for st in states[k]]
    valid = ~torch.cat((mask, torch.zeros(split_num * int(mask.sum()), dtype=torch.bool)))
    for k in params:
        params[k] = params[k][valid]
        states[k] = [st[valid] for st in states[k]]
    return params, states
