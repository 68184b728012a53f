"""Densify / prune with optimizer-state remap (pointrix_amd.densify.densify,
prune) against the restatement of pointrix's controller
(oracle/densify_ref.densify_and_split): same rows in the same order, the
same parameter values and Adam moments (a gather: bit for bit; split
positions and scalings, computed, within 1e-6), moments of new rows zero,
and the optimizer holding the new leaves."""
import pytest
import torch

from oracle import densify_ref as D
from pointrix_amd import scene as S
from pointrix_amd.densify import DensifyStats, densify, prune
from pointrix_amd.optim import FusedAdam

pytestmark = pytest.mark.gpu


def test_densify_matches_controller(device):
    sc = S.random_scene(4000, 21, sh_degree=3, log_scale=(-4.6, 1.0))
    names = list(sc.named_params())
    host = {k: v.detach().clone() for k, v in sc.named_params().items()}
    params = {k: torch.nn.Parameter(v.to(device)) for k, v in host.items()}
    opt = FusedAdam([{"params": [params[k]], "lr": 1e-3, "name": k} for k in names], eps=1e-15)
    g = torch.Generator().manual_seed(4)
    for _ in range(2):                           # populate Adam moments
        for k in names:
            params[k].grad = torch.randn(params[k].shape, generator=g).to(device)
        opt.step()
    states = {k: (opt.state[params[k]]["exp_avg"].cpu(), opt.state[params[k]]["exp_avg_sq"].cpu())
              for k in names}
    host = {k: params[k].detach().cpu() for k in names}
    grads = torch.rand(4000, 1, generator=g) * 4e-4
    extent, pd, thr = 1.0, 0.01, 2e-4
    scaling = torch.exp(host["scaling"])
    split = (grads.reshape(-1) >= thr) & (scaling.max(1).values > pd * extent)
    ns = int(split.sum())
    assert ns > 10 and int(((grads.reshape(-1) >= thr) & ~split).sum()) > 10
    samples = torch.randn(2 * ns, 3, generator=g)
    ref_p, ref_s = D.densify_and_split(host, states, grads, thr, pd, extent, 2, samples)
    leaves, masks = densify(params, opt, grads.to(device), scaling.to(device), thr, pd, extent,
                            split_num=2, samples=samples.to(device))
    for k in names:
        got = leaves[k].detach().cpu()
        assert got.shape == ref_p[k].shape, k
        if k in ("position", "scaling"):
            assert torch.allclose(got, ref_p[k], atol=1e-6, rtol=1e-6), k
        else:
            assert torch.equal(got, ref_p[k]), k
        st = opt.state[leaves[k]]
        assert torch.equal(st["exp_avg"].cpu(), ref_s[k][0]), k
        assert torch.equal(st["exp_avg_sq"].cpu(), ref_s[k][1]), k
        assert any(p is leaves[k] for grp in opt.param_groups for p in grp["params"])
    # prune the low-opacity half, statistics along
    n = leaves["position"].shape[0]
    stats = DensifyStats(n, device, 64, 64)
    stats.grad_accum += torch.arange(n, device=device, dtype=torch.float32)[:, None]
    keep = torch.sigmoid(leaves["opacity"].detach().reshape(-1)) > 0.5
    before = {k: v.detach().clone() for k, v in leaves.items()}
    m_before = opt.state[leaves["features_rest"]]["exp_avg"].clone()
    pruned = prune(leaves, opt, keep, stats)
    for k in names:
        assert torch.equal(pruned[k].detach(), before[k][keep]), k
    assert torch.equal(opt.state[pruned["features_rest"]]["exp_avg"], m_before[keep])
    assert torch.equal(stats.grad_accum.reshape(-1),
                       torch.nonzero(keep).reshape(-1).float())
    opt.step()   # the optimizer keeps working on the new leaves
