"""Fused Adam (pointrix_amd.optim.FusedAdam) against torch.optim.Adam on the
CPU in float64 (the same update; the library routine is the oracle):
several groups with their own lr (pointrix's per-attribute groups, eps
1e-15), ragged sizes, a tensor whose gradient stays zero, a parameter
without a gradient, six steps; also the state keys pointrix's densify code
reads, and the in-kernel zero_grad."""
import pytest
import torch

from pointrix_amd.optim import FusedAdam

pytestmark = pytest.mark.gpu

LRS = {"position": 1.6e-4, "features": 2.5e-3, "features_rest": 1.25e-4, "scaling": 5e-3,
       "rotation": 1e-3, "opacity": 5e-2}
SHAPES = {"position": (1001, 3), "features": (1001, 1, 3), "features_rest": (1001, 15, 3),
          "scaling": (1001, 3), "rotation": (1001, 4), "opacity": (1001, 1)}


def make(device, dtype, seed=3):
    g = torch.Generator().manual_seed(seed)
    return {k: torch.randn(s, generator=g).to(device=device, dtype=dtype).requires_grad_(True)
            for k, s in SHAPES.items()}


@pytest.mark.parametrize("zero_in_step", [False, True])
def test_fused_adam_matches_torch(device, zero_in_step):
    pg = make(device, torch.float32)
    pc = make("cpu", torch.float64)
    spare = torch.zeros(5, device=device, requires_grad=True)   # never gets a gradient
    opt_g = FusedAdam([{"params": [pg[k]], "lr": LRS[k], "name": k} for k in pg] +
                      [{"params": [spare], "lr": 1.0}], eps=1e-15,
                      zero_grad_in_step=zero_in_step)
    opt_c = torch.optim.Adam([{"params": [pc[k]], "lr": LRS[k]} for k in pc], eps=1e-15)
    gen = torch.Generator().manual_seed(11)
    for step in range(6):
        for k in pg:
            grad = torch.randn(SHAPES[k], generator=gen) * (0.0 if k == "opacity" else 1.0)
            pg[k].grad = grad.to(device)
            pc[k].grad = grad.double()
        opt_g.step()
        opt_c.step()
        if zero_in_step:
            assert all(float(pg[k].grad.abs().max()) == 0.0 for k in pg)
        for k in pg:
            d = (pg[k].detach().double().cpu() - pc[k].detach()).abs().max().item()
            # float32 rounds p once per step (<= 2^-24 |p|) and the update's
            # ratio m / sqrt(v) to ~1e-7 relative
            bound = (1e-6 * LRS[k] + 1.2e-7 * float(pc[k].abs().max())) * (step + 1)
            assert d <= bound, (step, k, d, bound)
    assert spare.grad is None and len(opt_g.state[spare]) == 0
    st = opt_g.state[pg["scaling"]]
    assert set(st) == {"step", "exp_avg", "exp_avg_sq"} and float(st["step"]) == 6
    ref = opt_c.state[pc["scaling"]]
    e = (st["exp_avg_sq"].double().cpu() - ref["exp_avg_sq"]).abs().max().item()
    assert e <= 1e-6 * float(ref["exp_avg_sq"].abs().max())
