"""Fused L1 + SSIM (ops.l1_ssim, ops.l1_ssim_loss) against the FP64
restatement of pointrix's losses (oracle/loss_ref.py): values within 1e-6,
gradients (FP64 autograd of the restatement) within 1e-4 of max |g|, on
ragged sizes, several planes, and a graph capture replayed against eager."""
import pytest
import torch

from oracle import loss_ref as L
from pointrix_amd import ops

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(3, 37, 53), (2, 3, 64, 48), (1, 1, 7, 5), (3, 822, 1237)])
def test_l1_ssim_matches_fp64(device, shape):
    g = torch.Generator().manual_seed(sum(shape))
    x = torch.rand(shape, generator=g)
    y = (x + 0.2 * torch.randn(shape, generator=g)).clamp(0, 1)
    xd = x.to(device).requires_grad_(True)
    out = ops.l1_ssim_loss(xd, y.to(device), lambda_ssim=0.2)
    out["loss"].backward()
    x64 = x.double().requires_grad_(True)
    ref = L.loss_dict(x64, y.double(), 0.2)
    ref["loss"].backward()
    for k in ("loss", "L1_loss", "ssim_loss"):
        assert abs(float(out[k]) - float(ref[k])) < 1e-6, (k, float(out[k]), float(ref[k]))
    gd, gr = xd.grad.double().cpu(), x64.grad
    err = float((gd - gr).abs().max() / gr.abs().max())
    assert err < 1e-4, err


def test_l1_ssim_graph_replay(device):
    g = torch.Generator().manual_seed(5)
    x = torch.rand(3, 96, 80, generator=g).to(device).requires_grad_(True)
    y = torch.rand(3, 96, 80, generator=g).to(device)

    def step():
        x.grad = None if x.grad is None else x.grad.zero_()
        loss = ops.l1_ssim_loss(x, y)["loss"]
        loss.backward()
        return loss
    l0 = float(step())
    g0 = x.grad.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        loss = step()
    for _ in range(3):
        graph.replay()
        torch.cuda.synchronize()
        assert float(loss) == l0
        assert torch.equal(x.grad, g0)
        y.mul_(1.0)   # eager work between replays
