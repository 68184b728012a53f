"""The densification-statistics restatement (oracle/densify_ref.py) on the
host: hand-computed cases -- the views' gradients are summed before the norm,
normalize_grad scales x by W/2 and y by H/2, invisible splats are untouched,
max_radii keeps the running maximum."""
import torch

from oracle import densify_ref as D


def test_hand_case():
    ga, st, mr = torch.zeros(3, 1), torch.zeros(3, 1), torch.zeros(3)
    g1 = torch.tensor([[3.0, 0.0], [1.0, 1.0], [5.0, 5.0]])
    g2 = torch.tensor([[0.0, 4.0], [-1.0, -1.0], [5.0, 5.0]])
    vis = torch.tensor([True, True, False])
    rad = torch.tensor([7, 2, 9], dtype=torch.int32)
    D.update(ga, st, mr, [g1, g2], vis, rad, 2, 2)      # W/2 = H/2 = 1
    assert torch.equal(ga[:, 0], torch.tensor([5.0, 0.0, 0.0]))   # |(3,4)|, |(0,0)|
    assert torch.equal(st[:, 0], torch.tensor([1.0, 1.0, 0.0]))
    assert torch.equal(mr, torch.tensor([7.0, 2.0, 0.0]))
    rad2 = torch.tensor([3, 4, 1], dtype=torch.int32)
    D.update(ga, st, mr, [g1], vis, rad2, 4, 6)           # scale (2, 3)
    assert torch.allclose(ga[:, 0], torch.tensor([5.0 + 6.0, (4.0 + 9.0) ** 0.5, 0.0]))
    assert torch.equal(st[:, 0], torch.tensor([2.0, 2.0, 0.0]))
    assert torch.equal(mr, torch.tensor([7.0, 4.0, 0.0]))


def test_no_normalize():
    ga, st, mr = torch.zeros(1, 1), torch.zeros(1, 1), torch.zeros(1)
    D.update(ga, st, mr, [torch.tensor([[0.6, 0.8]])], torch.tensor([True]),
             torch.tensor([1], dtype=torch.int32), 1000, 1000, normalize_grad=False)
    assert torch.allclose(ga, torch.tensor([[1.0]]))
