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


def test_densify_and_split_closed_form():
    """densify_clone + densify_split (controller/gs.py:182-247) on a hand-made
    cloud, checked against what the procedure fixes without restating it:
    the row count N + clones + (split_num - 1) * splits; kept rows and their
    Adam moments unchanged and in order; clones equal their source rows with
    zero moments; split children at R(q) (s * xi) + mu with R from scipy,
    log-scale s - ln(0.8 split_num), other attributes copied, zero moments."""
    import math

    import numpy as np
    from scipy.spatial.transform import Rotation

    g = torch.Generator().manual_seed(3)
    N, split_num, thr, pd, extent = 7, 3, 0.5, 0.1, 2.0
    # grads: rows 0, 1, 3, 4, 6 over the threshold; scales: rows 0, 4 small
    # (clone), rows 1, 3 large (split), row 6 small too (clone)
    grads = torch.tensor([[0.9], [0.8], [0.1], [0.7], [0.6], [0.2], [0.55]], dtype=torch.float64)
    small, large = math.log(0.05), math.log(0.5)
    sc = torch.full((N, 3), small, dtype=torch.float64)
    sc[[1, 3]] = torch.tensor([large, small, small], dtype=torch.float64)
    sc[2, 0] = sc[5, 0] = large
    params = {"position": torch.randn(N, 3, generator=g, dtype=torch.float64),
              "scaling": sc,
              "rotation": torch.randn(N, 4, generator=g, dtype=torch.float64),
              "opacity": torch.randn(N, 1, generator=g, dtype=torch.float64),
              "features": torch.randn(N, 1, 3, generator=g, dtype=torch.float64)}
    states = {k: (torch.randn_like(v), torch.rand_like(v)) for k, v in params.items()}
    clones, splits = [0, 4, 6], [1, 3]
    xi = torch.randn(split_num * len(splits), 3, generator=g, dtype=torch.float64)
    out, st = D.densify_and_split(params, states, grads, thr, pd, extent, split_num, xi)
    n_out = N + len(clones) + (split_num - 1) * len(splits)
    assert all(v.shape[0] == n_out for v in out.values())
    kept = [i for i in range(N) if i not in splits]
    for k, v in params.items():
        assert torch.equal(out[k][:len(kept)], v[kept])
        for s_out, s_in in zip(st[k], states[k]):
            assert torch.equal(s_out[:len(kept)], s_in[kept])
        c0 = len(kept)
        assert torch.equal(out[k][c0:c0 + len(clones)], v[clones])
        for s_out in st[k]:
            assert torch.all(s_out[c0:] == 0)
    first = len(kept) + len(clones)
    for j in range(split_num * len(splits)):
        src = splits[j % len(splits)]            # children are the split rows repeated
        q = params["rotation"][src].numpy()
        R = Rotation.from_quat(np.array([q[1], q[2], q[3], q[0]]) / np.linalg.norm(q))
        s = torch.exp(params["scaling"][src]).numpy()
        want = R.as_matrix() @ (s * xi[j].numpy()) + params["position"][src].numpy()
        row = first + j
        assert np.allclose(out["position"][row].numpy(), want, atol=1e-12)
        assert torch.allclose(out["scaling"][row],
                              params["scaling"][src] - math.log(0.8 * split_num), atol=1e-12)
        for k in ("rotation", "opacity", "features"):
            assert torch.equal(out[k][row], params[k][src])
