"""Scene files in pointrix's PLY layout (pointrix_amd.ply), host only:
property names and order as PointCloud.list_of_attributes writes them,
a save -> load round trip bit for bit, an ASCII file read, and a hand-built
header checked byte for byte."""
import numpy as np
import torch

from pointrix_amd import ply
from pointrix_amd import scene as S


def test_names():
    names = ply.attribute_names(16)
    assert names[:6] == ["x", "y", "z", "nx", "ny", "nz"]
    assert names[6:9] == ["features_0", "features_1", "features_2"]
    assert names[9] == "features_rest_0" and names[9 + 44] == "features_rest_44"
    assert names[-8:] == ["scaling_0", "scaling_1", "scaling_2", "rotation_0", "rotation_1",
                          "rotation_2", "rotation_3", "opacity_0"]
    assert len(names) == 6 + 3 + 45 + 3 + 4 + 1


def test_round_trip(tmp_path):
    sc = S.random_scene(1000, 3, sh_degree=3)
    p = tmp_path / "pc" / "point_cloud.ply"
    ply.save_ply(str(p), sc)
    raw = p.read_bytes()
    head = raw[:raw.index(b"end_header\n") + len(b"end_header\n")].decode()
    assert head.startswith("ply\nformat binary_little_endian 1.0\nelement vertex 1000\n")
    assert len(raw) - len(head) == 1000 * 62 * 4
    back = ply.load_ply(str(p))
    for k, a in sc.named_params().items():
        b = back.named_params()[k]
        assert a.shape == b.shape, k
        assert torch.equal(a.float(), b), k
    v = ply.read_vertices(str(p))
    assert np.all(v["nx"] == 0) and v["rotation_3"].dtype == np.float32


def test_ascii_and_sh0(tmp_path):
    p = tmp_path / "a.ply"
    rows = [[1, 2, 3, 0, 0, 0, 0.1, 0.2, 0.3, -1, -2, -3, 1, 0, 0, 0, 0.5],
            [4, 5, 6, 0, 0, 0, 0.4, 0.5, 0.6, -4, -5, -6, 0, 1, 0, 0, -0.5]]
    names = ply.attribute_names(1)
    assert len(names) == len(rows[0])
    txt = "ply\nformat ascii 1.0\ncomment test\nelement vertex 2\n"
    txt += "".join(f"property float {n}\n" for n in names) + "end_header\n"
    txt += "".join(" ".join(str(x) for x in r) + "\n" for r in rows)
    p.write_text(txt)
    sc = ply.load_ply(str(p))
    assert sc.features_rest.shape == (2, 0, 3)
    assert torch.allclose(sc.position, torch.tensor([[1., 2, 3], [4, 5, 6]]))
    assert torch.allclose(sc.opacity, torch.tensor([[0.5], [-0.5]]))
    assert torch.allclose(sc.rotation[1], torch.tensor([0., 1, 0, 0]))
