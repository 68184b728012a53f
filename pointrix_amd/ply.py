"""Scene files: PLY load/save with pointrix's attribute layout.

pointrix saves a Gaussian point cloud as one PLY ``vertex`` element of float32
properties (PointCloud.list_of_attributes / save_ply / load_ply,
pointrix/model/point_cloud/points.py:359-426): ``x, y, z, nx, ny, nz`` (zero
normals), then every other registered attribute flattened row-major as
``{name}_{i}`` in registration order -- for GaussianPointCloud
(gaussian_points.py:28-51) ``features`` [N,1,3], ``features_rest`` [N,K-1,3],
``scaling`` [N,3] (log), ``rotation`` [N,4] (unnormalised wxyz), ``opacity``
[N,1] (logit).  This module reads and writes that layout without the plyfile
package: binary little-endian (what pointrix writes) and ASCII on reading,
any scalar property type on reading, float32 on writing.  Loading streams
the vertex block through numpy, so multi-million-splat scenes load at disk
speed into a resident scene for rendering or serving.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from .scene import GaussianScene

_TYPES = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2",
          "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4",
          "uint32": "u4", "float": "f4", "float32": "f4", "double": "f8", "float64": "f8"}

ATTRIBUTES = ("features", "features_rest", "scaling", "rotation", "opacity")


def attribute_names(sh_coeffs: int) -> list:
    """The property names pointrix writes for a cloud with ``sh_coeffs``
    SH coefficients per channel (points.py:359-369)."""
    names = ["x", "y", "z", "nx", "ny", "nz"]
    sizes = {"features": 3, "features_rest": 3 * (sh_coeffs - 1), "scaling": 3, "rotation": 4,
             "opacity": 1}
    for a in ATTRIBUTES:
        names += [f"{a}_{i}" for i in range(sizes[a])]
    return names


def save_ply(path, scene: GaussianScene) -> None:
    """Write ``scene`` (raw parameters) in pointrix's layout (points.py:371-393)."""
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    n = len(scene)
    k = 1 + (0 if scene.features_rest is None else scene.features_rest.shape[1])
    cols = [scene.position.detach().float().cpu().numpy().reshape(n, 3),
            np.zeros((n, 3), np.float32)]
    for a in ATTRIBUTES:
        t = getattr(scene, a)
        if t is None or t.numel() == 0:
            continue
        cols.append(t.detach().float().cpu().numpy().reshape(n, -1))
    data = np.ascontiguousarray(np.concatenate(cols, axis=1), dtype="<f4")
    names = attribute_names(k)
    assert data.shape[1] == len(names)
    header = ["ply", "format binary_little_endian 1.0", f"element vertex {n}"]
    header += [f"property float {p}" for p in names]
    header += ["end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(header) + "\n").encode("ascii"))
        f.write(data.tobytes())


def _read_header(f):
    line = f.readline().strip()
    if line != b"ply":
        raise ValueError("not a PLY file")
    fmt, elements, cur = None, [], None
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header without end_header")
        tok = line.decode("ascii").split()
        if not tok or tok[0] in ("comment", "obj_info"):
            continue
        if tok[0] == "end_header":
            break
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            cur = {"name": tok[1], "count": int(tok[2]), "props": []}
            elements.append(cur)
        elif tok[0] == "property":
            if tok[1] == "list":
                raise ValueError("PLY list properties are not supported")
            cur["props"].append((tok[2], _TYPES[tok[1]]))
    return fmt, elements


def read_vertices(path) -> dict:
    """{property name: 1-D numpy array} of the ``vertex`` element."""
    with open(path, "rb") as f:
        fmt, elements = _read_header(f)
        if not elements or elements[0]["name"] != "vertex":
            raise ValueError("the first PLY element must be 'vertex'")
        el = elements[0]
        n = el["count"]
        if fmt == "ascii":
            arr = np.loadtxt(f, dtype=np.float64, max_rows=n, ndmin=2)
            return {p: arr[:, i] for i, (p, _) in enumerate(el["props"])}
        order = {"binary_little_endian": "<", "binary_big_endian": ">"}[fmt]
        dt = np.dtype([(p, order + t) for p, t in el["props"]])
        rec = np.fromfile(f, dtype=dt, count=n)
        if rec.shape[0] != n:
            raise ValueError(f"PLY truncated: {rec.shape[0]} of {n} vertices")
        return {p: rec[p] for p, _ in el["props"]}


def load_ply(path) -> GaussianScene:
    """Read a pointrix Gaussian PLY into raw parameters (points.py:395-426):
    position [N,3], features [N,1,3], features_rest [N,K-1,3], scaling [N,3],
    rotation [N,4], opacity [N,1]; float32, on the host."""
    v = read_vertices(path)

    def stack(prefix):
        keys = sorted((k for k in v if k.startswith(prefix + "_") and
                       k[len(prefix) + 1:].isdigit()), key=lambda k: int(k[len(prefix) + 1:]))
        if not keys:
            return None
        return np.stack([np.asarray(v[k], np.float32) for k in keys], axis=1)

    pos = np.stack([np.asarray(v[c], np.float32) for c in ("x", "y", "z")], axis=1)
    n = pos.shape[0]
    feats = stack("features")
    rest = stack("features_rest")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a))  # noqa: E731
    return GaussianScene(
        t(pos), t(feats.reshape(n, 1, 3)),
        t(rest.reshape(n, -1, 3)) if rest is not None else torch.zeros(n, 0, 3),
        t(stack("scaling")), t(stack("rotation")), t(stack("opacity")))
