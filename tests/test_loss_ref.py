"""Pins of the SSIM restatement (oracle/loss_ref.py), host only:
SSIM(x, x) = 1; for constant images the closed form
(2ab + C1)/(a^2 + b^2 + C1) away from the zero-padded border; the window is
the normalised 11-tap Gaussian of sigma 1.5; and a separable direct
evaluation (explicit double loops, written independently of conv2d) agrees."""
import math

import torch

from oracle import loss_ref as L


def test_identity():
    g = torch.Generator().manual_seed(1)
    x = torch.rand(2, 3, 24, 31, generator=g, dtype=torch.float64)
    assert abs(float(L.ssim(x, x)) - 1.0) < 1e-12


def test_window():
    g = L.gaussian(11, 1.5)
    assert abs(float(g.sum()) - 1.0) < 1e-15
    assert abs(float(g[5] / g[4]) - math.exp(1 / 4.5)) < 1e-12


def test_constant_images_interior():
    a, b = 0.3, 0.7
    x = torch.full((1, 1, 40, 40), a, dtype=torch.float64)
    y = torch.full((1, 1, 40, 40), b, dtype=torch.float64)
    C1, C2 = 1e-4, 9e-4
    # interior pixels (>= 5 from the border) have sigma = 0: map = (2ab+C1)/(a^2+b^2+C1)
    expect = (2 * a * b + C1) / (a * a + b * b + C1)
    w = L.create_window(11, 1)
    import torch.nn.functional as F
    mu1 = F.conv2d(x, w, padding=5)
    assert abs(float(mu1[0, 0, 20, 20]) - a) < 1e-14
    full = float(L.ssim(x, y))
    # the mean includes the border ring; the interior value bounds it from one side
    assert abs(full - expect) < 0.2
    x2 = torch.full((1, 1, 10, 10), a, dtype=torch.float64)
    assert abs(float(L.ssim(x2, x2)) - 1.0) < 1e-12


def test_direct_evaluation():
    g = torch.Generator().manual_seed(2)
    x = torch.rand(1, 1, 9, 12, generator=g, dtype=torch.float64)
    y = torch.rand(1, 1, 9, 12, generator=g, dtype=torch.float64)
    w = L.gaussian(11, 1.5).tolist()
    H, W = 9, 12

    def blur(img):
        out = [[0.0] * W for _ in range(H)]
        for i in range(H):
            for j in range(W):
                s = 0.0
                for u in range(11):
                    for v in range(11):
                        ii, jj = i + u - 5, j + v - 5
                        if 0 <= ii < H and 0 <= jj < W:
                            s += w[u] * w[v] * img[ii][jj]
                out[i][j] = s
        return out
    X, Y = x[0, 0].tolist(), y[0, 0].tolist()
    m1, m2 = blur(X), blur(Y)
    s11 = blur([[v * v for v in r] for r in X])
    s22 = blur([[v * v for v in r] for r in Y])
    s12 = blur([[a * b for a, b in zip(r, q)] for r, q in zip(X, Y)])
    tot = 0.0
    for i in range(H):
        for j in range(W):
            mu1, mu2 = m1[i][j], m2[i][j]
            v1, v2, v12 = s11[i][j] - mu1 ** 2, s22[i][j] - mu2 ** 2, s12[i][j] - mu1 * mu2
            tot += ((2 * mu1 * mu2 + 1e-4) * (2 * v12 + 9e-4)) / \
                ((mu1 ** 2 + mu2 ** 2 + 1e-4) * (v1 + v2 + 9e-4))
    assert abs(tot / (H * W) - float(L.ssim(x, y))) < 1e-12
