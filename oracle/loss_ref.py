"""Restatement of pointrix's image losses (test infrastructure): l1_loss
(pointrix/model/loss.py:27-46), gaussian / create_window / _ssim
(loss.py:67-125) and the get_loss_dict combination (model/base_model.py:
112-124), in torch on the CPU at any dtype (float64 for the oracle)."""
import math

import torch
import torch.nn.functional as F


def gaussian(window_size, sigma, dtype=torch.float64):
    g = torch.tensor([math.exp(-(x - window_size // 2) ** 2 / float(2 * sigma ** 2))
                      for x in range(window_size)], dtype=dtype)
    return g / g.sum()


def create_window(window_size, channel, dtype=torch.float64):
    w1 = gaussian(window_size, 1.5, dtype).unsqueeze(1)
    w2 = w1.mm(w1.t()).unsqueeze(0).unsqueeze(0)
    return w2.expand(channel, 1, window_size, window_size).contiguous()


def ssim(img1, img2, window_size=11):
    """Mean SSIM of [B, C, H, W] (or [C, H, W]) images (loss.py:73-116)."""
    if img1.dim() == 3:
        img1, img2 = img1[None], img2[None]
    channel = img1.size(-3)
    window = create_window(window_size, channel, img1.dtype)
    pad = window_size // 2
    mu1 = F.conv2d(img1, window, padding=pad, groups=channel)
    mu2 = F.conv2d(img2, window, padding=pad, groups=channel)
    mu1_sq, mu2_sq, mu1_mu2 = mu1.pow(2), mu2.pow(2), mu1 * mu2
    sigma1_sq = F.conv2d(img1 * img1, window, padding=pad, groups=channel) - mu1_sq
    sigma2_sq = F.conv2d(img2 * img2, window, padding=pad, groups=channel) - mu2_sq
    sigma12 = F.conv2d(img1 * img2, window, padding=pad, groups=channel) - mu1_mu2
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    ssim_map = ((2 * mu1_mu2 + C1) * (2 * sigma12 + C2)) / \
        ((mu1_sq + mu2_sq + C1) * (sigma1_sq + sigma2_sq + C2))
    return ssim_map.mean()


def loss_dict(pred, gt, lambda_ssim=0.2):
    """base_model.py:112-124."""
    L1 = torch.abs(pred - gt).mean()
    ssim_loss = 1.0 - ssim(pred, gt)
    return {"loss": (1.0 - lambda_ssim) * L1 + lambda_ssim * ssim_loss, "L1_loss": L1,
            "ssim_loss": ssim_loss}
