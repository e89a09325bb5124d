"""L1 + SSIM objective in FP64 (test infrastructure only).

Restates utils/loss_utils.py:17-62 (l1_loss, gaussian window sigma 1.5,
11x11 depthwise conv2d with zero padding, C1 = 0.01^2, C2 = 0.03^2) and the
weighting of train_geo_stage3.py:122-124.  Gradients come from torch autograd
of this FP64 forward.

Pins (tests/test_oracle_loss.py): SSIM(x, x) = 1 for any image; the window
sums to 1 and is symmetric; on a constant image pair the SSIM map has the
closed form (2 a b + C1) / (a^2 + b^2 + C1) away from the borders.
"""
import math

import torch
import torch.nn.functional as F


def window(size=11, sigma=1.5):
    g = torch.tensor([math.exp(-(x - size // 2) ** 2 / float(2 * sigma ** 2)) for x in range(size)],
                     dtype=torch.float64)
    g = g / g.sum()
    return torch.outer(g, g)


def ssim_map(x, y, size=11):
    C = x.shape[-3]
    w = window(size).expand(C, 1, size, size).contiguous()
    pad = size // 2
    x4, y4 = x.unsqueeze(0), y.unsqueeze(0)
    mu1 = F.conv2d(x4, w, padding=pad, groups=C)
    mu2 = F.conv2d(y4, w, padding=pad, groups=C)
    s11 = F.conv2d(x4 * x4, w, padding=pad, groups=C) - mu1 * mu1
    s22 = F.conv2d(y4 * y4, w, padding=pad, groups=C) - mu2 * mu2
    s12 = F.conv2d(x4 * y4, w, padding=pad, groups=C) - mu1 * mu2
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    return (((2 * mu1 * mu2 + C1) * (2 * s12 + C2)) / ((mu1 * mu1 + mu2 * mu2 + C1) * (s11 + s22 + C2)))[0]


def dssim_l1(x, y, lam=0.2):
    x = x.to(torch.float64)
    y = y.to(torch.float64)
    return (1 - lam) * (x - y).abs().mean() + lam * (1 - ssim_map(x, y).mean())


def dssim_l1_with_grad(x, y, lam=0.2):
    xx = x.detach().cpu().to(torch.float64).requires_grad_(True)
    val = dssim_l1(xx, y.detach().cpu())
    (g,) = torch.autograd.grad(val, xx)
    return val.item(), g
