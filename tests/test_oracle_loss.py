"""Pins of the FP64 loss oracle (utils/loss_utils.py:17-62): closed forms
that fix the window, C1 and C2 independently of the oracle's SSIM formula."""
import math

import numpy as np
import torch

from oracle import loss as OL

F64 = torch.float64


def test_identical_images():
    x = torch.rand(3, 30, 40, generator=torch.Generator().manual_seed(0), dtype=F64)
    assert abs(OL.ssim_map(x, x).mean().item() - 1.0) < 1e-12
    assert abs(OL.dssim_l1(x, x).item()) < 1e-12


def test_window_is_the_normalised_sigma_1_5_gaussian():
    w = OL.window()
    assert w.shape == (11, 11)
    assert abs(w.sum().item() - 1.0) < 1e-14 and torch.allclose(w, w.T) and torch.allclose(w, w.flip(0))
    g = w.sum(1)                                  # the separable 1D factor
    assert torch.allclose(w, torch.outer(g, g), atol=1e-16)
    # neighbouring taps differ by exp((2k + 1) / (2 sigma^2)) with sigma = 1.5
    for k in range(5):
        assert abs((g[5 + k] / g[6 + k]).item() - math.exp((2 * k + 1) / 4.5)) < 1e-12


def test_constant_images_fix_c1():
    """Constant images a, b: away from the border the variances vanish and
    SSIM = (2ab + C1) / (a^2 + b^2 + C1) with C1 = 0.01^2."""
    a, b = 0.03, 0.02
    sm = OL.ssim_map(torch.full((1, 30, 30), a, dtype=F64), torch.full((1, 30, 30), b, dtype=F64))
    C1 = 0.01 ** 2
    assert abs(sm[0, 15, 15].item() - (2 * a * b + C1) / (a * a + b * b + C1)) < 1e-12
    # a plausible slip (C1 = 0.01) would give a visibly different value
    assert abs((2 * a * b + 0.01) / (a * a + b * b + 0.01) - sm[0, 15, 15].item()) > 1e-3


def test_opposite_ramps_fix_c2():
    """x a linear ramp s (j - j0) along the columns, y = -x.  At column j0,
    away from the border: mu_x = mu_y = 0, sigma_x^2 = sigma_y^2 = s^2 m2 and
    sigma_xy = -s^2 m2, where m2 = sum_k g_k k^2 is the second moment of the
    11-tap window's 1D factor.  The luminance term is C1 / C1 = 1 and
    SSIM = (C2 - 2 s^2 m2) / (C2 + 2 s^2 m2) with C2 = 0.03^2."""
    g = OL.window().sum(1).numpy()
    k = np.arange(-5, 6, dtype=np.float64)
    m2 = float((g * k * k).sum())
    assert 2.0 < m2 < 2.25                        # a sigma 1.5 Gaussian, truncated at +-5
    C2 = 0.03 ** 2
    s = math.sqrt(C2 / (6 * m2))                  # 2 s^2 m2 = C2 / 3: SSIM = 0.5
    H, W, j0 = 24, 40, 20
    ramp = torch.tensor([[s * (j - j0) for j in range(W)]] * H, dtype=F64)[None]
    sm = OL.ssim_map(ramp, -ramp)
    exp = (C2 - 2 * s * s * m2) / (C2 + 2 * s * s * m2)
    assert abs(exp - 0.5) < 1e-12
    assert abs(sm[0, 12, j0].item() - exp) < 1e-10
    # off the zero column the luminance term (C1 - 2 mu^2) / (C1 + 2 mu^2) enters
    mu = s * 3
    C1 = 0.01 ** 2
    assert abs(sm[0, 12, j0 + 3].item() - exp * (C1 - 2 * mu * mu) / (C1 + 2 * mu * mu)) < 1e-10
    # C2 written as 0.03 instead of 0.03^2 would give 0.98
    wrong = (0.03 - 2 * s * s * m2) / (0.03 + 2 * s * s * m2)
    assert abs(wrong - sm[0, 12, j0].item()) > 0.4


def test_loss_weights_and_l1():
    """dssim_l1 = 0.8 L1 + 0.2 (1 - SSIM) (train_geo_stage3.py:122-124): with
    y = x + c (a constant) L1 = |c| and, away from the border, the SSIM map is
    the luminance term only."""
    g = torch.Generator().manual_seed(4)
    x = torch.rand(1, 40, 40, generator=g, dtype=F64)
    y = x + 0.1
    l1 = (x - y).abs().mean().item()
    assert abs(l1 - 0.1) < 1e-12
    val = OL.dssim_l1(x, y, 0.2).item()
    assert abs(val - (0.8 * l1 + 0.2 * (1 - OL.ssim_map(x, y).mean().item()))) < 1e-12
    assert abs(OL.dssim_l1(x, y, 0.0).item() - 0.1) < 1e-12
