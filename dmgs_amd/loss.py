"""Image losses of the training step (utils/loss_utils.py:17-62).

``l1_loss`` and ``ssim`` (11x11 Gaussian window, sigma 1.5, C1 = 0.01^2,
C2 = 0.03^2, depthwise conv2d, zero padding) with the reference's semantics;
``dssim_l1`` is the stage-3 / plain-3DGS objective
(1 - lambda) L1 + lambda (1 - SSIM) (train_geo_stage3.py:122-124).
"""
import math

import torch
import torch.nn.functional as F

_WINDOWS = {}


def l1_loss(network_output, gt):
    return torch.abs(network_output - gt).mean()


def l2_loss(network_output, gt):
    return ((network_output - gt) ** 2).mean()


def _window(window_size, channel, device, dtype):
    key = (window_size, channel, str(device), dtype)
    w = _WINDOWS.get(key)
    if w is None:
        g = torch.tensor([math.exp(-(x - window_size // 2) ** 2 / float(2 * 1.5 ** 2)) for x in range(window_size)])
        g = (g / g.sum()).unsqueeze(1)
        w2 = g.mm(g.t()).float().unsqueeze(0).unsqueeze(0)
        w = w2.expand(channel, 1, window_size, window_size).contiguous().to(device=device, dtype=dtype)
        _WINDOWS[key] = w
    return w


def ssim(img1, img2, window_size=11, size_average=True):
    channel = img1.size(-3)
    window = _window(window_size, channel, img1.device, img1.dtype)
    pad = window_size // 2
    mu1 = F.conv2d(img1, window, padding=pad, groups=channel)
    mu2 = F.conv2d(img2, window, padding=pad, groups=channel)
    mu1_sq, mu2_sq, mu1_mu2 = mu1.pow(2), mu2.pow(2), mu1 * mu2
    sigma1_sq = F.conv2d(img1 * img1, window, padding=pad, groups=channel) - mu1_sq
    sigma2_sq = F.conv2d(img2 * img2, window, padding=pad, groups=channel) - mu2_sq
    sigma12 = F.conv2d(img1 * img2, window, padding=pad, groups=channel) - mu1_mu2
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    ssim_map = ((2 * mu1_mu2 + C1) * (2 * sigma12 + C2)) / ((mu1_sq + mu2_sq + C1) * (sigma1_sq + sigma2_sq + C2))
    if size_average:
        return ssim_map.mean()
    return ssim_map.mean(1).mean(1).mean(1)


def dssim_l1(image, gt, lambda_dssim=0.2):
    """Reference formulation (torch ops, depthwise conv2d)."""
    return (1.0 - lambda_dssim) * l1_loss(image, gt) + lambda_dssim * (1.0 - ssim(image, gt))


class _FusedDssimL1(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, gt, lambda_dssim):
        from . import _native as N
        img = image.detach().to(torch.float32).contiguous()
        ref = gt.detach().to(device=img.device, dtype=torch.float32).contiguous()
        if img.dim() == 4 and img.shape[0] == 1:
            img, ref = img[0], ref.reshape(img.shape[1:])
        if img.dim() != 3 or img.shape != ref.shape:
            raise ValueError("fused_dssim_l1 expects two (C, H, W) images of the same shape")
        C, H, W = img.shape
        lib = N.load()
        ws = torch.empty(lib.dmgs_dssim_l1_workspace_size(C, H, W), dtype=torch.uint8, device=img.device)
        out = torch.empty(3, dtype=torch.float32, device=img.device)
        N.check(lib.dmgs_dssim_l1_forward(N.ptr(img), N.ptr(ref), C, H, W, float(lambda_dssim), N.ptr(ws),
                                          N.ptr(out), N.stream_handle(img.device)), "dmgs_dssim_l1_forward")
        ctx.save_for_backward(img, ref, ws)
        ctx.shape = image.shape
        ctx.lam = float(lambda_dssim)
        return out[0]

    @staticmethod
    def backward(ctx, grad_loss):
        from . import _native as N
        img, ref, ws = ctx.saved_tensors
        C, H, W = img.shape
        g = grad_loss.detach().to(torch.float32).reshape(1).contiguous()
        dimg = torch.empty_like(img)
        N.check(N.load().dmgs_dssim_l1_backward(N.ptr(img), N.ptr(ref), C, H, W, ctx.lam, N.ptr(ws), N.ptr(g),
                                                N.ptr(dimg), N.stream_handle(img.device)),
                "dmgs_dssim_l1_backward")
        return dimg.reshape(ctx.shape), None, None


def fused_dssim_l1(image, gt, lambda_dssim=0.2):
    """(1 - lambda) L1 + lambda (1 - SSIM) as two gfx950 kernels (forward and
    backward), same value and gradient as ``dssim_l1`` within FP32 rounding."""
    return _FusedDssimL1.apply(image, gt, lambda_dssim)
