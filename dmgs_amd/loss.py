"""Fused training objective of the rendering step on gfx950.

``fused_dssim_l1`` computes (1 - lambda) L1 + lambda (1 - SSIM) — the
objective of train_geo_stage3.py:122-124 with SSIM as utils/loss_utils.py:17-62
defines it (11x11 Gaussian window, sigma 1.5, zero padding, C1 = 0.01^2,
C2 = 0.03^2, mean over C*H*W) — as one forward and one backward HIP kernel
instead of five depthwise convolutions plus autograd.
"""
import torch


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
    """(1 - lambda) L1 + lambda (1 - SSIM) of two (C, H, W) images as two
    gfx950 kernels (forward and backward); differentiable in ``image``."""
    return _FusedDssimL1.apply(image, gt, lambda_dssim)
