"""GPU parity of the fused L1 + SSIM loss against the FP64 oracle
(utils/loss_utils.py:17-62, train_geo_stage3.py:122-124)."""
import pytest
import torch

from dmgs_amd import loss as L
from oracle import loss as OL
from tests.helpers import grad_rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("C,H,W", [(3, 64, 48), (3, 200, 200), (1, 37, 53), (3, 800, 800)])
def test_fused_dssim_l1_matches_oracle(C, H, W):
    print(f"case: fused dssim C={C} H={H} W={W}", flush=True)
    g = torch.Generator().manual_seed(H * 1000 + W)
    x = torch.rand(C, H, W, generator=g)
    y = torch.rand(C, H, W, generator=g)
    x[:, : H // 3] = 0.5 * x[:, : H // 3] + 0.25   # some structure: smoother region
    ref_val, ref_grad = OL.dssim_l1_with_grad(x, y, 0.2)
    xg = x.cuda().requires_grad_(True)
    val = L.fused_dssim_l1(xg, y.cuda(), 0.2)
    val.backward()
    assert abs(val.item() - ref_val) < 1e-5, (val.item(), ref_val)
    e = grad_rel_err(xg.grad, ref_grad)
    print(f"  value diff {abs(val.item() - ref_val):.2e}, grad rel err {e:.2e}")
    assert e < 1e-3


def test_fused_dssim_l1_identical_images():
    print("case: fused dssim identical", flush=True)
    x = torch.rand(3, 40, 40, generator=torch.Generator().manual_seed(3)).cuda()
    v = L.fused_dssim_l1(x, x.clone(), 0.2)
    assert abs(v.item()) < 1e-5   # L1 = 0 and SSIM = 1


def test_grouped_loss_equals_per_image_calls():
    """G images in one grouped launch pair give each image's own loss and
    gradient (scaled by its dloss) as separate single-image calls."""
    import ctypes
    from dmgs_amd import _native as N
    print("case: grouped loss, G=4 images 3x70x90", flush=True)
    lib = N.load()
    g = torch.Generator().manual_seed(12)
    G, H, W = 4, 70, 90
    img = torch.rand(G * 3, H, W, generator=g).cuda()
    gt = torch.rand(G * 3, H, W, generator=g).cuda()
    dloss = torch.tensor([1.0, 0.5, 2.0, -1.0]).cuda()
    ws = torch.empty(lib.dmgs_dssim_l1_workspace_size(G * 3, H, W), dtype=torch.uint8, device="cuda")
    out = torch.empty(G, 3, device="cuda")
    dimg = torch.empty_like(img)
    s = N.stream_handle()
    N.check(lib.dmgs_dssim_l1_forward_groups(N.ptr(img), N.ptr(gt), G, 3, H, W, 0.2, N.ptr(ws), N.ptr(out), s), "f")
    N.check(lib.dmgs_dssim_l1_backward_groups(N.ptr(img), N.ptr(gt), G, 3, H, W, 0.2, N.ptr(ws), N.ptr(dloss),
                                              N.ptr(dimg), s), "b")
    for k in range(G):
        x = img[3 * k:3 * k + 3].contiguous()
        y = gt[3 * k:3 * k + 3].contiguous()
        ws1 = torch.empty(lib.dmgs_dssim_l1_workspace_size(3, H, W), dtype=torch.uint8, device="cuda")
        o1 = torch.empty(3, device="cuda")
        d1 = torch.empty_like(x)
        N.check(lib.dmgs_dssim_l1_forward(N.ptr(x), N.ptr(y), 3, H, W, 0.2, N.ptr(ws1), N.ptr(o1), s), "f1")
        N.check(lib.dmgs_dssim_l1_backward(N.ptr(x), N.ptr(y), 3, H, W, 0.2, N.ptr(ws1), N.ptr(dloss[k:k + 1]),
                                           N.ptr(d1), s), "b1")
        torch.cuda.synchronize()
        assert torch.allclose(out[k], o1, rtol=1e-6, atol=1e-7), (k, out[k], o1)
        assert torch.allclose(dimg[3 * k:3 * k + 3], d1, rtol=1e-6, atol=1e-12), k


@pytest.mark.parametrize("G,H,W", [(4, 70, 90), (2, 33, 200), (1, 800, 800)])
def test_fixed_target_statistics_give_identical_loss(G, H, W):
    """The forward on precomputed target maps (G*y, G*y^2) gives bit-identical
    losses and backward coefficients to the forward that blurs all five maps,
    and the maps themselves are the 11x11 Gaussian blur of y and y^2."""
    from dmgs_amd import _native as N
    print(f"case: fixed-target statistics, G={G} 3x{H}x{W}", flush=True)
    lib = N.load()
    g = torch.Generator().manual_seed(21 + H)
    img = torch.rand(G * 3, H, W, generator=g).cuda()
    gt = torch.rand(G * 3, H, W, generator=g).cuda()
    dloss = torch.linspace(0.5, 1.5, G).cuda()
    s = N.stream_handle()
    stats = torch.empty(2, G * 3, H, W, device="cuda")
    N.check(lib.dmgs_dssim_target_stats(N.ptr(gt), G * 3, H, W, N.ptr(stats), s), "stats")
    outs, grads = [], []
    for use in (False, True):
        ws = torch.empty(lib.dmgs_dssim_l1_workspace_size(G * 3, H, W), dtype=torch.uint8, device="cuda")
        out = torch.empty(G, 3, device="cuda")
        dimg = torch.empty_like(img)
        if use:
            N.check(lib.dmgs_dssim_l1_forward_groups_stats(N.ptr(img), N.ptr(gt), N.ptr(stats), G, 3, H, W, 0.2,
                                                           N.ptr(ws), N.ptr(out), s), "f stats")
        else:
            N.check(lib.dmgs_dssim_l1_forward_groups(N.ptr(img), N.ptr(gt), G, 3, H, W, 0.2, N.ptr(ws), N.ptr(out),
                                                     s), "f")
        N.check(lib.dmgs_dssim_l1_backward_groups(N.ptr(img), N.ptr(gt), G, 3, H, W, 0.2, N.ptr(ws), N.ptr(dloss),
                                                  N.ptr(dimg), s), "b")
        outs.append(out)
        grads.append(dimg)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1])
    assert torch.equal(grads[0], grads[1])
    # the maps: depthwise conv2d with the normalised 11-tap Gaussian (sigma 1.5), zero padding 5
    x = torch.arange(11, dtype=torch.float64) - 5
    w1 = torch.exp(-x * x / (2 * 1.5 ** 2))
    w1 = w1 / w1.sum()
    win = (w1[:, None] * w1[None, :]).expand(G * 3, 1, 11, 11).contiguous()
    y64 = gt.double().cpu()[None]
    ref_mu = torch.nn.functional.conv2d(y64, win, padding=5, groups=G * 3)[0]
    ref_sq = torch.nn.functional.conv2d(y64 * y64, win, padding=5, groups=G * 3)[0]
    assert torch.allclose(stats[0].double().cpu(), ref_mu, atol=2e-6)
    assert torch.allclose(stats[1].double().cpu(), ref_sq, atol=2e-6)
