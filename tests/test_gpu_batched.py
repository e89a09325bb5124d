"""GPU parity of the batched multi-view pipeline (MI355X).

* Images of every view equal the drop-in single-view ``render`` of the same
  camera (1e-4 abs; they differ only in where the activations are computed),
  and the FP64 oracle's.
* Raw-parameter gradients equal the sum over views of the single-view
  gradients chained through the model activations by autograd, and the
  oracle's chained the same way (1e-3 norm-relative per tensor).
* The instance count equals the sum of the single-view counts; a capacity
  below it is reported by ``overflowed`` and the truncated run stays in bounds.
* Every blend variant gives bit-identical images.
* A captured HIP graph of the whole step reproduces the eager step on every
  replay, with eager work and synchronisation between replays.
"""
import ctypes
import math

import numpy as np
import pytest
import torch

from dmgs_amd import batched as BT
from dmgs_amd import cameras as C
from dmgs_amd import scenes as S
from dmgs_amd.gaussian_renderer import PipelineParams, render
from oracle import raster as OR
from tests.helpers import grad_rel_err

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-4
GRAD_TOL = 1e-3
RAW = ("_xyz", "_scaling", "_rotation", "_opacity", "_features_dc", "_features_rest")


def _cameras(sizes, seed=2, radius=4.0, device="cuda"):
    pos = C.sphere_positions(len(sizes), radius, seed)
    return [C.camera_at(p, w, h, fov_deg=50.0, device=device) for p, (w, h) in zip(pos, sizes)]


def _bgs(n):
    g = torch.Generator().manual_seed(5)
    return [torch.rand(3, generator=g) for _ in range(n)]


def _dL(renderer, seed=3):
    g = torch.Generator().manual_seed(seed)
    return (torch.randn(3 * renderer.total_pixels, generator=g) * 1e-3).to(renderer.device)


def _plain(model_cpu, cams, bgs, dL_views):
    """Single-view drop-in render per camera; autograd sums the gradients."""
    m = model_cpu.to("cuda").requires_grad_(True)
    pipe = PipelineParams()
    imgs = []
    counts = []
    for cam, bg, dl in zip(cams, bgs, dL_views):
        out = render(cam, m, pipe, bg.cuda())
        (out["render"] * dl).sum().backward()
        imgs.append(out["render"].detach())
        counts.append(out.get("num_rendered"))
    return imgs, {k: getattr(m, k).grad for k in RAW}, counts


def _ambiguous(model_cpu, cam, bg, sh_degree=3):
    """Pixels the FP64 oracle flags as blend threshold ties for this view
    (alpha ~ 1/255, T ~ 1e-4, power ~ 0): FP32 paths that differ by an ulp
    may resolve them differently, so they stay out of the comparison."""
    m = model_cpu
    st = OR.settings_for(_cpu(cam), bg.double(), sh_degree)
    _, _, state = OR.rasterize(m.get_xyz, m.get_opacity, st, shs=m.get_features, scales=m.get_scaling,
                               rotations=m.get_rotation)
    return state["ambiguous"]


def _cpu(cam):
    class Cam:
        pass
    c = Cam()
    for k in ("image_width", "image_height", "FoVx", "FoVy"):
        setattr(c, k, getattr(cam, k))
    for k in ("world_view_transform", "full_proj_transform", "camera_center"):
        setattr(c, k, getattr(cam, k).detach().cpu())
    return c


def _batched(model_cpu, cams, bgs, variant=0, cell_mode=None):
    r = BT.BatchedRenderer(cams, [b.cuda() for b in bgs], model_cpu.num, sh_degree=model_cpu.max_sh_degree,
                           variant=variant, cell_mode=cell_mode)
    params = BT.pack_model(model_cpu.to("cuda"))
    r.plan(params)
    return r, params


@pytest.mark.parametrize("n,sizes", [(20000, [(128, 96), (200, 200), (160, 90)]),
                                     (5000, [(64, 48)]),
                                     (30000, [(256, 256), (96, 160), (33, 17), (200, 120)]),
                                     # more than 16 views: global cell lists and plain depth keys
                                     (8000, [(48, 40), (40, 48), (64, 32)] * 6 + [(24, 24), (56, 40)])])
def test_batched_matches_single_view(n, sizes):
    print(f"case: batched vs single-view n={n} sizes={sizes}", flush=True)
    model = S.free_scene(n, seed=0, layout="mixture", shell_radius=3.0)
    _check_against_single_view(model, _cameras(sizes))


def _check_against_single_view(model, cams):
    bgs = _bgs(len(cams))
    r, params = _batched(model, cams, bgs)
    dL = _dL(r)
    amb = [_ambiguous(model, c, b, sh_degree=model.max_sh_degree) for c, b in zip(cams, bgs)]
    for v in range(r.B):  # keep threshold ties out of the gradient comparison
        r.image(dL, v)[:, amb[v].to(r.device)] = 0.0
    images = r.forward(params)
    assert not r.overflowed()
    grads = r.backward(params, dL)
    torch.cuda.synchronize()
    imgs, pgrads, _ = _plain(model, cams, bgs, [r.image(dL, v) for v in range(r.B)])
    for v in range(r.B):
        ok = ~amb[v].to(r.device)
        err = (r.image(images, v) - imgs[v]).abs()[:, ok].max().item()
        print(f"view {v}: {int((~ok).sum())} tie pixels, max image error {err:.3g}", flush=True)
        assert err < IMG_TOL, f"view {v}: image error {err}"
    bg = BT.unpack_grads(grads, (model.max_sh_degree + 1) ** 2)
    for k in RAW:
        e = grad_rel_err(bg[k], pgrads[k])
        assert e < GRAD_TOL, f"{k}: rel err {e}"
    return r, images, bg


def _away_camera(w, h, device="cuda"):
    """A camera outside the scene's shell that looks away from it: nothing is visible."""
    p = np.array([0.0, 0.0, 8.0])
    R = C.look_at(p, target=tuple(2.0 * p))
    f = math.radians(50.0)
    return C.SimpleCamera(R, -R.T @ p, f, f, w, h, device=device)


@pytest.mark.parametrize("kinds", [("scene", "away", "scene", "scene", "scene"), ("away", "away")])
def test_batched_empty_and_tiny_views(kinds):
    """Views that see nothing and 1-pixel-high views in one batch; a batch in
    which no view sees anything renders the backgrounds and has zero gradients."""
    sizes = [(96, 64), (64, 48), (1, 1), (17, 1), (1, 23)]
    print(f"case: empty and tiny views {kinds}", flush=True)
    model = S.free_scene(4000, seed=6, layout="mixture", shell_radius=3.0)
    pos = C.sphere_positions(len(kinds), 4.0, 9)
    cams = [C.camera_at(p, w, h, fov_deg=50.0) if k == "scene" else _away_camera(w, h)
            for k, p, (w, h) in zip(kinds, pos, sizes)]
    r, images, grads = _check_against_single_view(model, cams)
    bgs = _bgs(len(cams))
    for v, k in enumerate(kinds):
        if k == "away":
            img = r.image(images, v)
            assert torch.equal(img, bgs[v].to(img.device)[:, None, None].expand_as(img))
    if "scene" not in kinds:
        for name in RAW:
            assert grads[name].abs().max().item() == 0.0, name


def test_clamped_colour_channel_has_zero_sh_gradient():
    """A colour channel that comes out of the SH negative is clamped to zero in
    the forward and passes no gradient to its SH coefficients (the reference's
    per-channel clamp flags).  With degree-0 SH a Gaussian's colour is the same
    in every view, so Gaussians whose green channel is negative get exactly
    zero green coefficient gradients, from the batched engine and from the
    single-view path, while their red and blue channels still get gradient."""
    print("case: clamped colour channel, 3000 Gaussians, 3 views, SH degree 0", flush=True)
    model = S.free_scene(3000, seed=8, layout="mixture", shell_radius=3.0, sh_degree=0)
    n_neg = 1500
    model._features_dc[:n_neg, 0, 1] = -3.0  # green = C0 * (-3) + 0.5 < 0
    cams = _cameras([(96, 64), (80, 80), (64, 96)], seed=12)
    r, images, bgrads = _check_against_single_view(model, cams)
    dc = bgrads["_features_dc"]
    assert torch.count_nonzero(dc[:n_neg, 0, 1]).item() == 0
    assert torch.count_nonzero(dc[:n_neg, 0, 0]).item() > n_neg // 10
    assert torch.count_nonzero(dc[n_neg:, 0, 1]).item() > 0
    g = torch.Generator().manual_seed(13)
    dL_views = [torch.randn(3, c.image_height, c.image_width, generator=g).cuda() for c in cams]
    _, pgrads, _ = _plain(model, cams, _bgs(len(cams)), dL_views)
    assert torch.count_nonzero(pgrads["_features_dc"][:n_neg, 0, 1]).item() == 0
    assert torch.count_nonzero(pgrads["_features_dc"][:n_neg, 0, 2]).item() > n_neg // 10


@pytest.mark.parametrize("deg", [1, 2])
def test_batched_lower_sh_degrees(deg):
    """Models of SH degree 1 and 2 (4 and 9 coefficients per channel, the rest
    of the 48 packed channels unused): images and every raw-parameter
    gradient match the single-view path."""
    print(f"case: batched vs single-view, SH degree {deg}, 4000 Gaussians, 2 views", flush=True)
    model = S.free_scene(4000, seed=10 + deg, layout="mixture", shell_radius=3.0, sh_degree=deg)
    _, _, g = _check_against_single_view(model, _cameras([(112, 80), (72, 104)], seed=15))
    assert g["_features_rest"].shape[1:] == ((deg + 1) ** 2 - 1, 3)
    assert torch.count_nonzero(g["_features_rest"]).item() > 0


def test_batched_active_degree_below_model_degree():
    """The batched engine at active SH degree 1 on a degree-3 model (all 48
    channels packed) equals the single-view path at the same active degree,
    and the coefficients above degree 1 get exactly zero gradient."""
    print("case: batched active SH degree 1 of 3, 4000 Gaussians, 2 views", flush=True)
    model = S.free_scene(4000, seed=16, layout="mixture", shell_radius=3.0)
    cams = _cameras([(112, 80), (72, 104)], seed=17)
    bgs = _bgs(len(cams))
    r = BT.BatchedRenderer(cams, [b.cuda() for b in bgs], model.num, sh_degree=1)
    params = BT.pack_model(model.to("cuda"))
    r.plan(params)
    dL = _dL(r)
    amb = [_ambiguous(model, c, b, sh_degree=1) for c, b in zip(cams, bgs)]
    for v in range(r.B):
        r.image(dL, v)[:, amb[v].to(r.device)] = 0.0
    images = r.forward(params)
    assert not r.overflowed()
    grads = BT.unpack_grads(r.backward(params, dL))
    m = model.to("cuda").requires_grad_(True)
    m.active_sh_degree = 1
    for v, (cam, bg) in enumerate(zip(cams, bgs)):
        out = render(cam, m, PipelineParams(), bg.cuda())
        (out["render"] * r.image(dL, v)).sum().backward()
        ok = ~amb[v].to(r.device)
        err = (r.image(images, v) - out["render"].detach()).abs()[:, ok].max().item()
        print(f"view {v}: max image error {err:.3g}", flush=True)
        assert err < IMG_TOL
    for k in RAW:
        e = grad_rel_err(grads[k], getattr(m, k).grad)
        assert e < GRAD_TOL, f"{k}: rel err {e}"
    assert torch.count_nonzero(grads["_features_rest"][:, 3:]).item() == 0
    assert torch.count_nonzero(grads["_features_rest"][:, :3]).item() > 0


def test_batched_scale_modifier():
    """The batched engine with scale_modifier 0.6 equals the single-view
    render(scaling_modifier=0.6): images off ties and every raw-parameter
    gradient, including the log-scales'."""
    print("case: batched scale modifier 0.6, 4000 Gaussians, 2 views", flush=True)
    model = S.free_scene(4000, seed=18, layout="mixture", shell_radius=3.0)
    cams = _cameras([(112, 80), (72, 104)], seed=19)
    bgs = _bgs(len(cams))
    r = BT.BatchedRenderer(cams, [b.cuda() for b in bgs], model.num, sh_degree=3, scale_modifier=0.6)
    params = BT.pack_model(model.to("cuda"))
    r.plan(params)
    dL = _dL(r)
    amb = []
    for c, b in zip(cams, bgs):
        st = OR.settings_for(_cpu(c), b.double(), 3, scale_modifier=0.6)
        _, _, ost = OR.rasterize(model.get_xyz, model.get_opacity, st, shs=model.get_features,
                                 scales=model.get_scaling, rotations=model.get_rotation)
        amb.append(ost["ambiguous"])
    for v in range(r.B):
        r.image(dL, v)[:, amb[v].to(r.device)] = 0.0
    images = r.forward(params)
    assert not r.overflowed()
    grads = BT.unpack_grads(r.backward(params, dL))
    m = model.to("cuda").requires_grad_(True)
    for v, (cam, bg) in enumerate(zip(cams, bgs)):
        out = render(cam, m, PipelineParams(), bg.cuda(), scaling_modifier=0.6)
        (out["render"] * r.image(dL, v)).sum().backward()
        ok = ~amb[v].to(r.device)
        err = (r.image(images, v) - out["render"].detach()).abs()[:, ok].max().item()
        print(f"view {v}: max image error {err:.3g}", flush=True)
        assert err < IMG_TOL
    for k in RAW:
        e = grad_rel_err(grads[k], getattr(m, k).grad)
        assert e < GRAD_TOL, f"{k}: rel err {e}"


def test_huge_splats_cover_every_cell():
    """Three faint Gaussians of standard deviation 2.5 at the centre cover
    every tile and cell of every view (hundreds of instances each, which the
    load-balanced duplicate spreads over several chunks), among 3000 small
    ones: the batched engine still matches the single-view path."""
    print("case: huge splats, 3000 Gaussians, 3 views", flush=True)
    model = S.free_scene(3000, seed=9, layout="mixture", shell_radius=3.0)
    model._xyz[:3] = torch.tensor([[0.0, 0.0, 0.0], [0.3, -0.2, 0.1], [-0.2, 0.1, 0.3]])
    model._scaling[:3] = math.log(2.5)
    model._opacity[:3] = -2.5  # sigmoid: 0.076, so the splats behind them still show
    cams = _cameras([(200, 150), (128, 96), (96, 160)], seed=14)
    r, images, _ = _check_against_single_view(model, cams)
    radii = r.radii().reshape(r.B, r.P)[:, :3]
    assert (radii > 100).all(), radii


def test_batched_matches_single_view_at_large_sort_size():
    """600 000 Gaussians x 8 views: 4.8M entries put the depth sort on its
    large-input block size (4096 keys).  No oracle at this size, so threshold
    ties are not masked: at most one pixel in 10^4 may differ by more than the
    tolerance."""
    print("case: batched vs single-view, 600000 Gaussians, 8 views of 160x120", flush=True)
    model = S.free_scene(600_000, seed=1, layout="mixture", shell_radius=3.0)
    cams = _cameras([(160, 120)] * 8, seed=11)
    bgs = _bgs(len(cams))
    r, params = _batched(model, cams, bgs)
    assert r.B * r.P >= 4 << 20
    images = r.forward(params)
    assert not r.overflowed()
    m = model.to("cuda")
    with torch.no_grad():
        for v, (cam, bg) in enumerate(zip(cams, bgs)):
            ref = render(cam, m, PipelineParams(), bg.cuda())["render"]
            err = (r.image(images, v) - ref).abs()
            bad = int((err > IMG_TOL).sum())
            print(f"view {v}: {bad} pixels above {IMG_TOL}, max {err.max().item():.3g}", flush=True)
            assert bad <= err.numel() // 10_000


def test_batched_against_oracle():
    print("case: batched vs FP64 oracle, 800 Gaussians, 2 views", flush=True)
    model = S.free_scene(800, seed=4, layout="uniform")
    sizes = [(64, 48), (40, 56)]
    cams = _cameras(sizes, seed=7, radius=3.5, device="cpu")
    bgs = _bgs(2)
    r, params = _batched(model, cams, bgs)
    dL = _dL(r, seed=9)
    # oracle on FP64 activated parameters; activations chained by autograd
    raw = {k: getattr(model, k).detach().double().requires_grad_(True) for k in RAW}
    act = dict(means3D=raw["_xyz"], opac=torch.sigmoid(raw["_opacity"]),
               shs=torch.cat((raw["_features_dc"], raw["_features_rest"]), 1), scales=torch.exp(raw["_scaling"]),
               rots=torch.nn.functional.normalize(raw["_rotation"]))
    runs = []
    for v, cam in enumerate(cams):
        st = OR.settings_for(cam, bgs[v].double(), 3)
        color, _, state = OR.rasterize(act["means3D"].detach(), act["opac"].detach(), st,
                                       shs=act["shs"].detach(), scales=act["scales"].detach(),
                                       rotations=act["rots"].detach())
        r.image(dL, v)[:, state["ambiguous"].to(r.device)] = 0.0  # threshold ties stay out
        runs.append((st, color, state))
    images = r.forward(params)
    grads = BT.unpack_grads(r.backward(params, dL))
    torch.cuda.synchronize()
    total = {k: torch.zeros_like(v) for k, v in act.items()}
    for v, (st, color, state) in enumerate(runs):
        ok = ~state["ambiguous"]
        err = (r.image(images, v).cpu().double() - color).abs()[:, ok].max().item()
        assert err < IMG_TOL, f"view {v}: image error vs oracle {err}"
        g = OR.rasterize_backward(state, st, act["means3D"].detach(), r.image(dL, v).cpu().double(),
                                  shs=act["shs"].detach(), scales=act["scales"].detach(),
                                  rotations=act["rots"].detach())
        dmeans3D, _, dsh, _, dop, dscales, drot = g[:7]
        total["means3D"] += dmeans3D
        total["shs"] += dsh
        total["opac"] += dop.reshape(total["opac"].shape)
        total["scales"] += dscales
        total["rots"] += drot
    torch.autograd.backward([act[k] for k in act], [total[k] for k in act])
    for k in RAW:
        e = grad_rel_err(grads[k], raw[k].grad)
        assert e < GRAD_TOL, f"{k}: rel err vs oracle {e}"


def test_instance_count_and_overflow():
    print("case: instance count and overflow detection", flush=True)
    model = S.free_scene(20000, seed=1, layout="mixture")
    sizes = [(128, 128), (96, 64)]
    cams = _cameras(sizes, seed=11)
    bgs = _bgs(2)
    r, params = _batched(model, cams, bgs)
    K = r.count_instances(params)
    # the (splat, cell) instances of the drop-in path's tile rectangles: a
    # rectangle [x0, x1) x [y0, y1) of 16-pixel tiles touches the cells
    # x0 // 4 .. (x1 - 1) // 4 (and the same in y)
    from dmgs_amd import rasterizer as RZ
    from tests.helpers import settings_from_cam, workspace_views
    m = model.to("cuda")
    Ks, Ktiles = [], []
    for (w, h), cam, bg in zip(sizes, cams, bgs):
        rs = settings_from_cam(cam, bg, 3)
        _, radii, st = RZ.forward_raw(m.get_xyz, m.get_opacity, m.get_features, None, m.get_scaling,
                                      m.get_rotation, None, rs)
        wv = workspace_views(st, w, h)
        vis = radii.cpu().numpy() > 0
        x0, y0, x1, y1 = OR.rect_f32(wv["xy"][vis], radii.cpu().numpy()[vis], w, h)
        Ks.append(int((((x1 - 1) // 4 - x0 // 4 + 1) * ((y1 - 1) // 4 - y0 // 4 + 1)).sum()))
        Ktiles.append(st["K"])
    assert K == sum(Ks)
    assert K < sum(Ktiles)
    images_full = r.forward(params).clone()
    assert not r.overflowed()
    # a capacity below the count: reported, truncated, stays in bounds
    r.plan(params, K_cap=K // 2)
    images_cut = r.forward(params)
    torch.cuda.synchronize()
    assert r.overflowed()
    assert torch.isfinite(images_cut).all()
    r.plan(params)
    images_again = r.forward(params)
    assert not r.overflowed()
    assert torch.equal(images_again, images_full)


def test_render_variants_identical():
    print("case: blend variants", flush=True)
    model = S.free_scene(20000, seed=2, layout="mixture")
    cams = _cameras([(200, 150), (77, 53)], seed=13)
    bgs = _bgs(2)
    r, params = _batched(model, cams, bgs)
    dL = _dL(r)
    ref_img = r.forward(params).clone()
    ref_grad = r.backward(params, dL).clone()
    for v in range(1, len(BT.RENDER_VARIANTS)):
        r.variant = v
        img = r.forward(params)
        grad = r.backward(params, dL)
        torch.cuda.synchronize()
        assert torch.equal(img, ref_img), f"variant {v} image differs"
        assert grad_rel_err(grad, ref_grad) < 1e-5, f"variant {v} gradient differs"
    times = r.autotune(params, dL, reps=1)
    assert set(times) == set(range(len(BT.RENDER_VARIANTS)))


def test_graph_replay_matches_eager():
    print("case: HIP graph replay of the batched step", flush=True)
    model = S.free_scene(20000, seed=3, layout="mixture")
    sizes = [(160, 120), (128, 128), (90, 60)]
    cams = _cameras(sizes, seed=17)
    bgs = _bgs(3)
    r, params = _batched(model, cams, bgs)
    targets = [S.target_image(h, w, seed=20 + i) for i, (w, h) in enumerate(sizes)]
    step = BT.BatchedStep(r, BT.BatchedLoss(r, targets), params)
    eager_loss = step.run().clone()
    eager_img = step.images.clone()
    eager_grad = step.grads.clone()
    # reference: the single-view path with the fused loss, per view
    from dmgs_amd.loss import fused_dssim_l1
    m = model.to("cuda").requires_grad_(True)
    for v, (cam, bg) in enumerate(zip(cams, bgs)):
        out = render(cam, m, PipelineParams(), bg.cuda())
        loss = fused_dssim_l1(out["render"], targets[v].cuda(), 0.2)
        assert abs(loss.item() - eager_loss[v, 0].item()) < 1e-5
        loss.backward()
    g = BT.unpack_grads(eager_grad)
    for k in RAW:
        assert grad_rel_err(g[k], getattr(m, k).grad) < GRAD_TOL, k
    step.capture()
    for rep in range(4):
        step.images.zero_()
        step.grads.zero_()
        _ = (torch.randn(1000, 1000, device="cuda") @ torch.randn(1000, 1000, device="cuda")).sum().item()
        step.replay()
        torch.cuda.synchronize()
        assert torch.equal(step.loss.losses, eager_loss), f"replay {rep}: losses differ"
        assert torch.equal(step.images, eager_img), f"replay {rep}: images differ"
        assert grad_rel_err(step.grads, eager_grad) < 1e-5, f"replay {rep}: gradients differ"
        step.run()  # eager work between replays
        torch.cuda.synchronize()
    assert not r.overflowed()


@pytest.mark.parametrize("cell_mode,sizes", [(0, [(200, 150), (77, 53), (256, 256)]), (0, [(33, 17)]),
                                              (1, [(200, 150), (77, 53), (256, 256)]), (1, [(33, 17)]),
                                              (1, [(1100, 1000), (64, 64)])])
def test_cell_lists_equal_reference_tile_lists(cell_mode, sizes):
    """Every tile's list, read from its binning cell's list filtered by the
    tile's tag bit, equals the reference (tile | depth)-sorted list of that
    tile (oracle binning of the same FP32 splat positions, radii and
    depths): bit-exact entry ids in the same order.  Cell mode 0 keys lists
    by global cell id; mode 1 by (view, cell within the view) with one sort
    pass for views of up to 256 cells and two beyond (the 1100x1000 view has
    323 cells)."""
    print(f"case: cell lists vs reference tile lists, cell_mode={cell_mode} sizes={sizes}", flush=True)
    model = S.free_scene(20000, seed=6, layout="mixture")
    cams = _cameras(sizes, seed=19)
    r, params = _batched(model, cams, _bgs(len(cams)), cell_mode=cell_mode)
    assert r.cell_mode == cell_mode
    r.forward(params)
    torch.cuda.synchronize()
    assert not r.overflowed()
    ranges, keys, vals = r.cell_lists()
    ranges, keys, vals = ranges.numpy(), keys.numpy().view(np.uint32), vals.numpy().view(np.uint32)
    if cell_mode == 0:
        assert np.all(np.diff((keys & 0xFFFF).astype(np.int64)) >= 0), "cell ids must be sorted"
    else:
        assert np.all(np.diff((keys & 0xFFF).astype(np.int64)) >= 0), "cells within a view must be sorted"
        view = (keys >> 12) & 0xF
        assert np.array_equal(view, vals // r.P), "key view and entry view differ"
        base = np.array([r.views[v].cell_base for v in range(r.B)], dtype=np.int64)
        glob = base[view] + (keys & 0xFFF)
        # each (view, cell) run is contiguous and is exactly its range
        for c in np.unique(glob):
            pos = np.nonzero(glob == c)[0]
            assert pos[-1] - pos[0] + 1 == pos.size and tuple(ranges[c]) == (pos[0], pos[-1] + 1)
    o = r._offsets[5]
    splats = r.state[o:o + 48 * r.B * r.P].view(torch.float32).view(r.B * r.P, 12).cpu().numpy()
    ids = vals
    tags = keys >> 16
    keys = keys & 0xFFFF
    all_radii = r.radii().cpu().numpy()  # splat records of culled entries are not written
    checked = 0
    for v in range(r.B):
        d = r.views[v]
        H, W = r.sizes[v]
        sl = slice(v * r.P, (v + 1) * r.P)
        xy = splats[sl, 0:2].copy()
        radii = all_radii[v]
        assert np.array_equal(radii[radii > 0], splats[sl, 10].copy().view(np.int32)[radii > 0])
        depth = splats[sl, 9].copy()
        _, ovals, oranges = OR.bin_gaussians(xy, radii, depth, W, H)
        gx, gy = d.grid_x, d.grid_y
        cgx = (gx + 3) // 4  # cells of 4x4 tiles
        for t in range(gx * gy):
            tx, ty = t % gx, t // gx
            cell = d.cell_base + (ty // 4) * cgx + tx // 4
            bit = 1 << ((ty & 3) * 4 + (tx & 3))
            s0, s1 = ranges[cell]
            mine = ids[s0:s1][(tags[s0:s1] & bit) != 0].astype(np.int64) - v * r.P
            ref = ovals[oranges[t, 0]:oranges[t, 1]]
            assert np.array_equal(mine, ref), f"view {v} tile {t}: {mine[:8]} vs {ref[:8]}"
            checked += ref.size
    assert checked > 0


def test_view_major_far_depth_falls_back():
    """A visible splat at depth >= 2^27 cannot be ordered by view-major keys:
    the forward reports it, the renderer switches to global cell lists, and
    the re-rendered images equal those of a cell-mode-0 renderer."""
    print("case: view-major far depth fallback", flush=True)
    model = S.free_scene(4000, seed=8, layout="uniform")
    cams = _cameras([(96, 64), (64, 64)], seed=3)
    # one huge splat far along the first camera's optical axis
    c = cams[0]
    Rt = c.world_view_transform.detach().cpu().double()
    center, fwd = c.camera_center.detach().cpu().double(), Rt[:3, 2]
    with torch.no_grad():
        model._xyz[0] = (center + 6e8 * fwd).float()  # 2^27 = 1.3e8
        model._scaling[0] = math.log(4e7)
        model._opacity[0] = 5.0
    bgs = _bgs(len(cams))
    r, params = _batched(model, cams, bgs, cell_mode=1)
    img = r.forward(params)
    torch.cuda.synchronize()
    assert int(r.radii()[0, 0]) > 0, "the far splat must be visible"
    assert r.depth_clamped()
    assert r.overflowed() and r.cell_mode == 0
    img = r.forward(params)
    assert not r.overflowed()
    r0, _ = _batched(model, cams, bgs, cell_mode=0)
    assert torch.equal(img, r0.forward(params))


def test_densification_stats_match_per_view_consumer():
    """The statistics a batched backward accumulates equal the reference
    consumer (add_densification_stats + the max_radii2D update) run after
    each single-view render + backward, views in order: the norm sums within
    1e-4 of their largest value, view counts and radii exactly."""
    print("case: densification statistics", flush=True)
    model = S.free_scene(20000, seed=12, layout="mixture", shell_radius=3.0)
    cams = _cameras([(160, 120), (128, 128), (200, 100)], seed=4)
    bgs = _bgs(len(cams))
    r, params = _batched(model, cams, bgs)
    r.track_densification()
    dL = _dL(r, seed=5)
    for v in range(r.B):  # keep threshold ties out of the comparison
        r.image(dL, v)[:, _ambiguous(model, cams[v], bgs[v]).to(r.device)] = 0.0
    r.forward(params)
    r.backward(params, dL)
    r.backward(params, dL)  # a second step keeps accumulating
    torch.cuda.synchronize()
    acc, den, mr = r.densification_stats()
    m = model.to("cuda")
    pipe = PipelineParams()
    for _ in range(2):
        for v, (cam, bg) in enumerate(zip(cams, bgs)):
            mv = model.to("cuda").requires_grad_(True)
            out = render(cam, mv, pipe, bg.cuda())
            (out["render"] * r.image(dL, v)).sum().backward()
            vis = out["visibility_filter"]
            m.update_max_radii2D(out["radii"], vis)
            m.add_densification_stats(out["viewspace_points"], vis)
    assert torch.equal(den, m.denom)
    assert torch.equal(mr, m.max_radii2D)
    err = (acc - m.xyz_gradient_accum).abs().max().item()
    scale = m.xyz_gradient_accum.abs().max().item()
    print(f"max |accum diff| {err:.3g} of {scale:.3g}", flush=True)
    assert scale > 0 and err <= 1e-4 * scale


def test_training_steps_replan_on_overflow():
    """Five captured training steps with the parameters updated in place
    between replays (every log-scale grows by 0.25 per step, so the instance
    count outgrows the planned capacity): each step's check re-plans and
    redoes the step, and every step's images, losses and gradients equal an
    eager step of a renderer planned for exactly those parameters."""
    print("case: training steps with capacity re-planning", flush=True)
    model = S.free_scene(20000, seed=21, layout="mixture", shell_radius=3.0)
    sizes = [(160, 120), (128, 128), (96, 64)]
    cams = _cameras(sizes, seed=23)
    bgs = _bgs(3)
    r, params = _batched(model, cams, bgs)
    targets = [S.target_image(h, w, seed=40 + i) for i, (w, h) in enumerate(sizes)]
    step = BT.BatchedStep(r, BT.BatchedLoss(r, targets), params)
    step.capture()
    caps = [r.K_cap]
    for it in range(5):
        with torch.no_grad():
            params[BT.SCALE:BT.SCALE + 3] += 0.25
        losses = step.replay().clone()
        torch.cuda.synchronize()
        assert not step.overflowed() and not r.overflowed()
        caps.append(r.K_cap)
        fresh = BT.BatchedRenderer(cams, [b.cuda() for b in bgs], model.num, sh_degree=3)
        fresh.plan(params)
        ref = BT.BatchedStep(fresh, BT.BatchedLoss(fresh, targets), params)
        ref_losses = ref.run().clone()
        torch.cuda.synchronize()
        assert torch.equal(step.images, ref.images), f"step {it}: images differ"
        assert torch.equal(losses, ref_losses), f"step {it}: losses differ"
        assert grad_rel_err(step.grads, ref.grads) < 1e-5, f"step {it}: gradients differ"
        print(f"  step {it}: instances {int(r.instance_count().item())}, capacity {r.K_cap}, re-plans "
              f"{step.replans}", flush=True)
    assert step.replans >= 2 and caps[-1] > caps[0]


def test_65536_cells_keep_one_run_per_cell():
    """65535 views of 16x16 pixels and one of 64x64: 65536 binning cells,
    global cell lists.  The sort must not reach the first tile-mask bit above
    the 16-bit cell id, or the big view's cell would split into two runs:
    the big view and two small ones equal the drop-in single-view render."""
    print("case: 65536 cells (65535 views of 16x16 + one of 64x64)", flush=True)
    model = S.free_scene(48, seed=31, layout="uniform")
    pos = np.array([0.3, -0.2, 4.0])
    small = C.camera_at(pos, 16, 16, fov_deg=50.0, device="cpu")
    big = C.camera_at(pos, 64, 64, fov_deg=50.0, device="cpu")
    cams = [small] * 65535 + [big]
    bg = torch.tensor([0.2, 0.3, 0.4])
    r = BT.BatchedRenderer(cams, bg.cuda(), model.num, sh_degree=3)
    assert r.total_cells == 65536 and r.cell_mode == 0
    params = BT.pack_model(model.to("cuda"))
    r.plan(params, backward=False)
    assert r.lib.dmgs_batched_cell_sort_bits(ctypes.byref(r._desc())) == 16
    images = r.forward(params)
    torch.cuda.synchronize()
    assert not r.overflowed()
    m = model.to("cuda")
    with torch.no_grad():
        for v, cam in ((65535, big), (0, small), (40000, small)):
            ref = render(cam, m, PipelineParams(), bg.cuda())["render"]
            amb = _ambiguous(model, cam, bg).to("cuda")
            err = (r.image(images, v) - ref).abs()[:, ~amb].max().item()
            print(f"  view {v}: max image error {err:.3g}", flush=True)
            assert err < IMG_TOL
