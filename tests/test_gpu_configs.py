"""GPU parity at the benchmark workloads' full sizes (MI355X).

* 10k Gaussians, one 200x200 view, L1 against a seeded target: every tile,
  image and every raw-parameter gradient against the FP64 oracle.
* 1M Gaussians, one 800x800 view (drop-in ``render``) and the same scene in
  an 8-view batch of 800x800 views: lists bit-exact on every tile, preprocess
  on every Gaussian, image and gradients on sampled tiles, and the
  0.8 L1 + 0.2 (1 - SSIM) loss value and gradient.
* 3M Gaussians, 16 orbit views of uneven size (512^2 .. 1920x1080) in one
  batch: instance counts of every view, lists of sampled tiles bit-exact,
  image and summed gradients on sampled tiles.
* The serving mix: four resident scenes of 0.5M / 1M / 2M / 3M Gaussians,
  requests from 256^2 to 1920x1080 served concurrently and alone: every image
  bit-identical both ways, and equal to the oracle on sampled tiles.

Sampled-tile method and tolerances: tests/sampled.py.
"""
import numpy as np
import pytest
import torch

from dmgs_amd import batched as BT
from dmgs_amd import cameras as C
from dmgs_amd import rasterizer as RZ
from dmgs_amd import scenes as S
from dmgs_amd import serving as SV
from dmgs_amd.gaussian_renderer import PipelineParams, render
from dmgs_amd.loss import fused_dssim_l1
from oracle import loss as OL
from oracle import raster as OR
from tests import sampled as SP
from tests.helpers import oracle_settings, settings_from_cam

pytestmark = pytest.mark.gpu

UNEVEN_SIZES = [(512, 512), (800, 800), (1280, 720), (1600, 900), (1920, 1080)]


def _cpu_cam(cam):
    class Cam:
        pass
    c = Cam()
    for k in ("image_width", "image_height", "FoVx", "FoVy"):
        setattr(c, k, getattr(cam, k))
    for k in ("world_view_transform", "full_proj_transform", "camera_center"):
        setattr(c, k, getattr(cam, k).detach().cpu())
    return c


def _activated_device(m):
    """The drop-in render's rasterizer inputs for a model on the device."""
    with torch.no_grad():
        return dict(means3D=m.get_xyz.contiguous(), opac=m.get_opacity.contiguous(), shs=m.get_features.contiguous(),
                    scales=m.get_scaling.contiguous(), rots=m.get_rotation.contiguous())


def _oracle_chain(act, st, dm2, dcon, dop, dcol, radii_gpu, subset=None):
    """Oracle raster-input gradients -> raw parameters through the FP64
    activations (autograd).  ``subset``: Gaussian ids the geometry backward
    runs on (others have zero blend gradients)."""
    if subset is None:
        pre = OR.preprocess(act["means3D"].detach(), act["opac"].detach(), st, shs=act["shs"].detach(),
                            scales=act["scales"].detach(), rotations=act["rots"].detach())
        g = SP.oracle_geometry(pre, st, act["means3D"].detach(), dm2, dcon, dcol, radii_gpu,
                               shs=act["shs"].detach(), scales=act["scales"].detach(),
                               rotations=act["rots"].detach())
        grads = dict(means3D=g["dmeans3D"], shs=g["dsh"], opac=dop.reshape(act["opac"].shape), scales=g["dscales"],
                     rots=g["drotations"])
        return pre, grads
    idx = torch.as_tensor(subset, dtype=torch.int64)
    sub = {k: v.detach()[idx] for k, v in act.items()}
    pre = OR.preprocess(sub["means3D"], sub["opac"], st, shs=sub["shs"], scales=sub["scales"], rotations=sub["rots"])
    g = SP.oracle_geometry(pre, st, sub["means3D"], dm2[idx], dcon[idx], dcol[idx], np.asarray(radii_gpu)[subset],
                           shs=sub["shs"], scales=sub["scales"], rotations=sub["rots"])
    full = {k: torch.zeros_like(v, dtype=torch.float64) for k, v in act.items()}
    full["means3D"][idx] = g["dmeans3D"]
    full["shs"][idx] = g["dsh"]
    full["scales"][idx] = g["dscales"]
    full["rots"][idx] = g["drotations"]
    full["opac"] = dop.reshape(act["opac"].shape)
    return pre, full


# --------------------------------------------------------------------------
# 10k Gaussians, 200x200, L1: every tile
# --------------------------------------------------------------------------
def test_small_view_l1_every_tile():
    print("case: 10000 Gaussians, one 200x200 view, L1 against a seeded target, every tile", flush=True)
    W = H = 200
    model = S.free_scene(10_000, seed=0, layout="uniform")
    cam = C.front_camera(W, H, device="cpu")
    bg = torch.ones(3)
    target = S.target_image(H, W, seed=1)
    leaves, act = SP.fp64_activations(SP.raw_params(model))
    st = OR.settings_for(cam, bg.double(), 3)
    ocol, orad, ost = OR.rasterize(act["means3D"].detach(), act["opac"].detach(), st, shs=act["shs"].detach(),
                                   scales=act["scales"].detach(), rotations=act["rots"].detach())
    m = model.to("cuda").requires_grad_(True)
    out = render(C.front_camera(W, H, device="cuda"), m, PipelineParams(), bg.cuda())
    img = out["render"]
    ok = ~ost["ambiguous"]
    err = (img.detach().cpu().double() - ocol).abs().amax(0)[ok].max().item()
    print(f"  image max |diff| {err:.3e} on {int(ok.sum())} untied pixels", flush=True)
    assert err < SP.IMG_TOL
    # L1 = mean |img - target|: dL/dimg = sign / (3 H W), threshold-tie pixels left out
    dL = torch.sign(img.detach().cpu() - target) / (3 * H * W)
    dL[:, ost["ambiguous"]] = 0
    img.backward(dL.cuda())
    og = OR.rasterize_backward(ost, st, act["means3D"].detach(), dL.double(), shs=act["shs"].detach(),
                               scales=act["scales"].detach(), rotations=act["rots"].detach())
    dmeans3D, dmeans2D, dsh, _, dop, dscales, drot = og[:7]
    torch.autograd.backward([act["means3D"], act["shs"], act["opac"], act["scales"], act["rots"]],
                            [dmeans3D, dsh, dop.reshape(act["opac"].shape), dscales, drot])
    SP.assert_grads([(k, getattr(m, k).grad, leaves[k].grad) for k in SP.RAW], "c0")
    SP.assert_grads([("means2D (NDC)", out["viewspace_points"].grad[:, :2], dmeans2D[:, :2])], "c0")


# --------------------------------------------------------------------------
# 1M Gaussians, one 800x800 view through the drop-in render
# --------------------------------------------------------------------------
def test_headline_single_view_1M_800():
    print("case: 1M Gaussians (60/40 ball / r=3 shell), one 800x800 view, drop-in render", flush=True)
    W = H = 800
    model = S.free_scene(1_000_000, seed=0, layout="mixture", shell_radius=3.0)
    cam = C.front_camera(W, H, device="cpu")
    bg = torch.ones(3)
    m = model.to("cuda")
    act_d = _activated_device(m)
    rs = settings_from_cam(cam, bg, 3)
    color, radii, state = RZ.forward_raw(act_d["means3D"], act_d["opac"], act_d["shs"], None, act_d["scales"],
                                         act_d["rots"], None, rs)
    sp = SP.single_view_splats(state, W, H)
    sp["radii"] = radii.cpu().numpy()
    # lists: every tile, bit-exact
    keys, vals, ranges = OR.bin_gaussians(sp["xy"], sp["radii"], sp["depth"], W, H)
    assert state["K"] == keys.size
    assert np.array_equal(sp["values"].astype(np.int64), vals)
    assert np.array_equal(sp["ranges"], ranges)
    assert np.array_equal(sp["tile_ids"].astype(np.uint64), keys >> np.uint64(32))
    print(f"  lists bit-exact: {keys.size} instances over {ranges.shape[0]} tiles", flush=True)
    # preprocess: every Gaussian against FP64
    leaves, act = SP.fp64_activations(SP.raw_params(model))
    st = oracle_settings(rs)
    pre = OR.preprocess(act["means3D"].detach(), act["opac"].detach(), st, shs=act["shs"].detach(),
                        scales=act["scales"].detach(), rotations=act["rots"].detach())
    nvis, ntied = SP.check_preprocess(sp, pre, "c2")
    print(f"  preprocess: {nvis} visible, {ntied} radius/rectangle ties", flush=True)
    # sampled tiles: staged forward, end-to-end on agreeing lists
    tiles = SP.pick_tiles(ranges[:, 1] - ranges[:, 0], 300, seed=0)
    amb, _ = SP.staged_forward(sp, ranges, vals, bg, W, H, tiles, color, "c2")
    SP.end_to_end_forward(pre, bg, W, H, tiles, ranges, vals, color, "c2")
    # backward: dL on the sampled tiles' untied pixels, through the drop-in render + autograd
    dL = SP.masked_dL(tiles, amb, W, H, seed=7)
    mg = model.to("cuda").requires_grad_(True)
    out = render(C.front_camera(W, H, device="cuda"), mg, PipelineParams(), bg.cuda())
    assert torch.equal(out["render"].detach(), color), "render and forward_raw images differ"
    out["render"].backward(dL.cuda())
    dm2, dcon, dop, dcol = SP.staged_backward(sp, ranges, vals, bg, W, H, dL, tiles)
    _, og = _oracle_chain(act, st, dm2, dcon, dop, dcol, sp["radii"])
    torch.autograd.backward([act[k] for k in og], [og[k] for k in og])
    SP.assert_grads([(k, getattr(mg, k).grad, leaves[k].grad) for k in SP.RAW], "c2")
    SP.assert_grads([("means2D (NDC)", out["viewspace_points"].grad[:, :2], dm2[:, :2])], "c2")
    # the training objective on this image: value and gradient against FP64
    target = S.target_image(H, W, seed=1)
    xg = color.detach().clone().requires_grad_(True)
    val = fused_dssim_l1(xg, target.cuda(), 0.2)
    val.backward()
    ref_val, ref_grad = OL.dssim_l1_with_grad(color.detach().cpu(), target, 0.2)
    print(f"  loss {val.item():.7f} vs {ref_val:.7f}", flush=True)
    assert abs(val.item() - ref_val) < 1e-5
    SP.assert_grads([("dloss/dimage", xg.grad, ref_grad)], "c2")


# --------------------------------------------------------------------------
# batched views
# --------------------------------------------------------------------------
def _batched_check(model, cams, bgs, tiles_per_view, full_lists, seed, tag):
    """Batched forward + backward of ``cams`` against the oracle (see module
    docstring); returns the renderer and the images."""
    r = BT.BatchedRenderer(cams, [b.cuda() for b in bgs], model.num, sh_degree=3)
    params = BT.pack_model(model.to("cuda"))
    r.plan(params)
    images = r.forward(params)
    torch.cuda.synchronize()
    assert not r.overflowed()
    K_cell, K_tile = SP.batched_counts(r)
    print(f"  {tag}: {r.B} views, {K_cell} cell instances, {K_tile} tile instances, cell mode {r.cell_mode}",
          flush=True)
    ktot = [0, 0]
    per_view = []
    lists = SP.batched_lists(r)
    for v in range(r.B):
        d = r.views[v]
        H, W = r.sizes[v]
        sp = SP.batched_view_splats(r, v)
        assert np.array_equal(sp["radii"][sp["radii"] > 0], sp["radius_bits"][sp["radii"] > 0])
        kt, kc = SP.tile_rect_counts(sp["xy"], sp["radii"], W, H)
        ktot[0] += kc
        ktot[1] += kt
        cell, mask, ids, _ = SP.batched_cell_entries(r, v, lists)
        if full_lists:
            _, ovals, oranges = OR.bin_gaussians(sp["xy"], sp["radii"], sp["depth"], W, H)
            granges, gvals = SP.expand_cell_lists(cell, mask, ids, d)
            assert np.array_equal(granges, oranges), f"{tag} view {v}: tile ranges differ"
            assert np.array_equal(gvals, ovals), f"{tag} view {v}: tile lists differ"
            lengths = oranges[:, 1] - oranges[:, 0]
            tiles = SP.pick_tiles(lengths, tiles_per_view, seed + v)
            ranges, vals = oranges, ovals
        else:
            tiles = SP.pick_tiles(SP.tile_lengths(cell, mask, d), tiles_per_view, seed + v)
            ranges, vals = SP.sampled_lists(sp["xy"], sp["radii"], sp["depth"], W, H, tiles)
            granges, gvals = SP.expand_cell_lists(cell, mask, ids, d, tiles=tiles)
            assert np.array_equal(granges, ranges), f"{tag} view {v}: sampled tile ranges differ"
            assert np.array_equal(gvals, vals), f"{tag} view {v}: sampled tile lists differ"
        amb, _ = SP.staged_forward(sp, ranges, vals, bgs[v], W, H, tiles, r.image(images, v), f"{tag} view {v}")
        per_view.append((sp, ranges, vals, tiles, amb))
    assert (K_cell, K_tile) == tuple(ktot), f"{tag}: counts {(K_cell, K_tile)} vs reference rectangles {ktot}"
    # backward of sum_v <dL_v, image_v> with dL on the sampled tiles' untied pixels
    dL = r.new_images().zero_()
    for v, (sp, ranges, vals, tiles, amb) in enumerate(per_view):
        H, W = r.sizes[v]
        r.image(dL, v).copy_(SP.masked_dL(tiles, amb, W, H, seed=100 + v).cuda())
    grads = BT.unpack_grads(r.backward(params, dL))
    torch.cuda.synchronize()
    leaves, act = SP.fp64_activations(SP.raw_params(model))
    total = None
    for v, (sp, ranges, vals, tiles, amb) in enumerate(per_view):
        H, W = r.sizes[v]
        dm2, dcon, dop, dcol = SP.staged_backward(sp, ranges, vals, bgs[v], W, H, r.image(dL, v).cpu(), tiles)
        touched = np.nonzero((dm2.abs().sum(1) + dcon.abs().sum(1) + dcol.abs().sum(1) + dop.abs().reshape(-1))
                             .numpy() > 0)[0]
        st = OR.settings_for(_cpu_cam(cams[v]), bgs[v].double(), 3)
        _, og = _oracle_chain(act, st, dm2, dcon, dop, dcol, sp["radii"], subset=touched)
        total = og if total is None else {k: total[k] + og[k] for k in og}
    torch.autograd.backward([act[k] for k in total], [total[k] for k in total])
    SP.assert_grads([(k, grads[k], leaves[k].grad) for k in SP.RAW], tag)
    return r, images


def test_headline_batch_1M_8x800():
    print("case: 1M Gaussians, 8 orbit views of 800x800 in one batch", flush=True)
    model = S.free_scene(1_000_000, seed=0, layout="mixture", shell_radius=3.0)
    pos = C.sphere_positions(8, 4.0, seed=2)
    cams = [C.camera_at(p, 800, 800, fov_deg=50.0, device="cpu") for p in pos]
    bgs = [torch.ones(3)] * 8
    r, images = _batched_check(model, cams, bgs, tiles_per_view=40, full_lists=True, seed=10, tag="c2x8")
    # the per-view training objective of the batch against FP64, on two views
    targets = [S.target_image(800, 800, seed=1 + i) for i in range(8)]
    loss = BT.BatchedLoss(r, targets, 0.2)
    losses = loss.forward(images).cpu()
    dimg = loss.backward(images, r.new_images())
    torch.cuda.synchronize()
    for v in (0, 5):
        ref_val, ref_grad = OL.dssim_l1_with_grad(r.image(images, v).cpu(), targets[v], 0.2)
        print(f"  view {v} loss {losses[v, 0].item():.7f} vs {ref_val:.7f}", flush=True)
        assert abs(losses[v, 0].item() - ref_val) < 1e-5
        SP.assert_grads([(f"view {v} dloss/dimage", r.image(dimg, v), ref_grad)], "c2x8")


def test_uneven_batch_3M_16_views():
    print("case: 3M Gaussians (r=5 shell), 16 orbit views of 512^2 .. 1920x1080 in one batch", flush=True)
    model = S.free_scene(3_000_000, seed=0, layout="mixture", shell_radius=5.0)
    pos = C.sphere_positions(16, 4.0, seed=2)
    g = np.random.default_rng(2)
    sizes = [UNEVEN_SIZES[i] for i in g.integers(0, len(UNEVEN_SIZES), size=16)]
    assert (1920, 1080) in sizes
    cams = [C.camera_at(p, w, h, fov_deg=50.0, device="cpu") for p, (w, h) in zip(pos, sizes)]
    bgs = [torch.ones(3)] * 16
    _batched_check(model, cams, bgs, tiles_per_view=12, full_lists=False, seed=30, tag="c3")


# --------------------------------------------------------------------------
# serving
# --------------------------------------------------------------------------
def test_serving_scenes_to_1080p():
    """Four resident scenes (0.5M..3M Gaussians); 24 requests at 256^2 ..
    1920x1080 served concurrently are bit-identical to the same requests
    served one at a time; four of them (one per scene, the 1080p ones first)
    equal the oracle on sampled tiles."""
    print("case: serving, 4 scenes of 0.5M/1M/2M/3M, 256^2 .. 1920x1080", flush=True)
    scenes = {}
    models = {}
    for i, n in enumerate((500_000, 1_000_000, 2_000_000, 3_000_000)):
        models[f"s{n}"] = S.free_scene(n, seed=10 + i, layout="mixture", shell_radius=3.0)
        scenes[f"s{n}"] = BT.pack_model(models[f"s{n}"].to("cuda"))
    names = sorted(scenes)
    srv = SV.RenderServer(scenes, max_views=8, num_streams=4)
    mix = SV.poisson_requests(names, 24, None, seed=5, device="cuda")
    done = srv.run(mix)
    assert all(q.status == "done" for q in done) and len(done) == 24
    assert max(q.batch_size for q in done) > 1, "the mix must batch requests"
    sizes = {q.resolution for q in done}
    assert (1920, 1080) in sizes and (256, 256) in sizes
    imgs = {q.id: q.image.clone() for q in done}
    alone = SV.poisson_requests(names, 24, None, seed=5, device="cuda")
    for q in alone:
        d = srv.run([q])
        assert d[0].status == "done" and d[0].batch_size == 1
    for q_mix, q_alone in zip(sorted(mix, key=lambda q: q.id), alone):
        assert torch.equal(imgs[q_mix.id], q_alone.image), f"request {q_mix.id}: mix and alone differ"
    # oracle on sampled tiles for one request per scene (1080p first)
    picked = {}
    for q in sorted(mix, key=lambda q: -q.resolution[0] * q.resolution[1]):
        picked.setdefault(q.scene, q)
    for name, q in sorted(picked.items()):
        W, H = q.resolution
        r = BT.BatchedRenderer([q.camera], [q.bg.cuda()], scenes[name].shape[1], sh_degree=3)
        r.plan(scenes[name], backward=False)
        img = r.forward(scenes[name])
        torch.cuda.synchronize()
        assert torch.equal(r.image(img, 0), imgs[q.id]), f"{name}: served image differs from its batched render"
        sp = SP.batched_view_splats(r, 0)
        cell, mask, ids, _ = SP.batched_cell_entries(r, 0, SP.batched_lists(r))
        tiles = SP.pick_tiles(SP.tile_lengths(cell, mask, r.views[0]), 12, seed=len(name))
        ranges, vals = SP.sampled_lists(sp["xy"], sp["radii"], sp["depth"], W, H, tiles)
        granges, gvals = SP.expand_cell_lists(cell, mask, ids, r.views[0], tiles=tiles)
        assert np.array_equal(granges, ranges) and np.array_equal(gvals, vals), f"{name}: sampled lists differ"
        SP.staged_forward(sp, ranges, vals, q.bg, W, H, tiles, imgs[q.id], f"serve {name} {W}x{H}")
        # preprocess of this view against FP64 (every Gaussian of the scene)
        _, act = SP.fp64_activations(SP.raw_params(models[name]))
        st = OR.settings_for(_cpu_cam(q.camera), q.bg.double(), 3)
        pre = OR.preprocess(act["means3D"].detach(), act["opac"].detach(), st, shs=act["shs"].detach(),
                            scales=act["scales"].detach(), rotations=act["rots"].detach())
        SP.check_preprocess(sp, pre, f"serve {name}")
        del r, img
