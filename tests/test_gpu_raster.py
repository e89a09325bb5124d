"""GPU parity of the HIP rasterizer against the FP64 oracle (MI355X).

Bars (from the hot-path contract): tile lists, sorted keys and per-tile ranges
bit-exact; images within 1e-4 abs (FP32 blending) on every pixel the oracle
does not flag as a threshold tie (alpha ~ 1/255, T ~ 1e-4, power ~ 0);
gradients within 1e-3 norm-relative per tensor.
"""
import math

import numpy as np
import pytest
import torch

from dmgs_amd import cameras as C
from dmgs_amd import rasterizer as RZ
from dmgs_amd import scenes as S
from oracle import raster as OR
from tests.helpers import grad_rel_err, oracle_settings, settings_from_cam, workspace_views

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-4
GRAD_TOL = 1e-3


def _scene(n, seed=0, layout="uniform", sh_degree=3):
    m = S.free_scene(n, seed=seed, layout=layout, sh_degree=sh_degree)
    return dict(means3D=m.get_xyz.contiguous(), opac=m.get_opacity.contiguous(),
                shs=m.get_features.contiguous(), scales=m.get_scaling.contiguous(),
                rots=m.get_rotation.contiguous(), model=m)


def _run(sc, rs, use_cov=False, use_colors=False):
    dev = "cuda"
    t = {k: v.to(dev) for k, v in sc.items() if k != "model"}
    cov = sc["model"].get_covariance().to(dev) if use_cov else None
    cols = (torch.sigmoid(sc["shs"][:, 0, :]) if use_colors else None)
    color, radii, state = RZ.forward_raw(t["means3D"], t["opac"], None if use_colors else t["shs"],
                                         None if not use_colors else cols.to(dev),
                                         None if use_cov else t["scales"], None if use_cov else t["rots"],
                                         cov, rs)
    return color, radii, state, cov, cols


def _cam(W, H, fov=50.0, dist=4.0):
    return C.front_camera(W, H, fov_deg=fov, distance=dist, device="cpu")


@pytest.mark.parametrize("W,H,n", [(200, 200, 10000), (77, 53, 3000)])
def test_preprocess_binning_forward(W, H, n):
    print(f"case: preprocess/binning/forward W={W} H={H} n={n}", flush=True)
    sc = _scene(n)
    cam = _cam(W, H)
    bg = torch.ones(3)
    rs = settings_from_cam(cam, bg, 3)
    color, radii, state, _, _ = _run(sc, rs)
    v = workspace_views(state, W, H)
    ors = oracle_settings(rs)
    pre = OR.preprocess(sc["means3D"], sc["opac"], ors, shs=sc["shs"], scales=sc["scales"], rotations=sc["rots"])

    # --- preprocess ---
    r_gpu = radii.cpu().numpy()
    r_ora = pre["radii"].numpy()
    tied = pre["ambiguous"].numpy()
    mism = (r_gpu != r_ora) & ~tied
    assert mism.sum() == 0, f"radii differ on {mism.sum()} untied Gaussians"
    vis = (r_gpu > 0) & (r_ora > 0)
    assert np.abs(v["xy"][vis] - pre["xy"].numpy()[vis]).max() < 1e-3
    con_o = pre["conic"].numpy()[vis]
    assert (np.abs(v["conic"][vis] - con_o) / (np.abs(con_o) + 1e-6)).max() < 1e-4 or \
        np.abs(v["conic"][vis] - con_o).max() < 1e-6
    assert np.abs(v["rgb"][vis] - pre["rgb"].numpy()[vis]).max() < 1e-5
    assert np.abs(v["depth"][vis] - pre["depth"].numpy()[vis]).max() < 1e-5
    assert (v["radius_bits"] == r_gpu).all()

    # --- binning, bit-exact against the oracle fed the device's FP32 outputs ---
    keys, vals, ranges = OR.bin_gaussians(v["xy"], r_gpu, v["depth"], W, H)
    K = state["K"]
    assert K == keys.size
    assert (v["values"].astype(np.int64) == vals).all()
    assert (v["tile_ids"].astype(np.uint64) == (keys >> np.uint64(32))).all()
    dbits = v["depth"].view(np.uint32)[v["values"]].astype(np.uint64)
    assert ((v["tile_ids"].astype(np.uint64) << np.uint64(32) | dbits) == keys).all()
    assert (v["ranges"] == ranges).all()

    # --- blend forward on the device's own preprocess + lists ---
    out, fT, nc, amb = OR.render_forward(v["ranges"], v["values"], torch.from_numpy(v["xy"]),
                                         torch.from_numpy(v["conic"]), torch.from_numpy(v["opacity"]),
                                         torch.from_numpy(v["rgb"]), bg, W, H)
    ok = ~amb.numpy()
    diff = (color.cpu().double() - out).abs().numpy()
    print(f"  staged forward: max |diff| {diff.max(0)[ok].max():.3e}, flagged pixels {int((~ok).sum())}")
    assert diff.max(0)[ok].max() < IMG_TOL
    assert (v["n_contrib"][ok] == nc.numpy()[ok]).all()
    assert np.abs(v["final_T"][ok] - fT.numpy()[ok]).max() < IMG_TOL

    # --- end to end against the oracle's own FP64 preprocess ---
    ocol, _, ost = OR.rasterize(sc["means3D"], sc["opac"], ors, shs=sc["shs"], scales=sc["scales"],
                                rotations=sc["rots"])
    ok2 = ~ost["ambiguous"].numpy()
    d2 = (color.cpu().double() - ocol).abs().numpy().max(0)
    print(f"  end-to-end: max |diff| {d2[ok2].max():.3e} on {int(ok2.sum())} untied pixels")
    assert d2[ok2].max() < IMG_TOL


@pytest.mark.parametrize("mode", ["sh_scale_rot", "colors_cov"])
def test_backward_matches_oracle(mode):
    print(f"case: backward mode={mode}", flush=True)
    W, H, n = 96, 80, 2000
    sc = _scene(n, seed=3)
    cam = _cam(W, H)
    bg = torch.tensor([0.2, 0.5, 1.0])
    rs = settings_from_cam(cam, bg, 3)
    use = mode == "colors_cov"
    color, radii, state, cov, cols = _run(sc, rs, use_cov=use, use_colors=use)
    ors = oracle_settings(rs)
    kw = dict(colors_precomp=cols, cov3D_precomp=cov.cpu()) if use else dict(shs=sc["shs"], scales=sc["scales"],
                                                                            rotations=sc["rots"])
    ocol, _, ost = OR.rasterize(sc["means3D"], sc["opac"], ors, **kw)
    g = torch.Generator().manual_seed(7)
    dL = torch.randn(3, H, W, generator=g)
    dL[:, ost["ambiguous"]] = 0.0   # keep threshold ties out of the gradient comparison
    gg = RZ.backward_raw(state, radii, dL.cuda())
    og = OR.rasterize_backward(ost, ors, sc["means3D"], dL.double(), **kw)
    names = ["means3D", "means2D", "sh", "colors", "opacity", "scales", "rotations", "cov3D"]
    for i, nm in enumerate(names):
        ref = og[i]
        if ref is None:
            continue
        got = gg[nm]
        if nm == "means2D":
            got = got[:, :2]
            ref = ref[:, :2]
        e = grad_rel_err(got, ref)
        print(f"  {nm}: rel err {e:.2e}")
        assert e < GRAD_TOL, f"{nm} gradient rel err {e}"
    e = grad_rel_err(gg["conic"].reshape(-1, 4)[:, [0, 1, 3]], og[8][:, [0, 1, 3]])
    print(f"  conic: rel err {e:.2e}")
    assert e < GRAD_TOL


def _fwd_bwd_against_oracle(sc, cam, rs, seed):
    """Forward and backward of scene ``sc`` under settings ``rs`` against the
    oracle (image off threshold ties, every gradient); returns the device
    gradients and the oracle image."""
    W, H = rs.image_width, rs.image_height
    color, radii, state, _, _ = _run(sc, rs)
    ors = oracle_settings(rs)
    kw = dict(shs=sc["shs"], scales=sc["scales"], rotations=sc["rots"])
    ocol, _, ost = OR.rasterize(sc["means3D"], sc["opac"], ors, **kw)
    ok = ~ost["ambiguous"].numpy()
    d = (color.cpu().double() - ocol).abs().numpy().max(0)
    print(f"  image: max |diff| {d[ok].max():.3e} on {int(ok.sum())} untied pixels")
    assert d[ok].max() < IMG_TOL
    g = torch.Generator().manual_seed(seed)
    dL = torch.randn(3, H, W, generator=g)
    dL[:, ost["ambiguous"]] = 0.0
    gg = RZ.backward_raw(state, radii, dL.cuda())
    og = OR.rasterize_backward(ost, ors, sc["means3D"], dL.double(), **kw)
    for i, nm in enumerate(["means3D", "means2D", "sh", "colors", "opacity", "scales", "rotations"]):
        ref, got = og[i], gg[nm]
        if ref is None:
            continue
        if nm == "means2D":
            got, ref = got[:, :2], ref[:, :2]
        e = grad_rel_err(got, ref)
        print(f"  {nm}: rel err {e:.2e}")
        assert e < GRAD_TOL, f"{nm} gradient rel err {e}"
    return gg, ocol


def test_active_sh_degree_below_model_degree():
    """Training raises the active SH degree step by step while the features
    keep all 16 coefficients: with settings.sh_degree = 1 on a degree-3 model
    the colours use the first 4 coefficients only (image against the oracle
    evaluated at degree 1), every gradient matches the oracle, and the
    coefficients above the active degree get exactly zero gradient."""
    print("case: active SH degree 1 of 3", flush=True)
    sc = _scene(2000, seed=5)
    cam = _cam(96, 80)
    bg = torch.tensor([0.3, 0.1, 0.6])
    gg, ocol = _fwd_bwd_against_oracle(sc, cam, settings_from_cam(cam, bg, 1), seed=9)
    kw = dict(shs=sc["shs"], scales=sc["scales"], rotations=sc["rots"])
    ocol3, _, _ = OR.rasterize(sc["means3D"], sc["opac"], oracle_settings(settings_from_cam(cam, bg, 3)), **kw)
    assert (ocol3 - ocol).abs().max().item() > 1e-3  # degree 3 would give another image
    assert torch.count_nonzero(gg["sh"][:, 4:]).item() == 0
    assert torch.count_nonzero(gg["sh"][:, 1:4]).item() > 0


def test_scale_modifier_matches_oracle():
    """The scaling modifier multiplies every scale before the 3D covariance
    (the viewer's size slider): at 0.6 the image and every gradient match the
    oracle, and the image differs from the unmodified one."""
    print("case: scale modifier 0.6", flush=True)
    sc = _scene(2000, seed=6)
    cam = _cam(96, 80)
    bg = torch.tensor([0.4, 0.4, 0.1])
    _, ocol = _fwd_bwd_against_oracle(sc, cam, settings_from_cam(cam, bg, 3, scale_modifier=0.6), seed=10)
    kw = dict(shs=sc["shs"], scales=sc["scales"], rotations=sc["rots"])
    ocol1, _, _ = OR.rasterize(sc["means3D"], sc["opac"], oracle_settings(settings_from_cam(cam, bg, 3)), **kw)
    assert (ocol1 - ocol).abs().max().item() > 1e-2


def test_sort_pairs_stable():
    from dmgs_amd import _native as N
    rng = np.random.default_rng(0)
    for n, bits in [(1, 8), (1000, 12), (70000, 13), (300001, 32), (4096 * 3 + 17, 16), (5000, 0)]:
        print(f"case: sort n={n} bits={bits}", flush=True)
        hi = (1 << bits) if bits < 32 else (1 << 32)
        k = rng.integers(0, max(hi, 1), size=n, dtype=np.uint64).astype(np.uint32)
        if bits == 0:
            k[:] = 0
        v = np.arange(n, dtype=np.uint32)
        kt = torch.from_numpy(k.view(np.int32)).cuda()
        vt = torch.from_numpy(v.view(np.int32)).cuda()
        k1, v1 = torch.empty_like(kt), torch.empty_like(vt)
        lib = N.load()
        scratch = torch.empty(lib.dmgs_sort_scratch_size(n), dtype=torch.uint8, device="cuda")
        N.check(lib.dmgs_sort_pairs(N.ptr(kt), N.ptr(vt), N.ptr(k1), N.ptr(v1), n, bits, N.ptr(scratch),
                                    N.stream_handle()), "sort")
        order = np.argsort(k, kind="stable")
        assert (kt.cpu().numpy().view(np.uint32) == k[order]).all()
        assert (vt.cpu().numpy().view(np.uint32) == order).all()


def test_inclusive_scan():
    from dmgs_amd import _native as N
    rng = np.random.default_rng(1)
    # up to 2048 blocks of 4096 each block adds up its own prefix; past that
    # a spine launch scans the block sums
    for n in [1, 255, 4096, 4097, 1_000_003, 2048 * 4096, 2048 * 4096 + 12_345]:
        print(f"case: scan n={n}", flush=True)
        x = rng.integers(0, 50, size=n, dtype=np.uint32)
        xt = torch.from_numpy(x.view(np.int32)).cuda()
        out = torch.empty_like(xt)
        lib = N.load()
        scratch = torch.empty(lib.dmgs_scan_scratch_size(n), dtype=torch.uint8, device="cuda")
        N.check(lib.dmgs_inclusive_scan_u32(N.ptr(xt), N.ptr(out), n, N.ptr(scratch), N.stream_handle()), "scan")
        assert (out.cpu().numpy().view(np.uint32) == np.cumsum(x).astype(np.uint32)).all()


def test_edge_cases():
    print("case: edge cases", flush=True)
    dev = "cuda"
    cam = _cam(64, 48)
    bg = torch.tensor([0.1, 0.2, 0.3])
    rs = settings_from_cam(cam, bg, 0)
    # P == 0 returns zeros (reference behaviour)
    e = torch.zeros(0, 3, device=dev)
    col, radii, st = RZ.forward_raw(e, torch.zeros(0, 1, device=dev), torch.zeros(0, 1, 3, device=dev), None,
                                    e, torch.zeros(0, 4, device=dev), None, rs)
    assert col.abs().max().item() == 0 and radii.numel() == 0
    # everything behind the camera: image is the background
    m = torch.tensor([[0.0, 0.0, -10.0], [0.1, 0.1, -5.0]], device=dev)
    col, radii, st = RZ.forward_raw(m, torch.full((2, 1), 0.5, device=dev), torch.ones(2, 1, 3, device=dev),
                                    None, torch.full((2, 3), 0.1, device=dev),
                                    torch.tensor([[1.0, 0, 0, 0]] * 2, device=dev), None, rs)
    assert (radii == 0).all()
    assert torch.allclose(col, bg.to(dev)[:, None, None].expand_as(col))
    # one centred opaque Gaussian: centre pixel = min(.99, o) c + (1 - .) bg
    m = torch.tensor([[0.0, 0.0, 0.0]], device=dev)
    sh = torch.zeros(1, 1, 3, device=dev)
    sh[0, 0] = torch.tensor([1.0, -1.0, 0.5]) / 0.28209479177387814  # rgb = sh*C0 + 0.5
    col, radii, st = RZ.forward_raw(m, torch.tensor([[0.7]], device=dev), sh, None,
                                    torch.full((1, 3), 0.3, device=dev), torch.tensor([[1.0, 0, 0, 0]], device=dev),
                                    None, rs)
    v = workspace_views(st, 64, 48)
    cx, cy = v["xy"][0]
    px, py = int(round(cx)), int(round(cy))
    # exact alpha at that pixel from the conic
    a, b, c = v["conic"][0]
    dx, dy = cx - px, cy - py
    alpha = min(0.99, 0.7 * math.exp(-0.5 * (a * dx * dx + c * dy * dy) - b * dx * dy))
    rgb = np.array([1.5, 0.0, 1.0])  # clamped at 0
    exp = alpha * rgb + (1 - alpha) * bg.numpy()
    assert np.abs(col[:, py, px].cpu().numpy() - exp).max() < 1e-5


def test_mark_visible():
    print("case: markVisible", flush=True)
    cam = _cam(64, 64)
    rs = settings_from_cam(cam, torch.ones(3), 0)
    pts = torch.tensor([[0.0, 0.0, 0.0], [0.0, 0.0, -3.9], [0.0, 0.0, -3.7], [5.0, 5.0, -10.0]], device="cuda")
    vis = RZ.GaussianRasterizer(rs).markVisible(pts).cpu().tolist()
    assert vis == [True, False, True, False]
