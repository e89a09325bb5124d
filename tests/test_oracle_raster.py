"""Pins of the FP64 rasterizer oracle (CPU): brute-force binning, a textbook
per-pixel compositing loop, closed forms, autograd and finite differences."""
import math

import numpy as np
import pytest
import torch

from oracle import camera as OC
from oracle import raster as OR
from oracle import raster_autograd as RA

F64 = torch.float64


def _scene(n, seed, spread=0.8, log_scale=-1.6, opacity_max=0.9, shs=True):
    g = torch.Generator().manual_seed(seed)
    means = (torch.rand(n, 3, generator=g) * 2 - 1) * spread
    scales = torch.exp(torch.randn(n, 3, generator=g) * 0.3 + log_scale)
    rots = torch.randn(n, 4, generator=g)
    rots = rots / rots.norm(dim=1, keepdim=True)
    opac = torch.sigmoid(torch.randn(n, 1, generator=g)) * opacity_max
    sh = torch.randn(n, 16, 3, generator=g) * 0.3
    return means, opac, sh, scales, rots


def _settings(W, H, bg=(1.0, 1.0, 1.0), mod=1.0, fov=50.0):
    cam = OC.front_camera(W, H, fov_deg=fov)
    return OR.settings_for(cam, torch.tensor(bg), 3, scale_modifier=mod)


# ---------------------------------------------------------------- binning ---

def test_rect_hand_cases():
    # pixel (8, 8) radius 3 -> tile [0, 1) x [0, 1); (16.5, 8) radius 1 -> x tiles [0, 2)
    xy = np.array([[8.0, 8.0], [15.5, 8.0], [16.5, 8.0], [-40.0, 8.0], [100.0, 100.0]], np.float32)
    r = np.array([3, 1, 1, 10, 20])
    x0, y0, x1, y1 = OR.rect_f32(xy, r, 64, 48)
    assert (x0[0], x1[0], y0[0], y1[0]) == (0, 1, 0, 1)
    assert (x0[1], x1[1]) == (0, 1)
    assert (x0[2], x1[2]) == (0, 2)
    assert x1[3] == 0                      # entirely left of the image: empty
    assert (x1[4], y1[4]) == (4, 3)        # clamped to the 4 x 3 grid


def _brute_force_bins(xy, radii, depth, W, H):
    gx, gy = OR.tile_grid(W, H)
    items = []
    for i in range(len(radii)):
        if radii[i] <= 0:
            continue
        x0, y0, x1, y1 = [int(v[0]) for v in OR.rect_f32(xy[i:i + 1], radii[i:i + 1], W, H)]
        dbits = int(np.float32(depth[i]).view(np.uint32))
        for ty in range(y0, y1):
            for tx in range(x0, x1):
                items.append((ty * gx + tx, dbits, i))
    items.sort()
    return items


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_binning_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    n, W, H = 300, 77, 53
    xy = rng.uniform(-20, 100, size=(n, 2)).astype(np.float32)
    radii = rng.integers(0, 30, size=n)
    depth = rng.uniform(0.3, 5, size=n).astype(np.float32)
    depth[::7] = depth[0]                  # equal depths: ties keep index order
    keys, vals, ranges = OR.bin_gaussians(xy, radii, depth, W, H)
    ref = _brute_force_bins(xy, radii, depth, W, H)
    assert len(ref) == keys.size
    assert [(int(k >> np.uint64(32)), int(k & np.uint64(0xFFFFFFFF)), int(v)) for k, v in zip(keys, vals)] == ref
    gx, gy = OR.tile_grid(W, H)
    for t in range(gx * gy):
        idx = [j for j, it in enumerate(ref) if it[0] == t]
        if idx:
            assert tuple(ranges[t]) == (idx[0], idx[-1] + 1)
        else:
            assert tuple(ranges[t]) == (0, 0)
    assert OR.higher_msb(2500) == 12 and OR.higher_msb(8160) == 13 and OR.higher_msb(1) == 1


# ----------------------------------------------------------------- blend ---

def _textbook_pixel(px, py, ids, xy, conic, opac, col, bg):
    """Front-to-back 'over' compositing of one pixel, written out per splat."""
    T, C, last = 1.0, np.zeros(col.shape[1]), 0
    for cnt, i in enumerate(ids, start=1):
        dx, dy = xy[i, 0] - px, xy[i, 1] - py
        a, b, c = conic[i]
        power = -0.5 * (a * dx * dx + c * dy * dy) - b * dx * dy
        if power > 0:
            continue
        alpha = min(0.99, opac[i] * math.exp(power))
        if alpha < 1.0 / 255.0:
            continue
        if T * (1 - alpha) < 1e-4:
            break
        C += col[i] * alpha * T
        T *= 1 - alpha
        last = cnt
    return C + T * bg, T, last


def test_blend_matches_textbook_loop():
    W = H = 32
    means, opac, sh, scales, rots = _scene(48, 5, spread=0.5)
    st = _settings(W, H, bg=(0.2, 0.4, 0.6))
    color, radii, state = OR.rasterize(means, opac, st, shs=sh, scales=scales, rotations=rots)
    pre = state["pre"]
    xy, conic = pre["xy"].numpy(), pre["conic"].numpy()
    o, col = pre["opacity"].numpy(), pre["rgb"].numpy()
    bg = st.bg.double().numpy()
    gx, _ = OR.tile_grid(W, H)
    vals, ranges = state["vals"], state["ranges"]
    for py in range(H):
        for px in range(W):
            t = (py // 16) * gx + px // 16
            ids = vals[ranges[t, 0]:ranges[t, 1]]
            ref, T, last = _textbook_pixel(px, py, ids, xy, conic, o, col, bg)
            np.testing.assert_allclose(color[:, py, px].numpy(), ref, rtol=0, atol=1e-12)
            assert abs(state["final_T"][py, px].item() - T) < 1e-12
            assert state["n_contrib"][py, px].item() == last


def test_blend_closed_forms():
    W = H = 16
    bg = torch.tensor([0.1, 0.2, 0.3], dtype=F64)
    gx = 1
    ranges = np.zeros((1, 2), np.int64)
    # empty tile -> background
    out, T, n, _ = OR.render_forward(ranges, np.zeros(0, np.int64), torch.zeros(0, 2), torch.zeros(0, 3),
                                     torch.zeros(0), torch.zeros(0, 3), bg, W, H)
    assert torch.allclose(out, bg[:, None, None].expand(3, H, W)) and (T == 1).all() and (n == 0).all()
    # one splat centred on pixel (5, 7): alpha = min(.99, o) there
    xy = torch.tensor([[5.0, 7.0]], dtype=F64)
    conic = torch.tensor([[0.1, 0.0, 0.1]], dtype=F64)
    col = torch.tensor([[1.0, 0.5, 0.0]], dtype=F64)
    ranges[0] = (0, 1)
    for o, a in ((0.6, 0.6), (0.999, 0.99)):
        out, T, n, _ = OR.render_forward(ranges, np.array([0]), xy, conic, torch.tensor([o], dtype=F64), col, bg,
                                         W, H)
        exp = a * col[0] + (1 - a) * bg
        assert torch.allclose(out[:, 7, 5], exp, atol=1e-14) and n[7, 5] == 1
    # saturation: three 0.99 splats on top of each other -> the third stops
    xy3, conic3 = xy.repeat(3, 1), conic.repeat(3, 1)
    col3 = torch.tensor([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]], dtype=F64)
    ranges[0] = (0, 3)
    out, T, n, _ = OR.render_forward(ranges, np.array([0, 1, 2]), xy3, conic3, torch.full((3,), 0.999, dtype=F64),
                                     col3, bg, W, H)
    t1 = 1 - 0.99
    assert n[7, 5] == 2 and abs(T[7, 5] - t1 * t1) < 1e-15
    assert torch.allclose(out[:, 7, 5], torch.tensor([0.99, 0.99 * t1, 0.0], dtype=F64) + t1 * t1 * bg, atol=1e-14)


# ------------------------------------------------------------ preprocess ---

def test_ewa_isotropic_on_axis_closed_form():
    W = H = 64
    st = _settings(W, H, fov=60.0)
    s = 0.05
    means = torch.tensor([[0.0, 0.0, 1.0]])        # view depth 5
    pre = OR.preprocess(means, torch.ones(1, 1) * 0.5, st, colors_precomp=torch.ones(1, 3),
                        scales=torch.full((1, 3), s), rotations=torch.tensor([[1.0, 0, 0, 0]]))
    f = W / (2 * float(np.float32(st.tanfovx)))
    exp = (f * float(np.float32(s)) / 5.0) ** 2 + 0.3
    cov = pre["cov2D"][0]
    assert abs(cov[0] - exp) < 1e-9 and abs(cov[2] - exp) < 1e-9 and abs(cov[1]) < 1e-12
    assert abs(pre["xy"][0, 0] - (W - 1) / 2) < 1e-9     # principal point: ((0 + 1) W - 1) / 2


def test_cov3d_invariants():
    g = torch.Generator().manual_seed(2)
    s = torch.exp(torch.randn(20, 3, generator=g, dtype=F64))
    q = torch.randn(20, 4, generator=g, dtype=F64)
    q = q / q.norm(dim=1, keepdim=True)
    c = OR.cov3d_from_scale_rot(s, q, 1.7)
    S = OR.six_to_sym(c)
    ev = torch.linalg.eigvalsh(S)
    assert torch.allclose(ev.sort(1).values, ((1.7 * s) ** 2).sort(1).values, rtol=1e-10)
    assert torch.allclose(OR.cov3d_from_scale_rot(s, -q, 1.7), c, atol=1e-12)


def _anisotropic_preprocess(n=400, seed=3):
    """Strongly anisotropic Gaussians in front of the camera (random
    rotations, scales spread over two decades)."""
    st = _settings(160, 120, fov=55.0)
    g = torch.Generator().manual_seed(seed)
    means = (torch.rand(n, 3, generator=g, dtype=F64) * 2 - 1) * torch.tensor([0.8, 0.6, 1.5], dtype=F64)
    scales = torch.exp(torch.rand(n, 3, generator=g, dtype=F64) * 4.6 - 5.5)
    rots = torch.randn(n, 4, generator=g, dtype=F64)
    pre = OR.preprocess(means, torch.full((n, 1), 0.5, dtype=F64), st, colors_precomp=torch.ones(n, 3),
                        scales=scales, rotations=rots)
    return st, pre


def test_conic_is_the_inverse_of_cov2d():
    """conic @ cov2D = I for every Gaussian (2x2 matrix product, independent
    of the closed-form adjugate the oracle uses)."""
    _, pre = _anisotropic_preprocess()
    vis = pre["visible"]
    assert vis.sum() > 100
    a, b, c = pre["cov2D"][vis].unbind(1)
    A = torch.stack([a, b, b, c], 1).reshape(-1, 2, 2)
    ca, cb, cc = pre["conic"][vis].unbind(1)
    Q = torch.stack([ca, cb, cb, cc], 1).reshape(-1, 2, 2)
    eye = torch.eye(2, dtype=F64).expand_as(A)
    assert torch.allclose(Q @ A, eye, atol=1e-9)
    assert torch.allclose(A @ Q, eye, atol=1e-9)
    # the conic is symmetric positive definite (a > 0, det > 0)
    assert (ca > 0).all() and (ca * cc - cb * cb > 0).all()


def test_radius_is_three_sigma_of_the_largest_eigenvalue():
    """radius = ceil(3 sqrt(lambda_max(cov2D))) from a symmetric eigen
    solver wherever the eigenvalue gap is wide ((a - c)^2 / 4 + b^2 >= 0.1);
    where it is narrower the reference floors the root at sqrt(0.1), so
    radius = ceil(3 sqrt(mid + sqrt(0.1)))."""
    _, pre = _anisotropic_preprocess()
    vis = pre["visible"]
    a, b, c = pre["cov2D"][vis].unbind(1)
    A = torch.stack([a, b, b, c], 1).reshape(-1, 2, 2).numpy()
    lam_max = np.linalg.eigvalsh(A)[:, 1]
    gap = (a - c) ** 2 / 4 + b * b
    wide = (gap >= 0.1).numpy()
    assert wide.sum() > 100 and (~wide).sum() >= 5
    r = pre["radius_f"][vis].numpy()
    np.testing.assert_array_equal(r[wide], np.ceil(3 * np.sqrt(lam_max[wide])))
    mid = ((a + c) / 2).numpy()
    np.testing.assert_array_equal(r[~wide], np.ceil(3 * np.sqrt(mid[~wide] + math.sqrt(0.1))))
    # the radius is never below the largest eigenvalue's three sigma
    assert (r >= 3 * np.sqrt(lam_max) - 1e-9).all()


def test_radius_floor_special_case():
    """A point-like Gaussian far away has cov2D = 0.3 I (the dilation only):
    mid^2 - det = 0, the floor gives lambda = 0.3 + sqrt(0.1) and radius
    ceil(3 sqrt(0.6162)) = 3, where the true eigenvalue 0.3 would give 2."""
    st = _settings(64, 64, fov=60.0)
    pre = OR.preprocess(torch.tensor([[0.0, 0.0, 40.0]]), torch.full((1, 1), 0.5), st,
                        colors_precomp=torch.ones(1, 3), scales=torch.full((1, 3), 1e-7),
                        rotations=torch.tensor([[1.0, 0, 0, 0]]))
    assert torch.allclose(pre["cov2D"][0], torch.tensor([0.3, 0.0, 0.3], dtype=F64), atol=1e-12)
    assert int(pre["radii"][0]) == 3
    assert math.ceil(3 * math.sqrt(0.3)) == 2


# -------------------------------------------------------------- backward ---

def _grads_hand(state, st, means, dL, kw):
    return OR.rasterize_backward(state, st, means, dL, **kw)


def _leaves(means, opac, sh, scales, rots):
    return [t.to(F64).clone().requires_grad_(True) for t in (means, opac, sh, scales, rots)]


@pytest.mark.parametrize("mode", ["sh", "precomp"])
def test_backward_matches_autograd(mode):
    W, H = 40, 32
    means, opac, sh, scales, rots = _scene(30, 11, spread=0.6, log_scale=-1.2)
    st = _settings(W, H, bg=(0.3, 0.6, 0.9))
    if mode == "sh":
        kw = dict(shs=sh, scales=scales, rotations=rots)
    else:
        cols = torch.rand(30, 3, generator=torch.Generator().manual_seed(3))
        cov = OR.cov3d_from_scale_rot(scales.double(), rots.double(), 1.0).float()
        kw = dict(colors_precomp=cols, cov3D_precomp=cov)
    _, _, state = OR.rasterize(means, opac, st, **kw)
    assert state["pre"]["cov2D"][state["pre"]["visible"]][:, [0, 2]].min() > 1.0   # det^2 >> 1e-7
    dL = torch.randn(3, H, W, generator=torch.Generator().manual_seed(4), dtype=F64)
    hand = _grads_hand(state, st, means, dL, kw)
    m, o = means.to(F64).clone().requires_grad_(True), opac.to(F64).clone().requires_grad_(True)
    m2 = torch.zeros(30, 3, dtype=F64, requires_grad=True)
    if mode == "sh":
        s_, sc, r_ = [t.to(F64).clone().requires_grad_(True) for t in (sh, scales, rots)]
        out = RA.rasterize_autograd(m, m2, o, st, shs=s_, scales=sc, rotations=r_, state=state)
        ag = torch.autograd.grad(out, [m, m2, o, s_, sc, r_], dL)
        pairs = [(hand[0], ag[0]), (hand[1], ag[1]), (hand[4], ag[2]), (hand[2], ag[3]), (hand[5], ag[4]),
                 (hand[6], ag[5])]
    else:
        c_ = kw["colors_precomp"].to(F64).clone().requires_grad_(True)
        cv = kw["cov3D_precomp"].to(F64).clone().requires_grad_(True)
        out = RA.rasterize_autograd(m, m2, o, st, colors_precomp=c_, cov3D_precomp=cv, state=state)
        ag = torch.autograd.grad(out, [m, m2, o, c_, cv], dL)
        pairs = [(hand[0], ag[0]), (hand[1], ag[1]), (hand[4], ag[2]), (hand[3], ag[3]), (hand[7], ag[4])]
    for h, a in pairs:
        h = h.reshape(a.shape)
        assert (h - a).norm() <= 1e-7 * a.norm() + 1e-12


def test_autograd_formulation_matches_finite_differences():
    """Where no quirk is active (alpha below the 0.99 clamp, no tan(fov) clamp,
    modifier 1), autograd of the differentiable forward equals central
    differences of it, with the blend membership held fixed."""
    W, H = 32, 32
    means, opac, sh, scales, rots = _scene(12, 21, spread=0.4, log_scale=-1.5, opacity_max=0.8)
    st = _settings(W, H)
    _, _, state = OR.rasterize(means, opac, st, shs=sh, scales=scales, rotations=rots)
    dL = torch.randn(3, H, W, generator=torch.Generator().manual_seed(9), dtype=F64)
    leaves = _leaves(means, opac, sh, scales, rots)
    m2 = torch.zeros(12, 3, dtype=F64, requires_grad=True)
    f = lambda: (RA.rasterize_autograd(leaves[0], m2, leaves[1], st, shs=leaves[2], scales=leaves[3],
                                       rotations=leaves[4], state=state) * dL).sum()
    grads = torch.autograd.grad(f(), leaves + [m2])
    rng = np.random.default_rng(0)
    h = 1e-6
    with torch.no_grad():
        for li, leaf in enumerate(leaves + [m2]):
            flat = leaf.view(-1)
            for j in rng.choice(flat.numel(), size=min(6, flat.numel()), replace=False):
                if li == 5 and j % 3 == 2:
                    continue                                # means2D.z is unused
                old = flat[j].item()
                flat[j] = old + h
                fp = f().item()
                flat[j] = old - h
                fm = f().item()
                flat[j] = old
                fd = (fp - fm) / (2 * h)
                ad = grads[li].reshape(-1)[j].item()
                assert abs(fd - ad) <= 1e-5 * max(1.0, abs(ad)), (li, j, fd, ad)


def test_opacity_clamp_quirk():
    """The 0.99 clamp is not masked in the backward: dL/do = G dL/dalpha even
    where o G > 0.99 (the true derivative there is 0)."""
    W = H = 16
    bg = torch.zeros(3, dtype=F64)
    xy = torch.tensor([[5.0, 7.0]], dtype=F64)
    conic = torch.tensor([[0.5, 0.0, 0.5]], dtype=F64)
    col = torch.ones(1, 3, dtype=F64)
    ranges = np.array([[0, 1]])
    dpix = torch.zeros(3, H, W, dtype=F64)
    dpix[:, 7, 5] = 1.0
    _, _, dop, _ = OR.render_backward(ranges, np.array([0]), xy, conic, torch.tensor([0.999], dtype=F64), col, bg,
                                      W, H, dpix)
    assert abs(dop[0, 0].item() - 3.0) < 1e-12          # G = 1 at the centre, dC/dalpha = 3 (white, black bg)
