"""FP64 CPU restatement of the tile-based Gaussian rasterizer (test
infrastructure only).

The rasterizer lives in the un-vendored submodule ``diff-gaussian-rasterization``
(reference .gitmodules:4-6, no pinned commit; call sites
gaussian_renderer/__init__.py:36-51,86-94).  Its semantics are restated from the
published upstream algorithm ([U]); every quirk the reference gradients carry
is kept and named below.

Forward [U]
  preprocess  : near cull z_view <= 0.2; p_proj = p_hom / (w + 1e-7); cov3D =
                R(q) S^2 R(q)^T with q NOT normalised (or precomputed);
                EWA cov2D = J W Sigma W^T J^T + 0.3 I with tx/tz, ty/tz clamped
                to +-1.3 tan(fov/2); det == 0 culls; conic = inverse;
                radius = ceil(3 sqrt(max(l1, l2))), l = mid +- sqrt(max(0.1,
                mid^2 - det)); pixel = ((ndc + 1) S - 1) / 2; 16x16 tile rect
                (getRect, truncating float division, clamped to the grid), zero
                area culls; SH colour + 0.5 clamped at 0.
  binning     : key = (tile << 32) | float_bits(depth), stable sort; ranges.
                Done in integer numpy here — bit-exact.
  blend       : front-to-back; skip power > 0; alpha = min(0.99, o exp(power));
                skip alpha < 1/255; stop (without adding) once T (1 - alpha)
                < 1e-4; out = C + T bg.
Backward [U] (quirks kept)
  * dL/dalpha has the background term -T_final / (1 - alpha) (bg . dpix);
  * the 0.99 clamp is not masked: dalpha/do = G, dalpha/dG = o;
  * means2D gradient is w.r.t. NDC: dpix/dndc = (W/2, H/2); only x, y filled;
  * the conic off-diagonal accumulator is half the true derivative and the
    cov2D backward doubles it back; 1 / (det^2 + 1e-7) replaces 1 / det^2;
  * the tan(fov) clamp zeroes d/dtx (d/dty) and the clamped t.x is treated as
    a constant inside d/dtz;
  * no quaternion-normalisation term; dL/dscale is taken w.r.t. s = mod*scale
    (no extra factor mod);
  * SH gradient is zeroed per channel where the colour was clamped.

Pins (tests/test_oracle_raster.py): binning = brute force integer sort;
blend = per-pixel textbook "over" compositing loop and closed-form single
Gaussian / empty tile / saturation cases; backward = autograd of the
independent differentiable formulation in ``raster_autograd`` plus central
finite differences; EWA = isotropic on-axis closed form.  The upstream CUDA's
exact FP32 rounding is "parity unpinned" (no source or binary to compare).
"""
import math

import numpy as np
import torch

from . import sh as sh_mod

BLOCK_X = 16
BLOCK_Y = 16
F64 = torch.float64


class Settings:
    """Mirror of GaussianRasterizationSettings (gaussian_renderer/__init__.py:36-49)."""

    def __init__(self, image_height, image_width, tanfovx, tanfovy, bg, scale_modifier,
                 viewmatrix, projmatrix, sh_degree, campos, prefiltered=False, debug=False):
        self.image_height = int(image_height)
        self.image_width = int(image_width)
        self.tanfovx = float(tanfovx)
        self.tanfovy = float(tanfovy)
        self.bg = bg
        self.scale_modifier = float(scale_modifier)
        self.viewmatrix = viewmatrix
        self.projmatrix = projmatrix
        self.sh_degree = int(sh_degree)
        self.campos = campos
        self.prefiltered = prefiltered
        self.debug = debug


def settings_for(cam, bg, sh_degree, scale_modifier=1.0):
    """The settings ``render`` builds (gaussian_renderer/__init__.py:33-49)."""
    return Settings(cam.image_height, cam.image_width, math.tan(cam.FoVx * 0.5),
                    math.tan(cam.FoVy * 0.5), bg, scale_modifier, cam.world_view_transform,
                    cam.full_proj_transform, sh_degree, cam.camera_center)


def _d(t):
    return t.detach().cpu().to(F64)


def tile_grid(W, H):
    return (W + BLOCK_X - 1) // BLOCK_X, (H + BLOCK_Y - 1) // BLOCK_Y


def rect_f32(xy32, radii, W, H):
    """getRect [U] in exact FP32 semantics (numpy float32 rounds each op like
    the device): min = trunc((p - r) / 16), max = trunc((((p + r) + 16) - 1) / 16),
    both clamped to [0, grid]."""
    gx, gy = tile_grid(W, H)
    xy32 = np.asarray(xy32, dtype=np.float32)
    r = np.asarray(radii).astype(np.float32)
    b = np.float32(BLOCK_X)
    one = np.float32(1.0)
    with np.errstate(invalid="ignore", over="ignore"):
        x0 = np.trunc((xy32[:, 0] - r) / b)
        y0 = np.trunc((xy32[:, 1] - r) / b)
        x1 = np.trunc((((xy32[:, 0] + r) + b) - one) / b)
        y1 = np.trunc((((xy32[:, 1] + r) + b) - one) / b)
    clip = lambda v, hi: np.clip(np.nan_to_num(v, nan=0.0, posinf=hi, neginf=0), 0, hi).astype(np.int64)
    return clip(x0, gx), clip(y0, gy), clip(x1, gx), clip(y1, gy)


def _rect_f64(xy, radius, W, H):
    """getRect on FP64 values, plus a flag when a truncation sits within 1e-4
    pixel-tiles of an integer (an FP32 device may land on the other side)."""
    gx, gy = tile_grid(W, H)
    vals = [(xy[:, 0] - radius) / BLOCK_X, (xy[:, 1] - radius) / BLOCK_Y,
            (xy[:, 0] + radius + BLOCK_X - 1) / BLOCK_X, (xy[:, 1] + radius + BLOCK_Y - 1) / BLOCK_Y]
    tie = torch.zeros(xy.shape[0], dtype=torch.bool)
    out = []
    for v, hi in zip(vals, (gx, gy, gx, gy)):
        t = torch.trunc(v)
        frac = (v - torch.round(v)).abs()
        tie |= (frac < 1e-4) & (torch.round(v) >= 0) & (torch.round(v) <= hi)
        out.append(t.clamp(0, hi).to(torch.int64))
    return out[0], out[1], out[2], out[3], tie


def quat_to_rot(q):
    """Rotation matrix of an un-normalised quaternion (r, x, y, z) [U] (same
    polynomial as utils/general_utils.py:89-97 after normalisation)."""
    r, x, y, z = q.unbind(-1)
    R = torch.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y),
        2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x),
        2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y)], -1)
    return R.reshape(q.shape[:-1] + (3, 3))


def cov3d_from_scale_rot(scales, rots, mod):
    """Sigma = R S S^T R^T as the 6-vector (xx, xy, xz, yy, yz, zz) [U]
    (strip order utils/general_utils.py:64-76)."""
    R = quat_to_rot(rots)
    s = scales * mod
    M = R * s.unsqueeze(-2)            # R @ diag(s)
    S = M @ M.transpose(-1, -2)
    return sym_to6(S)


def sym_to6(S):
    return torch.stack([S[..., 0, 0], S[..., 0, 1], S[..., 0, 2], S[..., 1, 1], S[..., 1, 2],
                        S[..., 2, 2]], -1)


def six_to_sym(c):
    return torch.stack([c[..., 0], c[..., 1], c[..., 2], c[..., 1], c[..., 3], c[..., 4],
                        c[..., 2], c[..., 4], c[..., 5]], -1).reshape(c.shape[:-1] + (3, 3))


def _homog(p, m):
    """[p, 1] @ m for the row-vector-convention matrices the reference stores."""
    return p @ m[:3, :] + m[3, :]


def preprocess(means3D, opacities, settings, shs=None, colors_precomp=None, scales=None,
               rotations=None, cov3D_precomp=None):
    """preprocessCUDA [U] in FP64.  Returns a dict of per-Gaussian tensors."""
    W, H = settings.image_width, settings.image_height
    P = means3D.shape[0]
    p = _d(means3D)
    vm, pm = _d(settings.viewmatrix), _d(settings.projmatrix)
    tanx = float(np.float32(settings.tanfovx))
    tany = float(np.float32(settings.tanfovy))
    fx = W / (2.0 * tanx)
    fy = H / (2.0 * tany)
    p_hom = _homog(p, pm)
    p_w = 1.0 / (p_hom[:, 3] + 1e-7)
    p_proj = p_hom[:, :3] * p_w[:, None]
    t = _homog(p, vm)[:, :3]
    in_front = t[:, 2] > 0.2

    if cov3D_precomp is not None and cov3D_precomp.numel() > 0:
        cov3 = _d(cov3D_precomp).reshape(P, 6)
    else:
        cov3 = cov3d_from_scale_rot(_d(scales), _d(rotations), settings.scale_modifier)

    cov2, J, Wv, txc, tyc = ewa_cov2d(t, cov3, vm, fx, fy, tanx, tany)
    a, b, c = cov2[:, 0], cov2[:, 1], cov2[:, 2]
    det = a * c - b * b
    ok = in_front & (det != 0)
    det_safe = torch.where(det == 0, torch.ones_like(det), det)
    conic = torch.stack([c / det_safe, -b / det_safe, a / det_safe], -1)
    mid = 0.5 * (a + c)
    disc = torch.sqrt(torch.clamp(mid * mid - det, min=0.1))
    lam = torch.maximum(mid + disc, mid - disc)
    r_real = 3.0 * torch.sqrt(lam)
    radius = torch.ceil(r_real)
    radius_tie = (r_real - torch.round(r_real)).abs() < 1e-5 * r_real.clamp(min=1)
    xy = torch.stack([((p_proj[:, 0] + 1.0) * W - 1.0) * 0.5,
                      ((p_proj[:, 1] + 1.0) * H - 1.0) * 0.5], -1)
    x0, y0, x1, y1, rect_tie = _rect_f64(xy, radius, W, H)
    tiles = (x1 - x0) * (y1 - y0)
    visible = ok & (tiles > 0)

    out = dict(xy=xy, conic=conic, depth=t[:, 2], cov3D=cov3, cov2D=cov2, p_view=t, p_proj=p_proj,
               p_hom=p_hom, opacity=_d(opacities).reshape(P), radius_f=radius,
               visible=visible, J=J, Wv=Wv, tx_clamped=txc, ty_clamped=tyc, fx=fx, fy=fy)
    out["radii"] = torch.where(visible, radius, torch.zeros_like(radius)).to(torch.int32)
    out["tiles_touched"] = torch.where(visible, tiles, torch.zeros_like(tiles))
    out["ambiguous"] = visible & (radius_tie | rect_tie)
    if colors_precomp is not None and colors_precomp.numel() > 0:
        out["rgb"] = _d(colors_precomp).reshape(P, -1)
        out["clamped"] = torch.zeros(P, 3, dtype=torch.bool)
    else:
        sh = _d(shs)
        dirs = p - _d(settings.campos).reshape(1, 3)
        dirs = dirs / dirs.norm(dim=1, keepdim=True)
        # sh is (P, M, 3): coefficient-major, RGB-minor [U]
        raw = sh_mod.eval_sh(settings.sh_degree, sh.transpose(1, 2), dirs) + 0.5
        out["clamped"] = raw < 0
        out["rgb"] = torch.clamp(raw, min=0.0)
        out["dirs_unnorm"] = p - _d(settings.campos).reshape(1, 3)
    return out


def ewa_cov2d(t, cov3, vm, fx, fy, tanx, tany):
    """computeCov2D [U]: returns (P,3) [a, b, c] incl. the +0.3 dilation, J, W."""
    limx, limy = 1.3 * tanx, 1.3 * tany
    tz = t[:, 2]
    txtz, tytz = t[:, 0] / tz, t[:, 1] / tz
    txc = (txtz < -limx) | (txtz > limx)
    tyc = (tytz < -limy) | (tytz > limy)
    tx = torch.clamp(txtz, -limx, limx) * tz
    ty = torch.clamp(tytz, -limy, limy) * tz
    P = t.shape[0]
    J = torch.zeros(P, 2, 3, dtype=F64)
    J[:, 0, 0] = fx / tz
    J[:, 0, 2] = -(fx * tx) / (tz * tz)
    J[:, 1, 1] = fy / tz
    J[:, 1, 2] = -(fy * ty) / (tz * tz)
    Wv = vm[:3, :3].T                      # p_view = Wv p + trans (column-vector form)
    M = J @ Wv
    S2 = M @ six_to_sym(cov3) @ M.transpose(1, 2)
    cov2 = torch.stack([S2[:, 0, 0] + 0.3, S2[:, 0, 1], S2[:, 1, 1] + 0.3], -1)
    return cov2, J, Wv, txc, tyc


# ----------------------------------------------------------------------------
# Binning (integer, bit-exact)
# ----------------------------------------------------------------------------

def bin_gaussians(xy32, radii, depth32, W, H):
    """duplicateWithKeys + SortPairs(bits [0, 32+msb(#tiles))) + identifyTileRanges [U]
    on the FP32 per-Gaussian outputs.  Returns (keys uint64 sorted, values int64,
    ranges (T,2) int64).  Ties in the key keep Gaussian-index order (stable sort)."""
    xy32 = np.asarray(xy32, dtype=np.float32)
    radii = np.asarray(radii).astype(np.int64)
    depth32 = np.ascontiguousarray(np.asarray(depth32, dtype=np.float32))
    gx, gy = tile_grid(W, H)
    T = gx * gy
    vis = np.nonzero(radii > 0)[0]
    x0, y0, x1, y1 = rect_f32(xy32[vis], radii[vis], W, H)
    nx, ny = x1 - x0, y1 - y0
    cnt = nx * ny
    K = int(cnt.sum())
    ranges = np.zeros((T, 2), dtype=np.int64)
    if K == 0:
        return np.zeros(0, np.uint64), np.zeros(0, np.int64), ranges
    rep = np.repeat(np.arange(vis.size), cnt)
    starts = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    local = np.arange(K) - np.repeat(starts, cnt)
    ty = y0[rep] + local // nx[rep]
    tx = x0[rep] + local % nx[rep]
    tile = (ty * gx + tx).astype(np.uint64)
    dbits = depth32.view(np.uint32)[vis][rep].astype(np.uint64)
    keys = (tile << np.uint64(32)) | dbits
    vals = vis[rep].astype(np.int64)
    order = np.argsort(keys, kind="stable")
    keys, vals = keys[order], vals[order]
    tiles_sorted = (keys >> np.uint64(32)).astype(np.int64)
    t_ids = np.arange(T)
    ranges[:, 0] = np.searchsorted(tiles_sorted, t_ids, side="left")
    ranges[:, 1] = np.searchsorted(tiles_sorted, t_ids, side="right")
    empty = ranges[:, 0] == ranges[:, 1]
    ranges[empty] = 0
    return keys, vals, ranges


def higher_msb(n):
    """Number of bits the tile id occupies: smallest b with n >> b == 0 [U]."""
    b = 0
    while (n >> b) != 0:
        b += 1
    return b


# ----------------------------------------------------------------------------
# Blend forward / backward (vectorised per tile, FP64)
# ----------------------------------------------------------------------------

ALPHA_MIN = 1.0 / 255.0
T_STOP = 1e-4


def _tile_pixels(tile, W, H):
    gx, _ = tile_grid(W, H)
    tx, ty = tile % gx, tile // gx
    xs = np.arange(tx * BLOCK_X, min(tx * BLOCK_X + BLOCK_X, W))
    ys = np.arange(ty * BLOCK_Y, min(ty * BLOCK_Y + BLOCK_Y, H))
    px = np.repeat(xs[None, :], ys.size, 0).reshape(-1)
    py = np.repeat(ys[:, None], xs.size, 1).reshape(-1)
    return px, py


def _tile_eval(ids, px, py, xy, conic, opac):
    dx = xy[ids, 0][None, :] - torch.from_numpy(px).to(F64)[:, None]
    dy = xy[ids, 1][None, :] - torch.from_numpy(py).to(F64)[:, None]
    a, b, c = conic[ids, 0][None, :], conic[ids, 1][None, :], conic[ids, 2][None, :]
    power = -0.5 * (a * dx * dx + c * dy * dy) - b * dx * dy
    G = torch.exp(power)
    raw = opac[ids][None, :] * G
    alpha = torch.clamp(raw, max=0.99)
    valid = (power <= 0) & (alpha >= ALPHA_MIN)
    amb = (power.abs() < 1e-7) | ((alpha / ALPHA_MIN - 1).abs() < 1e-5)
    return dx, dy, power, G, alpha, valid, amb


def _tile_walk(valid, alpha):
    """Front-to-back transmittance walk with the stop rule; returns
    (include mask, T_before, T_final, n_contrib, stop_idx, test_T)."""
    one_m = torch.where(valid, 1 - alpha, torch.ones_like(alpha))
    Tinc = torch.cumprod(one_m, dim=1)
    Tbef = torch.cat([torch.ones_like(Tinc[:, :1]), Tinc[:, :-1]], 1)
    stop = valid & (Tinc < T_STOP)
    L = valid.shape[1]
    has = stop.any(1)
    sidx = torch.where(has, stop.to(torch.int8).argmax(1), torch.full_like(has, L, dtype=torch.int64))
    ar = torch.arange(L)[None, :]
    include = valid & (ar < sidx[:, None])
    Tfin = torch.where(has, Tbef.gather(1, sidx.clamp(max=L - 1)[:, None])[:, 0], Tinc[:, -1])
    last = torch.where(include, ar + 1, torch.zeros_like(ar)).amax(1)
    return include, Tbef, Tfin, last, sidx, Tinc


def _tiles_to_visit(ranges, tiles):
    nonempty = np.nonzero(ranges[:, 1] > ranges[:, 0])[0]
    if tiles is None:
        return nonempty
    return np.intersect1d(nonempty, np.asarray(tiles))


def render_forward(ranges, vals, xy, conic, opacity, colors, bg, W, H, tiles=None):
    """renderCUDA forward [U] in FP64 (inputs may be FP32 device outputs).
    Returns (out_color (C,H,W), final_T (H,W), n_contrib (H,W), ambiguous (H,W)).
    ``tiles`` restricts the work to a subset of tile ids (for timing samples)."""
    xy, conic, opac, col = _d(xy), _d(conic), _d(opacity).reshape(-1), _d(colors)
    bg = _d(bg).reshape(-1)
    C = col.shape[1]
    out = bg[:, None, None].repeat(1, H, W).clone()
    final_T = torch.ones(H, W, dtype=F64)
    n_contrib = torch.zeros(H, W, dtype=torch.int64)
    ambig = torch.zeros(H, W, dtype=torch.bool)
    vals_t = torch.as_tensor(np.asarray(vals), dtype=torch.int64)
    for tile in _tiles_to_visit(ranges, tiles):
        s, e = int(ranges[tile, 0]), int(ranges[tile, 1])
        ids = vals_t[s:e]
        px, py = _tile_pixels(int(tile), W, H)
        _, _, power, G, alpha, valid, amb = _tile_eval(ids, px, py, xy, conic, opac)
        include, Tbef, Tfin, last, sidx, Tinc = _tile_walk(valid, alpha)
        w = torch.where(include, alpha * Tbef, torch.zeros_like(alpha))
        Cpix = w @ col[ids]                                    # (npx, C)
        L = ids.numel()
        ar = torch.arange(L)[None, :]
        upto = ar <= sidx[:, None]
        stop_amb = valid & upto & ((Tinc / T_STOP - 1).abs() < 1e-4)
        amb_pix = ((amb | stop_amb) & upto).any(1)
        pyt, pxt = torch.from_numpy(py), torch.from_numpy(px)
        out[:, pyt, pxt] = (Cpix + Tfin[:, None] * bg[None, :]).T
        final_T[pyt, pxt] = Tfin
        n_contrib[pyt, pxt] = last
        ambig[pyt, pxt] = amb_pix
    return out, final_T, n_contrib, ambig


def render_backward(ranges, vals, xy, conic, opacity, colors, bg, W, H, dL_dpix, tiles=None):
    """renderCUDA backward [U] in FP64, restated in suffix-sum form instead of the
    device's reverse traversal.  Returns per-Gaussian (dL_dmean2D (P,3) NDC units,
    dL_dconic (P,4) with the reference's half-valued .y, dL_dopacity (P,1),
    dL_dcolors (P,C))."""
    xy, conic, opac, col = _d(xy), _d(conic), _d(opacity).reshape(-1), _d(colors)
    bg = _d(bg).reshape(-1)
    dpix = _d(dL_dpix)
    P, C = col.shape
    dmean2D = torch.zeros(P, 3, dtype=F64)
    dconic = torch.zeros(P, 4, dtype=F64)
    dopac = torch.zeros(P, dtype=F64)
    dcol = torch.zeros(P, C, dtype=F64)
    vals_t = torch.as_tensor(np.asarray(vals), dtype=torch.int64)
    for tile in _tiles_to_visit(ranges, tiles):
        s, e = int(ranges[tile, 0]), int(ranges[tile, 1])
        ids = vals_t[s:e]
        px, py = _tile_pixels(int(tile), W, H)
        dx, dy, power, G, alpha, valid, _ = _tile_eval(ids, px, py, xy, conic, opac)
        include, Tbef, Tfin, last, sidx, _ = _tile_walk(valid, alpha)
        inc = include.to(F64)
        g = dpix[:, torch.from_numpy(py), torch.from_numpy(px)].T      # (npx, C)
        w = alpha * Tbef * inc                                          # (npx, L)
        cg = col[ids]                                                   # (L, C)
        # colour gradient: sum_px alpha T dpix
        dcol.index_add_(0, ids, w.T @ g)
        # accum_rec_i = sum_{j>i} w_j c_j / (T_i (1 - alpha_i))
        wc = w.unsqueeze(-1) * cg.unsqueeze(0)                          # (npx, L, C)
        behind = torch.flip(torch.cumsum(torch.flip(wc, [1]), 1), [1]) - wc
        Taft = Tbef * (1 - alpha)
        accum = behind / torch.where(inc > 0, Taft, torch.ones_like(Taft)).unsqueeze(-1)
        dalpha = Tbef * ((cg.unsqueeze(0) - accum) * g.unsqueeze(1)).sum(-1)
        bgdot = (g * bg[None, :]).sum(-1)
        dalpha = dalpha - Tfin[:, None] / (1 - alpha) * bgdot[:, None]
        dalpha = dalpha * inc
        dG = opac[ids][None, :] * dalpha
        a, b, c = conic[ids, 0][None, :], conic[ids, 1][None, :], conic[ids, 2][None, :]
        gdx, gdy = G * dx, G * dy
        dGdx = -gdx * a - gdy * b
        dGdy = -gdy * c - gdx * b
        dmean2D[:, 0].index_add_(0, ids, (dG * dGdx * (0.5 * W)).sum(0))
        dmean2D[:, 1].index_add_(0, ids, (dG * dGdy * (0.5 * H)).sum(0))
        dconic[:, 0].index_add_(0, ids, (-0.5 * gdx * dx * dG).sum(0))
        dconic[:, 1].index_add_(0, ids, (-0.5 * gdx * dy * dG).sum(0))
        dconic[:, 3].index_add_(0, ids, (-0.5 * gdy * dy * dG).sum(0))
        dopac.index_add_(0, ids, (G * dalpha).sum(0))
    return dmean2D, dconic, dopac.reshape(P, 1), dcol


# ----------------------------------------------------------------------------
# Geometry backward (cov2D + projection + SH + cov3D), FP64
# ----------------------------------------------------------------------------

def geometry_backward(pre, settings, means3D, dmean2D, dconic, dcolors, shs=None,
                      scales=None, rotations=None, cov3D_precomp=None):
    """computeCov2DCUDA + preprocessCUDA backward [U] for visible Gaussians.
    Returns dict of gradients (FP64)."""
    P = means3D.shape[0]
    vis = pre["visible"]
    p = _d(means3D)
    cov2 = pre["cov2D"]
    a, b, c = cov2[:, 0], cov2[:, 1], cov2[:, 2]
    det = a * c - b * b
    inv2 = 1.0 / (det * det + 1e-7)
    # matrix gradient of the conic with independent entries
    Gq = torch.stack([dconic[:, 0], dconic[:, 1], dconic[:, 1], dconic[:, 3]], -1).reshape(P, 2, 2)
    adj = torch.stack([c, -b, -b, a], -1).reshape(P, 2, 2)
    G2 = -(adj @ Gq @ adj) * inv2[:, None, None]          # dL/dSigma2 (independent entries)
    J, Wv = pre["J"], pre["Wv"]
    M = J @ Wv
    G3 = M.transpose(1, 2) @ G2 @ M                        # dL/dSigma3 (independent entries)
    dcov6 = torch.stack([G3[:, 0, 0], 2 * G3[:, 0, 1], 2 * G3[:, 0, 2], G3[:, 1, 1],
                         2 * G3[:, 1, 2], G3[:, 2, 2]], -1)
    S3 = six_to_sym(pre["cov3D"])
    dM = 2 * G2 @ M @ S3
    dJ = dM @ Wv.T
    t = pre["p_view"]
    tz = t[:, 2]
    fx, fy = pre["fx"], pre["fy"]
    limx, limy = 1.3 * float(np.float32(settings.tanfovx)), 1.3 * float(np.float32(settings.tanfovy))
    txu = torch.clamp(t[:, 0] / tz, -limx, limx) * tz
    tyu = torch.clamp(t[:, 1] / tz, -limy, limy) * tz
    dtx = torch.where(pre["tx_clamped"], torch.zeros_like(tz), -fx / tz ** 2 * dJ[:, 0, 2])
    dty = torch.where(pre["ty_clamped"], torch.zeros_like(tz), -fy / tz ** 2 * dJ[:, 1, 2])
    dtz = (-fx / tz ** 2 * dJ[:, 0, 0] - fy / tz ** 2 * dJ[:, 1, 1]
           + 2 * fx * txu / tz ** 3 * dJ[:, 0, 2] + 2 * fy * tyu / tz ** 3 * dJ[:, 1, 2])
    dmean = torch.stack([dtx, dty, dtz], -1) @ Wv           # W^T dt
    # projection (means2D is in NDC)
    pm = _d(settings.projmatrix)
    ph = pre["p_hom"]
    mw = 1.0 / (ph[:, 3] + 1e-7)
    for k in range(2):
        dndc = dmean2D[:, k]
        dmean = dmean + dndc[:, None] * (pm[:3, k][None, :] * mw[:, None]
                                          - pm[:3, 3][None, :] * (ph[:, k] * mw * mw)[:, None])
    out = {}
    # SH
    if shs is not None and shs.numel() > 0:
        sh = _d(shs)
        v = pre["dirs_unnorm"].clone().requires_grad_(True)
        shc = sh.clone().requires_grad_(True)
        dirs = v / v.norm(dim=1, keepdim=True)
        rgb = sh_mod.eval_sh(settings.sh_degree, shc.transpose(1, 2), dirs)
        gcol = torch.where(pre["clamped"], torch.zeros_like(dcolors), dcolors)
        gcol = torch.where(vis[:, None], gcol, torch.zeros_like(gcol))
        dsh, dv = torch.autograd.grad(rgb, (shc, v), gcol)
        dmean = dmean + dv
        out["dsh"] = dsh
    # cov3D
    if cov3D_precomp is not None and cov3D_precomp.numel() > 0:
        out["dcov3D"] = torch.where(vis[:, None], dcov6, torch.zeros_like(dcov6))
    else:
        s = _d(scales) * settings.scale_modifier
        q = _d(rotations)
        s.requires_grad_(True)
        q.requires_grad_(True)
        R = quat_to_rot(q)
        Mx = R * s.unsqueeze(-2)
        c6 = sym_to6(Mx @ Mx.transpose(-1, -2))
        gc6 = torch.where(vis[:, None], dcov6, torch.zeros_like(dcov6))
        ds, dq = torch.autograd.grad(c6, (s, q), gc6)
        out["dscales"], out["drotations"] = ds, dq
    out["dmeans3D"] = torch.where(vis[:, None], dmean, torch.zeros_like(dmean))
    return out


# ----------------------------------------------------------------------------
# Whole-pipeline convenience wrappers (same argument order as _C.rasterize_gaussians)
# ----------------------------------------------------------------------------

def rasterize(means3D, opacities, settings, shs=None, colors_precomp=None, scales=None,
              rotations=None, cov3D_precomp=None):
    """Forward: returns (color (3,H,W) f64, radii int32, state)."""
    pre = preprocess(means3D, opacities, settings, shs, colors_precomp, scales, rotations,
                     cov3D_precomp)
    W, H = settings.image_width, settings.image_height
    xy32 = pre["xy"].to(torch.float32).numpy()
    d32 = pre["depth"].to(torch.float32).numpy()
    keys, vals, ranges = bin_gaussians(xy32, pre["radii"].numpy(), d32, W, H)
    color, final_T, n_contrib, amb = render_forward(ranges, vals, pre["xy"], pre["conic"],
                                                    pre["opacity"], pre["rgb"], settings.bg, W, H)
    state = dict(pre=pre, keys=keys, vals=vals, ranges=ranges, final_T=final_T,
                 n_contrib=n_contrib, ambiguous=amb)
    return color, pre["radii"], state


def rasterize_backward(state, settings, means3D, dL_dcolor, shs=None, colors_precomp=None,
                       scales=None, rotations=None, cov3D_precomp=None):
    """Backward: returns the gradient tuple in _RasterizeGaussians.backward order:
    (means3D, means2D, sh, colors_precomp, opacities, scales, rotations, cov3D)."""
    pre = state["pre"]
    W, H = settings.image_width, settings.image_height
    dm2, dcon, dop, dcol = render_backward(state["ranges"], state["vals"], pre["xy"], pre["conic"],
                                           pre["opacity"], pre["rgb"], settings.bg, W, H, dL_dcolor)
    g = geometry_backward(pre, settings, means3D, dm2, dcon, dcol, shs, scales, rotations,
                          cov3D_precomp)
    return (g["dmeans3D"], dm2, g.get("dsh"), dcol, dop, g.get("dscales"), g.get("drotations"),
            g.get("dcov3D"), dcon)
