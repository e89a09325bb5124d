"""Parity of the HIP rasterizer against the FP64 oracle at full workload
sizes, on sampled tiles.

At 1M-3M Gaussians and 800x800-1920x1080 the oracle cannot blend every tile
in the seconds a test may take, so the checks are:

* lists: bit-exact against the oracle's integer binning of the device's own
  FP32 splat positions, radii and depths — every tile of a view where the
  oracle's full binning fits (``OR.bin_gaussians``), else every sampled tile
  (``sampled_lists``) plus the exact instance counts of the whole view;
* preprocess: every Gaussian's radius (threshold ties excepted), pixel
  position, conic, colour and depth against the FP64 preprocess;
* blend forward on sampled tiles: the oracle's FP64 blend run on the device's
  FP32 splats and lists ("staged"), max |diff| < 1e-4 on every pixel the
  oracle does not flag as a threshold tie; and the oracle's own FP64 chain
  (its own lists) on the sampled tiles whose list equals the device's;
* backward: dL/dimage random on the sampled tiles' untied pixels and zero
  elsewhere, so only those tiles contribute; the oracle's blend backward on
  the same tiles feeds its FP64 geometry backward, and every gradient is
  compared norm-relative (1e-3).
"""
import ctypes

import numpy as np
import torch

from dmgs_amd import _native as N
from oracle import raster as OR

IMG_TOL = 1e-4
GRAD_TOL = 1e-3
F64 = torch.float64


# --------------------------------------------------------------------------
# tiles and masks
# --------------------------------------------------------------------------
def pick_tiles(lengths, n, seed):
    """n tile ids among those with a non-empty list (seeded), sorted."""
    nz = np.nonzero(np.asarray(lengths) > 0)[0]
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(nz, size=min(n, nz.size), replace=False))


def tile_pixel_mask(tiles, W, H):
    """(H, W) bool: pixels of the given 16x16 tiles."""
    gx, _ = OR.tile_grid(W, H)
    m = np.zeros((H, W), dtype=bool)
    for t in np.asarray(tiles):
        tx, ty = int(t) % gx, int(t) // gx
        m[ty * 16:ty * 16 + 16, tx * 16:tx * 16 + 16] = True
    return m


def masked_dL(tiles, amb, W, H, seed, scale=1.0):
    """(3, H, W) float32: N(0, scale^2) on the given tiles' untied pixels, 0 elsewhere."""
    g = torch.Generator().manual_seed(seed)
    dL = torch.randn(3, H, W, generator=g) * scale
    keep = torch.from_numpy(tile_pixel_mask(tiles, W, H) & ~np.asarray(amb))
    return dL * keep


def sampled_lists(xy32, radii, depth32, W, H, tiles):
    """The reference per-tile lists of only the given tiles: every Gaussian
    whose FP32 tile rectangle covers the tile, ordered by (depth bits,
    index).  Returns (ranges (T, 2), vals) in the layout of
    ``OR.bin_gaussians`` (other tiles empty)."""
    gx, gy = OR.tile_grid(W, H)
    radii = np.asarray(radii)
    vis = np.nonzero(radii > 0)[0]
    x0, y0, x1, y1 = OR.rect_f32(np.asarray(xy32)[vis], radii[vis], W, H)
    dbits = np.ascontiguousarray(np.asarray(depth32, dtype=np.float32)).view(np.uint32)[vis]
    ranges = np.zeros((gx * gy, 2), dtype=np.int64)
    vals = []
    pos = 0
    for t in np.sort(np.asarray(tiles)):
        tx, ty = int(t) % gx, int(t) // gx
        sel = np.nonzero((x0 <= tx) & (tx < x1) & (y0 <= ty) & (ty < y1))[0]
        order = np.lexsort((vis[sel], dbits[sel]))          # depth bits, then index (stable sort order)
        ids = vis[sel][order]
        ranges[t] = (pos, pos + ids.size) if ids.size else (0, 0)
        vals.append(ids)
        pos += ids.size
    return ranges, (np.concatenate(vals) if vals else np.zeros(0, np.int64)).astype(np.int64)


def tile_rect_counts(xy32, radii, W, H):
    """(tile instances, cell instances) of the reference tile rectangles."""
    radii = np.asarray(radii)
    vis = radii > 0
    x0, y0, x1, y1 = OR.rect_f32(np.asarray(xy32)[vis], radii[vis], W, H)
    k_tile = int(((x1 - x0) * (y1 - y0)).sum())
    live = (x1 > x0) & (y1 > y0)
    x0, y0, x1, y1 = x0[live], y0[live], x1[live], y1[live]
    k_cell = int((((x1 - 1) // 4 - x0 // 4 + 1) * ((y1 - 1) // 4 - y0 // 4 + 1)).sum())
    return k_tile, k_cell


# --------------------------------------------------------------------------
# device state
# --------------------------------------------------------------------------
def single_view_splats(state, W, H):
    """Per-Gaussian FP32 outputs and the lists of a single-view forward."""
    from tests.helpers import workspace_views
    return workspace_views(state, W, H)


def batched_view_splats(r, v):
    """Per-Gaussian FP32 splat fields of view ``v`` of a batched forward."""
    o = r._offsets[5]
    P = r.P
    rec = r.state[o + 48 * v * P:o + 48 * (v + 1) * P].view(torch.float32).view(P, 12).cpu().numpy()
    radii = r.radii()[v].cpu().numpy()
    return dict(xy=rec[:, 0:2].copy(), conic=rec[:, 2:5].copy(), opacity=rec[:, 5].copy(), rgb=rec[:, 6:9].copy(),
                depth=rec[:, 9].copy(), radii=radii, radius_bits=rec[:, 10].copy().view(np.int32))


def batched_lists(r):
    """All cell-list entries of the last batched forward: (keys, ids,
    ranges, global cell of every entry).  Checks that every global cell's
    range holds exactly the entries keyed to that cell, that the ranges are
    disjoint and cover the count, and (view-major lists) that each key's view
    bits name the entry's view.  In view-major mode the sort orders by the
    cell within the view first, so a view's cells interleave with the other
    views' cells: only each (view, cell) run is contiguous."""
    ko, vo = N.c_size_t(0), N.c_size_t(0)
    N.check(r.lib.dmgs_batched_list_offsets(ctypes.byref(r._desc()), ctypes.byref(ko), ctypes.byref(vo)), "offsets")
    K = min(int(r.instance_count().item()) & 0xFFFFFFFF, r.K_cap)
    keys = r.lists[ko.value:ko.value + 4 * K].view(torch.int32).cpu().numpy().view(np.uint32)
    ids = r.lists[vo.value:vo.value + 4 * K].view(torch.int32).cpu().numpy().view(np.uint32)
    o = r._offsets[2]
    rng = r.state[o:o + 8 * r.total_cells].view(torch.int32).view(-1, 2).cpu().numpy().astype(np.int64)
    k = keys.astype(np.int64)
    if r.cell_mode == 0:
        glob = k & 0xFFFF
    else:
        view = (k >> 12) & 0xF
        assert np.array_equal(view, ids.astype(np.int64) // r.P), "key view bits and entry view differ"
        base = np.array([r.views[v].cell_base for v in range(r.B)], dtype=np.int64)
        glob = base[view] + (k & 0xFFF)
    lens = rng[:, 1] - rng[:, 0]
    assert (lens >= 0).all() and int(lens.sum()) == K, "cell ranges do not cover the instance count"
    nz = np.nonzero(lens > 0)[0]
    order = nz[np.argsort(rng[nz, 0])]
    assert np.all(rng[order[:-1], 1] <= rng[order[1:], 0]), "cell ranges overlap"
    idx = _range_positions(rng[nz, 0], lens[nz])
    assert np.array_equal(glob[idx], np.repeat(nz, lens[nz])), "a cell range holds another cell's entries"
    return dict(keys=keys, ids=ids, ranges=rng)


def _range_positions(starts, lens):
    """Concatenation of arange(s, s + n) over (starts, lens)."""
    total = int(lens.sum())
    if total == 0:
        return np.zeros(0, np.int64)
    offs = np.repeat(starts - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens)
    return offs + np.arange(total)


def batched_cell_entries(r, v, lists):
    """(cell ids within the view, tile masks, Gaussian ids, list positions)
    of view ``v``'s cell lists, cell by cell, each in list (depth) order."""
    d = r.views[v]
    ncell = ((d.grid_x + 3) // 4) * ((d.grid_y + 3) // 4)
    rng = lists["ranges"][d.cell_base:d.cell_base + ncell]
    lens = rng[:, 1] - rng[:, 0]
    pos = _range_positions(rng[:, 0], lens)
    cell = np.repeat(np.arange(ncell), lens)
    keys = lists["keys"][pos].astype(np.int64)
    ids = lists["ids"][pos].astype(np.int64)
    assert np.all(ids // r.P == v), f"view {v}: an entry of another view in its cells"
    return cell, (keys >> 16) & 0xFFFF, ids - v * r.P, pos


def tile_lengths(cell, mask, d):
    """Per-tile list lengths of a view from its cell lists (set mask bits)."""
    gx, gy = d.grid_x, d.grid_y
    cgx = (gx + 3) // 4
    out = np.zeros(gx * gy, dtype=np.int64)
    cx, cy = cell % cgx, cell // cgx
    for b in range(16):
        sel = (mask & (1 << b)) != 0
        tx = cx[sel] * 4 + (b & 3)
        ty = cy[sel] * 4 + (b >> 2)
        assert np.all((tx < gx) & (ty < gy)), "a tile-mask bit outside the view's grid"
        out += np.bincount(ty * gx + tx, minlength=gx * gy)
    return out


def expand_cell_lists(cell, mask, ids, d, tiles=None):
    """Per-tile lists from cell lists: (ranges (T, 2), vals) in the layout of
    ``OR.bin_gaussians``; restricted to ``tiles`` if given."""
    gx, gy = d.grid_x, d.grid_y
    cgx = (gx + 3) // 4
    T = gx * gy
    if tiles is not None:  # only the entries of the sampled tiles' cells
        t = np.asarray(tiles)
        keep = np.isin(cell, np.unique((t // gx) // 4 * cgx + (t % gx) // 4))
        cell, mask, ids = cell[keep], mask[keep], ids[keep]
    cx, cy = cell % cgx, cell // cgx
    tl, pl = [], []
    for b in range(16):
        sel = np.nonzero(mask & (1 << b))[0]
        tx = cx[sel] * 4 + (b & 3)
        ty = cy[sel] * 4 + (b >> 2)
        assert np.all((tx < gx) & (ty < gy)), "a tile-mask bit outside the view's grid"
        tl.append(ty * gx + tx)
        pl.append(sel)
    tile = np.concatenate(tl)
    pos = np.concatenate(pl)
    if tiles is not None:
        keep = np.isin(tile, np.asarray(tiles))
        tile, pos = tile[keep], pos[keep]
    order = np.lexsort((pos, tile))
    tile, pos = tile[order], pos[order]
    vals = ids[pos]
    ranges = np.zeros((T, 2), dtype=np.int64)
    t_ids = np.arange(T)
    ranges[:, 0] = np.searchsorted(tile, t_ids, side="left")
    ranges[:, 1] = np.searchsorted(tile, t_ids, side="right")
    ranges[ranges[:, 0] == ranges[:, 1]] = 0
    return ranges, vals


def same_lists(ra, va, rb, vb, tiles=None):
    """Tiles (of ``tiles`` or all) whose lists are equal, as a bool array."""
    idx = np.arange(ra.shape[0]) if tiles is None else np.asarray(tiles)
    out = np.zeros(idx.size, dtype=bool)
    for k, t in enumerate(idx):
        a = va[ra[t, 0]:ra[t, 1]]
        b = vb[rb[t, 0]:rb[t, 1]]
        out[k] = a.size == b.size and np.array_equal(a, b)
    return out


# --------------------------------------------------------------------------
# comparisons
# --------------------------------------------------------------------------
def check_preprocess(sp, pre, tag=""):
    """Device FP32 preprocess outputs against the FP64 oracle (all Gaussians)."""
    r_gpu = np.asarray(sp["radii"])
    r_ora = pre["radii"].numpy()
    tied = pre["ambiguous"].numpy()
    mism = (r_gpu != r_ora) & ~tied
    assert mism.sum() == 0, f"{tag}: radii differ on {int(mism.sum())} untied Gaussians"
    vis = (r_gpu > 0) & (r_ora > 0)
    assert vis.sum() > 0
    assert np.abs(sp["xy"][vis] - pre["xy"].numpy()[vis]).max() < 2e-3, tag
    # the conic is the inverse of cov2D: FP32 inversion loses precision in
    # proportion to cov2D's condition number, so the bound scales with it
    a, b, c = pre["cov2D"].numpy()[vis].T
    mid, half = 0.5 * (a + c), np.sqrt(0.25 * (a - c) ** 2 + b * b)
    cond = (mid + half) / (mid - half)
    qo = pre["conic"].numpy()[vis]
    dq = sp["conic"][vis] - qo
    fro = lambda q: np.sqrt(q[:, 0] ** 2 + 2 * q[:, 1] ** 2 + q[:, 2] ** 2)
    err = fro(dq) / fro(qo)
    bad = err > 1e-5 + 2e-7 * cond
    assert not bad.any(), f"{tag}: conic error {err[bad].max():.3g} on {int(bad.sum())} Gaussians"
    assert np.abs(sp["rgb"][vis] - pre["rgb"].numpy()[vis]).max() < 1e-5, tag
    do = pre["depth"].numpy()[vis]
    assert (np.abs(sp["depth"][vis] - do) <= 2e-6 + 1e-6 * do).all(), tag
    return int(vis.sum()), int(tied.sum())


def staged_forward(sp, ranges, vals, bg, W, H, tiles, image_gpu, tag=""):
    """Oracle FP64 blend of the sampled tiles on the device's FP32 splats and
    lists; the device image must match within 1e-4 on untied pixels.
    Returns the (H, W) tie mask and the max error."""
    out, fT, nc, amb = OR.render_forward(ranges, vals, torch.from_numpy(sp["xy"]), torch.from_numpy(sp["conic"]),
                                         torch.from_numpy(sp["opacity"]), torch.from_numpy(sp["rgb"]),
                                         torch.as_tensor(bg), W, H, tiles=tiles)
    amb = amb.numpy()
    m = tile_pixel_mask(tiles, W, H)
    ok = m & ~amb
    img = image_gpu.detach().double().cpu().numpy()
    err = np.abs(img - out.numpy()).max(0)[ok].max()
    print(f"  {tag} staged forward on {len(tiles)} tiles: max |diff| {err:.3e}, tie pixels "
          f"{int((m & amb).sum())} of {int(m.sum())}", flush=True)
    assert err < IMG_TOL, f"{tag}: staged image error {err}"
    return amb, err


def end_to_end_forward(pre, bg, W, H, tiles, ranges_gpu, vals_gpu, image_gpu, tag="", min_frac=0.2):
    """The oracle's own chain (FP64 preprocess, its lists from the FP32
    rounding of its positions and depths) on the sampled tiles whose list
    equals the device's: max |diff| < 1e-4 on untied pixels."""
    xy32 = pre["xy"].to(torch.float32).numpy()
    d32 = pre["depth"].to(torch.float32).numpy()
    ro, vo = sampled_lists(xy32, pre["radii"].numpy(), d32, W, H, tiles)
    agree = same_lists(ro, vo, ranges_gpu, vals_gpu, tiles)
    use = np.asarray(tiles)[agree]
    print(f"  {tag} end-to-end: {int(agree.sum())} of {len(tiles)} sampled tiles have the oracle's own list",
          flush=True)
    assert agree.mean() >= min_frac, f"{tag}: only {agree.mean():.2f} of the sampled lists agree"
    out, _, _, amb = OR.render_forward(ro, vo, pre["xy"], pre["conic"], pre["opacity"], pre["rgb"],
                                       torch.as_tensor(bg), W, H, tiles=use)
    ok = tile_pixel_mask(use, W, H) & ~amb.numpy()
    img = image_gpu.detach().double().cpu().numpy()
    err = np.abs(img - out.numpy()).max(0)[ok].max()
    print(f"  {tag} end-to-end max |diff| {err:.3e} on {int(ok.sum())} pixels", flush=True)
    assert err < IMG_TOL, f"{tag}: end-to-end image error {err}"
    return err


def staged_backward(sp, ranges, vals, bg, W, H, dL, tiles):
    """Oracle blend backward of the sampled tiles on the device's splats:
    per-Gaussian (dmean2D (P,3) NDC, dconic (P,4), dopacity (P,1), dcolour (P,3))."""
    return OR.render_backward(ranges, vals, torch.from_numpy(sp["xy"]), torch.from_numpy(sp["conic"]),
                              torch.from_numpy(sp["opacity"]), torch.from_numpy(sp["rgb"]), torch.as_tensor(bg), W,
                              H, dL.double(), tiles=tiles)


def oracle_geometry(pre, settings, means3D, dm2, dcon, dcol, radii_gpu, **kw):
    """FP64 geometry backward with the device's visibility (a Gaussian whose
    radius sits on a rounding tie is visible on one side only)."""
    p = dict(pre)
    p["visible"] = torch.from_numpy(np.asarray(radii_gpu) > 0)
    return OR.geometry_backward(p, settings, means3D, dm2, dcon, dcol, **kw)


def rel(got, ref):
    got = torch.as_tensor(got).double().cpu().reshape(-1)
    ref = torch.as_tensor(ref).double().cpu().reshape(-1)
    den = ref.norm().item()
    return (got - ref).norm().item() / den if den > 0 else (got - ref).norm().item()


def assert_grads(pairs, tag=""):
    for name, got, ref in pairs:
        e = rel(got, ref)
        print(f"  {tag} {name}: rel err {e:.2e}", flush=True)
        assert torch.as_tensor(ref).abs().max() > 0, f"{tag} {name}: reference gradient is zero"
        assert e < GRAD_TOL, f"{tag} {name}: gradient rel err {e}"


def fp64_activations(raw):
    """FP64 model activations (scene/gaussian_model.py:95-118) of raw CPU
    parameters, as autograd leaves: returns (leaves, activated)."""
    leaves = {k: v.detach().cpu().double().requires_grad_(True) for k, v in raw.items()}
    act = dict(means3D=leaves["_xyz"], opac=torch.sigmoid(leaves["_opacity"]),
               shs=torch.cat((leaves["_features_dc"], leaves["_features_rest"]), 1),
               scales=torch.exp(leaves["_scaling"]), rots=torch.nn.functional.normalize(leaves["_rotation"]))
    return leaves, act


RAW = ("_xyz", "_scaling", "_rotation", "_opacity", "_features_dc", "_features_rest")


def raw_params(model):
    return {k: getattr(model, k).detach().cpu() for k in RAW}


def batched_counts(r):
    """(cell instances, tile instances) of the last batched forward, from
    the device lists (tile instances = set bits of the tile masks)."""
    ko, vo = N.c_size_t(0), N.c_size_t(0)
    N.check(r.lib.dmgs_batched_list_offsets(ctypes.byref(r._desc()), ctypes.byref(ko), ctypes.byref(vo)), "offsets")
    K = int(r.instance_count().item()) & 0xFFFFFFFF
    keys = r.lists[ko.value:ko.value + 4 * K].view(torch.int32)
    mask = (keys >> 16) & 0xFFFF
    bits = torch.zeros_like(mask)
    for b in range(16):
        bits += (mask >> b) & 1
    return K, int(bits.sum().item())


def views_of(r):
    return [r.views[v] for v in range(r.B)]

