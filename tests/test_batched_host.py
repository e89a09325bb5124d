"""Host-side logic of the batched pipeline (CPU only, no compute call):
view preparation, the tile -> view map, workspace sizing, argument checks,
and the component-major parameter packing."""
import ctypes
import math
import os

import pytest
import torch

from dmgs_amd import _native as N
from dmgs_amd import batched as BT
from dmgs_amd import cameras as C
from dmgs_amd import scenes as S


def _lib():
    if not os.path.exists(N.library_path()):
        pytest.skip("library not built (run __graft_entry__.build())")
    return N.load()


def _views(sizes):
    cams = [C.camera_at(p, w, h, device="cpu") for p, (w, h) in zip(C.sphere_positions(len(sizes), 4.0, 0), sizes)]
    return cams, (N.DmgsView * len(cams))(*[BT._view_record(c, torch.ones(3)) for c in cams])


def test_prepare_views_and_tile_map():
    lib = _lib()
    sizes = [(800, 800), (1920, 1080), (33, 17), (16, 16)]
    cams, views = _views(sizes)
    tt, tc, mc, tp = N.c_int32(0), N.c_int32(0), N.c_int32(0), N.c_int64(0)
    N.check(lib.dmgs_batched_prepare_views(views, len(sizes), ctypes.byref(tt), ctypes.byref(tc), ctypes.byref(mc),
                                           ctypes.byref(tp)), "prepare")
    tiles = cells = pixels = max_cells = 0
    for v, (w, h) in enumerate(sizes):
        d = views[v]
        assert (d.grid_x, d.grid_y) == ((w + 15) // 16, (h + 15) // 16)
        assert d.tile_base == tiles and d.pixel_base == pixels and d.cell_base == cells
        assert d.focal_x == pytest.approx(w / (2 * math.tan(cams[v].FoVx / 2)), rel=1e-6)
        assert d.focal_y == pytest.approx(h / (2 * math.tan(cams[v].FoVy / 2)), rel=1e-6)
        tiles += d.grid_x * d.grid_y
        vc = ((w + 63) // 64) * ((h + 63) // 64)  # cells of 4x4 tiles = 64x64 pixels
        cells += vc
        max_cells = max(max_cells, vc)
        pixels += w * h
    assert (tt.value, tc.value, mc.value, tp.value) == (tiles, cells, max_cells, pixels)
    tv = torch.full((tiles,), -1, dtype=torch.int32)
    N.check(lib.dmgs_batched_tile_view(views, len(sizes), ctypes.c_void_p(tv.data_ptr()), tiles), "tile_view")
    expect = torch.cat([torch.full((views[v].grid_x * views[v].grid_y,), v, dtype=torch.int32)
                        for v in range(len(sizes))])
    assert torch.equal(tv, expect)
    # a tile map shorter than the views' tiles is rejected
    assert lib.dmgs_batched_tile_view(views, len(sizes), ctypes.c_void_p(tv.data_ptr()), tiles - 1) == -1


def test_prepare_rejects_bad_views():
    lib = _lib()
    _, views = _views([(64, 64), (32, 32)])
    views[1].width = 0
    tt, tc, mc, tp = N.c_int32(0), N.c_int32(0), N.c_int32(0), N.c_int64(0)
    assert lib.dmgs_batched_prepare_views(views, 2, ctypes.byref(tt), ctypes.byref(tc), ctypes.byref(mc),
                                          ctypes.byref(tp)) == -1
    assert "view 1" in N.last_error()


def test_workspace_and_offsets():
    lib = _lib()
    B, P, K, T, X = 3, 100000, 2_000_000, 7500, 1_920_000
    n = lib.dmgs_batched_state_size(B, P, T, X)
    # at least every compulsory buffer: 48 B splat + 7 x 4 B per entry, 8 B range per tile, 8 B per pixel
    assert n >= B * P * (48 + 28) + T * 8 + X * 8
    assert lib.dmgs_batched_lists_size(K) >= 16 * K
    # one packed 36-byte accumulator row per (view, Gaussian), whole 256-byte blocks
    assert lib.dmgs_batched_backward_size(B, P) == (B * P * 36 + 255) // 256 * 256
    assert lib.dmgs_batched_state_size(0, P, T, X) == 0
    assert lib.dmgs_batched_lists_size(0) == 0
    offs = (N.c_size_t * 9)()
    N.check(lib.dmgs_batched_offsets(B, P, T, X, K, offs), "offsets")
    state_offs = list(offs[:6]) + [offs[8]]
    assert all(o % 4 == 0 and o < n for o in state_offs)
    assert len(set(state_offs)) == 7
    assert offs[6] < offs[7] < lib.dmgs_batched_lists_size(K)
    assert lib.dmgs_batched_offsets(B, 0, T, X, K, offs) == -1
    assert lib.dmgs_batched_offsets(B, 1 << 30, T, X, K, offs) == -1  # more than 2^31 - 1 entries


def _desc(**kw):
    d = N.DmgsBatch()
    d.views = d.tile_view = d.state = d.lists = d.bwd_ws = 0x1000
    d.num_views, d.sh_degree, d.num_gaussians = 2, 3, 100
    d.scale_modifier, d.variant, d.total_tiles, d.total_pixels, d.capacity = 1.0, 0, 4, 100, 10
    d.total_cells = 1
    d.cell_mode, d.max_view_cells = 0, 1
    for k, v in kw.items():
        setattr(d, k, v)
    return ctypes.byref(d)


def test_batch_descriptor_layout():
    # struct dmgs_batch: 2 pointers, 2 int32, int64, float, 3 int32, 2 int64, 3 pointers, 2 int32, pointer
    assert ctypes.sizeof(N.DmgsBatch) == 104 and N.DmgsBatch.dens.offset == 96
    assert N.DmgsBatch.capacity.offset == 56 and N.DmgsBatch.state.offset == 64
    assert N.DmgsBatch.cell_mode.offset == 88 and N.DmgsBatch.max_view_cells.offset == 92
    assert N.DmgsBatch.total_cells.offset == 44 and N.DmgsView.cell_base.offset == 196
    assert ctypes.sizeof(N.DmgsView) == 208


def test_cell_sort_width():
    """The cell-list sort covers exactly the bits of the largest cell id:
    global cell ids 0 .. total_cells - 1 in cell mode 0 (65536 cells need 16
    bits, so bit 16 — the first bit of the tile mask above the cell id —
    never enters the sort and a cell's entries stay one run), the cell within
    the view in mode 1."""
    lib = _lib()
    for cells, bits in ((1, 1), (2, 1), (3, 2), (256, 8), (257, 9), (32768, 15), (32769, 16), (65535, 16),
                        (65536, 16)):
        got = lib.dmgs_batched_cell_sort_bits(_desc(total_cells=cells, total_tiles=16 * cells, num_views=20))
        assert got == bits, (cells, got)
    for vc, bits in ((1, 1), (169, 8), (256, 8), (257, 9), (4096, 12)):
        got = lib.dmgs_batched_cell_sort_bits(_desc(cell_mode=1, max_view_cells=vc, total_cells=2 * vc,
                                                    total_tiles=32 * vc))
        assert got == bits, (vc, got)
    assert lib.dmgs_batched_cell_sort_bits(_desc(total_cells=65537, total_tiles=16 * 65537)) == -1


def test_entry_points_validate_before_launch():
    lib = _lib()
    fake = ctypes.c_void_p(0x1000)
    A = BT.STAGES_ALL
    assert lib.dmgs_batched_forward(None, fake, fake, A, None) == -1
    assert "batch is null" in N.last_error()
    assert lib.dmgs_batched_forward(_desc(), None, fake, A, None) == -1
    assert lib.dmgs_batched_forward(_desc(views=None), fake, fake, A, None) == -1
    # too many entries for 32-bit indices
    assert lib.dmgs_batched_forward(_desc(num_views=4096, num_gaussians=1 << 20), fake, fake, A, None) == -1
    assert "invalid batch" in N.last_error()
    assert lib.dmgs_batched_forward(_desc(num_views=2, num_gaussians=1 << 30), fake, fake, A, None) == -1
    assert lib.dmgs_batched_forward(_desc(total_tiles=1 << 20, total_cells=(1 << 16) + 1), fake, fake, A,
                                    None) == -1  # cell ids are 16 bits
    assert lib.dmgs_batched_forward(_desc(sh_degree=4), fake, fake, A, None) == -1
    assert "sh_degree" in N.last_error()
    assert lib.dmgs_batched_forward(_desc(state=0x1004), fake, fake, A, None) == -1
    assert "aligned" in N.last_error()
    assert lib.dmgs_batched_forward(_desc(total_cells=0), fake, fake, A, None) == -1
    assert lib.dmgs_batched_forward(_desc(total_cells=5), fake, fake, A, None) == -1  # more cells than tiles
    # view-major lists: at most 16 views of at most 4096 cells
    assert lib.dmgs_batched_forward(_desc(dens=0x1002), fake, fake, A, None) == -1
    assert "dens" in N.last_error()
    assert lib.dmgs_batched_forward(_desc(cell_mode=2), fake, fake, A, None) == -1
    assert "cell_mode" in N.last_error()
    assert lib.dmgs_batched_forward(_desc(cell_mode=1, num_views=17), fake, fake, A, None) == -1
    assert lib.dmgs_batched_forward(_desc(cell_mode=1, max_view_cells=0), fake, fake, A, None) == -1
    assert lib.dmgs_batched_forward(_desc(cell_mode=1, total_tiles=1 << 16, total_cells=8192, max_view_cells=4097),
                                    fake, fake, A, None) == -1
    # sorted lists: in the first buffers, but in the second after a one-pass (<= 256 cells per view) sort
    ko, vo = N.c_size_t(0), N.c_size_t(0)
    offs = (N.c_size_t * 9)()
    N.check(lib.dmgs_batched_offsets(2, 100, 4, 100, 10, offs), "offsets")
    N.check(lib.dmgs_batched_list_offsets(_desc(), ctypes.byref(ko), ctypes.byref(vo)), "list offsets")
    assert (ko.value, vo.value) == (offs[6], offs[7])
    N.check(lib.dmgs_batched_list_offsets(_desc(cell_mode=1), ctypes.byref(ko), ctypes.byref(vo)), "list offsets")
    assert len({ko.value, vo.value, offs[6], offs[7]}) == 4
    for bad in (0, 8, -1):
        assert lib.dmgs_batched_forward(_desc(), fake, fake, bad, None) == -1
        assert "stage mask" in N.last_error()
    # binning / blend need lists and a capacity; the blend needs images
    assert lib.dmgs_batched_forward(_desc(capacity=0), fake, fake, BT.STAGE_BINNING, None) == -1
    assert lib.dmgs_batched_forward(_desc(lists=None), fake, fake, BT.STAGE_BLEND, None) == -1
    assert lib.dmgs_batched_forward(_desc(), fake, None, BT.STAGE_BLEND, None) == -1
    # backward needs the accumulator, a capacity and its buffers
    B2 = BT.STAGE_BLEND | BT.STAGE_GEOMETRY
    assert lib.dmgs_batched_backward(_desc(bwd_ws=None), fake, fake, fake, B2, None) == -1
    assert lib.dmgs_batched_backward(_desc(capacity=0), fake, fake, fake, B2, None) == -1
    assert lib.dmgs_batched_backward(_desc(), fake, None, fake, B2, None) == -1
    assert lib.dmgs_batched_backward(_desc(), fake, fake, None, B2, None) == -1
    assert lib.dmgs_batched_backward(_desc(), fake, fake, fake, BT.STAGE_BINNING, None) == -1


def test_pack_unpack_round_trip():
    m = S.free_scene(257, seed=3, sh_degree=3)
    p = BT.pack_model(m)
    assert p.shape == (BT.CHANNELS, 257) and p.is_contiguous()
    assert torch.equal(p[BT.OPACITY], m._opacity[:, 0])
    assert torch.equal(p[BT.SH + 3 * 5 + 2], m._features_rest[:, 4, 2])
    back = BT.unpack_grads(p)
    for k in ("_xyz", "_scaling", "_rotation", "_opacity", "_features_dc", "_features_rest"):
        assert torch.equal(back[k], getattr(m, k)), k
    # lower SH degrees are zero-padded to 16 coefficients
    m1 = S.free_scene(10, seed=4, sh_degree=1)
    p1 = BT.pack_model(m1)
    assert torch.equal(p1[BT.SH + 12:], torch.zeros(48 - 12, 10))
    assert torch.equal(BT.unpack_grads(p1, sh_coeffs=4)["_features_rest"], m1._features_rest)
