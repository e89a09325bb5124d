"""The C-ABI library loads, exports every symbol include/dmgs.h declares and
rejects bad arguments before launching anything (CPU only: no compute call)."""
import ctypes
import os
import re

import pytest

from dmgs_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    with open(os.path.join(ROOT, "include", "dmgs.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dmgs_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(N.library_path()):
        pytest.skip("library not built (run __graft_entry__.build())")
    lib = N.load()
    declared = _declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), f"{name} declared in include/dmgs.h but not exported"
    assert set(N.exported_symbols()) <= set(declared)
    assert lib.dmgs_version().startswith(b"dmgs-gfx950")


def test_argument_validation_without_gpu():
    if not os.path.exists(N.library_path()):
        pytest.skip("library not built")
    lib = N.load()
    nul = [None] * 10
    assert lib.dmgs_forward_geometry(None, 10, *nul, None) == -1
    assert "settings" in N.last_error()
    s = N.DmgsSettings()
    assert lib.dmgs_forward_geometry(ctypes.byref(s), 10, *nul, None) == -1
    assert "image size" in N.last_error()
    s.image_width, s.image_height, s.tanfovx, s.tanfovy = 64, 48, 0.5, 0.5
    assert lib.dmgs_forward_geometry(ctypes.byref(s), 10, *nul, None) == -1
    assert "device pointers" in N.last_error()
    fake = ctypes.c_void_p(0x1000)
    s.viewmatrix = s.projmatrix = s.campos = s.bg = 0x1000
    # exactly one of shs / colors_precomp
    args = [fake, fake, fake, fake, fake, fake, None, fake, fake, fake]
    assert lib.dmgs_forward_geometry(ctypes.byref(s), 10, *args, None) == -1
    assert "SHs or precomputed colors" in N.last_error()
    # exactly one of (scales, rotations) / cov3D
    args = [fake, fake, None, fake, fake, fake, fake, fake, fake, fake]
    assert lib.dmgs_forward_geometry(ctypes.byref(s), 10, *args, None) == -1
    assert "scale/rotation pair" in N.last_error()
    s.sh_degree = 4
    assert lib.dmgs_forward_geometry(ctypes.byref(s), 10, *nul, None) == -1
    assert "sh_degree" in N.last_error()
    assert lib.dmgs_sort_pairs(fake, fake, fake, fake, 10, 40, fake, None) == -1
    assert lib.dmgs_inclusive_scan_u32(None, None, 5, None, None) == -1
    # the allocation-free scan checks its size limit and pointers before any HIP call
    assert lib.dmgs_inclusive_scan_u32_noalloc(fake, fake, (1 << 30) + 1, None) == -1
    assert "maximum of 1073741824" in N.last_error()
    assert lib.dmgs_inclusive_scan_u32_noalloc(None, fake, 5, None) == -1
    assert lib.dmgs_inclusive_scan_u32_noalloc(fake, fake, -1, None) == -1
    assert lib.dmgs_inclusive_scan_u32_noalloc(None, None, 0, None) == 0
    assert lib.dmgs_dssim_l1_forward(fake, fake, 0, 4, 4, 0.2, fake, fake, None) == -1
    assert "image shape" in N.last_error()


def test_workspace_sizes():
    if not os.path.exists(N.library_path()):
        pytest.skip("library not built")
    lib = N.load()
    g1, g2 = lib.dmgs_geometry_workspace_size(1000), lib.dmgs_geometry_workspace_size(2000)
    assert 0 < g1 < g2 and g1 % 256 == 0
    assert lib.dmgs_binning_workspace_size(0) > 0
    assert lib.dmgs_binning_workspace_size(10_000_000) >= 16 * 10_000_000
    assert lib.dmgs_image_workspace_size(800, 800) >= 8 * 800 * 800
    assert lib.dmgs_backward_workspace_size(10) >= 10 * 36  # one 36-byte accumulator row per Gaussian
    offs = (ctypes.c_size_t * 10)()
    assert lib.dmgs_workspace_offsets(100, 1000, 64, 48, offs) == 0
    assert offs[0] == 0 and all(o % 256 == 0 for o in offs)
