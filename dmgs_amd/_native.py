"""ctypes binding of the gfx950 HIP library (C-ABI declared in include/dmgs.h).

The library is built in-tree by ``dmgs_amd.build``.  There is no fallback:
if the library is missing or fails to load, every entry point raises.
"""
import ctypes
import os
import threading

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

from . import build as _build

c_int32, c_int64, c_size_t, c_void_p, c_float = (ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t,
                                                  ctypes.c_void_p, ctypes.c_float)


class DmgsSettings(ctypes.Structure):
    """struct dmgs_settings (include/dmgs.h)."""
    _fields_ = [
        ("image_width", c_int32), ("image_height", c_int32),
        ("tanfovx", c_float), ("tanfovy", c_float), ("scale_modifier", c_float),
        ("sh_degree", c_int32), ("sh_coeffs", c_int32), ("prefiltered", c_int32), ("debug", c_int32),
        ("viewmatrix", c_void_p), ("projmatrix", c_void_p), ("campos", c_void_p), ("bg", c_void_p),
    ]


class DmgsView(ctypes.Structure):
    """struct dmgs_view (include/dmgs.h): one camera of a batched launch."""
    _fields_ = [
        ("viewmatrix", c_float * 16), ("projmatrix", c_float * 16), ("campos", c_float * 4), ("bg", c_float * 4),
        ("tanfovx", c_float), ("tanfovy", c_float), ("focal_x", c_float), ("focal_y", c_float),
        ("width", c_int32), ("height", c_int32), ("grid_x", c_int32), ("grid_y", c_int32),
        ("tile_base", c_int32), ("cell_base", c_int32), ("pixel_base", c_int64),
    ]


class DmgsBatch(ctypes.Structure):
    """struct dmgs_batch (include/dmgs.h): a prepared batch and its workspaces."""
    _fields_ = [
        ("views", c_void_p), ("tile_view", c_void_p), ("num_views", c_int32), ("sh_degree", c_int32),
        ("num_gaussians", c_int64), ("scale_modifier", c_float), ("variant", c_int32), ("total_tiles", c_int32),
        ("total_cells", c_int32), ("total_pixels", c_int64), ("capacity", c_int64),
        ("state", c_void_p), ("lists", c_void_p), ("bwd_ws", c_void_p),
        ("cell_mode", c_int32), ("max_view_cells", c_int32), ("dens", c_void_p),
    ]


# name -> (restype, argtypes)
_SIGNATURES = {
    "dmgs_last_error": (ctypes.c_char_p, []),
    "dmgs_version": (ctypes.c_char_p, []),
    "dmgs_geometry_workspace_size": (c_size_t, [c_int64]),
    "dmgs_binning_workspace_size": (c_size_t, [c_int64]),
    "dmgs_image_workspace_size": (c_size_t, [c_int32, c_int32]),
    "dmgs_backward_workspace_size": (c_size_t, [c_int64]),
    "dmgs_workspace_offsets": (c_int32, [c_int64, c_int64, c_int32, c_int32, ctypes.POINTER(c_size_t)]),
    "dmgs_forward_geometry": (c_int32, [ctypes.POINTER(DmgsSettings), c_int64] + [c_void_p] * 10 + [c_void_p]),
    "dmgs_forward_binning": (c_int32, [ctypes.POINTER(DmgsSettings), c_int64, c_int64] + [c_void_p] * 3 + [c_void_p]),
    "dmgs_render_forward": (c_int32, [ctypes.POINTER(DmgsSettings), c_int64, c_int64] + [c_void_p] * 4 + [c_void_p]),
    "dmgs_forward_render": (c_int32, [ctypes.POINTER(DmgsSettings), c_int64, c_int64] + [c_void_p] * 4 + [c_void_p]),
    "dmgs_render_backward": (c_int32, [ctypes.POINTER(DmgsSettings), c_int64, c_int64] + [c_void_p] * 5 + [c_void_p]),
    "dmgs_geometry_backward": (c_int32, [ctypes.POINTER(DmgsSettings), c_int64] + [c_void_p] * 17 + [c_void_p]),
    "dmgs_backward": (c_int32, [ctypes.POINTER(DmgsSettings), c_int64, c_int64] + [c_void_p] * 21 + [c_void_p]),
    "dmgs_mark_visible": (c_int32, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dmgs_ply_rows_to_params": (c_int32, [c_void_p, c_int64, c_int32, ctypes.POINTER(c_int32), c_void_p, c_void_p]),
    "dmgs_params_to_ply_rows": (c_int32, [c_void_p, c_int64, c_int32, ctypes.POINTER(c_int32), c_void_p, c_void_p]),
    "dmgs_image_rgb8": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p]),
    "dmgs_sort_scratch_size": (c_size_t, [c_int64]),
    "dmgs_sort_pairs": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_void_p]),
    "dmgs_bind_forward": (c_int32, [c_int64, c_void_p, c_int64, c_void_p, c_int32, c_void_p, c_int32, c_void_p,
                                    c_void_p, c_float, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dmgs_bind_backward": (c_int32, [c_int64, c_void_p, c_int64, c_void_p, c_int32, c_void_p, c_int32, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p]),
    "dmgs_bind_cov_forward": (c_int32, [c_int64, c_void_p, c_int64, c_void_p, c_int32, c_void_p, c_int32, c_void_p,
                                        c_float, c_float, c_int32, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dmgs_bind_cov_backward": (c_int32, [c_int64, c_void_p, c_int64, c_void_p, c_int32, c_void_p, c_int32,
                                         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dmgs_faces_in_frustum": (c_int32, [c_int64, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dmgs_dssim_l1_workspace_size": (c_size_t, [c_int32, c_int32, c_int32]),
    "dmgs_dssim_l1_forward": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_int32, c_float, c_void_p, c_void_p,
                                        c_void_p]),
    "dmgs_dssim_l1_backward": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_int32, c_float, c_void_p,
                                         c_void_p, c_void_p, c_void_p]),
    "dmgs_dssim_l1_forward_groups": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32, c_float,
                                               c_void_p, c_void_p, c_void_p]),
    "dmgs_dssim_l1_backward_groups": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32, c_float,
                                                c_void_p, c_void_p, c_void_p, c_void_p]),
    "dmgs_dssim_target_stats": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p]),
    "dmgs_dssim_l1_forward_groups_stats": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32,
                                                     c_int32, c_float, c_void_p, c_void_p, c_void_p]),
    "dmgs_scan_scratch_size": (c_size_t, [c_int64]),
    "dmgs_inclusive_scan_u32": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    "dmgs_inclusive_scan_u32_noalloc": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p]),
    "dmgs_batched_prepare_views": (c_int32, [ctypes.POINTER(DmgsView), c_int32, ctypes.POINTER(c_int32),
                                             ctypes.POINTER(c_int32), ctypes.POINTER(c_int32),
                                             ctypes.POINTER(c_int64)]),
    "dmgs_batched_list_offsets": (c_int32, [ctypes.POINTER(DmgsBatch), ctypes.POINTER(c_size_t),
                                            ctypes.POINTER(c_size_t)]),
    "dmgs_batched_cell_sort_bits": (c_int32, [ctypes.POINTER(DmgsBatch)]),
    "dmgs_batched_tile_view": (c_int32, [ctypes.POINTER(DmgsView), c_int32, c_void_p, c_int32]),
    "dmgs_batched_state_size": (c_size_t, [c_int32, c_int64, c_int32, c_int64]),
    "dmgs_batched_lists_size": (c_size_t, [c_int64]),
    "dmgs_batched_backward_size": (c_size_t, [c_int32, c_int64]),
    "dmgs_batched_offsets": (c_int32, [c_int32, c_int64, c_int32, c_int64, c_int64, ctypes.POINTER(c_size_t)]),
    "dmgs_batched_forward": (c_int32, [ctypes.POINTER(DmgsBatch), c_void_p, c_void_p, c_int32, c_void_p]),
    "dmgs_batched_backward": (c_int32, [ctypes.POINTER(DmgsBatch), c_void_p, c_void_p, c_void_p, c_int32,
                                        c_void_p]),
}

_lib = None
_lock = threading.Lock()


def library_path():
    """The in-tree build, or the library named by DMGS_LIBRARY (the CPU
    tests point it at the host-side AddressSanitizer build)."""
    return os.environ.get("DMGS_LIBRARY") or _build.library_path()


def load():
    """Load (once) and return the ctypes handle; raises if the library is absent."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(
                f"dmgs_amd HIP library not found at {path}; build it with "
                "`python -m dmgs_amd.build` (there is no CPU fallback)")
        lib = ctypes.CDLL(path)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def exported_symbols():
    return sorted(_SIGNATURES)


def last_error():
    return load().dmgs_last_error().decode()


def check(rc, what):
    if rc != 0:
        kind = "invalid argument" if rc == -1 else "HIP error"
        raise RuntimeError(f"{what} failed ({kind}): {last_error()}")


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_handle(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
