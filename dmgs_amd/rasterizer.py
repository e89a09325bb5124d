"""Drop-in replacement for the ``diff_gaussian_rasterization`` Python module.

Same names and argument meaning as the reference's rasterizer API, as called
from gaussian_renderer/__init__.py:14,36-51,86-94 (the module itself is the
un-vendored submodule named in .gitmodules:4-6):

* ``GaussianRasterizationSettings`` — 12 fields, NamedTuple;
* ``GaussianRasterizer(settings)(means3D, means2D, opacities, shs=None,
  colors_precomp=None, scales=None, rotations=None, cov3D_precomp=None)``
  -> ``(color (3,H,W) float32, radii (P,) int32)``, plus ``markVisible``;
* ``rasterize_gaussians`` / ``_RasterizeGaussians`` (torch.autograd.Function)
  whose backward returns gradients in input order, ``means2D.grad`` in NDC
  units with z = 0.

Errors follow the reference: an ``Exception`` unless exactly one of {shs,
colors_precomp} and exactly one of {(scales, rotations), cov3D_precomp} is
given; P == 0 returns an all-zero image.  With ``debug`` every kernel is
synchronised and checked, and a failing call saves its CPU-copied arguments to
``snapshot_fw.dump`` / ``snapshot_bw.dump`` before re-raising.

All compute runs in the gfx950 HIP library through ``dmgs_amd._native``.
"""
import ctypes
from typing import NamedTuple

import torch
import torch.nn as nn

from . import _native as N


class GaussianRasterizationSettings(NamedTuple):
    image_height: int
    image_width: int
    tanfovx: float
    tanfovy: float
    bg: torch.Tensor
    scale_modifier: float
    viewmatrix: torch.Tensor
    projmatrix: torch.Tensor
    sh_degree: int
    campos: torch.Tensor
    prefiltered: bool
    debug: bool


def _f32(t, device):
    t = t.to(device=device, dtype=torch.float32).contiguous()
    if t.data_ptr() % 16:  # the kernels use 16-byte vector loads
        t = t.clone()
    return t


def _present(t):
    return t is not None and t.numel() > 0


class _Frame:
    """Device-side copies of the per-view constants and the C settings struct."""

    def __init__(self, rs, device, sh_coeffs):
        self.view = _f32(rs.viewmatrix, device)
        self.proj = _f32(rs.projmatrix, device)
        self.campos = _f32(rs.campos, device).reshape(-1)
        self.bg = _f32(rs.bg, device).reshape(-1)
        if self.bg.numel() != 3:
            raise ValueError("bg must have 3 channels")
        s = N.DmgsSettings()
        s.image_width = int(rs.image_width)
        s.image_height = int(rs.image_height)
        s.tanfovx = float(rs.tanfovx)
        s.tanfovy = float(rs.tanfovy)
        s.scale_modifier = float(rs.scale_modifier)
        s.sh_degree = int(rs.sh_degree)
        s.sh_coeffs = int(sh_coeffs)
        s.prefiltered = int(bool(rs.prefiltered))
        s.debug = int(bool(rs.debug))
        s.viewmatrix = self.view.data_ptr()
        s.projmatrix = self.proj.data_ptr()
        s.campos = self.campos.data_ptr()
        s.bg = self.bg.data_ptr()
        self.c = s


class _NullStage:
    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


_NULL_STAGE = _NullStage()
_stage_hook = None


def set_stage_hook(hook):
    """Install ``hook(name)`` returning a context manager entered around every
    C-ABI stage call on the launching stream (used by the benchmark to time
    stages with device events).  ``None`` removes it."""
    global _stage_hook
    _stage_hook = hook


def _stage(name):
    return _NULL_STAGE if _stage_hook is None else _stage_hook(name)


def _cpu_copy(args):
    return tuple(a.detach().cpu() if isinstance(a, torch.Tensor) else a for a in args)


def rasterize_gaussians(means3D, means2D, sh, colors_precomp, opacities, scales, rotations, cov3Ds_precomp,
                        raster_settings):
    return _RasterizeGaussians.apply(means3D, means2D, sh, colors_precomp, opacities, scales, rotations,
                                     cov3Ds_precomp, raster_settings)


def forward_raw(means3D, opacities, sh, colors_precomp, scales, rotations, cov3Ds_precomp, rs):
    """Run the forward pipeline; returns (color, radii, state) where state holds
    everything the backward needs (the reference's three byte buffers)."""
    dev = means3D.device
    if dev.type != "cuda":
        raise RuntimeError("dmgs_amd rasterizer needs CUDA/HIP tensors")
    N.load()
    P = means3D.shape[0]
    if means3D.dim() != 2 or means3D.shape[1] != 3:
        raise RuntimeError("means3D must have dimensions (num_points, 3)")
    H, W = int(rs.image_height), int(rs.image_width)
    M = sh.shape[1] if _present(sh) else 0
    radii = torch.zeros(P, dtype=torch.int32, device=dev)
    if P == 0:
        return torch.zeros(3, H, W, dtype=torch.float32, device=dev), radii, None
    if _present(sh) == _present(colors_precomp):
        raise Exception("Please provide excatly one of either SHs or precomputed colors!")
    if (_present(scales) or _present(rotations)) == _present(cov3Ds_precomp) or \
            (_present(scales) != _present(rotations)):
        raise Exception("Please provide exactly one of either scale/rotation pair or precomputed 3D covariance!")
    fr = _Frame(rs, dev, M)
    t = dict(
        means3D=_f32(means3D, dev), opacities=_f32(opacities, dev),
        shs=_f32(sh, dev) if _present(sh) else None,
        colors=_f32(colors_precomp, dev) if _present(colors_precomp) else None,
        scales=_f32(scales, dev) if _present(scales) else None,
        rotations=_f32(rotations, dev) if _present(rotations) else None,
        cov3D=_f32(cov3Ds_precomp, dev) if _present(cov3Ds_precomp) else None,
    )
    lib = N.load()
    stream = N.stream_handle(dev)
    geom = torch.empty(lib.dmgs_geometry_workspace_size(P), dtype=torch.uint8, device=dev)
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    with _stage("geometry"):
        N.check(lib.dmgs_forward_geometry(ctypes_ref(fr.c), P, N.ptr(t["means3D"]), N.ptr(t["opacities"]),
                                          N.ptr(t["shs"]), N.ptr(t["colors"]), N.ptr(t["scales"]),
                                          N.ptr(t["rotations"]), N.ptr(t["cov3D"]), N.ptr(radii), N.ptr(geom),
                                          N.ptr(count), stream), "dmgs_forward_geometry")
    K = int(count.item())  # the one host synchronisation of the plain path (as in the reference)
    binning = torch.empty(max(1, lib.dmgs_binning_workspace_size(K)), dtype=torch.uint8, device=dev)
    image = torch.empty(lib.dmgs_image_workspace_size(W, H), dtype=torch.uint8, device=dev)
    color = torch.empty(3, H, W, dtype=torch.float32, device=dev)
    with _stage("binning"):
        N.check(lib.dmgs_forward_binning(ctypes_ref(fr.c), P, K, N.ptr(geom), N.ptr(binning), N.ptr(image),
                                         stream), "dmgs_forward_binning")
    with _stage("render_forward"):
        N.check(lib.dmgs_render_forward(ctypes_ref(fr.c), P, K, N.ptr(geom), N.ptr(binning), N.ptr(image),
                                        N.ptr(color), stream), "dmgs_render_forward")
    state = dict(frame=fr, tensors=t, geom=geom, binning=binning, image=image, K=K, P=P, M=M)
    return color, radii, state


def backward_raw(state, radii, grad_color):
    """Run the backward pipeline for a forward ``state``; returns a dict of gradients."""
    fr, t = state["frame"], state["tensors"]
    P, K, M = state["P"], state["K"], state["M"]
    dev = t["means3D"].device
    lib = N.load()
    f = lambda *shape: torch.empty(*shape, dtype=torch.float32, device=dev)
    out = dict(means2D=f(P, 3), colors=f(P, 3), opacity=f(P, 1), conic=f(P, 2, 2), means3D=f(P, 3),
               cov3D=f(P, 6), sh=f(P, M, 3) if t["shs"] is not None else None,
               scales=f(P, 3) if t["scales"] is not None else None,
               rotations=f(P, 4) if t["rotations"] is not None else None)
    bws = torch.empty(lib.dmgs_backward_workspace_size(P), dtype=torch.uint8, device=dev)
    g = grad_color.to(dtype=torch.float32).contiguous()
    stream = N.stream_handle(dev)
    with _stage("render_backward"):
        N.check(lib.dmgs_render_backward(ctypes_ref(fr.c), P, K, N.ptr(state["geom"]), N.ptr(state["binning"]),
                                         N.ptr(state["image"]), N.ptr(bws), N.ptr(g), stream), "dmgs_render_backward")
    with _stage("geometry_backward"):
        N.check(lib.dmgs_geometry_backward(ctypes_ref(fr.c), P, N.ptr(t["means3D"]), N.ptr(t["shs"]),
                                           N.ptr(t["colors"]), N.ptr(t["scales"]), N.ptr(t["rotations"]),
                                           N.ptr(t["cov3D"]), N.ptr(radii), N.ptr(bws), N.ptr(out["means2D"]),
                                           N.ptr(out["colors"]), N.ptr(out["opacity"]), N.ptr(out["conic"]),
                                           N.ptr(out["means3D"]), N.ptr(out["cov3D"]), N.ptr(out["sh"]),
                                           N.ptr(out["scales"]), N.ptr(out["rotations"]), stream),
                "dmgs_geometry_backward")
    return out


def ctypes_ref(struct):
    return ctypes.byref(struct)


class _RasterizeGaussians(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means3D, means2D, sh, colors_precomp, opacities, scales, rotations, cov3Ds_precomp,
                raster_settings):
        args = (means3D, opacities, sh, colors_precomp, scales, rotations, cov3Ds_precomp, raster_settings)
        if raster_settings.debug:
            try:
                color, radii, state = forward_raw(*args)
            except Exception:
                torch.save(_cpu_copy(args[:-1]), "snapshot_fw.dump")
                print("\nAn error occured in forward. Please forward snapshot_fw.dump for debugging.")
                raise
        else:
            color, radii, state = forward_raw(*args)
        ctx.state = state
        ctx.raster_settings = raster_settings
        ctx.num_rendered = 0 if state is None else state["K"]
        ctx.shapes = [None if x is None else x.shape for x in
                      (means3D, means2D, sh, colors_precomp, opacities, scales, rotations, cov3Ds_precomp)]
        ctx.save_for_backward(radii)
        ctx.mark_non_differentiable(radii)
        return color, radii

    @staticmethod
    def backward(ctx, grad_out_color, _grad_radii):
        (radii,) = ctx.saved_tensors
        state = ctx.state
        shapes = ctx.shapes
        needs = ctx.needs_input_grad
        if state is None:
            z = [None if (s is None or not needs[i]) else torch.zeros(s, device=grad_out_color.device)
                 for i, s in enumerate(shapes)]
            return tuple(z) + (None,)
        if ctx.raster_settings.debug:
            try:
                g = backward_raw(state, radii, grad_out_color)
            except Exception:
                torch.save(_cpu_copy((grad_out_color,) + tuple(state["tensors"].values())), "snapshot_bw.dump")
                print("\nAn error occured in backward. Writing snapshot_bw.dump for debugging.\n")
                raise
        else:
            g = backward_raw(state, radii, grad_out_color)
        ordered = [g["means3D"], g["means2D"], g["sh"], g["colors"], g["opacity"], g["scales"], g["rotations"],
                   g["cov3D"]]
        grads = []
        for i, (gr, s) in enumerate(zip(ordered, shapes)):
            if not needs[i] or s is None or gr is None or s.numel() == 0:
                grads.append(None)
            else:
                grads.append(gr.reshape(s))
        ctx.state = None
        return tuple(grads) + (None,)


class GaussianRasterizer(nn.Module):
    def __init__(self, raster_settings):
        super().__init__()
        self.raster_settings = raster_settings

    def markVisible(self, positions):
        with torch.no_grad():
            rs = self.raster_settings
            pos = _f32(positions, positions.device)
            P = pos.shape[0]
            present = torch.zeros(P, dtype=torch.uint8, device=pos.device)
            view = _f32(rs.viewmatrix, pos.device)
            N.check(N.load().dmgs_mark_visible(P, N.ptr(pos), N.ptr(view), N.ptr(present),
                                               N.stream_handle(pos.device)), "dmgs_mark_visible")
            return present.bool()

    def forward(self, means3D, means2D, opacities, shs=None, colors_precomp=None, scales=None, rotations=None,
                cov3D_precomp=None):
        raster_settings = self.raster_settings
        if (shs is None and colors_precomp is None) or (shs is not None and colors_precomp is not None):
            raise Exception('Please provide excatly one of either SHs or precomputed colors!')
        if ((scales is None or rotations is None) and cov3D_precomp is None) or \
                ((scales is not None or rotations is not None) and cov3D_precomp is not None):
            raise Exception('Please provide exactly one of either scale/rotation pair or precomputed 3D covariance!')
        empty = lambda: torch.Tensor([])
        shs = empty() if shs is None else shs
        colors_precomp = empty() if colors_precomp is None else colors_precomp
        scales = empty() if scales is None else scales
        rotations = empty() if rotations is None else rotations
        cov3D_precomp = empty() if cov3D_precomp is None else cov3D_precomp
        return rasterize_gaussians(means3D, means2D, shs, colors_precomp, opacities, scales, rotations,
                                   cov3D_precomp, raster_settings)
