"""Shared test helpers: run the HIP pipeline and read its intermediate state
out of the workspaces, and build oracle inputs from the same bytes."""
import ctypes

import numpy as np
import torch

from dmgs_amd import _native as N
from dmgs_amd import rasterizer as RZ
from oracle import raster as OR


def settings_from_cam(cam, bg, sh_degree, device="cuda", scale_modifier=1.0, debug=False):
    import math
    return RZ.GaussianRasterizationSettings(
        image_height=cam.image_height, image_width=cam.image_width, tanfovx=math.tan(cam.FoVx * 0.5),
        tanfovy=math.tan(cam.FoVy * 0.5), bg=bg.to(device), scale_modifier=scale_modifier,
        viewmatrix=cam.world_view_transform.to(device), projmatrix=cam.full_proj_transform.to(device),
        sh_degree=sh_degree, campos=cam.camera_center.to(device), prefiltered=False, debug=debug)


def oracle_settings(rs):
    return OR.Settings(rs.image_height, rs.image_width, rs.tanfovx, rs.tanfovy, rs.bg.cpu(), rs.scale_modifier,
                       rs.viewmatrix.cpu(), rs.projmatrix.cpu(), rs.sh_degree, rs.campos.cpu())


def workspace_views(state, W, H):
    """Intermediate buffers of a forward ``state`` as CPU numpy arrays."""
    P, K = state["P"], state["K"]
    offs = (ctypes.c_size_t * 10)()
    N.check(N.load().dmgs_workspace_offsets(P, K, W, H, offs), "offsets")
    geom = state["geom"]
    binning = state["binning"]
    image = state["image"]
    torch.cuda.synchronize()

    def view(buf, off, count, dtype):
        nbytes = count * np.dtype(dtype).itemsize
        return buf[off:off + nbytes].cpu().numpy().view(dtype)

    gx, gy = OR.tile_grid(W, H)
    T = gx * gy
    splats = view(geom, offs[0], P * 12, np.float32).reshape(P, 12)
    return dict(
        splats=splats,
        xy=splats[:, 0:2].copy(), conic=np.stack([splats[:, 2], splats[:, 3], splats[:, 4]], 1),
        opacity=splats[:, 5].copy(), rgb=np.stack([splats[:, 6], splats[:, 7], splats[:, 8]], 1),
        depth=splats[:, 9].copy(), radius_bits=splats[:, 10].copy().view(np.int32),
        depth_keys=view(geom, offs[1], P, np.uint32), order=view(geom, offs[2], P, np.uint32),
        tiles_touched=view(geom, offs[3], P, np.uint32), offsets=view(geom, offs[4], P, np.uint32),
        tile_ids=view(binning, offs[5], K, np.uint32), values=view(binning, offs[6], K, np.uint32),
        ranges=view(image, offs[7], T * 2, np.uint32).reshape(T, 2).astype(np.int64),
        final_T=view(image, offs[8], W * H, np.float32).reshape(H, W),
        n_contrib=view(image, offs[9], W * H, np.uint32).reshape(H, W),
    )


def grad_rel_err(gpu, ref):
    """Norm-relative error ||gpu - ref|| / ||ref|| (0 if both are zero)."""
    gpu = torch.as_tensor(gpu).double().cpu().reshape(-1)
    ref = torch.as_tensor(ref).double().cpu().reshape(-1)
    den = ref.norm().item()
    num = (gpu - ref).norm().item()
    return num / den if den > 0 else num
