"""Independent differentiable FP64 formulation of the rasterizer forward
(test infrastructure only; small inputs).

Used to pin ``oracle.raster``'s hand-written backward: torch autograd of this
forward must equal it, and central finite differences of this forward must
equal autograd wherever no quirk is active.  The reference's gradient quirks
([U], listed in ``oracle.raster``) are expressed as value-preserving
``detach`` constructions so autograd reproduces them:

* alpha = raw - relu(raw - 0.99).detach()        (clamp not masked)
* tx' = (clamp(tx/tz) tz).detach() when clamped   (tan(fov) clamp)
* s = scale + (mod scale - scale).detach()        (dL/dscale w.r.t. mod*scale)
* ndc.xy = p_proj.xy + means2D.xy                 (means2D gradient in NDC)

The 1 / (det^2 + 1e-7) factor of the cov2D backward is not modelled; pins
use splats with det >= 1, where it changes gradients by < 1e-7 relative.
Blend membership (tile lists, skip and stop masks) is taken from the
non-differentiable oracle and held constant.
"""
import numpy as np
import torch

from . import raster as R
from . import sh as sh_mod

F64 = torch.float64


def rasterize_autograd(means3D, means2D, opacities, settings, shs=None, colors_precomp=None,
                       scales=None, rotations=None, cov3D_precomp=None, state=None):
    """Differentiable forward; inputs are FP64 leaves.  ``state`` is the
    non-differentiable oracle state (tile lists); computed if not given."""
    W, H = settings.image_width, settings.image_height
    vm = settings.viewmatrix.detach().to(F64)
    pm = settings.projmatrix.detach().to(F64)
    if state is None:
        _, _, state = R.rasterize(means3D.detach(), opacities.detach(), settings,
                                  None if shs is None else shs.detach(),
                                  None if colors_precomp is None else colors_precomp.detach(),
                                  None if scales is None else scales.detach(),
                                  None if rotations is None else rotations.detach(),
                                  None if cov3D_precomp is None else cov3D_precomp.detach())
    pre = state["pre"]
    P = means3D.shape[0]
    p = means3D
    p_hom = p @ pm[:3, :] + pm[3, :]
    p_w = 1.0 / (p_hom[:, 3] + 1e-7)
    ndc = p_hom[:, :2] * p_w[:, None] + means2D[:, :2]
    t = p @ vm[:3, :3] + vm[3, :3]
    if cov3D_precomp is not None:
        S3 = R.six_to_sym(cov3D_precomp)
    else:
        mod = settings.scale_modifier
        s = scales + (scales * mod - scales).detach()
        Rm = R.quat_to_rot(rotations)
        M3 = Rm * s.unsqueeze(-2)
        S3 = M3 @ M3.transpose(-1, -2)
    tanx = float(np.float32(settings.tanfovx))
    tany = float(np.float32(settings.tanfovy))
    fx, fy = W / (2.0 * tanx), H / (2.0 * tany)
    limx, limy = 1.3 * tanx, 1.3 * tany
    tz = t[:, 2]
    cx = ((t[:, 0] / tz).abs() > limx).detach()
    cy = ((t[:, 1] / tz).abs() > limy).detach()
    txc = (torch.clamp(t[:, 0] / tz, -limx, limx) * tz).detach()
    tyc = (torch.clamp(t[:, 1] / tz, -limy, limy) * tz).detach()
    tx = torch.where(cx, txc, t[:, 0])
    ty = torch.where(cy, tyc, t[:, 1])
    zero = torch.zeros_like(tz)
    J = torch.stack([fx / tz, zero, -(fx * tx) / (tz * tz),
                     zero, fy / tz, -(fy * ty) / (tz * tz)], -1).reshape(P, 2, 3)
    M = J @ vm[:3, :3].T
    S2 = M @ S3 @ M.transpose(1, 2)
    a, b, c = S2[:, 0, 0] + 0.3, S2[:, 0, 1], S2[:, 1, 1] + 0.3
    det = a * c - b * b
    det = torch.where(pre["visible"], det, torch.ones_like(det))
    conic = torch.stack([c / det, -b / det, a / det], -1)
    xy = torch.stack([((ndc[:, 0] + 1.0) * W - 1.0) * 0.5, ((ndc[:, 1] + 1.0) * H - 1.0) * 0.5], -1)
    if shs is not None:
        dirs = p - settings.campos.detach().to(F64).reshape(1, 3)
        dirs = dirs / dirs.norm(dim=1, keepdim=True)
        rgb = torch.clamp(sh_mod.eval_sh(settings.sh_degree, shs.transpose(1, 2), dirs) + 0.5, min=0.0)
    else:
        rgb = colors_precomp
    opac = opacities.reshape(P)
    bg = settings.bg.detach().to(F64).reshape(-1)
    Cn = rgb.shape[1]

    ranges, vals = state["ranges"], torch.as_tensor(np.asarray(state["vals"]), dtype=torch.int64)
    out = bg[:, None, None].expand(Cn, H, W).clone()
    for tile in np.nonzero(ranges[:, 1] > ranges[:, 0])[0]:
        s0, e0 = int(ranges[tile, 0]), int(ranges[tile, 1])
        ids = vals[s0:e0]
        px, py = R._tile_pixels(int(tile), W, H)
        # masks from the detached walk
        with torch.no_grad():
            _, _, _, _, alpha_d, valid, _ = R._tile_eval(ids, px, py, xy.detach(), conic.detach(),
                                                          opac.detach())
            include = R._tile_walk(valid, alpha_d)[0]
        dx = xy[ids, 0][None, :] - torch.from_numpy(px).to(F64)[:, None]
        dy = xy[ids, 1][None, :] - torch.from_numpy(py).to(F64)[:, None]
        qa, qb, qc = conic[ids, 0][None, :], conic[ids, 1][None, :], conic[ids, 2][None, :]
        power = -0.5 * (qa * dx * dx + qc * dy * dy) - qb * dx * dy
        raw = opac[ids][None, :] * torch.exp(power)
        alpha = raw - torch.relu(raw - 0.99).detach()
        a_inc = torch.where(include, alpha, torch.zeros_like(alpha))
        one_m = 1 - a_inc
        Tinc = torch.cumprod(one_m, dim=1)
        Tbef = torch.cat([torch.ones_like(Tinc[:, :1]), Tinc[:, :-1]], 1)
        Cpix = (a_inc * Tbef) @ rgb[ids]
        Tfin = Tinc[:, -1]
        val = Cpix + Tfin[:, None] * bg[None, :]
        pyt, pxt = torch.from_numpy(py), torch.from_numpy(px)
        out = out.index_put((torch.arange(Cn)[:, None], pyt[None, :], pxt[None, :]), val.T)
    return out
