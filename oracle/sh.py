"""Real spherical harmonics, degree 0-3, FP64 (test infrastructure only).

Restates utils/sh_utils.py:26-110 (eval_sh, constants :41-58; the CUDA
rasterizer [U] uses the same constants and polynomials, then adds 0.5 and
clamps at 0).

Pin (tests/test_oracle_sh_camera.py): every basis function is compared against
``scipy.special.sph_harm`` turned into real harmonics with the Condon-Shortley
phase removed, which fixes each constant, sign and polynomial independently of
this file; degree-0 and axis directions are checked as special cases.
"""
import torch

C0 = 0.28209479177387814
C1 = 0.4886025119029199
C2 = (1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
      -1.0925484305920792, 0.5462742152960396)
C3 = (-0.5900435899266435, 2.890611442640554, -0.4570457994644658,
      0.3731763325901154, -0.4570457994644658, 1.445305721320277,
      -0.5900435899266435)


def sh_basis(deg, dirs):
    """Basis values (..., (deg+1)^2) at unit directions ``dirs`` (..., 3)."""
    x, y, z = dirs[..., 0], dirs[..., 1], dirs[..., 2]
    out = [torch.full_like(x, C0)]
    if deg > 0:
        out += [-C1 * y, C1 * z, -C1 * x]
    if deg > 1:
        xx, yy, zz = x * x, y * y, z * z
        xy, yz, xz = x * y, y * z, x * z
        out += [C2[0] * xy, C2[1] * yz, C2[2] * (2.0 * zz - xx - yy), C2[3] * xz,
                C2[4] * (xx - yy)]
    if deg > 2:
        out += [C3[0] * y * (3 * xx - yy), C3[1] * xy * z, C3[2] * y * (4 * zz - xx - yy),
                C3[3] * z * (2 * zz - 3 * xx - 3 * yy), C3[4] * x * (4 * zz - xx - yy),
                C3[5] * z * (xx - yy), C3[6] * x * (xx - 3 * yy)]
    return torch.stack(out, dim=-1)


def eval_sh(deg, sh, dirs):
    """eval_sh (utils/sh_utils.py:26-110): sh (..., C, M) coefficients,
    dirs (..., 3) unit -> (..., C).  Only the first (deg+1)^2 coefficients are read."""
    n = (deg + 1) ** 2
    basis = sh_basis(deg, dirs)            # (..., n)
    return (sh[..., :n] * basis.unsqueeze(-2)).sum(-1)
