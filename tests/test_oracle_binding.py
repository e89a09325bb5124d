"""Pins of the FP64 binding oracle (CPU): invariants, closed forms, finite differences."""
import numpy as np
import torch

from dmgs_amd import binding as PB
from dmgs_amd import scenes as S
from oracle import binding as OB

F64 = torch.float64


def _mesh(sub=2, seed=0, G=6, per_gaussian=True):
    g = torch.Generator().manual_seed(seed)
    v, f = S.icosphere(sub)
    v = (v + torch.randn(v.shape, generator=g) * 0.02).to(F64)
    n = f.shape[0] * G
    if per_gaussian:
        e = -torch.log(torch.rand(n, 3, generator=g, dtype=F64))
        bc = e / e.sum(1, keepdim=True)
    else:
        bc = PB.barycentric_v2(G)[0].to(F64)
    rot2d = torch.randn(n, 2, generator=g, dtype=F64)
    scale2d = torch.randn(n, 2, generator=g, dtype=F64) * 0.3 - 3
    return v, f, bc, rot2d, scale2d


def test_stage3_invariants():
    v, f, bc, rot2d, scale2d = _mesh()
    xyz, q, scale, R = OB.stage3(v, f, bc, rot2d, scale2d, 1e-4)
    I = torch.eye(3, dtype=F64)
    assert torch.allclose(R.transpose(1, 2) @ R, I.expand_as(R), atol=1e-12)
    assert torch.allclose(torch.linalg.det(R), torch.ones(R.shape[0], dtype=F64), atol=1e-12)
    R_t2w, fv = OB.face_frame(v, f)
    e1, e2 = fv[:, 1] - fv[:, 0], fv[:, 2] - fv[:, 0]
    n = R_t2w[:, :, 2]
    assert (n * e1).sum(-1).abs().max() < 1e-12 and (n * e2).sum(-1).abs().max() < 1e-12
    assert torch.allclose(R_t2w[:, :, 0], e1 / e1.norm(dim=1, keepdim=True))
    # R(q) = R and the quaternion is standardised
    assert torch.allclose(OB.quat_to_rot(q), R, atol=1e-12) and (q[:, 0] >= 0).all()
    # means lie in their triangle with the given barycentrics
    G = 6
    tri = fv.repeat_interleave(G, 0)
    rec = torch.linalg.lstsq(tri.transpose(1, 2), xyz.unsqueeze(-1)).solution.squeeze(-1)
    assert torch.allclose(rec, bc, atol=1e-9)
    assert torch.allclose(scale[:, :2], torch.exp(scale2d)) and (scale[:, 2] == 1e-4).all()


def test_matrix_to_quaternion_branches():
    # rotations by pi about each axis exercise the four candidate branches
    for axis in range(3):
        R = -torch.eye(3, dtype=F64)
        R[axis, axis] = 1.0
        q = OB.matrix_to_quaternion(R[None])[0]
        assert torch.allclose(OB.quat_to_rot(q[None])[0], R, atol=1e-12)
        assert abs(abs(q[axis + 1].item()) - 1.0) < 1e-12
    q = OB.matrix_to_quaternion(torch.eye(3, dtype=F64)[None])[0]
    assert torch.allclose(q, torch.tensor([1.0, 0, 0, 0], dtype=F64))


def test_stage2_equilateral_and_eigen():
    s3 = 3 ** 0.5
    v = torch.tensor([[0.0, 0, 0], [1.0, 0, 0], [0.5, s3 / 2, 0]], dtype=F64)
    f = torch.tensor([[0, 1, 2]])
    bc, rad = PB.barycentric_v2(6)
    L = OB.triangle_L(v, f, rad, 1e-6)
    # the equilateral triangle is the reference shape: M = I
    assert torch.allclose(L[0, :2, :2], torch.eye(2, dtype=F64) * rad, atol=1e-12)
    k = torch.tensor(1.3, dtype=F64)
    xyz, cov, L = OB.stage2(v, f, bc.to(F64), 6, k, rad, 1e-6)
    S = torch.tensor([[cov[0, 0], cov[0, 1], cov[0, 2]], [cov[0, 1], cov[0, 3], cov[0, 4]],
                      [cov[0, 2], cov[0, 4], cov[0, 5]]])
    ev = torch.linalg.eigvalsh(S)
    exp = torch.tensor(sorted([(1e-6 * 1.3) ** 2, (rad * 1.3) ** 2, (rad * 1.3) ** 2]), dtype=F64)
    assert torch.allclose(ev, exp, rtol=1e-9, atol=1e-20)
    assert torch.allclose(cov, cov[:1].expand_as(cov))     # shared by the face's Gaussians


def test_binding_autograd_matches_finite_differences():
    v, f, bc, rot2d, scale2d = _mesh(sub=1, seed=3)
    g = torch.Generator().manual_seed(4)
    n = rot2d.shape[0]
    wx, wq, ws = (torch.randn(n, 3, generator=g, dtype=F64), torch.randn(n, 4, generator=g, dtype=F64),
                  torch.randn(n, 3, generator=g, dtype=F64))
    leaves = [v.clone().requires_grad_(True), rot2d.clone().requires_grad_(True),
              scale2d.clone().requires_grad_(True)]

    def f_():
        xyz, q, s, _ = OB.stage3(leaves[0], f, bc, leaves[1], leaves[2], 1e-4)
        return (xyz * wx).sum() + (q * wq).sum() + (s * ws).sum()

    grads = torch.autograd.grad(f_(), leaves)
    rng = np.random.default_rng(0)
    h = 1e-6
    with torch.no_grad():
        for li, leaf in enumerate(leaves):
            flat = leaf.view(-1)
            for j in rng.choice(flat.numel(), size=8, replace=False):
                old = flat[j].item()
                flat[j] = old + h
                fp = f_().item()
                flat[j] = old - h
                fm = f_().item()
                flat[j] = old
                fd = (fp - fm) / (2 * h)
                assert abs(fd - grads[li].reshape(-1)[j].item()) <= 1e-6 * max(1.0, abs(fd))


def test_in_frustum_matches_definition():
    from oracle import camera as OC
    cam = OC.front_camera(64, 64)
    v, f = S.icosphere(2)
    v = v * 3.0
    m = OB.in_frustum(cam.full_proj_transform.double(), v.double(), f)
    assert 0 < m.sum() < f.shape[0]
