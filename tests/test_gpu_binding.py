"""GPU parity of the mesh binding kernels against the FP64 oracle, alone and
chained through the rasterizer into the mesh vertices."""
import pytest
import torch

from dmgs_amd import binding as PB
from dmgs_amd import cameras as C
from dmgs_amd import scenes as S
from dmgs_amd.gaussian_renderer import PipelineParams, render
from oracle import binding as OB
from oracle import raster as OR
from tests.helpers import grad_rel_err, oracle_settings, settings_from_cam

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _mesh(sub, seed, G=6, per_gaussian=True):
    g = torch.Generator().manual_seed(seed)
    v, f = S.icosphere(sub)
    v = v + torch.randn(v.shape, generator=g) * 0.02
    n = f.shape[0] * G
    if per_gaussian:
        e = -torch.log(torch.rand(n, 3, generator=g))
        bc = e / e.sum(1, keepdim=True)
    else:
        bc = PB.barycentric_v2(G)[0]
    rot2d = torch.randn(n, 2, generator=g)
    scale2d = torch.randn(n, 2, generator=g) * 0.3 - 3
    return v, f, bc, rot2d, scale2d


@pytest.mark.parametrize("per_gaussian,sub", [(True, 3), (False, 3), (True, 6)])
def test_stage3_forward_backward(per_gaussian, sub):
    """sub 6 is the mesh-bound workload's size: 81 920 faces x 6 Gaussians."""
    print(f"case: bind stage3 per_gaussian={per_gaussian} subdivisions={sub}", flush=True)
    v, f, bc, rot2d, scale2d = _mesh(sub, 1, per_gaussian=per_gaussian)
    n = rot2d.shape[0]
    ov, orot, osc = [t.double().requires_grad_(True) for t in (v, rot2d, scale2d)]
    xyz_o, q_o, s_o, R_o = OB.stage3(ov, f, bc.double(), orot, osc, 1e-4)
    gv, grot, gsc = [t.cuda().requires_grad_(True) for t in (v, rot2d, scale2d)]
    xyz, q, s = PB.bind_stage3(gv, f.cuda(), bc.cuda(), grot, gsc, 1e-4)
    assert (xyz.cpu().double() - xyz_o).abs().max() < 1e-5
    assert (s.cpu().double() - s_o).abs().max() < 1e-6
    # the sign of q is a convention; R(q) is what the rasterizer uses
    qq = q.detach().cpu().double()
    assert (OB.quat_to_rot(qq) - R_o).abs().max() < 1e-5
    sign = torch.sign((qq * q_o).sum(1, keepdim=True))
    assert (qq - sign * q_o.detach()).abs().max() < 1e-5
    g = torch.Generator().manual_seed(2)
    wx, wq, ws = torch.randn(n, 3, generator=g), torch.randn(n, 4, generator=g), torch.randn(n, 3, generator=g)
    ref = torch.autograd.grad((xyz_o * wx).sum() + (q_o * (sign * wq)).sum() + (s_o * ws).sum(), [ov, orot, osc])
    got = torch.autograd.grad((xyz * wx.cuda()).sum() + (q * wq.cuda()).sum() + (s * ws.cuda()).sum(),
                              [gv, grot, gsc])
    for name, a, b in zip(("verts", "rotation", "scaling"), got, ref):
        e = grad_rel_err(a, b)
        print(f"  {name}: rel err {e:.2e}")
        assert e < 1e-3


def test_stage2_forward_backward():
    print("case: bind stage2", flush=True)
    v, f, _, _, _ = _mesh(3, 5)
    bc, rad = PB.barycentric_v2(6)
    sf = torch.tensor([0.4])
    ov, osf = v.double().requires_grad_(True), sf.double().requires_grad_(True)
    k_o = torch.tanh(osf.reshape(())) * 2.0
    xyz_o, cov_o, L_o = OB.stage2(ov, f, bc.double(), 6, k_o, rad, 2.5e-6)
    gv, gsf = v.cuda().requires_grad_(True), sf.cuda().requires_grad_(True)
    xyz, cov, L = PB.bind_stage2(gv, f.cuda(), bc.cuda(), gsf, rad, max_scale=2.0, spatial_lr_scale=2.5)
    assert (xyz.cpu().double() - xyz_o).abs().max() < 1e-5
    assert grad_rel_err(cov.detach(), cov_o.detach()) < 1e-5
    assert grad_rel_err(L, L_o) < 1e-5
    g = torch.Generator().manual_seed(3)
    wx, wc = torch.randn(xyz.shape, generator=g), torch.randn(cov.shape, generator=g) * 1e4
    ref = torch.autograd.grad((xyz_o * wx).sum() + (cov_o * wc).sum(), [ov, osf])
    got = torch.autograd.grad((xyz * wx.cuda()).sum() + (cov * wc.cuda()).sum(), [gv, gsf])
    for name, a, b in zip(("verts", "scale_factor"), got, ref):
        e = grad_rel_err(a, b)
        print(f"  {name}: rel err {e:.2e}")
        assert e < 1e-3


def test_faces_in_frustum():
    print("case: faces in frustum", flush=True)
    cam = C.front_camera(64, 64, device="cpu")
    v, f = S.icosphere(3)
    v = v * 3.0
    ref = OB.in_frustum(cam.full_proj_transform.double(), v.double(), f)
    got = PB.faces_in_frustum(v.cuda(), f.cuda(), cam.full_proj_transform.cuda()).cpu()
    assert (got == ref).float().mean() > 0.999


def test_mesh_bound_render_into_vertices():
    """Binding -> drop-in render -> backward into the vertices and SH,
    against the oracle chain (oracle rasterizer gradients pushed through the
    FP64 binding by autograd)."""
    print("case: mesh-bound render", flush=True)
    W, H = 120, 96
    model = PB.mesh_scene(subdivisions=3, seed=0)
    cam = C.front_camera(W, H, device="cpu")
    bg = torch.ones(3)
    g = torch.Generator().manual_seed(5)
    dL = torch.randn(3, H, W, generator=g)
    # oracle chain
    v = model.verts.double().requires_grad_(True)
    xyz_o, q_o, s_o, _ = OB.stage3(v, model.faces, model.bc_coords.double(), model._rotation.double(),
                                   model._scaling.double(), model.thin_z_scale)
    feats = torch.cat([model._features_dc, model._features_rest], 1)
    opac = torch.sigmoid(model._opacity)
    rs = settings_from_cam(cam, bg, 3, device="cpu")
    ors = oracle_settings(rs)
    ocol, _, ost = OR.rasterize(xyz_o.detach(), opac, ors, shs=feats, scales=s_o.detach(), rotations=q_o.detach())
    dL[:, ost["ambiguous"]] = 0
    og = OR.rasterize_backward(ost, ors, xyz_o.detach(), dL.double(), shs=feats, scales=s_o.detach(),
                               rotations=q_o.detach())
    (dv_ref,) = torch.autograd.grad([xyz_o, q_o], [v], [og[0], og[6]])   # scales do not depend on verts
    # device chain
    gm = PB.mesh_scene(subdivisions=3, seed=0, device="cuda").requires_grad_(True)
    gm.renew_gaussian()
    out = render(C.front_camera(W, H, device="cuda"), gm, PipelineParams(), bg.cuda())
    ok = ~ost["ambiguous"]
    err = (out["render"].detach().cpu().double() - ocol).abs().amax(0)[ok].max().item()
    print(f"  image max err {err:.2e}")
    assert err < 1e-4
    out["render"].backward(dL.cuda())
    e_v = grad_rel_err(gm.verts.grad, dv_ref)
    e_sh = grad_rel_err(gm._features_rest.grad, og[2][:, 1:, :])
    print(f"  verts grad rel err {e_v:.2e}, SH grad rel err {e_sh:.2e}")
    assert e_v < 1e-3 and e_sh < 1e-3


def _oracle_python_colors(shs, xyz, campos, deg):
    """sigmoid(eval_sh) colours of the convert_SHs_python path
    (gaussian_renderer/__init__.py:72-78) in FP64."""
    from oracle import sh as OS
    dirs = xyz - campos.reshape(1, 3)
    dirs = dirs / dirs.norm(dim=1, keepdim=True)
    return torch.sigmoid(OS.eval_sh(deg, shs.transpose(1, 2), dirs))


@pytest.mark.parametrize("python_sh", [False, True])
def test_mesh_bound_full_size(python_sh):
    """The mesh-bound workload at full size: 81 920 faces x 6 Gaussians,
    one 800x800 view, drop-in render (kernel SH, or sigmoid(eval_sh) colours
    as the stage-3 training script sets convert_SHs_python,
    train_geo_stage3.py:260-261) + backward into the vertices, rotations,
    scales, SH and opacities.  Lists bit-exact on every tile, image and
    gradients on sampled tiles against the FP64 oracle chain."""
    import numpy as np
    from dmgs_amd import rasterizer as RZ
    from dmgs_amd.gaussian_renderer import _python_colors
    from tests import sampled as SP
    print(f"case: mesh-bound 81920 faces x 6, 800x800, python_sh={python_sh}", flush=True)
    W = H = 800
    bg = torch.ones(3)
    cam = C.front_camera(W, H, device="cpu")
    gm = PB.mesh_scene(subdivisions=6, seed=0, device="cuda").requires_grad_(True)
    gm.renew_gaussian()
    pipe = PipelineParams(convert_SHs_python=python_sh)
    # the rasterizer inputs the drop-in render builds, for the staged oracle
    with torch.no_grad():
        xyz, opac, scl, rot = gm.get_xyz, gm.get_opacity, gm.get_scaling, gm.get_rotation
        feats = gm.get_features
        cols = None
        if python_sh:
            cols = _python_colors(feats.transpose(1, 2).view(-1, 3, 16), xyz, C.front_camera(W, H), 3)
        rs = settings_from_cam(cam, bg, 3)
        color, radii, state = RZ.forward_raw(xyz, opac, None if python_sh else feats, cols, scl, rot, None, rs)
    sp = SP.single_view_splats(state, W, H)
    sp["radii"] = radii.cpu().numpy()
    keys, vals, ranges = OR.bin_gaussians(sp["xy"], sp["radii"], sp["depth"], W, H)
    assert np.array_equal(sp["values"].astype(np.int64), vals) and np.array_equal(sp["ranges"], ranges)
    print(f"  lists bit-exact: {keys.size} instances", flush=True)
    tiles = SP.pick_tiles(ranges[:, 1] - ranges[:, 0], 200, seed=3)
    amb, _ = SP.staged_forward(sp, ranges, vals, bg, W, H, tiles, color, "c1")
    dL = SP.masked_dL(tiles, amb, W, H, seed=11)
    out = render(C.front_camera(W, H), gm, pipe, bg.cuda())
    assert torch.equal(out["render"].detach(), color)
    out["render"].backward(dL.cuda())
    # oracle chain from the raw parameters in FP64
    leaves = dict(verts=gm.verts, _rotation=gm._rotation, _scaling=gm._scaling, _features_dc=gm._features_dc,
                  _features_rest=gm._features_rest, _opacity=gm._opacity)
    L = {k: v.detach().cpu().double().requires_grad_(True) for k, v in leaves.items()}
    xyz_o, q_o, s_o, _ = OB.stage3(L["verts"], gm.faces.cpu(), gm.bc_coords.cpu().double(), L["_rotation"],
                                   L["_scaling"], gm.thin_z_scale)
    shs_o = torch.cat([L["_features_dc"], L["_features_rest"]], 1)
    opac_o = torch.sigmoid(L["_opacity"])
    ors = oracle_settings(rs)
    kw = dict(scales=s_o.detach(), rotations=q_o.detach())
    if python_sh:
        cols_o = _oracle_python_colors(shs_o, xyz_o, ors.campos.double(), 3)
        kw["colors_precomp"] = cols_o.detach()
    else:
        kw["shs"] = shs_o.detach()
    pre = OR.preprocess(xyz_o.detach(), opac_o.detach(), ors, **kw)
    SP.check_preprocess(sp, pre, "c1")
    dm2, dcon, dop, dcol = SP.staged_backward(sp, ranges, vals, bg, W, H, dL, tiles)
    kw.pop("colors_precomp", None)
    g = SP.oracle_geometry(pre, ors, xyz_o.detach(), dm2, dcon, dcol, sp["radii"], **kw)
    outs = [xyz_o, q_o, s_o, opac_o]
    grads = [g["dmeans3D"], g["drotations"], g["dscales"], dop.reshape(opac_o.shape)]
    if python_sh:
        outs.append(cols_o)
        grads.append(dcol)
    else:
        outs.append(shs_o)
        grads.append(g["dsh"])
    torch.autograd.backward(outs, grads)
    # the in-plane scales are equal and R_t = I, so turning a Gaussian about
    # its face normal leaves its covariance unchanged: the rotation gradient
    # vanishes in exact arithmetic and both sides hold rounding noise only
    ref_scale = L["_scaling"].grad.norm().item()
    for name, gr in (("device", gm._rotation.grad), ("oracle", L["_rotation"].grad)):
        frac = gr.double().norm().item() / ref_scale
        print(f"  c1 _rotation ({name}): |grad| {frac:.2e} of the scaling gradient's", flush=True)
        assert frac < 1e-5
    SP.assert_grads([(k, getattr(gm, k).grad, L[k].grad) for k in L if k != "_rotation"], "c1")


def test_stage2_render_dyn():
    """Stage-2 training path: bind_stage2 -> render_dyn with precomputed
    covariances and sigmoid(eval_sh) colours (compute_cov3D_python and
    convert_SHs_python, train_geo_stage2.py:101,268-269) -> gradients into the
    vertices, the scale factor, the SH features and the opacities, against
    the FP64 binding and rasterizer oracles on every tile."""
    from dmgs_amd.gaussian_renderer import render_dyn
    print("case: stage-2 binding + render_dyn", flush=True)
    W, H = 160, 120
    v, f, _, _, _ = _mesh(3, 9)
    bc, rad = PB.barycentric_v2(6)
    n = f.shape[0] * 6
    g = torch.Generator().manual_seed(13)
    sf = torch.tensor([0.3])
    feats = torch.randn(n, 3, 16, generator=g) * 0.3
    opl = torch.randn(n, 1, generator=g)
    cam = C.front_camera(W, H, device="cpu")
    bg = torch.tensor([0.3, 0.6, 0.9])
    # oracle chain (FP64)
    L = dict(verts=v.double().requires_grad_(True), sf=sf.double().requires_grad_(True),
             feats=feats.double().requires_grad_(True), opl=opl.double().requires_grad_(True))
    k_o = torch.tanh(L["sf"].reshape(())) * 2.0
    xyz_o, cov_o, _ = OB.stage2(L["verts"], f, bc.double(), 6, k_o, rad, 2.5e-6)
    opac_o = torch.sigmoid(L["opl"])
    rs = settings_from_cam(cam, bg, 3, device="cpu")
    ors = oracle_settings(rs)
    cols_o = _oracle_python_colors(L["feats"].transpose(1, 2), xyz_o, ors.campos.double(), 3)
    ocol, _, ost = OR.rasterize(xyz_o.detach(), opac_o.detach(), ors, colors_precomp=cols_o.detach(),
                                cov3D_precomp=cov_o.detach())
    dL = torch.randn(3, H, W, generator=g)
    dL[:, ost["ambiguous"]] = 0
    og = OR.rasterize_backward(ost, ors, xyz_o.detach(), dL.double(), colors_precomp=cols_o.detach(),
                               cov3D_precomp=cov_o.detach())
    torch.autograd.backward([xyz_o, cols_o, opac_o, cov_o], [og[0], og[3], og[4].reshape(opac_o.shape), og[7]])
    # device: the product path
    gv, gsf = v.cuda().requires_grad_(True), sf.cuda().requires_grad_(True)
    gfe, gop = feats.cuda().requires_grad_(True), opl.cuda().requires_grad_(True)
    xyz, cov, _ = PB.bind_stage2(gv, f.cuda(), bc.cuda(), gsf, rad, max_scale=2.0, spatial_lr_scale=2.5)
    info = dict(xyz=xyz, opacity=torch.sigmoid(gop), covariance=cov, features=gfe, active_sh_degree=3,
                max_sh_degree=3)
    out = render_dyn(C.front_camera(W, H), info, PipelineParams(compute_cov3D_python=True, convert_SHs_python=True),
                     bg.cuda())
    ok = ~ost["ambiguous"]
    err = (out["render"].detach().cpu().double() - ocol).abs().amax(0)[ok].max().item()
    print(f"  image max err {err:.2e}", flush=True)
    assert err < 1e-4
    out["render"].backward(dL.cuda())
    for name, got, ref in (("verts", gv.grad, L["verts"].grad), ("scale_factor", gsf.grad, L["sf"].grad),
                           ("features", gfe.grad, L["feats"].grad), ("opacity", gop.grad, L["opl"].grad)):
        e = grad_rel_err(got, ref)
        print(f"  {name}: rel err {e:.2e}", flush=True)
        assert ref.abs().max() > 0 and e < 1e-3
