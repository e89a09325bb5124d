"""GPU tests of the scene and frame conversions at the path's edges.

* PLY: a file written from device parameters has the bytes of the
  reference writer (scene/gaussian_model.py:191-210, restated with numpy
  from the raw attributes); reading it back gives the same parameters
  bit-exactly; a degree-1 file and a file with a different property order
  load with the same channel meaning; the loaded model renders the same
  image as the original.
* Frames: dmgs_image_rgb8 equals uint8(clamp(x, 0, 1) * 255) of train.py:60.
* Wire: concurrent viewer-protocol clients get, for every request in order,
  the reply bytes of their camera's render.
"""
import socket
import struct
import threading
import time

import numpy as np
import pytest
import torch

from dmgs_amd import batched as BT
from dmgs_amd import cameras as C
from dmgs_amd import ply, scenes as S, serving as SV, wire
from dmgs_amd.gaussian_renderer import PipelineParams, render

pytestmark = pytest.mark.gpu


def _reference_rows(model):
    """save_ply's attribute matrix (gaussian_model.py:196-206) in numpy."""
    P = model.num
    xyz = model._xyz.detach().cpu().numpy()
    f_dc = model._features_dc.detach().transpose(1, 2).reshape(P, -1).cpu().numpy()
    f_rest = model._features_rest.detach().transpose(1, 2).reshape(P, -1).cpu().numpy()
    return np.concatenate((xyz, np.zeros_like(xyz), f_dc, f_rest, model._opacity.detach().cpu().numpy(),
                           model._scaling.detach().cpu().numpy(), model._rotation.detach().cpu().numpy()),
                          axis=1).astype("<f4")


def test_ply_bytes_round_trip_and_render(tmp_path):
    print("case: PLY write/read, 5000 Gaussians", flush=True)
    model = S.free_scene(5000, seed=31, layout="mixture")
    params = BT.pack_model(model.to("cuda"))
    path = str(tmp_path / "scene.ply")
    ply.save_params(path, params, sh_degree=3)
    with open(path, "rb") as f:
        data = f.read()
    hb = ply.header_bytes(model.num, ply.attribute_names(3))
    assert data[:len(hb)] == hb
    assert data[len(hb):] == _reference_rows(model).tobytes()
    back, deg = ply.load_params(path)
    assert deg == 3 and torch.equal(back, params)
    m2 = ply.load_model(path)
    cam = C.camera_at((0.3, -0.5, 3.5), 160, 120, device="cuda")
    bg = torch.ones(3, device="cuda")
    a = render(cam, model.to("cuda"), PipelineParams(), bg)["render"]
    b = render(cam, m2, PipelineParams(), bg)["render"]
    assert torch.equal(a, b)


def test_ply_lower_degree_and_other_order(tmp_path):
    print("case: PLY degree 1, shuffled properties", flush=True)
    g = torch.Generator().manual_seed(4)
    P = 1000
    model = S.free_scene(P, seed=32, layout="uniform")
    rest1 = model._features_rest[:, :3].contiguous()  # degree 1: 3 rest coefficients
    m1 = S.FreeGaussians(model._xyz, model._features_dc, rest1, model._scaling, model._rotation, model._opacity,
                         sh_degree=1)
    names = ply.attribute_names(1)
    rows = _reference_rows(m1)
    perm = torch.randperm(len(names), generator=g).numpy()
    path = str(tmp_path / "deg1.ply")
    with open(path, "wb") as f:
        f.write(ply.header_bytes(P, [names[j] for j in perm]))
        f.write(np.ascontiguousarray(rows[:, perm]).tobytes())
    params, deg = ply.load_params(path)
    assert deg == 1
    expect = BT.pack_params(m1._xyz, m1._scaling, m1._rotation, m1._opacity, m1._features_dc, m1._features_rest).cuda()
    assert torch.equal(params, expect)
    assert torch.count_nonzero(params[11 + 3 * 4:]) == 0  # coefficients past degree 1 are zero


def test_rgb8_matches_reference_conversion():
    print("case: rgb8 frame conversion", flush=True)
    g = torch.Generator().manual_seed(6)
    img = (torch.rand(3, 37, 53, generator=g) * 1.4 - 0.2).cuda()
    img[0, 0, :5] = torch.tensor([0.0, 1.0, 1.0 / 255.0, 254.5 / 255.0, -0.0])
    out = wire.image_rgb8(img)
    ref = (torch.clamp(img, min=0, max=1.0) * 255).byte().permute(1, 2, 0).contiguous()
    assert torch.equal(out, ref)


def _client(addr, msgs, replies, errors):
    try:
        with socket.create_connection(addr, timeout=60) as s:
            for m in msgs:
                s.sendall(wire.encode_message(m))
            for m in msgs:
                w, h = m["resolution_x"], m["resolution_y"]
                img = wire.recv_exact(s, w * h * 3) if w and h else None
                n = struct.unpack("<I", wire.recv_exact(s, 4))[0]
                replies.append((img, wire.recv_exact(s, n).decode("ascii")))
    except Exception as e:  # reported by the main thread
        errors.append(repr(e))


def _wire_message(cam):
    wv = cam.world_view_transform.detach().cpu().clone()
    full = cam.full_proj_transform.detach().cpu().clone()
    for M in (wv, full):  # the viewer sends columns 1 and 2 negated
        M[:, 1] = -M[:, 1]
        M[:, 2] = -M[:, 2]
    return {"resolution_x": cam.image_width, "resolution_y": cam.image_height, "train": False,
            "fov_y": cam.FoVy, "fov_x": cam.FoVx, "z_near": 0.01, "z_far": 100.0, "shs_python": False,
            "rot_scale_python": False, "keep_alive": True, "scaling_modifier": 1.0,
            "view_matrix": wv.reshape(-1).tolist(), "view_projection_matrix": full.reshape(-1).tolist()}


def test_wire_server_concurrent_clients():
    print("case: wire server, 3 clients", flush=True)
    scene = BT.pack_model(S.free_scene(20000, seed=33, layout="mixture", shell_radius=3.0).to("cuda"))
    rs = SV.RenderServer({"s": scene}, max_views=4, num_streams=3)
    srv = wire.WireServer(rs, verify="dataset/path")
    pos = C.sphere_positions(6, 4.0, seed=8)
    sizes = [(96, 64), (128, 128), (200, 120)]
    cams = [C.camera_at(pos[k], *sizes[k % 3], device="cuda") for k in range(6)]
    per_client = [[cams[0], cams[1], None], [cams[2], cams[3]], [cams[4], cams[5]]]
    replies = [[] for _ in per_client]
    errors = []
    threads = []
    sent = []
    for k, cl in enumerate(per_client):
        msgs = [_wire_message(c) if c is not None else dict(_wire_message(cams[0]), resolution_x=0) for c in cl]
        sent.append(msgs)
        t = threading.Thread(target=_client, args=(srv.address, msgs, replies[k], errors), daemon=True)
        t.start()
        threads.append(t)
    total = sum(len(cl) for cl in per_client)
    t0 = time.time()
    srv.serve(lambda: srv.served >= total or time.time() - t0 > 60)
    for t in threads:
        t.join(timeout=30)
    srv.close()
    assert not errors, errors
    assert srv.served == total
    for msgs, rep in zip(sent, replies):
        assert len(rep) == len(msgs)
        for m, (img, verify) in zip(msgs, rep):
            assert verify == "dataset/path"
            cam = wire.decode_request(m)[0]  # the camera the server built from the message
            if cam is None:
                assert img is None
                continue
            solo = SV.RenderServer({"s": scene}, max_views=1, num_streams=1).run([SV.Request("s", cam)])[0]
            expect = (torch.clamp(solo.image, min=0, max=1.0) * 255).byte().permute(1, 2, 0).contiguous()
            assert img == expect.cpu().numpy().tobytes()
