"""Host side of the viewer wire format and of PLY scene files (CPU only).

* Requests decode as gaussian_renderer/network_gui.py:receive() does: the
  camera matrices arrive with columns 1 and 2 negated, a zero resolution asks
  for no image, and the reply is image bytes + 4-byte length + verify string
  (network_gui.py:51-55).
* PLY attribute names and order are the reference writer's
  (scene/gaussian_model.py:176-189); property names map to the 59 scene
  channels with the colour-major SH rest layout; headers round-trip and
  unsupported files are refused with a message.
"""
import io
import struct

import numpy as np
import pytest
import torch

from dmgs_amd import ply, wire


def _message(w=64, h=48, seed=0):
    g = np.random.default_rng(seed)
    return {"resolution_x": w, "resolution_y": h, "train": True, "fov_y": 0.7, "fov_x": 0.9, "z_near": 0.01,
            "z_far": 100.0, "shs_python": False, "rot_scale_python": False, "keep_alive": True,
            "scaling_modifier": 1.0, "view_matrix": g.normal(size=16).tolist(),
            "view_projection_matrix": g.normal(size=16).tolist()}


def test_request_decodes_like_network_gui():
    msg = _message()
    raw = wire.encode_message(msg)
    assert struct.unpack("<I", raw[:4])[0] == len(raw) - 4
    cam, train, shs_py, rs_py, keep, mod = wire.decode_request(msg, device="cpu")
    wv = torch.tensor(msg["view_matrix"], dtype=torch.float32).reshape(4, 4)
    full = torch.tensor(msg["view_projection_matrix"], dtype=torch.float32).reshape(4, 4)
    sign = torch.tensor([1.0, -1.0, -1.0, 1.0])
    assert torch.equal(cam.world_view_transform, wv * sign)
    assert torch.equal(cam.full_proj_transform, full * sign)
    assert torch.equal(cam.camera_center, torch.inverse(wv * sign)[3][:3])
    assert (cam.image_width, cam.image_height, cam.FoVx, cam.FoVy) == (64, 48, 0.9, 0.7)
    assert (train, shs_py, rs_py, keep, mod) == (True, False, False, True, 1.0)
    assert wire.decode_request(dict(msg, resolution_x=0), device="cpu") == (None,) * 6


def test_reply_layout():
    assert wire.encode_reply(b"\x01\x02\x03", "src") == b"\x01\x02\x03" + struct.pack("<I", 3) + b"src"
    assert wire.encode_reply(None, "") == struct.pack("<I", 0)


def test_ply_attribute_order_and_channel_map():
    names = ply.attribute_names(3)
    assert len(names) == 62
    assert names[:9] == ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
    assert names[9] == "f_rest_0" and names[53] == "f_rest_44"
    assert names[54:] == ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    m = 15
    chans = [ply._channel_of(n, m) for n in names]
    assert chans[:3] == [0, 1, 2] and chans[3:6] == [None] * 3
    assert chans[6:9] == [11, 12, 13]  # DC of r, g, b
    # f_rest_{c*15 + k-1}: coefficient k of colour c -> 11 + 3k + c
    for c in range(3):
        for k in range(1, 16):
            assert ply._channel_of(f"f_rest_{c * 15 + k - 1}", m) == 11 + 3 * k + c
    assert chans[54:] == [10, 3, 4, 5, 6, 7, 8, 9]
    assert sorted(c for c in chans if c is not None) == list(range(59))
    assert len(ply.attribute_names(0)) == 17 and len(ply.attribute_names(1)) == 26


def test_ply_header_round_trip_and_refusals():
    names = ply.attribute_names(2)
    hb = ply.header_bytes(7, names)
    f = io.BytesIO(hb + b"\0" * (7 * 4 * len(names)))
    count, got, offset = ply.read_header(f)
    assert (count, got, offset) == (7, names, len(hb))
    assert ply._sh_rest_per_colour(got) == (8, 2)
    with pytest.raises(ValueError, match="binary_little_endian"):
        ply.read_header(io.BytesIO(hb.replace(b"binary_little_endian", b"ascii")))
    with pytest.raises(ValueError, match="float32"):
        ply.read_header(io.BytesIO(hb.replace(b"property float x", b"property double x")))
    with pytest.raises(ValueError, match="not a PLY"):
        ply.read_header(io.BytesIO(b"obj\n"))
    bad = ply.attribute_names(3)[:-5] + ["opacity"]  # 44 rest coefficients
    with pytest.raises(ValueError):
        ply._sh_rest_per_colour([n for n in bad if n != "f_rest_44"])


def test_scene_getters_reuse_values_only_without_autograd():
    """FreeGaussians getters under no_grad return the previous value while the
    parameters are unchanged and recompute after an in-place update; with
    autograd on they always build a fresh graph."""
    from dmgs_amd import scenes
    m = scenes.free_scene(64, seed=3)
    with torch.no_grad():
        a = m.get_features
        assert m.get_features is a
        s0 = m.get_scaling
        m._scaling.add_(1.0)  # in-place update: new version
        s1 = m.get_scaling
        assert s1 is not s0 and torch.allclose(s1, torch.exp(m._scaling))
        m._features_dc = m._features_dc.clone()  # replaced parameter
        assert m.get_features is not a
    m.requires_grad_(True)
    f1, f2 = m.get_features, m.get_features
    assert f1 is not f2 and f1.requires_grad
    assert torch.equal(m.get_opacity, torch.sigmoid(m._opacity))
