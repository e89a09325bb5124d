"""Viewer wire format in front of the render server.

The reference's viewer link (gaussian_renderer/network_gui.py:26-85, driven
by train.py:52-65) is one TCP connection served between training steps, one
blocking request at a time:

* request: 4-byte little-endian length, then that many bytes of UTF-8 JSON
  with ``resolution_x/y``, ``fov_x/y``, ``z_near/far``, ``view_matrix`` and
  ``view_projection_matrix`` (16 floats each, column 1 and 2 negated on
  receipt), ``train``, ``shs_python``, ``rot_scale_python``, ``keep_alive``,
  ``scaling_modifier``; a zero resolution asks for no image;
* reply: the image as H x W x 3 bytes of uint8(clamp(x, 0, 1) * 255) (absent
  for a zero resolution), then a 4-byte little-endian length and that many
  ASCII bytes of a verification string (the dataset's source path).

``WireServer`` speaks the same bytes to any number of concurrent
connections and feeds every request into ``RenderServer``'s continuous
batching, so clients of different scenes and resolutions share batches; the
float image is turned into bytes on the device (``dmgs_image_rgb8``) and
only 3 bytes per pixel cross PCIe.
"""
import json
import selectors
import socket
import struct

import torch

from . import _native as N
from .cameras import MiniCam
from .serving import Request


def recv_exact(conn, n):
    """Read exactly n bytes (a short read ends the connection)."""
    buf = bytearray()
    while len(buf) < n:
        chunk = conn.recv(n - len(buf))
        if not chunk:
            raise ConnectionError("connection closed")
        buf += chunk
    return bytes(buf)


def encode_message(obj):
    """A request as the viewer sends it (length-prefixed JSON)."""
    data = json.dumps(obj).encode("utf-8")
    return struct.pack("<I", len(data)) + data


def decode_request(message, device="cuda"):
    """network_gui.receive() on a decoded JSON message: returns (camera,
    do_training, shs_python, rot_scale_python, keep_alive, scaling_modifier),
    all None for a zero resolution."""
    width, height = int(message["resolution_x"]), int(message["resolution_y"])
    if width == 0 or height == 0:
        return None, None, None, None, None, None
    wv = torch.tensor(message["view_matrix"], dtype=torch.float32).reshape(4, 4)
    wv[:, 1] = -wv[:, 1]
    wv[:, 2] = -wv[:, 2]
    full = torch.tensor(message["view_projection_matrix"], dtype=torch.float32).reshape(4, 4)
    full[:, 1] = -full[:, 1]
    full[:, 2] = -full[:, 2]
    cam = MiniCam(width, height, message["fov_y"], message["fov_x"], message["z_near"], message["z_far"],
                  wv.to(device), full.to(device))
    return (cam, bool(message["train"]), bool(message["shs_python"]), bool(message["rot_scale_python"]),
            bool(message["keep_alive"]), message["scaling_modifier"])


def encode_reply(image_bytes, verify):
    """network_gui.send(): image bytes (or nothing), 4-byte length, ASCII verify string."""
    v = verify.encode("ascii")
    return (image_bytes or b"") + struct.pack("<I", len(v)) + v


def image_rgb8(image, stream=None):
    """(3, H, W) float32 device image -> (H, W, 3) uint8 device tensor of
    uint8(clamp(x, 0, 1) * 255) (train.py:60), one kernel."""
    if image.dtype != torch.float32 or image.dim() != 3 or image.shape[0] != 3:
        raise ValueError("image must be a float32 (3, H, W) tensor")
    image = image.contiguous()
    H, W = image.shape[1], image.shape[2]
    out = torch.empty(H, W, 3, dtype=torch.uint8, device=image.device)
    N.check(N.load().dmgs_image_rgb8(N.ptr(image), H, W, N.ptr(out),
                                     stream if stream is not None else N.stream_handle(image.device)),
            "dmgs_image_rgb8")
    return out


class _Conn:
    def __init__(self, sock, scene):
        self.sock, self.scene = sock, scene
        self.inbuf = bytearray()
        self.outbuf = bytearray()
        self.waiting = []  # requests in flight, in arrival order (replies keep that order)


class WireServer:
    """Viewer-protocol front end of a ``RenderServer``.

    Every accepted connection is bound to one resident scene (``scene_for``
    picks it from the peer address; default: the first scene).  ``step()``
    accepts, reads complete requests, submits them, advances the render
    server and writes finished replies; replies on one connection keep the
    order of its requests."""

    def __init__(self, render_server, host="127.0.0.1", port=0, verify="", background=(1.0, 1.0, 1.0),
                 scene_for=None):
        self.rs = render_server
        self.verify = verify
        self.bg = background
        self.scene_for = scene_for or (lambda addr: next(iter(self.rs.scenes)))
        self.listener = socket.socket(socket.AF_INET, socket.SOCK_STREAM)
        self.listener.setsockopt(socket.SOL_SOCKET, socket.SO_REUSEADDR, 1)
        self.listener.bind((host, port))
        self.listener.listen()
        self.listener.setblocking(False)
        self.address = self.listener.getsockname()
        self.sel = selectors.DefaultSelector()
        self.sel.register(self.listener, selectors.EVENT_READ, None)
        self.conns = []
        self.served = 0

    def close(self):
        for c in self.conns:
            self.sel.unregister(c.sock)
            c.sock.close()
        self.conns = []
        self.sel.unregister(self.listener)
        self.listener.close()

    def _accept(self):
        sock, addr = self.listener.accept()
        sock.setblocking(False)
        c = _Conn(sock, self.scene_for(addr))
        self.conns.append(c)
        self.sel.register(sock, selectors.EVENT_READ, c)

    def _drop(self, c):
        self.sel.unregister(c.sock)
        c.sock.close()
        self.conns.remove(c)

    def _read(self, c):
        try:
            data = c.sock.recv(1 << 16)
        except (BlockingIOError, InterruptedError):
            return
        except OSError:
            data = b""
        if not data:
            self._drop(c)
            return
        c.inbuf += data
        while len(c.inbuf) >= 4:
            n = struct.unpack_from("<I", c.inbuf)[0]
            if len(c.inbuf) < 4 + n:
                break
            msg = json.loads(bytes(c.inbuf[4:4 + n]).decode("utf-8"))
            del c.inbuf[:4 + n]
            cam, _, _, _, _, mod = decode_request(msg, self.rs.device)
            if cam is None:
                c.waiting.append(None)  # empty reply slot: verify string only
                continue
            c.waiting.append(self.rs.submit(Request(c.scene, cam, bg=self.bg, scaling_modifier=mod)))

    def _flush_ready(self, c):
        while c.waiting and (c.waiting[0] is None or c.waiting[0].status in ("done", "rejected")):
            q = c.waiting.pop(0)
            payload = None
            if q is not None and q.status == "done":
                payload = image_rgb8(q.image).cpu().numpy().tobytes()
            c.outbuf += encode_reply(payload, self.verify)
            self.served += 1
        if c.outbuf:
            try:
                sent = c.sock.send(c.outbuf)
                del c.outbuf[:sent]
            except (BlockingIOError, InterruptedError):
                pass

    def step(self, timeout=0.0):
        for key, _ in self.sel.select(timeout):
            if key.data is None:
                self._accept()
            else:
                self._read(key.data)
        self.rs.poll()
        self.rs.take_completed()
        for c in list(self.conns):
            self._flush_ready(c)

    def serve(self, until):
        """Run ``step`` until ``until()`` is true."""
        while not until():
            self.step(0.0005)
