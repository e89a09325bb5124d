"""PLY scenes for the batched path and the server.

File format of scene/gaussian_model.py:176-256 (written there through
plyfile): ``binary_little_endian`` with one ``vertex`` element of float32
properties ``x y z nx ny nz f_dc_0..2 f_rest_0..(3m-1) opacity scale_0..2
rot_0..3``, where the SH rest coefficients are stored colour-major
(``f_rest_{c*m + k-1}`` is coefficient k of colour c, m = (deg+1)^2 - 1).
Properties are found by name, as the reference's ``load_ply`` does, so other
orders and extra properties load too.

The vertex block is memory-mapped, copied to the device as-is and turned
into the 59 component-major channels by one transpose kernel
(``dmgs_ply_rows_to_params``); writing runs the reverse kernel and streams
the rows back.  The header is plain text handled here.
"""
import ctypes
import math

import numpy as np
import torch

from . import _native as N
from .batched import CHANNELS, unpack_grads

_SH_BASE = 11


def _channel_of(name, m):
    """Scene channel of a PLY property name (None: not a scene channel)."""
    fixed = {"x": 0, "y": 1, "z": 2, "opacity": 10}
    if name in fixed:
        return fixed[name]
    for prefix, base, count in (("scale_", 3, 3), ("rot_", 6, 4), ("f_dc_", None, 3), ("f_rest_", None, 3 * m)):
        if name.startswith(prefix):
            try:
                i = int(name[len(prefix):])
            except ValueError:
                return None
            if not 0 <= i < count:
                return None
            if prefix == "f_dc_":
                return _SH_BASE + i  # coefficient 0 of colour i
            if prefix == "f_rest_":
                c, k = divmod(i, m)
                return _SH_BASE + 3 * (k + 1) + c
            return base + i
    return None


def _sh_rest_per_colour(names):
    rest = sum(1 for n in names if n.startswith("f_rest_"))
    if rest % 3:
        raise ValueError(f"{rest} f_rest properties is not a multiple of 3")
    m = rest // 3
    deg = int(round(math.sqrt(m + 1))) - 1
    if (deg + 1) ** 2 - 1 != m or deg > 3:
        raise ValueError(f"{m} SH rest coefficients per colour is not (deg+1)^2 - 1 for a degree <= 3")
    return m, deg


def read_header(f):
    """Parse a PLY header from a binary file object; returns (vertex count,
    property names, byte offset of the data)."""
    first = f.readline()
    if first.strip() != b"ply":
        raise ValueError("not a PLY file")
    count, names, in_vertex, fmt = None, [], False, None
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header has no end_header")
        tok = line.decode("ascii", "replace").split()
        if not tok or tok[0] in ("comment", "obj_info"):
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            in_vertex = tok[1] == "vertex"
            if in_vertex:
                count = int(tok[2])
            elif count is not None:
                raise ValueError("elements after 'vertex' are not supported")
        elif tok[0] == "property" and in_vertex:
            if tok[1] == "list":
                raise ValueError("list properties are not supported in the vertex element")
            if tok[1] not in ("float", "float32"):
                raise ValueError(f"vertex property {tok[-1]} is {tok[1]}; only float32 is supported")
            names.append(tok[2])
        elif tok[0] == "end_header":
            break
    if fmt != "binary_little_endian":
        raise ValueError(f"PLY format {fmt} is not supported (binary_little_endian only)")
    if count is None:
        raise ValueError("PLY file has no vertex element")
    return count, names, f.tell()


def header_bytes(count, names):
    lines = ["ply", "format binary_little_endian 1.0", f"element vertex {count}"]
    lines += [f"property float {n}" for n in names]
    lines.append("end_header")
    return ("\n".join(lines) + "\n").encode("ascii")


def attribute_names(sh_degree=3):
    """construct_list_of_attributes (scene/gaussian_model.py:176-189)."""
    m = (sh_degree + 1) ** 2 - 1
    return (["x", "y", "z", "nx", "ny", "nz"] + [f"f_dc_{i}" for i in range(3)] +
            [f"f_rest_{i}" for i in range(3 * m)] + ["opacity"] + [f"scale_{i}" for i in range(3)] +
            [f"rot_{i}" for i in range(4)])


def load_params(path, device="cuda", stream=None):
    """Read a PLY scene into a (59, P) float32 component-major tensor on
    ``device``.  Returns (params, sh_degree)."""
    lib = N.load()
    with open(path, "rb") as f:
        P, names, offset = read_header(f)
    m, deg = _sh_rest_per_colour(names)
    cols = [-1] * CHANNELS
    for j, n in enumerate(names):
        c = _channel_of(n, m)
        if c is not None:
            cols[c] = j
    for need in ("x", "y", "z", "opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3",
                 "f_dc_0", "f_dc_1", "f_dc_2"):
        if need not in names:
            raise ValueError(f"PLY vertex element lacks property {need!r}")
    R = len(names)
    params = torch.empty(CHANNELS, P, dtype=torch.float32, device=device)
    if P == 0:
        return params, deg
    rows = np.memmap(path, dtype="<f4", mode="c", offset=offset, shape=(P, R))  # copy-on-write: no host copy
    rows_dev = torch.from_numpy(np.ascontiguousarray(rows)).to(device)
    cmap = (ctypes.c_int32 * CHANNELS)(*cols)
    N.check(lib.dmgs_ply_rows_to_params(N.ptr(rows_dev), P, R, cmap, N.ptr(params),
                                        stream if stream is not None else N.stream_handle(params.device)),
            "dmgs_ply_rows_to_params")
    return params, deg


def save_params(path, params, sh_degree=3, stream=None):
    """Write (59, P) component-major parameters as the reference's PLY
    (normals zero, ``sh_degree`` rest coefficients per colour)."""
    lib = N.load()
    if params.dtype != torch.float32 or params.dim() != 2 or params.shape[0] != CHANNELS:
        raise ValueError(f"params must be a float32 ({CHANNELS}, P) tensor")
    params = params.contiguous()
    P = params.shape[1]
    names = attribute_names(sh_degree)
    m = (sh_degree + 1) ** 2 - 1
    chans = [_channel_of(n, m) for n in names]
    chans = [-1 if c is None else c for c in chans]
    R = len(names)
    rows = torch.empty(P, R, dtype=torch.float32, device=params.device)
    if P:
        cmap = (ctypes.c_int32 * R)(*chans)
        N.check(lib.dmgs_params_to_ply_rows(N.ptr(params), P, R, cmap, N.ptr(rows),
                                            stream if stream is not None else N.stream_handle(params.device)),
                "dmgs_params_to_ply_rows")
    host = rows.cpu().numpy()
    with open(path, "wb") as f:
        f.write(header_bytes(P, names))
        f.write(np.ascontiguousarray(host, dtype="<f4").tobytes())


def load_model(path, device="cuda"):
    """A FreeGaussians scene (raw parameters, as GaussianModel.load_ply
    leaves them) from a PLY file."""
    from .scenes import FreeGaussians
    params, deg = load_params(path, device)
    g = unpack_grads(params)
    m = (deg + 1) ** 2
    rest = g["_features_rest"][:, :m - 1].contiguous()
    return FreeGaussians(g["_xyz"].contiguous(), g["_features_dc"].contiguous(), rest, g["_scaling"].contiguous(),
                         g["_rotation"].contiguous(), g["_opacity"].contiguous(), sh_degree=deg)
