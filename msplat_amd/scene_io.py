"""On-disk scenes: the binary PLY layout of 3D Gaussian splatting files.

The reference keeps its Gaussians in memory only (its tutorials write a
video, tutorials/gs_3d.py:232-242); a server that keeps scenes resident in
HBM needs them on disk.  This module reads and writes the layout that 3D
Gaussian-splatting tools exchange, so trained scenes from elsewhere load
directly:

  element vertex N, float32 properties, binary little endian:
    x y z                    centre
    nx ny nz                 unused normals (written as 0)
    f_dc_0 .. f_dc_{C-1}     SH coefficient 0 of each channel
    f_rest_0 ..              SH coefficients 1..D-1, channel-major:
                             f_rest_{c (D-1) + k - 1} = shs[:, c, k]
    feat_0 .. feat_{F-1}     plain feature channels (this library's
                             extension: features rendered as they are)
    opacity                  logit of the opacity
    scale_0 .. scale_2       log of the scale
    rot_0 .. rot_3           quaternion (w, x, y, z), not normalised

``load_ply`` returns the rasterizer's inputs: scale = exp, opacity =
sigmoid, rotation normalised (``activate=True``), or the stored parameters
as they are.  Files are read with one memory map and column copies, so a
multi-million-Gaussian scene loads at disk speed.
"""

from __future__ import annotations

import math
import os
import re

import numpy as np
import torch

_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2", "ushort": "u2",
    "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4", "float": "f4", "float32": "f4",
    "double": "f8", "float64": "f8",
}


def _header(f):
    """(vertex count, [(name, numpy type)], byte offset of the vertex data)."""
    if f.readline().strip() != b"ply":
        raise ValueError("not a PLY file")
    fmt = None
    n = None
    props: list[tuple[str, str]] = []
    elements = []
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header has no end_header")
        tok = line.decode("ascii", errors="replace").split()
        if not tok or tok[0] in ("comment", "obj_info"):
            continue
        if tok[0] == "end_header":
            break
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            elements.append(tok[1])
            if tok[1] == "vertex":
                n = int(tok[2])
        elif tok[0] == "property":
            if tok[1] == "list":
                raise ValueError("PLY list properties are not supported")
            if len(elements) == 1 and elements[0] == "vertex":
                if tok[1] not in _PLY_TYPES:
                    raise ValueError(f"PLY property type {tok[1]!r}")
                props.append((tok[2], _PLY_TYPES[tok[1]]))
    if fmt != "binary_little_endian":
        raise ValueError(f"PLY format {fmt!r}: only binary_little_endian is read")
    if not elements or elements[0] != "vertex" or n is None:
        raise ValueError("PLY file has no leading vertex element")
    return n, props, f.tell()


def _indexed(names, prefix):
    pat = re.compile(re.escape(prefix) + r"(\d+)$")
    idx = sorted(int(m.group(1)) for m in (pat.match(s) for s in names) if m)
    if idx != list(range(len(idx))):
        raise ValueError(f"PLY properties {prefix}*: indices {idx} are not 0..{len(idx) - 1}")
    return [f"{prefix}{i}" for i in idx]


def load_ply(path: str, device="cpu", activate: bool = True) -> dict:
    """Read a Gaussian scene.  Returns ``xyz [N,3]``, ``scale [N,3]``,
    ``rotate [N,4]``, ``opacity [N,1]`` and, as present, ``shs [N,C,D]``
    (with ``sh_degree``) and ``feature [N,F]``, fp32 on ``device``."""
    with open(path, "rb") as f:
        n, props, offset = _header(f)
    names = [p[0] for p in props]
    dtype = np.dtype([(name, "<" + t) for name, t in props])
    size = os.path.getsize(path)
    if offset + n * dtype.itemsize > size:
        raise ValueError(f"PLY file {path}: {n} vertices of {dtype.itemsize} bytes do not fit the file")
    data = np.memmap(path, dtype=dtype, mode="r", offset=offset, shape=(n,))

    def cols(fields):
        if not fields:
            return np.zeros((n, 0), np.float32)
        return np.stack([np.asarray(data[c], dtype=np.float32) for c in fields], axis=1)

    for need in ("x", "y", "z", "opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"):
        if need not in names:
            raise ValueError(f"PLY file {path}: no {need!r} property")
    out = {
        "xyz": cols(["x", "y", "z"]),
        "scale": cols(["scale_0", "scale_1", "scale_2"]),
        "rotate": cols(["rot_0", "rot_1", "rot_2", "rot_3"]),
        "opacity": cols(["opacity"]),
    }
    dc = _indexed(names, "f_dc_")
    rest = _indexed(names, "f_rest_")
    if dc:
        C = len(dc)
        if len(rest) % C:
            raise ValueError(f"PLY file {path}: {len(rest)} f_rest properties for {C} channels")
        D = 1 + len(rest) // C
        deg = int(round(math.sqrt(D))) - 1
        if (deg + 1) ** 2 != D:
            raise ValueError(f"PLY file {path}: {D} SH coefficients per channel is not a square")
        shs = np.empty((n, C, D), np.float32)
        shs[:, :, 0] = cols(dc)
        if D > 1:
            shs[:, :, 1:] = cols(rest).reshape(n, C, D - 1)
        out["shs"] = shs
        out["sh_degree"] = deg
    feat = _indexed(names, "feat_")
    if feat:
        out["feature"] = cols(feat)
    res = {}
    for k, v in out.items():
        res[k] = v if k == "sh_degree" else torch.from_numpy(np.ascontiguousarray(v))
    if activate:
        res["scale"] = torch.exp(res["scale"])
        res["opacity"] = torch.sigmoid(res["opacity"])
        res["rotate"] = torch.nn.functional.normalize(res["rotate"], dim=1)
    return {k: (v.to(device) if isinstance(v, torch.Tensor) else v) for k, v in res.items()}


def _logit(p: np.ndarray) -> np.ndarray:
    p = p.astype(np.float64)
    with np.errstate(divide="ignore"):
        return np.log(p) - np.log1p(-p)


def save_ply(path: str, xyz, scale, rotate, opacity, shs=None, feature=None, activated: bool = True) -> None:
    """Write a Gaussian scene.  With ``activated`` the inputs are the
    rasterizer's (scale > 0, opacity in (0, 1)) and the file stores log
    scale and logit opacity, as 3D Gaussian-splatting files do; otherwise
    they are written as given."""
    npf = lambda t: np.asarray(t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else t, np.float32)
    xyz, scale, rotate = npf(xyz), npf(scale), npf(rotate)
    opacity = npf(opacity).reshape(-1, 1)
    n = xyz.shape[0]
    if activated:
        scale = np.log(scale.astype(np.float64)).astype(np.float32)
        opacity = _logit(opacity).astype(np.float32)
    fields: list[tuple[str, np.ndarray]] = [("x", xyz[:, 0]), ("y", xyz[:, 1]), ("z", xyz[:, 2])]
    zeros = np.zeros(n, np.float32)
    fields += [("nx", zeros), ("ny", zeros), ("nz", zeros)]
    if shs is not None:
        s = npf(shs)
        C, D = s.shape[1], s.shape[2]
        fields += [(f"f_dc_{c}", s[:, c, 0]) for c in range(C)]
        rest = s[:, :, 1:].reshape(n, C * (D - 1))
        fields += [(f"f_rest_{j}", rest[:, j]) for j in range(C * (D - 1))]
    if feature is not None:
        fe = npf(feature).reshape(n, -1)
        fields += [(f"feat_{j}", fe[:, j]) for j in range(fe.shape[1])]
    fields += [("opacity", opacity[:, 0])]
    fields += [(f"scale_{j}", scale[:, j]) for j in range(3)]
    fields += [(f"rot_{j}", rotate[:, j]) for j in range(4)]
    arr = np.empty(n, dtype=np.dtype([(name, "<f4") for name, _ in fields]))
    for name, col in fields:
        arr[name] = col
    head = ["ply", "format binary_little_endian 1.0", f"element vertex {n}"]
    head += [f"property float {name}" for name, _ in fields]
    head += ["end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode("ascii"))
        arr.tofile(f)
