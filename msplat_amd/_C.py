"""The twelve functions of msplat's native module ``msplat._C``
(msplat/src/ext.cpp:14-25), with its argument order, return tuples, shapes
and dtypes, backed by this library's C ABI.

A msplat checkout that keeps its own Python files and swaps only the native
layer installs this module as ``msplat._C`` (for instance
``sys.modules["msplat._C"] = msplat_amd._C`` before ``import msplat``, or a
one-line ``msplat/_C.py`` that star-imports it).  Outputs are allocated
here with torch on the inputs' device and every kernel runs on torch's
current stream; camera gradients are computed only for cameras that require
them and are zero otherwise, as projectPointsBackward and EWAProjectBackward
return them (msplat/src/project_point.cu:202-208, ewa_project.cu:318).
"""

from __future__ import annotations

import torch

from . import _lib
from .alpha_blending import blend_backward, blend_forward
from .sort_gaussian import compute_gaussian_key as _keys
from .sort_gaussian import compute_tile_gaussian_range as _ranges


def _dev(t):
    return torch.cuda.device(t.device)


def project_point_forward(xyz, intr, extr, W, H, nearest, extent):  # project_point.cu:148-180
    x = _lib.f32("xyz", xyz)
    n = x.shape[0]
    uv = torch.empty((n, 2), dtype=torch.float32, device=x.device)
    depth = torch.empty((n, 1), dtype=torch.float32, device=x.device)
    with _dev(x):
        _lib.call("msplat_project_point_forward", _lib.ptr(x), n, _lib.ptr(_lib.f32("intr", intr)),
                  _lib.ptr(_lib.f32("extr", extr)), int(W), int(H), float(nearest), float(extent), _lib.ptr(uv),
                  _lib.ptr(depth), _lib.stream_ptr())
    return uv, depth


def project_point_backward(xyz, intr, extr, W, H, uv, depth, dL_duv, dL_ddepth):  # project_point.cu:182-228
    del W, H, uv
    x = _lib.f32("xyz", xyz)
    n = x.shape[0]
    dev = x.device
    dxyz = torch.empty((n, 3), dtype=torch.float32, device=dev)
    dintr = torch.zeros(4, dtype=torch.float32, device=dev)
    dextr = torch.zeros((3, 4), dtype=torch.float32, device=dev)
    with _dev(x):
        _lib.call("msplat_project_point_backward", _lib.ptr(x), n, _lib.ptr(_lib.f32("intr", intr)),
                  _lib.ptr(_lib.f32("extr", extr)), _lib.ptr(_lib.f32("depth", depth)),
                  _lib.ptr(_lib.f32("dL_duv", dL_duv)),
                  _lib.ptr(None if dL_ddepth is None else _lib.f32("dL_ddepth", dL_ddepth)), _lib.ptr(dxyz),
                  _lib.ptr(dintr) if intr.requires_grad else None, _lib.ptr(dextr) if extr.requires_grad else None,
                  _lib.stream_ptr())
    return dxyz, dintr, dextr


def compute_cov3d_forward(scales, uquats, visible):  # compute_cov3d.cu:149-171
    s = _lib.f32("scales", scales)
    n = s.shape[0]
    cov = torch.empty((n, 6), dtype=torch.float32, device=s.device)
    with _dev(s):
        _lib.call("msplat_compute_cov3d_forward", _lib.ptr(s), _lib.ptr(_lib.f32("uquats", uquats)),
                  _lib.ptr(_lib.u8("visible", visible)), n, _lib.ptr(cov), _lib.stream_ptr())
    return cov


def compute_cov3d_backward(scales, uquats, visible, dL_dcov3Ds):  # compute_cov3d.cu:173-199
    s = _lib.f32("scales", scales)
    n = s.shape[0]
    ds = torch.empty((n, 3), dtype=torch.float32, device=s.device)
    dq = torch.empty((n, 4), dtype=torch.float32, device=s.device)
    with _dev(s):
        _lib.call("msplat_compute_cov3d_backward", _lib.ptr(s), _lib.ptr(_lib.f32("uquats", uquats)),
                  _lib.ptr(_lib.u8("visible", visible)), _lib.ptr(_lib.f32("dL_dcov3Ds", dL_dcov3Ds)), n,
                  _lib.ptr(ds), _lib.ptr(dq), _lib.stream_ptr())
    return ds, dq


def ewa_project_forward(xyz, cov3d, intr, extr, uv, W, H, visible):  # ewa_project.cu:255-298
    x = _lib.f32("xyz", xyz)
    n = x.shape[0]
    dev = x.device
    conic = torch.empty((n, 3), dtype=torch.float32, device=dev)
    radius = torch.empty(n, dtype=torch.int32, device=dev)
    tiles = torch.empty(n, dtype=torch.int32, device=dev)
    with _dev(x):
        _lib.call("msplat_ewa_project_forward", _lib.ptr(x), _lib.ptr(_lib.f32("cov3d", cov3d)),
                  _lib.ptr(_lib.f32("intr", intr)), _lib.ptr(_lib.f32("extr", extr)), _lib.ptr(_lib.f32("uv", uv)),
                  _lib.ptr(_lib.u8("visible", visible)), n, int(W), int(H), _lib.ptr(conic), _lib.ptr(radius),
                  _lib.ptr(tiles), _lib.stream_ptr())
    return conic, radius, tiles


def ewa_project_backward(xyz, cov3d, intr, extr, radius, dL_dconic):  # ewa_project.cu:300-345
    x = _lib.f32("xyz", xyz)
    n = x.shape[0]
    dev = x.device
    dxyz = torch.empty((n, 3), dtype=torch.float32, device=dev)
    dcov = torch.empty((n, 6), dtype=torch.float32, device=dev)
    dintr = torch.zeros(4, dtype=torch.float32, device=dev)
    dextr = torch.zeros((3, 4), dtype=torch.float32, device=dev)
    with _dev(x):
        _lib.call("msplat_ewa_project_backward", _lib.ptr(x), _lib.ptr(_lib.f32("cov3d", cov3d)),
                  _lib.ptr(_lib.f32("intr", intr)), _lib.ptr(_lib.f32("extr", extr)),
                  _lib.ptr(_lib.i32("radius", radius)), _lib.ptr(_lib.f32("dL_dconic", dL_dconic)), n,
                  _lib.ptr(dxyz), _lib.ptr(dcov), _lib.ptr(dintr) if intr.requires_grad else None,
                  _lib.ptr(dextr) if extr.requires_grad else None, _lib.stream_ptr())
    return dxyz, dcov, dintr, dextr


def compute_gaussian_key(uv, depth, W, H, radius, tiles):  # sort_gaussian.cu:75-113
    return _keys(uv, depth, W, H, radius, tiles)


def compute_tile_gaussian_range(W, H, tiles, key_sorted):  # sort_gaussian.cu:115-142
    return _ranges(W, H, tiles, key_sorted)


def compute_sh_forward(shs, view_dirs, visible):  # compute_sh.cu:1696-1722
    s = _lib.f32("shs", shs)
    n, C, D = s.shape
    value = torch.empty((n, C), dtype=torch.float32, device=s.device)
    with _dev(s):
        _lib.call("msplat_compute_sh_forward", _lib.ptr(s), _lib.ptr(_lib.f32("view_dirs", view_dirs)),
                  _lib.ptr(_lib.u8("visible", visible)), n, C, D, _lib.ptr(value), _lib.stream_ptr())
    return value


def compute_sh_backward(shs, view_dirs, visible, dL_dvalue):  # compute_sh.cu:1724-1753
    s = _lib.f32("shs", shs)
    n, C, D = s.shape
    dshs = torch.empty((n, C, D), dtype=torch.float32, device=s.device)
    ddirs = torch.empty((n, 3), dtype=torch.float32, device=s.device)
    with _dev(s):
        _lib.call("msplat_compute_sh_backward", _lib.ptr(s), _lib.ptr(_lib.f32("view_dirs", view_dirs)),
                  _lib.ptr(_lib.u8("visible", visible)), _lib.ptr(_lib.f32("dL_dvalue", dL_dvalue)), n, C, D,
                  _lib.ptr(dshs), _lib.ptr(ddirs), _lib.stream_ptr())
    return dshs, ddirs


def alpha_blending_forward(uv, conic, opacity, feature, idx_sorted, tile_range, bg, W, H):  # alpha_blending.cu:249-394
    return blend_forward(uv, conic, opacity, feature, idx_sorted, tile_range, bg, W, H)


def alpha_blending_backward(uv, conic, opacity, feature, idx_sorted, tile_range, bg, W, H, final_T, ncontrib,
                            dL_drendered):  # alpha_blending.cu:397-573
    duv, dconic, dopac, dfeat = blend_backward(uv, conic, opacity, feature, idx_sorted, tile_range, bg, W, H, final_T,
                                               ncontrib, dL_drendered)
    return duv, dconic, dopac.view_as(opacity), dfeat
