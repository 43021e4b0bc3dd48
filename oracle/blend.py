"""ctypes wrapper of the C blending oracle (test infrastructure only).

Restates msplat/src/alpha_blending.cu:16-246; see blend_body.h.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import build as _build

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_build.OUT):
            _build.build()
        lib = ctypes.CDLL(_build.OUT)
        P, I64, I32, F32 = ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_float
        for sfx in ("_f32", "_f64"):
            f = getattr(lib, "blend_forward" + sfx)
            f.restype = I64
            f.argtypes = [P, P, P, P, I64, I32, P, P, F32, I32, I32, P, P, P]
            b = getattr(lib, "blend_backward" + sfx)
            b.restype = None
            b.argtypes = [P, P, P, P, I64, I32, P, P, F32, I32, I32, P, P, P, P, P, P, P]
            fm = getattr(lib, "blend_forward_mt" + sfx)
            fm.restype = I64
            fm.argtypes = f.argtypes + [I32]
            bm = getattr(lib, "blend_backward_mt" + sfx)
            bm.restype = None
            bm.argtypes = b.argtypes + [I32]
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _prep(uv, conic, opacity, feature, idx, rng, dt):
    uv = np.ascontiguousarray(uv, dt)
    conic = np.ascontiguousarray(conic, dt)
    opacity = np.ascontiguousarray(np.asarray(opacity).reshape(-1), dt)
    feature = np.ascontiguousarray(feature, dt)
    idx = np.ascontiguousarray(idx, np.int32)
    rng = np.ascontiguousarray(rng, np.int32)
    return uv, conic, opacity, feature, idx, rng


def blend_forward(uv, conic, opacity, feature, idx_sorted, tile_range, bg, W, H, dtype=np.float32, threads=1):
    """Returns (rendered [C,H,W], final_T [H,W], ncontrib [H,W] int32, pairs evaluated).
    ``threads`` > 1 shares tile rows among OpenMP threads (same outputs)."""
    lib = _load()
    dt = np.dtype(dtype)
    uv, conic, opacity, feature, idx, rng = _prep(uv, conic, opacity, feature, idx_sorted, tile_range, dt)
    n, C = feature.shape
    out = np.zeros((C, H, W), dt)
    fT = np.zeros((H, W), dt)
    nc = np.zeros((H, W), np.int32)
    args = (_p(uv), _p(conic), _p(opacity), _p(feature), n, C, _p(idx), _p(rng), float(bg), W, H, _p(out), _p(fT),
            _p(nc))
    if threads > 1:
        fn = lib.blend_forward_mt_f32 if dt == np.float32 else lib.blend_forward_mt_f64
        ev = fn(*args, int(threads))
    else:
        fn = lib.blend_forward_f32 if dt == np.float32 else lib.blend_forward_f64
        ev = fn(*args)
    return out, fT, nc, int(ev)


def blend_backward(uv, conic, opacity, feature, idx_sorted, tile_range, bg, W, H, final_T, ncontrib, dout,
                   dtype=np.float32, threads=1):
    """Returns dL/duv [N,2], dL/dconic [N,3], dL/dopacity [N], dL/dfeature [N,C].
    ``threads`` > 1 shares tile rows among OpenMP threads, each with gradient
    arrays of its own summed afterwards (another summation order)."""
    lib = _load()
    dt = np.dtype(dtype)
    uv, conic, opacity, feature, idx, rng = _prep(uv, conic, opacity, feature, idx_sorted, tile_range, dt)
    n, C = feature.shape
    fT = np.ascontiguousarray(final_T, dt)
    nc = np.ascontiguousarray(ncontrib, np.int32)
    g = np.ascontiguousarray(dout, dt)
    duv = np.zeros((n, 2), dt)
    dco = np.zeros((n, 3), dt)
    dop = np.zeros(n, dt)
    dfe = np.zeros((n, C), dt)
    args = (_p(uv), _p(conic), _p(opacity), _p(feature), n, C, _p(idx), _p(rng), float(bg), W, H, _p(fT), _p(nc),
            _p(g), _p(duv), _p(dco), _p(dop), _p(dfe))
    if threads > 1:
        fn = lib.blend_backward_mt_f32 if dt == np.float32 else lib.blend_backward_mt_f64
        fn(*args, int(threads))
    else:
        fn = lib.blend_backward_f32 if dt == np.float32 else lib.blend_backward_f64
        fn(*args)
    return duv, dco, dop, dfe
