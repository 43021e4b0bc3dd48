"""ctypes binding of ``libmsplat_amd.so`` (the C ABI in include/msplat_amd.h).

The library is loaded lazily, after torch, so that it binds to the HIP
runtime torch already loaded.  There is no fallback: if the library is
missing, or a tensor is not on a HIP device, calls raise RuntimeError.
"""

from __future__ import annotations

import ctypes
import os
import threading

import torch

from . import profiling

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libmsplat_amd.so")

P = ctypes.c_void_p
I64 = ctypes.c_int64
I32 = ctypes.c_int
F32 = ctypes.c_float
SZ = ctypes.c_size_t

# name -> (restype, argtypes); mirrors include/msplat_amd.h.
SIGNATURES = {
    "msplat_version": (I32, []),
    "msplat_last_error": (ctypes.c_char_p, []),
    "msplat_wait_status": (I32, [P, I32, P]),
    "msplat_debug_force_expiry": (I32, [I32, P]),
    "msplat_project_point_forward": (I32, [P, I64, P, P, I32, I32, F32, F32, P, P, P]),
    "msplat_project_point_backward": (I32, [P, I64, P, P, P, P, P, P, P, P, P]),
    "msplat_compute_cov3d_forward": (I32, [P, P, P, I64, P, P]),
    "msplat_compute_cov3d_backward": (I32, [P, P, P, P, I64, P, P, P]),
    "msplat_ewa_project_forward": (I32, [P, P, P, P, P, P, I64, I32, I32, P, P, P, P]),
    "msplat_ewa_project_backward": (I32, [P, P, P, P, P, P, I64, P, P, P, P, P]),
    "msplat_preprocess_forward": (I32, [P, P, P, I64, P, P, I32, I32, F32, F32, P, P, P, P, P, P]),
    "msplat_preprocess_backward": (I32, [P, P, P, I64, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "msplat_preprocess_forward_col": (I32, [P, P, P, I64, P, P, I32, I32, F32, F32, P, P, P, P, P, P, I32, P, P]),
    "msplat_preprocess_backward_ld": (I32, [P, P, P, I64, P, P, P, P, P, P, I32, P, P, P, P, P, P, P]),
    "msplat_compute_sh_forward": (I32, [P, P, P, I64, I32, I32, P, P]),
    "msplat_compute_sh_backward": (I32, [P, P, P, P, I64, I32, I32, P, P, P]),
    "msplat_compute_sh_points_forward": (I32, [P, P, P, P, I64, I32, I32, P, I32, P]),
    "msplat_compute_sh_points_backward": (I32, [P, P, P, P, P, I32, I64, I32, I32, P, P, P]),
    "msplat_compute_sh_points_backward_add": (I32, [P, P, P, P, P, I32, I64, I32, I32, P, P, P]),
    "msplat_sort_prepare_workspace_bytes": (SZ, [I64, I32, I32]),
    "msplat_sort_prepare": (I32, [P, P, P, P, I64, I32, I32, P, SZ, P, P]),
    "msplat_sort_emit_workspace_bytes": (SZ, [I64, I64, I32, I32]),
    "msplat_sort_emit": (I32, [P, P, P, I64, I32, I32, P, I64, P, SZ, P, P, P, P]),
    "msplat_gaussian_keys": (I32, [P, P, P, P, I64, I32, I32, I64, P, P, P]),
    "msplat_tile_range_from_keys": (I32, [P, I64, I32, I32, P, P]),
    "msplat_alpha_blending_forward": (I32, [P, P, P, P, I64, I32, P, P, F32, I32, I32, P, P, P, P]),
    "msplat_alpha_blending_backward": (
        I32, [P, P, P, P, I64, I32, P, P, F32, I32, I32, P, P, P, P, P, P, P, P]),
    "msplat_timing_enable": (I32, [I32]),
    "msplat_timing_read": (I32, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(I64)]),
    "msplat_timing_only": (I32, [ctypes.c_char_p]),
    "msplat_bin_prepare_workspace_bytes": (SZ, [I64, I32, I32, I32]),
    "msplat_bin_prepare": (I32, [P, P, P, I64, I32, I32, I32, P, P, SZ, P, P]),
    "msplat_bin_emit_workspace_bytes": (SZ, [I64, I64, I32, I32, I32]),
    "msplat_bin_emit": (I32, [P, P, P, I64, I32, I32, I32, P, I64, P, P, SZ, P, P, P]),
    "msplat_preprocess_views_forward": (I32, [P, P, P, I64, I32, P, P, P, P, P, P, F32, F32, P, P, P, P, P, P]),
    "msplat_preprocess_views_backward": (I32, [P, P, P, I64, I32, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "msplat_sq_diff_sum_workspace_bytes": (SZ, []),
    "msplat_sq_diff_sum": (I32, [P, P, I64, P, SZ, P, P]),
    "msplat_sq_diff_grad": (I32, [P, P, I64, P, P, P]),
    "msplat_bin_views_prepare_workspace_bytes": (SZ, [I64]),
    "msplat_bin_views_prepare": (I32, [P, P, P, I64, I32, P, P, P, P, I32, P, P, SZ, P, P]),
    "msplat_bin_views_count": (I32, [P, P, P, I64, I32, P, P, P, P, I32, P, P, SZ, P, P]),
    "msplat_bin_views_order": (I32, [P, P, P, I64, I32, P, P, P, P, I32, P, P, SZ, P, P]),
    "msplat_bin_views_emit_workspace_bytes": (SZ, [I64]),
    "msplat_bin_views_emit": (I32, [P, P, P, I64, I32, P, P, P, P, I32, P, I64, P, P, SZ, P, P, P]),
    "msplat_rasterize_views_forward": (
        I32, [P, P, P, P, I64, I32, I32, I32, P, P, P, P, P, P, P, I32, F32, I32, I32, P, P, P, P, P]),
    "msplat_rasterize_views_backward": (
        I32, [P, P, P, P, I64, I32, I32, I32, P, P, P, P, P, I32, F32, I32, I32, P, P, P, P, P, P, P, P, P]),
    "msplat_rasterize_forward": (I32, [P, P, P, P, I64, I32, P, P, P, I32, F32, I32, I32, I32, I32, P, P, P, P, P]),
    "msplat_rasterize_backward": (
        I32, [P, P, P, P, I64, I32, P, P, P, I32, F32, I32, I32, I32, I32, P, P, P, P, P, P, P, P, P]),
    "msplat_rasterize_l2_workspace_bytes": (SZ, [I32, I32, I32]),
    "msplat_rasterize_forward_l2": (
        I32, [P, P, P, P, I64, I32, P, P, P, I32, F32, I32, I32, I32, I32, P, P, SZ, P, P, P, P, P, P]),
    "msplat_rasterize_backward_l2": (
        I32, [P, P, P, P, I64, I32, P, P, P, I32, F32, I32, I32, I32, I32, P, P, P, P, P, P, P, P, P, P, P, P]),
    "msplat_adam_step": (I32, [I32, P, P, P, P, P, P, P, F32, F32, F32, P]),
    "msplat_densify_accumulate": (I32, [P, P, I64, P, P, P]),
    "msplat_prune_clone_workspace_bytes": (SZ, [I64]),
    "msplat_prune_clone": (I32, [P, F32, P, P, I64, I32, P, P, P, SZ, P, P]),
    "msplat_noalloc_scratch_bytes": (SZ, []),
    "msplat_rasterize_forward_l2_noalloc": (
        I32, [P, P, P, P, I64, I32, P, P, P, I32, F32, I32, I32, I32, I32, P, P, P, P, P, P, P]),
}

_lib = None
_lock = threading.Lock()


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load (once) and return the library with typed entry points."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise RuntimeError(
                    f"msplat_amd HIP library not built ({path} missing); run "
                    "`python -m msplat_amd.build` or __graft_entry__.build()")
            lib = ctypes.CDLL(path)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def call(name: str, *args) -> int:
    """Call an entry point returning a status; raise RuntimeError on failure."""
    lib = load()
    timer = profiling.active()
    token = timer.begin(name) if timer is not None else None
    status = getattr(lib, name)(*args)
    if token is not None:
        timer.end(token)
    if status != 0:
        msg = lib.msplat_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed ({status}): {msg}")
    return status


def query(name: str, *args):
    return getattr(load(), name)(*args)


def timing_read(kernel: str) -> tuple[float, int]:
    """(total milliseconds, launches) the library recorded for `kernel` since the last read."""
    total = ctypes.c_double(0.0)
    launches = I64(0)
    call("msplat_timing_read", kernel.encode(), ctypes.byref(total), ctypes.byref(launches))
    return total.value, launches.value


def stream_ptr(device: torch.device | None = None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def check_device(name: str, t: torch.Tensor) -> None:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise RuntimeError(f"{name} must be a HIP (cuda) tensor")


def f32(name: str, t: torch.Tensor) -> torch.Tensor:
    """Validate device and return a contiguous float32 view/copy."""
    check_device(name, t)
    return t.detach().to(torch.float32).contiguous()


def i32(name: str, t: torch.Tensor) -> torch.Tensor:
    check_device(name, t)
    return t.detach().to(torch.int32).contiguous()


def u8(name: str, t: torch.Tensor | None) -> torch.Tensor | None:
    if t is None:
        return None
    check_device(name, t)
    return t.detach().to(torch.bool).contiguous().view(torch.uint8)
