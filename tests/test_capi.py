"""The C ABI without a GPU: the library loads, exports every entry point
include/msplat_amd.h declares (and the Python binding types all of them),
and rejects bad arguments with a status and a message before any launch."""

import ctypes
import os
import re
import subprocess
import sys

import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "msplat_amd.h")


def header_symbols():
    text = open(HEADER).read()
    return set(re.findall(r"^\s*(?:int|size_t|const char \*)\s*(msplat_\w+)\s*\(", text, re.M))


@pytest.fixture(scope="module")
def lib():
    from msplat_amd import _lib
    return _lib.load()


def test_header_parses():
    syms = header_symbols()
    assert len(syms) >= 20, syms


def test_library_exports_header(lib):
    for s in sorted(header_symbols()):
        assert hasattr(lib, s), f"{s} declared in include/msplat_amd.h but not exported"


def test_binding_matches_header():
    from msplat_amd import _lib
    assert set(_lib.SIGNATURES) == header_symbols()


def test_nm_exports_are_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "msplat_amd", "_lib",
                                                                       "libmsplat_amd.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (msplat_\w+)$", out, re.M))
    assert header_symbols() <= exported


def test_version(lib):
    assert lib.msplat_version() >= 100


def _views_range_call(L, fn, first, count, n=10):
    """A one-view call of `fn` with Gaussian range [first, first + count) of n (no pointers: it must fail first)."""
    W, H = (ctypes.c_int * 1)(64), (ctypes.c_int * 1)(48)
    F, N = (ctypes.c_int64 * 1)(first), (ctypes.c_int64 * 1)(count)
    if fn == "bin_views_count":
        return L.msplat_bin_views_count(None, None, None, n, 1, W, H, F, N, 8, None, None, 0, None, None)
    if fn == "preprocess_views_forward":
        return L.msplat_preprocess_views_forward(None, None, None, n, 1, None, None, W, H, F, N,
                                                 ctypes.c_float(0.0), ctypes.c_float(1.3), None, None, None, None,
                                                 None, None)
    return L.msplat_rasterize_views_forward(None, None, None, None, n, 3, 0, 1, W, H, F, N, None, None, None, 8,
                                            ctypes.c_float(0.0), 0, 1, None, None, None, None, None)


@pytest.mark.parametrize("fn", ["bin_views_count", "preprocess_views_forward", "rasterize_views_forward"])
@pytest.mark.parametrize("first,count", [(5, 6), (11, 0), (0, 11), (3, (1 << 63) - 2)])
def test_view_range_past_the_gaussians(lib, fn, first, count):
    """A view whose Gaussian range runs past the n Gaussians is an argument error, raised before any launch."""
    status = _views_range_call(lib, fn, first, count)
    assert status == -1
    assert "exceeds the 10 Gaussians" in lib.msplat_last_error().decode()


@pytest.mark.parametrize("fn", ["bin_views_count", "preprocess_views_forward", "rasterize_views_forward"])
def test_view_range_at_the_end_passes_the_range_check(lib, fn):
    status = _views_range_call(lib, fn, 4, 6)  # [4, 10): in range, then fails on the null pointers
    assert "exceeds" not in lib.msplat_last_error().decode() or status == 0


@pytest.mark.parametrize("call,needle", [
    (lambda L: L.msplat_project_point_forward(None, -1, None, None, 8, 8, 0.0, 1.3, None, None, None), "n = -1"),
    (lambda L: L.msplat_compute_sh_forward(None, None, None, 4, 3, 17, None, None), "D = 17"),
    (lambda L: L.msplat_ewa_project_forward(None, None, None, None, None, None, 3, 0, 8, None, None, None, None),
     "image size"),
    (lambda L: L.msplat_preprocess_forward(None, None, None, 5, None, None, 8, 8, 0.0, 1.3, None, None, None,
                                           None, None, None), "null pointer"),
    (lambda L: L.msplat_sort_emit(None, None, None, 0, 8, 8, None, -1, None, 0, None, None, None, None),
     "disagree"),
    (lambda L: L.msplat_sort_emit(None, None, None, 0, 8, 8, None, 1 << 31, None, 0, None, None, None, None),
     "int32"),
    (lambda L: L.msplat_sort_prepare(None, None, None, None, 10, 8, 8, None, 0, ctypes.c_void_p(8), None),
     "workspace"),
])
def test_argument_errors(lib, call, needle):
    status = call(lib)
    assert status < 0
    assert needle in lib.msplat_last_error().decode()


def test_workspace_sizes_grow(lib):
    a = lib.msplat_sort_prepare_workspace_bytes(1000, 1920, 1080)
    b = lib.msplat_sort_prepare_workspace_bytes(1_000_000, 1920, 1080)
    assert 0 < a < b and b >= 1_000_000 * 28
    assert lib.msplat_sort_emit_workspace_bytes(10, 5_000_000, 1920, 1080) >= 5_000_000 * 16


def test_python_raises_without_device_tensors():
    import torch

    import msplat_amd as ms
    with pytest.raises(RuntimeError):
        ms.compute_cov3d(torch.rand(4, 3), torch.rand(4, 4))


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
