"""The no-allocation fused-loss forward (msplat_rasterize_forward_l2_noalloc)
against the workspace entry point it shares its kernels with
(msplat_rasterize_forward_l2), through the C ABI: rendered image, final_T,
ncontrib, walk record and loss bit for bit, on two calls in a row, the second
with a smaller scene and image; an image whose partials exceed the scratch
array fails before anything is launched."""

import pytest
import torch

from msplat_amd import _lib, binning
from msplat_amd.rasterization import preprocess_forward
from msplat_amd.sort_gaussian import tile_grid
from tests.helpers import small_scene

pytestmark = pytest.mark.gpu
DEV = "cuda"
CELL = 8


def _inputs(seed, n, W, H, C):
    g, cam = small_scene(seed, n=n, W=W, H=H, half_extent=1.5, C=C)
    d = {k: v.to(DEV) for k, v in g.items()}
    uv, depth, conic, radius, _ = preprocess_forward(d["xyz"], d["scale"], d["rotate"], cam.intr.to(DEV),
                                                     cam.extr.to(DEV), W, H)
    bins = binning.bin_gaussians(uv, depth, W, H, radius, CELL)
    target = torch.rand(C, H, W, generator=torch.Generator().manual_seed(seed + 1)).to(DEV)
    return uv, conic, d["opacity"].contiguous(), d["feature"].contiguous(), bins, target


def _outputs(W, H, C):
    gx, gy = tile_grid(W, H)
    return (torch.empty((C, H, W), device=DEV), torch.empty((H, W), device=DEV),
            torch.empty((H, W), dtype=torch.int32, device=DEV), torch.empty((gx * gy, 2), dtype=torch.int32,
                                                                                device=DEV),
            torch.empty((), device=DEV))


def _call(noalloc, uv, conic, opacity, feature, bins, target, W, H):
    n, C = feature.shape
    out = _outputs(W, H, C)
    common = (_lib.ptr(uv), _lib.ptr(conic), _lib.ptr(opacity), _lib.ptr(feature), n, C, _lib.ptr(bins.rect),
              _lib.ptr(bins.cell_ids), _lib.ptr(bins.cell_range), CELL, 0.25, W, H, 0, 1, _lib.ptr(target))
    outs = tuple(_lib.ptr(t) for t in out)
    if noalloc:
        _lib.call("msplat_rasterize_forward_l2_noalloc", *common, *outs, _lib.stream_ptr())
    else:
        nb = int(_lib.query("msplat_rasterize_l2_workspace_bytes", W, H, C))
        ws = torch.empty(nb, dtype=torch.uint8, device=DEV)
        _lib.call("msplat_rasterize_forward_l2", *common, _lib.ptr(ws), nb, *outs, _lib.stream_ptr())
        torch.cuda.synchronize()  # ws stays alive until the kernels have run
    return out


def test_noalloc_equals_workspace_entry_point():
    for seed, n, W, H, C in ((3, 6000, 256, 192, 4), (4, 2500, 96, 64, 3)):  # the second call smaller
        print(f"no-alloc forward: seed {seed} n={n} {W}x{H} C={C}", flush=True)
        args = _inputs(seed, n, W, H, C)
        a = _call(True, *args, W, H)
        torch.cuda.synchronize()
        b = _call(False, *args, W, H)
        for x, y, what in zip(a, b, ("rendered", "final_T", "ncontrib", "walk", "loss")):
            assert torch.equal(x, y), what
        assert float(a[4]) > 0


def test_noalloc_rejects_images_beyond_the_scratch():
    limit = int(_lib.query("msplat_noalloc_scratch_bytes"))
    W = H = 16384
    need = int(_lib.query("msplat_rasterize_l2_workspace_bytes", W, H, 4))
    assert need > limit
    dummy = torch.zeros(16, device=DEV)
    p = _lib.ptr(dummy)
    lib = _lib.load()
    status = lib.msplat_rasterize_forward_l2_noalloc(p, p, p, p, 1, 4, p, p, p, CELL, 0.0, W, H, 0, 1, p, p, p, p,
                                                     p, p, _lib.stream_ptr())
    assert status != 0
    assert b"scratch" in lib.msplat_last_error()
