"""End-to-end parity of rasterization() (fused path) with the CPU oracle and with
the reference's own chain of per-stage ops, forward and backward."""

import numpy as np
import pytest
import torch

import msplat_amd as ms
from oracle import pipeline
from tests.helpers import np_, rel_err, small_scene

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _chain(xyz, scale, rotate, opacity, feature, intr, extr, W, H, bg, ndc=None):
    """msplat/__init__.py:71-91 with this package's per-stage ops."""
    uv, depth = ms.project_point(xyz, intr, extr, W, H)
    visible = depth != 0
    cov3d = ms.compute_cov3d(scale, rotate, visible)
    conic, radius, tiles = ms.ewa_project(xyz, cov3d, intr, extr, uv, W, H, visible)
    idx, rng = ms.sort_gaussian(uv, depth, W, H, radius, tiles)
    return ms.alpha_blending(uv, conic, opacity, feature, idx, rng, bg, W, H, ndc)


def _leaves(g, cam):
    xyz = g["xyz"].to(DEV).requires_grad_()
    scale = g["scale"].to(DEV).requires_grad_()
    rot = g["rotate"].to(DEV).requires_grad_()
    op = g["opacity"].to(DEV).requires_grad_()
    feat = g["feature"].to(DEV).requires_grad_()
    intr = cam.intr.to(DEV).requires_grad_()
    extr = cam.extr.to(DEV).requires_grad_()
    return xyz, scale, rot, op, feat, intr, extr


@pytest.mark.parametrize("seed,n,W,H,ext", [(0, 3000, 128, 96, 1.0), (1, 6000, 160, 120, 3.0)])
def test_rasterization_vs_oracle(seed, n, W, H, ext):
    g, cam = small_scene(seed, n=n, W=W, H=H, half_extent=ext)
    leaves = _leaves(g, cam)
    bg = 0.1
    ndc = torch.zeros(n, 2, device=DEV, requires_grad=True)
    out = ms.rasterization(*leaves, W, H, bg, ndc)
    npg = {k: np_(v) for k, v in g.items()}
    o64, inter = pipeline.rasterize_forward(npg["xyz"], npg["scale"], npg["rotate"], npg["opacity"],
                                            npg["feature"], np_(cam.intr), np_(cam.extr), W, H, bg,
                                            dtype=np.float64)
    err = np.abs(np_(out) - o64).max()
    assert err < 1e-4, f"image max-abs error vs fp64 oracle {err}"
    gout = np.random.default_rng(seed).normal(size=out.shape).astype(np.float32)
    (out * torch.from_numpy(gout).to(DEV)).sum().backward()
    ref = pipeline.rasterize_backward(npg["xyz"], npg["scale"], npg["rotate"], npg["opacity"], npg["feature"],
                                      np_(cam.intr), np_(cam.extr), W, H, bg, inter, gout, dtype=np.float64)
    xyz, scale, rot, op, feat, intr, extr = leaves
    checks = [(xyz.grad, ref["dxyz"], "dxyz"), (scale.grad, ref["dscale"], "dscale"),
              (rot.grad, ref["drotate"], "drotate"), (op.grad, ref["dopacity"].reshape(-1, 1), "dopacity"),
              (feat.grad, ref["dfeature"], "dfeature"), (intr.grad, ref["dintr"], "dintr"),
              (extr.grad.reshape(-1), ref["dextr"], "dextr")]
    # dL/dndc (msplat/alpha_blending.py:107-110): pixel u = (ndc_x + 1) W/2 - 1/2, v likewise with H, so
    # dL/dndc = dL/duv * (W/2, H/2), with dL/duv the fp64 oracle's blending adjoint (pinned by finite differences)
    checks.append((ndc.grad, ref["duv"] * np.array([0.5 * W, 0.5 * H]), "dndc"))
    for a, b, w in checks:
        e = rel_err(np_(a), b)
        assert e < 1e-3, f"{w}: relative error {e}"


def test_fused_equals_chain():
    g, cam = small_scene(5, n=4000, W=144, H=100, half_extent=2.0)
    la = _leaves(g, cam)
    lb = _leaves(g, cam)
    ndc_a = torch.zeros(4000, 2, device=DEV, requires_grad=True)
    ndc_b = torch.zeros(4000, 2, device=DEV, requires_grad=True)
    oa = ms.rasterization(*la, 144, 100, 0.3, ndc_a)
    ob = _chain(*lb, 144, 100, 0.3, ndc_b)
    assert torch.equal(oa, ob)
    gout = torch.randn(oa.shape, generator=torch.Generator().manual_seed(3)).to(DEV)
    (oa * gout).sum().backward()
    (ob * gout).sum().backward()
    for a, b in zip(la + (ndc_a,), lb + (ndc_b,)):
        # same kernels; only the order of the fp32 atomic additions differs
        assert rel_err(np_(a.grad), np_(b.grad)) < 1e-4


def test_extr_4x4_gradient_shape():
    g, cam = small_scene(6, n=1000, W=64, H=64)
    leaves = list(_leaves(g, cam))
    e4 = torch.eye(4)
    e4[:3] = cam.extr
    leaves[6] = e4.to(DEV).requires_grad_()
    out = ms.rasterization(*leaves, 64, 64, 0.0)
    out.sum().backward()
    assert leaves[6].grad.shape == (4, 4)
    assert torch.all(leaves[6].grad[3] == 0)


def test_no_fallback_on_cpu_tensors():
    g, cam = small_scene(7, n=100, W=32, H=32)
    with pytest.raises(RuntimeError):
        ms.project_point(g["xyz"], cam.intr, cam.extr, 32, 32)
