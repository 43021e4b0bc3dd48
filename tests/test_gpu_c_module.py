"""msplat_amd._C, the stand-in for msplat's pybind module, driven in the
order msplat's own Python files call it: the forward chain of
msplat/__init__.py:71-91 through the per-stage functions, the sort of
msplat/sort_gaussian.py:42-52 (inclusive cumsum, keys, torch.sort, gather,
tile ranges) and every backward function.  Same kernels as msplat_amd's ops,
so images, keys and ranges are bit-exact with rasterization() and
sort_gaussian(), and each backward equals the op's autograd gradient (to
the order of fp32 atomic additions where the kernels accumulate with
atomics: blending gradients and camera gradients, 1e-5 relative)."""

import pytest
import torch

import msplat_amd as ms
from msplat_amd import _C
from tests.helpers import np_, rel_err, small_scene

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _chain_via_C(xyz, scale, rotate, opacity, feature, intr, extr, W, H, bg):
    uv, depth = _C.project_point_forward(xyz, intr, extr, W, H, 0.0, 1.3)
    visible = depth != 0
    cov3d = _C.compute_cov3d_forward(scale, rotate, visible)
    conic, radius, tiles = _C.ewa_project_forward(xyz, cov3d, intr, extr, uv, W, H, visible)
    csum = torch.cumsum(tiles, dim=0, dtype=torch.int32)
    key, idx = _C.compute_gaussian_key(uv, depth, W, H, radius, csum)
    key_sorted, order = torch.sort(key, stable=True)
    idx_sorted = torch.gather(idx, 0, order)
    tile_range = _C.compute_tile_gaussian_range(W, H, csum, key_sorted)
    out, final_T, ncontrib = _C.alpha_blending_forward(uv, conic, opacity, feature, idx_sorted, tile_range, bg, W, H)
    return dict(uv=uv, depth=depth, visible=visible, cov3d=cov3d, conic=conic, radius=radius, tiles=tiles,
                idx_sorted=idx_sorted, tile_range=tile_range, out=out, final_T=final_T, ncontrib=ncontrib)


def test_reference_call_sequence():
    W, H, bg = 160, 112, 0.2
    g, cam = small_scene(12, n=5000, W=W, H=H, half_extent=2.5, C=3)
    d = {k: v.to(DEV) for k, v in g.items()}
    intr, extr = cam.intr.to(DEV), cam.extr.to(DEV)
    r = _chain_via_C(d["xyz"], d["scale"], d["rotate"], d["opacity"], d["feature"], intr, extr, W, H, bg)
    assert r["depth"].shape == (5000, 1) and r["radius"].dtype == torch.int32 and r["ncontrib"].dtype == torch.int32
    with torch.no_grad():
        img = ms.rasterization(d["xyz"], d["scale"], d["rotate"], d["opacity"], d["feature"], intr, extr, W, H, bg)
        idx, rng = ms.sort_gaussian(r["uv"], r["depth"], W, H, r["radius"], r["tiles"])
    assert torch.equal(r["out"], img)
    assert torch.equal(r["idx_sorted"], idx) and torch.equal(r["tile_range"], rng)

    # every backward, against the autograd op of the same stage
    gen = torch.Generator().manual_seed(3)
    rnd = lambda *s: torch.randn(*s, generator=gen).to(DEV)
    dout = rnd(*r["out"].shape)
    duv, dconic, dop, dfeat = _C.alpha_blending_backward(r["uv"], r["conic"], d["opacity"], d["feature"],
                                                         r["idx_sorted"], r["tile_range"], bg, W, H, r["final_T"],
                                                         r["ncontrib"], dout)
    assert dop.shape == d["opacity"].shape
    leaves = [t.clone().requires_grad_() for t in (r["uv"], r["conic"], d["opacity"], d["feature"])]
    ms.alpha_blending(*leaves, r["idx_sorted"], r["tile_range"], bg, W, H).backward(dout)
    for a, l in zip((duv, dconic, dop, dfeat), leaves):  # fp32 atomics: the order of additions may differ
        assert rel_err(np_(a), np_(l.grad)) < 1e-5

    iv = intr.clone().requires_grad_()
    ev = extr.clone().requires_grad_()
    dx_e, dcov, di_e, de_e = _C.ewa_project_backward(d["xyz"], r["cov3d"], iv, ev, r["radius"], dconic)
    x2, c2 = d["xyz"].clone().requires_grad_(), r["cov3d"].clone().requires_grad_()
    conic2, _, _ = ms.ewa_project(x2, c2, iv, ev, r["uv"], W, H, r["visible"])
    conic2.backward(dconic)
    assert torch.equal(dx_e, x2.grad) and torch.equal(dcov, c2.grad)
    assert rel_err(np_(di_e), np_(iv.grad)) < 1e-5 and rel_err(np_(de_e), np_(ev.grad)) < 1e-5
    zi, ze = _C.ewa_project_backward(d["xyz"], r["cov3d"], intr, extr, r["radius"], dconic)[2:]
    assert not zi.any() and not ze.any() and ze.shape == (3, 4)  # cameras without requires_grad: zeros

    ds, dq = _C.compute_cov3d_backward(d["scale"], d["rotate"], r["visible"], dcov)
    s2, q2 = d["scale"].clone().requires_grad_(), d["rotate"].clone().requires_grad_()
    ms.compute_cov3d(s2, q2, r["visible"]).backward(dcov)
    assert torch.equal(ds, s2.grad) and torch.equal(dq, q2.grad)

    ddepth = rnd(5000, 1)
    iv.grad = ev.grad = None
    dx_p, di_p, de_p = _C.project_point_backward(d["xyz"], iv, ev, W, H, r["uv"], r["depth"], duv, ddepth)
    x3 = d["xyz"].clone().requires_grad_()
    uv3, dp3 = ms.project_point(x3, iv, ev, W, H)
    torch.autograd.backward([uv3, dp3], [duv, ddepth])
    assert torch.equal(dx_p, x3.grad)
    assert rel_err(np_(di_p), np_(iv.grad)) < 1e-5 and rel_err(np_(de_p), np_(ev.grad)) < 1e-5


def test_sh_functions():
    gen = torch.Generator().manual_seed(4)
    n = 3000
    shs = torch.randn(n, 3, 16, generator=gen).to(DEV)
    dirs = torch.nn.functional.normalize(torch.randn(n, 3, generator=gen), dim=1).to(DEV)
    vis = (torch.rand(n, generator=gen) < 0.8).to(DEV)
    v = _C.compute_sh_forward(shs, dirs, vis)
    s2, d2 = shs.clone().requires_grad_(), dirs.clone().requires_grad_()
    v2 = ms.compute_sh(s2, d2, vis)
    assert torch.equal(v, v2)
    dv = torch.randn(n, 3, generator=gen).to(DEV)
    dshs, ddirs = _C.compute_sh_backward(shs, dirs, vis, dv)
    v2.backward(dv)
    assert torch.equal(dshs, s2.grad) and torch.equal(ddirs, d2.grad)
