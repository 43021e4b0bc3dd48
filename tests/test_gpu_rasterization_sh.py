"""rasterization_sh (SH colours, extra channels and depth rendered in one op,
the SH kernels on a side stream) against the composition of separate ops:
compute_sh_points + project_point's depth + cat + rasterization.  Images bit
for bit; every gradient to float summation order."""

import pytest
import torch

import msplat_amd as ms
from tests.helpers import np_, rel_err, small_scene

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("deg,E,depth_channel,campos_grad", [(3, 0, True, True), (3, 2, True, True),
                                                            (1, 0, False, True), (7, 1, True, True),
                                                            (3, 0, True, False), (2, 1, False, False)])
def test_matches_composition(deg, E, depth_channel, campos_grad):
    W, H = 176, 128
    g, cam = small_scene(5 + deg + E, n=3000, W=W, H=H, half_extent=2.0, C=3)
    gen = torch.Generator().manual_seed(deg)
    D = (deg + 1) ** 2
    shs = (torch.randn(3000, 3, D, generator=gen) * 0.3)
    extra = torch.rand(3000, E, generator=gen) if E else None
    intr = cam.intr.to(DEV)
    campos = cam.eye.to(DEV)

    def leaves():
        P = {k: g[k].to(DEV).clone().requires_grad_() for k in ("xyz", "scale", "rotate", "opacity")}
        P["shs"] = shs.to(DEV).clone().requires_grad_()
        P["extra"] = extra.to(DEV).clone().requires_grad_() if E else None
        P["extr"] = cam.extr.to(DEV).clone().requires_grad_()
        P["campos"] = campos.clone().requires_grad_(campos_grad)
        return P

    A, B = leaves(), leaves()
    out = ms.rasterization_sh(A["xyz"], A["scale"], A["rotate"], A["opacity"], A["shs"], intr, A["extr"],
                              A["campos"], W, H, 0.1, extra=A["extra"], depth_channel=depth_channel)
    parts = [ms.compute_sh_points(B["shs"], B["xyz"], B["campos"])]
    if E:
        parts.append(B["extra"])
    if depth_channel:
        parts.append(ms.project_point(B["xyz"], intr, B["extr"], W, H)[1])
    ref = ms.rasterization(B["xyz"], B["scale"], B["rotate"], B["opacity"], torch.cat(parts, 1), intr, B["extr"],
                           W, H, 0.1)
    assert out.shape == ref.shape == (3 + E + int(depth_channel), H, W)
    assert torch.equal(out, ref)
    dout = torch.randn(out.shape, generator=gen).to(DEV)
    (out * dout).sum().backward()
    (ref * dout).sum().backward()
    for k in ("xyz", "scale", "rotate", "opacity", "shs", "extra", "extr", "campos"):
        if A[k] is None or not A[k].requires_grad:
            continue
        assert rel_err(np_(A[k].grad), np_(B[k].grad)) < 1e-5, k


def test_sh_xyz_gradient_added_in_place_is_exact():
    """msplat_compute_sh_points_backward_add adds the SH xyz gradient into a
    given array: the same bits as storing it and adding the two arrays, the
    projected-only mask included, except that a row with no upstream gradient
    (its SH part is zero) is left as it is, so a -0 there stays -0 where the
    addition gives +0 (equal values); dL/dshs unchanged."""
    from msplat_amd import _lib
    n, C, D, ld = 5000, 3, 16, 4
    gen = torch.Generator().manual_seed(23)
    shs = (torch.randn(n, C, D, generator=gen) * 0.3).to(DEV)
    xyz = (torch.randn(n, 3, generator=gen) * 2.0).to(DEV)
    campos = torch.tensor([0.3, -0.2, 6.0]).to(DEV)
    dval = torch.randn(n, ld, generator=gen)
    dval[::3] = 0.0  # rows with no upstream gradient
    dval = dval.to(DEV)
    base = torch.randn(n, 3, generator=gen)
    base[::7] = -0.0
    base = base.to(DEV)
    for vis in (None, (torch.rand(n, generator=gen) > 0.3).to(torch.uint8).to(DEV)):
        dshs_a = torch.empty_like(shs)
        dxyz_sh = torch.empty_like(xyz)
        _lib.call("msplat_compute_sh_points_backward", _lib.ptr(shs), _lib.ptr(xyz), _lib.ptr(campos),
                  _lib.ptr(vis), _lib.ptr(dval), ld, n, C, D, _lib.ptr(dshs_a), _lib.ptr(dxyz_sh), _lib.stream_ptr())
        dshs_b = torch.empty_like(shs)
        acc = base.clone()
        _lib.call("msplat_compute_sh_points_backward_add", _lib.ptr(shs), _lib.ptr(xyz), _lib.ptr(campos),
                  _lib.ptr(vis), _lib.ptr(dval), ld, n, C, D, _lib.ptr(dshs_b), _lib.ptr(acc), _lib.stream_ptr())
        want = base + dxyz_sh
        assert torch.equal(acc, want)
        same_bits = acc.view(torch.int32) == want.view(torch.int32)
        assert bool(same_bits[want != 0].all())
        assert bool(((acc == 0) | same_bits).all())
        assert torch.equal(dshs_a, dshs_b)


@pytest.mark.parametrize("other_term,target_grad", [(False, False), (True, False), (False, True), (True, True)])
def test_l2_fused_matches_separate_loss(other_term, target_grad):
    """rasterization_sh_l2 = rasterization_sh + l2_loss: the image bit for bit,
    the loss to fp32 rounding, every gradient (the image's other terms and the
    target's included) to float summation order."""
    W, H = 176, 128
    g, cam = small_scene(31, n=3000, W=W, H=H, half_extent=2.0, C=3)
    gen = torch.Generator().manual_seed(31)
    shs = torch.randn(3000, 3, 16, generator=gen) * 0.3
    target = torch.rand(4, H, W, generator=gen).to(DEV)
    wimg = torch.randn(4, H, W, generator=gen).to(DEV)
    intr = cam.intr.to(DEV)

    def leaves():
        P = {k: g[k].to(DEV).clone().requires_grad_() for k in ("xyz", "scale", "rotate", "opacity")}
        P["shs"] = shs.to(DEV).clone().requires_grad_()
        P["extr"] = cam.extr.to(DEV).clone().requires_grad_()
        P["target"] = target.clone().requires_grad_(target_grad)
        return P

    A, B = leaves(), leaves()
    img_a, loss_a = ms.rasterization_sh_l2(A["xyz"], A["scale"], A["rotate"], A["opacity"], A["shs"], intr, A["extr"],
                                           cam.eye.to(DEV), W, H, 0.1, A["target"], depth_channel=True)
    img_b = ms.rasterization_sh(B["xyz"], B["scale"], B["rotate"], B["opacity"], B["shs"], intr, B["extr"],
                                cam.eye.to(DEV), W, H, 0.1, depth_channel=True)
    loss_b = ms.l2_loss(img_b, B["target"])
    assert torch.equal(img_a, img_b)
    ref64 = ((img_b.detach().double() - target.double()) ** 2).sum()
    assert abs(float(loss_a.detach()) - float(ref64)) <= 1e-6 * float(ref64)
    assert abs(float(loss_a.detach()) - float(loss_b.detach())) <= 1e-6 * float(ref64)
    total_a = loss_a * 0.75 + ((img_a * wimg).sum() if other_term else 0.0)
    total_b = loss_b * 0.75 + ((img_b * wimg).sum() if other_term else 0.0)
    total_a.backward()
    total_b.backward()
    for k in ("xyz", "scale", "rotate", "opacity", "shs", "extr") + (("target",) if target_grad else ()):
        assert rel_err(np_(A[k].grad), np_(B[k].grad)) < 1e-5, k


def test_l2_fused_image_only():
    """Gradients through the image alone (the loss unused) equal rasterization_sh's."""
    W, H = 128, 96
    g, cam = small_scene(32, n=2000, W=W, H=H, half_extent=2.0, C=3)
    gen = torch.Generator().manual_seed(32)
    shs = torch.randn(2000, 3, 4, generator=gen) * 0.3
    target = torch.rand(3, H, W, generator=gen).to(DEV)
    dout = torch.randn(3, H, W, generator=gen).to(DEV)
    A = {k: g[k].to(DEV).clone().requires_grad_() for k in ("xyz", "opacity")}
    B = {k: g[k].to(DEV).clone().requires_grad_() for k in ("xyz", "opacity")}
    args = (g["scale"].to(DEV), g["rotate"].to(DEV))
    img_a, _ = ms.rasterization_sh_l2(A["xyz"], *args, A["opacity"], shs.to(DEV), cam.intr.to(DEV),
                                      cam.extr.to(DEV), cam.eye.to(DEV), W, H, 0.0, target)
    img_b = ms.rasterization_sh(B["xyz"], *args, B["opacity"], shs.to(DEV), cam.intr.to(DEV), cam.extr.to(DEV),
                                cam.eye.to(DEV), W, H, 0.0)
    assert torch.equal(img_a, img_b)
    (img_a * dout).sum().backward()
    (img_b * dout).sum().backward()
    for k in ("xyz", "opacity"):
        assert rel_err(np_(A[k].grad), np_(B[k].grad)) < 1e-5, k
