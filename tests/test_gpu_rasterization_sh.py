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


@pytest.mark.parametrize("deg,E,depth_channel", [(3, 0, True), (3, 2, True), (1, 0, False), (7, 1, True)])
def test_matches_composition(deg, E, depth_channel):
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
        P["campos"] = campos.clone().requires_grad_()
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
        if A[k] is None:
            continue
        assert rel_err(np_(A[k].grad), np_(B[k].grad)) < 1e-5, k
