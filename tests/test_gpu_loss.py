"""The fused squared-error loss against torch: value to fp32 rounding (and
against an fp64 sum), gradient bit-identical to torch's 2 (x - t) dL,
independent of scheduling, odd sizes and misaligned views included."""

import pytest
import torch

import msplat_amd as ms

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("shape", [(4, 1080, 1920), (3, 37, 53), (1,), (1001,)])
def test_l2_loss_matches_torch(shape):
    gen = torch.Generator().manual_seed(sum(shape))
    x = torch.rand(shape, generator=gen).to(DEV).requires_grad_()
    t = torch.rand(shape, generator=gen).to(DEV)
    loss = ms.l2_loss(x, t)
    ref64 = ((x.detach().double() - t.double()) ** 2).sum()
    assert abs(float(loss.detach()) - float(ref64)) <= 1e-6 * float(ref64) + 1e-12
    loss.backward(torch.tensor(0.75, device=DEV))
    xr = x.detach().clone().requires_grad_()
    ((xr - t) ** 2).sum().backward(torch.tensor(0.75, device=DEV))
    assert torch.equal(x.grad, xr.grad)
    assert float(ms.l2_loss(x.detach(), t)) == float(loss.detach())  # the same bits every call


def test_l2_loss_misaligned_and_target_grad():
    gen = torch.Generator().manual_seed(3)
    base = torch.rand(10_003, generator=gen).to(DEV)
    t = torch.rand(10_002, generator=gen).to(DEV)
    xv = base[1:]  # a view 4 bytes past a 16-byte boundary: the unaligned path
    ref = ((xv.double() - t.double()) ** 2).sum()
    assert abs(float(ms.l2_loss(xv, t)) - float(ref)) <= 1e-6 * float(ref)
    x = xv.clone().requires_grad_()
    tt = t.clone().requires_grad_()
    ms.l2_loss(x, tt).backward()
    assert torch.equal(tt.grad, -x.grad)
