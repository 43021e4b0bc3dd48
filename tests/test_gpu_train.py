"""The training step next to the rasterizer, through the C ABI:

* msplat_amd.train.Adam against torch.optim.Adam (the reference's optimizer,
  tutorials/gs_3d.py:165-172) on the same tensors and gradients: gs_3d.py's
  five parameter groups with their learning rates, misaligned and odd-sized
  tensors, more than 8 tensors per step;
* DensifyStats (the dL/dndc consumer) bit-exact against the same fp32
  operations in numpy;
* prune_and_clone bit-exact against tutorials/gs_3d.py:193-225 restated with
  torch on the CPU (stable descending argsort, assignments in order),
  including pruned Gaussians that are also donors, tied scores and an empty
  prune set; an expired sort wait copies nothing and raises.

Adam tolerance: |p - p_torch| <= 1e-6 (1 + |p|) after 12 steps; the two
differ only where torch's kernels contract a multiply-add."""

import numpy as np
import pytest
import torch

from msplat_amd import _lib, train
from msplat_amd.binning import check_waits

pytestmark = pytest.mark.gpu
DEV = "cuda"

GS3D_LR = {"xyz": 0.0001, "rgb": 0.005, "opacity": 0.05, "scale": 0.01, "rot": 0.001}
WIDTH = {"xyz": 3, "rgb": 3, "opacity": 1, "scale": 3, "rot": 4}


def _params(n, seed):
    gen = torch.Generator().manual_seed(seed)
    return {k: torch.randn(n, w, generator=gen) for k, w in WIDTH.items()}


def _adam_pair(tensors, lrs, cls):
    ps = [torch.nn.Parameter(t.to(DEV).clone()) for t in tensors]
    opt = cls([{"params": [p], "lr": lr} for p, lr in zip(ps, lrs)])
    return ps, opt


def _run_adam(tensors, lrs, steps=12, seed=5):
    mine, o1 = _adam_pair(tensors, lrs, train.Adam)
    ref, o2 = _adam_pair(tensors, lrs, torch.optim.Adam)
    gen = torch.Generator().manual_seed(seed)
    for _ in range(steps):
        for a, b in zip(mine, ref):
            g = torch.randn(a.shape, generator=gen).to(DEV) * 0.1
            a.grad = g.clone()
            b.grad = g.clone()
        o1.step()
        o2.step()
    torch.cuda.synchronize()
    for a, b in zip(mine, ref):
        d = (a.detach() - b.detach()).abs()
        assert float((d - 1e-6 * (1 + b.detach().abs())).max()) <= 0, float(d.max())
        s1, s2 = o1.state[a], o2.state[b]
        assert s1["step"] == int(s2["step"])
        for key in ("exp_avg", "exp_avg_sq"):
            e = (s1[key] - s2[key]).abs().max()
            assert float(e) <= 1e-6 * (1 + float(s2[key].abs().max())), key
    return mine, ref


def test_adam_gs3d_groups():
    P = _params(100_003, 0)
    _run_adam([P[k] for k in GS3D_LR], list(GS3D_LR.values()))


def test_adam_misaligned_odd_and_many():
    gen = torch.Generator().manual_seed(1)
    # more than 8 tensors (two launches), odd sizes, one storage offset by one float
    tensors = [torch.randn(n, generator=gen) for n in (1, 3, 5, 7, 1023, 1025, 4097, 17, 6, 99, 64)]
    lrs = [0.01 * (i + 1) for i in range(len(tensors))]
    mine, o1 = _adam_pair(tensors, lrs, train.Adam)
    ref, o2 = _adam_pair(tensors, lrs, torch.optim.Adam)
    big = torch.zeros(1 + 1029, device=DEV)
    off = big[1:]  # 4-byte offset: the kernel takes its element-wise path
    off.copy_(torch.randn(1029, generator=gen).to(DEV))
    pm = torch.nn.Parameter(off)
    pr = torch.nn.Parameter(off.detach().clone())
    o1.add_param_group({"params": [pm], "lr": 0.02})
    o2.add_param_group({"params": [pr], "lr": 0.02})
    mine.append(pm)
    ref.append(pr)
    assert pm.data_ptr() % 16 != 0
    for _ in range(6):
        for a, b in zip(mine, ref):
            g = torch.randn(a.shape, generator=gen).to(DEV)
            a.grad = g.clone()
            b.grad = g.clone()
        o1.step()
        o2.step()
    torch.cuda.synchronize()
    for a, b in zip(mine, ref):
        d = (a.detach() - b.detach()).abs()
        assert float((d - 1e-6 * (1 + b.detach().abs())).max()) <= 0


def test_adam_skips_parameters_without_gradient():
    a = torch.nn.Parameter(torch.ones(10, device=DEV))
    b = torch.nn.Parameter(torch.ones(10, device=DEV))
    opt = train.Adam([a, b], lr=0.1)
    a.grad = torch.ones(10, device=DEV)
    opt.step()
    torch.cuda.synchronize()
    assert torch.all(a.detach() < 1) and torch.equal(b.detach(), torch.ones(10, device=DEV))
    assert b not in opt.state


def test_densify_accumulate_bitexact():
    gen = torch.Generator().manual_seed(2)
    n = 50_001
    st = train.DensifyStats(n, DEV)
    acc = np.zeros(n, np.float32)
    den = np.zeros(n, np.float32)
    for v in range(3):
        g = torch.randn(n, 2, generator=gen) * 1e-3
        g[torch.rand(n, generator=gen) < 0.3] = 0.0
        radius = (torch.rand(n, generator=gen) * 4).to(torch.int32) if v != 1 else None
        st.add(g.to(DEV), None if radius is None else radius.to(DEV))
        gn = g.numpy()
        seen = radius.numpy() > 0 if radius is not None else ((gn[:, 0] != 0) | (gn[:, 1] != 0))
        norm = np.sqrt(gn[:, 0] * gn[:, 0] + gn[:, 1] * gn[:, 1])
        acc[seen] = acc[seen] + norm[seen]
        den[seen] = den[seen] + np.float32(1)
    torch.cuda.synchronize()
    assert np.array_equal(st.accum.cpu().numpy(), acc)
    assert np.array_equal(st.denom.cpu().numpy(), den)
    m = st.mean().cpu().numpy()
    ok = den > 0
    assert np.allclose(m[ok], acc[ok] / den[ok], rtol=1e-6) and np.all(m[~ok] == 0)


def _gs3d_restatement(P, opacity_act, threshold, score=None):
    """tutorials/gs_3d.py:193-214 with torch on the CPU; argsort made stable,
    the assignment executed in index-list order (last write wins)."""
    prune_mask = opacity_act.reshape(-1) < threshold
    prune_indices = torch.nonzero(prune_mask).squeeze(-1)
    m = prune_indices.shape[0]
    out = {k: v.clone() for k, v in P.items()}
    if m == 0:
        return out, 0
    key = P["scale"].amax(dim=1) if score is None else score
    densify_indices = torch.argsort(key, descending=True, stable=True)[:m]
    all_indices = torch.cat([prune_indices, densify_indices]).tolist()
    for k in P:
        vals = P[k][densify_indices].repeat(2, 1)
        for j, i in enumerate(all_indices):
            out[k][i] = vals[j]
    return out, m


@pytest.mark.parametrize("case", ["plain", "overlap", "ties", "score", "none"])
def test_prune_and_clone_matches_gs3d(case):
    n = 20_000
    P = _params(n, 3)
    gen = torch.Generator().manual_seed(4)
    opacity_act = torch.rand(n, 1, generator=gen) * 0.2
    score = None
    if case == "overlap":  # large scales on low-opacity Gaussians: pruned rows that are donors
        low = (opacity_act.reshape(-1) < 0.01).nonzero().squeeze(-1)
        P["scale"][low[::2]] += 10.0
    if case == "ties":
        P["scale"] = torch.round(P["scale"] * 2) / 2
    if case == "score":
        score = torch.rand(n, generator=gen)
        score[::7] = 0.5  # ties
    if case == "none":
        opacity_act = opacity_act + 0.5
    ref, m = _gs3d_restatement(P, opacity_act, 0.01, score)
    if case == "overlap":
        pruned = set((opacity_act.reshape(-1) < 0.01).nonzero().squeeze(-1).tolist())
        donors = set(torch.argsort(P["scale"].amax(dim=1), descending=True, stable=True)[:m].tolist())
        assert pruned & donors
    D = {k: v.to(DEV).contiguous() for k, v in P.items()}
    got = train.prune_and_clone([D[k] for k in WIDTH], opacity_act.to(DEV), 0.01,
                                score=None if score is None else score.to(DEV),
                                scale=D["scale"] if score is None else None)
    assert got == m
    for k in WIDTH:
        assert torch.equal(D[k].cpu(), ref[k]), k


def test_prune_and_clone_expired_wait_copies_nothing():
    n = 9000
    P = _params(n, 6)
    opacity_act = torch.rand(n, 1) * 0.1
    D = {k: v.to(DEV).contiguous() for k, v in P.items()}
    s = _lib.stream_ptr()
    _lib.call("msplat_debug_force_expiry", 1, s)
    try:
        with pytest.raises(RuntimeError, match="expired"):
            train.prune_and_clone([D[k] for k in WIDTH], opacity_act.to(DEV), 0.01, scale=D["scale"])
    finally:
        _lib.call("msplat_debug_force_expiry", 0, s)
        torch.cuda.synchronize()
    for k in WIDTH:
        assert torch.equal(D[k].cpu(), P[k]), k
    check_waits()  # the word was cleared when the error was raised
