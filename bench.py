"""Benchmark: rasterization forward + backward throughput on one MI355X.

Workload (the headline configuration): N = 1,000,000 Gaussians in
[-3,3]^3, one 1920x1080 view (fx = fy = 1.171875 W, camera at distance 4),
features = compute_sh(degree 3, [N,3,16] ~ N(0,0.3)) concatenated with the
projected depth (C = 4), L2 loss against a U(0,1) target, backward to every
Gaussian parameter.  A step is one forward + backward of that view; inputs
are resident in HBM before the timed region.

Prints ONE JSON line (metric Mpix/s = W*H*steps / seconds / 1e6), with the
roofline figure of the dominant stage (HIP events on its launch stream
inside the timed region) and a CPU baseline (the oracle, on this host).
Per-stage timings go to stderr.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_VALU_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_step(cfg, dev):
    import torch

    import msplat_amd as ms
    from msplat_amd import scenes

    g = cfg["gaussians"]
    cam = cfg["camera"].to(dev)
    P = {k: v.to(dev).contiguous().requires_grad_() for k, v in g.items()}
    target = cfg["target"].to(dev)
    W, H = cam.W, cam.H

    def step():
        xyz = P["xyz"]
        _, depth = ms.project_point(xyz, cam.intr, cam.extr, W, H)
        rgb = ms.compute_sh(P["shs"], scenes.view_dirs(xyz, cam.eye))
        feats = torch.cat([rgb, depth], 1)
        out = ms.rasterization(xyz, P["scale"], P["rotate"], P["opacity"], feats, cam.intr, cam.extr, W, H, 0.0)
        loss = ((out - target) ** 2).sum()
        loss.backward()
        for p in P.values():
            p.grad = None
        return out

    return step, P, cam


def algorithmic_work(name, n, k, C, D):
    """Algorithmic bytes moved by one call of a stage (HBM-bound stages)."""
    table = {
        "msplat_project_point_forward": 24 * n,          # xyz in, uv + depth out
        "msplat_project_point_backward": 40 * n,         # xyz, depth, dL/duv, dL/ddepth in; dL/dxyz out
        "msplat_preprocess_forward": 72 * n,             # xyz, scale, rot in; uv, depth, conic, radius, tiles out
        "msplat_preprocess_backward": 108 * n,           # + depth, radius, dL/duv, dL/dconic in; 3 grads out
        "msplat_compute_sh_forward": n * (4 * C * D + 12 + 4 * C),
        "msplat_compute_sh_backward": n * (8 * C * D + 24 + 4 * C),
        # depth-key sort of n pairs: 4 passes x (hist read 4 + pair read 8 + pair write 8), + 28 B/Gaussian in/out
        "msplat_sort_prepare": n * (28 + 4 * 20 + 24),
        # per intersection: pair written once, 2 class passes x 20 B, range read 4; per Gaussian 28 B read
        "msplat_sort_emit": k * (8 + 2 * 20 + 4) + 28 * n,
    }
    return table.get(name)


def cpu_baseline(cfg, budget_rows: int = 128) -> dict:
    """Time the CPU oracle (numpy + C, one thread) on a bounded sample of the same frame."""
    import numpy as np

    from oracle import blend, geometry, sh, sort

    g = {k: v.numpy() for k, v in cfg["gaussians"].items()}
    cam = cfg["camera"]
    W, H = cam.W, cam.H
    intr, extr, eye = cam.intr.numpy(), cam.extr.numpy(), cam.eye.numpy()
    t0 = time.perf_counter()
    uv, depth = geometry.project_forward(g["xyz"], intr, extr, W, H)
    dirs = g["xyz"] - eye[None]
    dirs = dirs / np.linalg.norm(dirs, axis=1, keepdims=True)
    rgb = sh.sh_forward(g["shs"], dirs)
    feats = np.concatenate([rgb, depth[:, None]], 1).astype(np.float32)
    vis = depth != 0
    cov = geometry.cov3d_forward(g["scale"], g["rotate"], vis)
    conic, radius, tiles = geometry.ewa_forward(g["xyz"], cov, intr, extr, uv, W, H, vis)
    idx, rng, _ = sort.sort_gaussian(uv, depth, radius, W, H)
    t_geo_fwd = time.perf_counter() - t0
    # blending: a band of tile rows, forward + backward, scaled to the frame
    rows = min(H, budget_rows)
    band_rng = rng.reshape((H + 15) // 16, (W + 15) // 16, 2)[: (rows + 15) // 16].reshape(-1, 2)
    t1 = time.perf_counter()
    out, fT, nc, _ = blend.blend_forward(uv, conic, g["opacity"], feats, idx, band_rng, 0.0, W, rows)
    dout = np.ones_like(out)
    duv, dco, dop, dfe = blend.blend_backward(uv, conic, g["opacity"], feats, idx, band_rng, 0.0, W, rows, fT, nc,
                                              dout)
    t_blend_band = time.perf_counter() - t1
    t2 = time.perf_counter()
    dxyz_e, dcov, _, _ = geometry.ewa_backward(g["xyz"], cov, intr, extr, radius, dco)
    geometry.cov3d_backward(g["scale"], g["rotate"], dcov, vis)
    geometry.project_backward(g["xyz"], intr, extr, depth, duv)
    sh.sh_backward(g["shs"], dirs, dfe[:, :3])
    t_geo_bwd = time.perf_counter() - t2
    frame_s = t_geo_fwd + t_geo_bwd + t_blend_band * (H / rows)
    return dict(value=W * H / frame_s / 1e6, unit="Mpix/s", cores=1, kind="port",
                sample=(f"same frame (N={len(g['xyz'])}, {W}x{H}, C=4): projection, SH, covariance, EWA, sort and "
                        f"their adjoints timed in full ({t_geo_fwd + t_geo_bwd:.1f} s); blending forward+backward "
                        f"timed on the first {rows} rows ({t_blend_band:.1f} s) and scaled x{H / rows:.2f}"),
                seconds=time.perf_counter() - t0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=128)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards by view; "
              "see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch

    from msplat_amd import scenes
    from msplat_amd.profiling import stage_timer

    dev = torch.device("cuda:0")
    cfg = scenes.config_sh3_1080p(n=args.n)
    step, P, cam = build_step(cfg, dev)
    W, H = cam.W, cam.H
    log(f"workload: N={args.n} {W}x{H} C=4 (SH3 RGB + depth); warmup {args.warmup}, steps {args.steps}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    with stage_timer() as st:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    stages = st.summary()
    sec = t1 - t0
    value = W * H * args.steps / sec / 1e6

    # intersections of this view, for the sort's byte count
    from msplat_amd.rasterization import preprocess_forward
    with torch.no_grad():
        _, _, _, radius, tiles = preprocess_forward(P["xyz"], P["scale"], P["rotate"], cam.intr, cam.extr, W, H)
        K = int(tiles.sum().item())
    n, C, D = args.n, 3, 16
    total_stage_ms = sum(v["total_ms"] for v in stages.values())
    for name, v in sorted(stages.items(), key=lambda kv: -kv[1]["total_ms"]):
        w = algorithmic_work(name, n, K, C, D)
        bw = f"{w / (v['avg_ms'] * 1e-3) / 1e9:8.1f} GB/s" if w else ""
        log(f"  {name:40s} x{v['count'] // args.steps:<2d} avg {v['avg_ms']:8.3f} ms  "
            f"{100 * v['total_ms'] / total_stage_ms:5.1f}%  {bw}")
    log(f"K = {K} intersections; step {1e3 * sec / args.steps:.2f} ms")

    # dominant HBM-bound stage's roofline
    hbm = [(v["total_ms"], name) for name, v in stages.items() if algorithmic_work(name, n, K, C, D)]
    _, dom = max(hbm)
    w = algorithmic_work(dom, n, K, C, D)
    ach = w / (stages[dom]["avg_ms"] * 1e-3) / 1e9
    roofline = dict(kernel=dom, bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=None,
                    stage_share=round(stages[dom]["total_ms"] / total_stage_ms, 3))

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(cfg, args.cpu_rows)
        log(f"cpu baseline: {cpu['value']:.4f} Mpix/s in {cpu.pop('seconds'):.1f} s")

    line = dict(metric="fwd+bwd Mpix/s at N=1M,1080p", value=round(value, 3), unit="Mpix/s", n_gpus=1,
                steps=args.steps, warmup=args.warmup, ms_per_step=round(1e3 * sec / args.steps, 3),
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype="fp32", data="synthetic",
                config=dict(workload="N=1M Gaussians in [-3,3]^3, 1920x1080, SH deg 3 + depth (C=4), fwd+bwd L2",
                            gaussians=n, width=W, height=H, channels=4, intersections=K, parallelism="single GPU"),
                roofline=roofline, cpu_baseline=cpu,
                stages_ms={k: round(v["avg_ms"], 4) for k, v in stages.items()})
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
