"""Per-stage GPU timing with HIP events.

Inside ``with stage_timer() as st:``, every C-ABI call records a start and
an end event on the stream it launches on; ``st.summary()`` then returns
per-entry-point counts and average/total milliseconds.  Event recording is
asynchronous (no host synchronisation until ``summary``).
"""

from __future__ import annotations

import contextlib
from collections import defaultdict

import torch

_active: "StageTimer | None" = None


class StageTimer:
    def __init__(self):
        self.records: list[tuple[str, torch.cuda.Event, torch.cuda.Event]] = []

    def begin(self, name: str):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        return name, ev

    def end(self, token) -> None:
        name, ev0 = token
        ev1 = torch.cuda.Event(enable_timing=True)
        ev1.record()
        self.records.append((name, ev0, ev1))

    def summary(self) -> dict:
        torch.cuda.synchronize()
        agg = defaultdict(lambda: [0, 0.0])
        for name, a, b in self.records:
            agg[name][0] += 1
            agg[name][1] += a.elapsed_time(b)
        return {k: dict(count=c, total_ms=t, avg_ms=t / c) for k, (c, t) in agg.items()}


@contextlib.contextmanager
def stage_timer():
    global _active
    prev, _active = _active, StageTimer()
    try:
        yield _active
    finally:
        _active = prev


def active() -> "StageTimer | None":
    return _active


# ---- rocprofv3 output summaries (python -m msplat_amd.profiling ...) ----------
#
#   stats DIR                 kernel-trace statistics of a `rocprofv3 --kernel-trace
#                             --stats --output-format csv -d DIR` run
#   steps DIR PATTERN         per step (a step ends with a kernel whose name matches
#                             PATTERN): wall time on the device, busy time (sum of kernel
#                             durations), idle share and launches
#   counters DIR              per-launch counter averages of one `rocprofv3 --pmc ...`
#                             run and the utilisation figures they give
#   pmc FETCH_DIR WRITE_DIR MIB
#                             HBM bytes per launch of the library's kernels from two
#                             `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs of
#                             `bench.py --calibrate-copy-mib MIB`; each counter is scaled
#                             by what it reads for that copy of known size.

import csv as _csv
import glob as _glob
import re as _re
import sys as _sys

# library kernel -> device function pattern (in rocprof's demangled names)
KERNEL_PATTERNS = {
    "rasterize_bwd": r"blend_bwd_kernel<\d+, \d+, true(, \d+)?>",
    "rasterize_fwd": r"blend_fwd_kernel<\d+, \d+, true>",
    "blend_bwd": r"blend_bwd_kernel<\d+, \d+, false(, \d+)?>",
    "blend_fwd": r"blend_fwd_kernel<\d+, \d+, false>",
    "sh_fwd": r"sh_fwd_\w*kernel",
    "sh_bwd": r"sh_bwd_\w*kernel",
    "preprocess_fwd": r"preprocess_fwd_\w*kernel",
    "preprocess_bwd": r"preprocess_bwd_\w*kernel",
    "bin_prepare": r"bin_prepare_kernel",
    "emit": r"emit_(chunk|direct|cells)_kernel",
    # both radix sorts (depth order and cell class) launch the same kernels
    "radix_scatter": r"radix_scatter_tile_kernel",
    "radix_hist": r"(onesweep|radix)_hist_kernel",
}
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def _rows(d, suffix):
    out = []
    for f in sorted(_glob.glob(f"{d}/**/*{suffix}", recursive=True)):
        out.extend(_csv.DictReader(open(f)))
    return out


def summarize_stats(d, top=25):
    rows = _rows(d, "kernel_stats.csv")
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    lines = [f"{'kernel':72s} {'calls':>6s} {'avg_us':>10s} {'total_ms':>10s} {'share':>6s}"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:top]:
        name = _re.sub(r"\(.*", "", r["Name"])[:72]
        lines.append(f"{name:72s} {r['Calls']:>6s} {float(r['AverageNs']) / 1e3:10.1f} "
                     f"{float(r['TotalDurationNs']) / 1e6:10.3f} {100 * float(r['TotalDurationNs']) / total:5.1f}%")
    return "\n".join(lines)


def summarize_steps(d, pattern, last=5):
    rows = sorted(_rows(d, "kernel_trace.csv"), key=lambda r: int(r["Start_Timestamp"]))
    ends = [i for i, r in enumerate(rows) if _re.search(pattern, r["Kernel_Name"])]
    lines = [f"{'step':>4s} {'wall_us':>9s} {'busy_us':>9s} {'idle':>6s} {'launches':>8s}"]
    for j, (a, b) in enumerate(list(zip(ends[:-1], ends[1:]))[-last:]):
        seg = rows[a + 1:b + 1]
        wall = int(rows[b]["End_Timestamp"]) - int(rows[a]["End_Timestamp"])
        busy = sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in seg)
        lines.append(f"{j:4d} {wall / 1e3:9.1f} {busy / 1e3:9.1f} {100 * (1 - busy / wall):5.1f}% {len(seg):8d}")
    return "\n".join(lines)


def _per_dispatch(d, counter):
    """{dispatch id: (kernel name, summed counter value)} of one PMC run."""
    acc = {}
    for r in _rows(d, "counter_collection.csv"):
        if r.get("Counter_Name") != counter:
            continue
        key = r["Dispatch_Id"]
        name, v = acc.get(key, (r["Kernel_Name"], 0.0))
        acc[key] = (name, v + float(r["Counter_Value"]))
    return acc


def summarize_pmc(fetch_dir, write_dir, mib):
    copy_bytes = mib * (1 << 20)
    lines = []
    factors = {}
    per = {}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        disp = _per_dispatch(d, counter)
        copies = [v for name, v in disp.values() if COPY_KERNEL in name]
        if not copies:
            raise SystemExit(f"{counter}: no {COPY_KERNEL} dispatch in {d}")
        kib = max(copies)  # the calibration copy is the largest one
        factors[counter] = copy_bytes / (kib * 1024.0)
        lines.append(f"calibration {counter}: copy of {mib} MiB reads {kib:.0f} KiB -> factor {factors[counter]:.4f}")
        for k, pat in KERNEL_PATTERNS.items():
            vals = [v for name, v in disp.values() if _re.search(pat, name)]
            if vals:
                per.setdefault(k, {})[counter] = (sum(vals) / len(vals) * 1024.0 * factors[counter], len(vals))
    for k, c in per.items():
        f = c.get("FETCH_SIZE", (None, 0))[0]
        w = c.get("WRITE_SIZE", (None, 0))[0]
        if f is None or w is None:
            continue
        lines.append(f"traffic {k} bytes_per_launch={f + w:.0f} fetch={f:.0f} write={w:.0f} "
                     f"launches={c['FETCH_SIZE'][1]}")
    return "\n".join(lines)


# MI355X: 256 CUs (4 SIMDs each) in 8 XCDs.  GRBM_GUI_ACTIVE, summed over the
# dispatch's dimensions, counts each XCD's busy cycles, so / 8 is the
# kernel's own cycle count.
CUS, SIMDS, XCDS = 256, 1024, 8


def _derived(c):
    """Utilisation figures from one kernel's per-launch counter averages
    (only those whose counters the run holds)."""
    out = []
    cyc = c.get("GRBM_GUI_ACTIVE", 0.0) / XCDS
    if cyc > 0:
        out.append(("cycles", cyc))
        if "SQ_ACTIVE_INST_VALU" in c:  # 4 cycles per wave64 VALU issue on one SIMD
            out.append(("valu_busy", c["SQ_ACTIVE_INST_VALU"] * 4 / SIMDS / cyc))
        if "SQ_LDS_IDX_ACTIVE" in c:  # cycles the CU's LDS is busy with indexed ops
            out.append(("lds_util", c["SQ_LDS_IDX_ACTIVE"] / CUS / cyc))
    if c.get("SQ_LDS_IDX_ACTIVE", 0) > 0 and "SQ_LDS_BANK_CONFLICT" in c:
        out.append(("lds_bank_conflict", c["SQ_LDS_BANK_CONFLICT"] / c["SQ_LDS_IDX_ACTIVE"]))
    if c.get("SQ_WAVE_CYCLES", 0) > 0:
        for k, lab in (("SQ_WAIT_ANY", "wait_any"), ("SQ_WAIT_INST_LDS", "wait_lds"),
                       ("SQ_ACTIVE_INST_ANY", "issue_any")):
            if k in c:
                out.append((lab, c[k] / c["SQ_WAVE_CYCLES"]))
    return out


def summarize_counters(d):
    """Per-launch averages of every counter of one PMC run, for the kernels of
    KERNEL_PATTERNS (counters summed over a dispatch's dimensions first), then
    the utilisation figures they give:

      valu_busy         SQ_ACTIVE_INST_VALU * 4 / (1024 SIMDs * cycles)
      lds_util          SQ_LDS_IDX_ACTIVE / (256 CUs * cycles)
      lds_bank_conflict SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE
      wait_any, wait_lds, issue_any
                        SQ_WAIT_ANY, SQ_WAIT_INST_LDS, SQ_ACTIVE_INST_ANY per
                        SQ_WAVE_CYCLES (shares of the waves' resident cycles)

    with cycles = GRBM_GUI_ACTIVE / 8 XCDs."""
    per = {}
    names = []
    for r in _rows(d, "counter_collection.csv"):
        c = r.get("Counter_Name")
        if c not in names:
            names.append(c)
        key = (r["Dispatch_Id"], c)
        name, v = per.get(key, (r["Kernel_Name"], 0.0))
        per[key] = (name, v + float(r["Counter_Value"]))
    lines = [f"{'kernel':16s} {'launches':>8s} " + " ".join(f"{c:>20s}" for c in names)]
    derived = []
    for k, pat in KERNEL_PATTERNS.items():
        cols, n = [], 0
        for c in names:
            vals = [v for (disp, cn), (name, v) in per.items() if cn == c and _re.search(pat, name)]
            n = max(n, len(vals))
            cols.append(sum(vals) / len(vals) if vals else float("nan"))
        if n:
            lines.append(f"{k:16s} {n:8d} " + " ".join(f"{v:20.4g}" for v in cols))
            dv = _derived(dict(zip(names, cols)))
            if dv:
                derived.append(f"{k:16s} " + "  ".join(f"{lab} {v:.4g}" for lab, v in dv))
    if derived:
        lines += [""] + derived
    return "\n".join(lines)


if __name__ == "__main__":
    if len(_sys.argv) >= 3 and _sys.argv[1] == "stats":
        print(summarize_stats(_sys.argv[2]))
    elif len(_sys.argv) >= 4 and _sys.argv[1] == "steps":
        print(summarize_steps(_sys.argv[2], _sys.argv[3]))
    elif len(_sys.argv) >= 5 and _sys.argv[1] == "pmc":
        print(summarize_pmc(_sys.argv[2], _sys.argv[3], int(_sys.argv[4])))
    elif len(_sys.argv) >= 3 and _sys.argv[1] == "counters":
        print(summarize_counters(_sys.argv[2]))
    else:
        raise SystemExit(__doc__)
