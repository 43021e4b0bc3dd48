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
