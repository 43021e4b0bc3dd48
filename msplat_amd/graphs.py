"""HIP-graph capture of steps that rasterize.

The binning's entry count is data dependent; eagerly the host reads it back
(one 8-byte synchronisation per binning) to size the lists.  A captured graph
cannot synchronise, so ``GraphedStep`` first runs the step eagerly, recording
the count of every binning call in order, then captures it with each binning
call given a fixed capacity (the largest count seen times ``margin``): the
count stays on the device and the graph replays with no host round trip.
After a replay, ``overflowed()`` reports whether any binning exceeded its
capacity (its lists were then truncated and the step must be recaptured with
a larger margin).

Graph-safe by construction: the library clears memory with fill kernels (no
memset nodes), launches only on the capturing stream, and the kernel timing
events are off unless requested.

Disabled: on the MI355X test box the first capture of a single-view training
step ended in a segmentation fault inside the HIP runtime while the capture
was being closed and instantiated (torch.cuda.CUDAGraph.capture_end), after
the same step had run eagerly without error.  The cause is not identified, so
``GraphedStep`` raises instead of capturing; the capacity-mode binning it
relies on stays available.
"""


from __future__ import annotations

import math

import torch

from . import binning

CAPTURE_DISABLED = ("HIP graph capture of rasterization steps is disabled: a capture ended in a segmentation "
                    "fault inside the HIP runtime at capture end, cause not yet identified")


class _Plan:
    def __init__(self):
        self.mode = "record"
        self.seen: list[int] = []
        self.caps: list[int] = []
        self.counts: list[torch.Tensor] = []
        self.i = 0

    def begin(self):
        self.i = 0

    def capacity(self):
        if self.mode == "record":
            return None
        cap = self.caps[self.i]
        self.i += 1
        return cap

    def observe(self, count: int, count_tensor: torch.Tensor):
        if self.mode == "record":
            if self.i < len(self.seen):
                self.seen[self.i] = max(self.seen[self.i], count)
            else:
                self.seen.append(count)
            self.i += 1
        else:
            self.counts.append(count_tensor)


class GraphedStep:
    """Capture ``fn`` (a step with fixed shapes that may call rasterization) in a HIP graph.

    ``fn`` must read its inputs from, and write its results to, tensors that
    persist across calls (static buffers), as any graph-captured code must.
    """

    def __init__(self, fn, warmup: int = 2, margin: float = 1.25):
        raise RuntimeError(CAPTURE_DISABLED)
        self.fn = fn
        self.plan = _Plan()
        prev = binning.set_plan(self.plan)
        try:
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                for _ in range(max(1, warmup)):
                    self.plan.begin()
                    fn()
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            self.plan.caps = [max(1024, int(math.ceil(c * margin))) for c in self.plan.seen]
            self.plan.mode = "capture"
            self.plan.begin()
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self.out = fn()
        finally:
            binning.set_plan(prev)
        self.capacities = list(self.plan.caps)

    def __call__(self):
        self.graph.replay()
        return self.out

    def overflowed(self) -> bool:
        """True if a binning of the last replay exceeded its capacity (synchronises)."""
        return any(int(c.item()) > cap for c, cap in zip(self.plan.counts, self.plan.caps))
