"""Continuous-batching render server over scenes resident in HBM.

- ``SceneStore`` keeps every scene's Gaussians resident on the device, all
  scenes concatenated in one set of arrays, so views of different scenes can
  share one launch (each view names its scene's row range).
- ``RenderServer`` takes render requests (scene, camera, image size) as they
  arrive.  Each engine step admits queued requests while their estimated
  device memory fits the HBM budget (and a batch has view slots left), and
  launches them at once as one many-view, many-scene batch
  (``msplat_amd.views``: one preprocessing, one binning and one blending
  launch) on the next stream of a small pool; batches on different streams
  overlap on the device, so requests that arrive while others are on the GPU
  start as soon as memory and slots free up, not when the GPU drains.
  Finished batches are found by polling their completion events; their
  memory returns to the budget and each request's latency (completion -
  arrival) is recorded.
- ``run_poisson`` drives the server with Poisson arrivals and reports p50 /
  p99 latency beside throughput.

A request's image does not depend on what it was batched with: every view of
a launch is rendered by the same device code over its own lists, exactly as a
single-view ``rasterization`` call would.  The reference has no serving layer.
"""

from __future__ import annotations

import collections
import dataclasses
import math
import time

import torch

from . import tuning, views


@dataclasses.dataclass
class Scene:
    first: int
    count: int


class SceneStore:
    """Resident scenes: one concatenated set of per-Gaussian arrays on the device."""

    def __init__(self, device="cuda", channels: int = 3):
        self.device = torch.device(device)
        self.C = channels
        self.scenes: dict[str, Scene] = {}
        self.xyz = torch.empty((0, 3), device=self.device)
        self.scale = torch.empty((0, 3), device=self.device)
        self.rotate = torch.empty((0, 4), device=self.device)
        self.opacity = torch.empty((0, 1), device=self.device)
        self.feature = torch.empty((0, channels), device=self.device)

    def add(self, name: str, xyz, scale, rotate, opacity, feature) -> Scene:
        if feature.shape[1] != self.C:
            raise ValueError(f"scene {name}: {feature.shape[1]} channels, the store holds {self.C}")
        s = Scene(self.xyz.shape[0], xyz.shape[0])
        dev = self.device
        self.xyz = torch.cat([self.xyz, xyz.to(dev, torch.float32)])
        self.scale = torch.cat([self.scale, scale.to(dev, torch.float32)])
        self.rotate = torch.cat([self.rotate, rotate.to(dev, torch.float32)])
        self.opacity = torch.cat([self.opacity, opacity.to(dev, torch.float32).reshape(-1, 1)])
        self.feature = torch.cat([self.feature, feature.to(dev, torch.float32)])
        self.scenes[name] = s
        return s

    def nbytes(self) -> int:
        return sum(t.numel() * 4 for t in (self.xyz, self.scale, self.rotate, self.opacity, self.feature))


@dataclasses.dataclass
class Request:
    scene: str
    intr: torch.Tensor  # [4]
    extr: torch.Tensor  # [3,4] or [4,4]
    W: int
    H: int
    bg: float = 0.0
    t_arrival: float = 0.0
    t_done: float = 0.0
    image: torch.Tensor | None = None
    batch: int = -1

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival


def request_bytes(n: int, W: int, H: int, C: int, cell_entries_per_gaussian: float = 8.0) -> int:
    """Device memory a request holds while in flight (an upper estimate).

    Per (view, Gaussian) row: uv, depth, conic, radius, tiles, rect (40 B) and
    the binning workspace (about 40 B); cell entries (4 B plus 16 B of sort
    workspace each); the image planes (C + 2 floats per pixel)."""
    return int(n * 80 + n * cell_entries_per_gaussian * 20 + W * H * (C + 2) * 4)


@dataclasses.dataclass
class _Batch:
    bid: int
    reqs: list
    event: torch.cuda.Event
    out: torch.Tensor
    nbytes: int
    stream: torch.cuda.Stream


class RenderServer:
    def __init__(self, store: SceneStore, hbm_budget: int = 16 << 30, max_views: int = 16, streams: int = 3):
        if not 1 <= max_views <= views.MAX_VIEWS:
            raise ValueError(f"max_views must be in 1..{views.MAX_VIEWS}")
        self.store = store
        self.budget = int(hbm_budget)
        self.max_views = max_views
        self.queue: collections.deque[Request] = collections.deque()
        self.in_flight: list[_Batch] = []
        self.used = 0
        self.streams = [torch.cuda.Stream(device=store.device) for _ in range(streams)]  # default priority
        self.next_stream = 0
        self.next_batch = 0
        self.done: list[Request] = []

    def submit(self, req: Request) -> None:
        if req.scene not in self.store.scenes:
            raise KeyError(f"unknown scene {req.scene}")
        if req.t_arrival == 0.0:
            req.t_arrival = time.perf_counter()
        self.queue.append(req)

    def _cost(self, req: Request) -> int:
        return request_bytes(self.store.scenes[req.scene].count, req.W, req.H, self.store.C)

    def _admit(self) -> list[Request]:
        batch, nbytes = [], 0
        while self.queue and len(batch) < self.max_views:
            cost = self._cost(self.queue[0])
            if self.used + nbytes + cost > self.budget and (batch or self.in_flight):
                break  # wait for memory (a request alone is always admitted on an idle device)
            batch.append(self.queue.popleft())
            nbytes += cost
        return batch

    def _launch(self, reqs: list[Request]) -> None:
        st = self.store
        dev = st.device
        nbytes = sum(self._cost(r) for r in reqs)
        stream = self.streams[self.next_stream % len(self.streams)]
        self.next_stream += 1
        stream.wait_stream(torch.cuda.current_stream(dev))
        sizes = [(r.W, r.H) for r in reqs]
        first = [st.scenes[r.scene].first for r in reqs]
        count = [st.scenes[r.scene].count for r in reqs]
        cfg = tuning.config_for_views(sizes, st.C)
        with torch.cuda.stream(stream), torch.no_grad():
            intrs = torch.stack([r.intr.to(dev, torch.float32).reshape(4) for r in reqs])
            extrs = torch.stack([r.extr.to(dev, torch.float32).reshape(-1)[:12] for r in reqs])
            uv, depth, conic, radius, _ = views.preprocess_views_forward(st.xyz, st.scale, st.rotate, intrs, extrs,
                                                                         sizes, first=first, count=count)
            bins = views.bin_views(uv, depth, radius, st.xyz.shape[0], sizes, cfg.cell, first=first, count=count)
            bg = reqs[0].bg
            out, _, _, _ = views.rasterize_views_forward(uv, conic, st.opacity, st.feature, st.xyz.shape[0], sizes,
                                                         bins, bg, cfg.ppl_fwd, cfg.cull, first=first, count=count)
            ev = torch.cuda.Event()
            ev.record(stream)
        for r in reqs:
            r.batch = self.next_batch
        self.in_flight.append(_Batch(self.next_batch, reqs, ev, out, nbytes, stream))
        self.next_batch += 1
        self.used += nbytes

    def _launch_grouped(self, reqs: list[Request]) -> None:
        # one launch per background colour (bg is a per-launch constant)
        by_bg = collections.defaultdict(list)
        for r in reqs:
            by_bg[r.bg].append(r)
        for group in by_bg.values():
            self._launch(group)

    def _retire(self) -> int:
        keep, finished = [], 0
        for b in self.in_flight:
            if b.event.query():
                now = time.perf_counter()
                imgs = views.split_views(b.out, self.store.C, [(r.W, r.H) for r in b.reqs])
                for r, im in zip(b.reqs, imgs):
                    r.image = im
                    r.t_done = now
                    self.done.append(r)
                self.used -= b.nbytes
                finished += 1
            else:
                keep.append(b)
        self.in_flight = keep
        return finished

    def step(self) -> int:
        """One engine iteration: retire finished batches, admit and launch. Returns requests launched."""
        self._retire()
        reqs = self._admit()
        if reqs:
            self._launch_grouped(reqs)
        return len(reqs)

    def drain(self) -> None:
        while self.queue or self.in_flight:
            if not self.step():
                time.sleep(20e-6)

    def busy(self) -> bool:
        return bool(self.queue or self.in_flight)


def percentile(xs, q: float) -> float:
    xs = sorted(xs)
    if not xs:
        return float("nan")
    k = (len(xs) - 1) * q
    lo, hi = math.floor(k), math.ceil(k)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def run_poisson(server: RenderServer, make_request, rate: float, duration: float, seed: int = 0,
                warmup: int = 8) -> dict:
    """Submit Poisson(rate) arrivals for `duration` seconds of wall time; return latency / throughput stats.

    make_request(rng) -> Request (scene, camera, size).  The first `warmup`
    requests are served before timing starts and are not counted."""
    import random

    rng = random.Random(seed)
    for _ in range(warmup):
        server.submit(make_request(rng))
    server.drain()
    server.done.clear()
    t0 = time.perf_counter()
    t_next = t0 + rng.expovariate(rate)
    submitted = 0
    while True:
        now = time.perf_counter()
        while t_next <= now and t_next < t0 + duration:
            req = make_request(rng)
            req.t_arrival = t_next
            server.submit(req)
            submitted += 1
            t_next += rng.expovariate(rate)
        if not server.step():
            if t_next >= t0 + duration and not server.busy():
                break
            time.sleep(20e-6)
    t1 = time.perf_counter()
    lat = [r.latency * 1e3 for r in server.done]
    pix = sum(r.W * r.H for r in server.done)
    return dict(rate_req_s=rate, duration_s=round(t1 - t0, 3), requests=len(server.done),
                throughput_req_s=round(len(server.done) / (t1 - t0), 2),
                throughput_mpix_s=round(pix / (t1 - t0) / 1e6, 2),
                p50_ms=round(percentile(lat, 0.5), 3), p99_ms=round(percentile(lat, 0.99), 3),
                max_ms=round(max(lat), 3) if lat else None, batches=server.next_batch)
