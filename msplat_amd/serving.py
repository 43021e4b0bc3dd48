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
- A launch never waits for the device: cameras go up through pinned staging
  buffers (asynchronous copies on the batch's stream) and the binning runs
  in capacity mode (the entry count stays on the device).  The capacity is
  a running estimate of cell entries per (view, Gaussian) row; a batch whose
  lists overflowed it is detected when it retires, the estimate is raised
  and the batch is rendered again with the exact count.
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

from . import binning, tuning, views


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

    def add_ply(self, name: str, path: str, activate: bool = True) -> Scene:
        """Load a scene file (msplat_amd.scene_io's PLY layout) and keep it
        resident.  Its ``feat_*`` channels are rendered; a file with SH
        coefficients of degree 0 only renders their SH value (the constant
        basis function times the coefficient); higher degrees are view
        dependent and are refused."""
        from . import scene_io

        s = scene_io.load_ply(path, device="cpu", activate=activate)
        feat = s.get("feature")
        if feat is None and "shs" in s:
            if s["sh_degree"] != 0:
                raise ValueError(f"scene {name}: SH degree {s['sh_degree']} colours depend on the view; store "
                                 "rendered feature channels (feat_*) for serving")
            feat = s["shs"][:, :, 0] * 0.28209479177387814
        if feat is None:
            raise ValueError(f"scene {name}: the file holds no feature channels")
        return self.add(name, s["xyz"], s["scale"], s["rotate"], s["opacity"], feat)

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
    rows: int = 0
    capacity: int | None = None
    count: torch.Tensor | None = None  # [1] int64 on the device: the true cell-entry count


class RenderServer:
    def __init__(self, store: SceneStore, hbm_budget: int = 16 << 30, max_views: int = 16, streams: int = 3,
                 entries_per_row: float = 2.0, min_capacity: int = 4096):
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
        # capacity estimate (cell entries per row) and how often it was too small
        self.entries_per_row = float(entries_per_row)
        self.min_capacity = int(min_capacity)
        self.rerenders = 0
        # per stream: pinned host staging and device buffers for the cameras,
        # and the event after the last copy out of the staging buffers
        dev = store.device
        self._cam_host = [(torch.empty((max_views, 4), pin_memory=True), torch.empty((max_views, 12), pin_memory=True))
                          for _ in self.streams]
        self._cam_dev = [(torch.empty((max_views, 4), device=dev), torch.empty((max_views, 12), device=dev))
                         for _ in self.streams]
        self._cam_copied: list[torch.cuda.Event | None] = [None] * len(self.streams)

    def set_max_views(self, max_views: int) -> None:
        """Requests per launch from now on (the camera staging grows as needed)."""
        if not 1 <= max_views <= views.MAX_VIEWS:
            raise ValueError(f"max_views must be in 1..{views.MAX_VIEWS}")
        self.drain()
        if max_views > self._cam_host[0][0].shape[0]:
            dev = self.store.device
            self._cam_host = [(torch.empty((max_views, 4), pin_memory=True),
                               torch.empty((max_views, 12), pin_memory=True)) for _ in self.streams]
            self._cam_dev = [(torch.empty((max_views, 4), device=dev), torch.empty((max_views, 12), device=dev))
                             for _ in self.streams]
            self._cam_copied = [None] * len(self.streams)
        self.max_views = max_views

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

    def _cameras(self, slot: int, reqs: list[Request], stream) -> tuple[torch.Tensor, torch.Tensor]:
        """The batch's [B, 4] intrinsics and [B, 12] extrinsics on the device, without a host wait."""
        B = len(reqs)
        di, de = self._cam_dev[slot]
        if any(r.intr.is_cuda or r.extr.is_cuda for r in reqs):  # already on the device
            with torch.cuda.stream(stream):
                for i, r in enumerate(reqs):
                    di[i].copy_(r.intr.reshape(4))
                    de[i].copy_(r.extr.reshape(-1)[:12])
            return di[:B], de[:B]
        hi, he = self._cam_host[slot]
        if self._cam_copied[slot] is not None:
            self._cam_copied[slot].synchronize()  # the previous copy out of this staging has run
        for i, r in enumerate(reqs):
            hi[i].copy_(r.intr.reshape(4))
            he[i].copy_(r.extr.reshape(-1)[:12])
        with torch.cuda.stream(stream):
            di[:B].copy_(hi[:B], non_blocking=True)
            de[:B].copy_(he[:B], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
        self._cam_copied[slot] = ev
        return di[:B], de[:B]

    def _launch(self, reqs: list[Request], exact: bool = False) -> None:
        st = self.store
        dev = st.device
        nbytes = sum(self._cost(r) for r in reqs)
        slot = self.next_stream % len(self.streams)
        stream = self.streams[slot]
        self.next_stream += 1
        stream.wait_stream(torch.cuda.current_stream(dev))
        sizes = [(r.W, r.H) for r in reqs]
        first = [st.scenes[r.scene].first for r in reqs]
        count = [st.scenes[r.scene].count for r in reqs]
        rows = sum(count)
        capacity = None if exact else int(rows * self.entries_per_row) + self.min_capacity
        cfg = tuning.config_for_views(sizes, st.C)
        intrs, extrs = self._cameras(slot, reqs, stream)
        with torch.cuda.stream(stream), torch.no_grad():
            uv, depth, conic, radius, _ = views.preprocess_views_forward(st.xyz, st.scale, st.rotate, intrs, extrs,
                                                                         sizes, first=first, count=count)
            bins = views.bin_views(uv, depth, radius, st.xyz.shape[0], sizes, cfg.cell, capacity=capacity,
                                   first=first, count=count)
            bg = reqs[0].bg
            out, _, _, _ = views.rasterize_views_forward(uv, conic, st.opacity, st.feature, st.xyz.shape[0], sizes,
                                                         bins, bg, cfg.ppl_fwd, cfg.cull, first=first, count=count)
            ev = torch.cuda.Event()
            ev.record(stream)
        for r in reqs:
            r.batch = self.next_batch
        self.in_flight.append(_Batch(self.next_batch, reqs, ev, out, nbytes, stream, rows, capacity, bins.count))
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
        keep, again, finished = [], [], 0
        for b in self.in_flight:
            if not b.event.query():
                keep.append(b)
                continue
            self.used -= b.nbytes
            if b.capacity is not None:
                k, expired = (int(v) for v in b.count.tolist())  # the batch has finished: no wait
                if expired:  # an earlier batch's binning left undefined lists (its images are NaN)
                    binning.check_waits(self.store.device)
                if k > b.capacity:  # truncated lists: raise the estimate, render again exactly
                    self.entries_per_row = max(self.entries_per_row, 1.25 * k / max(b.rows, 1))
                    self.rerenders += 1
                    again.append(b.reqs)
                    continue
            now = time.perf_counter()
            imgs = views.split_views(b.out, self.store.C, [(r.W, r.H) for r in b.reqs])
            for r, im in zip(b.reqs, imgs):
                r.image = im
                r.t_done = now
                self.done.append(r)
            finished += 1
        self.in_flight = keep
        for reqs in again:
            self._launch(reqs, exact=True)
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
    # the client's arrival times and requests (cameras) are drawn before the
    # clock starts, so that making them does not delay the server loop
    arrivals, t = [], rng.expovariate(rate)
    while t < duration:
        arrivals.append((t, make_request(rng)))
        t += rng.expovariate(rate)
    b0, r0 = server.next_batch, server.rerenders
    t0 = time.perf_counter()
    nxt = 0
    while True:
        now = time.perf_counter()
        while nxt < len(arrivals) and t0 + arrivals[nxt][0] <= now:
            req = arrivals[nxt][1]
            req.t_arrival = t0 + arrivals[nxt][0]
            server.submit(req)
            nxt += 1
        if not server.step():
            if nxt >= len(arrivals) and not server.busy():
                break
            time.sleep(20e-6)
    t1 = time.perf_counter()
    lat = [r.latency * 1e3 for r in server.done]
    pix = sum(r.W * r.H for r in server.done)
    return dict(rate_req_s=rate, duration_s=round(t1 - t0, 3), requests=len(server.done),
                throughput_req_s=round(len(server.done) / (t1 - t0), 2),
                throughput_mpix_s=round(pix / (t1 - t0) / 1e6, 2),
                p50_ms=round(percentile(lat, 0.5), 3), p99_ms=round(percentile(lat, 0.99), 3),
                max_ms=round(max(lat), 3) if lat else None, batches=server.next_batch - b0,
                requests_per_batch=round(len(server.done) / max(server.next_batch - b0, 1), 2),
                rerendered_batches=server.rerenders - r0)
