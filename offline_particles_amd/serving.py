"""Serving form: continuous batching of particle releases on one device.

Requests (a release: seed positions and a number of AB3 steps) arrive over
wall-clock time.  ``ReleaseServer.run`` is the request loop:

- **admission by free HBM**: the particle buffers are sized once from the
  device memory free at start-up (``hbm_fraction`` of it, less the snapshot
  window); a queued request is admitted, first come first served, when its
  particles fit the free particle slots and a release id is free.  Slots of
  finished releases are reclaimed by the next re-bin (compaction);
- every admitted release starts at the **service clock**, the simulated time
  all running releases share, so they all read one resident (or streamed,
  3-slot) snapshot window;
- every running release advances in the same fused launch per step (graph
  replays of up to ``graph_steps`` steps); new requests join at chunk
  boundaries while the others keep running;
- a release that has run its steps is extracted (``opx_extract_release``)
  into pinned host memory, retired, and its completion event recorded.

Latency of a request = completion of its result copy (GPU event, on the wall
clock of the loop) - its arrival.  The reference has no serving layer; each
release's trajectory is exactly what ``AB3Engine`` gives it alone (the fused
step is per particle and its clock per release).
"""

from __future__ import annotations

import heapq
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from .engine import AB3Engine, SplitState

# device bytes per particle slot: state + spare state (re-bin target) + re-bin
# keys/values (4 x uint32) + radix-sort temporary (bounded by 8 B)
SLOT_BYTES = 2 * SplitState.BYTES_PER_PARTICLE + 16 + 8


@dataclass
class Request:
    """One release request.  ``arrival`` is in seconds from the start of the loop
    (wall clock); with ``arrival_step`` set instead, the request arrives when the
    service reaches that global step (deterministic mode, for tests)."""

    zidx: np.ndarray
    yidx: np.ndarray
    xidx: np.ndarray
    nsteps: int
    arrival: float = 0.0
    arrival_step: int | None = None
    # filled by the server
    release: int = -1
    start_step: int = -1
    start_time: float = float("nan")
    admitted_at: float = float("nan")
    completed_at: float = float("nan")
    result: dict | None = None
    _done: torch.cuda.Event | None = field(default=None, repr=False)

    @property
    def size(self) -> int:
        return int(len(self.zidx))

    @property
    def latency(self) -> float:
        return self.completed_at - self.arrival


class AdmissionQueue:
    """First-come-first-served admission against free particle slots and release ids."""

    def __init__(self, capacity: int, max_releases: int):
        self.capacity = int(capacity)
        self.max_releases = int(max_releases)
        self._q: deque = deque()

    def __len__(self) -> int:
        return len(self._q)

    def submit(self, req: Request) -> None:
        if req.size > self.capacity:
            raise ValueError(f"request of {req.size} particles exceeds the capacity of {self.capacity}")
        self._q.append(req)

    def head_fits(self, used_slots: int, used_ids: int) -> bool:
        """Whether the oldest queued request fits `used_slots` particle slots and
        `used_ids` release ids in use."""
        return bool(self._q) and used_ids < self.max_releases and used_slots + self._q[0].size <= self.capacity

    def pop(self) -> Request:
        return self._q.popleft()

    def head(self) -> Request | None:
        return self._q[0] if self._q else None


def capacity_from_free_hbm(device, *, hbm_fraction: float = 0.8, reserve_bytes: int = 0,
                           max_particles: int | None = None) -> int:
    """Particle slots that ``hbm_fraction`` of the free device memory (less
    ``reserve_bytes``) holds at SLOT_BYTES each, capped at ``max_particles``."""
    free, _ = torch.cuda.mem_get_info(device)
    cap = int(max(0, free * hbm_fraction - reserve_bytes) // SLOT_BYTES)
    cap = min(cap, (1 << 31) - 1)
    if max_particles is not None:
        cap = min(cap, int(max_particles))
    return cap


class ReleaseServer:
    """Request loop over one AB3Engine (see the module docstring)."""

    def __init__(self, fieldset, time_array, dt, *, storage_dtype=torch.bfloat16, device=None,
                 stream_slots: int | None = 3, hbm_fraction: float = 0.8, max_particles: int | None = None,
                 max_releases: int = 1024, rebin_every: int = 40, xlow: int = 4, graph_steps: int = 8,
                 layout: int | None = None, block: int = 128, start_time: float | None = None,
                 max_inflight_chunks: int = 3, launch_steps: int = 1, stream_tiles: int | None = None):
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        # the window (rings and packed stencil windows) is allocated first, then
        # the particle buffers from what is left
        self.engine = AB3Engine(fieldset, time_array, dt, storage_dtype=storage_dtype, device=self.device,
                                capacity=1, rebin_every=rebin_every, xlow=xlow, max_releases=max_releases,
                                stream_slots=stream_slots, stream_tiles=stream_tiles)
        eng = self.engine
        if layout is not None:
            eng.set_layout(layout)
        eng.block = int(block)
        eng.graph_steps = int(graph_steps)
        eng.launch_steps = int(launch_steps)
        bu, bv, bw, b2 = eng.packed_window_bytes()
        # the pair ring: one packed window per ring slot
        self.capacity = capacity_from_free_hbm(self.device, hbm_fraction=hbm_fraction,
                                               reserve_bytes=eng.nslots * (bu + bv + bw + b2) + (256 << 20),
                                               max_particles=max_particles)
        eng.reserve(self.capacity)
        self.queue = AdmissionQueue(self.capacity, max_releases)
        self.now = float(eng.time_array[0] if start_time is None else start_time)
        self.max_inflight_chunks = int(max_inflight_chunks)
        self.stats: dict = {}

    # the service clock advances as every release clock does: t += dt (timesteppers.py:157-161)
    def _advance_clock(self, k: int) -> None:
        for _ in range(k):
            self.now = self.now + self.engine.dt

    def _admit(self, req: Request, wall: float) -> None:
        eng = self.engine
        t_end = self.now + req.nsteps * eng.dt
        if t_end > eng.time_array[-1]:
            raise ValueError(f"a request of {req.nsteps} steps admitted at t = {self.now} would run past the end "
                             f"of the time array ({eng.time_array[-1]})")
        req.start_step = eng.global_step
        req.start_time = self.now
        req.admitted_at = wall
        req.release = eng.add_release(req.zidx, req.yidx, req.xidx, start_time=self.now, nsteps=req.nsteps,
                                      start_step=eng.global_step)

    def _finish(self, req: Request) -> None:
        """Copy the release's results to pinned host memory and retire it."""
        eng = self.engine
        z, y, x, st = eng.extract(req.release)
        out = {}
        for name, t in (("zidx", z), ("yidx", y), ("xidx", x), ("status", st)):
            h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
            h.copy_(t, non_blocking=True)
            out[name] = h
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        req.result = out
        req._done = ev
        eng.retire(req.release)

    def run(self, requests, *, realtime: bool = True) -> list:
        """Serve `requests` to completion; returns them with results and latencies.

        realtime: requests arrive at their ``arrival`` wall-clock second (the
        loop waits for the next one when idle); otherwise at ``arrival_step``.
        """
        eng = self.engine
        key = (lambda r: r.arrival) if realtime else (lambda r: r.arrival_step)
        pending = deque(sorted(requests, key=key))
        running: dict = {}
        ends: list = []  # heap of (end global step, release id)
        done: list = []
        inflight: deque = deque()
        peak_live = 0
        compactions = 0
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        ref = torch.cuda.Event(enable_timing=True)
        ref.record()
        while pending or len(self.queue) or running:
            wall = time.perf_counter() - t0
            while pending and (pending[0].arrival <= wall if realtime else pending[0].arrival_step <= eng.global_step):
                self.queue.submit(pending.popleft())
            # admission: FIFO while the head fits; compact first when retired
            # releases hold the slots or ids it needs
            while len(self.queue):
                if not self.queue.head_fits(eng.n, eng.releases_in_use) and eng.retired_releases:
                    eng.rebin()
                    compactions += 1
                if not self.queue.head_fits(eng.n, eng.releases_in_use):
                    if not running:
                        raise RuntimeError("the queue head does not fit an idle server")
                    break
                req = self.queue.pop()
                self._admit(req, wall)
                running[req.release] = req
                heapq.heappush(ends, (req.start_step + req.nsteps, req.release))
            peak_live = max(peak_live, eng.n - eng.retired_particles)
            if not running:
                if pending and realtime:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - (time.perf_counter() - t0))))
                elif pending:
                    # deterministic mode: the service clock runs with nothing to step
                    k = pending[0].arrival_step - eng.global_step
                    eng.step(k)
                    self._advance_clock(k)
                continue
            # run until the next release ends (or a graph chunk), then hand results back
            k = min(max(1, eng.graph_steps), ends[0][0] - eng.global_step)
            if pending and not realtime:
                k = min(k, max(1, pending[0].arrival_step - eng.global_step))
            eng.step(k)
            self._advance_clock(k)
            while ends and ends[0][0] <= eng.global_step:
                _, r = heapq.heappop(ends)
                req = running.pop(r)
                self._finish(req)
                done.append(req)
            # bound how far the host runs ahead of the device, so admissions
            # reach the device soon after the requests arrive
            ev = torch.cuda.Event()
            ev.record()
            inflight.append(ev)
            while len(inflight) > self.max_inflight_chunks:
                inflight.popleft().synchronize()
        torch.cuda.synchronize(self.device)
        elapsed = time.perf_counter() - t0
        for req in done:
            req.completed_at = ref.elapsed_time(req._done) / 1e3
            req._done = None
        eng.check_clock()
        lat = np.array([r.latency for r in done]) if done else np.zeros(1)
        work = sum(r.size * r.nsteps for r in done)
        self.stats = {
            "requests": len(done),
            "particles": int(sum(r.size for r in done)),
            "particle_steps": int(work),
            "seconds": elapsed,
            "particle_steps_per_s": work / elapsed if elapsed > 0 else 0.0,
            # the work the requests ask for per second of their arrival span: no
            # server delivers more than this over the span (realtime mode)
            "offered_particle_steps_per_s": (work / max(r.arrival for r in done)
                                             if realtime and done and max(r.arrival for r in done) > 0 else None),
            "last_arrival_s": float(max(r.arrival for r in done)) if realtime and done else None,
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_max_s": float(lat.max()),
            "queue_wait_p99_s": float(np.percentile([r.admitted_at - r.arrival for r in done], 99)) if done else 0.0,
            "global_steps": eng.global_step,
            "capacity_particles": self.capacity,
            "peak_live_particles": int(peak_live),
            "compactions": compactions,
            "snapshot_loads": getattr(eng, "snapshot_loads", 0),
            "snapshot_h2d_GBs": getattr(eng, "snapshot_bytes", 0) / elapsed / 1e9 if elapsed > 0 else 0.0,
            "snapshot_bytes": int(getattr(eng, "snapshot_bytes", 0)),
            "snapshot_full_bytes": int(getattr(eng, "snapshot_full_bytes", 0)),
            "stream_tiles": eng._tile,
            "tile_topups": int(getattr(eng, "tile_topups", 0)),
        }
        return done


def poisson_requests(wl, *, rate: float = 20.0, duration: float = 10.0, size_range=(1_000, 2_000_000),
                     steps_range=(720, 7200), box_cells: int = 10, seed: int = 7) -> list:
    """Requests arriving as a Poisson process (``rate`` per second over ``duration``
    seconds), sizes log-uniform on ``size_range``, steps uniform on ``steps_range``,
    each seeded uniformly in a random box of ``box_cells`` cells (10 km at 1 km
    spacing) over the full depth, boxes centred on water."""
    rng = np.random.default_rng(seed)
    times = []
    t = rng.exponential(1.0 / rate)
    while t < duration:
        times.append(t)
        t += rng.exponential(1.0 / rate)
    h = wl.fields["h"]
    water = np.argwhere(h > float(np.quantile(h, 0.05)))
    reqs = []
    lo, hi = np.log(size_range[0]), np.log(size_range[1])
    for t in times:
        m = int(np.exp(rng.uniform(lo, hi)))
        j, i = water[rng.integers(len(water))]
        y0 = float(np.clip(j - box_cells / 2, 1, wl.NY - 2 - box_cells))
        x0 = float(np.clip(i - box_cells / 2, 1, wl.NX - 2 - box_cells))
        reqs.append(Request(zidx=rng.uniform(0.5, wl.NZ - 1.5, m), yidx=y0 + rng.uniform(0, box_cells, m),
                            xidx=x0 + rng.uniform(0, box_cells, m),
                            nsteps=int(rng.integers(steps_range[0], steps_range[1] + 1)), arrival=float(t)))
    return reqs
