"""Served Redi steps: continuous batching of step requests on one GPU.

A request asks for ``nsteps`` Redi steps of one regional domain.  The
server keeps every admitted request's state resident in HBM and advances
all of them together, one batched step (``RediBatch.step``: one hipGraph
launch) per iteration:

  arrive   requests enter a FIFO wait queue at their arrival time;
  admit    before every iteration, waiting requests are admitted in arrival
           order while a batch slot is free and the request fits the HBM
           budget: its Phase-A state and step scratch in the batch's arena
           (carved, nothing allocated) and its tracers in the tracer budget;
           a request that does not fit yet waits while later, smaller ones
           may pass it, until the budget frees;
  step     one batched step of every admitted request: requests that arrived
           while others were on the GPU join at the next iteration, so their
           work overlaps on the device with the requests already running;
  retire   a request leaves the batch after its last step is enqueued; its
           completion time is taken when that step's event has completed,
           and its latency is completion - arrival.

Requests carry their inputs as a reference to one of a pool of seeded
regional domains resident on the device (h, metrics and face diffusivities
are read-only in a Redi step and are shared; each request gets its own copy
of the tracers), so the host does no per-request input generation or PCIe
transfer inside the served loop.

Baseline: the same request sequence served one request at a time by the
plain FP64 single-domain step (FIFO, no batching), with each request's
service time = its steps x the measured FP64 step time of its domain.
"""
from __future__ import annotations

import collections
import dataclasses
import math
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import inputs
from .batch import RediBatch, domain_bytes
from .redi import DeviceDomain


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float        # s after the start of the run
    nsteps: int
    domain: int           # index into the domain pool
    slot: int = -1
    remaining: int = 0
    admitted: float = math.nan
    done: float = math.nan
    state: Optional[DeviceDomain] = None
    result: Optional[np.ndarray] = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def pool_cases(n: int, seed: int = 0, lo: int = 64, hi: int = 720,
               nzs: Sequence[int] = (25, 50, 75), nt: int = 4) -> List[inputs.RediCase]:
    """``n`` seeded regional domains: square side log-uniform in [lo, hi],
    nz uniform over ``nzs`` (case_domain physics)."""
    sizes = inputs.batch_sizes(n, seed=1000 + seed, lo=lo, hi=hi, nzs=nzs, cfg=4)
    return [inputs.case_domain(4, seed * 1000 + i, s, s, k, nt=nt)
            for i, (s, k) in enumerate(sizes)]


def poisson_requests(rate: float, duration: float, npool: int, seed: int = 0,
                     max_steps: int = 20) -> List[Request]:
    """Poisson arrivals at ``rate``/s over ``duration`` s; each request asks
    for U{1..max_steps} steps of a uniformly drawn pool domain."""
    rng = np.random.default_rng([4, 77, seed])
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        out.append(Request(rid=len(out), arrival=t, nsteps=int(rng.integers(1, max_steps + 1)),
                           domain=int(rng.integers(0, npool))))
    return out


class RediServer:
    """Continuous-batching Redi server over a pool of device-resident domains."""

    def __init__(self, pool: Sequence[DeviceDomain], budget_bytes: int, form: str = "fp32_store",
                 max_slots: int = 256, shapes: Optional[Sequence[Dict]] = None,
                 device: str = "cuda", keep_results: bool = False):
        import torch
        self.pool = list(pool)
        self.form = form
        self.device = torch.device(device)
        self.keep_results = keep_results
        self.shapes = list(shapes) if shapes else [None] * len(self.pool)
        pool_bytes = sum(_field_bytes(d) for d in self.pool)
        free = budget_bytes - pool_bytes
        if free <= 0:
            raise ValueError("the domain pool alone exceeds the HBM budget")
        # arena (Phase-A state + scratch, ~280 B per cell in the FP32-store form)
        # and per-request tracer copies (16 B per cell) share what is left
        self.tracer_cap = int(free * 0.06)
        self.arena_bytes = free - self.tracer_cap
        self.need = [domain_bytes(d.case.nx, d.case.ny, d.case.nz, form, 4) for d in self.pool]
        self.tracer_bytes = [d.hTr.numel() * d.hTr.element_size() for d in self.pool]
        self.batch = RediBatch(form=form, group=4, max_domains=max_slots,
                               arena_bytes=self.arena_bytes, persist=1, device=device)
        self.max_slots = max_slots
        self.tracer_used = 0
        self.iterations = 0
        self.max_active = 0

    def close(self) -> None:
        self.batch.close()

    def _fits(self, r: Request, active: int) -> bool:
        if active >= self.max_slots:
            return False
        if self.tracer_used + self.tracer_bytes[r.domain] > self.tracer_cap:
            return False
        return self.batch.info()["largest_free_bytes"] >= self.need[r.domain]

    def _admit(self, r: Request, now: float) -> None:
        st = self.pool[r.domain].clone_tracers()
        r.slot = self.batch.add(st, self.shapes[r.domain])
        r.state = st
        r.remaining = r.nsteps
        r.admitted = now
        self.tracer_used += self.tracer_bytes[r.domain]

    def run(self, requests: Sequence[Request], log=None) -> Dict:
        """Serve ``requests`` (arrival times in s from now); returns stats."""
        import torch
        stream = torch.cuda.current_stream(self.device)
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting: List[Request] = []
        active: Dict[int, Request] = {}
        inflight = collections.deque()  # (event, requests whose last step it ends)
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0
        next_log = 5.0

        def retire(block: bool) -> None:
            while inflight and (block or inflight[0][0].query()):
                ev, fin = inflight.popleft()
                ev.synchronize()
                t = clock()
                for r in fin:
                    r.done = t
                    if self.keep_results:
                        r.result = r.state.tracers_host()
                    r.state = None
                    self.tracer_used -= self.tracer_bytes[r.domain]
                block = False

        while pending or waiting or active or inflight:
            now = clock()
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            admitted = []
            for r in waiting:
                if self._fits(r, len(active)):
                    self._admit(r, now)
                    active[r.slot] = r
                    admitted.append(r)
            if admitted:
                waiting = [r for r in waiting if r.slot < 0]
            if active:
                self.batch.step(1)
                ev = torch.cuda.Event()
                ev.record(stream)
                self.iterations += 1
                self.max_active = max(self.max_active, len(active))
                fin = []
                for slot, r in list(active.items()):
                    r.remaining -= 1
                    if r.remaining == 0:
                        fin.append(r)
                        self.batch.remove(slot)
                        del active[slot]
                inflight.append((ev, fin))
                # at most two steps in flight: the host prepares the next
                # iteration while the GPU runs the current one
                retire(block=len(inflight) > 2)
            elif inflight:
                retire(block=True)
            elif pending:
                time.sleep(max(0.0, min(pending[0].arrival - clock(), 0.01)))
            if log and clock() >= next_log:
                log(f"t={clock():6.1f}s active={len(active)} waiting={len(waiting)} "
                    f"pending={len(pending)} done={sum(1 for r in requests if not math.isnan(r.done))}")
                next_log += 5.0
        torch.cuda.synchronize()
        return summarize(requests, self.pool, clock())


def _field_bytes(d: DeviceDomain) -> int:
    ts = [d.h, d.hTr, d.dy_cu, d.dx_cv, d.idxCu, d.idyCv, d.areaT, d.wet_u, d.wet_v,
          d.khtr_u_ext, d.khtr_v_ext]
    return sum(t.numel() * t.element_size() for t in ts)


def cell_steps(r: Request, pool) -> int:
    c = pool[r.domain].case if hasattr(pool[r.domain], "case") else pool[r.domain]
    return c.interior_cells * r.nsteps


def summarize(requests: Sequence[Request], pool, wall: float) -> Dict:
    lat = np.array([r.latency for r in requests])
    work = sum(cell_steps(r, pool) for r in requests)
    span = max(r.done for r in requests) - min(r.arrival for r in requests)
    return {"requests": len(requests), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
            "max_ms": float(lat.max() * 1e3), "cell_steps": int(work),
            "value": work / span / 1e6, "unit": "Mcell-steps/s", "makespan_s": span, "wall_s": wall}


def fifo_serial(requests: Sequence[Request], step_s: Sequence[float], pool) -> Dict:
    """Latencies of serving ``requests`` one at a time in arrival order, each
    taking nsteps x step_s[domain] seconds (Lindley's recursion)."""
    free_at, lat = 0.0, []
    for r in sorted(requests, key=lambda r: r.arrival):
        start = max(free_at, r.arrival)
        free_at = start + r.nsteps * step_s[r.domain]
        lat.append(free_at - r.arrival)
    lat = np.array(lat)
    work = sum(cell_steps(r, pool) for r in requests)
    span = free_at - min(r.arrival for r in requests)
    return {"requests": len(requests), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
            "cell_steps": int(work), "value": work / span / 1e6, "unit": "Mcell-steps/s",
            "makespan_s": span}
