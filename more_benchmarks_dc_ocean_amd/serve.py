"""Served Redi steps: continuous batching of step requests on one GPU.

A request asks for ``nsteps`` Redi steps of one regional domain and carries
that domain's state: its seeded bathymetry (h) and T/S, its metrics and face
diffusivities -- held in pinned host memory, as a client would hand them
over -- and a seed from which its two passive tracers are drawn.  The
server keeps every admitted request's state resident in HBM and advances
all of them together, one batched step (``RediBatch.step``: one hipGraph
launch) per iteration:

  arrive   requests enter a FIFO wait queue at their arrival time;
  admit    before every iteration, waiting requests are admitted in arrival
           order while a batch slot is free and the request fits the HBM
           budget: its Phase-A state and step scratch in the batch's arena
           (carved, nothing allocated) and its fields in the field budget;
           a request that does not fit yet waits while later, smaller ones
           may pass it, until the budget frees;
  upload   an admitted request's fields are copied host -> HBM on a copy
           stream (overlapping the batched step already running) and its
           passive tracers drawn on the device from its seed; the compute
           stream waits for that copy before the request's first step;
  step     one batched step of every admitted request: requests that arrived
           while others were on the GPU join at the next iteration, so their
           work overlaps on the device with the requests already running;
  retire   a request leaves the batch after its last step is enqueued; its
           completion time is taken when that step's event has completed,
           and its latency is completion - arrival (so it includes the wait,
           the upload and every step).

The baseline (``run_serial``) serves the same requests one at a time, first
come first served, each uploaded and stepped alone by the plain FP64
single-domain entry, in real time against the same arrival times.

The requests' domains come from a pool of seeded regional domains generated
before the run (``pool_cases``), so that the host generates no fractal
bathymetry inside the served loop; each request copies its pool entry's
fields into HBM of its own and draws its own passive tracers.
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
from .redi import DeviceDomain, OceanRedi


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float        # s after the start of the run
    nsteps: int
    domain: int           # index into the domain pool
    seed: int = -1        # seed of its passive tracers (default: rid)
    slot: int = -1
    remaining: int = 0
    admitted: float = math.nan
    done: float = math.nan
    state: Optional[DeviceDomain] = None
    ready: object = None  # event: its upload has landed
    initial: Optional[np.ndarray] = None
    result: Optional[np.ndarray] = None

    def __post_init__(self):
        if self.seed < 0:
            self.seed = self.rid

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
                           domain=int(rng.integers(0, npool)), seed=1_000_000 * (seed + 1) + len(out)))
    return out


class PinnedDomain:
    """One pool domain's fields in pinned host memory, in the storage type of
    ``form``: h, hTr of T and S, and the FP64 2-D arrays
    (DeviceDomain.METRIC_NAMES)."""

    def __init__(self, case: inputs.RediCase, form: str):
        import torch
        if (case.idx_T, case.idx_S) != (1, 2) or case.nt < 2:
            raise ValueError("served domains carry T and S as tracers 1 and 2")
        self.case = case
        self.form = form
        fdt = torch.float32 if form == "fp32_store" else torch.float64
        self.h = torch.from_numpy(case.h).to(fdt).pin_memory()
        self.ts = torch.from_numpy(np.ascontiguousarray(case.hTr[:2])).to(fdt).pin_memory()
        self.metrics = [torch.from_numpy(np.ascontiguousarray(getattr(case, n))).pin_memory()
                        for n in DeviceDomain.METRIC_NAMES]

    def device_bytes(self) -> int:
        """HBM a request of this domain holds for its fields."""
        c = self.case
        return (self.h.numel() * self.h.element_size() * (1 + c.nt)
                + sum(m.numel() * m.element_size() for m in self.metrics))


def upload(pd: PinnedDomain, seed: int, device, stream=None) -> DeviceDomain:
    """The request's own copy of the domain on the device: h, T, S and the
    metrics copied from pinned host memory (asynchronously, on ``stream``),
    and tracers 3.. drawn on the device as U[0,1) x h from a generator seeded
    with ``seed`` (land, where h = 0, gets 0).  Tensors used on another
    stream must be recorded on it by the caller."""
    import torch
    c = pd.case
    dev = torch.device(device)
    with torch.cuda.stream(stream) if stream is not None else _nullcontext():
        h = pd.h.to(dev, non_blocking=True)
        hTr = torch.empty((c.nt,) + tuple(pd.h.shape), dtype=pd.h.dtype, device=dev)
        hTr[:2].copy_(pd.ts, non_blocking=True)
        if c.nt > 2:
            g = torch.Generator(device=dev)
            g.manual_seed(int(seed))
            u = torch.rand((c.nt - 2,) + tuple(pd.h.shape), generator=g, dtype=pd.h.dtype, device=dev)
            torch.mul(u, h, out=hTr[2:])
        metrics = [m.to(dev, non_blocking=True) for m in pd.metrics]
    return DeviceDomain.from_tensors(c, pd.form, dev, h, hTr, metrics)


class _nullcontext:
    def __enter__(self):
        return None

    def __exit__(self, *a):
        return False


class RediServer:
    """Continuous-batching Redi server: requests carry their domain's state
    (a PinnedDomain and a seed); admitted requests are uploaded and stepped
    together."""

    def __init__(self, pool: Sequence[PinnedDomain], budget_bytes: int, form: str = "fp32_store",
                 max_slots: int = 256, shapes: Optional[Sequence[Dict]] = None,
                 device: str = "cuda", keep_results: bool = False):
        import torch
        self.pool = list(pool)
        self.form = form
        self.device = torch.device(device)
        self.keep_results = keep_results
        self.shapes = list(shapes) if shapes else [None] * len(self.pool)
        self.need = [domain_bytes(p.case.nx, p.case.ny, p.case.nz, form, 4) for p in self.pool]
        self.field_bytes = [p.device_bytes() for p in self.pool]
        # the budget is split between the arena (Phase-A state + scratch) and
        # the requests' fields in the pool's mean proportion, with a margin
        # for the fields
        share = float(np.mean([f / (f + a) for f, a in zip(self.field_bytes, self.need)]))
        self.field_cap = int(budget_bytes * min(0.5, 1.25 * share))
        self.arena_bytes = int(budget_bytes - self.field_cap)
        self.batch = RediBatch(form=form, group=4, max_domains=max_slots,
                               arena_bytes=self.arena_bytes, persist=1, device=device)
        self.max_slots = max_slots
        self.field_used = 0
        self.iterations = 0
        self.max_active = 0
        self.copy_stream = torch.cuda.Stream(device=self.device)

    def close(self) -> None:
        self.batch.close()

    def _fits(self, r: Request, active: int) -> bool:
        if active >= self.max_slots:
            return False
        if self.field_used + self.field_bytes[r.domain] > self.field_cap:
            return False
        return self.batch.info()["largest_free_bytes"] >= self.need[r.domain]

    def _admit(self, r: Request, now: float, stream) -> None:
        import torch
        st = upload(self.pool[r.domain], r.seed, self.device, self.copy_stream)
        ev = torch.cuda.Event()
        ev.record(self.copy_stream)
        for t in (st.h, st.hTr, st.dy_cu, st.dx_cv, st.idxCu, st.idyCv, st.areaT, st.wet_u,
                  st.wet_v, st.khtr_u_ext, st.khtr_v_ext):
            t.record_stream(stream)  # used by the compute stream until the request retires
        r.ready = ev
        r.slot = self.batch.add(st, self.shapes[r.domain])
        r.state = st
        r.remaining = r.nsteps
        r.admitted = now
        self.field_used += self.field_bytes[r.domain]

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
                    self.field_used -= self.field_bytes[r.domain]
                block = False

        while pending or waiting or active or inflight:
            now = clock()
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            admitted = []
            for r in waiting:
                if self._fits(r, len(active)):
                    self._admit(r, now, stream)
                    active[r.slot] = r
                    admitted.append(r)
            if admitted:
                waiting = [r for r in waiting if r.slot < 0]
                for r in admitted:  # the step waits for the new requests' uploads
                    stream.wait_event(r.ready)
                    if self.keep_results:
                        r.initial = r.state.tracers_host()
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


def run_serial(requests: Sequence[Request], pool: Sequence[PinnedDomain], device: str = "cuda",
               log=None, keep_results: bool = False) -> Dict:
    """The baseline: the same requests served one at a time in arrival order
    (first come, first served) by the plain FP64 single-domain entry, in
    real time: each is uploaded and stepped alone once it has arrived and
    the previous one is done.  One ocean_redi_t per pool domain is created
    beforehand (its state sized for the domain), as a serial solver would
    keep one per domain size."""
    import torch
    if any(p.form != "fp64" for p in pool):
        raise ValueError("the serial baseline runs the plain FP64 form")
    solvers: Dict[int, OceanRedi] = {}
    for i in sorted({r.domain for r in requests}):
        o = OceanRedi(pool[i].case, device=device)
        o.configure(enable=True, khtr=pool[i].case.khtr)
        solvers[i] = o
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    clock = lambda: time.perf_counter() - t0
    next_log = 5.0
    for r in sorted(requests, key=lambda r: r.arrival):
        wait = r.arrival - clock()
        if wait > 0:
            time.sleep(wait)
        r.admitted = clock()
        dom = upload(pool[r.domain], r.seed, device)
        if keep_results:
            r.initial = dom.tracers_host()
        solvers[r.domain].step(dom, r.nsteps)
        torch.cuda.synchronize()
        r.done = clock()
        if keep_results:
            r.result = dom.tracers_host()
        del dom
        if log and clock() >= next_log:
            log(f"serial t={clock():6.1f}s done={r.rid + 1}")
            next_log += 5.0
    for o in solvers.values():
        o.close()
    return summarize(requests, pool, clock())


def cell_steps(r: Request, pool) -> int:
    c = pool[r.domain].case
    return c.interior_cells * r.nsteps


def summarize(requests: Sequence[Request], pool, wall: float) -> Dict:
    lat = np.array([r.latency for r in requests])
    work = sum(cell_steps(r, pool) for r in requests)
    span = max(r.done for r in requests) - min(r.arrival for r in requests)
    return {"requests": len(requests), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
            "max_ms": float(lat.max() * 1e3), "cell_steps": int(work),
            "value": work / span / 1e6, "unit": "Mcell-steps/s", "makespan_s": span, "wall_s": wall}
