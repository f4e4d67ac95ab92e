"""Serving loop: shot requests (reference parameter files) admitted by their
HBM footprint and run concurrently, each on its own handle and HIP stream.

A request moves through three stages:
  setup      on arrival, a CPU worker reads its parameter and model files
             (libsofi_host) and computes its device footprint
             (sofi_estimate_footprint); a request larger than the whole
             budget is rejected here
  admission  requests wait in arrival order; any waiting request whose
             footprint fits the free part of the HBM budget is admitted (FIFO
             with backfill), at most `gpu_workers` at a time
  run        a GPU worker opens the handle, uploads the medium and advances
             the shot in blocks of `block_steps` time steps; the wavefield,
             medium and seismograms stay resident on the device between
             blocks, and a cancelled request stops at the next block
             boundary; the seismograms are copied back at the end and the
             footprint is released when the handle closes
Requests that arrive while others run are admitted as soon as a worker and
memory are free, and their kernels overlap with the running ones on the
device (one stream per handle).  Latency is completion time minus arrival
time.  A request's results are those of the same request run alone: handles
share nothing but the GPU, and a block boundary does not change a step.

With worklist=True, small requests (one shot, at most lane_cells cells, a
rheology whose update kernels run one thread per cell by default: VEL_ISO,
VEL_VTI, AC_ISO) are admitted by footprint alone into a lane per FD order and
rheology instead of a GPU worker.  A lane advances all its requests in blocks
of lane_steps steps with sofi_run_many: each step's velocity and stress
updates of all of them are one launch over a (handle, tile) work-list.
Requests join a lane at the next block boundary and leave it when done, so
requests that arrive while others run are batched into the same launches.
Their results are still those of the request run alone, bit for bit.

The reference has no serving layer (it runs the shots of one parameter file
one after the other, src/sofi2D.c:320-395).
"""
from __future__ import annotations

import ctypes as C
import os
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Callable

import numpy as np


@dataclass
class Request:
    rid: int
    json_path: str
    arrival: float = 0.0                   # seconds after the start of the run
    footprint: int = 0
    ready: float | None = None             # setup done
    admitted: float | None = None
    end: float | None = None
    gpu_ms: float = 0.0
    cell_steps: float = 0.0
    sections: dict = field(default_factory=dict)
    error: str | None = None
    cancelled: bool = False

    @property
    def latency(self) -> float | None:
        return None if self.end is None else self.end - self.arrival


def device_free_bytes() -> int:
    """Free HBM of the current device (hipMemGetInfo)."""
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    if hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
        raise RuntimeError("hipMemGetInfo failed")
    return int(free.value)


def open_problem(req: Request):
    """Read the request's parameter, model and acquisition files.  The
    reference's parameter files name their inputs relative to the run
    directory: the host library resolves them against the request's
    directory, so setups of several requests run at once without touching
    the process's working directory."""
    from .sofi import Problem
    d = os.path.dirname(os.path.abspath(req.json_path))
    return Problem(os.path.basename(req.json_path), base_dir=d)


def gpu_prepare(req: Request):
    """Setup stage: the problem and its device footprint (one slot, the
    problem's receivers, which every shot shares, no snapshots)."""
    from . import _native as N
    P = open_problem(req)
    try:
        ntr = len(P.receivers()[0])
        fp = int(N.hip().sofi_estimate_footprint(C.byref(P.desc), 1, ntr, 0))
    except Exception:
        P.close()
        raise
    return fp, P


def gpu_execute(req: Request, P, should_stop: Callable[[], bool], launch: str = "auto",
                block_steps: int = 500) -> None:
    """Run stage: every shot of the request on one handle, in blocks of
    block_steps time steps; fills req.sections[ishot][comp]."""
    from .sofi import Simulation
    sim = None
    try:
        sim = Simulation(P)
        sim.set_launch(launch)
        for ishot in range(1, P.nshots + 1):
            if ishot > 1:
                P.set_shot(ishot)
            sim.load_shot(nsnap=0)
            nt = P.desc.nt
            for t0 in range(1, nt + 1, block_steps):
                if should_stop():
                    return
                req.gpu_ms += sim.run(t0, min(nt, t0 + block_steps - 1))
            req.cell_steps += float(P.desc.nx) * P.desc.ny * nt
            req.sections[ishot] = {c: sim.section(c) for c in P.seismo_comps()}
    finally:
        if sim is not None:
            sim.close()
        P.close()


class Server:
    """Admission by footprint and concurrent execution of shot requests.

    prepare(req) -> (footprint_bytes, state) and execute(req, state,
    should_stop) are the two stages (the GPU ones by default); execute owns
    state and releases it before returning."""

    def __init__(self, budget_bytes: int | None = None, gpu_workers: int = 4, setup_workers: int = 4,
                 launch: str = "auto", block_steps: int = 500,
                 prepare: Callable | None = None, execute: Callable | None = None,
                 worklist: bool = False, lane_cells: int = 1 << 21, lane_steps: int = 250, lane_size: int = 32):
        self.budget = int(budget_bytes if budget_bytes is not None else 0.85 * device_free_bytes())
        self.gpu_workers = gpu_workers
        self.setup_workers = setup_workers
        self.prepare = prepare or gpu_prepare
        self.execute = execute or (lambda r, s, stop: gpu_execute(r, s, stop, launch, block_steps))
        self.worklist = worklist and launch == "auto" and execute is None
        self.lane_cells = lane_cells
        self.lane_steps = lane_steps
        self.lane_size = lane_size
        self.lane_requests = 0                     # requests run in a lane
        self.used = 0
        self.peak_used = 0
        self.admit_order: list[int] = []
        self._cv = threading.Condition()
        self._cancel: set[int] = set()
        self._lanes: dict = {}                     # (fdorder, weq) -> requests waiting to join
        self._lane_n: dict = {}                    # (fdorder, weq) -> requests in the lane (joined or waiting)

    # -- work-list lanes -------------------------------------------------------
    def _lane_key(self, state):
        """The lane of a prepared request, or None when it runs on a worker."""
        if not self.worklist or state is None or not hasattr(state, "desc"):
            return None
        d = state.desc
        if (state.nshots != 1 or d.mode != 1 or d.fdorder_time != 2 or d.nx * d.ny > self.lane_cells
                or d.weq not in (0, 3, 5)):          # AC_ISO, VEL_ISO, VEL_VTI (one thread per cell by default)
            return None
        return (d.fdorder, d.weq)

    def _lane(self, key, t0: float):
        """Advance the lane's requests together until it has none left."""
        from .sofi import Simulation, run_many
        members = []                               # [req, problem, simulation, next step]

        def finish(m, err=None):
            req, P, sim = m[0], m[1], m[2]
            if err is None and sim is not None:
                req.sections[1] = {c: sim.section(c) for c in P.seismo_comps()}
                req.cell_steps += float(P.desc.nx) * P.desc.ny * P.desc.nt
            if sim is not None:
                sim.close()
            P.close()
            req.end = time.perf_counter() - t0
            with self._cv:
                if req.rid in self._cancel:
                    req.cancelled = True
                    err = err or "cancelled"
                req.error = err
                self.used -= req.footprint
                self._lane_n[key] -= 1
                self._done += 1
                self._cv.notify_all()

        while True:
            with self._cv:
                new = self._lanes.get(key, [])
                self._lanes[key] = []
                if not new and not members:
                    del self._lanes[key]
                    return
            for req, P in new:
                try:
                    sim = Simulation(P)
                    sim.load_shot(nsnap=0)
                    members.append([req, P, sim, 1])
                except Exception as e:  # noqa: BLE001
                    finish([req, P, None, 0], str(e))
            for m in [m for m in members if m[0].rid in self._cancel]:
                members.remove(m)
                finish(m, "cancelled")
            if not members:
                continue
            first = [m[3] for m in members]
            last = [min(m[1].desc.nt, m[3] + self.lane_steps - 1) for m in members]
            try:
                ms = run_many([m[2] for m in members], first, last)
            except Exception as e:  # noqa: BLE001
                for m in members:
                    finish(m, str(e))
                members = []
                continue
            for m, l in zip(members, last):
                m[0].gpu_ms += ms
                m[3] = l + 1
            for m in [m for m in members if m[3] > m[1].desc.nt]:
                members.remove(m)
                finish(m)

    def cancel(self, rid: int) -> None:
        """Stop request rid: dropped if it still waits, stopped at its next
        block boundary if it runs.  Its memory returns to the budget."""
        with self._cv:
            self._cancel.add(rid)
            self._cv.notify_all()

    # -- stages --------------------------------------------------------------
    def _setup(self, req: Request, t0: float, ready: list):
        state = None
        try:
            req.footprint, state = self.prepare(req)
            if req.footprint > self.budget:
                raise RuntimeError(f"request {req.rid} needs {req.footprint} B, "
                                   f"more than the {self.budget} B budget")
        except Exception as e:  # noqa: BLE001
            req.error = str(e)
            if state is not None and hasattr(state, "close"):
                state.close()
            state = None
        req.ready = time.perf_counter() - t0
        with self._cv:
            ready.append((req, state))
            self._cv.notify_all()

    def _run(self, req: Request, state, t0: float):
        try:
            self.execute(req, state, lambda: req.rid in self._cancel)
        except Exception as e:  # noqa: BLE001
            req.error = str(e)
        finally:
            req.end = time.perf_counter() - t0
            with self._cv:
                if req.rid in self._cancel:
                    req.cancelled = True
                    req.error = req.error or "cancelled"
                self.used -= req.footprint
                self._running -= 1
                self._done += 1
                self._cv.notify_all()

    # -- the loop --------------------------------------------------------------
    def serve(self, requests: list[Request]) -> dict:
        """Run the requests, each released at its arrival time; returns the
        throughput and latency statistics (requests are updated in place)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        ready: list = []
        self._running = 0
        self._done = 0
        t0 = time.perf_counter()
        setup_pool = ThreadPoolExecutor(self.setup_workers)
        gpu_pool = ThreadPoolExecutor(self.gpu_workers)

        def feeder():
            for r in reqs:
                dt = r.arrival - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                setup_pool.submit(self._setup, r, t0, ready)

        threading.Thread(target=feeder, daemon=True).start()
        with self._cv:
            while self._done < len(reqs):
                ready.sort(key=lambda rs: rs[0].arrival)
                k = 0
                while k < len(ready):
                    r, state = ready[k]
                    if r.error is not None or r.rid in self._cancel:
                        ready.pop(k)
                        if state is not None and hasattr(state, "close"):
                            state.close()
                        r.cancelled = r.rid in self._cancel
                        r.error = r.error or "cancelled"
                        r.end = time.perf_counter() - t0
                        self._done += 1
                        continue
                    # a small request into its work-list lane, by footprint alone
                    key = self._lane_key(state)
                    if (key is not None and self.used + r.footprint <= self.budget
                            and self._lane_n.get(key, 0) < self.lane_size):
                        ready.pop(k)
                        self.used += r.footprint
                        self.peak_used = max(self.peak_used, self.used)
                        r.admitted = time.perf_counter() - t0
                        self.admit_order.append(r.rid)
                        self.lane_requests += 1
                        self._lane_n[key] = self._lane_n.get(key, 0) + 1
                        if key in self._lanes:
                            self._lanes[key].append((r, state))
                        else:
                            self._lanes[key] = [(r, state)]
                            threading.Thread(target=self._lane, args=(key, t0), daemon=True).start()
                        continue
                    # first waiting request that fits, in arrival order (backfill)
                    if key is None and self._running < self.gpu_workers and self.used + r.footprint <= self.budget:
                        ready.pop(k)
                        self.used += r.footprint
                        self.peak_used = max(self.peak_used, self.used)
                        self._running += 1
                        r.admitted = time.perf_counter() - t0
                        self.admit_order.append(r.rid)
                        gpu_pool.submit(self._run, r, state, t0)
                        continue
                    k += 1
                self._cv.wait(0.05)
        setup_pool.shutdown()
        gpu_pool.shutdown()
        return summarize(reqs)


def summarize(reqs: list[Request]) -> dict:
    ok = [r for r in reqs if r.error is None]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    span = max(r.end for r in reqs) - min(r.arrival for r in reqs)
    return {
        "requests": len(reqs), "completed": len(ok), "span_s": span,
        "shots_per_s": sum(len(r.sections) for r in ok) / span if span > 0 else 0.0,
        "glups": sum(r.cell_steps for r in ok) / span / 1e9 if span > 0 else 0.0,
        "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
    }


def run_sequential(requests: list[Request], launch: str = "auto") -> dict:
    """The same requests one after the other (setup, run, results) as they
    arrive, with nothing overlapped: the one-at-a-time capacity."""
    reqs = sorted(requests, key=lambda q: q.arrival)
    t0 = time.perf_counter()
    for r in reqs:
        now = time.perf_counter() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        fp, P = gpu_prepare(r)
        r.footprint = fp
        r.admitted = time.perf_counter() - t0
        gpu_execute(r, P, lambda: False, launch, block_steps=1 << 30)
        r.end = time.perf_counter() - t0
    return summarize(reqs)
