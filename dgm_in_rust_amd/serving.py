"""Serving loop: flow cases arrive over time, are admitted into a :class:`~dgm_in_rust_amd.batch.Batch`
as capacity frees up, advance together on the GPU and leave when their lift coefficient has converged.

* Admission: a waiting request is added when its arena has a free range for it (its HBM footprint:
  state, RK stage buffers, mesh tables, fluxes, limiter buffers); requests are tried in arrival order
  and a smaller later one may pass a head that does not fit yet. Its mesh and state go to the device
  beside the queued steps; it takes its first step with the next chunk.
* Continuous batching: every chunk of ``chunk_steps`` steps advances all live cases in one graph
  launch per step. The host keeps ``depth`` chunks queued: while the GPU runs one, it reads the poll
  of the one before, retires finished cases and admits new ones, so the GPU does not wait for it.
* Convergence, on the device: the poll after each chunk computes every case's lift coefficient and stops
  a case (after the step it is at) when |Cl(n) - Cl(n - chunk_steps)| < cl_tol; a case also leaves at
  ``max_steps`` or when it stopped with an error status. A stopped case takes no further step, so the
  chunk already queued behind its last one leaves it as it was.
* Latency of a request = time its poll was read - arrival time (queueing included).

The reference has no serving layer; this module drives the dgm_batch_* entry points of libdgm.so.
"""
from __future__ import annotations

import collections
import dataclasses
import time
from typing import Callable, Dict, List, Optional

import numpy as np

from .batch import Batch, n_basis
from .meshgen import MeshInput


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float            # seconds after the start of the run
    mesh: MeshInput
    order: int
    mach: float
    aoa_deg: float


@dataclasses.dataclass
class Result:
    rid: int
    order: int
    n_elements: int
    arrival: float
    admitted: float
    finished: float
    steps: int
    converged: bool
    status: int
    cl: float
    state: Optional[np.ndarray] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def outcome(self) -> str:
        return "failed" if self.status != 0 else ("converged" if self.converged else "capped")


class Server:
    def __init__(self, batch: Batch, chunk_steps: int = 100, cl_tol: float = 1e-5, max_steps: int = 20000,
                 chord: float = 1.0, keep_states: bool = False, depth: int = 2):
        self.batch = batch
        self.chunk_steps = chunk_steps
        self.cl_tol = cl_tol
        self.max_steps = max_steps
        self.chord = chord
        self.keep_states = keep_states
        self.depth = max(1, min(2, depth))   # chunks kept queued (the batch holds two poll buffers)
        self.peak_live = 0
        self.dof_stages = 0.0     # sum over cases of n_elements * 4 Np * 3 * steps taken
        self.device_ms = 0.0      # device time of the step launches (HIP events), set at the end of run()

    def run(self, requests: List[Request], time_limit: Optional[float] = None,
            clock: Optional[Callable[[], float]] = None, idle: Optional[Callable[[float], None]] = None) -> List[Result]:
        """Serves the requests (sorted by arrival) until all are finished or time_limit passes.
        ``clock`` returns seconds since the start (default: wall clock); ``idle(dt)`` waits when nothing
        is live (default: time.sleep)."""
        b = self.batch
        if clock is None:
            t0 = time.perf_counter()
            clock = lambda: time.perf_counter() - t0  # noqa: E731
        if idle is None:
            idle = time.sleep
        dev0 = b.device_ms()
        todo = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting: List[Request] = []
        live: Dict[int, dict] = {}     # case id -> request, admission time, first poll, steps counted
        queued = collections.deque()  # tickets of the polls queued behind each chunk
        polls_queued = 0
        results: List[Result] = []

        def admit() -> int:
            nonlocal waiting
            still, n = [], 0
            for r in waiting:
                cid = b.add(r.mesh, r.order, r.mach, r.aoa_deg, final_step=self.max_steps)
                if cid is None:
                    still.append(r)
                    continue
                live[cid] = dict(req=r, admitted=clock(), first_poll=polls_queued,
                                 ne=int(sum(int(c) for c in r.mesh.el_count)), steps=0)
                n += 1
            waiting = still
            return n

        while todo or waiting or live or queued:
            now = clock()
            if time_limit is not None and now > time_limit:
                break
            while todo and todo[0].arrival <= now:
                waiting.append(todo.popleft())
            admit()
            self.peak_live = max(self.peak_live, len(live))
            # keep the GPU fed: queue chunks (each followed by its poll) up to the depth
            if live and len(queued) < self.depth:
                b.steps_async(self.chunk_steps)
                queued.append(b.poll_async(self.chord, self.cl_tol))
                polls_queued += 1
                continue
            if not queued:   # nothing live
                if waiting:
                    # the slots and ranges of removed cases come back once the device has passed the removal
                    b.synchronize()
                    if not admit():
                        raise RuntimeError(f"request {waiting[0].rid} does not fit in the empty batch")
                    continue
                if todo:
                    idle(max(0.0, todo[0].arrival - clock()))
                continue
            ticket = queued.popleft()
            info, cl, conv = b.poll_wait(ticket, max(live) + 1 if live else 1)
            now = clock()
            for cid in list(live):
                rec = live[cid]
                if rec["first_poll"] > ticket:   # admitted after this poll was queued
                    continue
                inf = info[cid]
                advanced = inf["steps"] - rec["steps"]
                rec["steps"] = inf["steps"]
                self.dof_stages += rec["ne"] * 4 * n_basis(rec["req"].order) * 3 * advanced
                if conv[cid] or not inf["active"]:
                    r = rec["req"]
                    state = b.state(cid) if self.keep_states else None
                    results.append(Result(r.rid, r.order, rec["ne"], r.arrival, rec["admitted"], now, inf["steps"],
                                          bool(conv[cid]) and inf["status"] == 0, int(inf["status"]), float(cl[cid]),
                                          state))
                    b.remove(cid)
                    del live[cid]
        b.synchronize()
        self.device_ms = b.device_ms() - dev0
        return results


def poisson_arrivals(rate: float, n: int, seed: int = 0) -> np.ndarray:
    """n arrival times of a Poisson process of the given rate (per second)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def _pct(x, q):
    return float(np.percentile(x, q)) if len(x) else None


def summary(results: List[Result], server: Server, wall: float) -> dict:
    """Latency percentiles over all requests and per outcome (converged, capped at max_steps, failed),
    throughput, aggregate GDOF * stage / s over the device time of the steps, and the GPU-busy fraction
    (device time of the step launches / wall time)."""
    lat = np.array([r.latency for r in results]) if results else np.zeros(0)
    by = {}
    for outcome in ("converged", "capped", "failed"):
        x = np.array([r.latency for r in results if r.outcome == outcome])
        st = [r.steps for r in results if r.outcome == outcome]
        by[outcome] = dict(count=int(len(x)), p50_latency_s=_pct(x, 50), p99_latency_s=_pct(x, 99),
                           mean_steps=float(np.mean(st)) if st else None)
    dev_s = server.device_ms / 1e3
    return dict(
        completed=len(results),
        converged=by["converged"]["count"],
        capped=by["capped"]["count"],
        failed=by["failed"]["count"],
        p50_latency_s=_pct(lat, 50),
        p99_latency_s=_pct(lat, 99),
        by_outcome=by,
        mean_steps=float(np.mean([r.steps for r in results])) if results else None,
        throughput_cases_per_s=len(results) / wall if wall > 0 else None,
        gdof_stage_per_s=server.dof_stages / dev_s / 1e9 if dev_s > 0 else None,
        gpu_busy_fraction=dev_s / wall if wall > 0 else None,
        peak_concurrent=server.peak_live,
        wall_s=wall,
    )


def main(argv=None):
    """python -m dgm_in_rust_amd.serving --rate 8 --requests 64 ...: Poisson arrivals of batched-sweep
    cases, served until converged; prints one JSON line with p50/p99 latency and throughput."""
    import argparse
    import json

    from . import cases, meshgen

    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=8.0)
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--tri-min", type=int, default=2000)
    ap.add_argument("--tri-max", type=int, default=60000)
    ap.add_argument("--max-steps", type=int, default=20000)
    ap.add_argument("--cl-tol", type=float, default=1e-5)
    ap.add_argument("--capacity", type=int, default=4_000_000, help="element capacity per order")
    ap.add_argument("--max-cases", type=int, default=1024)
    ap.add_argument("--time-limit", type=float, default=None)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    spec = cases.sweep_cases(a.requests, seed=a.seed, tri_min=a.tri_min, tri_max=a.tri_max)
    arr = poisson_arrivals(a.rate, a.requests, seed=a.seed + 1)
    reqs = [Request(i, float(t), meshgen.naca0012_ogrid(c["ns"], c["nl"], seed=c["mesh_seed"]), c["order"], c["mach"],
                    c["aoa_deg"]) for i, (c, t) in enumerate(zip(spec, arr))]
    with Batch({1: a.capacity, 2: a.capacity, 3: a.capacity}, precision=a.precision, max_cases=a.max_cases) as b:
        srv = Server(b, cl_tol=a.cl_tol, max_steps=a.max_steps)
        t0 = time.perf_counter()
        res = srv.run(reqs, time_limit=a.time_limit)
        wall = time.perf_counter() - t0
    out = summary(res, srv, wall)
    out.update(rate=a.rate, requests=a.requests, precision=a.precision, max_steps=a.max_steps, cl_tol=a.cl_tol,
               tri_range=[a.tri_min, a.tri_max])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
