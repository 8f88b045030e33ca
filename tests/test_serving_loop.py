"""The serving loop's host logic on the CPU, against a stand-in batch that keeps the device-side rules
of dgm_batch_*: a case admitted between chunks goes live with the next chunk; a poll snapshots every
record behind the queued steps and stops a running case whose lift changed by less than the tolerance
since the previous poll; polls complete in order, at most two outstanding; a removed case's slot and
capacity come back only once a later poll has completed (or after a synchronisation); a stopped case
takes no further step. Each request must leave with the outcome and step count these rules give it,
whenever it arrived, and the summary must split the latencies by outcome."""
import math
import types

import numpy as np
import pytest

from dgm_in_rust_amd.serving import Request, Server, summary

CHUNK, MAX_STEPS, TOL = 100, 2000, 1e-3


class FakeBatch:
    def __init__(self, capacity, max_cases, script):
        self.capacity, self.max_cases, self.script = capacity, max_cases, script
        self.cases = {}          # slot -> dict
        self.held = {}           # slot -> (poll index after which it is released, elements)
        self.pending = []        # admitted, not live yet
        self.polls, self.done = 0, 0
        self.snapshots = {}
        self.max_outstanding = 0
        self.chunks = 0
        self.ms = 0.0

    # the rules of the device side
    def _used(self):
        return sum(c["ne"] for c in self.cases.values()) + sum(ne for _, ne in self.held.values())

    def _flush(self):
        for s in self.pending:
            self.cases[s]["live"] = True
        self.pending = []

    def _release(self, upto):
        for s in [s for s, (after, _) in self.held.items() if after < upto]:
            del self.held[s]

    def add(self, mesh, order, mach, aoa, final_step=0):
        ne = int(mesh.el_count.sum())
        free = [s for s in range(self.max_cases) if s not in self.cases and s not in self.held]
        if not free or self._used() + ne > self.capacity:
            return None
        s = free[0]
        self.cases[s] = dict(rid=mesh.rid, ne=ne, live=False, steps=0, final=final_step, status=0,
                             clprev=0.0, conv=False)
        self.pending.append(s)
        return s

    def remove(self, s):
        self._flush()
        c = self.cases.pop(s)
        self.held[s] = (self.polls, c["ne"])

    def steps_async(self, n):
        self._flush()
        self.chunks += 1
        self.ms += 1.0
        for c in self.cases.values():
            if not c["live"]:
                continue
            fail_at = self.script[c["rid"]].get("fail_at")
            for _ in range(n):
                if c["status"] or c["steps"] >= c["final"]:
                    break
                c["steps"] += 1
                if fail_at is not None and c["steps"] == fail_at:
                    c["status"] = 2

    def poll_async(self, chord, tol):
        assert self.polls - self.done < 2, "two polls are outstanding"
        self._flush()
        snap = {}
        for s, c in self.cases.items():
            if not c["live"]:
                snap[s] = (dict(live=0, active=0, status=0, steps=0), float("nan"), False)
                continue
            cl = self.script[c["rid"]]["cl"](c["steps"])
            running = c["status"] == 0 and c["steps"] < c["final"] and c["steps"] > 0
            if running and not c["conv"] and abs(cl - c["clprev"]) < tol:
                c["conv"] = True
                c["final"] = c["steps"]
            c["clprev"] = cl
            active = int(c["status"] == 0 and c["steps"] < c["final"])
            snap[s] = (dict(live=1, active=active, status=c["status"], steps=c["steps"]), cl, c["conv"])
        self.snapshots[self.polls] = snap
        self.polls += 1
        self.max_outstanding = max(self.max_outstanding, self.polls - self.done)
        return self.polls - 1

    def poll_wait(self, ticket, n=None):
        assert ticket == self.done, "polls complete in order"
        snap = self.snapshots.pop(ticket)
        self.done += 1
        self._release(self.done)
        n = self.max_cases if n is None else n
        empty = (dict(live=0, active=0, status=0, steps=0), float("nan"), False)
        rows = [snap.get(i, empty) for i in range(n)]
        return [r[0] for r in rows], np.array([r[1] for r in rows]), np.array([r[2] for r in rows])

    def synchronize(self):
        self._flush()
        self.done = self.polls
        self.snapshots.clear()
        self.held.clear()

    def device_ms(self):
        return self.ms

    def state(self, s):
        return np.zeros(1)


def mesh(rid, ne):
    return types.SimpleNamespace(el_count=np.array([ne]), rid=rid)


def expected(script_entry):
    """Outcome and step count of a request under the rules above (polls every CHUNK steps)."""
    fail_at = script_entry.get("fail_at")
    prev = 0.0
    steps = 0
    while True:
        steps = min(steps + CHUNK, MAX_STEPS)
        if fail_at is not None and fail_at <= steps:
            return "failed", fail_at
        cl = script_entry["cl"](steps)
        if steps < MAX_STEPS and abs(cl - prev) < TOL:
            return "converged", steps
        if steps >= MAX_STEPS:
            return "capped", steps
        prev = cl


class Clock:
    def __init__(self, tick=0.01):
        self.t, self.tick = 0.0, tick

    def __call__(self):
        self.t += self.tick
        return self.t

    def idle(self, dt):
        self.t += dt


def make_script(n, seed):
    rng = np.random.default_rng(seed)
    script = {}
    for rid in range(n):
        tau = float(rng.uniform(100.0, 900.0))
        amp = float(rng.uniform(0.5, 5.0))
        entry = {"cl": (lambda tau, amp: (lambda s: 0.3 + amp * math.exp(-s / tau)))(tau, amp)}
        if rid % 7 == 3:
            entry["fail_at"] = int(rng.integers(50, 1500))
        script[rid] = entry
    return script


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_requests_leave_with_the_outcomes_of_the_device_rules(seed):
    n = 40
    script = make_script(n, seed)
    rng = np.random.default_rng(100 + seed)
    arrivals = np.cumsum(rng.exponential(0.05, n))
    sizes = rng.integers(100, 900, n)
    reqs = [Request(i, float(arrivals[i]), mesh(i, int(sizes[i])), 3, 0.5, 1.0) for i in range(n)]
    b = FakeBatch(capacity=2500, max_cases=6, script=script)
    clk = Clock()
    srv = Server(b, chunk_steps=CHUNK, cl_tol=TOL, max_steps=MAX_STEPS)
    res = srv.run(reqs, clock=clk, idle=clk.idle)
    assert sorted(r.rid for r in res) == list(range(n))
    assert b.max_outstanding == 2          # two chunks in flight
    assert srv.peak_live >= 3
    for r in res:
        outcome, steps = expected(script[r.rid])
        assert (r.outcome, r.steps) == (outcome, steps), r
        assert r.arrival <= r.admitted <= r.finished
    # cases waited for capacity while others ran
    assert any(r.admitted - r.arrival > 0.1 for r in res)
    s = summary(res, srv, clk.t)
    counts = {o: sum(1 for r in res if r.outcome == o) for o in ("converged", "capped", "failed")}
    assert {o: s["by_outcome"][o]["count"] for o in counts} == counts
    assert s["converged"] + s["capped"] + s["failed"] == n
    for o, c in counts.items():
        lat = sorted(r.latency for r in res if r.outcome == o)
        if c:
            assert s["by_outcome"][o]["p50_latency_s"] == pytest.approx(float(np.percentile(lat, 50)))
        else:
            assert s["by_outcome"][o]["p50_latency_s"] is None
    assert s["gpu_busy_fraction"] == pytest.approx(srv.device_ms / 1e3 / clk.t)


def test_a_request_larger_than_the_empty_batch_is_refused():
    script = {0: {"cl": lambda s: 1.0}}
    b = FakeBatch(capacity=100, max_cases=2, script=script)
    clk = Clock()
    with pytest.raises(RuntimeError, match="does not fit"):
        Server(b, chunk_steps=CHUNK, cl_tol=TOL, max_steps=MAX_STEPS).run(
            [Request(0, 0.0, mesh(0, 500), 3, 0.5, 1.0)], clock=clk, idle=clk.idle)


def test_removed_capacity_returns_only_after_a_completed_poll():
    """Two cases that each take the whole batch: the second is admitted only after the first has left
    and a poll queued after its removal has completed, and it still runs its full course."""
    script = {0: {"cl": lambda s: 0.3}, 1: {"cl": lambda s: 0.3 + 1.0 / (1.0 + s)}}
    b = FakeBatch(capacity=1000, max_cases=4, script=script)
    clk = Clock()
    srv = Server(b, chunk_steps=CHUNK, cl_tol=TOL, max_steps=MAX_STEPS)
    res = srv.run([Request(0, 0.0, mesh(0, 1000), 3, 0.5, 1.0), Request(1, 0.0, mesh(1, 1000), 3, 0.5, 1.0)],
                  clock=clk, idle=clk.idle)
    by = {r.rid: r for r in res}
    assert by[0].outcome == "converged" and by[0].steps == 200   # |dCl| = 0 at the second poll
    assert by[1].admitted > by[0].finished
    assert (by[1].outcome, by[1].steps) == expected(script[1])
