"""GPU: the serving loop. Requests arrive while others run, wait for capacity, are admitted between
chunks, and leave on lift convergence or their step cap; each served state must equal the same case
run alone for the same number of steps bit for bit, and its lift coefficient to 1e-12 (the wall-force
sum is accumulated with floating-point atomics, so its last bits depend on their order)."""
import numpy as np
import pytest

from dgm_in_rust_amd import Batch, Context, SpatialDisc, TemperalDisc, meshgen
from dgm_in_rust_amd.serving import Request, Server, summary

pytestmark = pytest.mark.gpu


class SimClock:
    """Deterministic clock: every reading advances it by `tick` seconds; idle(dt) jumps ahead."""

    def __init__(self, tick=0.01):
        self.t, self.tick = 0.0, tick

    def __call__(self):
        self.t += self.tick
        return self.t

    def idle(self, dt):
        self.t += dt


def alone(mesh, order, mach, aoa, steps):
    ctx = Context(mesh, order=order, mach=mach, aoa_deg=aoa, precision="fp32", limiter="jacobi")
    U = ctx.initial_solution(1.0, aoa, mach, 1.0e5)
    TemperalDisc(ctx, final_step=steps, final_time=1.0e300).time_march(SpatialDisc(ctx), ctx.zeros(), U)
    return U, ctx.lift_coefficient()


def test_served_results_equal_alone():
    spec = [  # arrival, (ns, nl, seed), order, mach, aoa
        (0.00, (40, 20, 51), 2, 0.5, 2.0),
        (0.00, (40, 20, 52), 2, 0.7, -1.0),
        (0.02, (40, 20, 53), 2, 0.6, 4.0),
        (0.05, (32, 12, 54), 1, 0.4, 1.0),
        (0.30, (24, 10, 55), 3, 0.8, 0.5),
        (0.31, (40, 20, 56), 2, 0.3, 3.0),
    ]
    reqs = [Request(i, t, meshgen.naca0012_ogrid(*m[:2], seed=m[2]), p, M, a) for i, (t, m, p, M, a) in enumerate(spec)]
    clk = SimClock()
    with Batch({1: 4096, 2: 4096, 3: 4096}, precision="fp32", limiter="jacobi", max_cases=8) as b:
        srv = Server(b, chunk_steps=20, cl_tol=2e-3, max_steps=100, keep_states=True)
        res = srv.run(reqs, clock=clk, idle=clk.idle)
        s = summary(res, srv, clk.t)
    assert len(res) == len(reqs) and all(r.status == 0 for r in res)
    assert srv.peak_live >= 3                                   # work of several requests overlapped
    assert any(r.admitted - r.arrival > 0.05 for r in res)      # someone waited for capacity
    assert s["p50_latency_s"] is not None and s["p99_latency_s"] >= s["p50_latency_s"]
    for r in res:
        assert r.steps % 20 == 0 and 20 <= r.steps <= 100
        t, m, p, M, a = spec[r.rid]
        U, cl = alone(meshgen.naca0012_ogrid(*m[:2], seed=m[2]), p, M, a, r.steps)
        assert np.array_equal(r.state, U), r.rid
        assert abs(r.cl - cl) <= 1e-12 * max(1.0, abs(cl)), (r.rid, r.cl, cl)


def test_poll_stops_converged_cases():
    """A poll with a tolerance no lift change meets stops nothing; one that every change meets stops each
    running case after the step it is at, so later steps leave it as it was. The poll's lift equals the
    batch lift read-out, and a case admitted after the steps were queued is not tested."""
    m1 = meshgen.naca0012_ogrid(40, 20, seed=61)
    m2 = meshgen.naca0012_ogrid(32, 12, seed=62)
    with Batch({2: 4096}, precision="fp32", limiter="jacobi", max_cases=4) as b:
        a = b.add(m1, 2, 0.6, 2.0, final_step=1000)
        b.steps_async(20)
        t0 = b.poll_async(1.0, 0.0)
        late = b.add(m2, 2, 0.5, 1.0, final_step=1000)   # goes live with the next steps
        t1 = b.poll_async(1.0, 1.0e9)                     # queued before any step of `late`
        info0, cl0, conv0 = b.poll_wait(t0)
        info1, cl1, conv1 = b.poll_wait(t1)
        assert info0[a]["steps"] == 20 and not conv0[a] and info0[late]["live"] == 0
        assert conv1[a] and not conv1[late] and info1[late]["steps"] == 0
        b.steps_async(20)
        b.synchronize()
        info = b.info(4)
        assert info[a]["steps"] == 20 and info[a]["active"] == 0 and info[a]["status"] == 0
        assert info[late]["steps"] == 20 and info[late]["active"] == 1
        lift = b.lift(1.0, 4)
        Ua = b.state(a)
    U, cl = alone(m1, 2, 0.6, 2.0, 20)
    assert np.array_equal(Ua, U)
    assert abs(cl0[a] - cl) <= 1e-12 * max(1.0, abs(cl)) and abs(lift[a] - cl) <= 1e-12 * max(1.0, abs(cl))
    assert abs(cl1[a] - cl0[a]) <= 1e-12 * max(1.0, abs(cl))   # same state; atomics may reorder the force sum
