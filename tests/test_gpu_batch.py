"""GPU: batches of cases (dgm_batch_*). Each case in a batch must give the same state, bit for bit,
as the same case alone in a single-case context: the per-element arithmetic does not depend on where
a case sits in its arena or on which cases share a launch, and the time-step minimum is order
independent. Also: per-case stop conditions, admission by free capacity (NO_SPACE), removal and
re-use of freed ranges between steps, per-case errors, lift, and the autotuner leaving states alone."""
import numpy as np
import pytest

from dgm_in_rust_amd import Batch, Context, SpatialDisc, TemperalDisc, meshgen

pytestmark = pytest.mark.gpu

CASES = [  # (surface points, layers, seed, order, mach, aoa)
    (40, 12, 1, 1, 0.5, 2.0),
    (64, 16, 2, 2, 0.8, 1.25),
    (48, 10, 3, 3, 0.7, -2.0),
    (32, 8, 4, 2, 0.3, 5.0),
    (24, 6, 5, 4, 0.4, 0.5),
]


def alone(mesh, order, mach, aoa, steps, precision):
    ctx = Context(mesh, order=order, mach=mach, aoa_deg=aoa, precision=precision, limiter="jacobi")
    U = ctx.initial_solution(1.0, aoa, mach, 1.0e5)
    st = TemperalDisc(ctx, final_step=steps, final_time=1.0e300).time_march(SpatialDisc(ctx), ctx.zeros(), U)
    return U, st, ctx


def make_batch(precision, cap=40000, max_cases=16):
    return Batch({1: cap, 2: cap, 3: cap, 4: cap}, precision=precision, limiter="jacobi", max_cases=max_cases)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_batch_matches_alone_bitwise(precision):
    steps = 25
    b = make_batch(precision)
    ids = []
    for ns, nl, seed, p, M, a in CASES:
        ids.append(b.add(meshgen.naca0012_ogrid(ns, nl, seed=seed), p, M, a, final_step=steps))
    assert None not in ids
    b.steps_async(steps + 3)   # extra steps are no-ops for finished cases
    b.synchronize()
    info = b.info(len(CASES))
    for cid, (ns, nl, seed, p, M, a) in zip(ids, CASES):
        U, st, _ = alone(meshgen.naca0012_ogrid(ns, nl, seed=seed), p, M, a, steps, precision)
        assert info[cid]["steps"] == st["steps"] == steps and info[cid]["status"] == 0
        assert info[cid]["time"] == st["time"]
        assert np.array_equal(b.state(cid), U), (cid, np.abs(b.state(cid) - U).max())
    b.close()


def test_continuous_batching_add_remove_midrun():
    """A case admitted while others run, and one placed into a range freed by a removal, still match
    their stand-alone runs exactly."""
    b = make_batch("fp32", cap=6000, max_cases=8)
    mA = meshgen.naca0012_ogrid(48, 20, seed=11)   # 1920 triangles
    mB = meshgen.naca0012_ogrid(40, 16, seed=12)
    mC = meshgen.naca0012_ogrid(40, 20, seed=13)
    a = b.add(mA, 2, 0.6, 1.0, final_step=40)
    b.steps_async(10)
    bb = b.add(mB, 2, 0.7, 3.0, final_step=30)
    b.steps_async(10)
    b.synchronize()
    assert b.info(8)[a]["steps"] == 20 and b.info(8)[bb]["steps"] == 10
    UA_mid = b.state(a)
    b.remove(a)
    b.synchronize()   # the removed case's ranges come back once the device has passed the removal
    c = b.add(mC, 2, 0.5, -1.0, final_step=15)
    assert c is not None
    b.steps_async(25)
    b.synchronize()
    ref_A, _, _ = alone(mA, 2, 0.6, 1.0, 20, "fp32")
    ref_B, _, _ = alone(mB, 2, 0.7, 3.0, 30, "fp32")
    ref_C, _, _ = alone(mC, 2, 0.5, -1.0, 15, "fp32")
    assert np.array_equal(UA_mid, ref_A)
    assert np.array_equal(b.state(bb), ref_B)
    assert np.array_equal(b.state(c), ref_C)
    info = b.info(8)
    assert info[bb]["steps"] == 30 and info[c]["steps"] == 15 and info[bb]["active"] == 0
    b.close()


def test_admission_by_capacity():
    b = Batch({2: 4096}, precision="fp32", limiter="jacobi", max_cases=4)
    m = meshgen.naca0012_ogrid(40, 20, seed=21)   # 1600 triangles -> 25 blocks of 64
    ids = [b.add(m, 2, 0.5, 0.0) for _ in range(3)]
    assert ids[0] is not None and ids[1] is not None and ids[2] is None   # 3 x 1600 > 4096
    hc = Context(m, order=2, host_only=True)
    npatch = len(m.eb_type)
    fp = b.footprint(2, hc.info.n_elements, hc.info.n_edges, npatch)
    assert hc.info.n_elements == 1600
    assert b.info(4)[ids[0]]["footprint_bytes"] == fp > 1600 * 100
    b.remove(ids[0])
    # the removed case's slot and ranges stay out of use until the device has passed the removal
    # (the next completed poll, or a synchronisation)
    assert b.add(m, 2, 0.5, 0.0) is None
    b.synchronize()
    assert b.add(m, 2, 0.5, 0.0) is not None
    with pytest.raises(Exception):
        b.add(m, 3, 0.5, 0.0)   # no arena for p = 3
    b.close()


def test_capacity_beyond_offset_range_is_rejected():
    """The kernels address a variable block of modes and the point-flux array with 32-bit byte offsets:
    an arena whose flux array would reach 4 GiB (p = 3, FP32: 40 M elements -> 64 M edges x 80 bytes)
    is refused with BAD_CONFIG before anything is allocated; a large arena within range is accepted."""
    from dgm_in_rust_amd._lib import DgmError
    with pytest.raises(DgmError) as ex:
        Batch({3: 40_000_000}, precision="fp32", limiter="jacobi", max_cases=4)
    assert ex.value.status == 4
    b = Batch({3: 1_000_000}, precision="fp32", limiter="jacobi", max_cases=4)
    b.close()


@pytest.mark.parametrize("good_order,bad_order", [(2, 2), (1, 2), (2, 1)])
def test_per_case_error_is_isolated(good_order, bad_order):
    """A case whose state fails the positivity check stops with its own status; the others go on
    and match their stand-alone runs. The failing case sits in the same arena as the good one, in a
    later one, or in the first one (whose stage-0 edge launch also runs every case's step control)."""
    b = make_batch("fp64", cap=8000, max_cases=4)
    good = meshgen.naca0012_ogrid(40, 12, seed=31)
    bad = meshgen.naca0012_ogrid(32, 8, seed=32)
    g = b.add(good, good_order, 0.6, 1.0, final_step=12)
    ctx = Context(bad, order=bad_order, mach=0.5, aoa_deg=0.0, limiter="jacobi")
    Ub = ctx.initial_solution(1.0, 0.0, 0.5, 1.0e5)
    Ub[5, 3, 0] = -1.0e5   # negative energy -> negative pressure at element 5
    x = b.add(bad, bad_order, 0.5, 0.0, U0=Ub, final_step=12)
    b.steps_async(12)
    b.synchronize()
    info = b.info(4)
    assert info[x]["status"] == 2 and info[x]["active"] == 0
    # the failure is found by the first step's edge pass, which runs beside that step's control: the
    # step had begun, so it is counted, and no later step runs
    assert info[x]["steps"] == 1
    assert info[g]["status"] == 0 and info[g]["steps"] == 12
    ref, _, _ = alone(good, good_order, 0.6, 1.0, 12, "fp64")
    assert np.array_equal(b.state(g), ref)
    b.close()


def test_batch_lift_matches_context():
    b = make_batch("fp64", cap=8000, max_cases=4)
    cases = [(meshgen.naca0012_ogrid(40, 12, seed=41), 2, 0.6, 2.0), (meshgen.naca0012_ogrid(48, 10, seed=42), 3, 0.5, -3.0)]
    ids = [b.add(m, p, M, a, final_step=8) for m, p, M, a in cases]
    b.steps_async(8)
    b.synchronize()
    cl = b.lift(1.0, 4)
    assert np.isnan(cl[2]) and np.isnan(cl[3])
    for cid, (m, p, M, a) in zip(ids, cases):
        _, _, ctx = alone(m, p, M, a, 8, "fp64")
        assert abs(cl[cid] - ctx.lift_coefficient()) <= 1e-12 * max(1.0, abs(cl[cid]))
    b.close()


def test_autotune_leaves_states_unchanged(tmp_path):
    steps = 12
    b = make_batch("fp32")
    ids = [b.add(meshgen.naca0012_ogrid(ns, nl, seed=seed), p, M, a, final_step=steps)
           for ns, nl, seed, p, M, a in CASES]
    b.steps_async(5)
    chosen = b.autotune(repeats=3, cache=str(tmp_path / "tune.json"))
    assert set(chosen) == {1, 2, 3, 4}
    for p, t in chosen.items():
        assert t["surface_block"] in (64, 128, 256)
        # the FP32-bulk element kernel variant is fixed by the order: one lane per element at p <= 2 (and
        # at p = 4, whose 25 cell points exceed the matrix-core tile, whatever element_waves says)
        assert t["element_waves"] == 0 if p <= 2 else t["element_waves"] in (1, 2, 4)
        assert b.tuning(p) == t
    again = b.autotune(cache=str(tmp_path / "tune.json"))   # served from the cache file
    assert again == chosen
    b.steps_async(steps)
    b.synchronize()
    for cid, (ns, nl, seed, p, M, a) in zip(ids, CASES):
        U, _, _ = alone(meshgen.naca0012_ogrid(ns, nl, seed=seed), p, M, a, steps, "fp32")
        assert np.array_equal(b.state(cid), U)
    b.close()


def test_context_autotune_leaves_results_unchanged():
    mesh = meshgen.naca0012_ogrid(64, 16, seed=61)
    ref, _, _ = alone(mesh, 3, 0.8, 1.25, 10, "fp32")
    ctx = Context(mesh, order=3, mach=0.8, aoa_deg=1.25, precision="fp32", limiter="jacobi")
    U = ctx.initial_solution(1.0, 1.25, 0.8, 1.0e5)
    ctx.upload(U)
    ctx.steps_async(4)
    ctx.synchronize()
    t = ctx.autotune(repeats=2)
    assert ctx.tuning() == t and t["detect_block"] in (64, 128, 256)
    ctx.steps_async(6)
    ctx.synchronize()
    assert np.array_equal(ctx.download(), ref)


def test_batch_profile_times_every_order_and_leaves_states_unchanged():
    """Per-kernel timing of a batch (eager steps, events per launch) reports every served order's edge,
    element, detector and rebuild passes, and leaves every case's state and clock as they were."""
    steps = 12
    b = make_batch("fp32")
    ids = [b.add(meshgen.naca0012_ogrid(ns, nl, seed=seed), p, M, a, final_step=steps)
           for ns, nl, seed, p, M, a in CASES]
    b.steps_async(5)
    before = [i["steps"] for i in b.info(16)]
    prof = b.profile_steps(3)
    assert set(prof) == {1, 2, 3, 4}
    for p, t in prof.items():
        assert t["surface_ms"] > 0 and t["element_ms"] > 0 and t["detect_ms"] > 0 and t["rebuild_ms"] > 0, (p, t)
        assert t["step_ms"] > t["element_ms"]
    assert [i["steps"] for i in b.info(16)] == before
    b.steps_async(steps)
    b.synchronize()
    for cid, (ns, nl, seed, p, M, a) in zip(ids, CASES):
        U, _, _ = alone(meshgen.naca0012_ogrid(ns, nl, seed=seed), p, M, a, steps, "fp32")
        assert np.array_equal(b.state(cid), U)
    b.close()


def test_sweep_full_size_cases_match_alone():
    """The benchmark's batched sweep at its full size: 256 cases of the sweep distribution (4.39 M elements,
    p = 1-3, 2k-60k triangles each, FP32-bulk, Jacobi WENO) in one batch for 100 steps. The smallest and
    the largest case of each order, and case 69 (which the bench's 1000-step sweep stops on a positivity
    loss), equal the same cases run alone, bit for bit, with the same step count and status; every case
    has run its 100 steps or stopped with an error status."""
    from dgm_in_rust_amd import cases
    steps = 100
    spec = cases.sweep_cases(256, seed=0)
    meshes = [meshgen.naca0012_ogrid(c["ns"], c["nl"], seed=c["mesh_seed"]) for c in spec]
    cap = {p: 0 for p in (1, 2, 3)}
    for c, m in zip(spec, meshes):
        cap[c["order"]] += (int(m.el_count.sum()) + 63) // 64 * 64
    assert sum(cap.values()) >= 4391624
    b = Batch(cap, precision="fp32", limiter="jacobi", max_cases=256)
    ids = [b.add(m, c["order"], c["mach"], c["aoa_deg"], final_step=steps) for c, m in zip(spec, meshes)]
    assert None not in ids
    b.steps_async(steps)
    b.synchronize()
    info = b.info(256)
    for i in ids:
        assert (info[i]["status"] == 0 and info[i]["steps"] == steps) or info[i]["status"] != 0, info[i]
    sample = {69}
    for p in (1, 2, 3):
        sizes = [(int(m.el_count.sum()), k) for k, (c, m) in enumerate(zip(spec, meshes)) if c["order"] == p]
        sample |= {min(sizes)[1], max(sizes)[1]}
    for k in sorted(sample):
        c = spec[k]
        ctx = Context(meshes[k], order=c["order"], mach=c["mach"], aoa_deg=c["aoa_deg"], precision="fp32",
                      limiter="jacobi")
        U = ctx.initial_solution(1.0, c["aoa_deg"], c["mach"], 1.0e5)
        try:
            st = TemperalDisc(ctx, final_step=steps, final_time=1.0e300).time_march(SpatialDisc(ctx), ctx.zeros(), U)
            status, nsteps = 0, st["steps"]
        except Exception as err:  # a positivity loss: the march reports its status
            status, nsteps = getattr(err, "status", -1), None
        print(f"sweep case {k}: p={c['order']}, {int(meshes[k].el_count.sum())} triangles, batch status "
              f"{info[ids[k]]['status']} steps {info[ids[k]]['steps']}, alone status {status}", flush=True)
        assert info[ids[k]]["status"] == status
        if status == 0:
            assert info[ids[k]]["steps"] == nsteps == steps
            assert np.array_equal(b.state(ids[k]), U), (k, np.abs(b.state(ids[k]) - U).max())
        ctx.close()
    b.close()
