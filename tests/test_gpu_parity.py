"""GPU parity: the HIP path through the C ABI against the FP64 oracle on the same seeded inputs.

Bar: FP64 mode within 1e-12 (max-norm relative per conserved variable); FP32-bulk mode within
the tolerances stated per test. Troubled-cell sets must match exactly.
"""
import numpy as np
import pytest

from dgm_in_rust_amd import Context, SpatialDisc, TemperalDisc, meshgen
from dgm_in_rust_amd._lib import DgmError
from oracle.refcpu import RefCase
from tests.helpers import flow_state, rel_err, rel_l2_per_var

pytestmark = pytest.mark.gpu

LIM = {"off": 0, "sequential": 1, "jacobi": 2}
TOL64 = 1e-12
# FP32-bulk against the FP64 oracle, per conserved variable (momentum components against the momentum
# norm): the residual of one state (measured up to 2.0e-5, on the O-grid's transverse momentum), and
# a 10-step march with WENO active (measured up to 6.1e-7)
FP32_RESID_TOL = 4e-5
FP32_WENO_TOL = 1e-6


def pair(mesh, order, limiter="sequential", precision="fp64", **kw):
    ref = RefCase(mesh, order=order, limiter=LIM[limiter], **kw)
    ctx = Context(mesh, order=order, limiter=limiter, precision=precision, **kw)
    return ref, ctx


def meshes(airfoil):
    return {
        "airfoil_p2": (airfoil, 2, dict(mach=0.8, aoa_deg=1.5)),
        "ogrid_p3": (meshgen.naca0012_ogrid(40, 10, seed=3), 3, dict(mach=0.8, aoa_deg=1.25)),
        "ogrid_p1": (meshgen.naca0012_ogrid(32, 8, seed=4), 1, dict(mach=0.5, aoa_deg=1.25)),
        "periodic_p4": (meshgen.periodic_square(6, seed=5), 4, dict(mach=0.5, aoa_deg=0.0)),
        "box_p2": (meshgen.box_farfield(5, seed=6), 2, dict(mach=0.5, aoa_deg=1.25)),
    }


@pytest.fixture(scope="module")
def all_meshes(airfoil_mesh):
    return meshes(airfoil_mesh)


@pytest.mark.parametrize("name", ["airfoil_p2", "ogrid_p3", "ogrid_p1", "periodic_p4", "box_p2"])
def test_residual_fp64(all_meshes, name):
    mesh, order, kw = all_meshes[name]
    ref, ctx = pair(mesh, order, **kw)
    U = flow_state(ref, seed=11, shock=True, mach=kw["mach"])
    Rref, s, _ = ref.compute_residuals(U)
    assert s == 0
    R = SpatialDisc(ctx).residuals(U)
    assert rel_err(R, Rref) <= TOL64, rel_err(R, Rref)


def test_residual_accumulate_semantics(all_meshes):
    mesh, order, kw = all_meshes["airfoil_p2"]
    ref, ctx = pair(mesh, order, **kw)
    U = flow_state(ref, seed=12, mach=0.8)
    R0 = np.random.default_rng(1).standard_normal(ref.shape) * 1e3
    Rref, s, _ = ref.compute_residuals(U, R0)
    R = R0.copy()
    SpatialDisc(ctx).compute_residuals(R, U)
    assert rel_err(R, Rref) <= TOL64


@pytest.mark.parametrize("name", ["airfoil_p2", "ogrid_p3", "periodic_p4"])
def test_time_step(all_meshes, name):
    mesh, order, kw = all_meshes[name]
    ref, ctx = pair(mesh, order, **kw)
    U = flow_state(ref, seed=13, shock=True, mach=kw["mach"])
    dref = ref.compute_time_step(U)
    d = TemperalDisc(ctx).compute_time_step(U)
    assert abs(d - dref) <= 1e-14 * abs(dref)


@pytest.mark.parametrize("limiter", ["sequential", "jacobi"])
@pytest.mark.parametrize("name", ["airfoil_p2", "ogrid_p3", "ogrid_p1", "periodic_p4"])
def test_limiter_fp64(all_meshes, name, limiter):
    mesh, order, kw = all_meshes[name]
    ref, ctx = pair(mesh, order, limiter=limiter, **kw)
    U = flow_state(ref, seed=14, shock=True, mach=kw["mach"], noise=0.02)
    Uref, tref, s = ref.apply_limiter(U)
    assert s == 0
    assert tref.sum() > 0, "state must trouble some cells"
    V = U.copy()
    t = np.zeros(ctx.ne, np.uint8)
    SpatialDisc(ctx).apply_limiter(V, t)
    assert np.array_equal(t.astype(bool), tref)
    assert rel_err(V, Uref) <= TOL64, rel_err(V, Uref)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", ["airfoil_p2", "ogrid_p3", "ogrid_p1", "periodic_p4"])
def test_sequential_walk_matches_level_launches(all_meshes, name, precision, monkeypatch):
    """The element-order sweep in one wavefront (after a detector and rebuild pass over every cell)
    against one launch per level of the schedule: the same troubled cells and the same state, on a
    state where cells with a rebuilt lower-numbered neighbour are troubled (so the walk re-detects and
    re-rebuilds them) and where the element order changes the result (it differs from Jacobi's)."""
    mesh, order, kw = all_meshes[name]
    ref = RefCase(mesh, order=order, limiter=1, **kw)
    U = flow_state(ref, seed=14, shock=True, mach=kw["mach"], noise=0.05)
    out = {}
    for mode, lim, walk in (("walk", "sequential", "1"), ("levels", "sequential", "0"), ("jacobi", "jacobi", "1")):
        monkeypatch.setenv("DGM_SEQUENTIAL_WALK", walk)
        ctx = Context(mesh, order=order, limiter=lim, precision=precision, **kw)
        V = U.copy()
        t = np.zeros(ctx.ne, np.uint8)
        SpatialDisc(ctx).apply_limiter(V, t)
        out[mode] = (V, t.astype(bool))
        nbr = ctx.connectivity()["ineighbours"]
    (Vw, tw), (Vl, tl), (Vj, _) = out["walk"], out["levels"], out["jacobi"]
    assert np.array_equal(tw, tl)
    # the walk and the per-level launches run the same detector and rebuild code on the same inputs
    assert np.array_equal(Vw, Vl), rel_err(Vw, Vl)
    chained = [e for e in np.nonzero(tw)[0] if any(0 <= j < e and tw[j] for j in nbr[e])]
    assert len(chained) > 0, "some troubled cell must follow a troubled lower-numbered neighbour"
    assert rel_err(Vw, Vj) > 1e-9, "the element order must matter on this state"
    if precision == "fp64":
        Uref, tref, s = ref.apply_limiter(U)
        assert s == 0 and np.array_equal(tw, tref)
        assert rel_err(Vw, Uref) <= TOL64, rel_err(Vw, Uref)


@pytest.mark.parametrize("name", ["airfoil_p2", "ogrid_p3", "ogrid_p1"])
def test_limiter_fp32_bulk_detection(all_meshes, name):
    """FP32-bulk detection (traces, jumps and indicators in FP32, on the FP32 shadow of the state)
    against the FP64 decision on that same rounded state: equal for every cell whose decision FP32
    rounding cannot move (oracle/numpy_ref.py kxrcf_margin), and those cells are nearly all of them."""
    from oracle.numpy_ref import NumpyRef
    mesh, order, kw = all_meshes[name]
    ref, ctx = pair(mesh, order, limiter="jacobi", precision="fp32", **kw)
    nref = NumpyRef(mesh, ref.tables(), order=order, fs=(kw["mach"], kw["aoa_deg"], 1.0, 1e5, 1.4),
                    limiter="jacobi")
    U = flow_state(ref, seed=14, shock=True, mach=kw["mach"], noise=0.02)
    U32 = U.astype(np.float32).astype(np.float64)
    t = np.zeros(ctx.ne, np.uint8)
    SpatialDisc(ctx).apply_limiter(U.copy(), t)
    dec = [nref.kxrcf_margin(U32, e) for e in range(ctx.ne)]
    want = np.array([d for d, _ in dec])
    sure = np.array([c for _, c in dec])
    got = t.astype(bool)
    assert want[sure].sum() > 0, "state must trouble some cells"
    assert np.array_equal(got[sure], want[sure]), np.nonzero(got[sure] != want[sure])[0][:20]
    assert sure.mean() >= 0.9, sure.mean()


@pytest.mark.parametrize("name", ["airfoil_p2", "ogrid_p3"])
def test_rk3_step_fp64(all_meshes, name):
    mesh, order, kw = all_meshes[name]
    ref, ctx = pair(mesh, order, **kw)
    U = flow_state(ref, seed=15, shock=True, mach=kw["mach"], amp=0.02)
    dt = ref.compute_time_step(U)
    Uref, s = ref.rk3(U, dt)
    assert s == 0
    V = U.copy()
    TemperalDisc(ctx).rk3(SpatialDisc(ctx), ctx.zeros(), V, dt)
    assert rel_err(V, Uref) <= TOL64, rel_err(V, Uref)


def test_time_march_reference_config(all_meshes, tmp_path):
    """airfoil.msh, p = 2, (6, 3) rules, HLLC, sequential WENO, reference CFL formula, 10 steps; the
    march writes its result file as the reference's does (step=10.vtk, cell means of the final state)."""
    from dgm_in_rust_amd import vtk
    mesh, order, kw = all_meshes["airfoil_p2"]
    ref, ctx = pair(mesh, order, **kw)
    U0 = ref.initial_solution(1.0, 1.5, 0.8, 1.0e5)
    Uref, s, n, t = ref.time_march(U0, 10, 5.0)
    assert s == 0 and n == 10
    V = U0.copy()
    td = TemperalDisc(ctx, final_step=10, final_time=5.0, output_dir=str(tmp_path))
    st = td.time_march(SpatialDisc(ctx), ctx.zeros(), V)
    assert st["steps"] == 10
    assert abs(st["time"] - t) <= 1e-14 * t
    assert rel_err(V, Uref) <= TOL64, rel_err(V, Uref)
    text = open(tmp_path / "step=10.vtk").read()
    assert text == vtk.vtk_text(mesh.vxy, ctx.connectivity()["element_vertices"], V, 1.4)


def test_time_march_ogrid_p3_jacobi(all_meshes):
    mesh, order, kw = all_meshes["ogrid_p3"]
    ref, ctx = pair(mesh, order, limiter="jacobi", cfl=0.3, **kw)
    U0 = ref.initial_solution(1.0, kw["aoa_deg"], kw["mach"], 1.0e5)
    Uref, s, n, t = ref.time_march(U0, 8, 100.0)
    assert s == 0
    V = U0.copy()
    st = TemperalDisc(ctx, final_step=8, final_time=100.0).time_march(SpatialDisc(ctx), ctx.zeros(), V)
    assert st["steps"] == n
    assert rel_err(V, Uref) <= TOL64, rel_err(V, Uref)


def test_final_time_clamp(all_meshes):
    mesh, order, kw = all_meshes["box_p2"]
    ref, ctx = pair(mesh, order, **kw)
    U0 = flow_state(ref, seed=16, mach=0.5)
    dt = ref.compute_time_step(U0)
    T = 2.5 * dt
    Uref, s, n, t = ref.time_march(U0, 100, T)
    V = U0.copy()
    st = TemperalDisc(ctx, final_step=100, final_time=T).time_march(SpatialDisc(ctx), ctx.zeros(), V)
    assert st["steps"] == n and abs(st["time"] - t) <= 1e-15 * T
    assert rel_err(V, Uref) <= TOL64


@pytest.mark.parametrize("variant", ["default", "other"])
@pytest.mark.parametrize("name", ["airfoil_p2", "ogrid_p3", "ogrid_p1", "periodic_p4", "box_p2"])
def test_residual_fp32_bulk(all_meshes, name, variant):
    """Both FP32-bulk element kernels (matrix-core: element_waves > 0; one lane per element: 0) against
    the oracle, at the order where each is the default and where it is not (p = 4 has one lane per
    element only: its 25 cell points exceed the matrix-core tile), on every mesh family."""
    mesh, order, kw = all_meshes[name]
    ref, ctx = pair(mesh, order, precision="fp32", **kw)
    if variant == "other":
        t = ctx.tuning()
        t["element_waves"] = 4 if t["element_waves"] == 0 else 0
        ctx.set_tuning(t)
    U = flow_state(ref, seed=17, shock=True, mach=kw["mach"])
    Rref, s, _ = ref.compute_residuals(U)
    R = SpatialDisc(ctx).residuals(U)
    # FP32 arithmetic on the state: each conserved variable's residual against its own scale
    err = rel_l2_per_var(R, Rref)
    print(f"{name} {variant}: residual rel L2 per variable {err}", flush=True)
    assert err.max() <= FP32_RESID_TOL, err


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name,order,waves", [("periodic", 1, None), ("periodic", 3, 4), ("periodic", 3, 0),
                                              ("periodic", 4, None), ("box", 2, None), ("box", 3, 4)])
def test_uniform_flow_is_steady(precision, name, order, waves):
    """A uniform state equal to the far-field state is a steady solution of the wall-free meshes (the
    periodic square, the far-field box): its residual vanishes up to rounding. Measured per variable
    against the residual of a 5 % perturbation of the same state. In FP32-bulk the volume terms are
    exactly zero (every point's deviation from the element mean is zero), so what is left is the FP32
    rounding of the edge fluxes; both element kernels at p = 3 (waves 4: matrix cores, 0: one lane).
    Measured: at most 1.0e-11 (FP64) and 1.8e-5 (FP32-bulk), the momentum on the far-field box, whose
    characteristic boundary flux rounds at the pressure's scale; the bounds sit ten times above that,
    while a flux or boundary term that does not preserve the uniform state gives ratios of 1e-2 to 1."""
    mesh = meshgen.periodic_square(6, seed=5) if name == "periodic" else meshgen.box_farfield(5, seed=6)
    kw = dict(mach=0.5, aoa_deg=1.25)
    ref, ctx = pair(mesh, order, limiter="off", precision=precision, **kw)
    if waves is not None:
        t = ctx.tuning()
        t["element_waves"] = waves
        ctx.set_tuning(t)
    U = ctx.initial_solution(1.0, kw["aoa_deg"], kw["mach"], 1.0e5)
    R = SpatialDisc(ctx).residuals(U)
    Rp = SpatialDisc(ctx).residuals(flow_state(ref, seed=21, amp=0.05, mach=kw["mach"]))
    ratio = np.array([np.abs(R[:, v]).max() / np.abs(Rp[:, v]).max() for v in range(4)])
    print(f"{name} p={order} {precision} waves={waves}: uniform / perturbed residual per variable {ratio}", flush=True)
    assert ratio.max() <= (1e-10 if precision == "fp64" else 2e-4), ratio


@pytest.mark.parametrize("name,cfl", [("ogrid_p3", 0.3), ("ogrid_p1", 0.0), ("periodic_p4", 0.0)])
def test_time_march_fp32_bulk(all_meshes, name, cfl):
    """20 FP32-bulk steps against the FP64 oracle (cfl 0: the reference's CFL formula for the order)."""
    mesh, order, kw = all_meshes[name]
    ref, ctx = pair(mesh, order, limiter="off", precision="fp32", cfl=cfl, **kw)
    U0 = flow_state(ref, seed=18, mach=kw["mach"], amp=0.02)
    Uref, s, n, t = ref.time_march(U0, 20, 100.0)
    assert s == 0 and n == 20
    V = U0.copy()
    TemperalDisc(ctx, final_step=20, final_time=100.0).time_march(SpatialDisc(ctx), ctx.zeros(), V)
    err = rel_l2_per_var(V, Uref)
    print(f"{name}: 20-step state rel L2 per variable {err}", flush=True)
    assert err.max() <= 1e-6, err


@pytest.mark.parametrize("name", ["ogrid_p3", "airfoil_p2"])
def test_time_march_fp32_bulk_weno(all_meshes, name):
    """FP32-bulk with the WENO limiter (Jacobi) against the FP64 oracle, from a shocked state whose
    troubled cells are rebuilt every stage: 10 steps, per conserved variable."""
    mesh, order, kw = all_meshes[name]
    ref, ctx = pair(mesh, order, limiter="jacobi", precision="fp32", **kw)
    U0 = flow_state(ref, seed=21, shock=True, mach=kw["mach"], amp=0.02)
    ref_tr = ref.apply_limiter(U0)[1]
    assert ref_tr.sum() > 0
    Uref, s, n, t = ref.time_march(U0, 10, 100.0)
    assert s == 0 and n == 10
    V = U0.copy()
    TemperalDisc(ctx, final_step=10, final_time=100.0).time_march(SpatialDisc(ctx), ctx.zeros(), V)
    err = rel_l2_per_var(V, Uref)
    print(f"{name}: 10-step WENO state rel L2 per variable {err}, {int(ref_tr.sum())} troubled", flush=True)
    assert err.max() <= FP32_WENO_TOL, err


def test_negative_density_reports_edge(all_meshes):
    mesh, order, kw = all_meshes["airfoil_p2"]
    ref, ctx = pair(mesh, order, **kw)
    U = flow_state(ref, seed=19, mach=0.8)
    U[1234, 0, 0] = -0.5  # negative mean density in one interior element
    _, s, bad = ref.compute_residuals(U)
    assert s == 1
    with pytest.raises(DgmError) as ei:
        SpatialDisc(ctx).residuals(U)
    assert ei.value.status == 1
    assert ei.value.edge == bad


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_negative_pressure_status(all_meshes, precision):
    """FP64 tests the reference's pressure; FP32-bulk tests the sign of 2 rho E - |rho u|^2 from exact FP64
    products of its FP32 traces. Both report the same status at the same (lowest) edge."""
    mesh, order, kw = all_meshes["ogrid_p3"]
    ref, ctx = pair(mesh, order, precision=precision, **kw)
    U = flow_state(ref, seed=20, mach=0.8)
    U[77, 3, 0] = 0.0  # zero energy: negative pressure on its edges
    _, s, bad = ref.compute_residuals(U)
    assert s == 2
    with pytest.raises(DgmError) as ei:
        SpatialDisc(ctx).residuals(U)
    assert ei.value.status == 2 and ei.value.edge == bad


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_march_stage0_failure_status(all_meshes, precision):
    """A march whose first edge pass fails (that pass runs beside the step control and records into a
    word of its own, which the element pass takes over) stops with the status and lowest failing edge
    of the stand-alone residual, and its state is the one it started from."""
    mesh, order, kw = all_meshes["ogrid_p3"]
    ref, ctx = pair(mesh, order, precision=precision, limiter="jacobi", **kw)
    U = flow_state(ref, seed=20, mach=0.8)
    U[77, 3, 0] = 0.0
    _, s, bad = ref.compute_residuals(U)
    assert s == 2
    V = U.copy()
    with pytest.raises(DgmError) as ei:
        TemperalDisc(ctx, final_step=5, final_time=1.0e300).time_march(SpatialDisc(ctx), ctx.zeros(), V)
    assert ei.value.status == 2 and ei.value.edge == bad
    assert np.array_equal(V, U)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_troubled_counts_and_profile_by_limiter_order(all_meshes, precision, monkeypatch):
    """The troubled-cell counts the library reports. The stand-alone limiter's count equals its flags.
    After two marched steps the per-stage counts agree between the one-wavefront walk and one launch
    per level (the walk corrects the detector pass's count to the cells it wrote). Per-kernel timing
    reports the edge, element and detector parts of a step in both forms."""
    mesh, order, kw = all_meshes["ogrid_p3"]
    ref = RefCase(mesh, order=order, limiter=1, **kw)
    U = flow_state(ref, seed=14, shock=True, mach=kw["mach"], noise=0.05)
    Um = flow_state(ref, seed=15, shock=True, mach=kw["mach"], amp=0.02)  # marchable: 24 troubled cells
    counts, states = {}, {}
    for walk in ("1", "0"):
        monkeypatch.setenv("DGM_SEQUENTIAL_WALK", walk)
        ctx = Context(mesh, order=order, limiter="sequential", precision=precision, **kw)
        t = np.zeros(ctx.ne, np.uint8)
        SpatialDisc(ctx).apply_limiter(U.copy(), t)
        c = ctx.troubled_counts()
        assert c[3] == int(t.astype(bool).sum()) > 0
        ctx.upload(Um)
        st = ctx.device_time_march(2, 1.0e9)
        assert st["steps"] == 2
        counts[walk] = ctx.troubled_counts()[:3]
        states[walk] = ctx.download()
        prof = ctx.profile_steps(2)
        assert prof["surface_ms"] > 0 and prof["element_ms"] > 0 and prof["detect_ms"] > 0, prof
    assert counts["1"] == counts["0"] and sum(counts["1"]) > 0, counts
    assert np.array_equal(states["1"], states["0"])
