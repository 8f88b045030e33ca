"""GPU: accuracy on the isentropic vortex (exact solution), FP32-bulk against FP64, the lift
coefficient, and the benchmark O-grid configuration against the oracle."""
import numpy as np
import pytest

from dgm_in_rust_amd import Context, SpatialDisc, TemperalDisc, cases, meshgen
from oracle.refcpu import RefCase
from tests.helpers import rel_err, rel_l2, rel_l2_per_var

pytestmark = pytest.mark.gpu


def march(ctx, U0, steps, final_time):
    V = U0.copy()
    st = TemperalDisc(ctx, final_step=steps, final_time=final_time).time_march(SpatialDisc(ctx), ctx.zeros(), V)
    return V, st


@pytest.mark.parametrize("order", [1, 2, 3])
def test_vortex_convergence(order):
    """The L2 density error at t = 0.5 falls at close to the design rate h^(p+1) under refinement."""
    errs = []
    for n in (8, 16, 32):
        ctx = Context(meshgen.periodic_square(n, jitter=0.1, seed=1), order=order, limiter="off",
                      mach=0.5, aoa_deg=0.0)
        U0 = cases.vortex_state(ctx, 0.0)
        U, st = march(ctx, U0, 100000, 0.5)
        assert abs(st["time"] - 0.5) < 1e-12
        errs.append(cases.vortex_error(ctx, U, 0.5)[0])
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert rates[-1] > order + 0.5, (errs, rates)


def test_vortex_fp64_matches_oracle():
    mesh = meshgen.periodic_square(8, seed=2)
    ref = RefCase(mesh, order=3, limiter=0)
    ctx = Context(mesh, order=3, limiter="off")
    U0 = cases.vortex_state(ctx, 0.0)
    Uref, s, n, t = ref.time_march(U0, 1000, 0.3)
    U, st = march(ctx, U0, 1000, 0.3)
    assert st["steps"] == n and abs(st["time"] - t) <= 1e-14
    assert rel_err(U, Uref) <= 1e-12


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_vortex_fp32_bulk_against_fp64(order):
    mesh = meshgen.periodic_square(24, seed=3)
    c64 = Context(mesh, order=order, limiter="off")
    c32 = Context(mesh, order=order, limiter="off", precision="fp32")
    U0 = cases.vortex_state(c64, 0.0)
    U64, _ = march(c64, U0, 100000, 1.0)
    U32, _ = march(c32, U0, 100000, 1.0)
    err = cases.relative_l2_per_variable(c64, U32, U64)
    print(f"p={order}: FP32-bulk vs FP64 rel L2 per variable {err}", flush=True)
    assert max(err.values()) <= 1e-6, err


@pytest.mark.parametrize("order,n", [(1, 512), (2, 256), (2, 512), (3, 128), (3, 256), (3, 512), (4, 128), (4, 256),
                                     (4, 512)])
def test_vortex_fp32_bulk_within_1e6_of_fp64(order, n):
    """The isentropic vortex on the N x N square to t = 2 (the accuracy study's case): with the residual
    in deviation form (point fluxes as differences from the mean's flux) and the stage-2 shadow rounded
    once, FP32-bulk stays within 1e-6 of FP64 per variable up to N = 512 at every order (at most 7.7e-7
    measured, E at p = 4, N = 512)."""
    mesh = meshgen.periodic_square(n, jitter=0.1, seed=1)
    c64 = Context(mesh, order=order, limiter="off", mach=0.5, aoa_deg=0.0)
    c32 = Context(mesh, order=order, limiter="off", mach=0.5, aoa_deg=0.0, precision="fp32")
    U0 = cases.vortex_state(c64, 0.0)
    U64, _ = march(c64, U0, 10 ** 8, 2.0)
    U32, _ = march(c32, U0, 10 ** 8, 2.0)
    err = cases.relative_l2_per_variable(c64, U32, U64)
    print(f"p={order}, N={n}: FP32-bulk vs FP64 rel L2 per variable {err}", flush=True)
    assert max(err.values()) <= 1e-6, err


def test_lift_zero_for_uniform_pressure():
    ctx = Context(meshgen.naca0012_ogrid(64, 16, seed=4), order=2, mach=0.5, aoa_deg=3.0)
    U = ctx.initial_solution(1.0, 3.0, 0.5, 1.0e5)
    ctx.upload(U)
    assert abs(ctx.lift_coefficient()) < 1e-10


def test_lift_matches_host_integral():
    """Cl from the device wall-pressure integral equals the same integral evaluated on the host."""
    mesh = meshgen.naca0012_ogrid(48, 12, seed=5)
    ctx = Context(mesh, order=2, mach=0.6, aoa_deg=2.0)
    xy = ctx.cell_points_xy()
    rho = 1.0 + 0.1 * np.sin(2 * xy[..., 0]) * np.cos(xy[..., 1])
    p = 1.0e5 * (1.0 + 0.05 * np.cos(3 * xy[..., 0] + xy[..., 1]))
    U = ctx.project(cases.conservative(rho, 100.0 + 0 * rho, 10.0 + 0 * rho, p))
    ctx.upload(U)
    cl = ctx.lift_coefficient()
    con, geo, tab = ctx.connectivity(), ctx.geometry(), ctx.tables()
    fx = fy = 0.0
    nb = int(mesh.eb_count.sum())
    for k in range(nb):
        if mesh.eb_type[con["ipatch"][k]] != 0:   # wall blocks only
            continue
        e, lic = con["ielements"][k, 0], con["in_cell_indices"][k, 0]
        n = geo["normal"][k] * (1.0 if con["element_vertices"][e, lic] == con["ivertices"][k, 0] else -1.0)
        tr = U[e] @ tab["phi_edge"][lic].T        # [4, ne]
        pw = 0.4 * (tr[3] - 0.5 * (tr[1] ** 2 + tr[2] ** 2) / tr[0])
        fx += float(np.sum(pw * n[0] * tab["edge_w"])) * 0.5 * geo["edge_length"][k]
        fy += float(np.sum(pw * n[1] * tab["edge_w"])) * 0.5 * geo["edge_length"][k]
    a = np.radians(2.0)
    V = 0.6 * np.sqrt(1.4e5)
    expect = (fy * np.cos(a) - fx * np.sin(a)) / (0.5 * V * V)
    assert abs(cl - expect) <= 1e-10 * max(1.0, abs(expect))


def test_ogrid_16k_short_run_matches_oracle():
    """The 160 x 50 O-grid (16k triangles), p = 2, M = 0.5, AoA 1.25, CFL 0.3, sequential
    WENO, marched from freestream against the oracle.

    The impulsive start makes the wall-layer cells troubled, and WENO's nonlinear weights there
    amplify roundoff: perturbing the oracle's own initial state by one ulp moves its 5-step result
    by about 1e-12 in the per-variable max norm (the L2 norm moves by 3e-16). So after 3 steps the
    max norm is held to 1e-12, and after 5 steps the L2 norm to 1e-14 and the max norm to ten times
    the oracle's own one-ulp spread, measured here."""
    mesh = meshgen.naca0012_ogrid(160, 50, seed=0)
    kw = dict(order=2, cfl=0.3, mach=0.5, aoa_deg=1.25)
    ref = RefCase(mesh, limiter=1, **kw)
    ctx = Context(mesh, limiter="sequential", **kw)
    U0 = ref.initial_solution(1.0, 1.25, 0.5, 1.0e5)

    Uref, s, n, _ = ref.time_march(U0, 3, 1.0e9)
    U, st = march(ctx, U0, 3, 1.0e9)
    assert s == 0 and st["steps"] == n == 3
    assert rel_err(U, Uref) <= 1e-12

    Uref, s, n, _ = ref.time_march(U0, 5, 1.0e9)
    U, st = march(ctx, U0, 5, 1.0e9)
    assert s == 0 and st["steps"] == n == 5
    rng = np.random.default_rng(0)
    spread = 0.0
    for _ in range(3):
        V = U0 * (1.0 + rng.choice([-1.0, 0.0, 1.0], size=U0.shape) * 2.0 ** -52)
        spread = max(spread, rel_err(ref.time_march(V, 5, 1.0e9)[0], Uref))
    assert 0.0 < spread < 1e-10
    assert rel_l2(U, Uref) <= 1e-14
    assert rel_err(U, Uref) <= 10.0 * spread, (rel_err(U, Uref), spread)


def test_ogrid_16k_100_steps_within_oracle_sensitivity():
    """The 16k-triangle O-grid march (p = 2, M = 0.5, CFL 0.3) to step 100, before any limiter decision
    near the KXRCF threshold flips. The state is held to 1e-12 in relative L2, and in the per-variable max norm to 20 times the oracle's
    own spread under one-ulp perturbations of its initial state. (Between steps 100 and 200 a
    marginal troubled-cell decision flips and the trajectory becomes chaotic: the oracle's own one-ulp
    spread grows to about 1e-1 by step 500.)"""
    mesh = meshgen.naca0012_ogrid(160, 50, seed=0)
    kw = dict(order=2, cfl=0.3, mach=0.5, aoa_deg=1.25)
    ref = RefCase(mesh, limiter=1, nthreads=16, **kw)
    ctx = Context(mesh, limiter="sequential", **kw)
    U0 = ref.initial_solution(1.0, 1.25, 0.5, 1.0e5)
    Uref, s, n, _ = ref.time_march(U0, 100, 1.0e9)
    U, st = march(ctx, U0, 100, 1.0e9)
    assert s == 0 and st["steps"] == n == 100
    rng = np.random.default_rng(1)
    spread = 0.0
    for _ in range(2):
        V = U0 * (1.0 + rng.choice([-1.0, 0.0, 1.0], size=U0.shape) * 2.0 ** -52)
        spread = max(spread, rel_err(ref.time_march(V, 100, 1.0e9)[0], Uref))
    assert 0.0 < spread < 1e-9
    assert rel_l2(U, Uref) <= 1e-12
    assert rel_err(U, Uref) <= 20.0 * spread, (rel_err(U, Uref), spread)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ogrid_16k_walk_equals_level_launches(precision, monkeypatch):
    """The 160 x 50 O-grid march (16k triangles, p = 2, M = 0.5, CFL 0.3), 200 steps from freestream with the
    element-order limiter: the one-wavefront walk and one launch per level of the schedule give the
    same state bit for bit, in both precisions, through the impulsive start and past the step where
    marginal decisions begin to flip (so every re-detection the walk makes is exercised on the real
    workload, stage after stage)."""
    mesh = meshgen.naca0012_ogrid(160, 50, seed=0)
    kw = dict(order=2, cfl=0.3, mach=0.5, aoa_deg=1.25, precision=precision)
    out = []
    for walk in ("1", "0"):
        monkeypatch.setenv("DGM_SEQUENTIAL_WALK", walk)
        ctx = Context(mesh, limiter="sequential", **kw)
        U0 = ctx.initial_solution(1.0, 1.25, 0.5, 1.0e5)
        U, st = march(ctx, U0, 200, 1.0e9)
        assert st["steps"] == 200
        out.append(U)
    assert np.array_equal(out[0], out[1])


def test_benchmark_case_full_size():
    """The benchmark's case at its full size: the NACA 0012 O-grid of 640 x 100 quads (128 000 triangles),
    p = 3, M 0.8, AoA 1.25, Jacobi WENO. From the benchmark's shocked start, in FP64 against the oracle:
    the residual, each entry within 1e-12 of the magnitude of its element's volume, edge and boundary
    terms (the wall cells are far smaller than the shock's, so their terms, and the rounding of their
    sums, are far larger than the largest residual; a max-norm against that residual reads about 2e-11
    here); the limiter in the Jacobi order and in the reference's element order (the same troubled cells,
    the same rebuilt state) and two time steps within 1e-12 per variable; and the FP32-bulk residual within 1e-5 of its element's terms. From the
    benchmark's freestream start: 20 FP32-bulk steps within 1e-6 of 20 FP64 steps per variable (the
    bench compares 2100)."""
    import os
    mesh = meshgen.naca0012_ogrid(640, 100, seed=0)
    kw = dict(mach=0.8, aoa_deg=1.25)
    c64 = Context(mesh, order=3, limiter="jacobi", **kw)
    assert c64.ne == 128000
    ref = RefCase(mesh, order=3, limiter=2, nthreads=min(16, os.cpu_count() or 1), **kw)
    U0 = cases.shocked_start(c64, kw["mach"], kw["aoa_deg"])
    Rref, s, _ = ref.compute_residuals(U0)
    assert s == 0
    R = SpatialDisc(c64).residuals(U0)
    V, E, B, s = ref.terms(U0)
    assert s == 0
    # per element and variable, the largest mass-scaled term over the modes (mode 0's terms cancel to
    # nothing in a uniform region; the higher modes' volume terms still carry the flux's scale there)
    scale = ref.divide_by_mass(np.abs(V) + np.abs(E) + np.abs(B)).max(axis=2, keepdims=True)
    dev = float((np.abs(R - Rref) / np.maximum(scale, 1e-300)).max())
    assert dev <= 1e-12, dev
    Lref, tref, s = ref.apply_limiter(U0)
    assert s == 0 and tref.sum() > 100, tref.sum()
    L = U0.copy()
    t = np.zeros(c64.ne, np.uint8)
    SpatialDisc(c64).apply_limiter(L, t)
    assert np.array_equal(t.astype(bool), tref)
    assert rel_err(L, Lref) <= 1e-12, rel_err(L, Lref)
    # the reference's own limiter order (the in-place element-order sweep, one walk launch at this size)
    rseq = RefCase(mesh, order=3, limiter=1, nthreads=min(16, os.cpu_count() or 1), **kw)
    Sref, sref, s = rseq.apply_limiter(U0)
    assert s == 0
    cseq = Context(mesh, order=3, limiter="sequential", **kw)
    S = U0.copy()
    ts = np.zeros(cseq.ne, np.uint8)
    SpatialDisc(cseq).apply_limiter(S, ts)
    assert np.array_equal(ts.astype(bool), sref)
    assert rel_err(S, Sref) <= 1e-12, rel_err(S, Sref)
    print(f"128k shocked start, element-order limiter: {int(sref.sum())} troubled, {rel_err(S, Sref):.2e}",
          flush=True)
    cseq.close()
    Uref, s, n, _ = ref.time_march(U0.copy(), 2, 1.0e9)
    assert s == 0 and n == 2
    V, st = march(c64, U0, 2, 1.0e9)
    assert st["steps"] == 2
    print(f"128k shocked start: residual {dev:.2e} of its terms ({rel_err(R, Rref):.2e} max-norm), limiter "
          f"{rel_err(L, Lref):.2e} "
          f"({int(tref.sum())} troubled), two steps {rel_err(V, Uref):.2e}", flush=True)
    assert rel_err(V, Uref) <= 1e-12, rel_err(V, Uref)
    c32 = Context(mesh, order=3, limiter="jacobi", precision="fp32", **kw)
    # The FP32-bulk residual of the shocked start against the oracle, per entry against its element's
    # terms: FP32 rounding of the few dozen terms each entry sums, 1e-6 of them a priori, bounded at 1e-5.
    # (The relative L2 over the whole mesh, the smaller meshes' measure, reads up to 1.2e-4 here, on the
    # transverse momentum: the wall cells' residuals are small differences of very large terms. In the
    # march their time steps are as small, and the 20 steps below stay within 3e-10.)
    R32 = SpatialDisc(c32).residuals(U0)
    e32 = rel_l2_per_var(R32, Rref)
    dev32 = float((np.abs(R32 - Rref) / np.maximum(scale, 1e-300)).max())
    print(f"128k shocked start: FP32-bulk residual {dev32:.2e} of its terms (rel L2 per variable {e32})", flush=True)
    assert dev32 <= 1e-5, dev32
    F0 = c64.initial_solution(1.0, kw["aoa_deg"], kw["mach"], 1.0e5)
    U64, _ = march(c64, F0, 20, 1.0e9)
    U32, _ = march(c32, F0, 20, 1.0e9)
    err = cases.relative_l2_per_variable(c64, U32, U64)
    print(f"128k freestream start, 20 steps: FP32-bulk vs FP64 per variable {err}", flush=True)
    assert max(err.values()) <= 1e-6, err
