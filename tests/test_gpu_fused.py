"""GPU: the fused edge-and-element pass of FP32-bulk p = 3 marches (DGM_FUSED=1 at set-up: one wavefront
per 16 elements evaluates its elements' edge fluxes into LDS, then updates them) against the separate
edge and element passes (the default). The fluxes are the same operations on the same inputs, so the
states, clocks and troubled-cell counts must agree bit for bit, with the limiter's pending rebuilds in
play, on wall, far-field and periodic meshes; a failing edge stops the case with the same status and
edge."""
import os

import numpy as np
import pytest

from dgm_in_rust_amd import Context, SpatialDisc, TemperalDisc, cases, meshgen
from dgm_in_rust_amd._lib import DgmError

pytestmark = pytest.mark.gpu


def make(mesh, fused, **kw):
    saved = os.environ.get("DGM_FUSED")
    os.environ["DGM_FUSED"] = "1" if fused else "0"
    try:
        return Context(mesh, order=3, precision="fp32", **kw)
    finally:
        if saved is None:
            os.environ.pop("DGM_FUSED", None)
        else:
            os.environ["DGM_FUSED"] = saved


def run_both(mesh, start, steps, **kw):
    out = {}
    for fused in (True, False):
        ctx = make(mesh, fused, **kw)
        assert ctx.info.fused_pass == (1 if fused else 0)
        U = start(ctx)
        TemperalDisc(ctx, final_step=steps, final_time=1.0e300).time_march(SpatialDisc(ctx), ctx.zeros(), U)
        out[fused] = (U, ctx.clock(), ctx.troubled_counts())
        ctx.close()
    return out


def check_equal(out, steps):
    (Uf, cf, tf), (Us, cs, ts) = out[True], out[False]
    assert cf[1] == cs[1] == steps
    assert np.array_equal(Uf, Us)
    assert cf == cs and tf == ts
    return tf


def test_fused_equals_separate_with_shock_and_pending_rebuilds():
    mesh = meshgen.naca0012_ogrid(160, 50, seed=0)
    out = run_both(mesh, lambda c: cases.shocked_start(c, 0.8, 1.25), 30, mach=0.8, aoa_deg=1.25, limiter="jacobi")
    counts = check_equal(out, 30)
    assert sum(counts[:3]) > 0   # rebuilds were pending between the stages


def test_fused_equals_separate_on_the_reference_airfoil(airfoil_mesh):
    out = run_both(airfoil_mesh, lambda c: c.initial_solution(1.0, 1.5, 0.8, 1.0e5), 20, mach=0.8, aoa_deg=1.5,
                   limiter="jacobi")
    check_equal(out, 20)


def test_fused_equals_separate_periodic_vortex():
    mesh = meshgen.periodic_square(24, seed=2)
    out = run_both(mesh, lambda c: cases.vortex_state(c, 0.0), 20, mach=0.5, aoa_deg=0.0, limiter="off")
    check_equal(out, 20)


def test_fused_failure_reports_the_separate_status_and_edge():
    mesh = meshgen.naca0012_ogrid(40, 10, seed=3)
    got = {}
    for fused in (True, False):
        ctx = make(mesh, fused, mach=0.8, aoa_deg=1.25, limiter="jacobi")
        assert ctx.info.fused_pass == (1 if fused else 0)
        U = ctx.initial_solution(1.0, 1.25, 0.8, 1.0e5)
        U[77, 3, 0] = 0.0   # zero energy: negative pressure on the element's edges in the first stage
        V = U.copy()
        with pytest.raises(DgmError) as ei:
            TemperalDisc(ctx, final_step=5, final_time=1.0e300).time_march(SpatialDisc(ctx), ctx.zeros(), V)
        got[fused] = (ei.value.status, ei.value.edge)
        assert np.array_equal(V, U)   # a first-stage failure leaves u^n as it was
        ctx.close()
    assert got[True] == got[False] and got[True][0] == 2
