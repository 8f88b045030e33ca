"""GPU: the allocation-free residual, dgm_compute_residuals_device_noalloc, against the existing entry
point dgm_compute_residuals (accumulate = 0). Both run the same surface and element kernels on the same
state, so the residuals must agree bit for bit, in FP64 and FP32-bulk, on two calls in a row whose
second input is smaller than the first (the scratch array keeps the first call's contents). Error
status and offending edge must agree too, and a working set beyond the scratch array is refused.
The caller's device buffers come from the HIP runtime that libdgm.so itself is linked against, called
through ctypes (a second runtime in the process, such as torch's bundled one, would not see the GPU)."""
import ctypes as C
import os

import numpy as np
import pytest

from dgm_in_rust_amd import Context, SpatialDisc, _lib, meshgen
from tests.helpers import flow_state

pytestmark = pytest.mark.gpu

NO_SPACE, NEG_PRESSURE = 9, 2
H2D, D2H = 1, 2


def hip():
    _lib.lib()   # libdgm.so first: the soname below then resolves to the runtime it loaded
    path = "/opt/rocm/lib/libamdhip64.so.7"
    rt = C.CDLL(path if os.path.exists(path) else "libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


def noalloc(ctx, U):
    """Caller side: device copies of U and of an all-NaN R, the call, and R read back."""
    rt = hip()
    U = np.ascontiguousarray(U, np.float64)
    R = np.empty_like(U)
    n = U.nbytes
    dU, dR = C.c_void_p(), C.c_void_p()
    assert rt.hipMalloc(C.byref(dU), n) == 0 and rt.hipMalloc(C.byref(dR), n) == 0
    try:
        assert rt.hipMemcpy(dU, U.ctypes.data_as(C.c_void_p), n, H2D) == 0
        assert rt.hipMemset(dR, 0xFF, n) == 0   # all-ones bytes: NaN in every double
        assert rt.hipDeviceSynchronize() == 0   # the call runs on the context's own stream
        s = _lib.lib().dgm_compute_residuals_device_noalloc(ctx.h, dU, dR)
        assert rt.hipMemcpy(R.ctypes.data_as(C.c_void_p), dR, n, D2H) == 0
    finally:
        rt.hipFree(dU)
        rt.hipFree(dR)
    return s, R


def last_error(ctx):
    e = _lib.Error()
    _lib.lib().dgm_last_error(ctx.h, C.byref(e))
    return e


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_residual_equals_entry_point_twice(precision):
    cases = [  # the second input is smaller than the first (fewer elements, edges and modes)
        (meshgen.naca0012_ogrid(160, 50, seed=7), 3, 0.8),
        (meshgen.naca0012_ogrid(48, 12, seed=8), 2, 0.5),
    ]
    for i, (mesh, order, mach) in enumerate(cases):
        print(f"call {i}: {precision}, p = {order}, {mesh.n_elements} triangles", flush=True)
        ctx = Context(mesh, order=order, precision=precision, limiter="jacobi", mach=mach, aoa_deg=1.25)
        U = flow_state(ctx, seed=20 + i, shock=True, mach=mach)
        want = SpatialDisc(ctx).residuals(U)
        s, got = noalloc(ctx, U)
        assert s == 0, last_error(ctx).message
        assert np.array_equal(got, want), np.nanmax(np.abs(got - want))
        ctx.close()


def test_noalloc_reports_the_entry_points_error():
    mesh = meshgen.naca0012_ogrid(40, 10, seed=9)
    ctx = Context(mesh, order=2, limiter="jacobi", mach=0.5, aoa_deg=0.0)
    U = ctx.initial_solution(1.0, 0.0, 0.5, 1.0e5)
    U[17, 3, 0] = -1.0e5   # negative energy: negative pressure on the element's interior edges
    with pytest.raises(_lib.DgmError) as ex:
        SpatialDisc(ctx).residuals(U)
    s, _ = noalloc(ctx, U)
    assert s == ex.value.status == NEG_PRESSURE
    assert last_error(ctx).edge == ex.value.edge >= 0


def test_noalloc_refuses_a_working_set_beyond_the_scratch():
    """128 000 triangles at p = 4 in FP64 need about 154 MB of working arrays (state and residual
    61 MB each, point fluxes 31 MB), more than the 128 MiB scratch array."""
    ctx = Context(meshgen.naca0012_ogrid(640, 100, seed=0), order=4, limiter="off", mach=0.5)
    U = ctx.initial_solution(1.0, 0.0, 0.5, 1.0e5)
    s, got = noalloc(ctx, U)
    assert s == NO_SPACE
    assert b"scratch" in last_error(ctx).message
    assert np.isnan(got).all()   # nothing was written
