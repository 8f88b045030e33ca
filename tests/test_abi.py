"""C ABI surface (CPU): libdgm.so loads and exports every function include/dgm.h declares;
host-only contexts refuse compute calls; the product path has no CPU fallback."""
import ctypes
import os
import re

import numpy as np
import pytest

from dgm_in_rust_amd import Context, SpatialDisc, meshgen
from dgm_in_rust_amd import _lib
from dgm_in_rust_amd._lib import DgmError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "dgm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dgm_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for n in ("dgm_compute_residuals", "dgm_apply_limiter", "dgm_compute_time_step", "dgm_rk3", "dgm_time_march",
              "dgm_ctx_create", "dgm_ctx_destroy"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTED) <= set(declared_functions())


def test_host_only_context_refuses_compute():
    ctx = Context(meshgen.box_farfield(3), order=2, host_only=True)
    U = ctx.initial_solution(1.0, 0.0, 0.5, 1e5)
    with pytest.raises(DgmError) as e:
        SpatialDisc(ctx).residuals(U)
    assert e.value.status == 8


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libdgm.so"))
    with pytest.raises(RuntimeError, match="missing"):
        _lib.lib()


def test_initial_solution_matches_oracle():
    from oracle.refcpu import RefCase
    m = meshgen.naca0012_ogrid(16, 4, seed=1)
    for p in (1, 2, 3, 4):
        ref = RefCase(m, order=p)
        ctx = Context(m, order=p, host_only=True)
        a = ref.initial_solution(1.0, 1.25, 0.5, 1e5)
        b = ctx.initial_solution(1.0, 1.25, 0.5, 1e5)
        assert np.allclose(a, b, rtol=1e-14, atol=1e-9)
        xy = ctx.cell_points_xy()
        assert np.allclose(xy, ref.cell_points_xy(), rtol=0, atol=1e-14)
