"""Host-side mirror of the reference's solver surface, running on the HIP library.

Names, argument meaning and in-place semantics follow the Rust crate:

* :class:`SpatialDisc` — ``compute_residuals(residuals, solutions)`` (src/spatial_disc.rs:27-39)
  accumulates into ``residuals`` then divides by the mass; ``apply_limiter(solutions)``
  (src/spatial_disc/limiter.rs:7-9) limits in place.
* :class:`TemperalDisc` — ``compute_time_step`` (src/temporal_disc.rs:58-88), ``rk3``
  (src/temporal_disc/rk3.rs:7-26), ``time_march`` (src/temporal_disc.rs:21-57) and
  ``set_residual_to_zero`` (:89-93).
* :class:`Solver` — ``set_initial_solution`` (src/solver.rs:45-79) and ``solve`` (:80-82).

Arrays are numpy float64 in the reference layout [n_elements, 4, n_basis]. Where the reference
panics, these raise :class:`DgmError`. Everything runs in libdgm.so on the GPU; there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import DgmError, ptr
from .meshgen import MeshInput

_PREC = {"fp64": _lib.FP64, "fp32": _lib.FP32_BULK, "fp32_bulk": _lib.FP32_BULK}
_LIM = {"off": _lib.LIMITER_OFF, "sequential": _lib.LIMITER_SEQUENTIAL, "jacobi": _lib.LIMITER_JACOBI,
        None: _lib.LIMITER_OFF}


class Context:
    """One dgm_ctx: mesh, basis/quadrature tables and device buffers (initialize_mesh_basis_gauss_params)."""

    def __init__(self, mesh: MeshInput, order: int = 2, n_cell_gp: int = 0, n_edge_gp: int = 0, hcr: float = 1.4,
                 cfl: float = 0.0, mach: float = 0.8, aoa_deg: float = 1.5, rho: float = 1.0, p: float = 1.0e5,
                 fs_hcr: float = 1.4, precision: str = "fp64", limiter: Optional[str] = "sequential",
                 host_only: bool = False):
        L = _lib.lib()
        self._L = L
        m = mesh.contiguous()
        self.mesh = m
        self._mesh_struct = _lib.Mesh(m.vxy.shape[0], ptr(m.vxy), len(m.eb_type), ptr(m.eb_type, _lib._i32p),
                                      ptr(m.eb_count, _lib._i64p), ptr(m.eb_nodes, _lib._i64p), len(m.el_count),
                                      ptr(m.el_count, _lib._i64p), ptr(m.el_nodes, _lib._i64p),
                                      ptr(m.elem_xy) if m.elem_xy is not None else None)
        self.params = _lib.Params(order, n_cell_gp, n_edge_gp, hcr, cfl, mach, aoa_deg, rho, p, fs_hcr,
                                  _PREC[precision], _LIM[limiter])
        h = C.c_void_p()
        create = L.dgm_ctx_create_host if host_only else L.dgm_ctx_create
        s = create(C.byref(self._mesh_struct), C.byref(self.params), C.byref(h))
        if s != 0:
            raise DgmError(s, message="dgm_ctx_create failed")
        self.h = h
        info = _lib.Info()
        L.dgm_ctx_info(self.h, C.byref(info))
        self.info = info
        self.ne = info.n_elements
        self.nbasis = info.n_basis
        self.precision = precision
        self.limiter = limiter

    def close(self):
        if getattr(self, "h", None):
            self._L.dgm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def shape(self):
        return (self.ne, 4, self.nbasis)

    def zeros(self):
        return np.zeros(self.shape)

    def check(self, status: int):
        if status != 0:
            e = _lib.Error()
            self._L.dgm_last_error(self.h, C.byref(e))
            raise DgmError(status, e.edge, e.element, e.message.decode(errors="replace"))

    # --- set-up queries -------------------------------------------------------------------
    def connectivity(self):
        i = self.info
        n, ne = i.n_edges, i.n_elements
        d = dict(ielements=np.zeros((n, 2), np.int64), ivertices=np.zeros((n, 2), np.int64),
                 in_cell_indices=np.zeros((n, 2), np.int64), ipatch=np.zeros(n, np.int64),
                 element_vertices=np.zeros((ne, 3), np.int64), element_edges=np.zeros((ne, 3), np.int64),
                 ineighbours=np.zeros((ne, 3), np.int64), internal_edges=np.zeros(i.n_internal_edges, np.int64),
                 internal_elements=np.zeros(i.n_internal_elements, np.int64),
                 boundary_elements=np.zeros(i.n_boundary_elements, np.int64))
        self.check(self._L.dgm_get_connectivity(self.h, *[ptr(v, _lib._i64p) for v in d.values()]))
        return d

    def geometry(self):
        i = self.info
        d = dict(normal=np.zeros((i.n_edges, 2)), edge_length=np.zeros(i.n_edges), jacob_det=np.zeros(i.n_elements),
                 circumradius=np.zeros(i.n_elements), minimum_height=np.zeros(i.n_elements))
        self.check(self._L.dgm_get_geometry(self.h, *[ptr(v) for v in d.values()]))
        return d

    def tables(self):
        i = self.info
        nq, ng, nb = i.n_cell_gp, i.n_edge_gp, i.n_basis
        d = dict(phi_cell=np.zeros((nq, nb)), phi_edge=np.zeros((3, ng, nb)), mass=np.zeros(nb),
                 cell_pts=np.zeros((nq, 2)), cell_w=np.zeros(nq), edge_pts=np.zeros((3, ng, 2)), edge_w=np.zeros(ng))
        self.check(self._L.dgm_get_tables(self.h, *[ptr(v) for v in d.values()]))
        return d

    def initial_solution(self, density=1.0, aoa_deg=1.5, mach=0.8, pressure=1.0e5):
        U = self.zeros()
        self.check(self._L.dgm_initial_solution(self.h, density, aoa_deg, mach, pressure, ptr(U)))
        return U

    def project(self, q_at_points):
        q = np.ascontiguousarray(q_at_points, np.float64)
        U = self.zeros()
        self.check(self._L.dgm_project(self.h, ptr(q), ptr(U)))
        return U

    def cell_points_xy(self):
        xy = np.zeros((self.ne, self.info.n_cell_gp, 2))
        self.check(self._L.dgm_cell_points_xy(self.h, ptr(xy)))
        return xy

    # --- resident state -------------------------------------------------------------------
    def upload(self, U):
        U = np.ascontiguousarray(U, np.float64)
        self.check(self._L.dgm_state_upload(self.h, ptr(U)))

    def download(self):
        U = self.zeros()
        self.check(self._L.dgm_state_download(self.h, ptr(U)))
        return U

    def device_time_march(self, final_step, final_time):
        st = _lib.MarchStats()
        self.check(self._L.dgm_device_time_march(self.h, final_step, final_time, C.byref(st)))
        return dict(steps=st.steps, time=st.time, last_dt=st.last_dt, troubled=st.troubled_cells_last_stage)

    def steps_async(self, n):
        self.check(self._L.dgm_device_steps_async(self.h, n))

    def synchronize(self):
        self.check(self._L.dgm_synchronize(self.h))

    def stream(self):
        return self._L.dgm_stream(self.h)

    def clock(self):
        t = C.c_double(); n = C.c_int32(); dt = C.c_double()
        self.check(self._L.dgm_device_clock(self.h, C.byref(t), C.byref(n), C.byref(dt)))
        return t.value, n.value, dt.value

    def check_state(self):
        self.check(self._L.dgm_check_state(self.h))

    def profile_steps(self, n):
        """Run n steps eagerly with HIP events around every launch; mean ms per launch by kernel."""
        kt = _lib.KernelTimes()
        self.check(self._L.dgm_profile_steps(self.h, n, C.byref(kt)))
        return {k: getattr(kt, k) for k, _ in _lib.KernelTimes._fields_}

    def troubled_counts(self):
        c = (C.c_int32 * 4)()
        self.check(self._L.dgm_troubled_counts(self.h, c))
        return list(c)

    def tuning(self) -> dict:
        t = _lib.Tuning()
        self.check(self._L.dgm_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, values: dict):
        t = _lib.Tuning(**{k: int(values[k]) for k, _ in _lib.Tuning._fields_})
        self.check(self._L.dgm_set_tuning(self.h, C.byref(t)))

    def autotune(self, repeats: int = 10) -> dict:
        """Times the kernels' launch shapes on the resident state and keeps the fastest (the state and
        the clock are left as they are)."""
        t = _lib.Tuning()
        self.check(self._L.dgm_autotune(self.h, repeats, C.byref(t)))
        return t.as_dict()

    def lift_coefficient(self, chord=1.0):
        cl = C.c_double()
        self.check(self._L.dgm_lift_coefficient(self.h, chord, C.byref(cl)))
        return cl.value


class SpatialDisc:
    """SpatialDisc (src/spatial_disc.rs:17-39) on the GPU."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def compute_residuals(self, residuals: np.ndarray, solutions: np.ndarray) -> None:
        """Reference semantics: residuals <- M^-1 (residuals + V - S - B), in place."""
        _check_array(residuals, self.ctx)
        U = np.ascontiguousarray(solutions, np.float64)
        self.ctx.check(self.ctx._L.dgm_compute_residuals(self.ctx.h, ptr(residuals), ptr(U), 1))

    def residuals(self, solutions: np.ndarray) -> np.ndarray:
        """Overwrite form: returns M^-1 (V - S - B)."""
        R = self.ctx.zeros()
        U = np.ascontiguousarray(solutions, np.float64)
        self.ctx.check(self.ctx._L.dgm_compute_residuals(self.ctx.h, ptr(R), ptr(U), 0))
        return R

    def apply_limiter(self, solutions: np.ndarray, troubled: Optional[np.ndarray] = None) -> None:
        _check_array(solutions, self.ctx)
        t = None if troubled is None else troubled
        if t is not None and (t.dtype != np.uint8 or t.size != self.ctx.ne):
            raise ValueError("troubled must be a uint8 array of n_elements")
        self.ctx.check(self.ctx._L.dgm_apply_limiter(self.ctx.h, ptr(solutions), ptr(t, _lib._u8p)))


class TemperalDisc:
    """TemperalDisc (src/temporal_disc.rs:8-93) on the GPU (the reference's spelling is kept)."""

    def __init__(self, ctx: Context, final_step: int = 20, final_time: float = 5.0, output_dir: Optional[str] = None):
        self.ctx = ctx
        self.final_step = final_step
        self.final_time = final_time
        self.current_time = 0.0
        self.current_step = 0
        # the reference writes <dir>/step=N.vtk at the end of every march (src/temporal_disc.rs:56); here
        # only when a folder is given
        self.output_dir = output_dir

    def compute_time_step(self, solutions: np.ndarray) -> float:
        U = np.ascontiguousarray(solutions, np.float64)
        dt = C.c_double()
        self.ctx.check(self.ctx._L.dgm_compute_time_step(self.ctx.h, ptr(U), C.byref(dt)))
        return dt.value

    @staticmethod
    def set_residual_to_zero(residuals: np.ndarray) -> None:
        residuals[...] = 0.0

    def rk3(self, spatial_disc: SpatialDisc, residuals: np.ndarray, solutions: np.ndarray, time_step: float) -> None:
        _check_array(solutions, self.ctx)
        self.ctx.check(self.ctx._L.dgm_rk3(self.ctx.h, ptr(solutions), time_step))

    def time_march(self, spatial_disc: SpatialDisc, residuals: np.ndarray, solutions: np.ndarray) -> dict:
        _check_array(solutions, self.ctx)
        st = _lib.MarchStats()
        self.ctx.check(self.ctx._L.dgm_time_march(self.ctx.h, ptr(solutions), self.final_step, self.final_time,
                                                  C.byref(st)))
        self.current_step = st.steps
        self.current_time = st.time
        if self.output_dir is not None:
            from .vtk import write_solution
            write_solution(self.ctx, solutions, self.current_step, self.output_dir)
        return dict(steps=st.steps, time=st.time, last_dt=st.last_dt, troubled=st.troubled_cells_last_stage)


class Solver:
    """Solver (src/solver.rs:32-82): owns residuals / solutions and the two discretisations."""

    def __init__(self, ctx: Context, final_step: int = 20, final_time: float = 5.0):
        self.ctx = ctx
        self.residuals = ctx.zeros()
        self.solutions = ctx.zeros()
        self.spatial_disc = SpatialDisc(ctx)
        self.temporal_disc = TemperalDisc(ctx, final_step, final_time)

    def set_initial_solution(self, density=1.0, aoa_deg=1.5, mach=0.8, pressure=1.0e5) -> None:
        self.solutions[...] = self.ctx.initial_solution(density, aoa_deg, mach, pressure)

    def solve(self) -> dict:
        return self.temporal_disc.time_march(self.spatial_disc, self.residuals, self.solutions)


def _check_array(a: np.ndarray, ctx: Context) -> None:
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous and a.shape == ctx.shape):
        raise ValueError(f"expected a C-contiguous float64 array of shape {ctx.shape}")
