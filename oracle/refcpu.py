"""ctypes binding of the oracle library oracle/_build/librefcpu.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker. The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "librefcpu.so")

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


class RefParams(C.Structure):
    _fields_ = [("order", C.c_int32), ("n_cell_gp", C.c_int32), ("n_edge_gp", C.c_int32), ("hcr", C.c_double),
                ("cfl", C.c_double), ("fs_mach", C.c_double), ("fs_aoa_deg", C.c_double), ("fs_rho", C.c_double),
                ("fs_p", C.c_double), ("fs_hcr", C.c_double), ("limiter", C.c_int32), ("nthreads", C.c_int32),
                ("gemm_fma", C.c_int32)]


def plant_path(plant: int) -> str:
    return os.path.join(HERE, "_build", f"librefcpu_plant{int(plant)}.so")


def build(force: bool = False, plants: bool = False) -> str:
    """Builds the oracle (and, with plants, its planted-mistake variants) when a source is newer."""
    src = [os.path.join(HERE, f) for f in ("refcpu.cpp", "refcpu.h")]
    newest = max(os.path.getmtime(f) for f in src)
    targets = [LIB] + ([plant_path(n) for n in range(1, 10)] if plants else [])
    stale = [t for t in targets if not os.path.exists(t) or os.path.getmtime(t) < newest]
    if force or stale:
        subprocess.run(["make", "-s", "-j8", "-C", HERE, "all"] + (["plants"] if plants else []), check=True)
    return LIB


_libs = {}


def lib(plant: int = 0):
    """The oracle library, or with plant = n its planted-mistake variant n (tests only)."""
    if plant not in _libs:
        path = LIB if plant == 0 else plant_path(plant)
        if not os.path.exists(path):
            build(plants=plant != 0)
        L = C.CDLL(path)
        L.ref_gmsh_parse.restype = C.c_void_p
        L.ref_gmsh_parse.argtypes = [C.c_char_p]
        L.ref_gmsh_sizes.argtypes = [C.c_void_p] + [_i64p] * 5
        L.ref_gmsh_get.argtypes = [C.c_void_p, _f64p, _i32p, _i64p, _i64p, _i64p, _i64p]
        L.ref_gmsh_free.argtypes = [C.c_void_p]
        L.ref_create.restype = C.c_void_p
        L.ref_create.argtypes = [C.c_int64, _f64p, C.c_int64, _i32p, _i64p, _i64p, C.c_int64, _i64p, _i64p, _f64p,
                                 C.POINTER(RefParams), _i32p]
        L.ref_destroy.argtypes = [C.c_void_p]
        L.ref_sizes.argtypes = [C.c_void_p] + [_i64p] * 5 + [_i32p] * 3
        L.ref_get_edges.argtypes = [C.c_void_p, _i64p, _i64p, _i64p, _i64p, _f64p, _f64p]
        L.ref_get_elements.argtypes = [C.c_void_p, _i64p, _i64p, _i64p, _f64p, _f64p, _f64p]
        L.ref_get_lists.argtypes = [C.c_void_p, _i64p, _i64p, _i64p]
        L.ref_n_deriv_keys.argtypes = [C.c_void_p]
        L.ref_get_tables.argtypes = [C.c_void_p] + [_f64p] * 8
        L.ref_cfl.restype = C.c_double
        L.ref_cfl.argtypes = [C.c_void_p]
        L.ref_last_bad_index.restype = C.c_int64
        L.ref_last_bad_index.argtypes = [C.c_void_p]
        L.ref_phys_deriv.restype = C.c_double
        L.ref_phys_deriv.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        L.ref_initial_solution.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, _f64p]
        L.ref_project.argtypes = [C.c_void_p, _f64p, _f64p]
        L.ref_cell_points_xy.argtypes = [C.c_void_p, _f64p]
        L.ref_compute_residuals.argtypes = [C.c_void_p, _f64p, _f64p, _i64p]
        L.ref_volume_terms.argtypes = [C.c_void_p, _f64p, _f64p]
        L.ref_edge_terms.argtypes = [C.c_void_p, _f64p, _f64p, _i64p]
        L.ref_boundary_terms.argtypes = [C.c_void_p, _f64p, _f64p]
        L.ref_divide_by_mass.argtypes = [C.c_void_p, _f64p]
        L.ref_apply_limiter.argtypes = [C.c_void_p, _f64p, _u8p]
        L.ref_kxrcf.argtypes = [C.c_void_p, _f64p, C.c_int64]
        L.ref_weno_element.argtypes = [C.c_void_p, _f64p, C.c_int64, _f64p]
        L.ref_compute_time_step.restype = C.c_double
        L.ref_compute_time_step.argtypes = [C.c_void_p, _f64p]
        L.ref_rk3.argtypes = [C.c_void_p, _f64p, _f64p, C.c_double]
        L.ref_time_march.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int32, C.c_double, _i32p, _f64p]
        L.ref_hllc.argtypes = [_f64p, _f64p, _f64p, C.c_double, _f64p]
        L.ref_flux.argtypes = [_f64p, C.c_double, _f64p, _f64p]
        L.ref_eigenmatrix.argtypes = [_f64p, C.c_double, C.c_double, C.c_double, _f64p, _f64p]
        L.ref_rk3_linear.argtypes = [_f64p, C.c_int64, C.c_double, C.c_double]
        _libs[plant] = L
    return _libs[plant]


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def gmsh_parse(path: str):
    """GmshParser::parse restated in C++; returns the MeshInput arrays."""
    from dgm_in_rust_amd.meshgen import MeshInput
    L = lib()
    g = L.ref_gmsh_parse(path.encode())
    if not g:
        raise FileNotFoundError(path)
    s = [C.c_int64() for _ in range(5)]
    L.ref_gmsh_sizes(g, *[C.byref(x) for x in s])
    nv, neb, nen, nelb, nel = [x.value for x in s]
    vxy = np.zeros((nv, 2)); et = np.zeros(neb, np.int32); ec = np.zeros(neb, np.int64)
    en = np.zeros((nen, 2), np.int64); lc = np.zeros(nelb, np.int64); ln = np.zeros((nel, 3), np.int64)
    L.ref_gmsh_get(g, _p(vxy, _f64p), _p(et, _i32p), _p(ec, _i64p), _p(en, _i64p), _p(lc, _i64p), _p(ln, _i64p))
    L.ref_gmsh_free(g)
    return MeshInput(vxy, et, ec, en, lc, ln, None, os.path.basename(path))


class RefCase:
    """One oracle case: mesh + parameters, with the reference's solver operations."""

    def __init__(self, mesh, order=2, n_cell_gp=0, n_edge_gp=0, hcr=1.4, cfl=0.0, mach=0.8, aoa_deg=1.5, rho=1.0,
                 p=1.0e5, fs_hcr=1.4, limiter=1, nthreads=1, gemm_fma=1, plant=0):
        L = lib(plant)
        m = mesh.contiguous()
        self.mesh = m
        self.prm = RefParams(order, n_cell_gp, n_edge_gp, hcr, cfl, mach, aoa_deg, rho, p, fs_hcr, limiter, nthreads,
                             gemm_fma)
        st = C.c_int32(0)
        self.h = L.ref_create(m.vxy.shape[0], _p(m.vxy, _f64p), len(m.eb_type), _p(m.eb_type, _i32p),
                              _p(m.eb_count, _i64p), _p(m.eb_nodes, _i64p), len(m.el_count), _p(m.el_count, _i64p),
                              _p(m.el_nodes, _i64p), _p(m.elem_xy, _f64p), C.byref(self.prm), C.byref(st))
        if not self.h:
            raise ValueError(f"ref_create failed with status {st.value}")
        s64 = [C.c_int64() for _ in range(5)]
        s32 = [C.c_int32() for _ in range(3)]
        L.ref_sizes(self.h, *[C.byref(x) for x in s64 + s32])
        (self.ne, self.nedge, self.n_internal_edges, self.n_internal_elements,
         self.n_boundary_elements) = [x.value for x in s64]
        self.nbasis, self.ncgp, self.negp = [x.value for x in s32]
        self.order = order
        self.hcr = hcr
        self._L = L

    def __del__(self):
        if getattr(self, "h", None):
            self._L.ref_destroy(self.h)
            self.h = None

    @property
    def shape(self):
        return (self.ne, 4, self.nbasis)

    def zeros(self):
        return np.zeros(self.shape)

    def cfl(self):
        return self._L.ref_cfl(self.h)

    def edges(self):
        n = self.nedge
        ie = np.zeros((n, 2), np.int64); iv = np.zeros((n, 2), np.int64); ic = np.zeros((n, 2), np.int64)
        ip = np.zeros(n, np.int64); nm = np.zeros((n, 2)); jd = np.zeros(n)
        self._L.ref_get_edges(self.h, _p(ie, _i64p), _p(iv, _i64p), _p(ic, _i64p), _p(ip, _i64p), _p(nm, _f64p),
                              _p(jd, _f64p))
        return dict(ielements=ie, ivertices=iv, in_cell_indices=ic, ipatch=ip, normal=nm, jacob_det=jd)

    def elements(self):
        n = self.ne
        iv = np.zeros((n, 3), np.int64); ie = np.zeros((n, 3), np.int64); nb = np.zeros((n, 3), np.int64)
        jd = np.zeros(n); cr = np.zeros(n); mh = np.zeros(n)
        self._L.ref_get_elements(self.h, _p(iv, _i64p), _p(ie, _i64p), _p(nb, _i64p), _p(jd, _f64p), _p(cr, _f64p),
                                 _p(mh, _f64p))
        return dict(ivertices=iv, iedges=ie, ineighbours=nb, jacob_det=jd, circumradius=cr, minimum_height=mh)

    def lists(self):
        a = np.zeros(self.n_internal_edges, np.int64)
        b = np.zeros(self.n_internal_elements, np.int64)
        c = np.zeros(self.n_boundary_elements, np.int64)
        self._L.ref_get_lists(self.h, _p(a, _i64p), _p(b, _i64p), _p(c, _i64p))
        return dict(internal_edges=a, internal_elements=b, boundary_elements=c)

    def tables(self):
        nq, ng, np_ = self.ncgp, self.negp, self.nbasis
        nk = self._L.ref_n_deriv_keys(self.h)
        t = dict(phi_cell=np.zeros((nq, np_)), phi_edge=np.zeros((3, ng, np_)), mass=np.zeros(np_),
                 cell_pts=np.zeros((nq, 2)), cell_w=np.zeros(nq), edge_pts=np.zeros((3, ng, 2)), edge_w=np.zeros(ng),
                 dref=np.zeros((nk, nq, np_)))
        self._L.ref_get_tables(self.h, *[_p(t[k], _f64p) for k in ("phi_cell", "phi_edge", "mass", "cell_pts",
                                                                      "cell_w", "edge_pts", "edge_w", "dref")])
        return t

    def phys_deriv(self, e, igp, k, l1, l2):
        return self._L.ref_phys_deriv(self.h, e, igp, k, l1, l2)

    def initial_solution(self, density=1.0, aoa_deg=1.5, mach=0.8, pressure=1.0e5):
        U = self.zeros()
        self._L.ref_initial_solution(self.h, density, aoa_deg, mach, pressure, _p(U, _f64p))
        return U

    def cell_points_xy(self):
        xy = np.zeros((self.ne, self.ncgp, 2))
        self._L.ref_cell_points_xy(self.h, _p(xy, _f64p))
        return xy

    def project(self, q_at_points):
        q = np.ascontiguousarray(q_at_points, np.float64)
        U = self.zeros()
        self._L.ref_project(self.h, _p(q, _f64p), _p(U, _f64p))
        return U

    def compute_residuals(self, U, R=None):
        """Reference semantics: accumulate into R (zero if None), then divide by the mass."""
        U = np.ascontiguousarray(U, np.float64)
        R = self.zeros() if R is None else np.ascontiguousarray(R, np.float64).copy()
        bad = C.c_int64(-1)
        s = self._L.ref_compute_residuals(self.h, _p(R, _f64p), _p(U, _f64p), C.byref(bad))
        return R, s, bad.value

    def terms(self, U):
        U = np.ascontiguousarray(U, np.float64)
        V = self.zeros(); E = self.zeros(); B = self.zeros()
        self._L.ref_volume_terms(self.h, _p(V, _f64p), _p(U, _f64p))
        bad = C.c_int64(-1)
        s = self._L.ref_edge_terms(self.h, _p(E, _f64p), _p(U, _f64p), C.byref(bad))
        self._L.ref_boundary_terms(self.h, _p(B, _f64p), _p(U, _f64p))
        return V, E, B, s

    def divide_by_mass(self, R):
        R = np.ascontiguousarray(R, np.float64).copy()
        self._L.ref_divide_by_mass(self.h, _p(R, _f64p))
        return R

    def apply_limiter(self, U):
        U = np.ascontiguousarray(U, np.float64).copy()
        tr = np.zeros(self.ne, np.uint8)
        s = self._L.ref_apply_limiter(self.h, _p(U, _f64p), _p(tr, _u8p))
        return U, tr.astype(bool), s

    def kxrcf(self, U, e):
        U = np.ascontiguousarray(U, np.float64)
        return bool(self._L.ref_kxrcf(self.h, _p(U, _f64p), e))

    def weno_element(self, U, e):
        """[4, nb] WENO reconstruction of element e from U (status 0 on success)."""
        U = np.ascontiguousarray(U, np.float64)
        out = np.zeros((4, self.nbasis))
        s = self._L.ref_weno_element(self.h, _p(U, _f64p), e, _p(out, _f64p))
        return out, s

    def compute_time_step(self, U):
        U = np.ascontiguousarray(U, np.float64)
        return self._L.ref_compute_time_step(self.h, _p(U, _f64p))

    def rk3(self, U, dt):
        U = np.ascontiguousarray(U, np.float64).copy()
        R = self.zeros()
        s = self._L.ref_rk3(self.h, _p(R, _f64p), _p(U, _f64p), dt)
        return U, s

    def last_bad_index(self):
        """Edge at which the last failing residual stopped (the first failing interior edge in list order)."""
        return int(self._L.ref_last_bad_index(self.h))

    def time_march(self, U, final_step, final_time):
        U = np.ascontiguousarray(U, np.float64).copy()
        R = self.zeros()
        n = C.c_int32(0); t = C.c_double(0.0)
        s = self._L.ref_time_march(self.h, _p(R, _f64p), _p(U, _f64p), final_step, final_time, C.byref(n), C.byref(t))
        return U, s, n.value, t.value


def hllc(ql, qr, n, hcr=1.4, plant=0):
    L = lib(plant)
    ql = np.ascontiguousarray(ql, np.float64); qr = np.ascontiguousarray(qr, np.float64)
    n = np.ascontiguousarray(n, np.float64); f = np.zeros(4)
    s = L.ref_hllc(_p(ql, _f64p), _p(qr, _f64p), _p(n, _f64p), hcr, _p(f, _f64p))
    return f, s


def flux(q, hcr=1.4, plant=0):
    L = lib(plant)
    q = np.ascontiguousarray(q, np.float64); f = np.zeros(4); g = np.zeros(4)
    L.ref_flux(_p(q, _f64p), hcr, _p(f, _f64p), _p(g, _f64p))
    return f, g


def eigenmatrix(q, nx, ny, hcr=1.4, plant=0):
    L = lib(plant)
    q = np.ascontiguousarray(q, np.float64); Lm = np.zeros(16); Rm = np.zeros(16)
    L.ref_eigenmatrix(_p(q, _f64p), nx, ny, hcr, _p(Lm, _f64p), _p(Rm, _f64p))
    return Lm.reshape(4, 4), Rm.reshape(4, 4)


def rk3_linear(u, lam, dt, plant=0):
    L = lib(plant)
    u = np.ascontiguousarray(u, np.float64).copy()
    L.ref_rk3_linear(_p(u, _f64p), u.size, lam, dt)
    return u
