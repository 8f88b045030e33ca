"""Quadrature and Dubiner-basis tables: oracle pinned by exact rational arithmetic and closed forms;
product tables (libdgm.so host-side set-up, no GPU needed) checked against the oracle."""
from math import factorial

import numpy as np
import pytest

from dgm_in_rust_amd import Context, meshgen
from oracle.refcpu import RefCase
from tests import exact_dubiner as X

RULES = [(1, 6, 3), (2, 6, 3), (2, 4, 3), (2, 6, 5), (3, 16, 5), (4, 25, 5)]
# The reference's active 4-point rule (src/gauss_point.rs:67-68) uses the points 2/15, 11/15 with the
# Strang-Fix weights -27/96, 25/96 (sum 0.5): it integrates over the triangle's area (not normalised)
# and is exact to degree 1 only. Only the 6-point rule is consistent with the 0.5 J scaling.
DEGREE = {4: 1, 6: 4, 16: 8, 25: 9}


@pytest.fixture(scope="module")
def tiny_mesh():
    return meshgen.box_farfield(2, seed=0)


def oracle_tables(mesh, p, nq, ne):
    return RefCase(mesh, order=p, n_cell_gp=nq, n_edge_gp=ne).tables()


def exact_monomial(a, b):
    # (1/|T|) int_T xi^a eta^b over the reference triangle, |T| = 1/2
    return 2.0 * factorial(a) * factorial(b) / factorial(a + b + 2)


@pytest.mark.parametrize("p,nq,ne", RULES)
def test_triangle_rule_exactness(tiny_mesh, p, nq, ne):
    t = oracle_tables(tiny_mesh, p, nq, ne)
    w, pts = t["cell_w"], t["cell_pts"]
    deg = DEGREE[nq]
    # the reference's 6-point literals carry 14 digits (weights sum to 0.99999999999999)
    tol = 2e-14 if nq == 6 else 1e-15
    scale = 0.5 if nq == 4 else 1.0  # src/gauss_point.rs:68: the 4-point weights sum to 0.5 (a quirk of the reference's rule)
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            got = float(np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b))
            assert abs(got - scale * exact_monomial(a, b)) <= tol * 10, (a, b, got)
    # one degree higher is not integrated exactly (the rule's degree is what it claims)
    misses = [abs(float(np.sum(w * pts[:, 0] ** a * pts[:, 1] ** (deg + 1 - a))) - scale * exact_monomial(a, deg + 1 - a))
              for a in range(deg + 2)]
    assert max(misses) > 1e-8


@pytest.mark.parametrize("n", [3, 4, 5])
def test_gauss_legendre_exactness(tiny_mesh, n):
    t = oracle_tables(tiny_mesh, 2, 6, n)
    w = t["edge_w"]
    # recover the interval nodes from local edge 0: xi = (x + 1) / 2
    x = 2.0 * t["edge_pts"][0, :, 0] - 1.0
    for d in range(2 * n):
        exact = 0.0 if d % 2 else 2.0 / (d + 1)
        assert abs(float(np.sum(w * x ** d)) - exact) <= 2e-15
    assert abs(float(np.sum(w * x ** (2 * n))) - 2.0 / (2 * n + 1)) > 1e-6


@pytest.mark.parametrize("p,nq,ne", RULES)
def test_edge_point_parametrisation(tiny_mesh, p, nq, ne):
    """local edges e0 (0,0)->(1,0), e1 (1,0)->(0,1), e2 (0,1)->(0,0), src/gauss_point.rs:17-27"""
    t = oracle_tables(tiny_mesh, p, nq, ne)
    ep = t["edge_pts"]
    x = 2.0 * ep[0, :, 0] - 1.0
    assert np.allclose(ep[0, :, 1], 0.0)
    close = lambda a, b: np.allclose(a, b, rtol=0, atol=2e-16)
    assert close(ep[1, :, 0], (1.0 - x) / 2.0) and close(ep[1, :, 1], (1.0 + x) / 2.0)
    assert np.allclose(ep[2, :, 0], 0.0) and close(ep[2, :, 1], (x[::-1] + 1.0) / 2.0)
    # each edge runs from its first vertex to its second in point order
    assert ep[0, 0, 0] < ep[0, -1, 0] and ep[1, 0, 1] < ep[1, -1, 1] and ep[2, 0, 1] > ep[2, -1, 1]


@pytest.mark.parametrize("p,nq,ne", RULES)
def test_basis_values_exact(tiny_mesh, p, nq, ne):
    t = oracle_tables(tiny_mesh, p, nq, ne)
    md = X.modes(p)
    for i, (xi, eta) in enumerate(t["cell_pts"]):
        for b, (l, k) in enumerate(md):
            assert abs(t["phi_cell"][i, b] - X.eval_deriv(l, k, 0, 0, xi, eta)) <= 1e-13 * max(1.0, abs(t["phi_cell"][i, b]))
    for e in range(3):
        for i, (xi, eta) in enumerate(t["edge_pts"][e]):
            for b, (l, k) in enumerate(md):
                assert abs(t["phi_edge"][e, i, b] - X.eval_deriv(l, k, 0, 0, xi, eta)) <= 1e-13 * max(1.0, abs(t["phi_edge"][e, i, b]))


@pytest.mark.parametrize("p,nq,ne", RULES)
def test_basis_derivatives_exact(tiny_mesh, p, nq, ne):
    """p = 2: the reference's hard-coded tables (src/basis_function.rs:114-193); p != 2: the extension."""
    t = oracle_tables(tiny_mesh, p, nq, ne)
    md = X.modes(p)
    maxd = max(p, 2)
    keys = [(a, d - a) for d in range(1, maxd + 1) for a in range(d, -1, -1)]
    assert t["dref"].shape[0] == len(keys)
    for kk, (a, b) in enumerate(keys):
        ex = np.array([[X.eval_deriv(l, k, a, b, xi, eta) for (l, k) in md] for (xi, eta) in t["cell_pts"]])
        if p == 2:  # the reference's literal tables reproduce the analytic derivatives to rounding
            assert np.abs(t["dref"][kk] - ex).max() <= 1e-14 * max(1.0, np.abs(ex).max()), (a, b)
        else:       # extension: monomial expansion in double; error relative to the table's scale
            assert np.abs(t["dref"][kk] - ex).max() <= 1e-13 * max(1.0, np.abs(ex).max()), (a, b)


@pytest.mark.parametrize("p,nq,ne", [(1, 6, 3), (2, 6, 3), (3, 16, 5), (4, 25, 5)])
def test_mass_and_orthogonality(tiny_mesh, p, nq, ne):
    t = oracle_tables(tiny_mesh, p, nq, ne)
    md = X.modes(p)
    closed = np.array([X.mass_closed_form(l, k) for (l, k) in md])
    # The 6-point literals carry 14 digits. The 16/25-point rules integrate monomials to 2e-16 (checked
    # in exact arithmetic); phi^2 has monomial coefficients in the hundreds, so the rule's own double
    # representation leaves up to ~1e-13 relative error in the mass (measured 8.5e-14 for mode 7 at p = 3).
    assert np.allclose(t["mass"], closed, rtol=2e-13, atol=0), (t["mass"] - closed)
    if DEGREE[nq] >= 2 * p:  # the rule integrates phi_i phi_j exactly: the basis is orthogonal
        G = (t["phi_cell"] * t["cell_w"][:, None]).T @ t["phi_cell"]
        off = G - np.diag(np.diag(G))
        assert np.abs(off).max() <= 1e-12


def test_reference_p2_mode_order(tiny_mesh):
    """(l,k) = (0,0),(0,1),(1,0),(0,2),(1,1),(2,0) — src/basis_function.rs:86-92"""
    assert X.modes(2) == [(0, 0), (0, 1), (1, 0), (0, 2), (1, 1), (2, 0)]
    t = oracle_tables(tiny_mesh, 2, 6, 3)
    # hard-coded first derivatives of mode 4 (l=1,k=1): d/dxi = 20 eta - 4
    for i, (xi, eta) in enumerate(t["cell_pts"]):
        assert t["dref"][0, i, 4] == 20.0 * eta - 4.0


@pytest.mark.parametrize("p,nq,ne", [(1, 6, 3), (2, 6, 3), (3, 16, 5), (4, 25, 5)])
def test_product_tables_match_oracle(tiny_mesh, p, nq, ne):
    ref = oracle_tables(tiny_mesh, p, nq, ne)
    ctx = Context(tiny_mesh, order=p, n_cell_gp=nq, n_edge_gp=ne, host_only=True)
    t = ctx.tables()
    for k in ("cell_pts", "edge_pts", "edge_w"):
        assert np.array_equal(t[k], ref[k]), k
    for k in ("cell_w", "phi_cell", "phi_edge", "mass"):
        assert np.allclose(t[k], ref[k], rtol=2e-15, atol=2e-15), k


def constant_flux_gap(t, md, fx, fy, cell_w_scale=1.0):
    """Volume minus surface term of a constant flux (fx, fy) against every basis function on the
    reference triangle, from the given quadrature and edge-basis tables: 0.5 sum_q w_q (fx dphi/dxi + fy
    dphi/deta) - sum_faces |e|/2 sum_i w_i phi(face, i) (F . n). The gradients are exact (fractions)."""
    faces = [((0.0, -1.0), 1.0), ((1.0 / np.sqrt(2.0), 1.0 / np.sqrt(2.0)), np.sqrt(2.0)), ((-1.0, 0.0), 1.0)]
    gap = np.zeros(len(md))
    scale = np.zeros(len(md))
    for b, (l, k) in enumerate(md):
        vol = 0.0
        for q, (xi, eta) in enumerate(t["cell_pts"]):
            w = t["cell_w"][q] * cell_w_scale
            vol += 0.5 * w * (fx * X.eval_deriv(l, k, 1, 0, xi, eta) + fy * X.eval_deriv(l, k, 0, 1, xi, eta))
        sur = 0.0
        for f, ((nx, ny), le) in enumerate(faces):
            for i in range(len(t["edge_w"])):
                sur += 0.5 * le * t["edge_w"][i] * t["phi_edge"][f, i, b] * (fx * nx + fy * ny)
        gap[b] = vol - sur
        scale[b] = abs(vol) + abs(sur) + 1.0
    return gap, scale


@pytest.mark.parametrize("p,nq,ne", [(1, 6, 3), (2, 6, 3), (3, 16, 5), (4, 25, 5)])
def test_constant_flux_volume_equals_surface(tiny_mesh, p, nq, ne):
    """The FP32-bulk element passes subtract each element's mean flux from its point fluxes (deviation
    form); that leaves the residual unchanged only if, for a constant flux, the volume term equals the
    surface term for every basis function, as the divergence theorem gives and as the quadrature rules
    reproduce when they integrate the gradients (degree p - 1) and the traces (degree p) exactly. Checked
    on the product's tables for every supported rule; the 6-point rule's 14-digit literals leave about
    1e-14. A rule whose weights sum to 1/2 (as the reference's 4-point rule's do) would break it: scaling
    the weights so fails the check."""
    ctx = Context(tiny_mesh, order=p, n_cell_gp=nq, n_edge_gp=ne, host_only=True)
    t = ctx.tables()
    md = X.modes(p)
    rng = np.random.default_rng(p)
    for _ in range(3):
        fx, fy = rng.uniform(-3.0, 3.0, size=2)
        gap, scale = constant_flux_gap(t, md, fx, fy)
        assert np.all(np.abs(gap) <= 1e-13 * scale), gap
        bad, bscale = constant_flux_gap(t, md, fx, fy, cell_w_scale=0.5)
        assert np.any(np.abs(bad) > 1e-3 * bscale), bad
