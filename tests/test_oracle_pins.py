"""Closed-form pins of the oracle's quirk-bearing steps (CPU).

Each test builds a special case whose answer follows from the mathematics alone (a hand-made mesh, a
state chosen so that the step reduces to a formula), checks the oracle against it, and then checks that
the oracle variant with the matching planted mistake (oracle/Makefile `plants`, -DORACLE_PLANT=n)
fails the same check. The quirks and the reference lines they follow:

* time step from the sum of all modes (src/temporal_disc.rs:63-79)
* HLLC p* numerator c_L + c_R - ((zeta - 1)/2)(u_R - u_L) (src/spatial_disc/flux.rs:75-78)
* KXRCF: jumps over inflow faces only, scaled by h^1.5, surface of the inflow faces, the neighbour's
  trace on the target's own local edge in reversed point order, 0/0 -> untroubled
  (src/spatial_disc/limiter/weno.rs:14-84)
* WENO: linear weights where a boundary face i zeroes weights[i] (weno.rs:93-110), neighbour
  candidates with their mode 0 replaced by the target's (weno.rs:148-155) used as coefficient vectors in
  the target's basis, and the neighbour-area-weighted mean of the per-face reconstructions
  (weno.rs:111-118,196,260-263)
"""
import math

import numpy as np
import pytest

from dgm_in_rust_amd.meshgen import FARFIELD, MeshInput
from oracle import refcpu
from oracle.refcpu import RefCase
from tests import exact_dubiner as xd

G = 1.4
# planted-mistake variants (ORACLE_PLANT values in oracle/refcpu.cpp)
P_DT_MODE0, P_HLLC_PSTAR, P_KX_NB_EDGE, P_KX_H1, P_KX_OUTFLOW, P_KX_SURFACE, P_W_NEXT, P_W_NOSWAP, P_W_PLAIN = range(1, 10)


@pytest.fixture(scope="module", autouse=True)
def plants():
    refcpu.build(plants=True)


# ---- a hand-made mesh: target triangle T0 = (O, A, B) with legs on the axes and its three neighbours
O, A, B = (0.0, 0.0), (1.0, 0.0), (0.0, 1.0)
C, D, E = (2.0, 2.0), (0.5, -1.5), (-2.0, 0.3)
VXY = np.array([O, A, B, C, D, E])
iO, iA, iB, iC, iD, iE = range(6)
# neighbours: T1 across the hypotenuse AB (area 1.5), T2 across OA (area 0.75), T3 across BO (area 1.0)
T1, T2, T3 = (iA, iC, iB), (iO, iD, iA), (iO, iB, iE)
AREA = {1: 1.5, 2: 0.75, 3: 1.0}


def four_mesh(rot=0, drop_t3=False):
    """T0 with its vertex order rotated by rot (which moves its local edge numbering), then T1, T2 and
    (unless dropped) T3; every outer edge is far field, oriented along its triangle."""
    t0 = [iO, iA, iB]
    t0 = t0[rot:] + t0[:rot]
    tris = [t0, list(T1), list(T2)] + ([] if drop_t3 else [list(T3)])
    bnd = [(iA, iC), (iC, iB), (iO, iD), (iD, iA)]
    bnd += [(iB, iO)] if drop_t3 else [(iB, iE), (iE, iO)]
    return MeshInput(vxy=VXY.copy(), eb_type=np.array([FARFIELD], np.int32), eb_count=np.array([len(bnd)], np.int64),
                     eb_nodes=np.array(bnd, np.int64), el_count=np.array([len(tris)], np.int64),
                     el_nodes=np.array(tris, np.int64), name="four")


def local_edge(ref, e, a, b):
    """Local index of the edge {a, b} of element e, from the oracle's element vertex list."""
    iv = list(ref.elements()["ivertices"][e])
    for i in range(3):
        if {iv[i], iv[(i + 1) % 3]} == {a, b}:
            return i
    raise AssertionError("edge not in element")


def ref_mid(lic):
    """Midpoint of local edge lic on the reference triangle: e0 (0,0)->(1,0), e1 (1,0)->(0,1), e2 (0,1)->(0,0)."""
    return [(0.5, 0.0), (0.5, 0.5), (0.0, 0.5)][lic]


def mode_index(order, l, k):
    return xd.modes(order).index((l, k))


def uniform(ref, rho, u, v, p):
    U = ref.zeros()
    U[:, 0, 0] = rho
    U[:, 1, 0] = rho * u
    U[:, 2, 0] = rho * v
    U[:, 3, 0] = p / (G - 1.0) + 0.5 * rho * (u * u + v * v)
    return U


# ---- time step (Q5) ----------------------------------------------------------------------------

def one_triangle(xy):
    return MeshInput(vxy=np.array(xy, float), eb_type=np.array([FARFIELD], np.int32), eb_count=np.array([3], np.int64),
                     eb_nodes=np.array([(0, 1), (1, 2), (2, 0)], np.int64), el_count=np.array([1], np.int64),
                     el_nodes=np.array([(0, 1, 2)], np.int64), name="one")


def dt_closed_form(xy, s, order):
    """CFL h_min / (|V| + c) with the state s given by the sum of the modes; h_min is the smallest
    altitude (2 area / longest edge); CFL = 1 / ((2p + 1)(1 + 4 / (p + 2)^2))."""
    (x1, y1), (x2, y2), (x3, y3) = xy
    area = 0.5 * abs((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1))
    longest = max(math.dist(xy[0], xy[1]), math.dist(xy[1], xy[2]), math.dist(xy[2], xy[0]))
    hmin = 2.0 * area / longest
    cfl = 1.0 / ((2 * order + 1) * (1.0 + 4.0 / (order + 2) ** 2))
    u, v = s[1] / s[0], s[2] / s[0]
    p = (G - 1.0) * (s[3] - 0.5 * s[0] * (u * u + v * v))
    return cfl * hmin / (math.hypot(u, v) + math.sqrt(G * p / s[0]))


@pytest.mark.parametrize("order", [1, 2, 3])
def test_time_step_uses_the_sum_of_all_modes(order):
    xy = [(0.0, 0.0), (2.0, 0.3), (0.4, 1.1)]
    U0 = np.zeros((1, 4, (order + 1) * (order + 2) // 2))
    U0[0, :, 0] = [1.0, 0.1, 0.05, 2.5]
    U0[0, :, 1] = [0.2, 0.3, -0.1, 0.4]
    U0[0, :, -1] = [-0.05, 0.02, 0.2, 0.1]
    s = U0[0].sum(axis=1)
    want = dt_closed_form(xy, s, order)
    got = RefCase(one_triangle(xy), order=order, limiter=0).compute_time_step(U0)
    assert abs(got - want) <= 1e-14 * want
    bad = RefCase(one_triangle(xy), order=order, limiter=0, plant=P_DT_MODE0).compute_time_step(U0)
    assert abs(bad - want) > 1e-3 * want


# ---- HLLC p* (Q6) ------------------------------------------------------------------------------

def symmetric_hllc_normal_flux(rho, U, p):
    """Normal momentum flux of HLLC for the symmetric pair (rho, U, p) | (rho, -U, p) across n = (1, 0).
    By symmetry S* = 0, the mass and energy fluxes vanish and F = p + rho U c q_L with
    q_L = sqrt(1 + (gamma + 1)/(2 gamma) (p*/p - 1)) when p* > p (else 1), where the reference's p* is
    p (1 - ((zeta - 1)/2) (u_R - u_L) / (2c))^(1/zeta) = p (1 + (zeta - 1) U / (2c))^(1/zeta)."""
    c = math.sqrt(G * p / rho)
    zeta = (G - 1.0) / (2.0 * G)
    ratio = (1.0 + (zeta - 1.0) * U / (2.0 * c)) ** (1.0 / zeta)
    q = 1.0 if ratio <= 1.0 else math.sqrt(1.0 + (G + 1.0) / (2.0 * G) * (ratio - 1.0))
    return p + rho * U * c * q


@pytest.mark.parametrize("mach", [0.5, -0.5, 0.2, -0.8])
def test_hllc_symmetric_pair_closed_form(mach):
    rho, p = 1.3, 1.0e5
    U = mach * math.sqrt(G * p / rho)
    E = p / (G - 1.0) + 0.5 * rho * U * U
    ql, qr = np.array([rho, rho * U, 0.0, E]), np.array([rho, -rho * U, 0.0, E])
    f, s = refcpu.hllc(ql, qr, np.array([1.0, 0.0]))
    want = symmetric_hllc_normal_flux(rho, U, p)
    scale = p + rho * abs(U) * math.sqrt(G * p / rho)
    assert s == 0
    assert abs(f[1] - want) <= 1e-13 * scale
    assert abs(f[0]) <= 1e-12 * rho * abs(U) and abs(f[2]) <= 1e-9 and abs(f[3]) <= 1e-12 * E * abs(U)
    fb, _ = refcpu.hllc(ql, qr, np.array([1.0, 0.0]), plant=P_HLLC_PSTAR)
    assert abs(fb[1] - want) > 1e-3 * rho * abs(U) * math.sqrt(G * p / rho)


# ---- KXRCF (Q7, Q8) ----------------------------------------------------------------------------

def kx_threshold_check(make, a_star, rel=1e-9):
    """The detector flips from untroubled to troubled where the amplitude a crosses a_star."""
    below, above = make(a_star * (1.0 - rel)), make(a_star * (1.0 + rel))
    return (not below) and above


def kxrcf_piecewise_constant(plant=0):
    """Target T0 with rho = 1, velocity (0.8, 0.3) c: inflow faces are OA (outward normal (0, -1)) and
    BO ((-1, 0)); AB ((1, 1)/sqrt 2) is outflow. Neighbour densities 1 + a s_f with s = 1 (OA), 3 (BO),
    -0.5 (AB); total energy equal everywhere, so only the density indicator moves:
    ind = a |sum_inflow |e| s_f| / (h^1.5 sum_inflow |e|), h = circumradius = sqrt 2 / 2."""
    mesh = four_mesh()
    ref = RefCase(mesh, order=2, limiter=1, plant=plant)
    c = math.sqrt(G)
    base = uniform(ref, 1.0, 0.8 * c, 0.3 * c, 1.0)
    E0 = base[0, 3, 0]

    def detect(a):
        U = base.copy()
        for el, s in ((2, 1.0), (3, 3.0), (1, -0.5)):
            U[el, 0, 0] = 1.0 + a * s
        U[:, 3, 0] = E0
        return ref.kxrcf(U, 0)
    return detect


def test_kxrcf_piecewise_constant_threshold():
    h15 = (math.sqrt(2.0) / 2.0) ** 1.5
    a_star = h15 * (1.0 + 1.0) / abs(1.0 * 1.0 + 1.0 * 3.0)
    assert kx_threshold_check(kxrcf_piecewise_constant(), a_star)
    for plant in (P_KX_H1, P_KX_OUTFLOW, P_KX_SURFACE):
        assert not kx_threshold_check(kxrcf_piecewise_constant(plant), a_star), plant


def test_kxrcf_no_inflow_face_is_untroubled():
    """Zero velocity: no face is inflow, the indicator is 0/0 = NaN and NaN >= 1 is false, however large
    the neighbours' jumps."""
    ref = RefCase(four_mesh(), order=2, limiter=1)
    U = uniform(ref, 1.0, 0.0, 0.0, 1.0)
    U[1:, 0, 0] = [50.0, 0.01, 9.0]
    U[1:, 3, 0] = [1e3, 1e-3, 7.0]
    assert not ref.kxrcf(U, 0)


@pytest.mark.parametrize("rot", [0, 1, 2])
@pytest.mark.parametrize("lk", [(1, 0), (0, 1)])
def test_kxrcf_neighbour_trace_on_target_edge(rot, lk):
    """Only BO is inflow (velocity (0.6, -0.3) c). Its neighbour T3 carries a linear density mode
    rho = 1 + a psi. The reference evaluates the neighbour's trace with the TARGET's local edge index
    (reversed points): the jump is -|e| a psi(mid of the target's local edge), exact for a linear mode
    under Gauss-Legendre, so ind = a |psi(mid_t)| / h^1.5."""
    order = 2
    mesh = four_mesh(rot=rot)
    ref = RefCase(mesh, order=order, limiter=1)
    bad = RefCase(mesh, order=order, limiter=1, plant=P_KX_NB_EDGE)
    lic_t = local_edge(ref, 0, iB, iO)
    lic_n = local_edge(ref, 3, iB, iO)
    k = mode_index(order, *lk)
    c = math.sqrt(G)
    base = uniform(ref, 1.0, 0.6 * c, -0.3 * c, 1.0)

    def make(r):
        def detect(a):
            U = base.copy()
            U[3, 0, k] = a
            return r.kxrcf(U, 0)
        return detect
    psi_t = xd.eval_deriv(*lk, 0, 0, *ref_mid(lic_t))
    psi_n = xd.eval_deriv(*lk, 0, 0, *ref_mid(lic_n))
    h15 = (math.sqrt(2.0) / 2.0) ** 1.5
    if psi_t == 0.0:
        assert not make(ref)(1e6)
    else:
        assert kx_threshold_check(make(ref), h15 / abs(psi_t))
    if abs(abs(psi_n) - abs(psi_t)) > 1e-12:
        if psi_t == 0.0:
            assert make(bad)(1e6)
        else:
            assert not kx_threshold_check(make(bad), h15 / abs(psi_t))


# ---- WENO (Q9, Q10, Q12) -----------------------------------------------------------------------
#
# At zero mean velocity the characteristic rows of a face with normal n see a momentum perturbation
# (a, b) phi as (-(n.(a,b))/(2c), -n_y a + n_x b, 0, (n.(a,b))/(2c)) phi. Candidates whose non-constant
# part is a pure rho-u or rho-v perturbation of one linear mode phi then have:
# * on the hypotenuse AB (n = (1,1)/sqrt 2) exactly equal smoothness in every field, so the nonlinear
#   weights reduce to the normalised linear weights and the face's reconstruction is
#   m + sum_j gamma_j d_j;
# * on an axis-aligned face, a smoothness of exactly 0 in each field for the candidates that do not
#   see it; those take all the weight (eps^2 against beta^2 ~ 1e-12), so the face's reconstruction
#   is m, provided their linear weights are not all zero.
# The final reconstruction is the neighbour-area-weighted mean of the per-face ones.

AMP = 1.0e-3
RU, RV = 1, 2


def weno_state(ref, vecs, order=2, k=1):
    """Mean m = (1, 0, 0, 2.5) in mode 0 of the target; candidate j's non-constant part is AMP * phi_k in
    variable vecs[j] (0: target, 1..3: T1..T3); neighbours carry unrelated means, which the reference
    replaces by the target's."""
    U = ref.zeros()
    U[0, 0, 0], U[0, 3, 0] = 1.0, 2.5
    for el, var in vecs.items():
        if el > 0:
            U[el, :, 0] = [1.7 + 0.1 * el, 0.3 * el, -0.2, 3.1 + el]
        U[el, var, k] = AMP
    return U


def expected(ref, U, face_weights, k=1):
    """m + sum over the per-face perturbations (area-weighted), with face_weights = {element across the
    face: {candidate element: coefficient}} for faces whose reconstruction is not m."""
    out = np.zeros_like(U[0])
    out[:, 0] = U[0, :, 0]
    total = sum(AREA[el] for el in face_weights["areas"])
    for face_el, coeffs in face_weights["faces"].items():
        for cand, w in coeffs.items():
            var = face_weights["vecs"][cand]
            out[var, k] += AREA[face_el] / total * w * AMP
    return out


def weno_err(ref, U, exp):
    got, s = ref.weno_element(U, 0)
    assert s == 0
    return np.abs(got - exp).max() / AMP


def test_weno_interior_element_area_weighting_and_mean_swap():
    vecs = {0: RU, 1: RV, 2: RU, 3: RV}
    spec = {"vecs": vecs, "areas": [1, 2, 3], "faces": {1: {0: 0.997, 1: 0.001, 2: 0.001, 3: 0.001}}}
    for plant, ok in ((0, True), (P_W_PLAIN, False), (P_W_NOSWAP, False)):
        ref = RefCase(four_mesh(), order=2, limiter=1, plant=plant)
        U = weno_state(ref, vecs)
        err = weno_err(ref, U, expected(ref, U, spec))
        assert (err <= 1e-9) == ok, (plant, err)


@pytest.mark.parametrize("rot", [0, 1, 2])
def test_weno_boundary_element_linear_weights(rot):
    """T3 removed: face BO is a boundary face, at local index 2, 1, 0 for rot = 0, 1, 2. weights =
    [0.997, 0.001, 0.001, 0.001], weights[i] = 0 for the boundary face i, then weights[1] = 0.998.
    Candidate 0 is the target; the neighbour across local face f takes weights[f + 1]."""
    mesh = four_mesh(rot=rot, drop_t3=True)
    ref = RefCase(mesh, order=2, limiter=1)
    i_bnd = local_edge(ref, 0, iB, iO)
    lw = [0.997, 0.001, 0.001, 0.001]
    lw[i_bnd] = 0.0
    lw[1] = 0.998
    f_ab, f_oa = local_edge(ref, 0, iA, iB), local_edge(ref, 0, iO, iA)
    g = {0: lw[0], 1: lw[f_ab + 1], 2: lw[f_oa + 1]}      # target, T1 (across AB), T2 (across OA)
    vecs = {0: RU, 1: RV, 2: RU}
    tot = sum(g.values())
    faces = {1: {j: g[j] / tot for j in g}}
    # face OA (n = (0, -1)): field 1 is smooth for the rho-v candidate T1, fields 0 and 3 for the rho-u
    # candidates; with T1's weight 0 (boundary at local 2) field 1 falls to the rho-u candidates, whose
    # values agree, so the face keeps the rho-u perturbation whole
    if g[1] == 0.0:
        faces[2] = {0: 1.0}
    spec = {"vecs": vecs, "areas": [1, 2], "faces": faces}
    U = weno_state(ref, vecs)
    assert weno_err(ref, U, expected(ref, U, spec)) <= 1e-9
    # a boundary face at local index 1 zeroes weights[1], which the count then resets to 0.998: there the
    # quirk and the planted placement (weights[i + 1], the missing candidate's) give the same weights
    if i_bnd != 1:
        badref = RefCase(mesh, order=2, limiter=1, plant=P_W_NEXT)
        assert weno_err(badref, U, expected(ref, U, spec)) > 1e-3
