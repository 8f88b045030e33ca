"""Flow cases of the benchmark configurations and their error norms.

* Isentropic vortex (strength beta, centre (x0, y0), mean flow (u_inf, v_inf), rho = p = 1 at
  infinity) on the doubly periodic square: the exact solution is the initial field advected
  with the mean flow. Not part of the reference (which has no periodic boundary); it pins the
  whole scheme's accuracy.
* Uniform freestream (the reference's initial condition, src/solver.rs:45-79).
* The batched-sweep distribution of airfoil cases (:func:`sweep_cases`).

Norms are evaluated with the cell quadrature rule of the context: for a field q given at the
cell points, ||q||_2^2 = sum_e sum_q (J_e / 2) w_q q^2.
"""
from __future__ import annotations

import numpy as np


def vortex_primitive(x, y, t=0.0, beta=5.0, x0=5.0, y0=5.0, u_inf=1.0, v_inf=1.0, length=10.0, hcr=1.4):
    """Exact (rho, u, v, p) of the isentropic vortex at time t on a periodic square of side length."""
    xc = x0 + u_inf * t
    yc = y0 + v_inf * t
    # nearest periodic image of the centre
    dx = (x - xc + 0.5 * length) % length - 0.5 * length
    dy = (y - yc + 0.5 * length) % length - 0.5 * length
    r2 = dx * dx + dy * dy
    f = np.exp(0.5 * (1.0 - r2))
    du = -beta / (2.0 * np.pi) * f * dy
    dv = beta / (2.0 * np.pi) * f * dx
    T = 1.0 - (hcr - 1.0) * beta * beta / (8.0 * hcr * np.pi * np.pi) * f * f
    rho = T ** (1.0 / (hcr - 1.0))
    p = rho ** hcr
    return rho, u_inf + du, v_inf + dv, p


def conservative(rho, u, v, p, hcr=1.4):
    return np.stack([rho, rho * u, rho * v, p / (hcr - 1.0) + 0.5 * rho * (u * u + v * v)], axis=-1)


def vortex_state(ctx, t=0.0, **kw):
    """L2 projection of the exact vortex at time t onto the element modes of ``ctx``."""
    xy = ctx.cell_points_xy()
    q = conservative(*vortex_primitive(xy[..., 0], xy[..., 1], t=t, **kw), hcr=kw.get("hcr", 1.4))
    return ctx.project(q)


def point_values(ctx, U):
    """[ne, nq, 4] values of the modal state U at the cell points."""
    phi = ctx.tables()["phi_cell"]                 # [nq, nb]
    return np.einsum("evk,qk->eqv", U, phi)


def l2_norm_of_points(ctx, q):
    """Per-variable L2 norm of point values q [ne, nq, 4]."""
    w = ctx.tables()["cell_w"]
    J = ctx.geometry()["jacob_det"]
    return np.sqrt(np.einsum("e,q,eqv->v", 0.5 * J, w, q * q))


def vortex_error(ctx, U, t, **kw):
    """Per-variable L2 error of U against the exact vortex at time t, and the density error."""
    xy = ctx.cell_points_xy()
    exact = conservative(*vortex_primitive(xy[..., 0], xy[..., 1], t=t, **kw), hcr=kw.get("hcr", 1.4))
    return l2_norm_of_points(ctx, point_values(ctx, U) - exact)


def relative_l2(ctx, U, Uref):
    """||U - Uref||_2 / ||Uref||_2 over all variables, with the cell quadrature. Dominated by the total
    energy whenever it is the largest variable; see relative_l2_per_variable."""
    d = l2_norm_of_points(ctx, point_values(ctx, U - Uref))
    r = l2_norm_of_points(ctx, point_values(ctx, Uref))
    return float(np.sqrt((d * d).sum() / (r * r).sum()))


VARIABLES = ("rho", "rho_u", "rho_v", "E")


def relative_l2_per_variable(ctx, U, Uref, own_norm=False):
    """{variable: ||U_v - Uref_v||_2 / ||Uref_v||_2} for rho, rho u, rho v, E, with the cell quadrature. The
    momentum components are taken against the norm of the momentum vector (a component's own norm vanishes
    with the flow angle, the error it carries does not); own_norm=True divides each by its own norm."""
    d = l2_norm_of_points(ctx, point_values(ctx, U - Uref))
    r = l2_norm_of_points(ctx, point_values(ctx, Uref))
    if not own_norm:
        m = float(np.hypot(r[1], r[2]))
        r = np.array([r[0], m, m, r[3]])
    return {name: float(d[i] / r[i]) for i, name in enumerate(VARIABLES)}


def shocked_start(ctx, mach, aoa_deg, rho=1.0, p=1.0e5, hcr=1.4, x_shock=0.6, half_height=1.5, m_shock=1.3):
    """Freestream with a normal shock standing at x = x_shock for |y| < half_height, projected on the
    element modes: downstream of it the Rankine-Hugoniot state of a stream of Mach m_shock along x,
    scaled to the freestream density and pressure (rho2/rho1 = (g+1)M^2 / ((g-1)M^2 + 2),
    p2/p1 = 1 + 2g(M^2 - 1)/(g+1)), with the freestream's transverse velocity. The limiter then has a
    shock to capture from the first step, as on a transonic airfoil (the reference's freestream start
    develops one only over many flow-through times)."""
    xy = ctx.cell_points_xy()
    x, y = xy[..., 0], xy[..., 1]
    c = np.sqrt(hcr * p / rho)
    a = np.radians(aoa_deg)
    u = np.full_like(x, mach * c * np.cos(a))
    v = np.full_like(x, mach * c * np.sin(a))
    r = np.full_like(x, rho)
    pp = np.full_like(x, p)
    m2 = m_shock * m_shock
    rr = (hcr + 1.0) * m2 / ((hcr - 1.0) * m2 + 2.0)
    pr = 1.0 + 2.0 * hcr * (m2 - 1.0) / (hcr + 1.0)
    down = (x > x_shock) & (np.abs(y) < half_height)
    r = np.where(down, rho * rr, r)
    pp = np.where(down, p * pr, pp)
    u = np.where(down, u / rr, u)
    return ctx.project(conservative(r, u, v, pp, hcr))


def ogrid_shape(n_triangles, ratio):
    """(surface points, layers) of an O-grid with about n_triangles = 2 * ns * nl and ns / nl = ratio."""
    nl = max(8, int(round(np.sqrt(n_triangles / (2.0 * ratio)))))
    ns = max(16, int(round(n_triangles / (2.0 * nl))))
    return ns, nl


def sweep_cases(n, seed=0, tri_min=2000, tri_max=60000, orders=(1, 2, 3), mach=(0.3, 0.85), aoa=(-4.0, 6.0)):
    """n airfoil cases: mesh size log-uniform in [tri_min, tri_max] triangles (NACA 0012 O-grid with a
    random surface / layer split, ns / nl uniform in [3, 6.5]), order uniform over ``orders``, Mach and
    angle of attack uniform. Returns dicts with ns, nl, mesh_seed, order, mach, aoa_deg."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        tri = float(np.exp(rng.uniform(np.log(tri_min), np.log(tri_max))))
        ns, nl = ogrid_shape(tri, rng.uniform(3.0, 6.5))
        out.append(dict(ns=ns, nl=nl, mesh_seed=int(rng.integers(1 << 30)), order=int(rng.choice(orders)),
                        mach=float(rng.uniform(*mach)), aoa_deg=float(rng.uniform(*aoa))))
    return out
