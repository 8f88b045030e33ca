"""Shared test helpers: seeded flow states and error norms."""
from __future__ import annotations

import numpy as np


def prim_to_cons(rho, u, v, p, hcr=1.4):
    return np.stack([rho, rho * u, rho * v, p / (hcr - 1.0) + 0.5 * rho * (u * u + v * v)], axis=-1)


def flow_state(case, seed=0, amp=0.05, shock=False, mach=0.5, p0=1.0e5, rho0=1.0, hcr=1.4, noise=0.0):
    """Smooth (optionally shocked) state projected on the element modes of ``case``.

    ``case`` is an oracle RefCase or a product Context (both provide cell_points_xy / project).
    Density / velocity / pressure are a freestream plus seeded sine modes; with ``shock`` a density
    and pressure step across the line x = 0.3 (x + 0.2 y) makes KXRCF flag cells near it.
    """
    rng = np.random.default_rng(seed)
    xy = case.cell_points_xy()
    x, y = xy[..., 0], xy[..., 1]
    c0 = np.sqrt(hcr * p0 / rho0)
    k = rng.uniform(0.5, 2.0, size=6)
    ph = rng.uniform(0, 2 * np.pi, size=6)
    rho = rho0 * (1.0 + amp * np.sin(k[0] * x + ph[0]) * np.cos(k[1] * y + ph[1]))
    u = mach * c0 * (1.0 + amp * np.sin(k[2] * y + ph[2]))
    v = mach * c0 * amp * np.cos(k[3] * x + ph[3])
    p = p0 * (1.0 + amp * np.cos(k[4] * x + k[5] * y + ph[4]))
    if shock:
        s = (x + 0.2 * y) > 0.3
        rho = np.where(s, rho * 1.6, rho)
        p = np.where(s, p * 2.2, p)
        u = np.where(s, u * 0.6, u)
    q = prim_to_cons(rho, u, v, p, hcr)
    U = case.project(q)
    if noise > 0.0:
        U[:, :, 1:] += noise * rng.standard_normal(U[:, :, 1:].shape) * np.abs(U[:, :, :1])
    return U


def rel_err(a, b):
    """max over variables of ||a_v - b_v||_inf / ||b_v||_inf (variables on axis 1)."""
    a = np.asarray(a)
    b = np.asarray(b)
    num = np.abs(a - b).max(axis=(0, 2))
    den = np.maximum(np.abs(b).max(axis=(0, 2)), 1e-300)
    return float((num / den).max())


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def rel_l2_per_var(a, b):
    """[4] relative L2 difference of each conserved variable (variables on axis 1), coefficient norm: rho and
    E against their own norms, each momentum component against the norm of the momentum vector (a
    component's own norm vanishes with the flow angle, and the error it carries does not)."""
    a = np.asarray(a)
    b = np.asarray(b)
    num = np.sqrt(((a - b) ** 2).sum(axis=(0, 2)))
    own = np.sqrt((b ** 2).sum(axis=(0, 2)))
    den = np.array([own[0], np.hypot(own[1], own[2]), np.hypot(own[1], own[2]), own[3]])
    return num / np.maximum(den, 1e-300)
