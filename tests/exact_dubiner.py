"""Exact rational expansion of the Dubiner basis, independent of both C++ implementations.

psi_{l,k}(xi, eta) = 2^k P_k(2 xi / (1 - eta) - 1) (1 - eta)^k P_l^(2k+1, 0)(2 eta - 1)
(src/basis_function.rs:75-79) is a polynomial of degree l + k. Its monomial coefficients are
built here with fractions.Fraction from the textbook three-term Jacobi recurrence, so values
and derivatives of every order at a floating-point point are exact up to the final rounding.
"""
from __future__ import annotations

from fractions import Fraction as F
from functools import lru_cache
from typing import Dict, Tuple

Poly1 = Tuple[F, ...]
Poly2 = Dict[Tuple[int, int], F]


@lru_cache(maxsize=None)
def jacobi_coeffs(n: int, a: int, b: int) -> Poly1:
    """Monomial coefficients (in x) of P_n^(a,b)(x)."""
    a, b = F(a), F(b)
    p0 = [F(1)]
    if n == 0:
        return tuple(p0)
    p1 = [(a + 1) - (a + b + 2) / 2, (a + b + 2) / 2]
    for m in range(2, n + 1):
        c = 2 * m + a + b
        a1 = 2 * m * (m + a + b) * (c - 2)
        a2 = c - 1
        a3x, a30 = c * (c - 2), a * a - b * b
        a4 = 2 * (m + a - 1) * (m + b - 1) * c
        r = [F(0)] * (m + 1)
        for j, v in enumerate(p1):
            r[j + 1] += a2 * a3x * v / a1
            r[j] += a2 * a30 * v / a1
        for j, v in enumerate(p0):
            r[j] -= a4 * v / a1
        p0, p1 = p1, r
    return tuple(p1)


def _mul(p: Poly2, q: Poly2) -> Poly2:
    r: Poly2 = {}
    for (i, j), u in p.items():
        for (k, l), v in q.items():
            r[(i + k, j + l)] = r.get((i + k, j + l), F(0)) + u * v
    return r


def _pow(p: Poly2, n: int) -> Poly2:
    r: Poly2 = {(0, 0): F(1)}
    for _ in range(n):
        r = _mul(r, p)
    return r


@lru_cache(maxsize=None)
def dubiner_poly(l: int, k: int):
    s = {(1, 0): F(2), (0, 1): F(1), (0, 0): F(-1)}   # 2 xi + eta - 1
    om = {(0, 0): F(1), (0, 1): F(-1)}                 # 1 - eta
    A: Poly2 = {}
    for j, c in enumerate(jacobi_coeffs(k, 0, 0)):
        t = _mul(_pow(s, j), _pow(om, k - j))
        for key, v in t.items():
            A[key] = A.get(key, F(0)) + (2 ** k) * c * v
    x = {(0, 1): F(2), (0, 0): F(-1)}                  # 2 eta - 1
    B: Poly2 = {}
    for j, c in enumerate(jacobi_coeffs(l, 2 * k + 1, 0)):
        for key, v in _pow(x, j).items():
            B[key] = B.get(key, F(0)) + c * v
    return tuple(sorted(_mul(A, B).items()))


def modes(p: int):
    return [(l, d - l) for d in range(p + 1) for l in range(d + 1)]


def eval_deriv(l: int, k: int, da: int, db: int, xi: float, eta: float) -> float:
    X, Y = F(xi), F(eta)
    s = F(0)
    for (a, b), c in dubiner_poly(l, k):
        if a < da or b < db:
            continue
        f = c
        for i in range(da):
            f *= a - i
        for i in range(db):
            f *= b - i
        s += f * X ** (a - da) * Y ** (b - db)
    return float(s)


def mass_closed_form(l: int, k: int) -> float:
    """(1/|T|) int_T psi_{l,k}^2 = 4^k / ((2k+1)(l+k+1)) on the reference triangle."""
    return 4.0 ** k / ((2 * k + 1) * (l + k + 1))
