"""Second, independent restatement of the reference's step in numpy (small meshes only).

TEST INFRASTRUCTURE ONLY: used by tests/ to cross-check the C++ oracle (oracle/refcpu.cpp),
which is the checker for the HIP product. Written separately from refcpu.cpp, in array form:
its own process_mesh, basis values and derivatives from the exact rational Dubiner expansion
(tests/exact_dubiner.py), and the quadrature points/weights of the oracle's tables (pinned by
their exactness tests). Follows:
  process_mesh            src/initialization.rs:112-261
  geometry / neighbours   src/mesh.rs:15-147
  residual                src/spatial_disc.rs:27-69,161-211,349-357; boundary_condition.rs:9-120
  hllc / flux             src/spatial_disc/flux.rs:28-168
  eigenmatrix             src/spatial_disc/local_characteristics.rs:4-50
  KXRCF / WENO            src/spatial_disc/limiter/weno.rs:7-269
  time step / RK3         src/temporal_disc.rs:58-88, src/temporal_disc/rk3.rs:7-26
"""
from __future__ import annotations

import math

import numpy as np

from tests import exact_dubiner as X

EPS = np.finfo(np.float64).eps


class NumpyRef:
    def __init__(self, mesh, tables, order=2, hcr=1.4, cfl=0.0, fs=(0.8, 1.5, 1.0, 1.0e5, 1.4), limiter="sequential"):
        self.p, self.hcr = order, hcr
        self.limiter = limiter
        self.cw, self.cp = tables["cell_w"], tables["cell_pts"]
        self.ew, self.ep = tables["edge_w"], tables["edge_pts"]
        self.modes = X.modes(order)
        nb = len(self.modes)
        self.nb = nb
        self.phi = np.array([[X.eval_deriv(l, k, 0, 0, a, b) for (l, k) in self.modes] for (a, b) in self.cp])
        self.phie = np.array([[[X.eval_deriv(l, k, 0, 0, a, b) for (l, k) in self.modes] for (a, b) in self.ep[e]]
                              for e in range(3)])
        self.mass = (self.cw[:, None] * self.phi ** 2).sum(axis=0)
        maxd = max(order, 2)
        self.dref = {(a, d - a): np.array([[X.eval_deriv(l, k, a, d - a, xa, xb) for (l, k) in self.modes]
                                           for (xa, xb) in self.cp])
                     for d in range(1, maxd + 1) for a in range(d + 1)}
        self._mesh(mesh)
        if cfl > 0:
            self.cfl = cfl
        else:
            q = float(order)
            self.cfl = 1.0 / ((2 * q + 1) * (1 + 4 / (q + 2) ** 2))
        mach, aoa, rho, pr, fh = fs
        c = math.sqrt(fh * pr / rho)
        u, v = mach * c * math.cos(math.radians(aoa)), mach * c * math.sin(math.radians(aoa))
        self.qinf = np.array([rho, rho * u, rho * v, pr / (hcr - 1) + 0.5 * rho * (u * u + v * v)])

    # ---- mesh -----------------------------------------------------------------------------------
    def _mesh(self, m):
        V = m.vxy
        edges = []          # [v0, v1, left, right, lic_l, lic_r, patch]
        key = {}
        o = 0
        for b, cnt in enumerate(m.eb_count):
            for a, c in m.eb_nodes[o:o + cnt]:
                key[frozenset((int(a), int(c)))] = len(edges)
                edges.append([int(a), int(c), 0, -1, 0, -1, b])
            o += cnt
        self.patch_type = list(m.eb_type)
        tri = m.el_nodes
        for t in tri:
            for i in range(3):
                a, c = int(t[i]), int(t[(i + 1) % 3])
                if frozenset((a, c)) not in key:
                    key[frozenset((a, c))] = len(edges)
                    edges.append([a, c, 0, 0, 0, 0, -1])
        el_edges = np.zeros((len(tri), 3), int)
        bnd = []
        for e, t in enumerate(tri):
            isb = False
            for i in range(3):
                a, c = int(t[i]), int(t[(i + 1) % 3])
                k = key[frozenset((a, c))]
                el_edges[e, i] = k
                ed = edges[k]
                if ed[3] == -1:
                    ed[2], ed[4] = e, i
                    isb = True
                elif a == ed[0]:
                    ed[2], ed[4] = e, i
                else:
                    ed[3], ed[5] = e, i
            bnd.append(isb)
        self.E = np.array(edges, dtype=np.int64)
        self.el_edges, self.tri = el_edges, tri
        self.internal_edges = [k for k in range(len(edges)) if edges[k][6] == -1]
        self.boundary_elements = [e for e in range(len(tri)) if bnd[e]]
        self.internal_elements = [e for e in range(len(tri)) if not bnd[e]]
        xy = m.elem_xy if m.elem_xy is not None else V[tri]
        x, y = xy[..., 0], xy[..., 1]
        self.J = (x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) - (x[:, 2] - x[:, 0]) * (y[:, 1] - y[:, 0])
        self.rx = np.stack([(y[:, 2] - y[:, 0]) / self.J, -(x[:, 2] - x[:, 0]) / self.J,
                            -(y[:, 1] - y[:, 0]) / self.J, (x[:, 1] - x[:, 0]) / self.J], axis=1)
        a = np.hypot(x[:, 1] - x[:, 0], y[:, 1] - y[:, 0])
        b = np.hypot(x[:, 2] - x[:, 1], y[:, 2] - y[:, 1])
        c = np.hypot(x[:, 0] - x[:, 2], y[:, 0] - y[:, 2])
        s = (a + b + c) / 2
        self.cr = a * b * c / (4 * np.sqrt(s * (s - a) * (s - b) * (s - c)))
        self.hmin = np.minimum(np.minimum(self.J / a, self.J / b), self.J / c)
        n = len(edges)
        self.n = np.zeros((n, 2))
        self.len = np.zeros(n)
        for k, ed in enumerate(edges):
            if m.elem_xy is not None and ed[6] == -1:
                e, i = ed[2], ed[4]
                p0, p1 = xy[e, i], xy[e, (i + 1) % 3]
            else:
                p0, p1 = V[ed[0]], V[ed[1]]
            L = math.hypot(p1[0] - p0[0], p1[1] - p0[1])
            self.len[k] = L
            self.n[k] = [(p1[1] - p0[1]) / L, -(p1[0] - p0[0]) / L]
        self.nbr = -np.ones((len(tri), 3), int)
        for e in range(len(tri)):
            for i in range(3):
                ed = edges[el_edges[e, i]]
                if ed[3] == -1:
                    continue
                self.nbr[e, i] = ed[3] if tri[e, i] == ed[0] else ed[2]

    def dphys(self, e, l1, l2):
        """[nq, nb] physical derivative table d^(l1+l2) / dx^l1 dy^l2 of element e."""
        xx, xy, ex, ey = self.rx[e]
        out = np.zeros_like(self.phi)
        for i in range(l1 + 1):
            for j in range(l2 + 1):
                c = math.comb(l1, i) * math.comb(l2, j) * xx ** i * ex ** (l1 - i) * xy ** j * ey ** (l2 - j)
                out += c * self.dref[(i + j, l1 + l2 - i - j)]
        return out

    # ---- physics --------------------------------------------------------------------------------
    def flux(self, q):
        r, mu, mv, E = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
        u, v = mu / r, mv / r
        p = (self.hcr - 1) * (E - 0.5 * (mu * mu + mv * mv) / r)
        f = np.stack([mu, mu * u + p, mv * u, u * (E + p)], -1)
        g = np.stack([mv, mu * v, mv * v + p, v * (E + p)], -1)
        return f, g

    def hllc(self, L, R, n):
        g = self.hcr
        nx, ny = n
        rot = lambda q: np.array([q[0], q[1] * nx + q[2] * ny, -q[1] * ny + q[2] * nx, q[3]])
        ql, qr = rot(L), rot(R)
        ul, ur, vl, vr = ql[1] / ql[0], qr[1] / qr[0], ql[2] / ql[0], qr[2] / qr[0]
        pl = (g - 1) * (ql[3] - 0.5 * (ql[1] ** 2 + ql[2] ** 2) / ql[0])
        pr = (g - 1) * (qr[3] - 0.5 * (qr[1] ** 2 + qr[2] ** 2) / qr[0])
        if ql[0] < 0 or qr[0] < 0:
            raise FloatingPointError("density")
        if pl < 0 or pr < 0:
            raise FloatingPointError("pressure")
        cl, cr = math.sqrt(g * pl / ql[0]), math.sqrt(g * pr / qr[0])
        fl = np.array([ql[1], ql[1] * ul + pl, ql[2] * ul, ul * (ql[3] + pl)])
        fr = np.array([qr[1], qr[1] * ur + pr, qr[2] * ur, ur * (qr[3] + pr)])
        z = (g - 1) / (2 * g)
        base = (cl + cr - (z - 1) / 2 * (ur - ul)) / (cl / pl ** z + cr / pr ** z)
        ps = base ** (1 / z) if base >= 0 else float("nan")
        fac = lambda pp: 1.0 if ps <= pp else math.sqrt(1 + (g + 1) / (2 * g) * (ps / pp - 1))
        sl, sr = ul - cl * fac(pl), ur + cr * fac(pr)
        if sl >= 0:
            F = fl
        elif sr <= 0:
            F = fr
        else:
            ss = (pr - pl + ql[0] * ul * (sl - ul) - qr[0] * ur * (sr - ur)) / (ql[0] * (sl - ul) - qr[0] * (sr - ur))
            if ss >= 0:
                q0, uu, vv, P, S, FF = ql, ul, vl, pl, sl, fl
            else:
                q0, uu, vv, P, S, FF = qr, ur, vr, pr, sr, fr
            a = q0[0] * (S - uu) / (S - ss)
            qs = np.array([a, a * ss, a * vv, a * (q0[3] / q0[0] + (ss - uu) * (ss + P / (q0[0] * (S - uu))))])
            F = FF + S * (qs - q0)
        return np.array([F[0], F[1] * nx - F[2] * ny, F[1] * ny + F[2] * nx, F[3]])

    def eig(self, q, nx, ny):
        g = self.hcr
        u, v = q[1] / q[0], q[2] / q[0]
        p = (g - 1) * (q[3] - 0.5 * (q[1] ** 2 + q[2] ** 2) / q[0])
        c = math.sqrt(g * p / q[0])
        h = (q[3] + p) / q[0]
        b1 = (g - 1) / c ** 2
        b2 = b1 * (u * u + v * v) / 2
        vn = u * nx + v * ny
        Lm = np.array([[(b2 + vn / c) / 2, -(b1 * u + nx / c) / 2, -(b1 * v + ny / c) / 2, b1 / 2],
                       [ny * u - nx * v, -ny, nx, 0.0],
                       [1 - b2, b1 * u, b1 * v, -b1],
                       [(b2 - vn / c) / 2, -(b1 * u - nx / c) / 2, -(b1 * v - ny / c) / 2, b1 / 2]])
        Rm = np.array([[1.0, 0.0, 1.0, 1.0],
                       [u - c * nx, -ny, u, u + c * nx],
                       [v - c * ny, nx, v, v + c * ny],
                       [h - c * vn, -u * ny + v * nx, (u * u + v * v) / 2, h + c * vn]])
        return Lm, Rm

    # ---- residual -------------------------------------------------------------------------------
    def volume(self, U, e):
        uq = U[e] @ self.phi.T                          # [4, nq]
        f, g = self.flux(uq.T)                          # [nq, 4]
        dx, dy = self.dphys(e, 1, 0), self.dphys(e, 0, 1)
        w = self.cw * 0.5 * self.J[e]
        return (f * w[:, None]).T @ dx + (g * w[:, None]).T @ dy

    def residual(self, U):
        R = np.zeros_like(U)
        for e in self.internal_elements:
            R[e] += self.volume(U, e)
        ne = len(self.ew)
        for k in self.internal_edges:
            v0, v1, L, Rr, il, ir, _ = self.E[k]
            for q in range(ne):
                qa = U[L] @ self.phie[il, q]
                qb = U[Rr] @ self.phie[ir, ne - 1 - q]
                F = self.hllc(qa, qb, self.n[k])
                R[L] -= self.ew[q] * np.outer(F, self.phie[il, q]) * 0.5 * self.len[k]
                R[Rr] += self.ew[ne - 1 - q] * np.outer(F, self.phie[ir, ne - 1 - q]) * 0.5 * self.len[k]
        for e in self.boundary_elements:
            for i in range(3):
                k = self.el_edges[e, i]
                v0, v1, L, Rr, il, ir, patch = self.E[k]
                if patch < 0:
                    continue
                nx, ny = self.n[k] if self.tri[e, i] == v0 else -self.n[k]
                for q in range(ne):
                    ql = U[e] @ self.phie[il, q]
                    if self.patch_type[patch] == 0:
                        p = (self.hcr - 1) * (ql[3] - 0.5 * (ql[1] ** 2 + ql[2] ** 2) / ql[0])
                        F = np.array([0.0, p * nx, p * ny, 0.0])
                    else:
                        u, v = ql[1] / ql[0], ql[2] / ql[0]
                        p = (self.hcr - 1) * (ql[3] - 0.5 * (ql[1] ** 2 + ql[2] ** 2) / ql[0])
                        c = math.sqrt(self.hcr * p / ql[0])
                        vn = u * nx + v * ny
                        Lm, Rm = self.eig(ql, nx, ny)
                        lam = np.array([vn - c, vn, vn, vn + c])
                        qb = Rm @ np.where(lam >= 0, Lm @ ql, Lm @ self.qinf)
                        ub, vb = qb[1] / qb[0], qb[2] / qb[0]
                        vnb = ub * nx + vb * ny
                        pb = (self.hcr - 1) * (qb[3] - 0.5 * (qb[1] ** 2 + qb[2] ** 2) / qb[0])
                        F = np.array([qb[0] * vnb, qb[1] * vnb + pb * nx, qb[2] * vnb + pb * ny, (qb[3] + pb) * vnb])
                    R[e] -= self.ew[q] * np.outer(F, self.phie[il, q]) * 0.5 * self.len[k]
            R[e] += self.volume(U, e)
        return R / (self.mass[None, None, :] * 0.5 * self.J[:, None, None])

    # ---- limiter --------------------------------------------------------------------------------
    def kxrcf(self, U, e):
        ne = len(self.ew)
        dj = ej = surf = 0.0
        mr = me = 1.0e10
        for i in range(3):
            k = self.el_edges[e, i]
            v0, v1, L, Rr, il, ir, patch = self.E[k]
            if Rr == -1:
                continue
            nb = self.nbr[e, i]
            nx, ny = self.n[k] if self.tri[e, i] == v0 else -self.n[k]
            tr = U[e] @ self.phie[i].T               # [4, ne] traces on the element's own edge i
            mr, me = min(mr, tr[0].min()), min(me, tr[3].min())
            mom = tr[:, ne // 2]
            if mom[1] * nx + mom[2] * ny < 0:
                ext = U[nb] @ self.phie[i, ::-1].T   # neighbour modes on the TARGET's edge table (quirk)
                dj += float(np.sum(0.5 * self.len[k] * self.ew * (tr[0] - ext[0])))
                ej += float(np.sum(0.5 * self.len[k] * self.ew * (tr[3] - ext[3])))
                surf += self.len[k]
        h = self.cr[e] ** 1.5
        with np.errstate(invalid="ignore", divide="ignore"):
            return bool(abs(dj) / (h * surf * mr) >= 1.0 or abs(ej) / (h * surf * me) >= 1.0)

    def kxrcf_margin(self, U, e, eps=2.0 ** -24):
        """kxrcf() above, plus whether the decision is certain under FP32 evaluation of the same state:
        returns (troubled, certain). An evaluation in single precision (traces as sums of Np products,
        jumps as sums over the edge points) moves each indicator by at most the bound below; the decision
        is uncertain when an indicator lies within that bound of 1, or an inflow test within rounding of
        0. Checker for the FP32-bulk detector; the bound is a generous a-priori one (16 eps per sum)."""
        ne = len(self.ew)
        dj = ej = surf = 0.0
        edj = eej = 0.0
        mr = me = 1.0e10
        emr = eme = 0.0
        certain = True
        for i in range(3):
            k = self.el_edges[e, i]
            v0, v1, L, Rr, il, ir, patch = self.E[k]
            if Rr == -1:
                continue
            nb = self.nbr[e, i]
            nx, ny = self.n[k] if self.tri[e, i] == v0 else -self.n[k]
            tr = U[e] @ self.phie[i].T
            ta = np.abs(U[e]) @ np.abs(self.phie[i]).T            # size of the terms of each trace
            if tr[0].min() < mr:
                mr, emr = tr[0].min(), 16 * eps * ta[0][np.argmin(tr[0])]
            if tr[3].min() < me:
                me, eme = tr[3].min(), 16 * eps * ta[3][np.argmin(tr[3])]
            mom, mta = tr[:, ne // 2], ta[:, ne // 2]
            mn = mom[1] * nx + mom[2] * ny
            if abs(mn) <= 16 * eps * (mta[1] * abs(nx) + mta[2] * abs(ny)):
                certain = False
            if mn < 0:
                ext = U[nb] @ self.phie[i, ::-1].T
                xa = np.abs(U[nb]) @ np.abs(self.phie[i, ::-1]).T
                c = 0.5 * self.len[k] * self.ew
                dj += float(np.sum(c * (tr[0] - ext[0])))
                ej += float(np.sum(c * (tr[3] - ext[3])))
                edj += 16 * eps * float(np.sum(c * (ta[0] + xa[0])))
                eej += 16 * eps * float(np.sum(c * (ta[3] + xa[3])))
                surf += self.len[k]
        h = self.cr[e] ** 1.5
        if surf == 0.0:
            return False, certain
        out = []
        for jmp, err, mn_, emn in ((dj, edj, mr, emr), (ej, eej, me, eme)):
            ind = abs(jmp) / (h * surf * mn_)
            # rounding of the jump, of the minimum and of the products and the division
            tol = err / (h * surf * abs(mn_)) + ind * (emn / abs(mn_) + 16 * eps)
            out.append((ind >= 1.0, abs(ind - 1.0) > tol))
        troubled = out[0][0] or out[1][0]
        # a certain "troubled" from either indicator settles it; otherwise both must be certain
        sure = (out[0][0] and out[0][1]) or (out[1][0] and out[1][1]) or (out[0][1] and out[1][1])
        return troubled, certain and sure

    def weno(self, U, e):
        lw = [0.997, 0.001, 0.001, 0.001]
        bnd = [i for i in range(3) if self.nbr[e, i] < 0]
        for i in bnd:
            lw[i] = 0.0
        lw[len(bnd)] = [0.997, 0.998, 0.999][len(bnd)]
        keys = [(l1, l - l1) for l in range(1, self.p + 1) for l1 in range(l + 1)]
        D = {kk: self.dphys(e, *kk) for kk in keys}
        w = 0.5 * self.J[e] * self.cw

        def beta(P):
            return sum(self.J[e] ** (l1 + l2 - 1) * float(np.sum(w * (D[(l1, l2)] @ P) ** 2)) for (l1, l2) in keys)

        rec = np.zeros((4, self.nb))
        area = 0.0
        for i in range(3):
            if self.nbr[e, i] < 0:
                continue
            area += 0.5 * self.J[self.nbr[e, i]]
            Lm, Rm = self.eig(U[e][:, 0], *self.n[self.el_edges[e, i]])
            polys = [U[e]] + [None] * 3
            for j in range(3):
                if self.nbr[e, j] >= 0:
                    P = U[self.nbr[e, j]].copy()
                    P[:, 0] = U[e][:, 0]
                    polys[j + 1] = P
            proj = [None if P is None else Lm @ P for P in polys]
            nlw = np.array([[0.0 if proj[j] is None else lw[j] / (beta(proj[j][v]) + EPS) ** 2 for v in range(4)]
                            for j in range(4)])
            wts = nlw / nlw.sum(axis=0)
            new = sum(wts[j][:, None] * proj[j] for j in range(4) if proj[j] is not None)
            rec += (Rm @ new) * 0.5 * self.J[self.nbr[e, i]]
        return rec / area

    def apply_limiter(self, U):
        U = U.copy()
        flags = np.zeros(len(U), bool)
        if self.limiter == "sequential":
            for e in range(len(U)):
                if self.kxrcf(U, e):
                    flags[e] = True
                    U[e] = self.weno(U, e)
        elif self.limiter == "jacobi":
            flags = np.array([self.kxrcf(U, e) for e in range(len(U))])
            new = {e: self.weno(U, e) for e in np.flatnonzero(flags)}
            for e, v in new.items():
                U[e] = v
        return U, flags

    # ---- time -----------------------------------------------------------------------------------
    def time_step(self, U):
        s = U.sum(axis=2)
        u, v = s[:, 1] / s[:, 0], s[:, 2] / s[:, 0]
        p = (self.hcr - 1) * (s[:, 3] - 0.5 * s[:, 0] * (u * u + v * v))
        with np.errstate(invalid="ignore"):
            dt = self.cfl * self.hmin / (np.sqrt(u * u + v * v) + np.sqrt(self.hcr * p / s[:, 0]))
        dt = dt[~np.isnan(dt)]
        return float(min(1.0e10, dt.min())) if dt.size else 1.0e10

    def rk3(self, U, dt):
        u1, _ = self.apply_limiter(U + dt * self.residual(U))
        u2, _ = self.apply_limiter(0.75 * U + 0.25 * (u1 + dt * self.residual(u1)))
        u3, _ = self.apply_limiter(U / 3.0 + 2.0 / 3.0 * (u2 + dt * self.residual(u2)))
        return u3

    def time_march(self, U, final_step, final_time):
        t, n = 0.0, 0
        while n < final_step and t < final_time:
            dt = self.time_step(U)
            if t + dt > final_time:
                dt = final_time - t
            U = self.rk3(U, dt)
            t += dt
            n += 1
        return U, n, t
