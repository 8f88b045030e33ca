// Basis and quadrature tables for the device kernels.
//
// Quadrature (src/gauss_point.rs:39-95): the reference's 3/4/5-point Gauss-Legendre and
// 4/6-point triangle literals are reproduced digit for digit (they define the p = 2 discrete
// operator). For p = 3 and p = 4 the build adds a 16-point degree-8 symmetric triangle rule and
// a 25-point conical-product rule (degree 9).
//
// Basis (src/basis_function.rs:56-113): psi_{l,k}(xi, eta) = 2^k P_k(2 xi/(1-eta) - 1) (1-eta)^k
// P_l^(2k+1,0)(2 eta - 1), modes ordered by total degree, then l ascending. Values come from the
// three-term Jacobi recurrence. Reference-element derivatives of every order come from the
// polynomial's monomial coefficients, found here by interpolating psi on the degree-p principal
// lattice (an exact solve in extended precision), then differentiated term by term; at p = 2 they
// equal the reference's hard-coded tables (src/basis_function.rs:114-193).
#include <cmath>

#include "dgm_internal.h"

namespace dgm {

namespace {

const double kGL3x[3] = {-0.7745966692414834, 0.0, 0.7745966692414834};
const double kGL3w[3] = {0.5555555555555556, 0.8888888888888889, 0.5555555555555556};
const double kGL4x[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
const double kGL4w[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};
const double kGL5x[5] = {-0.9061798459386640, -0.5384693101056831, 0.0, 0.5384693101056831, 0.9061798459386640};
const double kGL5w[5] = {0.2369268850561891, 0.4786286704993665, 0.5688888888888889, 0.4786286704993665,
                         0.2369268850561891};

bool gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  const double *px, *pw;
  switch (n) {
    case 3: px = kGL3x; pw = kGL3w; break;
    case 4: px = kGL4x; pw = kGL4w; break;
    case 5: px = kGL5x; pw = kGL5w; break;
    default: return false;
  }
  x.assign(px, px + n);
  w.assign(pw, pw + n);
  return true;
}

void push(std::vector<double>& p, std::vector<double>& w, double x, double y, double wt) {
  p.push_back(x);
  p.push_back(y);
  w.push_back(wt);
}

bool triangle(int n, std::vector<double>& p, std::vector<double>& w) {
  p.clear();
  w.clear();
  switch (n) {
    case 4:
      push(p, w, 1.0 / 3.0, 1.0 / 3.0, -27.0 / 96.0);
      push(p, w, 2.0 / 15.0, 11.0 / 15.0, 25.0 / 96.0);
      push(p, w, 2.0 / 15.0, 2.0 / 15.0, 25.0 / 96.0);
      push(p, w, 11.0 / 15.0, 2.0 / 15.0, 25.0 / 96.0);
      return true;
    case 6: {
      const double a = 0.44594849091597, b = 0.10810301816807, c = 0.09157621350977, d = 0.81684757298046;
      const double wa = 0.22338158967801, wc = 0.10995174365532;
      push(p, w, a, a, wa);
      push(p, w, a, b, wa);
      push(p, w, b, a, wa);
      push(p, w, c, c, wc);
      push(p, w, c, d, wc);
      push(p, w, d, c, wc);
      return true;
    }
    case 16: {
      // symmetric degree-8 rule: centroid, three (a, a, 1-2a) orbits, one (b, c, 1-b-c) orbit
      push(p, w, 1.0 / 3.0, 1.0 / 3.0, 0.14431560767764748);
      const double a[3] = {0.4592925882926314, 0.17056930775165638, 0.05054722831703524};
      const double wa[3] = {0.09509163426737452, 0.10321737053473984, 0.03245849762321117};
      for (int i = 0; i < 3; ++i) {
        const double o = 1.0 - 2.0 * a[i];
        push(p, w, a[i], a[i], wa[i]);
        push(p, w, a[i], o, wa[i]);
        push(p, w, o, a[i], wa[i]);
      }
      const double b = 0.00839477740981607, c = 0.2631128296349502, d = 1.0 - b - c, wb = 0.02723031417439596;
      push(p, w, b, c, wb);
      push(p, w, c, b, wb);
      push(p, w, b, d, wb);
      push(p, w, d, b, wb);
      push(p, w, c, d, wb);
      push(p, w, d, c, wb);
      return true;
    }
    case 25: {
      // conical product: xi = s (1 - t), eta = t, s on 5-point Gauss-Legendre, t on 5-point Gauss-Jacobi(1,0)
      const double s[5] = {0.046910077030668018, 0.23076534494715845, 0.5, 0.7692346550528415,
                           0.95308992296933193};
      const double ws[5] = {0.11846344252809449, 0.23931433524968326, 0.2844444444444445, 0.23931433524968326,
                            0.11846344252809449};
      const double t[5] = {0.039809857051468722, 0.19801341787360821, 0.43797481024738616, 0.69546427335363614,
                           0.90146491420117358};
      const double wt[5] = {0.19356318045330295, 0.33434927618873939, 0.2927739741693397, 0.14781774014523336,
                            0.031495829043384597};
      for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) push(p, w, s[j] * (1.0 - t[i]), t[i], ws[j] * wt[i]);
      return true;
    }
    default:
      return false;
  }
}

double jacobi(int n, double a, double b, double x) {
  if (n == 0) return 1.0;
  double p0 = 1.0;
  double p1 = (a + 1.0) + (a + b + 2.0) * (x - 1.0) / 2.0;
  for (int m = 2; m <= n; ++m) {
    const double c = 2.0 * m + a + b;
    const double a1 = 2.0 * m * (m + a + b) * (c - 2.0);
    const double a2 = c - 1.0;
    const double a3 = c * (c - 2.0) * x + a * a - b * b;
    const double a4 = 2.0 * (m + a - 1.0) * (m + b - 1.0) * c;
    const double p2 = (a2 * a3 * p1 - a4 * p0) / a1;
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

double dubiner(int l, int k, double xi, double eta) {
  const double leg = jacobi(k, 0.0, 0.0, 2.0 * xi / (1.0 - eta) - 1.0);
  const double jac = jacobi(l, 2.0 * k + 1.0, 0.0, 2.0 * eta - 1.0);
  return std::ldexp(1.0, k) * leg * std::pow(1.0 - eta, k) * jac;
}

// Solve A x = b (n x n, row-major) in extended precision with partial pivoting.
bool solve_ld(int n, std::vector<long double> A, std::vector<long double>& b) {
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs((double)A[r * n + c]) > std::fabs((double)A[piv * n + c])) piv = r;
    if (A[piv * n + c] == 0.0L) return false;
    if (piv != c) {
      for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[piv * n + k]);
      std::swap(b[c], b[piv]);
    }
    for (int r = c + 1; r < n; ++r) {
      const long double f = A[r * n + c] / A[c * n + c];
      for (int k = c; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
      b[r] -= f * b[c];
    }
  }
  for (int c = n - 1; c >= 0; --c) {
    long double s = b[c];
    for (int k = c + 1; k < n; ++k) s -= A[c * n + k] * b[k];
    b[c] = s / A[c * n + c];
  }
  return true;
}

}  // namespace

int Tables::key(int a, int b) const {
  for (int i = 0; i < nkeys; ++i)
    if (key_a[i] == a && key_b[i] == b) return i;
  return -1;
}

int default_cell_rule(int p) { return p <= 2 ? 6 : (p == 3 ? 16 : 25); }
int default_edge_rule(int p) { return p <= 2 ? 3 : 5; }

bool build_tables(Tables& t, int p, int nq, int ne) {
  if (p < 1 || p > 4) return false;
  std::vector<double> gx;
  if (!triangle(nq, t.cell_pts, t.cell_w) || !gauss_legendre(ne, gx, t.edge_w)) return false;
  t.p = p;
  t.np = (p + 1) * (p + 2) / 2;
  t.nq = nq;
  t.ne = ne;
  // edge points of the local edges e0 (0,0)->(1,0), e1 (1,0)->(0,1), e2 (0,1)->(0,0)
  t.edge_pts.resize(3 * ne * 2);
  for (int i = 0; i < ne; ++i) {
    const double x = gx[i], xr = gx[ne - 1 - i];
    const double pts[3][2] = {{(x + 1.0) / 2.0, 0.0}, {(1.0 - x) / 2.0, (1.0 + x) / 2.0}, {0.0, (xr + 1.0) / 2.0}};
    for (int e = 0; e < 3; ++e) {
      t.edge_pts[(e * ne + i) * 2] = pts[e][0];
      t.edge_pts[(e * ne + i) * 2 + 1] = pts[e][1];
    }
  }
  std::vector<int> L, K;
  for (int d = 0; d <= p; ++d)
    for (int l = 0; l <= d; ++l) {
      L.push_back(l);
      K.push_back(d - l);
    }
  const int np = t.np;
  t.phi_cell.resize(nq * np);
  for (int q = 0; q < nq; ++q)
    for (int b = 0; b < np; ++b) t.phi_cell[q * np + b] = dubiner(L[b], K[b], t.cell_pts[2 * q], t.cell_pts[2 * q + 1]);
  t.phi_edge.resize(3 * ne * np);
  for (int e = 0; e < 3; ++e)
    for (int i = 0; i < ne; ++i)
      for (int b = 0; b < np; ++b)
        t.phi_edge[(e * ne + i) * np + b] =
            dubiner(L[b], K[b], t.edge_pts[(e * ne + i) * 2], t.edge_pts[(e * ne + i) * 2 + 1]);
  t.mass.assign(np, 0.0);
  for (int b = 0; b < np; ++b)
    for (int q = 0; q < nq; ++q) t.mass[b] += t.cell_w[q] * t.phi_cell[q * np + b] * t.phi_cell[q * np + b];

  // monomial coefficients by interpolation on the principal lattice (i/p, j/p), i + j <= p
  std::vector<int> ma, mb;  // monomial exponents, same count as modes
  for (int d = 0; d <= p; ++d)
    for (int a = d; a >= 0; --a) {
      ma.push_back(a);
      mb.push_back(d - a);
    }
  // lattice shrunk toward the centroid by 0.9 (still unisolvent, and clear of the collapsed vertex (0,1))
  std::vector<double> lx, ly;
  for (int j = 0; j <= p; ++j)
    for (int i = 0; i + j <= p; ++i) {
      lx.push_back(1.0 / 3.0 + 0.9 * ((double)i / p - 1.0 / 3.0));
      ly.push_back(1.0 / 3.0 + 0.9 * ((double)j / p - 1.0 / 3.0));
    }
  std::vector<long double> V(np * np);
  for (int r = 0; r < np; ++r)
    for (int c = 0; c < np; ++c) V[r * np + c] = std::pow((long double)lx[r], ma[c]) * std::pow((long double)ly[r], mb[c]);
  std::vector<std::vector<long double>> coef(np);
  for (int b = 0; b < np; ++b) {
    std::vector<long double> rhs(np);
    for (int r = 0; r < np; ++r) rhs[r] = dubiner(L[b], K[b], lx[r], ly[r]);
    if (!solve_ld(np, V, rhs)) return false;
    coef[b] = rhs;
  }
  t.maxd = p > 2 ? p : 2;
  t.key_a.clear();
  t.key_b.clear();
  for (int d = 1; d <= t.maxd; ++d)
    for (int a = d; a >= 0; --a) {
      t.key_a.push_back(a);
      t.key_b.push_back(d - a);
    }
  t.nkeys = (int)t.key_a.size();
  t.dref.assign((size_t)t.nkeys * nq * np, 0.0);
  for (int kk = 0; kk < t.nkeys; ++kk)
    for (int q = 0; q < nq; ++q)
      for (int b = 0; b < np; ++b) {
        const long double xi = t.cell_pts[2 * q], eta = t.cell_pts[2 * q + 1];
        long double s = 0.0L;
        for (int m = 0; m < np; ++m) {
          const int a = ma[m], c = mb[m];
          if (a < t.key_a[kk] || c < t.key_b[kk]) continue;
          long double f = coef[b][m];
          for (int i = 0; i < t.key_a[kk]; ++i) f *= (long double)(a - i);
          for (int i = 0; i < t.key_b[kk]; ++i) f *= (long double)(c - i);
          s += f * std::pow(xi, a - t.key_a[kk]) * std::pow(eta, c - t.key_b[kk]);
        }
        t.dref[((size_t)kk * nq + q) * np + b] = (double)s;
      }
  return true;
}

}  // namespace dgm
