// refcpu.cpp — FP64 CPU restatement of the Decem5150/DGM_in_rust explicit SSP-RK3 step.
//
// TEST INFRASTRUCTURE ONLY (the oracle). Loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker; the HIP product never links or calls it.
//
// Built with -O2 -ffp-contract=off: the reference is Rust, which never contracts a*b+c into
// an FMA, so every expression below keeps the reference's operation order. Where the
// reference calls ndarray's 2-D `dot` on f64 (which dispatches to matrixmultiply 0.3.8
// dgemm), the product is restated as an explicit FMA chain in k order (gemm_fma = 1, the
// kernel matrixmultiply selects on FMA-capable x86) or as a mul/add chain (gemm_fma = 0).
// `x.powf(2.0)` / `x.powi(2)` are restated as x*x (LLVM folds them so in optimised builds);
// `powf` with a run-time exponent is libm pow.
//
// Every function names the reference file:line it restates. Extensions beyond the
// reference (orders p != 2, quadrature rules for p = 3, 4, periodic meshes given by
// per-element coordinates, a Jacobi limiter order, function projection) are marked
// "extension"; their p = 2 / non-periodic / sequential paths are the restatement itself.
#include "refcpu.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum Status { OK = 0, NEG_DENSITY = 1, NEG_PRESSURE = 2, NONFINITE = 3, BAD_CONFIG = 4, BAD_MESH = 5 };

// Planted mistakes, for the tests that pin the quirk-bearing steps: a build with -DORACLE_PLANT=n
// (oracle/Makefile target `plants`, libraries librefcpu_plant<n>.so) replaces one quirk with the
// plausible alternative named below, and each closed-form test must fail on the matching build. The
// oracle itself is built with ORACLE_PLANT = 0, where every hook below compiles away.
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif
enum Plant {
  PLANT_NONE = 0,
  PLANT_DT_MODE0 = 1,             // time-step wave speed from mode 0 (the mean) instead of the sum of modes
  PLANT_HLLC_TEXTBOOK_PSTAR = 2,  // p* numerator c_L + c_R - ((gamma - 1)/2)(u_R - u_L)
  PLANT_KXRCF_NB_OWN_EDGE = 3,    // neighbour trace on the neighbour's own local edge
  PLANT_KXRCF_H1 = 4,             // indicator scaled by h instead of h^1.5
  PLANT_KXRCF_OUTFLOW = 5,        // jumps over outflow faces instead of inflow ones
  PLANT_KXRCF_ALL_SURFACE = 6,    // surface summed over every face with a neighbour
  PLANT_WENO_WEIGHT_NEXT = 7,     // a boundary face i zeroes weights[i + 1] (its neighbour's candidate)
  PLANT_WENO_NO_MEAN_SWAP = 8,    // neighbour candidates keep their own mode 0
  PLANT_WENO_PLAIN_MEAN = 9,      // unweighted mean of the per-face reconstructions
};
constexpr int kPlant = ORACLE_PLANT;

// ---------------------------------------------------------------------------------------
// Quadrature — src/gauss_point.rs
// ---------------------------------------------------------------------------------------

// gauss_points_interval, src/gauss_point.rs:39-61 (literals copied: they fix the p=2 tables bit for bit)
bool interval_rule(int n, std::vector<double>& x, std::vector<double>& w) {
  switch (n) {
    case 3:
      x = {-0.7745966692414834, 0.0, 0.7745966692414834};
      w = {0.5555555555555556, 0.8888888888888889, 0.5555555555555556};
      return true;
    case 4:
      x = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
      w = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};
      return true;
    case 5:
      x = {-0.9061798459386640, -0.5384693101056831, 0.0, 0.5384693101056831, 0.9061798459386640};
      w = {0.2369268850561891, 0.4786286704993665, 0.5688888888888889, 0.4786286704993665, 0.2369268850561891};
      return true;
    default:
      return false;  // reference: panic!("Number of points not supported")
  }
}

// gauss_points_triangle, src/gauss_point.rs:62-95 (4- and 6-point literals), plus two extension
// rules for p = 3, 4: the 16-point degree-8 symmetric rule (Dunavant 1985, orbit structure
// 1 + 3x3 + 6) and the 25-point conical-product rule (Gauss-Legendre x Gauss-Jacobi(1,0), degree 9).
// Weights are normalised to sum 1, like the reference's.
bool triangle_rule(int n, std::vector<double>& p, std::vector<double>& w) {
  p.clear();
  w.clear();
  if (n == 4) {
    p = {1.0 / 3.0, 1.0 / 3.0, 2.0 / 15.0, 11.0 / 15.0, 2.0 / 15.0, 2.0 / 15.0, 11.0 / 15.0, 2.0 / 15.0};
    w = {-27.0 / 96.0, 25.0 / 96.0, 25.0 / 96.0, 25.0 / 96.0};
    return true;
  }
  if (n == 6) {
    p = {0.44594849091597, 0.44594849091597, 0.44594849091597, 0.10810301816807, 0.10810301816807,
         0.44594849091597, 0.09157621350977, 0.09157621350977, 0.09157621350977, 0.81684757298046,
         0.81684757298046, 0.09157621350977};
    w = {0.22338158967801, 0.22338158967801, 0.22338158967801,
         0.10995174365532, 0.10995174365532, 0.10995174365532};
    return true;
  }
  if (n == 16) {  // extension
    const double w0 = 0.14431560767764748;
    const double a[3] = {0.4592925882926314, 0.17056930775165638, 0.05054722831703524};
    const double wa[3] = {0.09509163426737452, 0.10321737053473984, 0.03245849762321117};
    const double b = 0.00839477740981607, c = 0.2631128296349502, wb = 0.02723031417439596;
    p.push_back(1.0 / 3.0);
    p.push_back(1.0 / 3.0);
    w.push_back(w0);
    for (int i = 0; i < 3; ++i) {
      const double o = 1.0 - 2.0 * a[i];
      const double xy[3][2] = {{a[i], a[i]}, {a[i], o}, {o, a[i]}};
      for (auto& q : xy) {
        p.push_back(q[0]);
        p.push_back(q[1]);
        w.push_back(wa[i]);
      }
    }
    const double d = 1.0 - b - c;
    const double xy[6][2] = {{b, c}, {c, b}, {b, d}, {d, b}, {c, d}, {d, c}};
    for (auto& q : xy) {
      p.push_back(q[0]);
      p.push_back(q[1]);
      w.push_back(wb);
    }
    return true;
  }
  if (n == 25) {  // extension: conical product, computed to 17 digits from Gauss-Legendre / Gauss-Jacobi nodes
    static const double t[25][3] = {
        {0.045042593569803738, 0.039809857051468722, 0.022930160703185036},
        {0.22157860955237921, 0.039809857051468722, 0.046322443858996684},
        {0.48009507147426567, 0.039809857051468722, 0.055057971328939519},
        {0.73861153339615204, 0.039809857051468722, 0.046322443858996684},
        {0.91514754937872755, 0.039809857051468722, 0.022930160703185036},
        {0.037621252345111204, 0.19801341787360821, 0.039608166264094719},
        {0.18507071026738944, 0.19801341787360821, 0.080014574772320920},
        {0.40099329106319592, 0.19801341787360821, 0.095103794115908110},
        {0.61691587185900232, 0.19801341787360821, 0.080014574772320920},
        {0.76436532978128058, 0.19801341787360821, 0.039608166264094719},
        {0.026364644944470925, 0.43797481024738616, 0.034683012862731393},
        {0.12969593678225411, 0.43797481024738616, 0.070065009006743465},
        {0.28101259487630692, 0.43797481024738616, 0.083277930430389975},
        {0.43232925297035968, 0.43797481024738616, 0.070065009006743465},
        {0.53566054480814285, 0.43797481024738616, 0.034683012862731393},
        {0.014285794395571387, 0.69546427335363614, 0.017510998364327658},
        {0.070276292008281713, 0.69546427335363614, 0.035374904220966938},
        {0.15226786332318193, 0.69546427335363614, 0.042045934974644164},
        {0.23425943463808213, 0.69546427335363614, 0.035374904220966938},
        {0.29024993225079243, 0.69546427335363614, 0.017510998364327658},
        {0.0046222884650464298, 0.90146491420117358, 0.0037311043337556805},
        {0.022738483063764033, 0.90146491420117358, 0.0075374033906552525},
        {0.049267542899413208, 0.90146491420117358, 0.0089588135945627320},
        {0.075796602735062377, 0.90146491420117358, 0.0075374033906552525},
        {0.093912797333779982, 0.90146491420117358, 0.0037311043337556805}};
    for (auto& q : t) {
      p.push_back(q[0]);
      p.push_back(q[1]);
      w.push_back(q[2]);
    }
    return true;
  }
  return false;
}

// ---------------------------------------------------------------------------------------
// Dubiner basis — src/basis_function.rs
// ---------------------------------------------------------------------------------------

// __powidf2 (compiler-rt), which Rust's f64::powi with a run-time exponent lowers to.
double powi_rt(double a, int b) {
  const bool recip = b < 0;
  double r = 1;
  while (true) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1 / r : r;
}

// jacobi_poly, src/basis_function.rs:56-74
double jacobi_poly(int n, double alpha, double beta, double x) {
  if (n == 0) return 1.0;
  if (n == 1) return (alpha + 1.0) + (alpha + beta + 2.0) * (x - 1.0) / 2.0;
  const double nf = (double)n;
  const double a1 = 2.0 * nf * (nf + alpha + beta) * (2.0 * nf + alpha + beta - 2.0);
  const double a2 = 2.0 * nf + alpha + beta - 1.0;
  const double a3 = (2.0 * nf + alpha + beta) * (2.0 * nf + alpha + beta - 2.0) * x + alpha * alpha - beta * beta;
  const double a4 = 2.0 * (nf + alpha - 1.0) * (nf + beta - 1.0) * (2.0 * nf + alpha + beta);
  const double pn_1 = jacobi_poly(n - 1, alpha, beta, x);
  const double pn_2 = jacobi_poly(n - 2, alpha, beta, x);
  return (a2 * a3 * pn_1 - a4 * pn_2) / a1;
}

// dubiner_basis, src/basis_function.rs:75-79
double dubiner_basis(int l, int k, double xi, double eta) {
  const double legendre = jacobi_poly(k, 0.0, 0.0, 2.0 * xi / (1.0 - eta) - 1.0);
  const double jacobi = jacobi_poly(l, 2.0 * (double)k + 1.0, 0.0, 2.0 * eta - 1.0);
  return powi_rt(2.0, k) * legendre * powi_rt(1.0 - eta, k) * jacobi;
}

// Mode order of src/basis_function.rs:86-92: by total degree d, then l = 0..d (k = d - l).
void mode_list(int p, std::vector<int>& L, std::vector<int>& K) {
  L.clear();
  K.clear();
  for (int d = 0; d <= p; ++d)
    for (int l = 0; l <= d; ++l) {
      L.push_back(l);
      K.push_back(d - l);
    }
}

// Extension (p != 2): reference-element derivatives of psi_{l,k} from its exact monomial expansion.
// poly[a][b] is the coefficient of xi^a eta^b. Pinned by the p = 2 hard-coded tables of
// src/basis_function.rs:114-193, which it reproduces, and by finite differences of dubiner_basis.
typedef std::vector<std::vector<double>> Poly2;

std::vector<double> jacobi_coeffs(int n, double alpha, double beta) {  // monomials in x
  std::vector<std::vector<double>> P(n + 1);
  P[0] = {1.0};
  if (n >= 1) P[1] = {(alpha + 1.0) - (alpha + beta + 2.0) / 2.0, (alpha + beta + 2.0) / 2.0};
  for (int m = 2; m <= n; ++m) {
    const double nf = m;
    const double a1 = 2.0 * nf * (nf + alpha + beta) * (2.0 * nf + alpha + beta - 2.0);
    const double a2 = 2.0 * nf + alpha + beta - 1.0;
    const double c1 = (2.0 * nf + alpha + beta) * (2.0 * nf + alpha + beta - 2.0);
    const double c0 = alpha * alpha - beta * beta;
    const double a4 = 2.0 * (nf + alpha - 1.0) * (nf + beta - 1.0) * (2.0 * nf + alpha + beta);
    std::vector<double> r(m + 1, 0.0);
    for (size_t j = 0; j < P[m - 1].size(); ++j) {
      r[j + 1] += a2 * c1 * P[m - 1][j] / a1;
      r[j] += a2 * c0 * P[m - 1][j] / a1;
    }
    for (size_t j = 0; j < P[m - 2].size(); ++j) r[j] -= a4 * P[m - 2][j] / a1;
    P[m] = r;
  }
  return P[n];
}

Poly2 poly_zero(int P) { return Poly2(P + 1, std::vector<double>(P + 1, 0.0)); }
Poly2 poly_mul(const Poly2& a, const Poly2& b, int P) {
  Poly2 r = poly_zero(P);
  for (int i = 0; i <= P; ++i)
    for (int j = 0; j <= P; ++j) {
      if (a[i][j] == 0.0) continue;
      for (int k = 0; i + k <= P; ++k)
        for (int l = 0; j + l <= P; ++l) r[i + k][j + l] += a[i][j] * b[k][l];
    }
  return r;
}
Poly2 poly_pow(const Poly2& a, int n, int P) {
  Poly2 r = poly_zero(P);
  r[0][0] = 1.0;
  for (int i = 0; i < n; ++i) r = poly_mul(r, a, P);
  return r;
}

Poly2 dubiner_poly(int l, int k, int P) {
  // 2^k (1-eta)^k P_k((2xi - (1-eta))/(1-eta)) = 2^k sum_j c_j (2xi + eta - 1)^j (1-eta)^(k-j)
  Poly2 s = poly_zero(P), om = poly_zero(P);
  s[1][0] = 2.0;
  s[0][1] = 1.0;
  s[0][0] = -1.0;
  om[0][0] = 1.0;
  om[0][1] = -1.0;
  const std::vector<double> cl = jacobi_coeffs(k, 0.0, 0.0);
  Poly2 A = poly_zero(P);
  for (int j = 0; j <= k; ++j) {
    Poly2 t = poly_mul(poly_pow(s, j, P), poly_pow(om, k - j, P), P);
    for (int a = 0; a <= P; ++a)
      for (int b = 0; b <= P; ++b) A[a][b] += std::ldexp(cl[j], k) * t[a][b];
  }
  // P_l^(2k+1,0)(2 eta - 1)
  const std::vector<double> cj = jacobi_coeffs(l, 2.0 * k + 1.0, 0.0);
  Poly2 x = poly_zero(P), B = poly_zero(P);
  x[0][1] = 2.0;
  x[0][0] = -1.0;
  for (int j = 0; j <= l; ++j) {
    Poly2 t = poly_pow(x, j, P);
    for (int a = 0; a <= P; ++a)
      for (int b = 0; b <= P; ++b) B[a][b] += cj[j] * t[a][b];
  }
  return poly_mul(A, B, P);
}

double poly_deriv_eval(const Poly2& c, int P, int da, int db, double xi, double eta) {
  double s = 0.0;
  for (int a = da; a <= P; ++a)
    for (int b = db; b <= P; ++b) {
      if (c[a][b] == 0.0) continue;
      double f = c[a][b];
      for (int i = 0; i < da; ++i) f *= (double)(a - i);
      for (int i = 0; i < db; ++i) f *= (double)(b - i);
      s += f * std::pow(xi, a - da) * std::pow(eta, b - db);
    }
  return s;
}

// Reference-element derivative tables of src/basis_function.rs:114-193 (p = 2, hard-coded in the
// reference, restated literally). Keys (1,0),(0,1),(2,0),(1,1),(0,2) → slots 0..4.
double dphi_p2(int key, int ibasis, double xi, double eta) {
  static const double c[6][5] = {{0, 0, 0, 0, 0},     {4.0, 2.0, 0, 0, 0},    {0, 3.0, 0, 0, 0},
                                 {0, 0, 48.0, 24.0, 8.0}, {0, 0, 0.0, 20.0, 20.0}, {0, 0, 0.0, 0.0, 20.0}};
  if (ibasis == 3 && key == 0) return 48.0 * xi + 24.0 * eta - 24.0;
  if (ibasis == 3 && key == 1) return 24.0 * xi + 8.0 * eta - 8.0;
  if (ibasis == 4 && key == 0) return 20.0 * eta - 4.0;
  if (ibasis == 4 && key == 1) return 20.0 * xi + 20.0 * eta - 12.0;
  if (ibasis == 5 && key == 0) return 0.0;
  if (ibasis == 5 && key == 1) return 20.0 * eta - 8.0;
  return c[ibasis][key];
}

struct Tables {
  int p, np, nq, ne, nkeys, maxd;
  std::vector<double> cell_pts, cell_w, edge_pts, edge_w;  // edge_pts [3][ne][2]
  std::vector<double> phi_cell, phi_edge, mass;            // [nq][np], [3][ne][np], [np]
  std::vector<int> key_a, key_b;                           // derivative keys (a,b)
  std::vector<int> key_of;                                 // [(maxd+1)^2] -> key slot
  std::vector<double> dref;                                // [nkeys][nq][np]
  int key_index(int a, int b) const { return key_of[a * (maxd + 1) + b]; }
};

bool build_tables(Tables& T, int p, int nq_req, int ne_req) {
  T.p = p;
  T.np = (p + 1) * (p + 2) / 2;
  std::vector<double> ip;
  if (!triangle_rule(nq_req, T.cell_pts, T.cell_w)) return false;
  if (!interval_rule(ne_req, ip, T.edge_w)) return false;
  T.nq = nq_req;
  T.ne = ne_req;
  // GaussPoints::new, src/gauss_point.rs:13-37: local edges e0 (0,0)->(1,0), e1 (1,0)->(0,1), e2 (0,1)->(0,0)
  T.edge_pts.assign(3 * T.ne * 2, 0.0);
  for (int i = 0; i < T.ne; ++i) {
    const double x = ip[i];
    T.edge_pts[(0 * T.ne + i) * 2 + 0] = (x + 1.0) / 2.0;
    T.edge_pts[(0 * T.ne + i) * 2 + 1] = 0.0;
    T.edge_pts[(1 * T.ne + i) * 2 + 0] = (1.0 - x) / 2.0;
    T.edge_pts[(1 * T.ne + i) * 2 + 1] = (1.0 + x) / 2.0;
    T.edge_pts[(2 * T.ne + i) * 2 + 0] = 0.0;
    T.edge_pts[(2 * T.ne + i) * 2 + 1] = (ip[T.ne - 1 - i] + 1.0) / 2.0;
  }
  std::vector<int> L, K;
  mode_list(p, L, K);
  // compute_phi_cell / compute_phi_edge, src/basis_function.rs:80-113
  T.phi_cell.assign(T.nq * T.np, 0.0);
  for (int i = 0; i < T.nq; ++i)
    for (int b = 0; b < T.np; ++b)
      T.phi_cell[i * T.np + b] = dubiner_basis(L[b], K[b], T.cell_pts[2 * i], T.cell_pts[2 * i + 1]);
  T.phi_edge.assign(3 * T.ne * T.np, 0.0);
  for (int e = 0; e < 3; ++e)
    for (int i = 0; i < T.ne; ++i)
      for (int b = 0; b < T.np; ++b)
        T.phi_edge[(e * T.ne + i) * T.np + b] =
            dubiner_basis(L[b], K[b], T.edge_pts[(e * T.ne + i) * 2], T.edge_pts[(e * T.ne + i) * 2 + 1]);
  // derivative keys: src/basis_function.rs:114-153 stores (1,0),(0,1),(2,0),(1,1),(0,2) at p = 2
  T.maxd = p > 2 ? p : 2;
  T.key_a.clear();
  T.key_b.clear();
  for (int d = 1; d <= T.maxd; ++d)
    for (int a = d; a >= 0; --a) {
      T.key_a.push_back(a);
      T.key_b.push_back(d - a);
    }
  T.nkeys = (int)T.key_a.size();
  T.key_of.assign((T.maxd + 1) * (T.maxd + 1), -1);
  for (int i = 0; i < T.nkeys; ++i) T.key_of[T.key_a[i] * (T.maxd + 1) + T.key_b[i]] = i;
  T.dref.assign((size_t)T.nkeys * T.nq * T.np, 0.0);
  std::vector<Poly2> polys;
  if (p != 2)
    for (int b = 0; b < T.np; ++b) polys.push_back(dubiner_poly(L[b], K[b], T.maxd));
  for (int key = 0; key < T.nkeys; ++key)
    for (int i = 0; i < T.nq; ++i)
      for (int b = 0; b < T.np; ++b) {
        const double xi = T.cell_pts[2 * i], eta = T.cell_pts[2 * i + 1];
        double v;
        if (p == 2)
          v = dphi_p2(key, b, xi, eta);
        else
          v = poly_deriv_eval(polys[b], T.maxd, T.key_a[key], T.key_b[key], xi, eta);
        T.dref[((size_t)key * T.nq + i) * T.np + b] = v;
      }
  // compute_mass_mat, src/basis_function.rs:194-203
  T.mass.assign(T.np, 0.0);
  for (int b = 0; b < T.np; ++b)
    for (int i = 0; i < T.nq; ++i) {
      const double f = T.phi_cell[i * T.np + b];
      T.mass[b] += T.cell_w[i] * (f * f);
    }
  return true;
}

// ---------------------------------------------------------------------------------------
// Mesh — src/mesh.rs types, src/initialization.rs:112-261 process_mesh, src/mesh.rs:15-147 geometry
// ---------------------------------------------------------------------------------------

struct Edge {
  int64_t ielements[2];
  int64_t ivertices[2];
  double jacob_det;
  double normal[2];
  int64_t in_cell_indices[2];
  int64_t ipatch;
};
struct Element {
  int64_t ivertices[3];
  int64_t iedges[3];
  int64_t ineighbours[3];
  double jacob_det, circumradius, minimum_height;
  double dxi_dx, dxi_dy, deta_dx, deta_dy;  // metric of src/mesh.rs:90-97
};
struct Patch {
  int64_t first, last;
  int type;  // 0 wall, 1 farfield
  double rho, u, v, p;
};

struct Case {
  ref_params prm;
  Tables T;
  std::vector<double> vx, vy;
  std::vector<double> elem_xy;  // optional per-element coordinates [ne][3][2]
  std::vector<Edge> edges;
  std::vector<Element> elements;
  std::vector<Patch> patches;
  std::vector<int64_t> internal_edges, internal_elements, boundary_elements;
  double cfl;
  std::vector<double> tmp0, tmp1;  // TemperalDisc::temporary_solutions
  int64_t bad_index = -1;          // element/edge that made the last fold fail

  size_t nstate() const { return elements.size() * 4 * (size_t)T.np; }
  size_t idx(int64_t e, int v, int k) const { return ((size_t)e * 4 + v) * T.np + k; }

  void elem_vertex(int64_t e, int i, double& x, double& y) const {
    if (!elem_xy.empty()) {
      x = elem_xy[(e * 3 + i) * 2];
      y = elem_xy[(e * 3 + i) * 2 + 1];
    } else {
      x = vx[elements[e].ivertices[i]];
      y = vy[elements[e].ivertices[i]];
    }
  }
  // Edge end points. Interior edges of periodic meshes take the left element's own coordinates
  // (extension); otherwise the vertex coordinates, as in src/mesh.rs:15-33.
  void edge_vertices(const Edge& ed, double& x1, double& y1, double& x2, double& y2) const {
    if (!elem_xy.empty() && ed.ielements[1] != -1) {
      const int64_t e = ed.ielements[0];
      const int lic = (int)ed.in_cell_indices[0];
      elem_vertex(e, lic, x1, y1);
      elem_vertex(e, (lic + 1) % 3, x2, y2);
    } else {
      x1 = vx[ed.ivertices[0]];
      y1 = vy[ed.ivertices[0]];
      x2 = vx[ed.ivertices[1]];
      y2 = vy[ed.ivertices[1]];
    }
  }

  // physical derivative tables of src/mesh.rs:85-132 (recomputed on demand with the same
  // expressions the reference stores per element; order >= 3 is the extension's binomial expansion)
  double dphys(int64_t e, int igp, int k, int l1, int l2) const {
    const Element& el = elements[e];
    const int nq = T.nq, np = T.np;
    auto D = [&](int a, int b) { return T.dref[((size_t)T.key_index(a, b) * nq + igp) * np + k]; };
    const double xx = el.dxi_dx, xy = el.dxi_dy, ex = el.deta_dx, ey = el.deta_dy;
    const int n = l1 + l2;
    if (n == 1) {
      if (l1 == 1) return xx * D(1, 0) + ex * D(0, 1);
      return xy * D(1, 0) + ey * D(0, 1);
    }
    if (n == 2) {
      if (l1 == 2) return xx * xx * D(2, 0) + 2.0 * xx * ex * D(1, 1) + ex * ex * D(0, 2);
      if (l1 == 1) return xx * xy * D(2, 0) + (xx * ey + ex * xy) * D(1, 1) + ex * ey * D(0, 2);
      return xy * xy * D(2, 0) + 2.0 * xy * ey * D(1, 1) + ey * ey * D(0, 2);
    }
    // extension: d^a/dx^a d^b/dy^b = (xx d_xi + ex d_eta)^a (xy d_xi + ey d_eta)^b
    double s = 0.0;
    auto binom = [](int m, int r) {
      double c = 1.0;
      for (int i = 1; i <= r; ++i) c = c * (m - r + i) / i;
      return c;
    };
    for (int i = 0; i <= l1; ++i)
      for (int j = 0; j <= l2; ++j) {
        const double c = binom(l1, i) * binom(l2, j) * std::pow(xx, i) * std::pow(ex, l1 - i) * std::pow(xy, j) *
                         std::pow(ey, l2 - j);
        s += c * D(i + j, n - i - j);
      }
    return s;
  }
};

// process_mesh, src/initialization.rs:112-261 (edge numbering: boundary edges first in block
// order, then interior edges in first-visit element order; left = element whose local edge runs
// ivertices[0] -> ivertices[1]).
int process_mesh(Case& C, int64_t neb, const int32_t* eb_type, const int64_t* eb_count, const int64_t* eb_nodes,
                 int64_t nelb, const int64_t* el_count, const int64_t* el_nodes) {
  const int64_t nv = (int64_t)C.vx.size();
  auto key = [nv](int64_t a, int64_t b) { return (uint64_t)a * (uint64_t)nv + (uint64_t)b; };
  std::unordered_map<uint64_t, int64_t> edge_map;
  // BoundaryQuantity from FreeStreamCondition, src/initialization.rs:124-133
  const ref_params& P = C.prm;
  const double sound_speed = std::sqrt(P.fs_hcr * P.fs_p / P.fs_rho);
  const double rad = P.fs_aoa_deg * (M_PI / 180.0);  // f64::to_radians
  int64_t off = 0;
  for (int64_t b = 0; b < neb; ++b) {
    Patch pt;
    if (eb_type[b] == 0) {
      pt.type = 0;
      pt.rho = pt.u = pt.v = pt.p = 0.0;
    } else if (eb_type[b] == 1) {
      pt.type = 1;
      pt.rho = P.fs_rho;
      pt.u = P.fs_mach * sound_speed * std::cos(rad);
      pt.v = P.fs_mach * sound_speed * std::sin(rad);
      pt.p = P.fs_p;
    } else {
      return BAD_MESH;  // reference: panic!("Boundary type not implemented")
    }
    pt.first = (int64_t)C.edges.size();
    for (int64_t i = 0; i < eb_count[b]; ++i) {
      Edge ed;
      ed.ivertices[0] = eb_nodes[2 * (off + i)];
      ed.ivertices[1] = eb_nodes[2 * (off + i) + 1];
      ed.normal[0] = ed.normal[1] = 0.0;
      ed.jacob_det = 0.0;
      ed.ielements[0] = 0;
      ed.ielements[1] = -1;
      ed.in_cell_indices[0] = 0;
      ed.in_cell_indices[1] = -1;
      ed.ipatch = (int64_t)C.patches.size();
      C.edges.push_back(ed);
      edge_map[key(ed.ivertices[0], ed.ivertices[1])] = (int64_t)C.edges.size() - 1;
    }
    off += eb_count[b];
    pt.last = (int64_t)C.edges.size() - 1;
    C.patches.push_back(pt);
  }
  int64_t eoff = 0;
  for (int64_t blk = 0; blk < nelb; ++blk) {
    const int64_t n = el_count[blk];
    const int64_t* nodes = el_nodes + 3 * eoff;
    // create edges from connectivities of elements (:262-291)
    for (int64_t t = 0; t < n; ++t) {
      for (int i = 0; i < 3; ++i) {
        const int64_t a = nodes[3 * t + i], b = nodes[3 * t + (i + 1) % 3];
        if (edge_map.find(key(a, b)) == edge_map.end() && edge_map.find(key(b, a)) == edge_map.end()) {
          Edge ed;
          ed.ivertices[0] = a;
          ed.ivertices[1] = b;
          ed.normal[0] = ed.normal[1] = 0.0;
          ed.jacob_det = 0.0;
          ed.ielements[0] = ed.ielements[1] = 0;
          ed.in_cell_indices[0] = ed.in_cell_indices[1] = 0;
          ed.ipatch = -1;
          C.edges.push_back(ed);
          C.internal_edges.push_back((int64_t)C.edges.size() - 1);
          edge_map[key(a, b)] = (int64_t)C.edges.size() - 1;
        }
      }
    }
    // assign edges to elements (:292-350)
    for (int64_t t = 0; t < n; ++t) {
      Element el;
      bool is_internal = true;
      for (int i = 0; i < 3; ++i) {
        const int64_t a = nodes[3 * t + i], b = nodes[3 * t + (i + 1) % 3];
        el.ivertices[i] = a;
        auto it = edge_map.find(key(a, b));
        if (it == edge_map.end()) it = edge_map.find(key(b, a));
        el.iedges[i] = it->second;
      }
      const int64_t me = (int64_t)C.elements.size();
      for (int i = 0; i < 3; ++i) {
        Edge& ed = C.edges[el.iedges[i]];
        if (ed.ielements[1] == -1) {
          ed.ielements[0] = me;
          ed.in_cell_indices[0] = i;
          is_internal = false;
        } else if (el.ivertices[i] == ed.ivertices[0]) {
          ed.ielements[0] = me;
          ed.in_cell_indices[0] = i;
        } else {
          ed.ielements[1] = me;
          ed.in_cell_indices[1] = i;
        }
      }
      el.ineighbours[0] = el.ineighbours[1] = el.ineighbours[2] = -1;
      el.jacob_det = el.circumradius = el.minimum_height = 0.0;
      C.elements.push_back(el);
      if (is_internal)
        C.internal_elements.push_back(me);
      else
        C.boundary_elements.push_back(me);
    }
    eoff += n;
  }
  return OK;
}

// Mesh::compute_jacob_det, compute_normal, compute_dphi (metric), compute_circumradius,
// compute_minimum_height, set_neighbours — src/mesh.rs:15-147, in the order of src/initialization.rs:59-65.
void compute_geometry(Case& C) {
  for (Edge& ed : C.edges) {  // :25-33
    double x1, y1, x2, y2;
    C.edge_vertices(ed, x1, y1, x2, y2);
    ed.jacob_det = std::sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
  }
  for (size_t e = 0; e < C.elements.size(); ++e) {  // :34-41
    double x1, y1, x2, y2, x3, y3;
    C.elem_vertex(e, 0, x1, y1);
    C.elem_vertex(e, 1, x2, y2);
    C.elem_vertex(e, 2, x3, y3);
    C.elements[e].jacob_det = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
  }
  for (Edge& ed : C.edges) {  // :15-24
    double x1, y1, x2, y2;
    C.edge_vertices(ed, x1, y1, x2, y2);
    ed.normal[0] = (y2 - y1) / ed.jacob_det;
    ed.normal[1] = -(x2 - x1) / ed.jacob_det;
  }
  for (size_t e = 0; e < C.elements.size(); ++e) {
    Element& el = C.elements[e];
    double x1, y1, x2, y2, x3, y3;
    C.elem_vertex(e, 0, x1, y1);
    C.elem_vertex(e, 1, x2, y2);
    C.elem_vertex(e, 2, x3, y3);
    // :85-100
    const double dx_dxi = x2 - x1, dx_deta = x3 - x1, dy_dxi = y2 - y1, dy_deta = y3 - y1;
    el.dxi_dx = dy_deta / el.jacob_det;
    el.dxi_dy = -dx_deta / el.jacob_det;
    el.deta_dx = -dy_dxi / el.jacob_det;
    el.deta_dy = dx_dxi / el.jacob_det;
    // :43-57
    {
      const double a = std::sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
      const double b = std::sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2));
      const double c = std::sqrt((x1 - x3) * (x1 - x3) + (y1 - y3) * (y1 - y3));
      const double s = (a + b + c) / 2.0;
      el.circumradius = a * b * c / (4.0 * std::sqrt(s * (s - a) * (s - b) * (s - c)));
    }
    // :58-84
    {
      const double surface = 0.5 * el.jacob_det;
      const double a = std::sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
      const double b = std::sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2));
      const double c = std::sqrt((x1 - x3) * (x1 - x3) + (y1 - y3) * (y1 - y3));
      el.minimum_height = std::fmin(std::fmin(2.0 * surface / a, 2.0 * surface / b), 2.0 * surface / c);
    }
  }
  for (Element& el : C.elements) {  // set_neighbours :133-147
    for (int i = 0; i < 3; ++i) {
      const Edge& ed = C.edges[el.iedges[i]];
      if (ed.ielements[1] == -1) continue;
      if (el.ivertices[i] == ed.ivertices[0])
        el.ineighbours[i] = ed.ielements[1];
      else
        el.ineighbours[i] = ed.ielements[0];
    }
  }
}

// ---------------------------------------------------------------------------------------
// Fluxes — src/spatial_disc/flux.rs, src/spatial_disc/local_characteristics.rs
// ---------------------------------------------------------------------------------------

// flux::flux, src/spatial_disc/flux.rs:154-168
inline void euler_flux(const double* q, double hcr, double* f, double* g) {
  const double u = q[1] / q[0];
  const double v = q[2] / q[0];
  const double p = (hcr - 1.0) * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
  f[0] = q[1];
  f[1] = q[1] * u + p;
  f[2] = q[2] * u;
  f[3] = u * (q[3] + p);
  g[0] = q[2];
  g[1] = q[1] * v;
  g[2] = q[2] * v + p;
  g[3] = v * (q[3] + p);
}

// flux::hllc, src/spatial_disc/flux.rs:28-153 (including the altered p* numerator of :75-78)
int hllc(const double* lv, const double* rv, const double* normal, double hcr, double* flux) {
  const double nx = normal[0], ny = normal[1];
  const double ql[4] = {lv[0], lv[1] * nx + lv[2] * ny, -lv[1] * ny + lv[2] * nx, lv[3]};
  const double qr[4] = {rv[0], rv[1] * nx + rv[2] * ny, -rv[1] * ny + rv[2] * nx, rv[3]};
  const double ul = ql[1] / ql[0];
  const double ur = qr[1] / qr[0];
  const double vl = ql[2] / ql[0];
  const double vr = qr[2] / qr[0];
  const double pl = (hcr - 1.0) * (ql[3] - 0.5 * (ql[1] * ql[1] + ql[2] * ql[2]) / ql[0]);
  const double pr = (hcr - 1.0) * (qr[3] - 0.5 * (qr[1] * qr[1] + qr[2] * qr[2]) / qr[0]);
  if (ql[0] < 0.0 || qr[0] < 0.0) return NEG_DENSITY;
  if (pl < 0.0 || pr < 0.0) return NEG_PRESSURE;
  const double cl = std::sqrt(hcr * pl / ql[0]);
  const double cr = std::sqrt(hcr * pr / qr[0]);
  double fl[4], fr[4];
  fl[0] = ql[1];
  fl[1] = ql[1] * ul + pl;
  fl[2] = ql[2] * ul;
  fl[3] = ul * (ql[3] + pl);
  fr[0] = qr[1];
  fr[1] = qr[1] * ur + pr;
  fr[2] = qr[2] * ur;
  fr[3] = ur * (qr[3] + pr);
  const double zeta = (hcr - 1.0) / (2.0 * hcr);
  const double pcoef = kPlant == PLANT_HLLC_TEXTBOOK_PSTAR ? (hcr - 1.0) / 2.0 : (zeta - 1.0) / 2.0;
  const double p_star =
      std::pow((cl + cr - pcoef * (ur - ul)) / (cl / std::pow(pl, zeta) + cr / std::pow(pr, zeta)), 1.0 / zeta);
  double ql_lower, qr_lower;
  if (p_star <= pl)
    ql_lower = 1.0;
  else
    ql_lower = std::sqrt(1.0 + (hcr + 1.0) / (2.0 * hcr) * (p_star / pl - 1.0));
  const double sl = ul - cl * ql_lower;
  if (p_star <= pr)
    qr_lower = 1.0;
  else
    qr_lower = std::sqrt(1.0 + (hcr + 1.0) / (2.0 * hcr) * (p_star / pr - 1.0));
  const double sr = ur + cr * qr_lower;
  if (sl >= 0.0) {
    flux[0] = fl[0];
    flux[1] = fl[1] * nx - fl[2] * ny;
    flux[2] = fl[1] * ny + fl[2] * nx;
    flux[3] = fl[3];
  } else if (sr <= 0.0) {
    flux[0] = fr[0];
    flux[1] = fr[1] * nx - fr[2] * ny;
    flux[2] = fr[1] * ny + fr[2] * nx;
    flux[3] = fr[3];
  } else {
    const double s_star = (pr - pl + ql[0] * ul * (sl - ul) - qr[0] * ur * (sr - ur)) /
                          (ql[0] * (sl - ul) - qr[0] * (sr - ur));
    if (s_star >= 0.0) {
      double q_star[4];
      q_star[0] = ql[0] * (sl - ul) / (sl - s_star);
      q_star[1] = ql[0] * (sl - ul) / (sl - s_star) * s_star;
      q_star[2] = ql[0] * (sl - ul) / (sl - s_star) * vl;
      q_star[3] = ql[0] * (sl - ul) / (sl - s_star) * (ql[3] / ql[0] + (s_star - ul) * (s_star + pl / (ql[0] * (sl - ul))));
      flux[0] = fl[0] + sl * (q_star[0] - ql[0]);
      const double fn = fl[1] + sl * (q_star[1] - ql[1]);
      const double ft = fl[2] + sl * (q_star[2] - ql[2]);
      flux[1] = fn * nx - ft * ny;
      flux[2] = fn * ny + ft * nx;
      flux[3] = fl[3] + sl * (q_star[3] - ql[3]);
    } else {
      double q_star[4];
      q_star[0] = qr[0] * (sr - ur) / (sr - s_star);
      q_star[1] = qr[0] * (sr - ur) / (sr - s_star) * s_star;
      q_star[2] = qr[0] * (sr - ur) / (sr - s_star) * vr;
      q_star[3] = qr[0] * (sr - ur) / (sr - s_star) * (qr[3] / qr[0] + (s_star - ur) * (s_star + pr / (qr[0] * (sr - ur))));
      flux[0] = fr[0] + sr * (q_star[0] - qr[0]);
      const double fn = fr[1] + sr * (q_star[1] - qr[1]);
      const double ft = fr[2] + sr * (q_star[2] - qr[2]);
      flux[1] = fn * nx - ft * ny;
      flux[2] = fn * ny + ft * nx;
      flux[3] = fr[3] + sr * (q_star[3] - qr[3]);
    }
  }
  return OK;
}

// local_characteristics::compute_eigenmatrix, src/spatial_disc/local_characteristics.rs:4-50
void eigenmatrix(const double* q, double nx, double ny, double hcr, double* Lm, double* Rm) {
  const double u = q[1] / q[0];
  const double v = q[2] / q[0];
  const double p = (hcr - 1.0) * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
  const double c = std::sqrt(hcr * p / q[0]);
  const double h = (q[3] + p) / q[0];
  const double b1 = (hcr - 1.0) / (c * c);
  const double b2 = b1 * (u * u + v * v) / 2.0;
  Lm[0] = (b2 + (u * nx + v * ny) / c) / 2.0;
  Lm[1] = -(b1 * u + nx / c) / 2.0;
  Lm[2] = -(b1 * v + ny / c) / 2.0;
  Lm[3] = b1 / 2.0;
  Lm[4] = ny * u - nx * v;
  Lm[5] = -ny;
  Lm[6] = nx;
  Lm[7] = 0.0;
  Lm[8] = 1.0 - b2;
  Lm[9] = b1 * u;
  Lm[10] = b1 * v;
  Lm[11] = -b1;
  Lm[12] = (b2 - (u * nx + v * ny) / c) / 2.0;
  Lm[13] = -(b1 * u - nx / c) / 2.0;
  Lm[14] = -(b1 * v - ny / c) / 2.0;
  Lm[15] = b1 / 2.0;
  Rm[0] = 1.0;
  Rm[1] = 0.0;
  Rm[2] = 1.0;
  Rm[3] = 1.0;
  Rm[4] = u - c * nx;
  Rm[5] = -ny;
  Rm[6] = u;
  Rm[7] = u + c * nx;
  Rm[8] = v - c * ny;
  Rm[9] = nx;
  Rm[10] = v;
  Rm[11] = v + c * ny;
  Rm[12] = h - c * (u * nx + v * ny);
  Rm[13] = -u * ny + v * nx;
  Rm[14] = (u * u + v * v) / 2.0;
  Rm[15] = h + c * (u * nx + v * ny);
}

// ndarray 2-D x 2-D `dot` on f64 (matrixmultiply dgemm): 4x4 times 4xn, C-order.
void gemm4(const double* A, const double* B, int n, double* out, bool use_fma) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < n; ++j) {
      double s = A[i * 4 + 0] * B[0 * n + j];
      for (int k = 1; k < 4; ++k) s = use_fma ? std::fma(A[i * 4 + k], B[k * n + j], s) : s + A[i * 4 + k] * B[k * n + j];
      out[i * n + j] = s;
    }
}
// ndarray 2-D x 1-D `dot` (unrolled_dot with 4 terms: a sequential sum from 0).
void gemv4(const double* A, const double* x, double* y) {
  for (int i = 0; i < 4; ++i) {
    double s = 0.0;
    for (int k = 0; k < 4; ++k) s = s + A[i * 4 + k] * x[k];
    y[i] = s;
  }
}

// ---------------------------------------------------------------------------------------
// Residual — src/spatial_disc.rs, src/spatial_disc/boundary_condition.rs
// ---------------------------------------------------------------------------------------

int nthreads_of(const Case& C) { return C.prm.nthreads > 0 ? C.prm.nthreads : 1; }

// rayon par_iter().fold(zeros).reduce(+) over an index list, then `*residuals = &*residuals + sum`
// (src/spatial_disc.rs:46,68). Each thread folds a contiguous chunk into its own full-size array.
template <class F>
int fold_over(Case& C, const std::vector<int64_t>& list, double* R, F body) {
  const int nt = nthreads_of(C);
  const size_t ns = C.nstate();
  std::vector<std::vector<double>> acc(nt);
  std::vector<int> st(nt, OK);
  std::vector<int64_t> bad(nt, -1);
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
    acc[tid].assign(ns, 0.0);
    const size_t n = list.size();
    const size_t lo = n * tid / nt, hi = n * (tid + 1) / nt;
    for (size_t i = lo; i < hi; ++i) {
      const int s = body(list[i], acc[tid].data());
      if (s != OK && st[tid] == OK) {
        st[tid] = s;
        bad[tid] = list[i];
      }
    }
  }
  std::vector<double> sum(ns, 0.0);
  for (int t = 0; t < nt; ++t)
    for (size_t i = 0; i < ns; ++i) sum[i] = sum[i] + acc[t][i];
  for (size_t i = 0; i < ns; ++i) R[i] = R[i] + sum[i];
  for (int t = 0; t < nt; ++t)
    if (st[t] != OK) {
      C.bad_index = bad[t];
      return st[t];
    }
  return OK;
}

// volume integral of one element: src/spatial_disc.rs:47-66 (also src/spatial_disc/boundary_condition.rs:98-117)
inline void volume_element(const Case& C, int64_t e, const double* U, double* A) {
  const Tables& T = C.T;
  const double hcr = C.prm.hcr;
  const double J = C.elements[e].jacob_det;
  for (int igp = 0; igp < T.nq; ++igp) {
    double sol[4] = {0.0, 0.0, 0.0, 0.0};
    for (int v = 0; v < 4; ++v)
      for (int k = 0; k < T.np; ++k) sol[v] += U[C.idx(e, v, k)] * T.phi_cell[igp * T.np + k];
    double f[4], g[4];
    euler_flux(sol, hcr, f, g);
    for (int v = 0; v < 4; ++v)
      for (int k = 0; k < T.np; ++k) {
        const double dphi_dx = C.dphys(e, igp, k, 1, 0);
        const double dphi_dy = C.dphys(e, igp, k, 0, 1);
        A[C.idx(e, v, k)] += (f[v] * dphi_dx + g[v] * dphi_dy) * T.cell_w[igp] * 0.5 * J;
      }
  }
}

// integrate_over_cell_euler, src/spatial_disc.rs:40-69
int volume_terms(Case& C, double* R, const double* U) {
  return fold_over(C, C.internal_elements, R, [&](int64_t e, double* A) {
    volume_element(C, e, U, A);
    return (int)OK;
  });
}

// integrate_over_edges_euler, src/spatial_disc.rs:161-211
int edge_terms(Case& C, double* R, const double* U) {
  const Tables& T = C.T;
  const double hcr = C.prm.hcr;
  const int ngp = T.ne, np = T.np;
  return fold_over(C, C.internal_edges, R, [&](int64_t iedge, double* A) {
    const Edge& ed = C.edges[iedge];
    const int64_t ilic = ed.in_cell_indices[0], iric = ed.in_cell_indices[1];
    const int64_t L = ed.ielements[0], Rr = ed.ielements[1];
    for (int igp = 0; igp < ngp; ++igp) {
      double lv[4] = {0, 0, 0, 0}, rv[4] = {0, 0, 0, 0};
      for (int v = 0; v < 4; ++v)
        for (int k = 0; k < np; ++k) {
          lv[v] += U[C.idx(L, v, k)] * T.phi_edge[(ilic * ngp + igp) * np + k];
          rv[v] += U[C.idx(Rr, v, k)] * T.phi_edge[(iric * ngp + (ngp - 1 - igp)) * np + k];
        }
      double F[4];
      const int s = hllc(lv, rv, ed.normal, hcr, F);
      if (s != OK) return s;  // reference: panic!("Error in HLLC flux computation!")
      for (int v = 0; v < 4; ++v)
        for (int k = 0; k < np; ++k) {
          A[C.idx(L, v, k)] -= T.edge_w[igp] * (F[v]) * T.phi_edge[(ilic * ngp + igp) * np + k] * 0.5 * ed.jacob_det;
          A[C.idx(Rr, v, k)] += T.edge_w[ngp - 1 - igp] * (F[v]) *
                                T.phi_edge[(iric * ngp + (ngp - 1 - igp)) * np + k] * 0.5 * ed.jacob_det;
        }
    }
    return (int)OK;
  });
}

// compute_boundary_edges_values, src/spatial_disc/boundary_condition.rs:496-509
inline void boundary_edge_values(const Case& C, const Edge& ed, const double* U, int igp, double* out) {
  const Tables& T = C.T;
  const int64_t e = ed.ielements[0];
  const int64_t lic = ed.in_cell_indices[0];
  for (int v = 0; v < 4; ++v) {
    double s = 0.0;
    for (int k = 0; k < T.np; ++k) s += U[C.idx(e, v, k)] * T.phi_edge[(lic * T.ne + igp) * T.np + k];
    out[v] = s;
  }
}

// apply_bc_euler, src/spatial_disc/boundary_condition.rs:9-120
int boundary_terms(Case& C, double* R, const double* U) {
  const Tables& T = C.T;
  const double hcr = C.prm.hcr;
  const int ngp = T.ne, np = T.np;
  const bool use_fma = C.prm.gemm_fma != 0;
  return fold_over(C, C.boundary_elements, R, [&](int64_t e, double* A) {
    const Element& el = C.elements[e];
    for (int ilic = 0; ilic < 3; ++ilic) {
      const Edge& ed = C.edges[el.iedges[ilic]];
      if (ed.ipatch == -1) continue;
      double nx, ny;
      if (el.ivertices[ilic] == ed.ivertices[0]) {
        nx = ed.normal[0];
        ny = ed.normal[1];
      } else {
        nx = -ed.normal[0];
        ny = -ed.normal[1];
      }
      const Patch& pt = C.patches[ed.ipatch];
      if (pt.type == 0) {  // Wall :31-47
        for (int igp = 0; igp < ngp; ++igp) {
          double lv[4];
          boundary_edge_values(C, ed, U, igp, lv);
          const double pressure = (hcr - 1.0) * (lv[3] - 0.5 * (lv[1] * lv[1] + lv[2] * lv[2]) / lv[0]);
          const double bf[4] = {0.0, pressure * nx, pressure * ny, 0.0};
          for (int v = 0; v < 4; ++v)
            for (int k = 0; k < np; ++k)
              A[C.idx(e, v, k)] -= T.edge_w[igp] * bf[v] * T.phi_edge[(ilic * ngp + igp) * np + k] * 0.5 * ed.jacob_det;
        }
      } else {  // FarField :48-94
        const double cons[4] = {pt.rho, pt.rho * pt.u, pt.rho * pt.v,
                                pt.p / (hcr - 1.0) + 0.5 * pt.rho * (pt.u * pt.u + pt.v * pt.v)};
        std::vector<double> qnew_gps(ngp * 4);
        for (int igp = 0; igp < ngp; ++igp) {
          double lv[4];
          boundary_edge_values(C, ed, U, igp, lv);
          const double u = lv[1] / lv[0];
          const double v = lv[2] / lv[0];
          const double p = (hcr - 1.0) * (lv[3] - 0.5 * (lv[1] * lv[1] + lv[2] * lv[2]) / lv[0]);
          const double c = std::sqrt(hcr * p / lv[0]);
          const double vn = u * nx + v * ny;
          double Lm[16], Rm[16];
          eigenmatrix(lv, nx, ny, hcr, Lm, Rm);
          const double eig[4] = {vn - c, vn, vn, vn + c};
          double alpha[4], beta[4], qn[4], qo[4];
          gemv4(Lm, lv, alpha);                // 2-D x 1-D
          gemm4(Lm, cons, 1, beta, use_fma);   // 2-D x 2-D (4x1)
          for (int i = 0; i < 4; ++i) qn[i] = eig[i] >= 0.0 ? alpha[i] : beta[i];
          gemv4(Rm, qn, qo);
          for (int i = 0; i < 4; ++i) qnew_gps[igp * 4 + i] = qo[i];
        }
        for (int igp = 0; igp < ngp; ++igp) {
          const double* q = &qnew_gps[igp * 4];
          const double u_new = q[1] / q[0];
          const double v_new = q[2] / q[0];
          const double vn_new = u_new * nx + v_new * ny;
          const double p_new = (hcr - 1.0) * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
          const double h_new = q[3] + p_new;
          const double bf[4] = {q[0] * vn_new, q[1] * vn_new + p_new * nx, q[2] * vn_new + p_new * ny, h_new * vn_new};
          for (int v = 0; v < 4; ++v)
            for (int k = 0; k < np; ++k)
              A[C.idx(e, v, k)] -= T.edge_w[igp] * bf[v] * T.phi_edge[(ilic * ngp + igp) * np + k] * 0.5 * ed.jacob_det;
        }
      }
    }
    volume_element(C, e, U, A);  // :98-117
    return (int)OK;
  });
}

// divide_residual_by_mass_mat_diag, src/spatial_disc.rs:349-357
void divide_by_mass(const Case& C, double* R) {
  const int np = C.T.np;
  for (size_t e = 0; e < C.elements.size(); ++e)
    for (int v = 0; v < 4; ++v)
      for (int k = 0; k < np; ++k) R[C.idx(e, v, k)] /= C.T.mass[k] * 0.5 * C.elements[e].jacob_det;
}

// SpatialDisc::compute_residuals, src/spatial_disc.rs:27-39
int compute_residuals(Case& C, double* R, const double* U, int64_t* bad) {
  int s = volume_terms(C, R, U);
  if (s != OK) return s;
  s = edge_terms(C, R, U);
  if (s != OK) {
    if (bad) *bad = C.bad_index;
    return s;
  }
  s = boundary_terms(C, R, U);
  if (s != OK) return s;
  divide_by_mass(C, R);
  return OK;
}

// ---------------------------------------------------------------------------------------
// Limiter — src/spatial_disc/limiter/weno.rs
// ---------------------------------------------------------------------------------------

// modified_kxrcf, src/spatial_disc/limiter/weno.rs:7-85 (neighbour trace index quirk of :21-27)
bool modified_kxrcf(const Case& C, const double* U, int64_t ielem) {
  const Tables& T = C.T;
  const int np = T.np, ngp = T.ne;
  const Element& el = C.elements[ielem];
  double density_jump = 0.0, total_energy_jump = 0.0, surface = 0.0;
  double min_density = 1.0e10, min_total_energy = 1.0e10;
  for (int lic = 0; lic < 3; ++lic) {
    const Edge& ed = C.edges[el.iedges[lic]];
    if (ed.ielements[1] == -1) continue;
    const int64_t nb = el.ineighbours[lic];
    double nx, ny;
    int64_t nlic;
    if (el.ivertices[lic] == ed.ivertices[0]) {
      nx = ed.normal[0];
      ny = ed.normal[1];
      nlic = ed.in_cell_indices[kPlant == PLANT_KXRCF_NB_OWN_EDGE ? 1 : 0];
    } else {
      nx = -ed.normal[0];
      ny = -ed.normal[1];
      nlic = ed.in_cell_indices[kPlant == PLANT_KXRCF_NB_OWN_EDGE ? 0 : 1];
    }
    for (int igp = 0; igp < ngp; ++igp) {
      double rho = 0.0, rho_e = 0.0;
      for (int k = 0; k < np; ++k) {
        rho += U[C.idx(ielem, 0, k)] * T.phi_edge[(lic * ngp + igp) * np + k];
        rho_e += U[C.idx(ielem, 3, k)] * T.phi_edge[(lic * ngp + igp) * np + k];
      }
      if (rho < min_density) min_density = rho;
      if (rho_e < min_total_energy) min_total_energy = rho_e;
    }
    double rho_u = 0.0, rho_v = 0.0;
    const int cg = ngp / 2;
    for (int k = 0; k < np; ++k) {
      rho_u += U[C.idx(ielem, 1, k)] * T.phi_edge[(lic * ngp + cg) * np + k];
      rho_v += U[C.idx(ielem, 2, k)] * T.phi_edge[(lic * ngp + cg) * np + k];
    }
    const double center = rho_u * nx + rho_v * ny;
    if (kPlant == PLANT_KXRCF_ALL_SURFACE && !(center < 0.0)) surface += ed.jacob_det;
    if (kPlant == PLANT_KXRCF_OUTFLOW ? center > 0.0 : center < 0.0) {
      double rj = 0.0, ej = 0.0;
      for (int igp = 0; igp < ngp; ++igp) {
        double ir = 0.0, ie = 0.0, xr = 0.0, xe = 0.0;
        for (int k = 0; k < np; ++k) {
          ir += U[C.idx(ielem, 0, k)] * T.phi_edge[(lic * ngp + igp) * np + k];
          ie += U[C.idx(ielem, 3, k)] * T.phi_edge[(lic * ngp + igp) * np + k];
          xr += U[C.idx(nb, 0, k)] * T.phi_edge[(nlic * ngp + (ngp - 1 - igp)) * np + k];
          xe += U[C.idx(nb, 3, k)] * T.phi_edge[(nlic * ngp + (ngp - 1 - igp)) * np + k];
        }
        rj += 0.5 * ed.jacob_det * T.edge_w[igp] * (ir - xr);
        ej += 0.5 * ed.jacob_det * T.edge_w[igp] * (ie - xe);
      }
      density_jump += rj;
      total_energy_jump += ej;
      surface += ed.jacob_det;
    }
  }
  const double h = el.circumradius;
  const double hp = kPlant == PLANT_KXRCF_H1 ? h : std::pow(h, 1.5);
  const double ind_d = std::fabs(density_jump) / (hp * surface * min_density);
  const double ind_e = std::fabs(total_energy_jump) / (hp * surface * min_total_energy);
  return (ind_d >= 1.0) || (ind_e >= 1.0);
}

// WENO reconstruction of one troubled element, src/spatial_disc/limiter/weno.rs:93-263.
// Reads U (target and neighbours) and writes the reconstructed polynomial to out[4][np].
int weno_element(const Case& C, const double* U, int64_t ielem, double* out) {
  const Tables& T = C.T;
  const int np = T.np;
  const double hcr = C.prm.hcr;
  const bool use_fma = C.prm.gemm_fma != 0;
  const Element& el = C.elements[ielem];
  // linear weights :93-110 (boundary face i zeroes weights[i], quirk)
  double lw[4] = {0.997, 0.001, 0.001, 0.001};
  int nbnd = 0;
  for (int i = 0; i < 3; ++i)
    if (C.edges[el.iedges[i]].ielements[1] == -1) {
      nbnd += 1;
      lw[kPlant == PLANT_WENO_WEIGHT_NEXT ? i + 1 : i] = 0.0;
    }
  if (nbnd == 0)
    lw[0] = 0.997;
  else if (nbnd == 1)
    lw[1] = 0.998;
  else if (nbnd == 2)
    lw[2] = 0.999;
  else
    return BAD_MESH;  // reference: panic!("Invalid number of boundary neighbours!")
  double neighbour_area = 0.0;
  std::vector<double> tgt(4 * np), rec(4 * np, 0.0), proj(4 * 4 * np), nbp(4 * np), pn(4 * np), pn2(4 * np);
  for (int v = 0; v < 4; ++v)
    for (int k = 0; k < np; ++k) tgt[v * np + k] = U[C.idx(ielem, v, k)];
  // smoothness indicator of one projected polynomial row, :125-143
  auto beta_of = [&](const double* pol) {
    double beta = 0.0;
    for (int l = 1; l <= T.p; ++l)
      for (int l1 = 0; l1 <= l; ++l1) {
        const int l2 = l - l1;
        double integral = 0.0;
        for (int igp = 0; igp < T.nq; ++igp) {
          double value = 0.0;
          for (int k = 0; k < np; ++k) value += pol[k] * C.dphys(ielem, igp, k, l1, l2);
          integral += 0.5 * el.jacob_det * T.cell_w[igp] * (value * value);
        }
        beta += std::pow(el.jacob_det, (double)(l - 1)) * integral;
      }
    return beta;
  };
  const double EPS = 2.220446049250313e-16;  // f64::EPSILON
  for (int i = 0; i < 3; ++i) {
    const Edge& ei = C.edges[el.iedges[i]];
    if (ei.ielements[1] == -1) continue;
    const int64_t inb = el.ineighbours[i];
    neighbour_area += kPlant == PLANT_WENO_PLAIN_MEAN ? 1.0 : C.elements[inb].jacob_det * 0.5;
    double nlw[4][4] = {{0}};  // [poly j][var]
    std::fill(proj.begin(), proj.end(), 0.0);
    const double nx = ei.normal[0], ny = ei.normal[1];  // raw edge normal (:121-122)
    double mean[4] = {tgt[0 * np], tgt[1 * np], tgt[2 * np], tgt[3 * np]};
    double Lm[16], Rm[16];
    eigenmatrix(mean, nx, ny, hcr, Lm, Rm);
    gemm4(Lm, tgt.data(), np, &proj[0], use_fma);
    for (int v = 0; v < 4; ++v) {
      const double beta = beta_of(&proj[v * np]);
      nlw[0][v] = lw[0] / ((beta + EPS) * (beta + EPS));
    }
    for (int j = 0; j < 3; ++j) {
      const Edge& ej = C.edges[el.iedges[j]];
      if (ej.ielements[1] == -1) continue;
      const int64_t jnb = el.ineighbours[j];
      for (int v = 0; v < 4; ++v)
        for (int k = 0; k < np; ++k) nbp[v * np + k] = U[C.idx(jnb, v, k)];
      if (kPlant != PLANT_WENO_NO_MEAN_SWAP)
        for (int v = 0; v < 4; ++v) nbp[v * np + 0] = tgt[v * np + 0];  // :148-154
      gemm4(Lm, nbp.data(), np, &proj[(j + 1) * 4 * np], use_fma);
      for (int v = 0; v < 4; ++v) {
        const double beta = beta_of(&proj[((j + 1) * 4 + v) * np]);
        nlw[j + 1][v] = lw[j + 1] / ((beta + EPS) * (beta + EPS));
      }
    }
    double wts[4][4];
    for (int v = 0; v < 4; ++v) {
      double den = 0.0;  // sum_axis(Axis(0)): zeros + row 0 + row 1 + ...
      for (int j = 0; j < 4; ++j) den = den + nlw[j][v];
      for (int j = 0; j < 4; ++j) wts[j][v] = nlw[j][v] / den;
    }
    std::fill(pn.begin(), pn.end(), 0.0);
    for (int v = 0; v < 4; ++v)
      for (int j = 0; j < 4; ++j)
        for (int k = 0; k < np; ++k) pn[v * np + k] += wts[j][v] * proj[(j * 4 + v) * np + k];
    gemm4(Rm, pn.data(), np, pn2.data(), use_fma);
    const double Jn = kPlant == PLANT_WENO_PLAIN_MEAN ? 2.0 : C.elements[inb].jacob_det;
    for (int t = 0; t < 4 * np; ++t) rec[t] = rec[t] + pn2[t] * Jn * 0.5;
  }
  for (int t = 0; t < 4 * np; ++t) out[t] = rec[t] / neighbour_area;
  return OK;
}

// SpatialDisc::weno, src/spatial_disc/limiter/weno.rs:86-269 (mode 1: in place, element order;
// mode 2, extension: every element tested and reconstructed from the pre-limiter state).
int apply_limiter(Case& C, double* U, uint8_t* troubled) {
  const int mode = C.prm.limiter;
  const size_t ne = C.elements.size();
  const int np = C.T.np;
  if (troubled) std::memset(troubled, 0, ne);
  if (mode == 0) return OK;
  std::vector<double> out(4 * np);
  if (mode == 1) {
    for (size_t e = 0; e < ne; ++e) {
      if (!modified_kxrcf(C, U, e)) continue;
      if (troubled) troubled[e] = 1;
      const int s = weno_element(C, U, e, out.data());
      if (s != OK) return s;
      for (int t = 0; t < 4 * np; ++t) U[e * 4 * np + t] = out[t];
    }
    return OK;
  }
  std::vector<uint8_t> flag(ne, 0);
  const int nt = nthreads_of(C);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t e = 0; e < (int64_t)ne; ++e) flag[e] = modified_kxrcf(C, U, e) ? 1 : 0;
  std::vector<double> res;
  std::vector<int64_t> list;
  for (size_t e = 0; e < ne; ++e)
    if (flag[e]) list.push_back(e);
  res.assign(list.size() * 4 * np, 0.0);
  int status = OK;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 16)
  for (int64_t i = 0; i < (int64_t)list.size(); ++i) {
    const int s = weno_element(C, U, list[i], &res[i * 4 * np]);
    if (s != OK) status = s;
  }
  if (status != OK) return status;
  for (size_t i = 0; i < list.size(); ++i) {
    for (int t = 0; t < 4 * np; ++t) U[list[i] * 4 * np + t] = res[i * 4 * np + t];
    if (troubled) troubled[list[i]] = 1;
  }
  return OK;
}

// ---------------------------------------------------------------------------------------
// Temporal — src/temporal_disc.rs, src/temporal_disc/rk3.rs
// ---------------------------------------------------------------------------------------

// TemperalDisc::compute_time_step, src/temporal_disc.rs:58-88 (wave speed from the SUM of all modes)
double compute_time_step(const Case& C, const double* U) {
  const int np = C.T.np;
  const double hcr = C.prm.hcr;
  double time_step = 1.0e10;
  for (size_t e = 0; e < C.elements.size(); ++e) {
    double rho = 0.0, rho_u = 0.0, rho_v = 0.0, rho_e = 0.0;
    for (int k = 0; k < (kPlant == PLANT_DT_MODE0 ? 1 : np); ++k) {
      rho += U[C.idx(e, 0, k)];
      rho_u += U[C.idx(e, 1, k)];
      rho_v += U[C.idx(e, 2, k)];
      rho_e += U[C.idx(e, 3, k)];
    }
    const double u = rho_u / rho;
    const double v = rho_v / rho;
    const double p = (hcr - 1.0) * (rho_e - 0.5 * rho * (u * u + v * v));
    const double c = std::sqrt(hcr * p / rho);
    const double speed = std::sqrt(u * u + v * v) + c;
    const double dx = C.elements[e].minimum_height;
    const double dt = C.cfl * dx / speed;
    if (dt < time_step) time_step = dt;
  }
  return time_step;
}

// TemperalDisc::rk3, src/temporal_disc/rk3.rs:7-26
int rk3(Case& C, double* R, double* U, double dt) {
  const size_t n = C.nstate();
  std::vector<double>& t0 = C.tmp0;
  std::vector<double>& t1 = C.tmp1;
  t0.assign(n, 0.0);
  t1.assign(n, 0.0);
  int s;
  std::fill(R, R + n, 0.0);
  if ((s = compute_residuals(C, R, U, nullptr)) != OK) return s;
  for (size_t i = 0; i < n; ++i) t0[i] = U[i] + dt * R[i];
  if ((s = apply_limiter(C, t0.data(), nullptr)) != OK) return s;
  std::fill(R, R + n, 0.0);
  if ((s = compute_residuals(C, R, t0.data(), nullptr)) != OK) return s;
  for (size_t i = 0; i < n; ++i) t1[i] = 0.75 * U[i] + 0.25 * (t0[i] + dt * R[i]);
  if ((s = apply_limiter(C, t1.data(), nullptr)) != OK) return s;
  std::fill(R, R + n, 0.0);
  if ((s = compute_residuals(C, R, t1.data(), nullptr)) != OK) return s;
  for (size_t i = 0; i < n; ++i) U[i] = 1.0 / 3.0 * U[i] + 2.0 / 3.0 * (t1[i] + dt * R[i]);
  return apply_limiter(C, U, nullptr);
}

}  // namespace

struct ref_case {
  Case c;
};

// ---------------------------------------------------------------------------------------
// GmshParser — src/io/mesh_parser.rs:28-158
// ---------------------------------------------------------------------------------------
struct ref_gmsh {
  std::vector<double> vxy;
  std::vector<std::string> eb_name;
  std::vector<std::vector<int64_t>> eb_nodes;
  std::vector<std::vector<int64_t>> el_nodes;
};

static std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  std::istringstream is(s);
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}
static std::string lower(std::string s) {
  for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}

extern "C" {

ref_gmsh* ref_gmsh_parse(const char* path) {
  std::ifstream f(path);
  if (!f) return nullptr;
  enum St { Standby, Phys, PhysFirst, Ent, EntFirst, Nodes, NodesFirst, ElemFirst, EdgeBlocks, ElemBlocks, BlockFirst };
  St st = Standby;
  std::map<int64_t, std::string> phys;
  std::map<int64_t, int64_t> ent2phys;
  ref_gmsh* g = new ref_gmsh();
  int64_t line_count = 0, max_line_count = 0;
  std::string line;
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.rfind("$End", 0) == 0) { st = Standby; continue; }
    if (line.rfind("$PhysicalNames", 0) == 0) { st = PhysFirst; continue; }
    if (line.rfind("$Entities", 0) == 0) { st = EntFirst; continue; }
    if (line.rfind("$Nodes", 0) == 0) { st = NodesFirst; continue; }
    if (line.rfind("$Elements", 0) == 0) { st = ElemFirst; continue; }
    const auto v = split_ws(line);
    switch (st) {
      case Standby: break;
      case PhysFirst: st = Phys; break;
      case Phys:
        if (v.size() == 3) {
          std::string name = v[2];
          while (!name.empty() && name.front() == '"') name.erase(name.begin());
          while (!name.empty() && name.back() == '"') name.pop_back();
          phys[std::stoll(v[1])] = name;
        }
        break;
      case EntFirst: st = Ent; break;
      case Ent:
        if (v.size() >= 11) ent2phys[std::stoll(v[0])] = std::stoll(v[8]);
        break;
      case NodesFirst: st = Nodes; break;
      case Nodes:
        if (v.size() == 3) {
          g->vxy.push_back(std::strtod(v[0].c_str(), nullptr));
          g->vxy.push_back(std::strtod(v[1].c_str(), nullptr));
        }
        break;
      case ElemFirst: st = BlockFirst; break;
      case BlockFirst: {
        const int64_t dim = std::stoll(v[0]), tag = std::stoll(v[1]), num = std::stoll(v[3]);
        const std::string name = phys.at(ent2phys.at(tag));
        if (dim == 1) {
          g->eb_name.push_back(name);
          g->eb_nodes.emplace_back();
          st = EdgeBlocks;
        }
        if (dim == 2) {
          g->el_nodes.emplace_back();
          st = ElemBlocks;
        }
        line_count = 0;
        max_line_count = num;
        break;
      }
      case EdgeBlocks:
        line_count += 1;
        g->eb_nodes.back().push_back(std::stoll(v[1]) - 1);
        g->eb_nodes.back().push_back(std::stoll(v[2]) - 1);
        if (line_count == max_line_count) st = BlockFirst;
        break;
      case ElemBlocks:
        line_count += 1;
        g->el_nodes.back().push_back(std::stoll(v[1]) - 1);
        g->el_nodes.back().push_back(std::stoll(v[2]) - 1);
        g->el_nodes.back().push_back(std::stoll(v[3]) - 1);
        if (line_count == max_line_count) st = BlockFirst;
        break;
    }
  }
  return g;
}

void ref_gmsh_sizes(const ref_gmsh* g, int64_t* nv, int64_t* neb, int64_t* nen, int64_t* nelb, int64_t* nel) {
  *nv = (int64_t)g->vxy.size() / 2;
  *neb = (int64_t)g->eb_nodes.size();
  int64_t s = 0;
  for (auto& b : g->eb_nodes) s += (int64_t)b.size() / 2;
  *nen = s;
  *nelb = (int64_t)g->el_nodes.size();
  s = 0;
  for (auto& b : g->el_nodes) s += (int64_t)b.size() / 3;
  *nel = s;
}

void ref_gmsh_get(const ref_gmsh* g, double* vxy, int32_t* eb_type, int64_t* eb_count, int64_t* eb_nodes,
                  int64_t* el_count, int64_t* el_nodes) {
  std::memcpy(vxy, g->vxy.data(), g->vxy.size() * sizeof(double));
  size_t o = 0;
  for (size_t b = 0; b < g->eb_nodes.size(); ++b) {
    const std::string n = lower(g->eb_name[b]);
    eb_type[b] = n == "wall" ? 0 : (n == "farfield" ? 1 : -1);
    eb_count[b] = (int64_t)g->eb_nodes[b].size() / 2;
    std::memcpy(eb_nodes + o, g->eb_nodes[b].data(), g->eb_nodes[b].size() * sizeof(int64_t));
    o += g->eb_nodes[b].size();
  }
  o = 0;
  for (size_t b = 0; b < g->el_nodes.size(); ++b) {
    el_count[b] = (int64_t)g->el_nodes[b].size() / 3;
    std::memcpy(el_nodes + o, g->el_nodes[b].data(), g->el_nodes[b].size() * sizeof(int64_t));
    o += g->el_nodes[b].size();
  }
}

void ref_gmsh_free(ref_gmsh* g) { delete g; }

ref_case* ref_create(int64_t nv, const double* vxy, int64_t neb, const int32_t* eb_type, const int64_t* eb_count,
                     const int64_t* eb_nodes, int64_t nelb, const int64_t* el_count, const int64_t* el_nodes,
                     const double* elem_xy, const ref_params* prm, int32_t* status) {
  ref_case* rc = new ref_case();
  Case& C = rc->c;
  C.prm = *prm;
  *status = OK;
  // default rules (extension): the reference pair (6, 3) for p <= 2; (16, 5) for p = 3; (25, 5) for p = 4
  int nq = prm->n_cell_gp, ng = prm->n_edge_gp;
  if (nq <= 0) nq = prm->order <= 2 ? 6 : (prm->order == 3 ? 16 : 25);
  if (ng <= 0) ng = prm->order <= 2 ? 3 : 5;
  C.prm.n_cell_gp = nq;
  C.prm.n_edge_gp = ng;
  if (prm->order < 1 || prm->order > 4 || !build_tables(C.T, prm->order, nq, ng)) {
    *status = BAD_CONFIG;
    delete rc;
    return nullptr;
  }
  C.vx.resize(nv);
  C.vy.resize(nv);
  for (int64_t i = 0; i < nv; ++i) {
    C.vx[i] = vxy[2 * i];
    C.vy[i] = vxy[2 * i + 1];
  }
  int64_t ne = 0;
  for (int64_t b = 0; b < nelb; ++b) ne += el_count[b];
  if (elem_xy) C.elem_xy.assign(elem_xy, elem_xy + ne * 6);
  const int s = process_mesh(C, neb, eb_type, eb_count, eb_nodes, nelb, el_count, el_nodes);
  if (s != OK) {
    *status = s;
    delete rc;
    return nullptr;
  }
  compute_geometry(C);
  // CFL, src/initialization.rs:29-33 (the JSON cfl is ignored by the reference; a positive
  // ref_params.cfl overrides the formula, as the benchmark configurations require)
  if (prm->cfl > 0.0) {
    C.cfl = prm->cfl;
  } else {
    const double viscous_coefficient = 1.0;
    const double p = (double)prm->order;
    C.cfl = viscous_coefficient * 1.0 / ((2.0 * p + 1.0) * (1.0 + 4.0 / ((p + 2.0) * (p + 2.0))));
  }
  return rc;
}

void ref_destroy(ref_case* c) { delete c; }

void ref_sizes(const ref_case* rc, int64_t* ne, int64_t* ned, int64_t* nie, int64_t* nin, int64_t* nbn, int32_t* nb,
               int32_t* nq, int32_t* ng) {
  const Case& C = rc->c;
  *ne = (int64_t)C.elements.size();
  *ned = (int64_t)C.edges.size();
  *nie = (int64_t)C.internal_edges.size();
  *nin = (int64_t)C.internal_elements.size();
  *nbn = (int64_t)C.boundary_elements.size();
  *nb = C.T.np;
  *nq = C.T.nq;
  *ng = C.T.ne;
}

void ref_get_edges(const ref_case* rc, int64_t* ielements, int64_t* ivertices, int64_t* in_cell, int64_t* ipatch,
                   double* normal, double* jacob) {
  const Case& C = rc->c;
  for (size_t i = 0; i < C.edges.size(); ++i) {
    const Edge& e = C.edges[i];
    for (int j = 0; j < 2; ++j) {
      ielements[2 * i + j] = e.ielements[j];
      ivertices[2 * i + j] = e.ivertices[j];
      in_cell[2 * i + j] = e.in_cell_indices[j];
      normal[2 * i + j] = e.normal[j];
    }
    ipatch[i] = e.ipatch;
    jacob[i] = e.jacob_det;
  }
}

void ref_get_elements(const ref_case* rc, int64_t* iv, int64_t* ie, int64_t* inb, double* jacob, double* cr,
                      double* mh) {
  const Case& C = rc->c;
  for (size_t i = 0; i < C.elements.size(); ++i) {
    const Element& e = C.elements[i];
    for (int j = 0; j < 3; ++j) {
      iv[3 * i + j] = e.ivertices[j];
      ie[3 * i + j] = e.iedges[j];
      inb[3 * i + j] = e.ineighbours[j];
    }
    jacob[i] = e.jacob_det;
    cr[i] = e.circumradius;
    mh[i] = e.minimum_height;
  }
}

void ref_get_lists(const ref_case* rc, int64_t* a, int64_t* b, int64_t* c) {
  const Case& C = rc->c;
  std::memcpy(a, C.internal_edges.data(), C.internal_edges.size() * sizeof(int64_t));
  std::memcpy(b, C.internal_elements.data(), C.internal_elements.size() * sizeof(int64_t));
  std::memcpy(c, C.boundary_elements.data(), C.boundary_elements.size() * sizeof(int64_t));
}

int32_t ref_n_deriv_keys(const ref_case* rc) { return rc->c.T.nkeys; }

void ref_get_tables(const ref_case* rc, double* phi_cell, double* phi_edge, double* mass, double* cell_pts,
                    double* cell_w, double* edge_pts, double* edge_w, double* dref) {
  const Tables& T = rc->c.T;
  std::memcpy(phi_cell, T.phi_cell.data(), T.phi_cell.size() * 8);
  std::memcpy(phi_edge, T.phi_edge.data(), T.phi_edge.size() * 8);
  std::memcpy(mass, T.mass.data(), T.mass.size() * 8);
  std::memcpy(cell_pts, T.cell_pts.data(), T.cell_pts.size() * 8);
  std::memcpy(cell_w, T.cell_w.data(), T.cell_w.size() * 8);
  std::memcpy(edge_pts, T.edge_pts.data(), T.edge_pts.size() * 8);
  std::memcpy(edge_w, T.edge_w.data(), T.edge_w.size() * 8);
  std::memcpy(dref, T.dref.data(), T.dref.size() * 8);
}

double ref_cfl(const ref_case* rc) { return rc->c.cfl; }
int64_t ref_last_bad_index(const ref_case* rc) { return rc->c.bad_index; }

double ref_phys_deriv(const ref_case* rc, int64_t e, int32_t igp, int32_t k, int32_t l1, int32_t l2) {
  return rc->c.dphys(e, igp, k, l1, l2);
}

// Solver::set_initial_solution, src/solver.rs:45-79
void ref_initial_solution(const ref_case* rc, double density, double aoa_deg, double mach, double pressure, double* U) {
  const Case& C = rc->c;
  const double hcr = C.prm.hcr;
  const double sound_speed = std::sqrt(hcr * pressure / density);
  const double rad = aoa_deg * (M_PI / 180.0);
  const double x_mach = mach * std::cos(rad);
  const double y_mach = mach * std::sin(rad);
  const double xv = x_mach * sound_speed, yv = y_mach * sound_speed;
  const double E = pressure / (hcr - 1.0) + 0.5 * density * (xv * xv + yv * yv);
  const double q[4] = {density, density * xv, density * yv, E};
  const Tables& T = C.T;
  std::fill(U, U + C.nstate(), 0.0);
  for (size_t e = 0; e < C.elements.size(); ++e)
    for (int v = 0; v < 4; ++v)
      for (int b = 0; b < T.np; ++b)
        for (int i = 0; i < T.nq; ++i)
          U[C.idx(e, v, b)] += 0.5 * C.elements[e].jacob_det * T.cell_w[i] * q[v] * T.phi_cell[i * T.np + b];
  for (size_t e = 0; e < C.elements.size(); ++e)
    for (int v = 0; v < 4; ++v)
      for (int b = 0; b < T.np; ++b) U[C.idx(e, v, b)] /= T.mass[b] * 0.5 * C.elements[e].jacob_det;
}

// extension: the same projection (src/solver.rs:80-95) with a pointwise state per cell point
void ref_project(const ref_case* rc, const double* qp, double* U) {
  const Case& C = rc->c;
  const Tables& T = C.T;
  std::fill(U, U + C.nstate(), 0.0);
  for (size_t e = 0; e < C.elements.size(); ++e)
    for (int v = 0; v < 4; ++v)
      for (int b = 0; b < T.np; ++b)
        for (int i = 0; i < T.nq; ++i)
          U[C.idx(e, v, b)] +=
              0.5 * C.elements[e].jacob_det * T.cell_w[i] * qp[(e * T.nq + i) * 4 + v] * T.phi_cell[i * T.np + b];
  for (size_t e = 0; e < C.elements.size(); ++e)
    for (int v = 0; v < 4; ++v)
      for (int b = 0; b < T.np; ++b) U[C.idx(e, v, b)] /= T.mass[b] * 0.5 * C.elements[e].jacob_det;
}

void ref_cell_points_xy(const ref_case* rc, double* xy) {
  const Case& C = rc->c;
  const Tables& T = C.T;
  for (size_t e = 0; e < C.elements.size(); ++e) {
    double x1, y1, x2, y2, x3, y3;
    C.elem_vertex(e, 0, x1, y1);
    C.elem_vertex(e, 1, x2, y2);
    C.elem_vertex(e, 2, x3, y3);
    for (int i = 0; i < T.nq; ++i) {
      const double xi = T.cell_pts[2 * i], eta = T.cell_pts[2 * i + 1];
      xy[(e * T.nq + i) * 2] = x1 + (x2 - x1) * xi + (x3 - x1) * eta;
      xy[(e * T.nq + i) * 2 + 1] = y1 + (y2 - y1) * xi + (y3 - y1) * eta;
    }
  }
}

int32_t ref_compute_residuals(ref_case* rc, double* R, const double* U, int64_t* bad_edge) {
  return compute_residuals(rc->c, R, U, bad_edge);
}
int32_t ref_volume_terms(ref_case* rc, double* R, const double* U) { return volume_terms(rc->c, R, U); }
int32_t ref_edge_terms(ref_case* rc, double* R, const double* U, int64_t* bad_edge) {
  const int s = edge_terms(rc->c, R, U);
  if (s != OK && bad_edge) *bad_edge = rc->c.bad_index;
  return s;
}
int32_t ref_boundary_terms(ref_case* rc, double* R, const double* U) { return boundary_terms(rc->c, R, U); }
void ref_divide_by_mass(const ref_case* rc, double* R) { divide_by_mass(rc->c, R); }

int32_t ref_apply_limiter(ref_case* rc, double* U, uint8_t* troubled) { return apply_limiter(rc->c, U, troubled); }
int32_t ref_kxrcf(const ref_case* rc, const double* U, int64_t e) { return modified_kxrcf(rc->c, U, e) ? 1 : 0; }
int32_t ref_weno_element(const ref_case* rc, const double* U, int64_t e, double* out) {
  return weno_element(rc->c, U, e, out);
}
double ref_compute_time_step(const ref_case* rc, const double* U) { return compute_time_step(rc->c, U); }
int32_t ref_rk3(ref_case* rc, double* R, double* U, double dt) { return rk3(rc->c, R, U, dt); }

// TemperalDisc::time_march, src/temporal_disc.rs:21-57 (VTK output and printing omitted)
int32_t ref_time_march(ref_case* rc, double* R, double* U, int32_t final_step, double final_time, int32_t* steps_done,
                       double* t_done) {
  Case& C = rc->c;
  int32_t step = 0;
  double t = 0.0;
  while (step < final_step && t < final_time) {
    double dt = compute_time_step(C, U);
    if (t + dt > final_time) dt = final_time - t;
    const int s = rk3(C, R, U, dt);
    if (s != OK) {
      *steps_done = step;
      *t_done = t;
      return s;
    }
    t += dt;
    step += 1;
  }
  *steps_done = step;
  *t_done = t;
  return OK;
}

int32_t ref_hllc(const double* ql, const double* qr, const double* n, double hcr, double* f) {
  return hllc(ql, qr, n, hcr, f);
}
void ref_flux(const double* q, double hcr, double* f, double* g) { euler_flux(q, hcr, f, g); }
void ref_eigenmatrix(const double* q, double nx, double ny, double hcr, double* L, double* R) {
  eigenmatrix(q, nx, ny, hcr, L, R);
}

// rk3.rs:7-26 with R(u) = lambda*u (zeroed, "computed", then the same stage expressions)
void ref_rk3_linear(double* u, int64_t n, double lambda, double dt) {
  std::vector<double> t0(n), t1(n), r(n);
  for (int64_t i = 0; i < n; ++i) r[i] = lambda * u[i];
  for (int64_t i = 0; i < n; ++i) t0[i] = u[i] + dt * r[i];
  for (int64_t i = 0; i < n; ++i) r[i] = lambda * t0[i];
  for (int64_t i = 0; i < n; ++i) t1[i] = 0.75 * u[i] + 0.25 * (t0[i] + dt * r[i]);
  for (int64_t i = 0; i < n; ++i) r[i] = lambda * t1[i];
  for (int64_t i = 0; i < n; ++i) u[i] = 1.0 / 3.0 * u[i] + 2.0 / 3.0 * (t1[i] + dt * r[i]);
}

}  // extern "C"
