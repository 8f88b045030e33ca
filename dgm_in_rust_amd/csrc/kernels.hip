// HIP kernels (gfx950) of the DG-Euler explicit step and their host-side launch sequences.
//
// Per RK stage (src/temporal_disc/rk3.rs:7-26):
//   k_surface  one lane per edge: traces of both neighbours at the Ne edge points, HLLC
//              (src/spatial_disc/flux.rs:28-153) on interior edges, slip-wall / characteristic
//              far-field flux (src/spatial_disc/boundary_condition.rs:9-97) on boundary edges;
//              writes flux * |e|/2 per point. Each edge is evaluated once with the reference's
//              left/right orientation (the altered p* makes HLLC orientation-dependent).
//   k_element  one lane per element (k_element_mfma: 16 elements per wavefront on the matrix cores):
//              volume integral (src/spatial_disc.rs:40-69) with reference-element gradients, gather of
//              its three edges' fluxes and lift, the diagonal mass inverse (src/spatial_disc.rs:349-357),
//              and the SSP-RK3 stage combination. The flux-to-element step is a gather, so no atomics and
//              no edge colouring are needed. In FP32-bulk mode the residual runs in deviation form
//              (flux_deviation).
//   limiter    KXRCF detector (src/spatial_disc/limiter/weno.rs:7-85), one lane per element,
//              then one wavefront per troubled element for the characteristic WENO rebuild
//              (weno.rs:86-269) with FP64 smoothness indicators and weights.
// Per step: k_dt min-reduction (src/temporal_disc.rs:58-88) and device-side step control.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <memory>
#include <vector>

#include "kernels.h"
#include "solver.h"

namespace dgm {

namespace {

constexpr int np_of(int p) { return (p + 1) * (p + 2) / 2; }

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD b % 8). Renumber them so
// that each XCD works on one contiguous range of logical blocks: neighbouring edges / elements, which
// gather the same element data, then share that XCD's L2. A bijection on [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
  return x * q + (x < r ? x : r) + i;
}

// ---- arithmetic helpers: FP64 uses IEEE operations; the FP32 bulk path uses the hardware
// reciprocal / square root / log2 / exp2 (1 ulp class), which is its stated precision. ----
__device__ __forceinline__ double dv(double a, double b) { return a / b; }
__device__ __forceinline__ float dv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double sqrt_(double a) { return sqrt(a); }
__device__ __forceinline__ float sqrt_(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ __forceinline__ double pow_(double a, double b) { return pow(a, b); }
__device__ __forceinline__ float pow_(float a, float b) { return __builtin_amdgcn_exp2f(b * __builtin_amdgcn_logf(a)); }

template <class T>
__device__ __forceinline__ const T* geo_el(const DevMesh& m);
template <>
__device__ __forceinline__ const double* geo_el<double>(const DevMesh& m) { return m.el_geo; }
template <>
__device__ __forceinline__ const float* geo_el<float>(const DevMesh& m) { return m.el_geo32; }
template <class T>
__device__ __forceinline__ const T* geo_ed(const DevMesh& m);
template <>
__device__ __forceinline__ const double* geo_ed<double>(const DevMesh& m) { return m.ed_geo; }
template <>
__device__ __forceinline__ const float* geo_ed<float>(const DevMesh& m) { return m.ed_geo32; }

// Table offsets (bulk type T): PHIV, DXI, DETA [NQ][NP]; PHIE, LIFT [3][NE][NP]; KXW [3][NP], the edge-weighted
// basis of face i read in reversed point order, KXW[i][k] = sum_q w_q phi_i(NE - 1 - q, k): the neighbour's
// side of a KXRCF jump integral as one dot product with its modes.
template <int NP, int NQ, int NE>
struct TabOff {
  static constexpr int PHIV = 0, DXI = NQ * NP, DETA = 2 * NQ * NP, PHIE = 3 * NQ * NP, LIFT = 3 * NQ * NP + 3 * NE * NP;
  static constexpr int KXW = 3 * NQ * NP + 6 * NE * NP;
  static constexpr int N = KXW + 3 * NP;
};
// FP64 limiter tables: PHIE [3][NE][NP], EW [NE], CW [NQ], DREF [nkeys][NQ][NP]
template <int P, int NQ, int NE>
struct LimOff {
  static constexpr int NP = np_of(P);
  static constexpr int MAXD = P > 2 ? P : 2;
  static constexpr int NKEYS = MAXD * (MAXD + 3) / 2;
  static constexpr int PHIE = 0, EW = 3 * NE * NP, CW = EW + NE, DREF = CW + NQ;
  static constexpr int N = DREF + NKEYS * NQ * NP;
  static constexpr int NKW = P * (P + 3) / 2;  // physical keys (l1, l - l1), l = 1..P
};
__host__ __device__ constexpr int dref_key(int a, int b) { return (a + b) * (a + b + 1) / 2 - 1 + b; }

// ---- physics ----

// flux::flux, src/spatial_disc/flux.rs:154-168
template <class T>
__device__ __forceinline__ void euler_flux(const T q[4], T hcr, T f[4], T g[4]) {
  const T ri = dv(T(1), q[0]);
  const T u = q[1] * ri, v = q[2] * ri;
  const T p = (hcr - T(1)) * (q[3] - T(0.5) * (q[1] * q[1] + q[2] * q[2]) * ri);
  f[0] = q[1];
  f[1] = q[1] * u + p;
  f[2] = q[2] * u;
  f[3] = u * (q[3] + p);
  g[0] = q[2];
  g[1] = q[1] * v;
  g[2] = q[2] * v + p;
  g[3] = v * (q[3] + p);
}

// flux::hllc, src/spatial_disc/flux.rs:28-153, including the reference's p* numerator
// (c_L + c_R - ((zeta - 1)/2)(u_R - u_L)) of :75-78. Returns 0, or 1 / 2 for a negative density /
// pressure; the pressure test runs in FP64 whatever T is.
template <class T>
__device__ __forceinline__ int hllc(const T lv[4], const T rv[4], T nx, T ny, T hcr, T out[4]) {
  // multiply-adds are fused within one expression only (decided in the front end), so every kernel that
  // inlines the edge arithmetic (the edge pass, the fused edge-and-element pass) evaluates the same
  // operations
#pragma clang fp contract(on)
  const T ql0 = lv[0], ql1 = lv[1] * nx + lv[2] * ny, ql2 = -lv[1] * ny + lv[2] * nx, ql3 = lv[3];
  const T qr0 = rv[0], qr1 = rv[1] * nx + rv[2] * ny, qr2 = -rv[1] * ny + rv[2] * nx, qr3 = rv[3];
  const T il = dv(T(1), ql0), ir = dv(T(1), qr0);
  const T ul = ql1 * il, ur = qr1 * ir, vl = ql2 * il, vr = qr2 * ir;
  const T pl = (hcr - T(1)) * (ql3 - T(0.5) * (ql1 * ql1 + ql2 * ql2) * il);
  const T pr = (hcr - T(1)) * (qr3 - T(0.5) * (qr1 * qr1 + qr2 * qr2) * ir);
  if (ql0 < T(0) || qr0 < T(0)) return 1;
  if constexpr (sizeof(T) == 4) {
    // FP32 states: with gamma > 1 and rho >= 0, p < 0 iff 2 rho E < (rho u)^2 + (rho v)^2. The products of
    // FP32 values are exact in FP64, so this is the FP64 test without its two divisions (rho = 0: a
    // nonzero momentum fails, as -inf < 0 does; zero momentum passes, as NaN < 0 does).
    const double l = 2.0 * (double)ql0 * (double)ql3 - ((double)ql1 * (double)ql1 + (double)ql2 * (double)ql2);
    const double r = 2.0 * (double)qr0 * (double)qr3 - ((double)qr1 * (double)qr1 + (double)qr2 * (double)qr2);
    if (l < 0.0 || r < 0.0) return 2;
  } else {
    const double h = (double)hcr;
    const double pl64 = (h - 1.0) * ((double)ql3 - 0.5 * ((double)ql1 * (double)ql1 + (double)ql2 * (double)ql2) / (double)ql0);
    const double pr64 = (h - 1.0) * ((double)qr3 - 0.5 * ((double)qr1 * (double)qr1 + (double)qr2 * (double)qr2) / (double)qr0);
    if (pl64 < 0.0 || pr64 < 0.0) return 2;
  }
  const T cl = sqrt_(hcr * pl * il), cr = sqrt_(hcr * pr * ir);
  const T zeta = dv(hcr - T(1), T(2) * hcr);
  const T base = dv(cl + cr - (zeta - T(1)) * T(0.5) * (ur - ul), dv(cl, pow_(pl, zeta)) + dv(cr, pow_(pr, zeta)));
  const T ps = pow_(base, dv(T(1), zeta));
  const T gk = dv(hcr + T(1), T(2) * hcr);
  const T qlw = (ps <= pl) ? T(1) : sqrt_(T(1) + gk * (dv(ps, pl) - T(1)));
  const T qrw = (ps <= pr) ? T(1) : sqrt_(T(1) + gk * (dv(ps, pr) - T(1)));
  const T sl = ul - cl * qlw, sr = ur + cr * qrw;
  // The reference's four cases (:101-147): F(q_L) if S_L >= 0; else F(q_R) if S_R <= 0; else the star
  // state of side L if S* >= 0, of side R otherwise. Each is F(q_K) of one side K, plus, in the star
  // region, S_K (U*_K - q_K). The side's values are selected first and the expressions evaluated once,
  // the same operations as the case taken (lanes of a wavefront take different cases, and evaluating
  // every case under a mask would cost all four).
  const T ss = dv(pr - pl + ql0 * ul * (sl - ul) - qr0 * ur * (sr - ur), ql0 * (sl - ul) - qr0 * (sr - ur));
  const bool star = !(sl >= T(0)) && !(sr <= T(0));
  const bool left = (sl >= T(0)) || (star && ss >= T(0));
  const T k0 = left ? ql0 : qr0, k1 = left ? ql1 : qr1, k2 = left ? ql2 : qr2, k3 = left ? ql3 : qr3;
  const T uk = left ? ul : ur, vk = left ? vl : vr, pk = left ? pl : pr, ik = left ? il : ir, sk = left ? sl : sr;
  T f0 = k1, fn = k1 * uk + pk, ft = k2 * uk, f3 = uk * (k3 + pk);
  if (star) {
    const T a = dv(k0 * (sk - uk), sk - ss);
    const T e3 = a * (k3 * ik + (ss - uk) * (ss + dv(pk, k0 * (sk - uk))));
    f0 = k1 + sk * (a - k0);
    fn = fn + sk * (a * ss - k1);
    ft = ft + sk * (a * vk - k2);
    f3 = f3 + sk * (e3 - k3);
  }
  out[0] = f0;
  out[1] = fn * nx - ft * ny;
  out[2] = fn * ny + ft * nx;
  out[3] = f3;
  return 0;
}

// FP32-bulk residual in deviation form. The element passes subtract the flux of the element's mean state
// qm (its mode 0: the Dubiner mode 0 is 1 everywhere) from every point flux of the element: at the volume
// points as F(qm + dq) - F(qm), evaluated from the point's deviation dq (the sum over the modes above mode
// 0) by flux_deviation, and at the edge points as its normal component times the half length
// (surface_deviation). In exact arithmetic the residual is unchanged: the quadrature rules integrate a
// constant flux against the basis gradients and against the traces exactly, so its volume and surface
// terms cancel. In FP32 the point fluxes, their products with the tables and the sums then round at the
// scale of the deviations rather than of the full fluxes, which largely cancel in a smooth flow. What the
// volume terms leave out is the exact flux of qm, so the surface terms subtract it in FP64, where its
// rounding stays far below FP32's. Multiply-adds are fused within one expression only, so the separate
// and the fused element passes round alike.
struct MeanState {
  float u, v, h;  // velocity and total enthalpy per volume (E + p) of the mean state
};
__device__ __forceinline__ MeanState mean_state(const float qm[4], float hcr) {
#pragma clang fp contract(on)
  const float ri = dv(1.f, qm[0]);
  const float u = qm[1] * ri, v = qm[2] * ri;
  const float p = (hcr - 1.f) * (qm[3] - 0.5f * (qm[1] * qm[1] + qm[2] * qm[2]) * ri);
  return {u, v, qm[3] + p};
}
// F(qm + dq) - F(qm) and G(qm + dq) - G(qm) of the Euler flux (euler_flux) from the identities, with
// r' = r + dr the point's density, u' = u + du its velocity and (m, n) = r (u, v):
//   du = (dm - u dr) / r',  d(m u) = dm u' + m du,  d(m v) = dm v' + m dv,  dp = (g - 1)(dE - (d(m u) + d(n v)) / 2),
//   d((E + p) u) = (dE + dp) u' + (E + p) du,  and likewise in y;
// every term is a deviation times a mean quantity, so the rounding scales with the deviation
__device__ __forceinline__ void flux_deviation(const float qm[4], const MeanState& s, const float dq[4], float hcr,
                                               float df[4], float dg[4]) {
#pragma clang fp contract(on)
  const float ri = dv(1.f, qm[0] + dq[0]);
  const float du = (dq[1] - s.u * dq[0]) * ri, dw = (dq[2] - s.v * dq[0]) * ri;
  const float u1 = s.u + du, v1 = s.v + dw;
  const float dmu = dq[1] * u1 + qm[1] * du, dnv = dq[2] * v1 + qm[2] * dw;
  const float dnu = dq[2] * u1 + qm[2] * du, dmv = dq[1] * v1 + qm[1] * dw;
  const float dp = (hcr - 1.f) * (dq[3] - 0.5f * (dmu + dnv));
  const float dh = dq[3] + dp;
  df[0] = dq[1];
  df[1] = dmu + dp;
  df[2] = dnu;
  df[3] = dh * u1 + s.h * du;
  dg[0] = dq[2];
  dg[1] = dmv;
  dg[2] = dnv + dp;
  dg[3] = dh * v1 + s.h * dw;
}
// the metric-transformed flux a f + b g
__device__ __forceinline__ float metric(float a, float b, float f, float g) {
#pragma clang fp contract(on)
  return a * f + b * g;
}
// the flux of the mean state in FP64
__device__ __forceinline__ void mean_flux64(const float qm[4], double hcr, double fr[4], double gr[4]) {
  const double q[4] = {(double)qm[0], (double)qm[1], (double)qm[2], (double)qm[3]};
  euler_flux(q, hcr, fr, gr);
}
// the mean flux's normal component times the half length of an edge (FP64): what surface_deviation subtracts
__device__ __forceinline__ double mean_normal_flux(double fr, double gr, float nx, float ny, float hl) {
#pragma clang fp contract(on)
  return (fr * (double)nx + gr * (double)ny) * (double)hl;
}
// lift operand of an edge point: sc (the side's sign times 2/J) times the point flux less the mean flux's
// normal part, the difference taken in FP64 and rounded once
__device__ __forceinline__ float surface_deviation(float sc, float flux, double rn) {
  return sc * (float)((double)flux - rn);
}

// local_characteristics::compute_eigenmatrix, src/spatial_disc/local_characteristics.rs:4-50
template <class T>
__device__ __forceinline__ void eigenmatrix(const T q[4], T nx, T ny, T hcr, T L[16], T R[16]) {
  // multiply-adds are fused within one expression only (decided in the front end), so every kernel that
  // inlines the edge arithmetic (the edge pass, the fused edge-and-element pass) evaluates the same
  // operations
#pragma clang fp contract(on)
  const T ri = dv(T(1), q[0]);
  const T u = q[1] * ri, v = q[2] * ri;
  const T p = (hcr - T(1)) * (q[3] - T(0.5) * (q[1] * q[1] + q[2] * q[2]) * ri);
  const T c = sqrt_(hcr * p * ri);
  const T ic = dv(T(1), c);
  const T h = (q[3] + p) * ri;
  const T b1 = (hcr - T(1)) * ic * ic;
  const T b2 = b1 * (u * u + v * v) * T(0.5);
  const T vn = u * nx + v * ny;
  L[0] = (b2 + vn * ic) * T(0.5);
  L[1] = -(b1 * u + nx * ic) * T(0.5);
  L[2] = -(b1 * v + ny * ic) * T(0.5);
  L[3] = b1 * T(0.5);
  L[4] = ny * u - nx * v;
  L[5] = -ny;
  L[6] = nx;
  L[7] = T(0);
  L[8] = T(1) - b2;
  L[9] = b1 * u;
  L[10] = b1 * v;
  L[11] = -b1;
  L[12] = (b2 - vn * ic) * T(0.5);
  L[13] = -(b1 * u - nx * ic) * T(0.5);
  L[14] = -(b1 * v - ny * ic) * T(0.5);
  L[15] = b1 * T(0.5);
  R[0] = T(1);
  R[1] = T(0);
  R[2] = T(1);
  R[3] = T(1);
  R[4] = u - c * nx;
  R[5] = -ny;
  R[6] = u;
  R[7] = u + c * nx;
  R[8] = v - c * ny;
  R[9] = nx;
  R[10] = v;
  R[11] = v + c * ny;
  R[12] = h - c * vn;
  R[13] = -u * ny + v * nx;
  R[14] = (u * u + v * v) * T(0.5);
  R[15] = h + c * vn;
}

// characteristic far-field state, src/spatial_disc/boundary_condition.rs:48-94
template <class T>
__device__ __forceinline__ void farfield_flux(const T lv[4], const double* qinf, T nx, T ny, T hcr, T out[4]) {
  // multiply-adds are fused within one expression only (decided in the front end), so every kernel that
  // inlines the edge arithmetic (the edge pass, the fused edge-and-element pass) evaluates the same
  // operations
#pragma clang fp contract(on)
  const T ri = dv(T(1), lv[0]);
  const T u = lv[1] * ri, v = lv[2] * ri;
  const T p = (hcr - T(1)) * (lv[3] - T(0.5) * (lv[1] * lv[1] + lv[2] * lv[2]) * ri);
  const T c = sqrt_(hcr * p * ri);
  const T vn = u * nx + v * ny;
  T L[16], R[16];
  eigenmatrix(lv, nx, ny, hcr, L, R);
  const T lam[4] = {vn - c, vn, vn, vn + c};
  const T qi[4] = {(T)qinf[0], (T)qinf[1], (T)qinf[2], (T)qinf[3]};
  T w[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const T a = L[4 * i] * lv[0] + L[4 * i + 1] * lv[1] + L[4 * i + 2] * lv[2] + L[4 * i + 3] * lv[3];
    const T b = L[4 * i] * qi[0] + L[4 * i + 1] * qi[1] + L[4 * i + 2] * qi[2] + L[4 * i + 3] * qi[3];
    w[i] = lam[i] >= T(0) ? a : b;
  }
  T q[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) q[i] = R[4 * i] * w[0] + R[4 * i + 1] * w[1] + R[4 * i + 2] * w[2] + R[4 * i + 3] * w[3];
  const T qi0 = dv(T(1), q[0]);
  const T un = q[1] * qi0, vnw = q[2] * qi0;
  const T vnn = un * nx + vnw * ny;
  const T pn = (hcr - T(1)) * (q[3] - T(0.5) * (q[1] * q[1] + q[2] * q[2]) * qi0);
  out[0] = q[0] * vnn;
  out[1] = q[1] * vnn + pn * nx;
  out[2] = q[2] * vnn + pn * ny;
  out[3] = (q[3] + pn) * vnn;
}

// Element-array access through a 64-bit base in scalar registers plus a 32-bit unsigned byte offset per
// lane, which the memory instruction adds itself: no 64-bit address arithmetic per lane. The byte offset
// must stay below 4 GiB; Arena::init rejects capacities whose variable blocks or flux array would not.
template <class T>
__device__ __forceinline__ T ld32(const T* base, unsigned idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ void st32(T* base, unsigned idx, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + idx * (unsigned)sizeof(T)) = v;
}

__device__ __forceinline__ void record_error(CaseCtl* cc, int edge, int code) {
  atomicMin(&cc->error, ((unsigned long long)(unsigned)edge << 8) | (unsigned long long)code);
}

__device__ __forceinline__ bool case_runs(const CaseCtl& cc) { return cc.active && case_error(cc) == ~0ull; }

// Case of a 64-block, read through the scalar unit: the block index is uniform over the wavefront
// (every caller's wavefront covers one 64-block), so the case record that follows is a scalar load.
__device__ __forceinline__ int case_of(const DevMesh& m, const int* __restrict__ map, int index) {
  if (m.single_case >= 0) return m.single_case;
  return map[__builtin_amdgcn_readfirstlane(index >> 6)];
}

// store a rebuilt element (double [4][NP]) into the buffers of the given stage
template <class T, int NP, int STAGE>
__device__ __forceinline__ void store_limited(const StageBufs<T>& b, int nes, int e, int idx, double val) {
  const int v = idx / NP, k = idx % NP;
  const size_t ix = (size_t)(v * NP + k) * nes + e;
  if constexpr (sizeof(T) == 8) {
    if (STAGE == 2)
      b.un[ix] = val;
    else
      b.us[ix] = (T)val;
  } else {
    if (STAGE == 2) {
      b.un[ix] = val;
      b.sn[ix] = (T)val;
    } else {
      (STAGE == 0 ? b.s : b.s2)[ix] = (T)val;
      b.d[ix] = (T)(val - b.un[ix]);
    }
  }
}

// store a whole rebuilt element w[4 NP]: every load (the row, and u^n where the stage stores d = w - u^n)
// is issued before the first store, which the compiler could not otherwise move them past
template <class T, int NP, int STAGE>
__device__ __forceinline__ void store_limited_row(const StageBufs<T>& b, int nes, int e, const double* w) {
  constexpr bool DN = sizeof(T) == 4 && STAGE < 2;
  double v[4 * NP], un[DN ? 4 * NP : 1];
#pragma unroll
  for (int c = 0; c < 4 * NP; ++c) {
    v[c] = w[c];
    if constexpr (DN) un[c] = b.un[(size_t)c * nes + e];
  }
#pragma unroll
  for (int c = 0; c < 4 * NP; ++c) {
    const size_t ix = (size_t)c * nes + e;
    if constexpr (DN) {
      (STAGE == 0 ? b.s : b.s2)[ix] = (T)v[c];
      b.d[ix] = (T)(v[c] - un[c]);
    } else {
      store_limited<T, NP, STAGE>(b, nes, e, c, v[c]);
    }
  }
}

// A case that has stopped no longer reads its pending rebuilds, and later limiter passes of other cases
// reuse their slots: the first stage-0 element pass after it stops writes them back (u^n and the FP32
// shadow) and clears the marks, before this step's limiter reuses the slots.
template <class T, int NP>
__device__ __forceinline__ void write_back_pending(const StageBufs<T>& b, int e, int nes) {
  const int sl = b.pend[e];
  if (sl < 0) return;
  const double* w = b.pw + (size_t)sl * 4 * NP;
#pragma unroll 1
  for (int c = 0; c < 4 * NP; ++c) {
    b.un[(size_t)c * nes + e] = w[c];
    if constexpr (sizeof(T) == 4) b.sn[(size_t)c * nes + e] = (float)w[c];
  }
  const_cast<int*>(b.pend)[e] = -1;
}

// Appends the lanes with pred set to a list counter with one atomic per wavefront; returns each such
// lane's slot (-1 for the others). Every active lane of the wavefront must call it.
__device__ __forceinline__ int wave_append(int* counter, bool pred, int* also) {
  const unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return -1;
  const int lane = (int)(threadIdx.x & 63);
  const int leader = __ffsll((long long)mask) - 1;
  int base = 0;
  if (lane == leader) {
    const int n = __popcll(mask);
    base = atomicAdd(counter, n);
    atomicAdd(also, n);
  }
  base = __shfl(base, leader, 64);
  return pred ? base + __popcll(mask & ((1ull << lane) - 1ull)) : -1;
}

// dt slot of a case for the given step parity
__device__ __forceinline__ unsigned long long* dt_slot(const Ctl& ctl, int c, int parity) {
  return ctl.dtmin + (size_t)c * DT_CASE + (size_t)(parity & 1) * DT_SLOT;
}

// ---------------------------------------------------------------------------------------
// surface kernel: one lane per edge
// ---------------------------------------------------------------------------------------

// traces tr[q][v] at the NE points of local edge lic (point order reversed with rev) of the modes
// coef(v, k)
template <class T, int NP, int NE, class F>
__device__ __forceinline__ void edge_traces(F coef, const T* sphi, int lic, bool rev, T tr[NE][4]) {
  // multiply-adds are fused within one expression only (decided in the front end), so every kernel that
  // inlines the edge arithmetic (the edge pass, the fused edge-and-element pass) evaluates the same
  // operations
#pragma clang fp contract(on)
#pragma unroll
  for (int q = 0; q < NE; ++q)
#pragma unroll
    for (int v = 0; v < 4; ++v) tr[q][v] = T(0);
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const T c = coef(v, k);
#pragma unroll
      for (int q = 0; q < NE; ++q) tr[q][v] += c * sphi[(lic * NE + (rev ? NE - 1 - q : q)) * NP + k];
    }
}
__device__ __forceinline__ void step_control(const Ctl& ctl, int i, int lane);

// Stage 0 of a step (s0) runs beside the step control of that step: with nctl > 0 the last blocks of
// the grid run it for case slots 0 .. nctl - 1, one wavefront per slot. The edges of stage 0 then read
// the case's active flag as it stands (the step control may be clearing it: either value is safe, since
// a case computed in vain records its errors into error_s0, which the stage-0 element pass takes over
// only when the case runs the step) and never write error.
template <class T, int NP, int NQ, int NE>
__global__ __launch_bounds__(256) void k_surface(DevMesh m, const T* __restrict__ U, T* __restrict__ G,
                                                 const T* __restrict__ tab, Ctl ctl, int force, int remap,
                                                 const int* __restrict__ pend, const double* __restrict__ pw, int s0,
                                                 int nctl) {
  using O = TabOff<NP, NQ, NE>;
  __shared__ T sphi[3 * NE * NP];
  const int wpb = (int)(blockDim.x >> 6);
  const int nblk = (int)gridDim.x - (nctl + wpb - 1) / wpb;  // blocks of edges; the rest: step control
  if ((int)blockIdx.x >= nblk) {  // uniform per block, before any barrier
    const int i = ((int)blockIdx.x - nblk) * wpb + (int)(threadIdx.x >> 6);
    if (i < nctl) step_control(ctl, i, (int)(threadIdx.x & 63));
    return;
  }
  const int blk = remap ? xcd_block(blockIdx.x, nblk) : (int)blockIdx.x;
  const int ed = blk * blockDim.x + threadIdx.x;
  const int nes = m.nes, est = m.ed_stride;
  // the edge record (read at a clamped index) goes out before the table staging and its barrier, both
  // element ids included (the right one read later would add a dependent trip before its modes)
  const int ee = ed < m.nedge ? ed : 0;
  const int info = m.ed_info[ee];
  const int L = m.ed_elem[ee];
  const int R = m.ed_elem[est + ee];  // -1 on a boundary edge (not read there)
  const T* eg = geo_ed<T>(m);
  const T nx = eg[ee], ny = eg[est + ee], hl = eg[2 * est + ee];
  // and so does the case record (of case 0 for edges outside any case, which exit below)
  const int cs = case_of(m, m.ed_case64, ee);
  CaseCtl* cc = ctl.cases + (cs < 0 ? 0 : cs);
  const int c_active = s0 ? __hip_atomic_load(&cc->active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : cc->active;
  const int c_ed_end = cc->ed_end;
  const unsigned long long c_error = case_error(*cc);
  for (int i = threadIdx.x; i < 3 * NE * NP; i += blockDim.x) sphi[i] = tab[O::PHIE + i];
  __syncthreads();
  if (ed >= m.nedge || cs < 0) return;
  if (!force && !(c_active && c_error == ~0ull)) return;  // case_runs()
  if (ed >= c_ed_end) return;
  const int licL = info & 3, licR = (info >> 2) & 3, bc = (info >> 4) & 3, patch = info >> 8;
  const T hcr = (T)m.hcr;
  // an element rebuilt by the last limiter pass and not yet written back is read from the rebuild
  // (rare: the stored modes are loaded regardless, so the common path waits on no extra load)
  const int pl = pend ? pend[L] : -1;
  T ul[NE][4];
  edge_traces<T, NP, NE>([&](int v, int k) { return U[(size_t)(v * NP + k) * nes + L]; }, sphi, licL, false, ul);
  if (pl >= 0)
    edge_traces<T, NP, NE>([&](int v, int k) { return (T)pw[(size_t)pl * 4 * NP + v * NP + k]; }, sphi, licL, false, ul);
  T* g = G + (size_t)ed * NE * 4;
  if (bc == BC_INTERIOR) {
    const int pr = pend ? pend[R] : -1;
    T ur[NE][4];
    edge_traces<T, NP, NE>([&](int v, int k) { return U[(size_t)(v * NP + k) * nes + R]; }, sphi, licR, true, ur);
    if (pr >= 0)
      edge_traces<T, NP, NE>([&](int v, int k) { return (T)pw[(size_t)pr * 4 * NP + v * NP + k]; }, sphi, licR, true,
                             ur);
    int err = 0;
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      T f[4];
      const int s = hllc(ul[q], ur[q], nx, ny, hcr, f);
      if (s && !err) err = s;
#pragma unroll
      for (int v = 0; v < 4; ++v) g[q * 4 + v] = f[v] * hl;
    }
    if (err) {
      if (s0)
        atomicMin(&cc->error_s0, ((unsigned long long)(unsigned)(ed - cc->ed_off) << 8) | (unsigned long long)err);
      else
        record_error(cc, ed - cc->ed_off, err);
    }
  } else if (bc == BC_WALL) {
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      const T p = (hcr - T(1)) * (ul[q][3] - T(0.5) * dv(ul[q][1] * ul[q][1] + ul[q][2] * ul[q][2], ul[q][0]));
      g[q * 4 + 0] = T(0);
      g[q * 4 + 1] = p * nx * hl;
      g[q * 4 + 2] = p * ny * hl;
      g[q * 4 + 3] = T(0);
    }
  } else {
    const double* qinf = m.patch_q + 4 * patch;
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      T f[4];
      farfield_flux(ul[q], qinf, nx, ny, hcr, f);
#pragma unroll
      for (int v = 0; v < 4; ++v) g[q * 4 + v] = f[v] * hl;
    }
  }
}

// ---------------------------------------------------------------------------------------
// element kernel: one lane per element — volume integral + flux lift + mass inverse + RK stage
// ---------------------------------------------------------------------------------------
template <class T, int NP, int NQ, int NE, int OP>
__global__ __launch_bounds__(256) void k_element(DevMesh m, StageBufs<T> b, const T* __restrict__ G,
                                                 const T* __restrict__ tab, Ctl ctl, int force, int remap) {
  using O = TabOff<NP, NQ, NE>;
  // the limiter pass of this stage builds its troubled list from zero
  if (OP != OP_RESID && blockIdx.x == 0 && threadIdx.x == 0) ctl.count[OP - OP_S0] = 0;
  const int blk = remap ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int e = blk * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  const int cs = case_of(m, m.el_case64, e);
  if (cs < 0) return;
  const CaseCtl& cc = ctl.cases[cs];
  // stage 0: an error of this step's edge pass (error_s0) stops the case here and becomes its error
  const unsigned long long es0 = OP == OP_S0 ? cc.error_s0 : ~0ull;
  if (!force && !(case_runs(cc) && es0 == ~0ull)) {
    if (OP == OP_S0 && es0 != ~0ull && cc.active && case_error(cc) == ~0ull && (threadIdx.x & 63) == 0)
      atomicMin(&ctl.cases[cs].error, es0);
    if (OP == OP_S0 && b.pend && cc.live && e < cc.e_end) write_back_pending<T, NP>(b, e, m.nes);
    return;
  }
  if (e >= cc.e_end) return;
  const int nes = m.nes;
  T u[4][NP];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int k = 0; k < NP; ++k) u[v][k] = ld32(b.in + (size_t)v * NP * nes, (unsigned)(k * nes + e));
  // a rebuild still pending from the limiter pass before: use it, and write it back (rare)
  if (OP != OP_RESID && b.pend) {
    const int sl = b.pend[e];
    if (sl >= 0) {
      const double* w = b.pw + (size_t)sl * 4 * NP;
      constexpr int PREV = OP == OP_S0 ? 2 : (OP == OP_S1 ? 0 : 1);
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int k = 0; k < NP; ++k) {
          u[v][k] = (T)w[v * NP + k];
          store_limited<T, NP, PREV>(b, nes, e, v * NP + k, w[v * NP + k]);
        }
    }
  }
  const T* eg = geo_el<T>(m);
  const unsigned eu = (unsigned)e;
  const T xx = ld32(eg + (size_t)G_XX * nes, eu), xy = ld32(eg + (size_t)G_XY * nes, eu);
  const T ex = ld32(eg + (size_t)G_EX * nes, eu), ey = ld32(eg + (size_t)G_EY * nes, eu);
  const T j2 = ld32(eg + (size_t)G_2J * nes, eu);
  const T hcr = (T)m.hcr;
  // FP32-bulk: point fluxes as deviations from the flux of the element's mean state (mode 0; deviation
  // form, see flux_deviation); FP64 keeps the reference's plain sums
  constexpr bool DEV = sizeof(T) == 4;
  float qm[4] = {0.f, 0.f, 0.f, 0.f};
  MeanState ms{0.f, 0.f, 0.f};
  if constexpr (DEV) {
#pragma unroll
    for (int v = 0; v < 4; ++v) qm[v] = (float)u[v][0];
    ms = mean_state(qm, (float)hcr);
  }
  T r[4][NP];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int k = 0; k < NP; ++k) r[v][k] = T(0);
  // volume integral with reference-element gradients: sum_q (F~ dphi/dxi + G~ dphi/deta) w_q / m_k
#pragma unroll 2
  for (int q = 0; q < NQ; ++q) {
    // the point value (FP32-bulk: its deviation, the modes above mode 0)
    T w[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      T s = T(0);
#pragma unroll
      for (int k = DEV ? 1 : 0; k < NP; ++k) s += u[v][k] * tab[O::PHIV + q * NP + k];
      w[v] = s;
    }
    T f[4], g[4];
    if constexpr (DEV)
      flux_deviation(qm, ms, w, hcr, f, g);
    else
      euler_flux(w, hcr, f, g);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      T ft, gt;
      if constexpr (DEV) {
        ft = metric(xx, xy, f[v], g[v]);
        gt = metric(ex, ey, f[v], g[v]);
      } else {
        ft = xx * f[v] + xy * g[v];
        gt = ex * f[v] + ey * g[v];
      }
#pragma unroll
      for (int k = 0; k < NP; ++k) r[v][k] += ft * tab[O::DXI + q * NP + k] + gt * tab[O::DETA + q * NP + k];
    }
  }
  // surface: gather the three edges' point fluxes (reversed point order on the right side)
  double fr[4] = {0.0, 0.0, 0.0, 0.0}, gr[4] = {0.0, 0.0, 0.0, 0.0};
  if constexpr (DEV) mean_flux64(qm, m.hcr, fr, gr);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int es = ld32(m.el_edge, (unsigned)(i * nes + e));
    const int edge = es >> 1, side = es & 1;
    const unsigned gb = (unsigned)edge * (unsigned)(NE * 4);
    const T sg = side ? j2 : -j2;
    // the edge's normal and half length, for the mean flux's surface term (FP32-bulk)
    float fnx = 0.f, fny = 0.f, fhl = 0.f;
    if constexpr (DEV) {
      fnx = ld32(m.ed_geo32, (unsigned)edge);
      fny = ld32(m.ed_geo32 + m.ed_stride, (unsigned)edge);
      fhl = ld32(m.ed_geo32 + 2 * (size_t)m.ed_stride, (unsigned)edge);
    }
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      const int src = side ? (NE - 1 - q) : q;
      T gv[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        if constexpr (DEV)
          gv[v] = surface_deviation(sg, ld32(G, gb + src * 4 + v), mean_normal_flux(fr[v], gr[v], fnx, fny, fhl));
        else
          gv[v] = sg * ld32(G, gb + src * 4 + v);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int k = 0; k < NP; ++k) r[v][k] += gv[v] * tab[O::LIFT + (i * NE + q) * NP + k];
    }
  }
  if (OP == OP_RESID) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int k = 0; k < NP; ++k) st32(b.rout + (size_t)v * NP * nes, (unsigned)(k * nes + e), (double)r[v][k]);
    return;
  }
  const double dt = cc.dt;
  // The RK operands of one variable are read together before its stores (loads after a store to
  // another array would otherwise each wait for the one before).
  if constexpr (sizeof(T) == 8) {
    // src/temporal_disc/rk3.rs:10,16,22
    // batches of half a variable, kept below the flux gather (mixed into it they raise the register
    // count above what keeps two waves per SIMD)
    constexpr int HB = (NP + 1) / 2;
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (OP == OP_S0) {
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int k = 0; k < NP; ++k)
          st32(b.us + (size_t)v * NP * nes, (unsigned)(k * nes + e), (double)u[v][k] + dt * (double)r[v][k]);
    } else
#pragma unroll
    for (int c0 = 0; c0 < 4 * NP; c0 += HB) {
      double un0[HB];
      if (OP != OP_S0) {
#pragma unroll
        for (int c = 0; c < HB; ++c)
          if (c0 + c < 4 * NP) un0[c] = ld32(b.un + (size_t)((c0 + c) / NP) * NP * nes, (unsigned)(((c0 + c) % NP) * nes + e));
      }
#pragma unroll
      for (int c = 0; c < HB; ++c) {
        if (c0 + c >= 4 * NP) break;
        const int v = (c0 + c) / NP, k = (c0 + c) % NP;
        const size_t vb = (size_t)v * NP * nes;
        const unsigned ix = (unsigned)(k * nes + e);
        const double rr = (double)r[v][k], uu = (double)u[v][k];
        if (OP == OP_S0)
          st32(b.us + vb, ix, uu + dt * rr);
        else if (OP == OP_S1)
          st32(b.us + vb, ix, 0.75 * un0[c] + 0.25 * (uu + dt * rr));
        else
          st32(b.un + vb, ix, 1.0 / 3.0 * un0[c] + 2.0 / 3.0 * (uu + dt * rr));
      }
    }
  } else {
    // FP32-bulk in increment form: u^(1) = u^n + d1, d1 = dt R(u^n); u^(2) = u^n + d2,
    // d2 = (d1 + dt R(u^(1))) / 4; u^(n+1) = u^n + 2 (d2 + dt R(u^(2))) / 3 accumulated in FP64.
    const float dtf = (float)dt;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      float d0[NP];
      double un0[NP];
      if (OP != OP_S0) {
#pragma unroll
        for (int k = 0; k < NP; ++k) {
          d0[k] = ld32(b.d + (size_t)v * NP * nes, (unsigned)(k * nes + e));
          if (OP == OP_S2) un0[k] = ld32(b.un + (size_t)v * NP * nes, (unsigned)(k * nes + e));
        }
      }
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const size_t vb = (size_t)v * NP * nes;
        const unsigned ix = (unsigned)(k * nes + e);
        const float rr = (float)r[v][k];
        if (OP == OP_S0) {
          const float d = dtf * rr;
          st32(b.d + vb, ix, d);
          st32(b.s + vb, ix, (float)u[v][k] + d);
        } else if (OP == OP_S1) {
          const float d = 0.25f * (d0[k] + dtf * rr);
          st32(b.d + vb, ix, d);
          st32(b.s2 + vb, ix, (float)u[v][k] + (d - d0[k]));  // one rounding at the state's scale
        } else {
          const double d = 2.0 / 3.0 * ((double)d0[k] + dt * (double)rr);
          const double un = un0[k] + d;
          st32(b.un + vb, ix, un);
          st32(b.sn + vb, ix, (float)un);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// FP32-bulk element kernel on the matrix cores: one wavefront per 16 elements.
//
// With c = lane % 16 (the element e0 + c) and g = lane / 16, each contraction is a
// v_mfma_f32_16x16x4f32 product over a K slice of 4 (A: lane holds A[c][g]; B: lane holds B[g][c];
// D: lane holds D[4g + r][c], r = 0..3), written with the element index as the 16 columns:
//   values    u_v^T[q][e]  = sum_k PHI^T[q][k] U_v^T[k][e]          K slice s: k = 4g + s
//   volume    R_v^T[k][e] += sum_q DXI^T[k][q] Ft_v^T[q][e] + DETA^T[k][q] Gt_v^T[q][e]
//                                                                   K slice s: q = 4g + s
//   surface   R_v^T[k][e] += sum_m' LIFT^T[k][m'] S_v^T[m'][e]      K slice s: m' = (face, point) = 4g + s
// The K order of the volume product is chosen so that each lane's B operand in slice s is the
// flux it computed itself at point q = 4g + s (the output rows of the value product): no data
// moves between lanes. Lanes also hold exactly the (k = 4g + r, e) state entries the RK update
// needs. Requires Np <= 16, Nq <= 16, 3 Ne <= 16 (p <= 3).
// ---------------------------------------------------------------------------------------
typedef float f32x4 __attribute__((ext_vector_type(4)));

// x where p holds, else +0: a bit mask rather than a select, so that the compiler does not turn the
// load that produced x into a branch around it (it does so for selects with a load operand)
__device__ __forceinline__ float keep_if(bool p, float x) { return __int_as_float(__float_as_int(x) & -(int)p); }

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <int NP, int NQ, int NE, int OP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_element_mfma(DevMesh m, StageBufs<float> b, const float* __restrict__ G,
                                                      const float* __restrict__ tab, Ctl ctl, int force, int remap) {
  static_assert(NP <= 16 && NQ <= 16 && 3 * NE <= 16, "matrix-core element kernel needs p <= 3");
  using O = TabOff<NP, NQ, NE>;
  if (OP != OP_RESID && blockIdx.x == 0 && threadIdx.x == 0) ctl.count[OP - OP_S0] = 0;
  const int blk = remap ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int tid = blk * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int wave = tid >> 6;
  // the wavefront's 16 elements lie in one 64-element block, hence in one case (uniform exit)
  const int cs = case_of(m, m.el_case64, wave * 16);
  if (cs < 0) return;
  const CaseCtl& cc = ctl.cases[cs];
  // the fields of the case record this kernel uses, read first: the exit test below then waits on
  // them alone, not on the loads issued after them
  const int c_active = cc.active, c_live = cc.live, e_end = cc.e_end;
  const unsigned long long c_error = case_error(cc);
  // stage 0: an error of this step's edge pass (error_s0) stops the case here and becomes its error
  const unsigned long long c_es0 = OP == OP_S0 ? cc.error_s0 : ~0ull;
  const double dt = cc.dt;
  const int c = lane & 15, g = lane >> 4;
  const int e = wave * 16 + c;
  const int nes = m.nes;
  // Every load below reads from a clamped in-range index and the unused entries are selected away
  // afterwards, so the loads carry no branches and are issued together: the state, the table
  // operands, the edge ids of the lift gather and the metric go out before the case record is back.
  const int es = e < m.ne ? e : 0;
  // edge ids of the lane's (face, point) slots 4g + s of the lift product, first: the flux gather waits
  // on them alone
  int eid[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int mp = 4 * g + s;
    eid[s] = ld32(m.el_edge, (unsigned)((mp < 3 * NE ? mp / NE : 0) * nes + es));
  }
  // state entries held by this lane: U_v[e][k = 4g + s]
  float us[4][4];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int k = 4 * g + s;
      const float x = ld32(b.in + (size_t)v * NP * nes, (unsigned)((k < NP ? k : NP - 1) * nes + es));
      us[v][s] = keep_if(k < NP, x);
    }
  // per-lane table operands (A of the value and volume products; the lift operands are read later)
  float aPhi[4], aDxi[4], aDeta[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int k = 4 * g + s;  // value product: K index = mode; volume: K index = point
    const int kp = k < NP ? k : NP - 1, kq = k < NQ ? k : NQ - 1;
    const int cq = c < NQ ? c : NQ - 1, cp = c < NP ? c : NP - 1;
    const float ph = ld32(tab, (unsigned)(O::PHIV + cq * NP + kp));
    const float dx = ld32(tab, (unsigned)(O::DXI + kq * NP + cp)), de = ld32(tab, (unsigned)(O::DETA + kq * NP + cp));
    aPhi[s] = keep_if(k > 0 && k < NP && c < NQ, ph);  // the value product sums the modes above mode 0
    aDxi[s] = keep_if(k < NQ && c < NP, dx);
    aDeta[s] = keep_if(k < NQ && c < NP, de);
  }
  const float* eg = m.el_geo32;
  const unsigned eu = (unsigned)es;
  const float xx = ld32(eg + (size_t)G_XX * nes, eu), xy = ld32(eg + (size_t)G_XY * nes, eu);
  const float ex = ld32(eg + (size_t)G_EX * nes, eu), ey = ld32(eg + (size_t)G_EY * nes, eu);
  const float j2 = ld32(eg + (size_t)G_2J * nes, eu);
  const int pend_slot = (OP != OP_RESID && b.pend) ? ld32(b.pend, eu) : -1;
  // the RK operands (d, and u^n in stage 3) load after the volume product, below: held from here they
  // would take the registers the deviation form's mean state and flux deviations need (a pending
  // write-back stores d first, and the loads then read it)
  float d0[4][4];
  double un0[4][4];
  const bool runs = (force != 0) | ((c_active != 0) & (c_error == ~0ull) & (c_es0 == ~0ull));  // case_runs()
  if (!runs) {
    if (OP == OP_S0 && c_es0 != ~0ull && c_active && c_error == ~0ull && lane == 0)
      atomicMin(&ctl.cases[cs].error, c_es0);
    const int ew = wave * 16 + (lane & 15);
    if (OP == OP_S0 && b.pend && c_live && lane < 16 && ew < m.ne && ew < e_end)
      write_back_pending<float, NP>(b, ew, m.nes);
    return;
  }
  const bool valid = (e < m.ne) & (e < e_end);
  // surface: B operand = signed, 2/J-scaled point flux of (face, point) = 4g + s, point order reversed on
  // the right. The four gathers go out together as soon as the edge ids are in, ahead of the volume
  // product (unused slots read edge 0 and are masked to 0 when used).
  float4 fl[4];
  bool use[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int mp = 4 * g + s;
    use[s] = valid && mp < 3 * NE;
    // the address depends on the edge ids only (not on the case record): any in-range element's ids
    const int ed = (e < m.ne && mp < 3 * NE) ? eid[s] : 0;
    const int pt = mp % NE;
    const int src = (ed & 1) ? (NE - 1 - pt) : pt;
    fl[s] = ld32(reinterpret_cast<const float4*>(G), (unsigned)((ed >> 1) * NE + src));
  }
  // normal and half length of the (at most two) faces of the lane's slots, for the mean flux's surface
  // term: face a of slot 0, face b of slot 3
  float gxa, gya, gha, gxb, gyb, ghb;
  {
    const unsigned ea = (unsigned)(((e < m.ne) ? eid[0] : 0) >> 1);
    const unsigned eb = (unsigned)(((e < m.ne && 4 * g + 3 < 3 * NE) ? eid[3] : 0) >> 1);
    gxa = ld32(m.ed_geo32, ea);
    gya = ld32(m.ed_geo32 + m.ed_stride, ea);
    gha = ld32(m.ed_geo32 + 2 * (size_t)m.ed_stride, ea);
    gxb = ld32(m.ed_geo32, eb);
    gyb = ld32(m.ed_geo32 + m.ed_stride, eb);
    ghb = ld32(m.ed_geo32 + 2 * (size_t)m.ed_stride, eb);
  }
  // a rebuild still pending from the limiter pass before: use it and write it back, each lane the
  // modes it holds (rare)
  if (OP != OP_RESID && b.pend && valid) {
    const int sl = pend_slot;
    if (sl >= 0) {
      const double* w = b.pw + (size_t)sl * 4 * NP;
      constexpr int PREV = OP == OP_S0 ? 2 : (OP == OP_S1 ? 0 : 1);
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int k = 4 * g + s;
          if (k < NP) {
            us[v][s] = (float)w[v * NP + k];
            store_limited<float, NP, PREV>(b, nes, e, v * NP + k, w[v * NP + k]);
          }
        }
    }
  }
  // deviations of the values at the cell points q = 4g + r of element e from its mean (mode 0 is left
  // out of the product)
  f32x4 uq[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    uq[v] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; ++s) uq[v] = mfma4(aPhi[s], us[v][s], uq[v]);
  }
  const float hcr = (float)m.hcr;
  // the element's mean state (mode 0 sits in lane c, the g = 0 lane of element c): every point flux enters
  // as its deviation from the mean's flux (deviation form, see flux_deviation)
  float qm[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) qm[v] = __shfl(us[v][0], c, 64);
  // flux deviations at the lane's four points, in the metric-transformed form
  float ft[4][4], gt[4][4];  // [var][r]
  {
    const MeanState ms = mean_state(qm, hcr);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float dq[4] = {uq[0][r], uq[1][r], uq[2][r], uq[3][r]};
      const bool live = valid && (4 * g + r) < NQ;
      float f[4], gg[4];
      flux_deviation(qm, ms, dq, hcr, f, gg);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        ft[v][r] = live ? metric(xx, xy, f[v], gg[v]) : 0.f;
        gt[v][r] = live ? metric(ex, ey, f[v], gg[v]) : 0.f;
      }
    }
  }
  f32x4 acc[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    acc[v] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      acc[v] = mfma4(aDxi[s], ft[v][s], acc[v]);
      acc[v] = mfma4(aDeta[s], gt[v][s], acc[v]);
    }
  }
  if (OP == OP_S1 || OP == OP_S2) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = 4 * g + r;
        const unsigned ix = (unsigned)((k < NP ? k : NP - 1) * nes + es);
        d0[v][r] = ld32(b.d + (size_t)v * NP * nes, ix);
        if (OP == OP_S2) un0[v][r] = ld32(b.un + (size_t)v * NP * nes, ix);
      }
  }
  // lift operands: K index = face * NE + point
  float aLift[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int k = 4 * g + s;
    const int km = k < 3 * NE ? k : 3 * NE - 1, cp = c < NP ? c : NP - 1;
    aLift[s] = keep_if(k < 3 * NE && c < NP, ld32(tab, (unsigned)(O::LIFT + km * NP + cp)));
  }
  // the mean flux (FP64), whose normal component times the half length each slot subtracts: formed per
  // variable for the (at most two) faces a and b of the lane's slots, so that only two such values are
  // held at a time
  double fr[4], gr[4];
  mean_flux64(qm, m.hcr, fr, gr);
  float sv[4][4];  // [s][var]
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const double rna = mean_normal_flux(fr[v], gr[v], gxa, gya, gha);
    const double rnb = mean_normal_flux(fr[v], gr[v], gxb, gyb, ghb);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const float sc = (eid[s] & 1) ? j2 : -j2;
      // slot s lies on face a when it shares slot 0's face, else on face b
      const bool fa = (4 * g + s) / NE == (4 * g) / NE;
      const float f = v == 0 ? fl[s].x : (v == 1 ? fl[s].y : (v == 2 ? fl[s].z : fl[s].w));
      sv[s][v] = keep_if(use[s], surface_deviation(sc, f, fa ? rna : rnb));
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int s = 0; s < 4; ++s) acc[v] = mfma4(aLift[s], sv[s][v], acc[v]);
  if (!valid) return;
  // lane holds R_v[e][k = 4g + r]
  if (OP == OP_RESID) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = 4 * g + r;
        if (k < NP) st32(b.rout + (size_t)v * NP * nes, (unsigned)(k * nes + e), (double)acc[v][r]);
      }
    return;
  }
  const float dtf = (float)dt;
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = 4 * g + r;
      if (k >= NP) continue;
      const unsigned ix = (unsigned)(k * nes + e);
      const size_t vb = (size_t)v * NP * nes;
      const float rr = acc[v][r];
      if (OP == OP_S0) {
        const float d = dtf * rr;
        st32(b.d + vb, ix, d);
        st32(b.s + vb, ix, us[v][r] + d);
      } else if (OP == OP_S1) {
        // u^(2) shadow = u^(1) shadow + (d2 - d1) (= s^n + d2 up to float rounding); no s^n re-read. The
        // increments' difference is formed first, so that the sum rounds once at the state's scale
        const float d = 0.25f * (d0[v][r] + dtf * rr);
        st32(b.d + vb, ix, d);
        st32(b.s2 + vb, ix, us[v][r] + (d - d0[v][r]));
      } else {
        const double d = 2.0 / 3.0 * ((double)d0[v][r] + dt * (double)rr);
        const double un = un0[v][r] + d;
        st32(b.un + vb, ix, un);
        st32(b.sn + vb, ix, (float)un);
      }
    }
}

// ---------------------------------------------------------------------------------------
// Fused edge-and-element pass of a march stage (FP32-bulk, matrix cores): one wavefront per group of 16
// elements evaluates the numerical fluxes of the group's distinct edges into LDS, then runs the element
// pass of k_element_mfma on them. Nothing goes through the point-flux array, and the edge work is
// spread over the wavefront's 64 lanes by (edge, side, variable) for the traces and by (edge, point)
// for the fluxes, so the lanes stay busy whatever the group's edge count.
//  1. traces: item (j, side, v) is variable v of the left or right element of edge j at the edge points,
//     the operations of k_surface's edge_traces in the same order (a rebuild still pending from the
//     limiter pass before is read instead of the stored modes, as there);
//  2. fluxes: item (j, q) is point q of edge j: HLLC, the slip wall or the characteristic far field
//     (src/spatial_disc/flux.rs:28-153, src/spatial_disc/boundary_condition.rs:9-97) times the half
//     length, as k_surface evaluates them, written over the left trace of the same point;
//  3. the element pass (volume product, lift from LDS, mass inverse, RK stage) of k_element_mfma.
// An edge shared by two groups is evaluated by both from the same inputs with the same operations, so
// both elements receive the same flux, bit for bit equal to the separate edge pass's; the pass reads
// only its stage's input arrays (which the stage does not write) for the traces of other groups'
// elements. A failing edge is recorded into CaseCtl::error_f[stage], which no wavefront of the same
// pass reads: every wavefront of a case takes the same decision to run.
// ---------------------------------------------------------------------------------------
template <int NE>
struct FusedWave {
  float4 tr[GRP_EDGES][2][NE];  // traces of the left / right element at the edge points; the flux of a
                                // point then replaces its left trace
  int4 rec[GRP_EDGES];          // the group's edge records (grp_rec, grp_geo)
  float4 geo[GRP_EDGES];
  unsigned char err[GRP_EDGES][NE];
};

// LDS of one block of the fused pass: a FusedWave per wavefront, then the edge basis table
template <int NP, int NE>
constexpr size_t fused_lds_bytes(int waves) {
  return (size_t)waves * sizeof(FusedWave<NE>) + (size_t)3 * NE * NP * sizeof(float);
}

// completes a wavefront's LDS writes before its lanes read each other's
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NP, int NQ, int NE, int OP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_fused(DevMesh m, StageBufs<float> b,
                                                                                       const float* __restrict__ tab,
                                                                                       Ctl ctl, int force, int remap) {
  static_assert(NP <= 16 && NQ <= 16 && 3 * NE <= 16, "matrix-core element kernel needs p <= 3");
  static_assert(OP != OP_RESID, "the fused pass is a march stage");
  static_assert(GRP_EDGES <= 64, "one edge record per lane");
  using O = TabOff<NP, NQ, NE>;
  constexpr int STAGE = OP - OP_S0;
  extern __shared__ __align__(16) unsigned char fused_lds[];
  const int wpb = (int)(blockDim.x >> 6);
  FusedWave<NE>& sw = reinterpret_cast<FusedWave<NE>*>(fused_lds)[threadIdx.x >> 6];
  float* sphi = reinterpret_cast<float*>(fused_lds + (size_t)wpb * sizeof(FusedWave<NE>));
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl.count[STAGE] = 0;
  const int blk = remap ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int wave = (blk * (int)blockDim.x + (int)threadIdx.x) >> 6;  // the group: elements 16 wave ..
  const int nes = m.nes;
  // First, in one trip, everything that depends on the group alone (every index is in range: the
  // tables have the arena's capacity): the case, the edge count and records, the edge basis table, and
  // the element pass's face slots, metric and pending mark. The case record follows the case id.
  const int cs = case_of(m, m.el_case64, wave * 16);
  const int n = __builtin_amdgcn_readfirstlane(m.grp_n[wave]);
  const int4 rec = m.grp_rec[(size_t)wave * GRP_EDGES + (lane < GRP_EDGES ? lane : 0)];
  const float4 geo = m.grp_geo[(size_t)wave * GRP_EDGES + (lane < GRP_EDGES ? lane : 0)];
  constexpr int TN = 3 * NE * NP, TI = (TN + 63) / 64;  // table entries, and per thread at most
  float tphi[TI];
#pragma unroll
  for (int i = 0; i < TI; ++i) {
    const int ix = (int)threadIdx.x + (int)blockDim.x * i;
    tphi[i] = ix < TN ? tab[O::PHIE + ix] : 0.f;
  }
  const int c = lane & 15, g = lane >> 4;
  const int e = wave * 16 + c;
  const int es = e < m.ne ? e : 0;
  int gsl[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int mp = 4 * g + s;
    gsl[s] = ld32(m.el_gslot, (unsigned)((mp < 3 * NE ? mp / NE : 0) * nes + es));
  }
  const float* elg = m.el_geo32;
  const unsigned eu = (unsigned)es;
  const float xx = ld32(elg + (size_t)G_XX * nes, eu), xy = ld32(elg + (size_t)G_XY * nes, eu);
  const float ex = ld32(elg + (size_t)G_EX * nes, eu), ey = ld32(elg + (size_t)G_EY * nes, eu);
  const float j2 = ld32(elg + (size_t)G_2J * nes, eu);
  const int pend_slot = b.pend ? ld32(b.pend, eu) : -1;
  const CaseCtl& cc = ctl.cases[cs < 0 ? 0 : cs];
  const int c_active = cc.active, c_live = cc.live, e_end = cc.e_end, ed_off = cc.ed_off;
  unsigned long long c_error = cc.error;
  // the fused passes of the earlier stages of this step, never this one's own word
  if (STAGE >= 1) c_error = cc.error_f[0] < c_error ? cc.error_f[0] : c_error;
  if (STAGE >= 2) c_error = cc.error_f[1] < c_error ? cc.error_f[1] : c_error;
  const double dt = cc.dt;
  // the table, and the edge records for the lanes of the trace and flux items
#pragma unroll
  for (int i = 0; i < TI; ++i) {
    const int ix = (int)threadIdx.x + (int)blockDim.x * i;
    if (ix < TN) sphi[ix] = tphi[i];
  }
  if (lane < GRP_EDGES) {
    sw.rec[lane] = rec;
    sw.geo[lane] = geo;
  }
  __syncthreads();  // the only block barrier, before any exit
  if (cs < 0) return;  // the wavefront's 16 elements lie in one 64-element block, hence in one case
  const bool runs = (force != 0) | ((c_active != 0) & (c_error == ~0ull));  // case_runs()
  if (!runs) {
    // a stopped case writes back the rebuilds still pending from its last limiter pass (stage 0)
    const int ew = wave * 16 + lane;
    if (OP == OP_S0 && b.pend && c_live && lane < 16 && ew < m.ne && ew < e_end) write_back_pending<float, NP>(b, ew, nes);
    return;
  }
  // ---- 1. traces: items t = (v * 2 + side) * n + j, so that the lanes of one load cover consecutive
  // edges of one side and variable (t / n by a multiply, exact for t < 8 * GRP_EDGES) ----
  // (in two halves of three 64-item rounds, each half's loads issued together)
  constexpr int R = 3;
  static_assert(2 * R * 64 >= 8 * GRP_EDGES, "two halves cover every trace item");
  const unsigned mn = (65536u + (unsigned)n - 1u) / (unsigned)(n > 0 ? n : 1);
#pragma unroll 1
  for (int h = 0; h < 2 && 64 * R * h < 8 * n; ++h) {
    float cv[R][NP];
    int xs[R], ps[R], its[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      xs[r] = -1;
      ps[r] = -1;
      its[r] = 0;
      if (64 * (R * h + r) < 8 * n) {  // uniform
        const int t = lane + 64 * (R * h + r);
        const int vs = (int)(((unsigned)t * mn) >> 16), j = t - vs * n;
        const bool item = t < 8 * n;
        const int4 er = sw.rec[item ? j : 0];
        const int x = item ? ((vs & 1) ? er.y : er.x) : -1;
        const int xa = x >= 0 ? x : 0;
        const int v = (vs >> 1) & 3;
#pragma unroll
        for (int k = 0; k < NP; ++k) cv[r][k] = ld32(b.in, (unsigned)((v * NP + k) * nes + xa));
        xs[r] = x;
        ps[r] = b.pend ? b.pend[xa] : -1;
        its[r] = (vs << 8) | j;
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (xs[r] >= 0) {
        const int j = its[r] & 255, vs = its[r] >> 8, side = vs & 1, v = (vs >> 1) & 3;
        const int info = sw.rec[j].z;
        const int lic = side ? (info >> 2) & 3 : info & 3;
        float tr[NE];
#pragma unroll
        for (int q = 0; q < NE; ++q) tr[q] = 0.f;
        if (ps[r] >= 0) {
          // a rebuild still pending from the limiter pass before (rare)
          const double* w = b.pw + (size_t)ps[r] * 4 * NP + v * NP;
#pragma unroll
          for (int k = 0; k < NP; ++k) {
            const float cw = (float)w[k];
#pragma unroll
            for (int q = 0; q < NE; ++q) tr[q] += cw * sphi[(lic * NE + (side ? NE - 1 - q : q)) * NP + k];
          }
        } else {
#pragma unroll
          for (int k = 0; k < NP; ++k) {
#pragma unroll
            for (int q = 0; q < NE; ++q) tr[q] += cv[r][k] * sphi[(lic * NE + (side ? NE - 1 - q : q)) * NP + k];
          }
        }
#pragma unroll
        for (int q = 0; q < NE; ++q) reinterpret_cast<float*>(&sw.tr[j][side][q])[v] = tr[q];
      }
    }
  }
  // the element pass's state entries (lane: element c, modes 4g + s) and table operands go out here, to
  // arrive while the fluxes are computed
  float us[4][4];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int k = 4 * g + s;
      const float x = ld32(b.in + (size_t)v * NP * nes, (unsigned)((k < NP ? k : NP - 1) * nes + es));
      us[v][s] = keep_if(k < NP, x);
    }
  float aPhi[4], aDxi[4], aDeta[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int k = 4 * g + s;
    const int kp = k < NP ? k : NP - 1, kq = k < NQ ? k : NQ - 1;
    const int cq = c < NQ ? c : NQ - 1, cp = c < NP ? c : NP - 1;
    const float ph = ld32(tab, (unsigned)(O::PHIV + cq * NP + kp));
    const float dx = ld32(tab, (unsigned)(O::DXI + kq * NP + cp)), de = ld32(tab, (unsigned)(O::DETA + kq * NP + cp));
    aPhi[s] = keep_if(k > 0 && k < NP && c < NQ, ph);  // the value product sums the modes above mode 0
    aDxi[s] = keep_if(k < NQ && c < NP, dx);
    aDeta[s] = keep_if(k < NQ && c < NP, de);
  }
  wave_lds_sync();
  // ---- 2. fluxes: items t = (j, q) ----
  {
    const float hcr = (float)m.hcr;
    for (int t = lane; t < NE * n; t += 64) {
      const int j = t / NE, q = t - j * NE;
      const int info = sw.rec[j].z;
      const int bc = (info >> 4) & 3;
      const float4 ge = sw.geo[j];
      const float nx = ge.x, ny = ge.y, hl = ge.z;
      const float4 a = sw.tr[j][0][q];
      const float ul[4] = {a.x, a.y, a.z, a.w};
      int st = 0;
      float4 gq;
      if (bc == BC_INTERIOR) {
        const float4 r4 = sw.tr[j][1][q];
        const float ur[4] = {r4.x, r4.y, r4.z, r4.w};
        float f[4] = {0.f, 0.f, 0.f, 0.f};
        st = hllc(ul, ur, nx, ny, hcr, f);
        gq = make_float4(f[0] * hl, f[1] * hl, f[2] * hl, f[3] * hl);
      } else if (bc == BC_WALL) {
        const float p = (hcr - 1.f) * (ul[3] - 0.5f * dv(ul[1] * ul[1] + ul[2] * ul[2], ul[0]));
        gq = make_float4(0.f, p * nx * hl, p * ny * hl, 0.f);
      } else {
        float f[4];
        farfield_flux(ul, m.patch_q + 4 * (info >> 8), nx, ny, hcr, f);
        gq = make_float4(f[0] * hl, f[1] * hl, f[2] * hl, f[3] * hl);
      }
      sw.tr[j][0][q] = gq;
      sw.err[j][q] = (unsigned char)st;
    }
  }
  wave_lds_sync();
  // the first failing point of each failing edge, as k_surface records it
  if (lane < n) {
    int err = 0;
#pragma unroll
    for (int q = 0; q < NE; ++q) err = err ? err : (int)sw.err[lane][q];
    if (err)
      atomicMin(&ctl.cases[cs].error_f[STAGE],
                ((unsigned long long)(unsigned)(rec.w - ed_off) << 8) | (unsigned long long)err);
  }
  // ---- 3. the element pass (k_element_mfma, with the fluxes from LDS) ----
  // the RK operands load after the volume product, as in k_element_mfma
  float d0[4][4];
  double un0[4][4];
  const bool valid = (e < m.ne) & (e < e_end);
  float4 fl[4];
  bool use[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int mp = 4 * g + s;
    use[s] = valid && mp < 3 * NE;
    const int slot = use[s] ? (gsl[s] >> 1) : 0;
    const int pt = mp % NE;
    const int src = (gsl[s] & 1) ? (NE - 1 - pt) : pt;
    fl[s] = sw.tr[slot][0][src];
  }
  if (b.pend && valid) {
    const int sl = pend_slot;
    if (sl >= 0) {
      const double* w = b.pw + (size_t)sl * 4 * NP;
      constexpr int PREV = OP == OP_S0 ? 2 : (OP == OP_S1 ? 0 : 1);
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int k = 4 * g + s;
          if (k < NP) {
            us[v][s] = (float)w[v * NP + k];
            store_limited<float, NP, PREV>(b, nes, e, v * NP + k, w[v * NP + k]);
          }
        }
    }
  }
  f32x4 uq[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    uq[v] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; ++s) uq[v] = mfma4(aPhi[s], us[v][s], uq[v]);
  }
  const float hcr = (float)m.hcr;
  // the deviation form of k_element_mfma, from the same values
  float qm[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) qm[v] = __shfl(us[v][0], c, 64);
  float ft[4][4], gt[4][4];  // [var][r]
  {
    const MeanState ms = mean_state(qm, hcr);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float dq[4] = {uq[0][r], uq[1][r], uq[2][r], uq[3][r]};
      const bool live = valid && (4 * g + r) < NQ;
      float f[4], gg[4];
      flux_deviation(qm, ms, dq, hcr, f, gg);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        ft[v][r] = live ? metric(xx, xy, f[v], gg[v]) : 0.f;
        gt[v][r] = live ? metric(ex, ey, f[v], gg[v]) : 0.f;
      }
    }
  }
  f32x4 acc[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    acc[v] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      acc[v] = mfma4(aDxi[s], ft[v][s], acc[v]);
      acc[v] = mfma4(aDeta[s], gt[v][s], acc[v]);
    }
  }
  if (OP == OP_S1 || OP == OP_S2) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = 4 * g + r;
        const unsigned ix = (unsigned)((k < NP ? k : NP - 1) * nes + es);
        d0[v][r] = ld32(b.d + (size_t)v * NP * nes, ix);
        if (OP == OP_S2) un0[v][r] = ld32(b.un + (size_t)v * NP * nes, ix);
      }
  }
  float aLift[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int k = 4 * g + s;
    const int km = k < 3 * NE ? k : 3 * NE - 1, cp = c < NP ? c : NP - 1;
    aLift[s] = keep_if(k < 3 * NE && c < NP, ld32(tab, (unsigned)(O::LIFT + km * NP + cp)));
  }
  double fr[4], gr[4];
  mean_flux64(qm, m.hcr, fr, gr);
  float sv[4][4];  // [s][var]
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const float sc = (gsl[s] & 1) ? j2 : -j2;
    // the edge's normal and half length: lane `slot` holds the group's record of that edge
    const int slot = use[s] ? (gsl[s] >> 1) : 0;
    const float fnx = __shfl(geo.x, slot, 64), fny = __shfl(geo.y, slot, 64), fhl = __shfl(geo.z, slot, 64);
    sv[s][0] = keep_if(use[s], surface_deviation(sc, fl[s].x, mean_normal_flux(fr[0], gr[0], fnx, fny, fhl)));
    sv[s][1] = keep_if(use[s], surface_deviation(sc, fl[s].y, mean_normal_flux(fr[1], gr[1], fnx, fny, fhl)));
    sv[s][2] = keep_if(use[s], surface_deviation(sc, fl[s].z, mean_normal_flux(fr[2], gr[2], fnx, fny, fhl)));
    sv[s][3] = keep_if(use[s], surface_deviation(sc, fl[s].w, mean_normal_flux(fr[3], gr[3], fnx, fny, fhl)));
  }
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int s = 0; s < 4; ++s) acc[v] = mfma4(aLift[s], sv[s][v], acc[v]);
  if (!valid) return;
  const float dtf = (float)dt;
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = 4 * g + r;
      if (k >= NP) continue;
      const unsigned ix = (unsigned)(k * nes + e);
      const size_t vb = (size_t)v * NP * nes;
      const float rr = acc[v][r];
      if (OP == OP_S0) {
        const float d = dtf * rr;
        st32(b.d + vb, ix, d);
        st32(b.s + vb, ix, us[v][r] + d);
      } else if (OP == OP_S1) {
        const float d = 0.25f * (d0[v][r] + dtf * rr);
        st32(b.d + vb, ix, d);
        st32(b.s2 + vb, ix, us[v][r] + (d - d0[v][r]));
      } else {
        const double d = 2.0 / 3.0 * ((double)d0[v][r] + dt * (double)rr);
        const double un = un0[v][r] + d;
        st32(b.un + vb, ix, un);
        st32(b.sn + vb, ix, (float)un);
      }
    }
}

// ---------------------------------------------------------------------------------------
// time step: TemperalDisc::compute_time_step (wave speed from the sum of ALL modes, src/temporal_disc.rs:63-79)
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// dt of one element, src/temporal_disc.rs:63-81: wave speed of the state given by the SUM of all modes
template <int NP, class F>
__device__ __forceinline__ double element_dt(F coef, double hcr, double cfl, double hmin) {
  double s[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    double a = 0.0;
#pragma unroll
    for (int k = 0; k < NP; ++k) a += coef(v, k);
    s[v] = a;
  }
  const double u = s[1] / s[0], v = s[2] / s[0];
  const double p = (hcr - 1.0) * (s[3] - 0.5 * s[0] * (u * u + v * v));
  const double c = sqrt(hcr * p / s[0]);
  const double speed = sqrt(u * u + v * v) + c;
  return cfl * hmin / speed;
}

// block minimum of per-thread dt candidates, then one atomicMin per block (positive doubles order
// like their bit patterns); a NaN candidate is never taken, as in `if dt < time_step` of the reference
// every lane of the wavefront must reach this call
__device__ __forceinline__ void wave_min_dt(double dt, unsigned long long* slot_base, int wave_index) {
  dt = wave_min(dt);
  if ((threadIdx.x & 63) == 0 && dt < 1.0e10)
    atomicMin(&slot_base[(wave_index % DT_LANES) * DT_STRIDE], (unsigned long long)__double_as_longlong(dt));
}

// stand-alone time-step reduction into the slot of the step about to begin
// (one wavefront = one 64-element block = one case; every lane of it reaches the reduction)
template <class T, int NP>
__global__ __launch_bounds__(256) void k_dt(DevMesh m, const T* __restrict__ U, Ctl ctl, int force, int only) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if ((e >> 6) >= (m.nes >> 6)) return;
  const int cs = case_of(m, m.el_case64, e);
  if (cs < 0 || (only >= 0 && cs != only)) return;
  const CaseCtl& cc = ctl.cases[cs];
  if (!force && !cc.live) return;
  double dt = 1.0e10;
  if (e < m.ne && e < cc.e_end) {
    const int nes = m.nes;
    const double d = element_dt<NP>([&](int v, int k) { return (double)U[(size_t)(v * NP + k) * nes + e]; }, m.hcr,
                                    cc.cfl, m.el_geo[G_HMIN * nes + e]);
    if (d < dt) dt = d;
  }
  wave_min_dt(dt, dt_slot(ctl, cs, cc.step + cc.pending), e >> 6);
}

// Step control (src/temporal_disc.rs:22-34): count the step that just ran, take its successor's dt
// (clamped to final_time), and decide whether it runs (step < final_step && t < final_time).
// One wavefront per case slot i: lanes 0 .. DT_LANES - 1 read and reset the case's dt words in parallel.
// It reads error and the fused passes' words (case_error) but never error_s0, so the stage-0 edge pass,
// which records into error_s0, can run in the same launch (k_surface) without changing what it decides.
__device__ __forceinline__ void step_control(const Ctl& ctl, int i, int lane) {
  static_assert(2 * DT_LANES == 64, "one wavefront reads the dt words of both step parities");
  CaseCtl& c = ctl.cases[i];
  // The record and the dt words of both step parities are read together (lane l: parity l / 32, word
  // l % 32); the parity of the step about to begin then selects one half, which is reset. Nothing below
  // returns early, so no load waits behind another.
  unsigned long long* word = dt_slot(ctl, i, lane >> 5) + (size_t)(lane & (DT_LANES - 1)) * DT_STRIDE;
  const unsigned long long raw = *word;
  const int live = c.live, pending = c.pending, step0 = c.step, final_step = c.final_step, cnt2 = c.count[2];
  const double t0 = c.t, dt0 = c.dt, final_time = c.final_time;
  const unsigned long long error = case_error(c);
  const int step = step0 + (pending ? 1 : 0);  // the step about to begin
  const bool mine = live && (lane >> 5) == (step & 1);
  const double v = mine ? __longlong_as_double((long long)raw) : 1.0e10;
  if (mine) *word = (unsigned long long)__double_as_longlong(1.0e10);
  const double dtm = wave_min(v);
  if (lane != 0 || !live) return;
  int st = step0;
  double t = t0;
  if (pending) {
    t = t0 + dt0;
    st = step0 + 1;
    c.t = t;
    c.step = st;
    c.pending = 0;
    c.last_count = cnt2;
  }
  const bool go = st < final_step && t < final_time && error == ~0ull;
  c.active = go ? 1 : 0;
  if (go) {
    double dt = dtm;
    if (t + dt > final_time) dt = final_time - t;
    c.dt = dt;
    c.pending = 1;
    c.count[0] = c.count[1] = c.count[2] = 0;
  }
}

__global__ __launch_bounds__(64) void k_step_begin(Ctl ctl, int ncases) {
  if ((int)blockIdx.x < ncases) step_control(ctl, blockIdx.x, threadIdx.x);
}

__global__ void k_reset_count(Ctl ctl, int slot) { ctl.count[slot] = 0; }

// writes a whole control record from a kernel argument (no host-to-device copy)
__global__ void k_set_case(CaseCtl* dst, CaseCtl value) { *dst = value; }

// ---------------------------------------------------------------------------------------
// limiter
// ---------------------------------------------------------------------------------------

// modified_kxrcf, src/spatial_disc/limiter/weno.rs:7-85. The neighbour trace uses the target's own
// local edge index with reversed point order (:21-27, :63-64); faces without a neighbour are skipped;
// no inflow face gives 0/0 = NaN, which is not >= 1. The detection arithmetic runs in the bulk type
// (FP64 in FP64 mode; FP32 in FP32-bulk mode, where only the WENO indicators and weights and the
// positivity tests stay in FP64); the mode sums for the next time step are FP64 in both.
// after_loads() runs once every load of the element has been issued (the caller's own reads go there,
// so that they overlap the element's instead of preceding or following them).
template <class T, int P, int NQ, int NE, class AfterLoads>
__device__ __forceinline__ bool kxrcf(const DevMesh& m, const T* __restrict__ U, const double* __restrict__ lt,
                                      const T* __restrict__ tphie, int e, double& sum0, double& sum1, double& sum2,
                                      double& sum3, AfterLoads after_loads) {
  // multiply-adds are fused within one expression only, so every kernel that inlines the detector (the
  // detector pass, the per-level sweep, the walk) evaluates the same operations in the same order
#pragma clang fp contract(on)
  using LO = LimOff<P, NQ, NE>;
  constexpr int NP = LO::NP;
  const int nes = m.nes;
  // modes held in the bulk type (half the registers in FP32-bulk), arithmetic in FP64
  T rho[NP], ru[NP], rv[NP], en[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const unsigned o = (unsigned)(k * nes + e);
    rho[k] = ld32(U, o);
    ru[k] = ld32(U + (size_t)NP * nes, o);
    rv[k] = ld32(U + (size_t)2 * NP * nes, o);
    en[k] = ld32(U + (size_t)3 * NP * nes, o);
  }
  {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      a0 += (double)rho[k];
      a1 += (double)ru[k];
      a2 += (double)rv[k];
      a3 += (double)en[k];
    }
    sum0 = a0;
    sum1 = a1;
    sum2 = a2;
    sum3 = a3;
  }
  using A = T;  // detection arithmetic in the bulk type
  A djump = 0, ejump = 0, surface = 0, minr = A(1.0e10), mine = A(1.0e10);
  // edge basis values [3][NE][NP] in the bulk type (the same values as the FP64 table, rounded once on the
  // host in FP32-bulk mode, so the detector converts nothing)
  const T* phie = tphie;
  const double* ew = lt + LO::EW;
  // Connectivity and face geometry of all three faces, then (FP32) the neighbours' density and energy
  // modes of all three faces, each group issued together rather than face by face. A face without a
  // neighbour reads the element itself, unused.
  int nbr[3], eds[3];
#pragma unroll
  for (int lic = 0; lic < 3; ++lic) {
    nbr[lic] = ld32(m.el_nbr, (unsigned)(lic * nes + e));
    eds[lic] = ld32(m.el_edge, (unsigned)(lic * nes + e));
  }
  const A h15 = (A)ld32(m.el_geo + (size_t)G_CR15 * nes, (unsigned)e);
  A fnx[3], fny[3], flen[3];
#pragma unroll
  for (int lic = 0; lic < 3; ++lic) {
    const int edge = eds[lic] >> 1;
    const A sgn = (eds[lic] & 1) ? A(-1) : A(1);
    fnx[lic] = sgn * (A)ld32(m.ed_geo, (unsigned)edge);
    fny[lic] = sgn * (A)ld32(m.ed_geo + m.ed_stride, (unsigned)edge);
    flen[lic] = A(2) * (A)ld32(m.ed_geo + 2 * (size_t)m.ed_stride, (unsigned)edge);
  }
  constexpr bool PRELOAD = sizeof(T) == 4;
  T nrho[PRELOAD ? 3 : 1][NP], nen[PRELOAD ? 3 : 1][NP];
  if constexpr (PRELOAD) {
#pragma unroll
    for (int lic = 0; lic < 3; ++lic) {
      const int nb = nbr[lic] >= 0 ? nbr[lic] : e;
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        nrho[lic][k] = ld32(U, (unsigned)(k * nes + nb));
        nen[lic][k] = ld32(U + (size_t)3 * NP * nes, (unsigned)(k * nes + nb));
      }
    }
  }
  after_loads();
  // FP32: every face is evaluated and its terms are selected in (the same operations in the same
  // order as skipping it), so no branch separates the loads above from their uses. FP64 keeps the
  // branches, which keep the neighbour loads (and their registers) to the inflow faces.
#pragma unroll
  for (int lic = 0; lic < 3; ++lic) {
    const int nb = nbr[lic];
    if constexpr (!PRELOAD) {
      if (nb < 0) continue;
    }
    const bool has = nb >= 0;
    const A nx = fnx[lic], ny = fny[lic], len = flen[lic];
    A tr[NE], te[NE];
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      A a = 0, b = 0;
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const A ph = (A)phie[(lic * NE + q) * NP + k];
        a += (A)rho[k] * ph;
        b += (A)en[k] * ph;
      }
      tr[q] = a;
      te[q] = b;
      minr = (has && a < minr) ? a : minr;
      mine = (has && b < mine) ? b : mine;
    }
    A mu = 0, mv = 0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const A ph = (A)phie[(lic * NE + NE / 2) * NP + k];
      mu += (A)ru[k] * ph;
      mv += (A)rv[k] * ph;
    }
    const bool inflow = has && mu * nx + mv * ny < A(0);
    if constexpr (PRELOAD) {
      // jump = (len / 2) sum_q w_q (own_q - nb_q): the own side from the traces above, the neighbour's
      // (its modes on this face's basis in reversed point order) as one dot product with KXW
      const T* kxw = tphie + (TabOff<NP, NQ, NE>::KXW - TabOff<NP, NQ, NE>::PHIE);
      A ro = 0, eo = 0, rn = 0, en2 = 0;
#pragma unroll
      for (int q = 0; q < NE; ++q) {
        ro += (A)ew[q] * tr[q];
        eo += (A)ew[q] * te[q];
      }
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const A w = kxw[lic * NP + k];
        rn += (A)nrho[lic][k] * w;
        en2 += (A)nen[lic][k] * w;
      }
      const A rj = A(0.5) * len * (ro - rn), ej = A(0.5) * len * (eo - en2);
      djump = inflow ? djump + rj : djump;
      ejump = inflow ? ejump + ej : ejump;
      surface = inflow ? surface + len : surface;
    } else if (inflow) {
      A rj = 0, ej = 0;
#pragma unroll
      for (int q = 0; q < NE; ++q) {
        A xr = 0, xe = 0;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
          const A ph = (A)phie[(lic * NE + (NE - 1 - q)) * NP + k];
          xr += (A)ld32(U, (unsigned)(k * nes + nb)) * ph;
          xe += (A)ld32(U + (size_t)3 * NP * nes, (unsigned)(k * nes + nb)) * ph;
        }
        rj += A(0.5) * len * (A)ew[q] * (tr[q] - xr);
        ej += A(0.5) * len * (A)ew[q] * (te[q] - xe);
      }
      djump += rj;
      ejump += ej;
      surface += len;
    }
  }
  const A id = fabs(djump) / (h15 * surface * minr);
  const A ie = fabs(ejump) / (h15 * surface * mine);
  return (id >= A(1)) || (ie >= A(1));
}

// KXRCF detector, one lane per element. With fuse_dt (the last stage of a step) the next step's time
// step is reduced here for the cells that stay untouched; k_weno adds the rebuilt ones.
// (one wavefront = one 64-element block = one case: the exits below are uniform per wavefront)
template <class T, int P, int NQ, int NE>
__global__ __launch_bounds__(256) void k_kxrcf(DevMesh m, const T* __restrict__ U, const double* __restrict__ lt,
                                               const T* __restrict__ tphie, LimiterBufs lb, Ctl ctl, int slot,
                                               int force, int fuse_dt, int remap) {
  const int blk = remap ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int e = blk * blockDim.x + threadIdx.x;
  if ((e >> 6) >= (m.nes >> 6)) return;
  const int cs = case_of(m, m.el_case64, e);
  CaseCtl* cc = ctl.cases + (cs < 0 ? 0 : cs);
  // The case record is read once the element's own loads have gone out (in after_loads) and consulted
  // after them: every lane below the capacity loads in bounds (the arrays are capacity-sized, and the
  // connectivity of unused capacity is -1 or a former case's in-range ids), and the results of a lane
  // outside a running case's range are masked, not branched around. (Scalar loads return out of order,
  // so a wait for any one waits for all: nothing before the element's loads branches on a kernel
  // argument or waits for the record.)
  const double hmin = ld32(m.el_geo + (size_t)G_HMIN * m.nes, (unsigned)e);
  double s0, s1, s2, s3;
  int c_active, e_end, c_step;
  unsigned long long c_error;
  double cfl;
  const bool det = kxrcf<T, P, NQ, NE>(m, U, lt, tphie, e, s0, s1, s2, s3, [&]() {
    c_active = cc->active;
    e_end = cc->e_end;
    c_step = cc->step;
    c_error = case_error(*cc);
    cfl = cc->cfl;
  });
  const bool runs = (cs >= 0) & ((force != 0) | ((c_active != 0) & (c_error == ~0ull)));  // case_runs()
  const bool valid = runs & (e < m.ne) & (e < e_end);
  const bool t = det & valid;
  if (valid && lb.flags) lb.flags[e] = t ? 1 : 0;
  const int i = wave_append(&ctl.count[slot], t, &cc->count[slot]);
  if (i >= 0) lb.list[i] = e;
  if (valid && lb.pend) lb.pend[e] = i;
  if (fuse_dt && runs) {
    double dt = 1.0e10;
    if (valid && !t) {
      const double d = element_dt<1>([&](int v, int) { return v == 0 ? s0 : (v == 1 ? s1 : (v == 2 ? s2 : s3)); },
                                     m.hcr, cfl, hmin);
      if (d < dt) dt = d;
    }
    wave_min_dt(dt, dt_slot(ctl, cs, c_step + 1), e >> 6);
  }
}

// Smoothness matrix of one element: beta(P) = sum_{l=1..p} J^(l-1) sum_{l1=0..l} sum_q (J/2) w_q
// (d^l P / dx^l1 dy^(l-l1))(x_q)^2 (src/spatial_disc/limiter/weno.rs:125-143) is the quadratic form
// P^T B P with B[k][k'] = sum_l J^(l-1) sum_l1 sum_q (J/2) w_q D_{l1}[q][k] D_{l1}[q][k'], where D are the
// physical derivative tables of src/mesh.rs:101-112 (binomial expansion of
// (xx d_xi + ex d_eta)^l1 (xy d_xi + ey d_eta)^l2 for any order). B depends on geometry only, so it is
// built once per element at set-up (one 64-lane block per element) and stored packed (upper triangle).
template <int P, int NQ, int NE>
struct SmoothShared {
  static constexpr int NP = np_of(P);
  static constexpr int NKW = P * (P + 3) / 2;
  double D[NKW][NQ][NP];
};

template <int P, int NQ, int NE>
__global__ __launch_bounds__(64) void k_smoothness(DevMesh m, const double* __restrict__ lt, double* __restrict__ smat,
                                                  int e0) {
  using LO = LimOff<P, NQ, NE>;
  constexpr int NP = LO::NP;
  constexpr int NKW = LO::NKW;
  constexpr int NPK = NP * (NP + 1) / 2;
  __shared__ SmoothShared<P, NQ, NE> sh;
  const int e = e0 + (int)blockIdx.x;
  if (e >= m.nes) return;
  const int lane = threadIdx.x;
  const int nes = m.nes;
  const double J = m.el_geo[G_J * nes + e];
  const double xx = m.el_geo[G_XX * nes + e], xy = m.el_geo[G_XY * nes + e];
  const double ex = m.el_geo[G_EX * nes + e], ey = m.el_geo[G_EY * nes + e];
  const double* dref = lt + LO::DREF;
  for (int idx = lane; idx < NKW * NQ * NP; idx += 64) {
    const int kw = idx / (NQ * NP), q = (idx / NP) % NQ, k = idx % NP;
    int l = 1;
    while ((l + 1) * (l + 2) / 2 - 1 <= kw) ++l;
    const int l1 = kw - (l * (l + 1) / 2 - 1), l2 = l - l1;
    double s = 0.0;
    double cx = 1.0;  // C(l1, i)
    for (int i = 0; i <= l1; ++i) {
      double cy = 1.0;  // C(l2, jj)
      for (int jj = 0; jj <= l2; ++jj) {
        double c = cx * cy;
        for (int t = 0; t < i; ++t) c *= xx;
        for (int t = 0; t < l1 - i; ++t) c *= ex;
        for (int t = 0; t < jj; ++t) c *= xy;
        for (int t = 0; t < l2 - jj; ++t) c *= ey;
        const int a = i + jj, bb = l - i - jj;
        s += c * dref[(dref_key(a, bb) * NQ + q) * NP + k];
        cy = cy * (l2 - jj) / (jj + 1);
      }
      cx = cx * (l1 - i) / (i + 1);
    }
    sh.D[kw][q][k] = s;
  }
  __syncthreads();
  const double* cw = lt + LO::CW;
  for (int idx = lane; idx < NPK; idx += 64) {
    int k = 0, rem = idx;
    while (rem >= NP - k) {
      rem -= NP - k;
      ++k;
    }
    const int kk = k + rem;
    double s = 0.0;
    for (int kw = 0; kw < NKW; ++kw) {
      int l = 1;
      while ((l + 1) * (l + 2) / 2 - 1 <= kw) ++l;
      double jl = 1.0;
      for (int t = 0; t < l - 1; ++t) jl *= J;
      double a = 0.0;
      for (int q = 0; q < NQ; ++q) a += 0.5 * J * cw[q] * sh.D[kw][q][k] * sh.D[kw][q][kk];
      s += jl * a;
    }
    smat[(size_t)e * NPK + idx] = s;
  }
}

// WENO rebuild of one troubled element by one 64-lane wavefront (block of 64), FP64 throughout.
// Lane (face i, polynomial j, variable v) = 16 i + 4 j + v projects polynomial j (0: the target,
// 1..3: the neighbour across face j-1 with its mean replaced by the target's, weno.rs:148-155) on
// the characteristic row v of face i (eigenmatrix at the target mean and the raw edge normal,
// weno.rs:121-124), forms beta = P^T B P with the element's smoothness matrix, and the nonlinear
// weight linear_w / (beta + eps)^2 (weno.rs:142,173).
template <int P, int NQ, int NE>
struct WenoShared {
  static constexpr int NP = np_of(P);
  static constexpr int NPK = NP * (NP + 1) / 2;
  double src[4][4][NP];
  double B[NPK];
  double Pj[3][4][4][NP];
  double nlw[3][4][4];
  double Rrow[3][4][4];
  double PN[3][4][NP];
  double RES[3][4][NP];
  double area[3];
  double fin[4 * NP];  // the rebuilt polynomial, for the time step of the fused reduction
  int flag;
};

// after_loads() runs once the element's loads are issued, before the first barrier (the caller's own
// reads go there, so that they overlap the rebuild's instead of preceding them).
template <class T, int P, int NQ, int NE, class AfterLoads>
__device__ int weno_wave(const DevMesh& m, const T* __restrict__ U, const double* __restrict__ smat, double hcr, int e,
                         WenoShared<P, NQ, NE>& sh, double* __restrict__ out /* [4][NP] */, AfterLoads after_loads) {
  // fused within one expression only (see kxrcf): the walk's rebuilds equal the pass's and the levels'
#pragma clang fp contract(on)
  constexpr int NP = np_of(P);
  constexpr int NPK = NP * (NP + 1) / 2;
  const int lane = threadIdx.x;
  const int nes = m.nes;
  // Two memory trips. First what depends on the element only: its neighbour and edge ids (the same for
  // every lane: e is uniform over the wavefront), its smoothness matrix. Then, all together, what depends
  // on those ids: the candidates' modes, the neighbours' areas and the face normals. Every load reads a
  // valid index (the element itself where a face has no neighbour) and the unused values are masked
  // afterwards, so no branch holds a load back.
  const int nbr[3] = {m.el_nbr[e], m.el_nbr[nes + e], m.el_nbr[2 * nes + e]};
  const int edr[3] = {m.el_edge[e] >> 1, m.el_edge[nes + e] >> 1, m.el_edge[2 * nes + e] >> 1};
  constexpr int NBQ = (NPK + 63) / 64;
  double bq[NBQ];
#pragma unroll
  for (int it = 0; it < NBQ; ++it) {
    const int idx = lane + 64 * it;
    bq[it] = smat[(size_t)e * NPK + (idx < NPK ? idx : 0)];
  }
  // candidate j, variable v, mode k at idx = (j * 4 + v) * NP + k, NIT per lane
  constexpr int NSRC = 16 * NP, NIT = (NSRC + 63) / 64;
  double sv[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int idx0 = lane + 64 * it, idx = idx0 < NSRC ? idx0 : 0;
    const int j = idx / (4 * NP), v = (idx / NP) % 4, k = idx % NP;
    const int nj = j > 0 ? nbr[j - 1] : e;
    // mode 0 of every candidate is the element's own (weno.rs:148-155)
    const int src = (j == 0 || k == 0 || nj < 0) ? e : nj;
    const double val = (double)U[(size_t)(v * NP + k) * nes + src];
    sv[it] = (j > 0 && nj < 0) ? 0.0 : val;  // no neighbour across the face: no candidate
  }
  const int nbi = lane < 3 ? nbr[lane] : -1;
  const double ar = m.el_geo[G_J * nes + (nbi >= 0 ? nbi : e)];
  // face normal of the characteristic lanes (lane 16 i + ...: face i), used only where face i has a
  // neighbour
  const int fe = edr[lane < 48 ? lane >> 4 : 0];
  const double fnx = m.ed_geo[fe], fny = m.ed_geo[m.ed_stride + fe];
  after_loads();
  if (lane < 3) sh.area[lane] = nbi >= 0 ? ar * 0.5 : 0.0;
#pragma unroll
  for (int it = 0; it < NBQ; ++it)
    if (lane + 64 * it < NPK) sh.B[lane + 64 * it] = bq[it];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int idx = lane + 64 * it;
    if (idx < NSRC) sh.src[idx / (4 * NP)][(idx / NP) % 4][idx % NP] = sv[it];
  }
  // linear weights, weno.rs:93-110 (a boundary face i zeroes weights[i], then weights[count] is set)
  const bool b0 = nbr[0] < 0, b1 = nbr[1] < 0, b2 = nbr[2] < 0;
  const int nbnd = (int)b0 + (int)b1 + (int)b2;
  double lw0 = b0 ? 0.0 : 0.997, lw1 = b1 ? 0.0 : 0.001, lw2 = b2 ? 0.0 : 0.001, lw3 = 0.001;
  if (nbnd == 0) lw0 = 0.997;
  if (nbnd == 1) lw1 = 0.998;
  if (nbnd == 2) lw2 = 0.999;
  __syncthreads();
  if (nbnd > 2) return 5;  // the reference panics: "Invalid number of boundary neighbours!"
  const double EPS = 2.220446049250313e-16;
  if (lane < 48) {
    const int i = lane >> 4, j = (lane >> 2) & 3, v = lane & 3;
    double Pk[NP];
    double w = 0.0;
#pragma unroll
    for (int k = 0; k < NP; ++k) Pk[k] = 0.0;
    if (nbr[i] >= 0) {
      const double nx = fnx, ny = fny;
      const double mean[4] = {sh.src[0][0][0], sh.src[0][1][0], sh.src[0][2][0], sh.src[0][3][0]};
      double Lm[16], Rm[16];
      eigenmatrix(mean, nx, ny, hcr, Lm, Rm);
      double lr[4], rr[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        lr[c] = Lm[c];
        rr[c] = Rm[c];
      }
#pragma unroll
      for (int r = 1; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          lr[c] = v == r ? Lm[4 * r + c] : lr[c];
          rr[c] = v == r ? Rm[4 * r + c] : rr[c];
        }
      if (j == 0) {
#pragma unroll
        for (int c = 0; c < 4; ++c) sh.Rrow[i][v][c] = rr[c];
      }
      if (j == 0 || nbr[j - 1] >= 0) {
#pragma unroll
        for (int k = 0; k < NP; ++k)
          Pk[k] = lr[0] * sh.src[j][0][k] + lr[1] * sh.src[j][1][k] + lr[2] * sh.src[j][2][k] + lr[3] * sh.src[j][3][k];
        double beta = 0.0;
        int o = 0;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
          double t = sh.B[o] * Pk[k];
#pragma unroll
          for (int kk = k + 1; kk < NP; ++kk) t += 2.0 * sh.B[o + kk - k] * Pk[kk];
          o += NP - k;
          beta += Pk[k] * t;
        }
        const double lwj = j == 0 ? lw0 : (j == 1 ? lw1 : (j == 2 ? lw2 : lw3));
        w = lwj / ((beta + EPS) * (beta + EPS));
      }
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) sh.Pj[i][j][v][k] = Pk[k];
    sh.nlw[i][j][v] = w;
  }
  __syncthreads();
  // the two combinations below run over (face, variable, mode) items, one LDS round trip per item
  for (int it = lane; it < 12 * NP; it += 64) {
    const int i = it / (4 * NP), v = (it / NP) % 4, k = it % NP;
    if (nbr[i] >= 0) {
      const double n0 = sh.nlw[i][0][v], n1 = sh.nlw[i][1][v], n2 = sh.nlw[i][2][v], n3 = sh.nlw[i][3][v];
      const double p0 = sh.Pj[i][0][v][k], p1 = sh.Pj[i][1][v][k], p2 = sh.Pj[i][2][v][k], p3 = sh.Pj[i][3][v][k];
      const double den = ((n0 + n1) + n2) + n3;
      const double w0 = n0 / den, w1 = n1 / den, w2 = n2 / den, w3 = n3 / den;
      sh.PN[i][v][k] = w0 * p0 + w1 * p1 + w2 * p2 + w3 * p3;
    }
  }
  __syncthreads();
  for (int it = lane; it < 12 * NP; it += 64) {
    const int i = it / (4 * NP), v = (it / NP) % 4, k = it % NP;
    if (nbr[i] >= 0) {
      const double r0 = sh.Rrow[i][v][0], r1 = sh.Rrow[i][v][1], r2 = sh.Rrow[i][v][2], r3 = sh.Rrow[i][v][3];
      const double q0 = sh.PN[i][0][k], q1 = sh.PN[i][1][k], q2 = sh.PN[i][2][k], q3 = sh.PN[i][3][k];
      sh.RES[i][v][k] = r0 * q0 + r1 * q1 + r2 * q2 + r3 * q3;
    }
  }
  __syncthreads();
  // area-weighted mean over faces, weno.rs:111-118,196,260-263
  for (int idx = lane; idx < 4 * NP; idx += 64) {
    const int v = idx / NP, k = idx % NP;
    double rec = 0.0, area = 0.0;
    for (int i = 0; i < 3; ++i)
      if (nbr[i] >= 0) {
        rec += sh.RES[i][v][k] * sh.area[i];
        area += sh.area[i];
      }
    out[idx] = rec / area;
    sh.fin[idx] = rec / area;
  }
  return 0;
}


// The troubled list holds elements of running cases only (the detector skips the others). With defer
// (the pass ahead of the element-order walk) a failed rebuild is left in pend[e] as -2 - status for
// the walk to report or drop, instead of stopping the case here. The only failure of weno_wave is a
// cell with three boundary faces (status 5), and the detector never marks such a cell: with no
// neighbour face the inflow surface and both jumps are 0 and the indicators 0/0 = NaN, which is not
// >= 1. So neither this failure path nor the walk's handling of it (flag 2, erb, the -2 - p decoding)
// can run from a detector pass; they are kept so that a rebuild failure, should one become possible,
// still stops the case with its status instead of writing an unfinished polynomial.
template <class T, int P, int NQ, int NE>
__global__ __launch_bounds__(64) void k_weno(DevMesh m, const T* __restrict__ U, const double* __restrict__ smat,
                                             LimiterBufs lb, Ctl ctl, int slot, int fuse_dt, int defer) {
  __shared__ WenoShared<P, NQ, NE> sh;
  constexpr int NP = np_of(P);
  // the block's first list entry is read in the same trip as the count (at a clamped index: the list
  // holds nes entries, and an entry at or past the count is stale and left unused)
  int s = blockIdx.x;
  const int n = ctl.count[slot];
  int e = lb.list[s < m.nes ? s : m.nes - 1];
  asm volatile("" : "+s"(e));  // keeps the load ahead of the loop test (it would be sunk below it)
  for (; s < n; s += gridDim.x) {
    if (s != (int)blockIdx.x) e = lb.list[s];
    double* out = lb.w + (size_t)s * 4 * NP;
    // the case and what the fused time step needs are read once the rebuild's own loads are out
    int cs = 0, step = 0;
    double hmin = 0.0, cfl = 0.0;
    const int st = weno_wave<T, P, NQ, NE>(m, U, smat, m.hcr, e, sh, out, [&]() {
      cs = case_of(m, m.el_case64, e);
      hmin = m.el_geo[G_HMIN * m.nes + e];
      cfl = ctl.cases[cs].cfl;
      step = ctl.cases[cs].step;
    });
    CaseCtl* cc = ctl.cases + cs;
    if (st && threadIdx.x == 0) {
      if (defer) {
        lb.pend[e] = -2 - st;
        lb.flags[e] = 2;
      } else
        atomicMin(&cc->error, (0xFFFFFFull << 8) | (unsigned long long)st);
    }
    __syncthreads();
    if (fuse_dt && !st && threadIdx.x == 0) {
      const double d = element_dt<NP>([&](int v, int k) { return sh.fin[v * NP + k]; }, m.hcr, cfl, hmin);
      if (d < 1.0e10)
        atomicMin(dt_slot(ctl, cs, step + 1) + (e % DT_LANES) * DT_STRIDE, (unsigned long long)__double_as_longlong(d));
    }
    __syncthreads();
  }
}

template <class T, int NP, int STAGE>
__global__ __launch_bounds__(256) void k_apply(LimiterBufs lb, StageBufs<T> b, int nes, Ctl ctl, int slot) {
  const int n = ctl.count[slot];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n * 4 * NP; t += gridDim.x * blockDim.x) {
    const int s = t / (4 * NP), idx = t % (4 * NP);
    store_limited<T, NP, STAGE>(b, nes, lb.list[s], idx, lb.w[(size_t)s * 4 * NP + idx]);
    if (idx == 0 && lb.pend) lb.pend[lb.list[s]] = -1;
  }
}

// writes back the rebuilt elements still pending after a last-stage limiter pass (u^n and its FP32
// shadow) and clears their marks
template <class T, int NP>
__global__ void k_flush(int* __restrict__ pend, const double* __restrict__ w, double* __restrict__ un,
                        float* __restrict__ sn, int ne, int nes) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int sl = pend[e];
  if (sl < 0) return;
  for (int c = 0; c < 4 * NP; ++c) {
    const double v = w[(size_t)sl * 4 * NP + c];
    un[(size_t)c * nes + e] = v;
    if (sn) sn[(size_t)c * nes + e] = (float)v;
  }
  pend[e] = -1;
}

// element-order walk: largest mesh whose bit masks fit in LDS, and record chunks staged per trip
constexpr int WALK_MAX_E = 1 << 18;
constexpr int WALK_STAGE = 8;

// exact element-order sweep: one wavefront per element of one level; elements of a level are not
// adjacent, so each reads final values of lower-numbered neighbours and untouched higher ones.
template <class T, int P, int NQ, int NE, int STAGE>
__global__ __launch_bounds__(64) void k_limit_level(DevMesh m, const int* __restrict__ elems, int n, StageBufs<T> b,
                                                    const T* __restrict__ U, const double* __restrict__ lt,
                                                    const T* __restrict__ tphie, const double* __restrict__ smat,
                                                    LimiterBufs lb, Ctl ctl, int slot, int force) {
  __shared__ WenoShared<P, NQ, NE> sh;
  __shared__ double outv[4 * np_of(P)];
  constexpr int NP = np_of(P);
  const int s = blockIdx.x;
  if (s >= n) return;
  const int e = elems[s];
  const int cs = case_of(m, m.el_case64, e);
  if (cs < 0) return;
  CaseCtl* cc = ctl.cases + cs;
  if (!force && !case_runs(*cc)) return;
  if (threadIdx.x == 0) {
    double s0, s1, s2, s3;
    const bool t = kxrcf<T, P, NQ, NE>(m, U, lt, tphie, e, s0, s1, s2, s3, []() {});
    sh.flag = t ? 1 : 0;
    if (lb.flags) lb.flags[e] = t ? 1 : 0;
  }
  __syncthreads();
  if (!sh.flag) return;
  const int st = weno_wave<T, P, NQ, NE>(m, U, smat, m.hcr, e, sh, outv, []() {});
  __syncthreads();
  if (st) {
    if (threadIdx.x == 0)
      atomicMin(&cc->error, ((unsigned long long)(unsigned)(e - cc->e_off) << 8) | (unsigned long long)st);
    return;
  }
  for (int idx = threadIdx.x; idx < 4 * NP; idx += 64) store_limited<T, NP, STAGE>(b, m.nes, e, idx, outv[idx]);
  if (threadIdx.x == 0) {
    atomicAdd(&ctl.count[slot], 1);
    atomicAdd(&cc->count[slot], 1);
  }
}

// Element-order sweep in one launch (weno.rs:86-91: cells in index order, each seeing the final values
// of its lower-numbered neighbours and the untouched ones of its higher-numbered ones). The detector and
// the rebuild have first run on every cell from the stage state (k_kxrcf, then k_weno with defer: flags
// 1 = troubled and rebuilt, 2 = troubled and failed). A cell none of whose lower-numbered neighbours was
// rebuilt sees exactly that state, so its decision and its rebuild are those of the pass: the walk only
// marks it (k_walk_flush writes the rebuild back afterwards, where pend still holds its slot). A cell
// with a rebuilt lower-numbered neighbour is detected and rebuilt again from the current state, after
// writing back the pending rebuilds of those neighbours (a higher-numbered neighbour of a cell is never
// written before that cell's turn: it comes at a later level than the cell, and is only written back
// for cells of still later levels).
// One wavefront walks the level schedule: records {e, lower-numbered neighbours or -1} in level order,
// each level padded to whole 64-record chunks (e = -1), so a chunk holds cells of one level, none
// adjacent to another. Which cells are rebuilt, and what the pass found, are bit masks in LDS. Counts
// end as the number of cells written. A failure stops the case (the error names the lowest-numbered
// failing cell); the walk still finishes.
template <class T, int P, int NQ, int NE, int STAGE>
__global__ __launch_bounds__(64) void k_limit_walk(DevMesh m, const int4* __restrict__ recs, int nchunks,
                                                   StageBufs<T> b, const T* U, const double* __restrict__ lt,
                                                   const T* __restrict__ tphie, const double* __restrict__ smat,
                                                   LimiterBufs lb, Ctl ctl, int slot, int force) {
  constexpr int NP = np_of(P);
  __shared__ WenoShared<P, NQ, NE> sh;
  __shared__ double outv[4 * NP];
  __shared__ unsigned modb[WALK_MAX_E / 32], jtb[WALK_MAX_E / 32], erb[WALK_MAX_E / 32];
  __shared__ int4 rs[WALK_STAGE * 64];
  const int lane = threadIdx.x;
  CaseCtl* cc = ctl.cases + m.single_case;
  if (!force && !case_runs(*cc)) return;
  // the pass's outcome as bit masks: one 32-cell word per lane and trip
  const int ne = m.ne, nw = (ne + 31) >> 5;
  for (int w = lane; w < nw; w += 64) {
    unsigned ok = 0u, bad = 0u;
    if (32 * w + 32 <= ne) {
      const unsigned* f4 = reinterpret_cast<const unsigned*>(lb.flags + 32 * w);
      unsigned v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = f4[i];
#pragma unroll
      for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const unsigned f = (v[i] >> (8 * j)) & 0xffu;
          ok |= (f == 1u ? 1u : 0u) << (4 * i + j);
          bad |= (f == 2u ? 1u : 0u) << (4 * i + j);
        }
    } else {
      for (int e = 32 * w; e < ne; ++e) {
        const unsigned f = lb.flags[e];
        ok |= (f == 1u ? 1u : 0u) << (e & 31);
        bad |= (f == 2u ? 1u : 0u) << (e & 31);
      }
    }
    modb[w] = 0u;
    jtb[w] = ok;
    erb[w] = bad;
  }
  auto bit = [](const unsigned* mask, int j) { return (mask[j >> 5] >> (j & 31)) & 1u; };
  if (nchunks <= 0) return;
  // the next WALK_STAGE chunks' records are loaded into registers while the current ones are walked
  // (a chunk past the end reads the last one again, unused)
  typedef int v4i __attribute__((ext_vector_type(4)));  // a native vector stays in registers
  const v4i* rv = reinterpret_cast<const v4i*>(recs);
  v4i pf[WALK_STAGE];
#pragma unroll
  for (int i = 0; i < WALK_STAGE; ++i) pf[i] = rv[(size_t)(i < nchunks ? i : nchunks - 1) * 64 + lane];
  for (int c0 = 0; c0 < nchunks; c0 += WALK_STAGE) {
    const int nc = nchunks - c0 < WALK_STAGE ? nchunks - c0 : WALK_STAGE;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < WALK_STAGE; ++i) rs[i * 64 + lane] = make_int4(pf[i].x, pf[i].y, pf[i].z, pf[i].w);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < WALK_STAGE; ++i) {
      const int ch = c0 + WALK_STAGE + i < nchunks ? c0 + WALK_STAGE + i : nchunks - 1;
      pf[i] = rv[(size_t)ch * 64 + lane];
    }
    for (int k = 0; k < nc; ++k) {
      const int4 r = rs[k * 64 + lane];
      const int e = r.x;
      const bool live = e >= 0;
      const bool ly = r.y >= 0 && bit(modb, r.y), lz = r.z >= 0 && bit(modb, r.z), lw = r.w >= 0 && bit(modb, r.w);
      const bool dirty = live && (ly || lz || lw);
      const bool jok = live && bit(jtb, e), jbad = live && bit(erb, e);
      int delta = 0;  // cells written minus cells the pass counted
      if (!dirty && jok) atomicOr(&modb[e >> 5], 1u << (e & 31));
      if (!dirty && jbad) {
        const int p = lb.pend[e];
        atomicMin(&cc->error, ((unsigned long long)(unsigned)(e - cc->e_off) << 8) | (unsigned long long)(-2 - p));
        lb.flags[e] = 1;
        lb.pend[e] = -1;
        delta = -1;
      }
      const unsigned long long dm = __ballot(dirty);
      if (dm) {
        // pending rebuilds of the rebuilt lower-numbered neighbours go in place first
        const int nb[3] = {ly ? r.y : -1, lz ? r.z : -1, lw ? r.w : -1};
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          const int j = dirty ? nb[f] : -1;
          const int p = j >= 0 ? lb.pend[j] : -1;
          if (p >= 0) {
            store_limited_row<T, NP, STAGE>(b, m.nes, j, lb.w + (size_t)p * 4 * NP);
            lb.pend[j] = -1;
          }
        }
        // those and the rebuilds written at earlier levels are read below
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const int ef = __shfl(e, __builtin_ctzll(dm));
        double s0, s1, s2, s3;
        const bool t = kxrcf<T, P, NQ, NE>(m, U, lt, tphie, dirty ? e : ef, s0, s1, s2, s3, []() {});
        const bool td = dirty && t;
        if (dirty) lb.flags[e] = td ? 1 : 0;
        unsigned long long tm = __ballot(td), ok = 0ull;
        while (tm) {
          const int l = __builtin_ctzll(tm);
          tm &= tm - 1;
          const int ed = __shfl(e, l);
          // FP32-bulk stages 0 and 1 store d = w - u^n: u^n is read ahead of the rebuild (4 NP <= 64 lanes)
          constexpr bool DN = sizeof(T) == 4 && STAGE < 2;
          static_assert(4 * NP <= 64, "one mode per lane");
          const size_t ix = (size_t)(lane < 4 * NP ? lane : 0) * m.nes + ed;
          const double un = DN ? b.un[ix] : 0.0;
          const int st = weno_wave<T, P, NQ, NE>(m, U, smat, m.hcr, ed, sh, outv, []() {});
          __syncthreads();
          if (!st) {
            if constexpr (DN) {
              if (lane < 4 * NP) {
                (STAGE == 0 ? b.s : b.s2)[ix] = (T)outv[lane];
                b.d[ix] = (T)(outv[lane] - un);
              }
            } else {
              for (int idx = lane; idx < 4 * NP; idx += 64) store_limited<T, NP, STAGE>(b, m.nes, ed, idx, outv[idx]);
            }
            if (lane == 0) atomicOr(&modb[ed >> 5], 1u << (ed & 31));
            ok |= 1ull << l;
          } else if (lane == 0) {
            atomicMin(&cc->error, ((unsigned long long)(unsigned)(ed - cc->e_off) << 8) | (unsigned long long)st);
          }
          __syncthreads();
        }
        if (dirty) {
          delta = (int)((ok >> lane) & 1ull) - ((jok || jbad) ? 1 : 0);
          if (jok || jbad) lb.pend[e] = -1;
        }
      }
      if (delta) {
        atomicAdd(&ctl.count[slot], delta);
        atomicAdd(&cc->count[slot], delta);
      }
      // the next chunk's mask reads stay after this chunk's updates (one wavefront: LDS is in order)
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
}

// writes back the rebuilds the walk left pending (cells whose decision was the detector pass's)
template <class T, int NP, int STAGE>
__global__ __launch_bounds__(256) void k_walk_flush(int* __restrict__ pend, const double* __restrict__ w, StageBufs<T> b,
                                                    int ne, int nes) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int sl = pend[e];
  if (sl < 0) return;
  store_limited_row<T, NP, STAGE>(b, nes, e, w + (size_t)sl * 4 * NP);
  pend[e] = -1;
}

// ---------------------------------------------------------------------------------------
// layout conversion and checks
// ---------------------------------------------------------------------------------------
// host layout [ne][4][NP] (double) -> device [4 NP][nes]
// host layout [n][4][NP] (double) -> device [4 NP][nes] at elements e0 .. e0 + n - 1
template <class T>
__global__ void k_to_soa(const double* __restrict__ in, T* __restrict__ out, double* __restrict__ out64, int e0, int n,
                         int nes, int ncomp) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (e >= n) return;
  const double v = in[(size_t)e * ncomp + c];
  if (out) out[(size_t)c * nes + e0 + e] = (T)v;
  if (out64) out64[(size_t)c * nes + e0 + e] = v;
}
__global__ void k_from_soa(const double* __restrict__ in, double* __restrict__ out, int e0, int n, int nes, int ncomp) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (e >= n) return;
  out[(size_t)e * ncomp + c] = in[(size_t)c * nes + e0 + e];
}
__global__ void k_check_finite(const double* __restrict__ u, int e0, int n, int nes, int ncomp, int* flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int e = e0 + i;
  for (int c = 0; c < ncomp; ++c)
    if (!isfinite(u[(size_t)c * nes + e])) {
      atomicOr(flag, 1);
      return;
    }
}

// wall pressure force (Fx, Fy) per case from the mode coefficients: sum over wall edges and edge
// points of p * n * w * |e| / 2, with the normal oriented out of the fluid element; force[2 case + i].
// Edges ed0 .. ed0 + n - 1 with ed0 a multiple of 64 (one wavefront = one case).
template <class T, int P, int NQ, int NE>
__global__ void k_wall_force(DevMesh m, const double* __restrict__ U, const double* __restrict__ lt, Ctl ctl, int ed0,
                             int n, double* __restrict__ force) {
  using LO = LimOff<P, NQ, NE>;
  constexpr int NP = LO::NP;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int ed = ed0 + i;
  const int w0 = ed0 + (int)(blockIdx.x * blockDim.x) + (int)(threadIdx.x & ~63u);  // first edge of the wavefront
  if (w0 >= m.ed_stride) return;
  const int cs = case_of(m, m.ed_case64, w0);
  if (cs < 0) return;
  const double hcr = m.hcr;
  double fx = 0.0, fy = 0.0;
  if (i < n && ed < ctl.cases[cs].ed_end) {
    const int info = m.ed_info[ed];
    if (((info >> 4) & 3) == BC_WALL) {
      const int L = m.ed_elem[ed], lic = info & 3;
      const double nx = m.ed_geo[ed], ny = m.ed_geo[m.ed_stride + ed], hl = m.ed_geo[2 * m.ed_stride + ed];
      for (int q = 0; q < NE; ++q) {
        double s[4] = {0, 0, 0, 0};
        for (int v = 0; v < 4; ++v)
          for (int k = 0; k < NP; ++k) s[v] += U[(size_t)(v * NP + k) * m.nes + L] * lt[LO::PHIE + (lic * NE + q) * NP + k];
        const double p = (hcr - 1.0) * (s[3] - 0.5 * (s[1] * s[1] + s[2] * s[2]) / s[0]);
        fx += p * nx * lt[LO::EW + q] * hl;
        fy += p * ny * lt[LO::EW + q] * hl;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    fx += __shfl_xor(fx, o, 64);
    fy += __shfl_xor(fy, o, 64);
  }
  if ((threadIdx.x & 63) == 0 && (fx != 0.0 || fy != 0.0)) {
    atomicAdd(&force[2 * cs], fx);
    atomicAdd(&force[2 * cs + 1], fy);
  }
}

inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// One launch of a stage part. With the arena's profiling events set, the launch records them at the
// dispatch's own start and end (hipExtLaunchKernel), so their elapsed time is the kernel duration
// without the launch gaps a stream event pair would include.
template <typename... KArgs, typename... Args>
void go(const DeviceState& s, void (*kernel)(KArgs...), dim3 grid, dim3 block, hipStream_t st, Args... args) {
  if (s.prof_ev[0])
    hipExtLaunchKernelGGL(kernel, grid, block, 0, st, s.prof_ev[0], s.prof_ev[1], 0, static_cast<KArgs>(args)...);
  else
    hipLaunchKernelGGL(kernel, grid, block, 0, st, static_cast<KArgs>(args)...);
}
// the same with lds bytes of dynamic shared memory per block
template <typename... KArgs, typename... Args>
void go_lds(const DeviceState& s, void (*kernel)(KArgs...), dim3 grid, dim3 block, size_t lds, hipStream_t st,
            Args... args) {
  if (s.prof_ev[0])
    hipExtLaunchKernelGGL(kernel, grid, block, lds, st, s.prof_ev[0], s.prof_ev[1], 0, static_cast<KArgs>(args)...);
  else
    hipLaunchKernelGGL(kernel, grid, block, lds, st, static_cast<KArgs>(args)...);
}

}  // namespace

void launch_step_begin(Ctl ctl, int ncases, hipStream_t st) {
  if (ncases > 0) k_step_begin<<<ncases, 64, 0, st>>>(ctl, ncases);
}

namespace {
__global__ void k_case_poll_init(double* cpar, double* clprev, int* conv, int slot, double mach, double aoa_deg,
                                 double rho, double p, double hcr) {
  const double a = sqrt(hcr * p / rho), V = mach * a;
  const double rad = aoa_deg * (M_PI / 180.0);
  cpar[4 * slot + 0] = cos(rad);
  cpar[4 * slot + 1] = sin(rad);
  cpar[4 * slot + 2] = 1.0 / (0.5 * rho * V * V);
  cpar[4 * slot + 3] = 0.0;
  clprev[slot] = 0.0;
  conv[slot] = 0;
}

__global__ void k_converge(CaseCtl* cases, const double* __restrict__ force, const double* __restrict__ cpar,
                           double* clprev, int* conv, double* cl, int n, double chord, double tol) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  CaseCtl& c = cases[i];
  if (!c.live) return;
  const double v = (force[2 * i + 1] * cpar[4 * i] - force[2 * i] * cpar[4 * i + 1]) * cpar[4 * i + 2] / chord;
  cl[i] = v;
  // a case that has not stepped yet (admitted since the last steps) is not tested
  const bool running = c.active && case_error(c) == ~0ull && (c.step > 0 || c.pending);
  if (running && !conv[i] && fabs(v - clprev[i]) < tol) {
    conv[i] = 1;
    c.final_step = c.step + (c.pending ? 1 : 0);
  }
  clprev[i] = v;
}
}  // namespace

void launch_case_poll_init(double* cpar, double* clprev, int* conv, int slot, double mach, double aoa_deg, double rho,
                           double p, double hcr, hipStream_t st) {
  k_case_poll_init<<<1, 1, 0, st>>>(cpar, clprev, conv, slot, mach, aoa_deg, rho, p, hcr);
}

void launch_converge(CaseCtl* cases, const double* force, const double* cpar, double* clprev, int* conv, double* cl,
                     int ncases, double chord, double tol, hipStream_t st) {
  if (ncases > 0) k_converge<<<(ncases + 255) / 256, 256, 0, st>>>(cases, force, cpar, clprev, conv, cl, ncases, chord, tol);
}

void launch_set_case(CaseCtl* dst, const CaseCtl& value, hipStream_t st) { k_set_case<<<1, 1, 0, st>>>(dst, value); }

// ---------------------------------------------------------------------------------------
// Launch sequences for one (precision, order, rule) combination
// ---------------------------------------------------------------------------------------
template <class T, int P, int NQ, int NE>
class SolverImpl final : public Solver {
 public:
  static constexpr int NP = np_of(P);
  using O = TabOff<NP, NQ, NE>;
  using LO = LimOff<P, NQ, NE>;

  int table_len() const override { return O::N; }
  int limiter_table_len() const override { return LO::N; }
  bool fp32() const override { return sizeof(T) == 4; }
  bool has_mfma() const override { return sizeof(T) == 4 && NP <= 16 && NQ <= 16 && 3 * NE <= 16; }

  StageBufs<T> bufs(const DeviceState& s) const {
    StageBufs<T> b;
    b.in = nullptr;
    b.un = s.un;
    b.us = sizeof(T) == 8 ? (T*)s.us : nullptr;
    b.sn = sizeof(T) == 4 ? (T*)s.sn : nullptr;
    b.s = sizeof(T) == 4 ? (T*)s.s : nullptr;
    b.s2 = sizeof(T) == 4 ? (T*)s.s2 : nullptr;
    b.d = sizeof(T) == 4 ? (T*)s.d : nullptr;
    b.rout = s.rout;
    // the Jacobi sweep of a march leaves its rebuilds pending for the next stage's kernels to read
    b.pend = s.limiter == DGM_LIMITER_JACOBI ? s.lim_pend : nullptr;
    b.pw = s.lim_w;
    return b;
  }
  // stage input: FP64 reads u^n then the stage state; FP32-bulk reads the float shadows
  const T* stage_input(const DeviceState& s, int stage) const {
    if (sizeof(T) == 8) return stage == 0 ? (const T*)s.un : (const T*)s.us;
    return stage == 0 ? (const T*)s.sn : (stage == 1 ? (const T*)s.s : (const T*)s.s2);
  }
  const T* limiter_input(const DeviceState& s, int stage) const {
    if (sizeof(T) == 8) return stage == 2 ? (const T*)s.un : (const T*)s.us;
    return stage == 2 ? (const T*)s.sn : (stage == 0 ? (const T*)s.s : (const T*)s.s2);
  }

  // s0: the stage-0 edge pass of a march (errors into error_s0); nctl > 0: it also runs the step control
  void surface(const DeviceState& s, const StageBufs<T>& b, hipStream_t st, int force, int s0 = 0, int nctl = 0) const {
    if (s.mesh.nedge <= 0) {
      if (nctl > 0) launch_step_begin(s.ctl, nctl, st);
      return;
    }
    const int bs = s.tune.surface_block, wpb = bs / 64;
    const int extra = nctl > 0 ? (nctl + wpb - 1) / wpb : 0;
    go(s, k_surface<T, NP, NQ, NE>, blocks_for(s.mesh.nedge, bs) + extra, bs, st, s.mesh, b.in, (T*)s.flux,
       (const T*)s.tab, s.ctl, force, s.tune.remap, b.pend, b.pw, s0, nctl);
  }
  // the fused edge-and-element pass replaces a march stage's edge and element passes
  static bool fused_on(const DeviceState& s) {
    return sizeof(T) == 4 && NP <= 16 && NQ <= 16 && 3 * NE <= 16 && s.fused && s.use_mfma &&
           s.tune.element_waves > 0 && s.mesh.grp_rec;
  }
  // march: the element pass of a stage, fused with its edge pass when fused_on()
  template <int OP>
  void element(const DeviceState& s, StageBufs<T> b, hipStream_t st, int force, bool march = false) const {
    if (s.mesh.ne <= 0) return;
    if constexpr (sizeof(T) == 4 && NP <= 16 && NQ <= 16 && 3 * NE <= 16 && OP != OP_RESID) {
      if (march && fused_on(s)) {
        const int w = s.tune.element_waves;
        go_lds(s, k_fused<NP, NQ, NE, OP>, blocks_for(s.mesh.ne, 16 * w), 64 * w, fused_lds_bytes<NP, NE>(w), st,
               s.mesh, b, (const float*)s.tab, s.ctl, force, s.tune.remap);
        return;
      }
    }
    if constexpr (sizeof(T) == 4 && NP <= 16 && NQ <= 16 && 3 * NE <= 16) {
      if (s.use_mfma && s.tune.element_waves > 0) {
        const int w = s.tune.element_waves;
        go(s, k_element_mfma<NP, NQ, NE, OP>, blocks_for(s.mesh.ne, 16 * w), 64 * w, st, s.mesh, b,
           (const float*)s.flux, (const float*)s.tab, s.ctl, force, s.tune.remap);
        return;
      }
    }
    const int bs = s.tune.element_block;
    go(s, k_element<T, NP, NQ, NE, OP>, blocks_for(s.mesh.ne, bs), bs, st, s.mesh, b, (const T*)s.flux,
       (const T*)s.tab, s.ctl, force, s.tune.remap);
  }

  void residual(const DeviceState& s, hipStream_t st) const override {
    StageBufs<T> b = bufs(s);
    b.in = stage_input(s, 0);
    b.pend = nullptr;  // a stand-alone residual reads the uploaded state as it is
    surface(s, b, st, 1);
    element<OP_RESID>(s, b, st, 1);
  }

  void limiter(const DeviceState& s, hipStream_t st) const override {
    // stand-alone SpatialDisc::apply_limiter on the resident u^n
    k_reset_count<<<1, 1, 0, st>>>(s.ctl, 3);
    for (int part = PART_DETECT; part < N_PARTS; ++part) limiter_part<2>(s, st, 1, 3, part);
  }

  void stage(const DeviceState& s, int stage_idx, hipStream_t st, int nctl) const override {
    for (int part = 0; part < N_PARTS; ++part) stage_part(s, stage_idx, part, st, nctl);
  }

  // one launch (or, for the element-order sweep, one launch sequence) of a stage
  void stage_part(const DeviceState& s, int stage_idx, int part, hipStream_t st, int nctl) const override {
    StageBufs<T> b = bufs(s);
    b.in = stage_input(s, stage_idx);
    if (part == PART_SURFACE) {
      if (fused_on(s)) {
        // the edges are the fused pass's: stage 0 keeps only the step control (a part timed on its own
        // that launches nothing records its events back to back)
        if (stage_idx == 0 && nctl > 0) {
          go(s, k_step_begin, nctl, 64, st, s.ctl, nctl);
        } else if (s.prof_ev[0]) {
          (void)hipEventRecord(s.prof_ev[0], st);
          (void)hipEventRecord(s.prof_ev[1], st);
        }
      } else {
        surface(s, b, st, 0, stage_idx == 0 ? 1 : 0, stage_idx == 0 ? nctl : 0);
      }
    } else if (part == PART_ELEMENT) {
      if (stage_idx == 0)
        element<OP_S0>(s, b, st, 0, true);
      else if (stage_idx == 1)
        element<OP_S1>(s, b, st, 0, true);
      else
        element<OP_S2>(s, b, st, 0, true);
    } else if (s.limiter != DGM_LIMITER_OFF) {
      if (stage_idx == 0)
        limiter_part<0>(s, st, 0, 0, part);
      else if (stage_idx == 1)
        limiter_part<1>(s, st, 0, 1, part);
      else
        limiter_part<2>(s, st, 0, 2, part);
    }
  }

  template <int STAGE>
  void limiter_part(const DeviceState& s, hipStream_t st, int force, int slot, int part) const {
    if (s.mesh.ne <= 0) return;
    StageBufs<T> b = bufs(s);
    const T* U = limiter_input(s, STAGE);
    LimiterBufs lb{s.lim_list, s.lim_w, s.lim_flags, s.lim_pend};
    if (s.limiter == DGM_LIMITER_JACOBI) {
      // the last stage of a step also reduces the next step's time step (not the stand-alone call)
      const int fuse = (STAGE == 2 && slot == 2) ? 1 : 0;
      if (part == PART_DETECT) {
        const int bs = s.tune.detect_block;
        go(s, k_kxrcf<T, P, NQ, NE>, blocks_for(s.mesh.ne, bs), bs, st, s.mesh, U, s.ltab, (const T*)s.tab + O::PHIE,
           lb, s.ctl, slot, force, fuse, s.tune.remap);
      } else if (part == PART_REBUILD) {
        go(s, k_weno<T, P, NQ, NE>, s.tune.weno_grid, 64, st, s.mesh, U, s.smat, lb, s.ctl, slot, fuse, 0);
      } else if (part == PART_APPLY && slot == 3) {
        // only the stand-alone limiter writes back at once; in a march the rebuilds stay pending for
        // the next stage's surface and element kernels (and stage 0 writes back u^n)
        go(s, k_apply<T, NP, STAGE>, 1024, 256, st, lb, b, s.mesh.nes, s.ctl, slot);
      }
    } else if (s.limiter == DGM_LIMITER_SEQUENTIAL && part == PART_DETECT && s.walk && s.lim_flags &&
               s.mesh.ne <= WALK_MAX_E) {
      const int bs = s.tune.detect_block;
      k_kxrcf<T, P, NQ, NE><<<blocks_for(s.mesh.ne, bs), bs, 0, st>>>(s.mesh, U, s.ltab, (const T*)s.tab + O::PHIE, lb,
                                                                       s.ctl, slot, force, 0, s.tune.remap);
      k_weno<T, P, NQ, NE><<<s.tune.weno_grid, 64, 0, st>>>(s.mesh, U, s.smat, lb, s.ctl, slot, 0, 1);
      k_limit_walk<T, P, NQ, NE, STAGE><<<1, 64, 0, st>>>(s.mesh, (const int4*)s.walk, s.walk_chunks, b, U, s.ltab,
                                                          (const T*)s.tab + O::PHIE, s.smat, lb, s.ctl, slot, force);
      k_walk_flush<T, NP, STAGE><<<blocks_for(s.mesh.ne, 256), 256, 0, st>>>(s.lim_pend, s.lim_w, b, s.mesh.ne, s.mesh.nes);
    } else if (s.limiter == DGM_LIMITER_SEQUENTIAL && part == PART_DETECT) {
      for (int l = 0; l + 1 < (int)s.level_offsets.size(); ++l) {
        const int n = s.level_offsets[l + 1] - s.level_offsets[l];
        k_limit_level<T, P, NQ, NE, STAGE><<<n, 64, 0, st>>>(s.mesh, s.levels + s.level_offsets[l], n, b, U, s.ltab,
                                                             (const T*)s.tab + O::PHIE, s.smat, lb, s.ctl, slot,
                                                             force);
      }
    }
  }

  int smoothness_len() const override { return NP * (NP + 1) / 2; }
  void build_smoothness(const DeviceState& s, int e0, int n, hipStream_t st) const override {
    if (n > 0) k_smoothness<P, NQ, NE><<<n, 64, 0, st>>>(s.mesh, s.ltab, s.smat, e0);
  }

  void time_step(const DeviceState& s, hipStream_t st, int force, int only) const override {
    if (s.mesh.ne <= 0) return;
    const T* U = sizeof(T) == 8 ? (const T*)s.un : (const T*)s.sn;
    k_dt<T, NP><<<blocks_for(s.mesh.ne, 256), 256, 0, st>>>(s.mesh, U, s.ctl, force, only);
  }

  void upload(const DeviceState& s, const double* host_layout_dev, int e0, int n, hipStream_t st) const override {
    if (n <= 0) return;
    dim3 g(blocks_for(n, 256), 4 * NP);
    if (sizeof(T) == 8)
      k_to_soa<double><<<g, 256, 0, st>>>(host_layout_dev, nullptr, s.un, e0, n, s.mesh.nes, 4 * NP);
    else
      k_to_soa<float><<<g, 256, 0, st>>>(host_layout_dev, (float*)s.sn, s.un, e0, n, s.mesh.nes, 4 * NP);
  }
  void download(const DeviceState& s, const double* soa, double* host_layout_dev, int e0, int n,
                hipStream_t st) const override {
    if (n <= 0) return;
    dim3 g(blocks_for(n, 256), 4 * NP);
    k_from_soa<<<g, 256, 0, st>>>(soa, host_layout_dev, e0, n, s.mesh.nes, 4 * NP);
  }
  void check_finite(const DeviceState& s, int* flag, int e0, int n, hipStream_t st) const override {
    if (n > 0) k_check_finite<<<blocks_for(n, 256), 256, 0, st>>>(s.un, e0, n, s.mesh.nes, 4 * NP, flag);
  }
  void flush(const DeviceState& s, hipStream_t st) const override {
    if (s.mesh.ne <= 0 || s.limiter != DGM_LIMITER_JACOBI) return;
    k_flush<T, NP><<<blocks_for(s.mesh.ne, 256), 256, 0, st>>>(s.lim_pend, s.lim_w, s.un, (float*)s.sn, s.mesh.ne,
                                                              s.mesh.nes);
  }
  void wall_force(const DeviceState& s, double* force, int ed0, int n, hipStream_t st) const override {
    if (n > 0)
      k_wall_force<T, P, NQ, NE><<<blocks_for(n, 256), 256, 0, st>>>(s.mesh, s.un, s.ltab, s.ctl, ed0, n, force);
  }
};

std::unique_ptr<Solver> make_solver(int precision, int p, int nq, int ne) {
  const bool f32 = precision == DGM_FP32_BULK;
#define DGM_CASE(P_, NQ_, NE_)                                               \
  if (p == P_ && nq == NQ_ && ne == NE_) {                                   \
    if (f32) return std::unique_ptr<Solver>(new SolverImpl<float, P_, NQ_, NE_>()); \
    return std::unique_ptr<Solver>(new SolverImpl<double, P_, NQ_, NE_>());         \
  }
  DGM_CASE(1, 6, 3)
  DGM_CASE(2, 6, 3)
  DGM_CASE(3, 16, 5)
  DGM_CASE(4, 25, 5)
#undef DGM_CASE
  return nullptr;
}

}  // namespace dgm
