// Arena buffers, range allocation and stream-ordered case placement (see arena.h).
#include "arena.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>

namespace dgm {

// ---- RangeAlloc ----

void RangeAlloc::reset(int64_t capacity) {
  free_.clear();
  used_.clear();
  cap_ = capacity;
  if (capacity > 0) free_[0] = capacity;
}

int64_t RangeAlloc::alloc(int64_t n) {
  if (n <= 0) return -1;
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second < n) continue;
    const int64_t start = it->first, len = it->second;
    free_.erase(it);
    if (len > n) free_[start + n] = len - n;
    used_[start] = n;
    return start;
  }
  return -1;
}

void RangeAlloc::release(int64_t start, int64_t n) {
  auto u = used_.find(start);
  if (u == used_.end() || u->second != n) return;
  used_.erase(u);
  int64_t s = start, len = n;
  auto next = free_.lower_bound(s);
  if (next != free_.end() && next->first == s + len) {
    len += next->second;
    next = free_.erase(next);
  }
  if (next != free_.begin()) {
    auto prev = std::prev(next);
    if (prev->first + prev->second == s) {
      s = prev->first;
      len += prev->second;
      free_.erase(prev);
    }
  }
  free_[s] = len;
}

int64_t RangeAlloc::used_end() const { return used_.empty() ? 0 : used_.rbegin()->first + used_.rbegin()->second; }

int64_t RangeAlloc::free_units() const {
  int64_t f = 0;
  for (const auto& kv : free_) f += kv.second;
  return f;
}

int64_t RangeAlloc::largest_free() const {
  int64_t f = 0;
  for (const auto& kv : free_) f = std::max(f, kv.second);
  return f;
}

// ---- helpers ----

std::vector<double> farfield_states(const HostMesh& M, double mach, double aoa_deg, double rho, double p, double fs_hcr,
                                    double hcr) {
  const double cinf = std::sqrt(fs_hcr * p / rho);
  const double rad = aoa_deg * (M_PI / 180.0);
  const double uinf = mach * cinf * std::cos(rad), vinf = mach * cinf * std::sin(rad);
  std::vector<double> q;
  for (size_t b = 0; b < M.patch_type.size(); ++b) {
    q.push_back(rho);
    q.push_back(rho * uinf);
    q.push_back(rho * vinf);
    q.push_back(p / (hcr - 1.0) + 0.5 * rho * (uinf * uinf + vinf * vinf));
  }
  return q;
}

CaseCtl fresh_case(double cfl) {
  CaseCtl h;
  std::memset(&h, 0, sizeof(h));
  h.final_time = 1.0e300;
  h.cfl = cfl;
  h.error = ~0ull;
  h.error_s0 = ~0ull;
  for (int s = 0; s < 3; ++s) h.error_f[s] = ~0ull;
  h.final_step = 0x7fffffff;
  h.active = 1;
  h.live = 1;
  return h;
}

void initial_state(const Tables& T, const HostMesh& M, double density, double aoa_deg, double mach, double pressure,
                   double hcr, double* U) {
  const double a = std::sqrt(hcr * pressure / density);
  const double rad = aoa_deg * (M_PI / 180.0);
  const double u = mach * std::cos(rad) * a, v = mach * std::sin(rad) * a;
  const double q[4] = {density, density * u, density * v, pressure / (hcr - 1.0) + 0.5 * density * (u * u + v * v)};
  for (int64_t e = 0; e < M.ne; ++e) {
    const double hj = 0.5 * M.el_J[e];
    for (int var = 0; var < 4; ++var)
      for (int k = 0; k < T.np; ++k) {
        double s = 0.0;
        for (int i = 0; i < T.nq; ++i) s += hj * T.cell_w[i] * q[var] * T.phi_cell[i * T.np + k];
        U[(e * 4 + var) * T.np + k] = s / (T.mass[k] * hj);
      }
  }
}

double default_cfl(int order, double cfl) {
  if (cfl > 0.0) return cfl;
  const double p = (double)order;
  return 1.0 / ((2.0 * p + 1.0) * (1.0 + 4.0 / ((p + 2.0) * (p + 2.0))));
}

Tuning to_tuning(const dgm_tuning& t) {
  Tuning u;
  u.surface_block = t.surface_block;
  u.element_waves = t.element_waves;
  u.element_block = t.element_block;
  u.detect_block = t.detect_block;
  u.remap = t.remap;
  u.weno_grid = t.weno_grid;
  return u;
}
dgm_tuning from_tuning(const Tuning& u) {
  dgm_tuning t;
  t.surface_block = u.surface_block;
  t.element_waves = u.element_waves;
  t.element_block = u.element_block;
  t.detect_block = u.detect_block;
  t.remap = u.remap;
  t.weno_grid = u.weno_grid;
  return t;
}
static bool valid_block(int v) { return v == 64 || v == 128 || v == 256; }
bool valid_tuning(const dgm_tuning& t) {
  return valid_block(t.surface_block) &&
         (t.element_waves == 0 || t.element_waves == 1 || t.element_waves == 2 || t.element_waves == 4) &&
         valid_block(t.element_block) && valid_block(t.detect_block) && (t.remap == 0 || t.remap == 1) &&
         t.weno_grid >= 64 && t.weno_grid <= 65536;
}


// ---- Arena ----

Footprint Arena::footprint(int order, int nq, int ne_pts, int precision, int limiter, bool residual_buffer) {
  (void)nq;
  const int64_t np = (order + 1) * (order + 2) / 2;
  const int64_t nc = 4 * np;
  const bool f32 = precision == DGM_FP32_BULK;
  Footprint f;
  int64_t e = 0;
  e += nc * 8;                            // u^n
  e += f32 ? 4 * nc * 4 : nc * 8;         // shadows + increment, or the FP64 stage state
  if (residual_buffer) e += nc * 8;       // residual output
  e += 3 * 4 + 3 * 4;                     // edge ids, neighbours
  e += G_N * 8 + G_N * 4;                 // geometry (FP64 + FP32 copy)
  e += 4 + nc * 8 + 1;                    // troubled list, rebuilt polynomials, flags
  if (limiter != DGM_LIMITER_OFF) e += np * (np + 1) / 2 * 8;  // smoothness matrix
  if (f32 && order == 3) e += 3 * 4 + GRP_EDGES * 32 / 16 + 1;  // fused pass: face slots, group edge records
  e += 1;                                 // case map (4 bytes per 64 elements, rounded up)
  f.per_element = e;
  int64_t d = 0;
  d += 2 * 4 + 4;                         // elements, info word
  d += 3 * 8 + 3 * 4;                     // normal and half length (FP64 + FP32)
  d += (int64_t)ne_pts * 4 * (f32 ? 4 : 8);  // point fluxes
  d += 1;                                 // case map
  f.per_edge = d;
  f.per_patch = 4 * 8;
  return f;
}

template <class X>
int Arena::dalloc(X** p, size_t n, std::string& err) {
  *p = nullptr;
  if (n == 0) n = 1;
  void* v = nullptr;
  const hipError_t e = hipMalloc(&v, n * sizeof(X));
  if (e != hipSuccess) {
    err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return DGM_OUT_OF_MEMORY;
  }
  allocs_.push_back(v);
  bytes += (int64_t)(n * sizeof(X));
  *p = (X*)v;
  return DGM_OK;
}

#define ACHK(call)                                                   \
  do {                                                               \
    hipError_t e_ = (call);                                          \
    if (e_ != hipSuccess) {                                          \
      err = std::string(#call ": ") + hipGetErrorString(e_);         \
      return DGM_HIP_ERROR;                                          \
    }                                                                \
  } while (0)

Arena::~Arena() {
  if (stream) (void)hipStreamSynchronize(stream);
  if (hbuf_) (void)hipFree(hbuf_);
  for (void* p : allocs_) (void)hipFree(p);
  if (pinned_) (void)hipHostFree(pinned_);
}

int Arena::init(int order_, int nq, int ne_pts, int precision_, int limiter_, int64_t cap_elements, int64_t cap_edges,
                int64_t cap_patches, bool residual_buffer, CaseCtl* cases, unsigned long long* dtmin, hipStream_t st,
                std::string& err) {
  order = order_;
  precision = precision_;
  limiter = limiter_;
  stream = st;
  if (!build_tables(T, order, nq, ne_pts)) {
    err = "unsupported quadrature rule";
    return DGM_BAD_CONFIG;
  }
  solver = make_solver(precision, order, nq, ne_pts);
  if (!solver) {
    err = "no kernel instantiation for this order and rule";
    return DGM_BAD_CONFIG;
  }
  cap_e = std::max<int64_t>(64, (cap_elements + 63) / 64 * 64);
  cap_ed = std::max<int64_t>(64, (cap_edges + 63) / 64 * 64);
  cap_patch = std::max<int64_t>(1, cap_patches);
  if (cap_e >= (1 << 30) || cap_ed >= (1 << 29) || cap_patch >= (1 << 22)) {
    err = "arena capacity beyond the 32-bit index range";
    return DGM_BAD_CONFIG;
  }
  // The kernels address one variable's block of modes ([Np][cap_e], up to 8 bytes each) and the point
  // fluxes ([cap_ed][Ne][4]) with 32-bit byte offsets from a base pointer.
  {
    const uint64_t np = (uint64_t)(order + 1) * (order + 2) / 2, lim = 1ull << 32;
    const uint64_t flux_bytes = (uint64_t)cap_ed * ne_pts * 4 * (precision == DGM_FP32_BULK ? 4 : 8);
    if (np * (uint64_t)cap_e * 8 >= lim || flux_bytes >= lim) {
      err = "arena capacity beyond the 32-bit offset range of one variable block or of the flux array";
      return DGM_BAD_CONFIG;
    }
  }
  r_e.reset(cap_e / 64);
  r_ed.reset(cap_ed / 64);
  r_patch.reset(cap_patch);
  const int np = T.np, nqq = T.nq, nep = T.ne;
  const bool f32 = solver->fp32();

  // bulk tables: PHIV, DXI, DETA [nq][np] (derivatives pre-divided by the mass, times the weight);
  // PHIE, LIFT [3][ne][np]; KXW [3][np] (the detector's neighbour-side jump weights)
  std::vector<double> tab(solver->table_len(), 0.0);
  if ((int)tab.size() != 3 * nqq * np + 6 * nep * np + 3 * np) {
    err = "table size mismatch";
    return DGM_BAD_CONFIG;
  }
  const int kxi = T.key(1, 0), keta = T.key(0, 1);
  for (int q = 0; q < nqq; ++q)
    for (int k = 0; k < np; ++k) {
      tab[q * np + k] = T.phi_cell[q * np + k];
      tab[nqq * np + q * np + k] = T.cell_w[q] * T.dref[((size_t)kxi * nqq + q) * np + k] / T.mass[k];
      tab[2 * nqq * np + q * np + k] = T.cell_w[q] * T.dref[((size_t)keta * nqq + q) * np + k] / T.mass[k];
    }
  for (int i = 0; i < 3; ++i)
    for (int m = 0; m < nep; ++m)
      for (int k = 0; k < np; ++k) {
        const double ph = T.phi_edge[(i * nep + m) * np + k];
        tab[3 * nqq * np + (i * nep + m) * np + k] = ph;
        tab[3 * nqq * np + 3 * nep * np + (i * nep + m) * np + k] = T.edge_w[m] * ph / T.mass[k];
      }
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < np; ++k) {
      double w = 0.0;
      for (int m = 0; m < nep; ++m) w += T.edge_w[m] * T.phi_edge[(i * nep + (nep - 1 - m)) * np + k];
      tab[3 * nqq * np + 6 * nep * np + i * np + k] = w;
    }
  if (f32) {
    std::vector<float> t32(tab.begin(), tab.end());
    float* p;
    if (int s = dalloc(&p, t32.size(), err)) return s;
    ACHK(hipMemcpy(p, t32.data(), t32.size() * 4, hipMemcpyHostToDevice));
    S.tab = p;
  } else {
    double* p;
    if (int s = dalloc(&p, tab.size(), err)) return s;
    ACHK(hipMemcpy(p, tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
    S.tab = p;
  }
  // FP64 limiter tables: PHIE [3][ne][np], EW [ne], CW [nq], DREF
  std::vector<double> lt;
  lt.insert(lt.end(), T.phi_edge.begin(), T.phi_edge.end());
  lt.insert(lt.end(), T.edge_w.begin(), T.edge_w.end());
  lt.insert(lt.end(), T.cell_w.begin(), T.cell_w.end());
  lt.insert(lt.end(), T.dref.begin(), T.dref.end());
  if ((int)lt.size() != solver->limiter_table_len()) {
    err = "limiter table size mismatch";
    return DGM_BAD_CONFIG;
  }
  if (int s = dalloc(&S.ltab, lt.size(), err)) return s;
  ACHK(hipMemcpy(S.ltab, lt.data(), lt.size() * 8, hipMemcpyHostToDevice));

  // mesh arrays at capacity
  DevMesh& dm = S.mesh;
  dm.ne = 0;
  dm.nes = (int)cap_e;
  dm.nedge = 0;
  dm.ed_stride = (int)cap_ed;
  dm.npatch = (int)cap_patch;
  int *el_edge, *el_nbr, *ed_elem, *ed_info, *elc, *edc;
  double *el_geo, *ed_geo, *pq;
  float *el_geo32, *ed_geo32;
  if (int s = dalloc(&el_edge, 3 * cap_e, err)) return s;
  if (int s = dalloc(&el_nbr, 3 * cap_e, err)) return s;
  if (int s = dalloc(&el_geo, (size_t)G_N * cap_e, err)) return s;
  if (int s = dalloc(&el_geo32, (size_t)G_N * cap_e, err)) return s;
  if (int s = dalloc(&ed_elem, 2 * cap_ed, err)) return s;
  if (int s = dalloc(&ed_info, cap_ed, err)) return s;
  if (int s = dalloc(&ed_geo, 3 * cap_ed, err)) return s;
  if (int s = dalloc(&ed_geo32, 3 * cap_ed, err)) return s;
  if (int s = dalloc(&pq, 4 * cap_patch, err)) return s;
  if (int s = dalloc(&elc, cap_e / 64, err)) return s;
  if (int s = dalloc(&edc, cap_ed / 64, err)) return s;
  ACHK(hipMemset(el_edge, 0, 3 * cap_e * 4));
  ACHK(hipMemset(el_nbr, 0xff, 3 * cap_e * 4));
  ACHK(hipMemset(el_geo, 0, G_N * cap_e * 8));
  ACHK(hipMemset(el_geo32, 0, G_N * cap_e * 4));
  ACHK(hipMemset(ed_elem, 0xff, 2 * cap_ed * 4));
  ACHK(hipMemset(ed_info, 0, cap_ed * 4));
  ACHK(hipMemset(ed_geo, 0, 3 * cap_ed * 8));
  ACHK(hipMemset(ed_geo32, 0, 3 * cap_ed * 4));
  ACHK(hipMemset(pq, 0, 4 * cap_patch * 8));
  ACHK(hipMemset(elc, 0xff, cap_e / 64 * 4));
  ACHK(hipMemset(edc, 0xff, cap_ed / 64 * 4));
  dm.el_edge = el_edge;
  dm.el_nbr = el_nbr;
  dm.el_geo = el_geo;
  dm.el_geo32 = el_geo32;
  dm.ed_elem = ed_elem;
  dm.ed_info = ed_info;
  dm.ed_geo = ed_geo;
  dm.ed_geo32 = ed_geo32;
  dm.patch_q = pq;
  dm.el_case64 = elc;
  dm.ed_case64 = edc;
  dm.single_case = -1;

  // state and work buffers
  const size_t ns = (size_t)4 * np * cap_e;
  if (int s = dalloc(&S.un, ns, err)) return s;
  ACHK(hipMemset(S.un, 0, ns * 8));
  if (f32) {
    float *a, *b, *b2, *d;
    if (int s = dalloc(&a, ns, err)) return s;
    if (int s = dalloc(&b, ns, err)) return s;
    if (int s = dalloc(&b2, ns, err)) return s;
    if (int s = dalloc(&d, ns, err)) return s;
    ACHK(hipMemset(a, 0, ns * 4));
    ACHK(hipMemset(b, 0, ns * 4));
    ACHK(hipMemset(b2, 0, ns * 4));
    ACHK(hipMemset(d, 0, ns * 4));
    S.sn = a;
    S.s = b;
    S.s2 = b2;
    S.d = d;
  } else {
    double* a;
    if (int s = dalloc(&a, ns, err)) return s;
    ACHK(hipMemset(a, 0, ns * 8));
    S.us = a;
  }
  if (residual_buffer) {
    if (int s = dalloc(&S.rout, ns, err)) return s;
  }
  char* fl;
  if (int s = dalloc(&fl, (size_t)cap_ed * nep * 4 * (f32 ? 4 : 8), err)) return s;
  ACHK(hipMemset(fl, 0, (size_t)cap_ed * nep * 4 * (f32 ? 4 : 8)));
  S.flux = fl;
  if (int s = dalloc(&S.lim_list, cap_e, err)) return s;
  if (int s = dalloc(&S.lim_w, (size_t)cap_e * 4 * np, err)) return s;
  if (int s = dalloc(&S.lim_flags, cap_e, err)) return s;
  if (int s = dalloc(&S.lim_pend, cap_e, err)) return s;
  ACHK(hipMemset(S.lim_pend, 0xff, cap_e * 4));
  if (limiter != DGM_LIMITER_OFF) {
    if (int s = dalloc(&S.smat, (size_t)cap_e * solver->smoothness_len(), err)) return s;
  }
  int* count;
  if (int s = dalloc(&count, 4, err)) return s;
  ACHK(hipMemset(count, 0, 16));
  S.ctl.cases = cases;
  S.ctl.dtmin = dtmin;
  S.ctl.count = count;
  S.limiter = limiter;
  S.use_mfma = solver->has_mfma() ? 1 : 0;
  // FP32-bulk element kernel: the matrix-core one from p = 3; at p <= 2 its padded products cost more
  // than the one-lane-per-element kernel (measured on the 128k-triangle O-grid: p = 1 18.2 vs 9.0 us,
  // p = 2 19.5 vs 13.2 us, p = 3 23.3 vs 27.2 us per stage)
  if (order <= 2) S.tune.element_waves = 0;
  // The fused edge-and-element pass of a march (FP32-bulk at p = 3, on the matrix cores) and its group
  // edge records, with DGM_FUSED=1 in the environment at set-up. It gives the separate passes' results
  // bit for bit but is not the default: on the 128k-triangle O-grid it takes 43 us per stage against
  // 14.4 + 23.0 us for the edge and element kernels (a 16-element group evaluates 2.06 edges per element
  // against 1.5, and spends 1.6 times the vector instructions of the two kernels together).
  {
    const char* fz = std::getenv("DGM_FUSED");
    if (f32 && S.use_mfma && order == 3 && fz && fz[0] == '1') {
      int4* gr;
      float4* gg;
      int *gn, *gs;
      if (int s = dalloc(&gr, (size_t)cap_e / 16 * GRP_EDGES, err)) return s;
      if (int s = dalloc(&gg, (size_t)cap_e / 16 * GRP_EDGES, err)) return s;
      if (int s = dalloc(&gn, (size_t)cap_e / 16, err)) return s;
      if (int s = dalloc(&gs, (size_t)3 * cap_e, err)) return s;
      ACHK(hipMemset(gr, 0, (size_t)cap_e / 16 * GRP_EDGES * sizeof(int4)));
      ACHK(hipMemset(gg, 0, (size_t)cap_e / 16 * GRP_EDGES * sizeof(float4)));
      ACHK(hipMemset(gn, 0, (size_t)cap_e / 16 * 4));
      ACHK(hipMemset(gs, 0, (size_t)3 * cap_e * 4));
      dm.grp_rec = gr;
      dm.grp_geo = gg;
      dm.grp_n = gn;
      dm.el_gslot = gs;
      S.fused = 1;
    }
  }
  // rebuild wavefronts: enough for a troubled fraction of 1 / 32 in one pass, 2048 at least
  S.tune.weno_grid = (int)std::min<int64_t>(65536, std::max<int64_t>(2048, cap_e / 32));
  dm.hcr = 1.4;
  return DGM_OK;
}

// Copies (dst, src, bytes) pieces through one pinned staging buffer on the placement stream, then waits
// for that stream.
int Arena::stage_copy(const std::vector<std::pair<void*, std::pair<const void*, size_t>>>& pieces, std::string& err) {
  size_t total = 0;
  for (const auto& p : pieces) total += (p.second.second + 255) / 256 * 256;
  if (total > pinned_size_) {
    if (pinned_) ACHK(hipHostFree(pinned_));
    pinned_ = nullptr;
    ACHK(hipHostMalloc(&pinned_, total, hipHostMallocDefault));
    pinned_size_ = total;
  }
  size_t off = 0;
  for (const auto& p : pieces) {
    const size_t n = p.second.second;
    if (n == 0) continue;
    std::memcpy((char*)pinned_ + off, p.second.first, n);
    ACHK(hipMemcpyAsync(p.first, (char*)pinned_ + off, n, hipMemcpyHostToDevice, pstream()));
    off += (n + 255) / 256 * 256;
  }
  ACHK(hipStreamSynchronize(pstream()));
  return DGM_OK;
}

int Arena::place(const HostMesh& M, const std::vector<double>& patch_q, int case_id, Placement& out, std::string& err) {
  Placement pl;
  pl.ne = M.ne;
  pl.ned = M.nedge;
  pl.npatch = (int64_t)M.patch_type.size();
  pl.ne_alloc = (pl.ne + 63) / 64 * 64;
  pl.ned_alloc = (pl.ned + 63) / 64 * 64;
  const int64_t be = r_e.alloc(pl.ne_alloc / 64);
  if (be < 0) return DGM_NO_SPACE;
  const int64_t bd = r_ed.alloc(pl.ned_alloc / 64);
  if (bd < 0) {
    r_e.release(be, pl.ne_alloc / 64);
    return DGM_NO_SPACE;
  }
  const int64_t bp = pl.npatch > 0 ? r_patch.alloc(pl.npatch) : 0;
  if (bp < 0) {
    r_e.release(be, pl.ne_alloc / 64);
    r_ed.release(bd, pl.ned_alloc / 64);
    return DGM_NO_SPACE;
  }
  pl.e_off = be * 64;
  pl.ed_off = bd * 64;
  pl.patch_off = bp;
  const int64_t ne = pl.ne, ned = pl.ned, eo = pl.e_off, edo = pl.ed_off;
  const int64_t nes = cap_e, est = cap_ed;
  // element arrays, rows of the [k][capacity] layouts
  std::vector<int> el_edge(3 * ne), el_nbr(3 * ne);
  std::vector<double> el_geo((size_t)G_N * ne);
  for (int64_t e = 0; e < ne; ++e) {
    for (int i = 0; i < 3; ++i) {
      const int64_t id = M.el_edge[3 * e + i];
      const int side = (M.ed_elem[2 * id + 1] == e && M.ed_lic[2 * id + 1] == i) ? 1 : 0;
      el_edge[i * ne + e] = (int)((edo + id) << 1) | side;
      const int64_t nb = M.el_nbr[3 * e + i];
      el_nbr[i * ne + e] = nb < 0 ? -1 : (int)(eo + nb);
    }
    const double J = M.el_J[e];
    el_geo[G_J * ne + e] = J;
    el_geo[G_XX * ne + e] = M.el_metric[4 * e + 0];
    el_geo[G_XY * ne + e] = M.el_metric[4 * e + 1];
    el_geo[G_EX * ne + e] = M.el_metric[4 * e + 2];
    el_geo[G_EY * ne + e] = M.el_metric[4 * e + 3];
    el_geo[G_CR15 * ne + e] = std::pow(M.el_cr[e], 1.5);
    el_geo[G_HMIN * ne + e] = M.el_hmin[e];
    el_geo[G_2J * ne + e] = 2.0 / J;
  }
  std::vector<float> el_geo32(el_geo.begin(), el_geo.end());
  std::vector<int> ed_elem(2 * ned), ed_info(ned);
  std::vector<double> ed_geo(3 * ned);
  for (int64_t i = 0; i < ned; ++i) {
    ed_elem[i] = (int)(eo + M.ed_elem[2 * i]);
    ed_elem[ned + i] = M.ed_elem[2 * i + 1] < 0 ? -1 : (int)(eo + M.ed_elem[2 * i + 1]);
    const int licL = (int)M.ed_lic[2 * i];
    if (M.ed_patch[i] < 0) {
      ed_info[i] = licL | ((int)M.ed_lic[2 * i + 1] << 2) | (BC_INTERIOR << 4);
      ed_geo[i] = M.ed_normal[2 * i];
      ed_geo[ned + i] = M.ed_normal[2 * i + 1];
    } else {
      const int patch = (int)M.ed_patch[i];
      const int bc = M.patch_type[patch] == DGM_WALL ? BC_WALL : BC_FARFIELD;
      ed_info[i] = licL | (bc << 4) | ((int)(pl.patch_off + patch) << 8);
      ed_geo[i] = M.ed_bc_normal[2 * i];
      ed_geo[ned + i] = M.ed_bc_normal[2 * i + 1];
    }
    ed_geo[2 * ned + i] = 0.5 * M.ed_len[i];
  }
  std::vector<float> ed_geo32(ed_geo.begin(), ed_geo.end());
  DevMesh& dm = S.mesh;
  std::vector<std::pair<void*, std::pair<const void*, size_t>>> pc;
  for (int i = 0; i < 3; ++i) {
    pc.push_back({(void*)(dm.el_edge + i * nes + eo), {el_edge.data() + i * ne, ne * 4}});
    pc.push_back({(void*)(dm.el_nbr + i * nes + eo), {el_nbr.data() + i * ne, ne * 4}});
  }
  for (int g = 0; g < G_N; ++g) {
    pc.push_back({(void*)(dm.el_geo + g * nes + eo), {el_geo.data() + g * ne, ne * 8}});
    pc.push_back({(void*)(dm.el_geo32 + g * nes + eo), {el_geo32.data() + g * ne, ne * 4}});
  }
  for (int i = 0; i < 2; ++i) pc.push_back({(void*)(dm.ed_elem + i * est + edo), {ed_elem.data() + i * ned, ned * 4}});
  pc.push_back({(void*)(dm.ed_info + edo), {ed_info.data(), ned * 4}});
  for (int i = 0; i < 3; ++i) {
    pc.push_back({(void*)(dm.ed_geo + i * est + edo), {ed_geo.data() + i * ned, ned * 8}});
    pc.push_back({(void*)(dm.ed_geo32 + i * est + edo), {ed_geo32.data() + i * ned, ned * 4}});
  }
  if (pl.npatch > 0)
    pc.push_back({(void*)(dm.patch_q + 4 * pl.patch_off), {patch_q.data(), (size_t)pl.npatch * 4 * 8}});
  // group edge lists of the fused pass: per 16 elements (the groups of the 64-aligned range), the
  // distinct edges of the case's elements in increasing order, and each face's place in its list
  const int64_t ngrp = pl.ne_alloc / 16;
  std::vector<int4> grp_rec;
  std::vector<float4> grp_geo;
  std::vector<int> grp_n, el_gslot;
  if (dm.grp_rec) {
    grp_rec.assign((size_t)ngrp * GRP_EDGES, int4{0, -1, 0, 0});
    grp_geo.assign((size_t)ngrp * GRP_EDGES, float4{0.f, 0.f, 0.f, 0.f});
    grp_n.assign((size_t)ngrp, 0);
    el_gslot.assign((size_t)3 * ne, 0);
    for (int64_t gi = 0; gi < ngrp; ++gi) {
      const int64_t lo = gi * 16, hi = std::min<int64_t>(ne, lo + 16);
      std::vector<int64_t> ids;
      for (int64_t e = lo; e < hi; ++e)
        for (int i = 0; i < 3; ++i) ids.push_back(M.el_edge[3 * e + i]);
      std::sort(ids.begin(), ids.end());
      ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
      grp_n[gi] = (int)ids.size();
      for (size_t k = 0; k < ids.size(); ++k) {
        const int64_t id = ids[k];
        const bool interior = ((ed_info[id] >> 4) & 3) == BC_INTERIOR;
        grp_rec[(size_t)gi * GRP_EDGES + k] = int4{ed_elem[id], interior ? ed_elem[ned + id] : -1, ed_info[id], (int)(edo + id)};
        grp_geo[(size_t)gi * GRP_EDGES + k] = float4{ed_geo32[id], ed_geo32[ned + id], ed_geo32[2 * ned + id], 0.f};
      }
      for (int64_t e = lo; e < hi; ++e)
        for (int i = 0; i < 3; ++i) {
          const int k = (int)(std::lower_bound(ids.begin(), ids.end(), M.el_edge[3 * e + i]) - ids.begin());
          el_gslot[i * ne + e] = (k << 1) | (el_edge[i * ne + e] & 1);
        }
    }
    for (int i = 0; i < 3; ++i)
      pc.push_back({(void*)(const_cast<int*>(dm.el_gslot) + i * nes + eo), {el_gslot.data() + i * ne, ne * 4}});
    pc.push_back({(void*)(const_cast<int4*>(dm.grp_rec) + eo / 16 * GRP_EDGES), {grp_rec.data(), grp_rec.size() * sizeof(int4)}});
    pc.push_back({(void*)(const_cast<float4*>(dm.grp_geo) + eo / 16 * GRP_EDGES), {grp_geo.data(), grp_geo.size() * sizeof(float4)}});
    pc.push_back({(void*)(const_cast<int*>(dm.grp_n) + eo / 16), {grp_n.data(), grp_n.size() * 4}});
  }
  std::vector<int> elc((size_t)(pl.ne_alloc / 64), case_id), edc((size_t)(pl.ned_alloc / 64), case_id);
  if (int s = stage_copy(pc, err)) return s;
  if (limiter != DGM_LIMITER_OFF) {
    solver->build_smoothness(S, (int)eo, (int)ne, pstream());
    ACHK(hipGetLastError());
  }
  // the blocks become the case's last, after its data (stream order)
  if (int s = stage_copy({{(void*)(const_cast<int*>(dm.el_case64) + eo / 64), {elc.data(), elc.size() * 4}},
                          {(void*)(const_cast<int*>(dm.ed_case64) + edo / 64), {edc.data(), edc.size() * 4}}},
                         err))
    return s;
  out = pl;
  return DGM_OK;
}

int Arena::unplace(const Placement& pl, std::string& err, bool release) {
  DevMesh& dm = S.mesh;
  // -1 in every byte: the blocks belong to no case (kernels already queued still see the case)
  ACHK(hipMemsetAsync(S.lim_pend + pl.e_off, 0xff, pl.ne_alloc * 4, stream));
  ACHK(hipMemsetAsync(const_cast<int*>(dm.el_case64) + pl.e_off / 64, 0xff, pl.ne_alloc / 64 * 4, stream));
  ACHK(hipMemsetAsync(const_cast<int*>(dm.ed_case64) + pl.ed_off / 64, 0xff, pl.ned_alloc / 64 * 4, stream));
  if (release) {
    ACHK(hipStreamSynchronize(stream));
    release_ranges(pl);
  }
  return DGM_OK;
}

void Arena::release_ranges(const Placement& pl) {
  r_e.release(pl.e_off / 64, pl.ne_alloc / 64);
  r_ed.release(pl.ed_off / 64, pl.ned_alloc / 64);
  if (pl.npatch > 0) r_patch.release(pl.patch_off, pl.npatch);
}

int Arena::grow_staging(size_t n, std::string& err) {
  if (n <= hbuf_size_) return DGM_OK;
  ACHK(hipStreamSynchronize(stream));
  if (ustream) ACHK(hipStreamSynchronize(ustream));
  if (hbuf_) {
    ACHK(hipFree(hbuf_));
    bytes -= (int64_t)(hbuf_size_ * sizeof(double));
  }
  hbuf_ = nullptr;
  hbuf_size_ = 0;
  void* v = nullptr;
  const hipError_t e = hipMalloc(&v, n * sizeof(double));
  if (e != hipSuccess) {
    err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return DGM_OUT_OF_MEMORY;
  }
  hbuf_ = (double*)v;
  hbuf_size_ = n;
  bytes += (int64_t)(n * sizeof(double));
  return DGM_OK;
}

int Arena::upload_state(const double* U, const Placement& pl, std::string& err) {
  const size_t n = (size_t)pl.ne * 4 * T.np;
  if (int s = grow_staging(n, err)) return s;
  ACHK(hipMemcpyAsync(hbuf_, U, n * 8, hipMemcpyHostToDevice, pstream()));
  solver->upload(S, hbuf_, (int)pl.e_off, (int)pl.ne, pstream());
  // nothing of an earlier state stays pending over the uploaded one
  ACHK(hipMemsetAsync(S.lim_pend + pl.e_off, 0xff, pl.ne_alloc * 4, pstream()));
  ACHK(hipGetLastError());
  ACHK(hipStreamSynchronize(pstream()));
  return DGM_OK;
}

int Arena::download_state(const double* soa, double* U, const Placement& pl, std::string& err) {
  const size_t n = (size_t)pl.ne * 4 * T.np;
  if (int s = grow_staging(n, err)) return s;
  solver->flush(S, stream);
  solver->download(S, soa, hbuf_, (int)pl.e_off, (int)pl.ne, stream);
  ACHK(hipGetLastError());
  ACHK(hipMemcpyAsync(U, hbuf_, n * 8, hipMemcpyDeviceToHost, stream));
  ACHK(hipStreamSynchronize(stream));
  return DGM_OK;
}

void Arena::update_extent() {
  S.mesh.ne = (int)(r_e.used_end() * 64);
  S.mesh.nedge = (int)(r_ed.used_end() * 64);
}

}  // namespace dgm

namespace dgm {

// The state a step reads from earlier steps: u^n (and its FP32 shadow) of every arena, the case
// records and their dt slots. Everything else a step reads it writes first in the same step.
struct StepSnapshot {
  std::vector<std::pair<void*, void*>> pairs;  // (live buffer, copy)
  std::vector<size_t> sizes;
  ~StepSnapshot() {
    for (auto& p : pairs) (void)hipFree(p.second);
  }
  int add(void* live, size_t bytes, std::string& err) {
    void* c = nullptr;
    ACHK(hipMalloc(&c, bytes));
    pairs.push_back({live, c});
    sizes.push_back(bytes);
    return DGM_OK;
  }
  int copy(bool restore, hipStream_t st, std::string& err) {
    for (size_t i = 0; i < pairs.size(); ++i) {
      void* dst = restore ? pairs[i].first : pairs[i].second;
      void* src = restore ? pairs[i].second : pairs[i].first;
      ACHK(hipMemcpyAsync(dst, src, sizes[i], hipMemcpyDeviceToDevice, st));
    }
    return DGM_OK;
  }
};

int autotune_steps(const std::vector<Arena*>& arenas, CaseCtl* cases, unsigned long long* dtmin, int ncases,
                   hipStream_t st, const std::function<void()>& step, int reps, std::string& err) {
  // write back pending rebuilds first: the snapshot is then the whole state, and the marks the trials
  // leave are cleared at the end
  for (Arena* A : arenas) A->solver->flush(A->S, st);
  ACHK(hipGetLastError());
  ACHK(hipStreamSynchronize(st));
  StepSnapshot snap;
  for (Arena* A : arenas) {
    const size_t ns = (size_t)4 * A->T.np * A->cap_e;
    if (int s = snap.add(A->S.un, ns * 8, err)) return s;
    if (A->S.sn)
      if (int s = snap.add(A->S.sn, ns * 4, err)) return s;
  }
  if (int s = snap.add(cases, sizeof(CaseCtl) * ncases, err)) return s;
  if (int s = snap.add(dtmin, sizeof(unsigned long long) * DT_CASE * ncases, err)) return s;
  if (int s = snap.copy(false, st, err)) return s;
  hipEvent_t e0, e1;
  ACHK(hipEventCreate(&e0));
  ACHK(hipEventCreate(&e1));
  // mean ms of one step from the snapshot (one untimed step first)
  auto trial = [&](double& ms) -> int {
    if (int s = snap.copy(true, st, err)) return s;
    step();
    ACHK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) step();
    ACHK(hipEventRecord(e1, st));
    ACHK(hipEventSynchronize(e1));
    ACHK(hipGetLastError());
    float x = 0.f;
    ACHK(hipEventElapsedTime(&x, e0, e1));
    ms = x / reps;
    return DGM_OK;
  };
  const int blocks[3] = {64, 128, 256};
  const int waves[3] = {1, 2, 4};
  const int remaps[2] = {0, 1};
  int status = DGM_OK;
  for (Arena* A : arenas) {
    if (A->extent_elements() <= 0) continue;
    auto pick = [&](auto set, const int* vals, int n) -> int {
      double bt = 1e30;
      Tuning best = A->S.tune;
      for (int i = 0; i < n; ++i) {
        Tuning cand = best;
        set(cand, vals[i]);
        A->S.tune = cand;
        double ms = 0.0;
        if (int s = trial(ms)) return s;
        if (ms < bt) {
          bt = ms;
          set(best, vals[i]);
        }
      }
      A->S.tune = best;
      return DGM_OK;
    };
    // (with the fused edge-and-element pass no edge kernel runs in a march)
    if (!A->S.fused) status = pick([](Tuning& t, int v) { t.surface_block = v; }, blocks, 3);
    if (!status) {
      // the shape of the element kernel variant in use. The variant itself (matrix-core, or 0: one lane
      // per element) is not searched: the two round FP32 sums differently, and a case must give the
      // same state alone and in any batch whatever shapes their arenas were tuned to.
      if (A->S.use_mfma && A->solver->has_mfma() && A->S.tune.element_waves > 0)
        status = pick([](Tuning& t, int v) { t.element_waves = v; }, waves, 3);
      else
        status = pick([](Tuning& t, int v) { t.element_block = v; }, blocks, 3);
    }
    // (the rebuild's wavefront count is not searched: the troubled-cell count it must cover grows
    // as shocks form, so a value timed on the current state under-provisions later steps)
    if (!status && A->limiter == DGM_LIMITER_JACOBI)
      status = pick([](Tuning& t, int v) { t.detect_block = v; }, blocks, 3);
    if (!status) status = pick([](Tuning& t, int v) { t.remap = v; }, remaps, 2);
    if (status) break;
  }
  if (int s = snap.copy(true, st, err)) status = status ? status : s;
  for (Arena* A : arenas)
    if (hipMemsetAsync(A->S.lim_pend, 0xff, A->cap_e * 4, st) != hipSuccess && !status) status = DGM_HIP_ERROR;
  (void)hipStreamSynchronize(st);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return status;
}

int profile_step_parts(const std::vector<Arena*>& arenas, CaseCtl* cases, unsigned long long* dtmin, int ncases,
                       hipStream_t st, int n, std::vector<std::array<double, 5>>& ms, double& step_ms,
                       std::string& err) {
  constexpr int NPART = Solver::N_PARTS;
  ms.assign(arenas.size(), std::array<double, 5>{0, 0, 0, 0, 0});
  step_ms = 0.0;
  if (n < 1 || arenas.empty()) return DGM_OK;
  for (Arena* A : arenas) A->solver->flush(A->S, st);
  ACHK(hipGetLastError());
  ACHK(hipStreamSynchronize(st));
  StepSnapshot snap;
  for (Arena* A : arenas) {
    const size_t ns = (size_t)4 * A->T.np * A->cap_e;
    if (int s = snap.add(A->S.un, ns * 8, err)) return s;
    if (A->S.sn)
      if (int s = snap.add(A->S.sn, ns * 4, err)) return s;
  }
  if (int s = snap.add(cases, sizeof(CaseCtl) * ncases, err)) return s;
  if (int s = snap.add(dtmin, sizeof(unsigned long long) * DT_CASE * ncases, err)) return s;
  if (int s = snap.copy(false, st, err)) return s;
  // a start and a stop event per (arena, stage, part), and two around the step
  std::vector<hipEvent_t> ev(2 * arenas.size() * 3 * NPART + 2, nullptr);
  int status = DGM_OK;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) status = DGM_HIP_ERROR;
  const Ctl shared{cases, dtmin, nullptr};
  auto launches = [](const Arena* A, int part) {
    if (part <= Solver::PART_ELEMENT) return true;
    return A->limiter == DGM_LIMITER_JACOBI && part <= Solver::PART_REBUILD;
  };
  for (int it = 0; it < n && !status; ++it) {
    if (hipEventRecord(ev[ev.size() - 2], st) != hipSuccess) status = DGM_HIP_ERROR;
    for (Arena* A : arenas)
      if (A->limiter != DGM_LIMITER_JACOBI && A->extent_elements() > 0) A->solver->time_step(A->S, st, 0);
    launch_step_begin(shared, ncases, st);
    for (int s = 0; s < 3; ++s)
      for (size_t a = 0; a < arenas.size(); ++a)
        for (int part = 0; part < NPART; ++part) {
          Arena* A = arenas[a];
          if (!launches(A, part)) continue;
          hipEvent_t* pe = &ev[2 * ((a * 3 + s) * NPART + part)];
          if (A->extent_elements() > 0) {
            A->S.prof_ev[0] = pe[0];
            A->S.prof_ev[1] = pe[1];
            A->solver->stage_part(A->S, s, part, st, 0);
            A->S.prof_ev[0] = A->S.prof_ev[1] = nullptr;
          } else {
            (void)hipEventRecord(pe[0], st);
            (void)hipEventRecord(pe[1], st);
          }
        }
    if (hipEventRecord(ev[ev.size() - 1], st) != hipSuccess) status = DGM_HIP_ERROR;
    if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess) status = DGM_HIP_ERROR;
    float x = 0.f;
    for (size_t a = 0; a < arenas.size() && !status; ++a)
      for (int s = 0; s < 3; ++s)
        for (int part = 0; part < NPART; ++part) {
          if (!launches(arenas[a], part)) continue;
          hipEvent_t* pe = &ev[2 * ((a * 3 + s) * NPART + part)];
          if (hipEventElapsedTime(&x, pe[0], pe[1]) != hipSuccess) status = DGM_HIP_ERROR;
          ms[a][part] += x / (3.0 * n);
        }
    if (!status && hipEventElapsedTime(&x, ev[ev.size() - 2], ev[ev.size() - 1]) == hipSuccess) step_ms += x / n;
  }
  if (status) err = "kernel timing failed";
  if (int s = snap.copy(true, st, err)) status = status ? status : s;
  for (Arena* A : arenas)
    if (hipMemsetAsync(A->S.lim_pend, 0xff, A->cap_e * 4, st) != hipSuccess && !status) status = DGM_HIP_ERROR;
  (void)hipStreamSynchronize(st);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  return status;
}

}  // namespace dgm
