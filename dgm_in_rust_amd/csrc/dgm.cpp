// C ABI of include/dgm.h: context set-up (tables, mesh, device upload), host-buffer entry points
// mirroring SpatialDisc / TemperalDisc, and the device-resident time march replayed from a
// captured HIP graph of one SSP-RK3 step.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "dgm_internal.h"
#include "solver.h"

using namespace dgm;

struct dgm_ctx {
  dgm_params prm{};
  Tables T;
  HostMesh M;
  std::unique_ptr<Solver> solver;
  DeviceState S;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  int64_t bytes = 0;
  double* hbuf = nullptr;     // device staging buffer in host layout [ne][4][np]
  double* d_force = nullptr;  // [2]
  int* d_flag = nullptr;
  int* d_levels = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  dgm_error err{};
  double cfl = 0.0;
  int nes = 0;
};

namespace {

int fail(dgm_ctx* c, int status, const std::string& msg, int64_t edge = -1, int64_t element = -1) {
  if (c) {
    c->err.status = status;
    c->err.edge = edge;
    c->err.element = element;
    std::snprintf(c->err.message, sizeof(c->err.message), "%s", msg.c_str());
  }
  return status;
}

#define HIPCHK(c, call)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return fail((c), DGM_HIP_ERROR, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class X>
int dalloc(dgm_ctx* c, X** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  void* v = nullptr;
  const hipError_t e = hipMalloc(&v, n * sizeof(X));
  if (e != hipSuccess) return fail(c, DGM_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  c->allocs.push_back(v);
  c->bytes += (int64_t)(n * sizeof(X));
  *p = (X*)v;
  return DGM_OK;
}

template <class X>
int upload(dgm_ctx* c, X** p, const std::vector<X>& h) {
  int s = dalloc(c, p, h.size());
  if (s) return s;
  if (!h.empty()) HIPCHK(c, hipMemcpy(*p, h.data(), h.size() * sizeof(X), hipMemcpyHostToDevice));
  return DGM_OK;
}

size_t nstate(const dgm_ctx* c) { return (size_t)c->M.ne * 4 * c->T.np; }

int read_ctl(dgm_ctx* c, Ctl& h) {
  if (!c->stream) return fail(c, DGM_BAD_ARGUMENT, "host-only context");
  HIPCHK(c, hipMemcpyAsync(&h, c->S.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DGM_OK;
}
int write_ctl(dgm_ctx* c, const Ctl& h) {
  HIPCHK(c, hipMemcpyAsync(c->S.ctl, &h, sizeof(Ctl), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DGM_OK;
}

Ctl fresh_ctl(const dgm_ctx* c) {
  Ctl h;
  std::memset(&h, 0, sizeof(h));
  h.t = 0.0;
  h.dt = 0.0;
  h.final_time = 1.0e300;
  h.cfl = c->cfl;
  h.hcr = c->prm.hcr;
  const double big = 1.0e10;
  for (int sl = 0; sl < 2; ++sl)
    for (int i = 0; i < DT_LANES * DT_STRIDE; ++i) std::memcpy(&h.dtmin_bits[sl][i], &big, 8);
  h.pending = 0;
  h.error = ~0ull;
  h.step = 0;
  h.final_step = 0x7fffffff;
  h.active = 1;
  return h;
}

// status word of the device: lowest failing edge and its code
int ctl_status(dgm_ctx* c, const Ctl& h) {
  if (h.error == ~0ull) return DGM_OK;
  const int code = (int)(h.error & 0xff);
  const int64_t edge = (int64_t)(h.error >> 8);
  if (code == 1) return fail(c, DGM_NEG_DENSITY, "Negative density found in HLLC!", edge);
  if (code == 2) return fail(c, DGM_NEG_PRESSURE, "Negative pressure found in HLLC!", edge);
  if (code == 5) return fail(c, DGM_BAD_MESH, "Invalid number of boundary neighbours!", -1, edge == 0xFFFFFF ? -1 : edge);
  return fail(c, DGM_HIP_ERROR, "unknown device status");
}

int stage_in(dgm_ctx* c, const double* U) {
  if (!c->stream) return fail(c, DGM_BAD_ARGUMENT, "host-only context");
  HIPCHK(c, hipMemcpyAsync(c->hbuf, U, nstate(c) * 8, hipMemcpyHostToDevice, c->stream));
  c->solver->upload(c->S, c->hbuf, c->stream);
  HIPCHK(c, hipGetLastError());
  return DGM_OK;
}
int stage_out(dgm_ctx* c, const double* soa, double* U) {
  c->solver->download(c->S, soa, c->hbuf, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(U, c->hbuf, nstate(c) * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DGM_OK;
}

// The Jacobi sweep's last stage reduces the next step's time step itself; other modes run the
// stand-alone reduction at the start of each step.
bool fused_dt(const dgm_ctx* c) { return c->prm.limiter == DGM_LIMITER_JACOBI; }

void enqueue_step(dgm_ctx* c) {
  if (!fused_dt(c)) c->solver->time_step(c->S, c->stream);
  launch_step_begin(c->S.ctl, c->stream);
  for (int s = 0; s < 3; ++s) c->solver->stage(c->S, s, c->stream);
}

// fresh control block for a resident run, with the first step's time step already reduced
int start_run(dgm_ctx* c, int32_t final_step, double final_time) {
  Ctl h = fresh_ctl(c);
  h.final_step = final_step;
  h.final_time = final_time;
  if (int s = write_ctl(c, h)) return s;
  if (fused_dt(c)) c->solver->time_step(c->S, c->stream);
  HIPCHK(c, hipGetLastError());
  return DGM_OK;
}

// the step counted at the start of the next one (Ctl::pending) is already done
void effective_clock(const Ctl& h, double& t, int32_t& steps) {
  t = h.pending ? h.t + h.dt : h.t;
  steps = h.step + (h.pending ? 1 : 0);
}

int ensure_graph(dgm_ctx* c) {
  if (!c->stream) return fail(c, DGM_BAD_ARGUMENT, "host-only context");
  if (c->graph_exec) return DGM_OK;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  enqueue_step(c);
  HIPCHK(c, hipStreamEndCapture(c->stream, &c->graph));
  HIPCHK(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
  return DGM_OK;
}

int build_device(dgm_ctx* c) {
  const Tables& T = c->T;
  const HostMesh& M = c->M;
  const int np = T.np, nq = T.nq, ne = T.ne;
  const int nes = (int)((M.ne + 63) / 64 * 64);
  c->nes = nes;
  DeviceState& S = c->S;
  const bool f32 = c->solver->fp32();
  // bulk tables: PHIV, DXI, DETA [nq][np]; PHIE, LIFT [3][ne][np]
  std::vector<double> tab(c->solver->table_len(), 0.0);
  if ((int)tab.size() != 3 * nq * np + 6 * ne * np) return fail(c, DGM_BAD_CONFIG, "table size mismatch");
  const int kxi = T.key(1, 0), keta = T.key(0, 1);
  for (int q = 0; q < nq; ++q)
    for (int k = 0; k < np; ++k) {
      tab[q * np + k] = T.phi_cell[q * np + k];
      tab[nq * np + q * np + k] = T.cell_w[q] * T.dref[((size_t)kxi * nq + q) * np + k] / T.mass[k];
      tab[2 * nq * np + q * np + k] = T.cell_w[q] * T.dref[((size_t)keta * nq + q) * np + k] / T.mass[k];
    }
  for (int i = 0; i < 3; ++i)
    for (int m = 0; m < ne; ++m)
      for (int k = 0; k < np; ++k) {
        const double ph = T.phi_edge[(i * ne + m) * np + k];
        tab[3 * nq * np + (i * ne + m) * np + k] = ph;
        tab[3 * nq * np + 3 * ne * np + (i * ne + m) * np + k] = T.edge_w[m] * ph / T.mass[k];
      }
  if (f32) {
    std::vector<float> t32(tab.begin(), tab.end());
    float* p;
    if (int s = upload(c, &p, t32)) return s;
    S.tab = p;
  } else {
    double* p;
    if (int s = upload(c, &p, tab)) return s;
    S.tab = p;
  }
  // FP64 limiter tables: PHIE [3][ne][np], EW [ne], CW [nq], DREF
  std::vector<double> lt;
  lt.insert(lt.end(), T.phi_edge.begin(), T.phi_edge.end());
  lt.insert(lt.end(), T.edge_w.begin(), T.edge_w.end());
  lt.insert(lt.end(), T.cell_w.begin(), T.cell_w.end());
  lt.insert(lt.end(), T.dref.begin(), T.dref.end());
  if ((int)lt.size() != c->solver->limiter_table_len()) return fail(c, DGM_BAD_CONFIG, "limiter table size mismatch");
  if (int s = upload(c, &S.ltab, lt)) return s;

  // mesh, structure of arrays
  const int64_t nel = M.ne, ned = M.nedge;
  std::vector<int> el_edge(3 * (size_t)nes, 0), el_nbr(3 * (size_t)nes, -1);
  std::vector<double> el_geo((size_t)G_N * nes, 0.0);
  for (int64_t e = 0; e < nel; ++e) {
    for (int i = 0; i < 3; ++i) {
      const int64_t id = M.el_edge[3 * e + i];
      const int side = (M.ed_elem[2 * id + 1] == e && M.ed_lic[2 * id + 1] == i) ? 1 : 0;
      el_edge[i * nes + e] = (int)(id << 1) | side;
      el_nbr[i * nes + e] = (int)M.el_nbr[3 * e + i];
    }
    const double J = M.el_J[e];
    el_geo[G_J * nes + e] = J;
    el_geo[G_XX * nes + e] = M.el_metric[4 * e + 0];
    el_geo[G_XY * nes + e] = M.el_metric[4 * e + 1];
    el_geo[G_EX * nes + e] = M.el_metric[4 * e + 2];
    el_geo[G_EY * nes + e] = M.el_metric[4 * e + 3];
    el_geo[G_CR * nes + e] = M.el_cr[e];
    el_geo[G_HMIN * nes + e] = M.el_hmin[e];
    el_geo[G_2J * nes + e] = 2.0 / J;
  }
  std::vector<int> ed_elem(2 * ned), ed_info(ned);
  std::vector<double> ed_geo(3 * ned);
  for (int64_t i = 0; i < ned; ++i) {
    ed_elem[i] = (int)M.ed_elem[2 * i];
    ed_elem[ned + i] = (int)M.ed_elem[2 * i + 1];
    const int licL = (int)M.ed_lic[2 * i];
    if (M.ed_patch[i] < 0) {
      ed_info[i] = licL | ((int)M.ed_lic[2 * i + 1] << 2) | (BC_INTERIOR << 4);
      ed_geo[i] = M.ed_normal[2 * i];
      ed_geo[ned + i] = M.ed_normal[2 * i + 1];
    } else {
      const int patch = (int)M.ed_patch[i];
      const int bc = M.patch_type[patch] == DGM_WALL ? BC_WALL : BC_FARFIELD;
      ed_info[i] = licL | (bc << 4) | (patch << 8);
      ed_geo[i] = M.ed_bc_normal[2 * i];
      ed_geo[ned + i] = M.ed_bc_normal[2 * i + 1];
    }
    ed_geo[2 * ned + i] = 0.5 * M.ed_len[i];
  }
  // far-field conserved state per patch (src/initialization.rs:124-133, boundary_condition.rs:50-55)
  const dgm_params& P = c->prm;
  const double cinf = std::sqrt(P.fs_hcr * P.fs_p / P.fs_rho);
  const double rad = P.fs_aoa_deg * (M_PI / 180.0);
  const double uinf = P.fs_mach * cinf * std::cos(rad), vinf = P.fs_mach * cinf * std::sin(rad);
  std::vector<double> patch_q;
  for (size_t b = 0; b < M.patch_type.size(); ++b) {
    patch_q.push_back(P.fs_rho);
    patch_q.push_back(P.fs_rho * uinf);
    patch_q.push_back(P.fs_rho * vinf);
    patch_q.push_back(P.fs_p / (P.hcr - 1.0) + 0.5 * P.fs_rho * (uinf * uinf + vinf * vinf));
  }
  DevMesh& dm = S.mesh;
  dm.ne = (int)nel;
  dm.nes = nes;
  dm.nedge = (int)ned;
  dm.npatch = (int)M.patch_type.size();
  int* pi;
  double* pd;
  float* pf;
  if (int s = upload(c, &pi, el_edge)) return s;
  dm.el_edge = pi;
  if (int s = upload(c, &pi, el_nbr)) return s;
  dm.el_nbr = pi;
  if (int s = upload(c, &pd, el_geo)) return s;
  dm.el_geo = pd;
  if (int s = upload(c, &pf, std::vector<float>(el_geo.begin(), el_geo.end()))) return s;
  dm.el_geo32 = pf;
  if (int s = upload(c, &pi, ed_elem)) return s;
  dm.ed_elem = pi;
  if (int s = upload(c, &pi, ed_info)) return s;
  dm.ed_info = pi;
  if (int s = upload(c, &pd, ed_geo)) return s;
  dm.ed_geo = pd;
  if (int s = upload(c, &pf, std::vector<float>(ed_geo.begin(), ed_geo.end()))) return s;
  dm.ed_geo32 = pf;
  if (int s = upload(c, &pd, patch_q)) return s;
  dm.patch_q = pd;

  // state and work buffers
  const size_t ns = (size_t)4 * np * nes;
  const size_t eb = f32 ? 4 : 8;
  if (int s = dalloc(c, &S.un, ns)) return s;
  HIPCHK(c, hipMemset(S.un, 0, ns * 8));
  if (f32) {
    float *a, *b2, *d;
    if (int s = dalloc(c, &a, ns)) return s;
    if (int s = dalloc(c, &b2, ns)) return s;
    if (int s = dalloc(c, &d, ns)) return s;
    HIPCHK(c, hipMemset(a, 0, ns * 4));
    HIPCHK(c, hipMemset(b2, 0, ns * 4));
    HIPCHK(c, hipMemset(d, 0, ns * 4));
    S.sn = a;
    S.s = b2;
    S.d = d;
  } else {
    double* a;
    if (int s = dalloc(c, &a, ns)) return s;
    HIPCHK(c, hipMemset(a, 0, ns * 8));
    S.us = a;
  }
  if (int s = dalloc(c, &S.rout, ns)) return s;
  char* fl;
  if (int s = dalloc(c, &fl, (size_t)ned * ne * 4 * eb)) return s;
  S.flux = fl;
  if (int s = dalloc(c, &S.lim_list, (size_t)nel)) return s;
  if (int s = dalloc(c, &S.lim_w, (size_t)nel * 4 * np)) return s;
  if (int s = dalloc(c, &S.lim_flags, (size_t)nel)) return s;
  if (int s = dalloc(c, &S.ctl, 1)) return s;
  if (int s = dalloc(c, &c->hbuf, nstate(c))) return s;
  if (int s = dalloc(c, &c->d_force, 2)) return s;
  if (int s = dalloc(c, &c->d_flag, 1)) return s;
  std::vector<int> lv(M.level_elems.begin(), M.level_elems.end());
  if (int s = upload(c, &c->d_levels, lv)) return s;
  S.levels = c->d_levels;
  S.level_offsets.assign(M.level_offsets.begin(), M.level_offsets.end());
  S.limiter = c->prm.limiter;
  S.hcr = c->prm.hcr;
  {
    const char* v = std::getenv("DGM_ELEMENT_KERNEL");  // "valu" selects the one-lane-per-element kernel
    S.use_mfma = (v && std::strcmp(v, "valu") == 0) ? 0 : 1;
    const char* r = std::getenv("DGM_DETECT_REMAP");
    S.remap_detect = (r && r[0] == '0') ? 0 : 1;
  }
  if (c->prm.limiter != DGM_LIMITER_OFF) {
    if (int s = dalloc(c, &S.smat, (size_t)nel * c->solver->smoothness_len())) return s;
    c->solver->build_smoothness(S, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  S.weno_grid = 2048;
  S.apply_grid = 1024;
  return write_ctl(c, fresh_ctl(c));
}

int create(const dgm_mesh* mesh, const dgm_params* params, dgm_ctx** out, bool device) {
  if (!mesh || !params || !out) return DGM_BAD_ARGUMENT;
  *out = nullptr;
  std::unique_ptr<dgm_ctx> c(new dgm_ctx());
  c->prm = *params;
  dgm_params& P = c->prm;
  if (P.order < 1 || P.order > 4) return DGM_BAD_CONFIG;
  if (P.n_cell_gp <= 0) P.n_cell_gp = default_cell_rule(P.order);
  if (P.n_edge_gp <= 0) P.n_edge_gp = default_edge_rule(P.order);
  if (P.precision != DGM_FP64 && P.precision != DGM_FP32_BULK) return DGM_BAD_CONFIG;
  if (P.limiter < DGM_LIMITER_OFF || P.limiter > DGM_LIMITER_JACOBI) return DGM_BAD_CONFIG;
  if (!(P.hcr > 1.0)) return DGM_BAD_CONFIG;
  if (!build_tables(c->T, P.order, P.n_cell_gp, P.n_edge_gp)) return DGM_BAD_CONFIG;
  c->solver = make_solver(P.precision, P.order, P.n_cell_gp, P.n_edge_gp);
  if (!c->solver) return DGM_BAD_CONFIG;
  std::string err;
  if (int s = build_mesh(*mesh, c->M, err)) return s;
  if (c->M.ne >= (1 << 30) || c->M.nedge >= (1 << 29)) return DGM_BAD_MESH;
  if (P.cfl > 0.0) {
    c->cfl = P.cfl;
  } else {
    const double p = (double)P.order;  // src/initialization.rs:128-132
    c->cfl = 1.0 / ((2.0 * p + 1.0) * (1.0 + 4.0 / ((p + 2.0) * (p + 2.0))));
  }
  if (!device) {
    *out = c.release();
    return DGM_OK;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return DGM_HIP_ERROR;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return DGM_HIP_ERROR;
  if (int s = build_device(c.get())) {
    dgm_ctx_destroy(c.release());
    return s;
  }
  *out = c.release();
  return DGM_OK;
}

}  // namespace

extern "C" {

int dgm_ctx_create(const dgm_mesh* mesh, const dgm_params* params, dgm_ctx** out) {
  return create(mesh, params, out, true);
}

int dgm_ctx_create_host(const dgm_mesh* mesh, const dgm_params* params, dgm_ctx** out) {
  return create(mesh, params, out, false);
}

void dgm_ctx_destroy(dgm_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int dgm_ctx_info(const dgm_ctx* c, dgm_info* info) {
  if (!c || !info) return DGM_BAD_ARGUMENT;
  info->n_elements = c->M.ne;
  info->n_edges = c->M.nedge;
  info->n_internal_edges = (int64_t)c->M.internal_edges.size();
  info->n_internal_elements = (int64_t)c->M.internal_elements.size();
  info->n_boundary_elements = (int64_t)c->M.boundary_elements.size();
  info->n_basis = c->T.np;
  info->n_cell_gp = c->T.nq;
  info->n_edge_gp = c->T.ne;
  info->order = c->prm.order;
  info->precision = c->prm.precision;
  info->limiter = c->prm.limiter;
  info->n_limiter_levels = (int32_t)c->M.level_offsets.size() - 1;
  info->cfl = c->cfl;
  info->device_bytes = c->bytes;
  return DGM_OK;
}

int dgm_last_error(const dgm_ctx* c, dgm_error* e) {
  if (!c || !e) return DGM_BAD_ARGUMENT;
  *e = c->err;
  return DGM_OK;
}

int dgm_get_connectivity(const dgm_ctx* c, int64_t* edge_elements, int64_t* edge_vertices, int64_t* edge_in_cell,
                         int64_t* edge_patch, int64_t* element_vertices, int64_t* element_edges,
                         int64_t* element_neighbours, int64_t* internal_edges, int64_t* internal_elements,
                         int64_t* boundary_elements) {
  if (!c) return DGM_BAD_ARGUMENT;
  const HostMesh& M = c->M;
  auto cp = [](int64_t* dst, const std::vector<int64_t>& v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * 8);
  };
  cp(edge_elements, M.ed_elem);
  cp(edge_vertices, M.ed_vert);
  cp(edge_in_cell, M.ed_lic);
  cp(edge_patch, M.ed_patch);
  cp(element_vertices, M.el_vert);
  cp(element_edges, M.el_edge);
  cp(element_neighbours, M.el_nbr);
  cp(internal_edges, M.internal_edges);
  cp(internal_elements, M.internal_elements);
  cp(boundary_elements, M.boundary_elements);
  return DGM_OK;
}

int dgm_get_geometry(const dgm_ctx* c, double* edge_normal, double* edge_length, double* element_jacob,
                     double* element_circumradius, double* element_minimum_height) {
  if (!c) return DGM_BAD_ARGUMENT;
  const HostMesh& M = c->M;
  auto cp = [](double* dst, const std::vector<double>& v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * 8);
  };
  cp(edge_normal, M.ed_normal);
  cp(edge_length, M.ed_len);
  cp(element_jacob, M.el_J);
  cp(element_circumradius, M.el_cr);
  cp(element_minimum_height, M.el_hmin);
  return DGM_OK;
}

int dgm_get_tables(const dgm_ctx* c, double* phi_cell, double* phi_edge, double* mass, double* cell_points,
                   double* cell_weights, double* edge_points, double* edge_weights) {
  if (!c) return DGM_BAD_ARGUMENT;
  const Tables& T = c->T;
  auto cp = [](double* dst, const std::vector<double>& v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * 8);
  };
  cp(phi_cell, T.phi_cell);
  cp(phi_edge, T.phi_edge);
  cp(mass, T.mass);
  cp(cell_points, T.cell_pts);
  cp(cell_weights, T.cell_w);
  cp(edge_points, T.edge_pts);
  cp(edge_weights, T.edge_w);
  return DGM_OK;
}

// Solver::set_initial_solution, src/solver.rs:45-79
int dgm_initial_solution(dgm_ctx* c, double density, double aoa_deg, double mach, double pressure, double* U) {
  if (!c || !U) return DGM_BAD_ARGUMENT;
  const double hcr = c->prm.hcr;
  const double a = std::sqrt(hcr * pressure / density);
  const double rad = aoa_deg * (M_PI / 180.0);
  const double u = mach * std::cos(rad) * a, v = mach * std::sin(rad) * a;
  const double q[4] = {density, density * u, density * v, pressure / (hcr - 1.0) + 0.5 * density * (u * u + v * v)};
  const Tables& T = c->T;
  for (int64_t e = 0; e < c->M.ne; ++e) {
    const double hj = 0.5 * c->M.el_J[e];
    for (int var = 0; var < 4; ++var)
      for (int k = 0; k < T.np; ++k) {
        double s = 0.0;
        for (int i = 0; i < T.nq; ++i) s += hj * T.cell_w[i] * q[var] * T.phi_cell[i * T.np + k];
        U[(e * 4 + var) * T.np + k] = s / (T.mass[k] * hj);
      }
  }
  return DGM_OK;
}

int dgm_project(dgm_ctx* c, const double* qp, double* U) {
  if (!c || !qp || !U) return DGM_BAD_ARGUMENT;
  const Tables& T = c->T;
  for (int64_t e = 0; e < c->M.ne; ++e)
    for (int var = 0; var < 4; ++var)
      for (int k = 0; k < T.np; ++k) {
        double s = 0.0;
        for (int i = 0; i < T.nq; ++i) s += T.cell_w[i] * qp[(e * T.nq + i) * 4 + var] * T.phi_cell[i * T.np + k];
        U[(e * 4 + var) * T.np + k] = s / T.mass[k];
      }
  return DGM_OK;
}

int dgm_cell_points_xy(const dgm_ctx* c, double* xy) {
  if (!c || !xy) return DGM_BAD_ARGUMENT;
  const Tables& T = c->T;
  const HostMesh& M = c->M;
  for (int64_t e = 0; e < M.ne; ++e) {
    double x[3], y[3];
    for (int i = 0; i < 3; ++i) {
      if (!M.exy.empty()) {
        x[i] = M.exy[(e * 3 + i) * 2];
        y[i] = M.exy[(e * 3 + i) * 2 + 1];
      } else {
        x[i] = M.vx[M.el_vert[3 * e + i]];
        y[i] = M.vy[M.el_vert[3 * e + i]];
      }
    }
    for (int q = 0; q < T.nq; ++q) {
      const double xi = T.cell_pts[2 * q], eta = T.cell_pts[2 * q + 1];
      xy[(e * T.nq + q) * 2] = x[0] + (x[1] - x[0]) * xi + (x[2] - x[0]) * eta;
      xy[(e * T.nq + q) * 2 + 1] = y[0] + (y[1] - y[0]) * xi + (y[2] - y[0]) * eta;
    }
  }
  return DGM_OK;
}

int dgm_state_upload(dgm_ctx* c, const double* U) {
  if (!c || !U) return DGM_BAD_ARGUMENT;
  if (int s = stage_in(c, U)) return s;
  if (int s = start_run(c, 0x7fffffff, 1.0e300)) return s;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DGM_OK;
}

int dgm_state_download(dgm_ctx* c, double* U) {
  if (!c || !U) return DGM_BAD_ARGUMENT;
  return stage_out(c, c->S.un, U);
}

int dgm_compute_residuals(dgm_ctx* c, double* R, const double* U, int32_t accumulate) {
  if (!c || !R || !U) return DGM_BAD_ARGUMENT;
  std::vector<double> rin;
  if (accumulate) rin.assign(R, R + nstate(c));
  if (int s = stage_in(c, U)) return s;
  Ctl h = fresh_ctl(c);
  if (int s = write_ctl(c, h)) return s;
  c->solver->residual(c->S, c->stream);
  HIPCHK(c, hipGetLastError());
  if (int s = stage_out(c, c->S.rout, R)) return s;
  if (int s = read_ctl(c, h)) return s;
  if (int s = ctl_status(c, h)) return s;
  if (accumulate) {
    const int np = c->T.np;
    for (int64_t e = 0; e < c->M.ne; ++e)
      for (int v = 0; v < 4; ++v)
        for (int k = 0; k < np; ++k) {
          const size_t i = (e * 4 + v) * np + k;
          R[i] += rin[i] / (c->T.mass[k] * 0.5 * c->M.el_J[e]);
        }
  }
  return DGM_OK;
}

int dgm_apply_limiter(dgm_ctx* c, double* U, uint8_t* troubled) {
  if (!c || !U) return DGM_BAD_ARGUMENT;
  if (int s = stage_in(c, U)) return s;
  if (int s = write_ctl(c, fresh_ctl(c))) return s;
  if (troubled) HIPCHK(c, hipMemsetAsync(c->S.lim_flags, 0, c->M.ne, c->stream));
  c->solver->limiter(c->S, c->stream);
  HIPCHK(c, hipGetLastError());
  if (int s = stage_out(c, c->S.un, U)) return s;
  if (troubled) HIPCHK(c, hipMemcpy(troubled, c->S.lim_flags, c->M.ne, hipMemcpyDeviceToHost));
  Ctl h;
  if (int s = read_ctl(c, h)) return s;
  return ctl_status(c, h);
}

int dgm_compute_time_step(dgm_ctx* c, const double* U, double* dt) {
  if (!c || !U || !dt) return DGM_BAD_ARGUMENT;
  if (int s = stage_in(c, U)) return s;
  Ctl h = fresh_ctl(c);
  if (int s = write_ctl(c, h)) return s;
  c->solver->time_step(c->S, c->stream);
  HIPCHK(c, hipGetLastError());
  if (int s = read_ctl(c, h)) return s;
  double m = 1.0e10;
  for (int i = 0; i < DT_LANES; ++i) {
    double v;
    std::memcpy(&v, &h.dtmin_bits[0][i * DT_STRIDE], 8);
    m = std::fmin(m, v);
  }
  *dt = m;
  return DGM_OK;
}

int dgm_rk3(dgm_ctx* c, double* U, double dt) {
  if (!c || !U) return DGM_BAD_ARGUMENT;
  if (int s = stage_in(c, U)) return s;
  Ctl h = fresh_ctl(c);
  h.dt = dt;
  if (int s = write_ctl(c, h)) return s;
  for (int st = 0; st < 3; ++st) c->solver->stage(c->S, st, c->stream);
  HIPCHK(c, hipGetLastError());
  if (int s = read_ctl(c, h)) return s;
  if (int s = ctl_status(c, h)) return s;
  return stage_out(c, c->S.un, U);
}

int dgm_device_time_march(dgm_ctx* c, int32_t final_step, double final_time, dgm_march_stats* stats) {
  if (!c) return DGM_BAD_ARGUMENT;
  if (int s = ensure_graph(c)) return s;
  if (int s = start_run(c, final_step, final_time)) return s;
  const int batch = 16;
  Ctl h;
  double t = 0.0;
  int32_t steps = 0;
  while (true) {
    for (int i = 0; i < batch; ++i) HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
    if (int s = read_ctl(c, h)) return s;
    effective_clock(h, t, steps);
    if (h.error != ~0ull || !h.active || steps >= final_step || !(t < final_time)) break;
  }
  if (stats) {
    stats->steps = steps;
    stats->time = t;
    stats->last_dt = h.dt;
    stats->troubled_cells_last_stage = h.pending ? h.count[2] : h.last_count;
  }
  return ctl_status(c, h);
}

int dgm_time_march(dgm_ctx* c, double* U, int32_t final_step, double final_time, dgm_march_stats* stats) {
  if (!c || !U) return DGM_BAD_ARGUMENT;
  if (int s = stage_in(c, U)) return s;
  const int st = dgm_device_time_march(c, final_step, final_time, stats);
  if (int s = stage_out(c, c->S.un, U)) return s;
  return st;
}

int dgm_device_steps_async(dgm_ctx* c, int32_t n) {
  if (!c) return DGM_BAD_ARGUMENT;
  if (int s = ensure_graph(c)) return s;
  for (int i = 0; i < n; ++i) HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
  return DGM_OK;
}

int dgm_synchronize(dgm_ctx* c) {
  if (!c) return DGM_BAD_ARGUMENT;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  Ctl h;
  if (int s = read_ctl(c, h)) return s;
  return ctl_status(c, h);
}

void* dgm_stream(dgm_ctx* c) { return c ? (void*)c->stream : nullptr; }

int dgm_check_state(dgm_ctx* c) {
  if (!c || !c->stream) return DGM_BAD_ARGUMENT;
  HIPCHK(c, hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
  c->solver->check_finite(c->S, c->d_flag, c->stream);
  int f = 0;
  HIPCHK(c, hipMemcpyAsync(&f, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return f ? fail(c, DGM_NONFINITE, "non-finite state") : DGM_OK;
}

int dgm_device_clock(dgm_ctx* c, double* t, int32_t* steps, double* last_dt) {
  if (!c) return DGM_BAD_ARGUMENT;
  Ctl h;
  if (int s = read_ctl(c, h)) return s;
  double te;
  int32_t se;
  effective_clock(h, te, se);
  if (t) *t = te;
  if (steps) *steps = se;
  if (last_dt) *last_dt = h.dt;
  return DGM_OK;
}

int dgm_profile_steps(dgm_ctx* c, int32_t n, dgm_kernel_times* out) {
  if (!c || !c->stream || !out || n < 1) return DGM_BAD_ARGUMENT;
  std::memset(out, 0, sizeof(*out));
  const int nev = 3 + 3 * Solver::N_PARTS;
  std::vector<hipEvent_t> ev(nev);
  for (auto& e : ev) HIPCHK(c, hipEventCreate(&e));
  double acc[Solver::N_PARTS] = {0, 0, 0, 0, 0};
  double t_dt = 0.0, t_ctl = 0.0, t_step = 0.0;
  for (int it = 0; it < n; ++it) {
    int k = 0;
    HIPCHK(c, hipEventRecord(ev[k++], c->stream));
    if (!fused_dt(c)) c->solver->time_step(c->S, c->stream);
    HIPCHK(c, hipEventRecord(ev[k++], c->stream));
    launch_step_begin(c->S.ctl, c->stream);
    HIPCHK(c, hipEventRecord(ev[k++], c->stream));
    for (int st = 0; st < 3; ++st)
      for (int part = 0; part < Solver::N_PARTS; ++part) {
        c->solver->stage_part(c->S, st, part, c->stream);
        HIPCHK(c, hipEventRecord(ev[k++], c->stream));
      }
    HIPCHK(c, hipEventSynchronize(ev[k - 1]));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, ev[0], ev[1]));
    t_dt += ms;
    HIPCHK(c, hipEventElapsedTime(&ms, ev[1], ev[2]));
    t_ctl += ms;
    int idx = 2;
    for (int st = 0; st < 3; ++st)
      for (int part = 0; part < Solver::N_PARTS; ++part, ++idx) {
        HIPCHK(c, hipEventElapsedTime(&ms, ev[idx], ev[idx + 1]));
        acc[part] += ms;
      }
    HIPCHK(c, hipEventElapsedTime(&ms, ev[0], ev[idx]));
    t_step += ms;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  out->time_step_ms = t_dt / n;
  out->control_ms = t_ctl / n;
  out->surface_ms = acc[0] / (3.0 * n);
  out->element_ms = acc[1] / (3.0 * n);
  out->detect_ms = acc[2] / (3.0 * n);
  out->rebuild_ms = acc[3] / (3.0 * n);
  out->apply_ms = acc[4] / (3.0 * n);
  out->step_ms = t_step / n;
  out->steps = n;
  Ctl h;
  if (int s = read_ctl(c, h)) return s;
  return ctl_status(c, h);
}

int dgm_troubled_counts(dgm_ctx* c, int32_t counts[4]) {
  if (!c || !counts) return DGM_BAD_ARGUMENT;
  Ctl h;
  if (int s = read_ctl(c, h)) return s;
  for (int i = 0; i < 4; ++i) counts[i] = h.count[i];
  return DGM_OK;
}

int dgm_lift_coefficient(dgm_ctx* c, double chord, double* cl) {
  if (!c || !c->stream || !cl || !(chord > 0.0)) return DGM_BAD_ARGUMENT;
  HIPCHK(c, hipMemsetAsync(c->d_force, 0, 2 * sizeof(double), c->stream));
  c->solver->wall_force(c->S, c->d_force, c->stream);
  double f[2];
  HIPCHK(c, hipMemcpyAsync(f, c->d_force, sizeof(f), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const dgm_params& P = c->prm;
  const double a = std::sqrt(P.fs_hcr * P.fs_p / P.fs_rho);
  const double V = P.fs_mach * a;
  const double rad = P.fs_aoa_deg * (M_PI / 180.0);
  *cl = (f[1] * std::cos(rad) - f[0] * std::sin(rad)) / (0.5 * P.fs_rho * V * V * chord);
  return DGM_OK;
}

}  // extern "C"
