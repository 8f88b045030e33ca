// C ABI of include/dgm.h: context set-up (tables, mesh, device upload), host-buffer entry points
// mirroring SpatialDisc / TemperalDisc, and the device-resident time march replayed from a
// captured HIP graph of one SSP-RK3 step.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "arena.h"
#include "dgm_internal.h"
#include "solver.h"

using namespace dgm;

// Working memory of the allocation-free residual (dgm_compute_residuals_device_noalloc): the
// structure-of-arrays state and residual and the edge point fluxes of one call. Every GPU the library
// runs on holds its own copy until the process ends; calls that overlap on one GPU share it and race.
// Host code never uses this name as an address: each device's address comes from
// hipGetSymbolAddress (scratch_address below) and is passed to the kernels as a plain pointer.
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct dgm_ctx {
  dgm_params prm{};
  Tables T;        // host copy for the host-side entry points
  HostMesh M;
  hipStream_t stream = nullptr;
  std::unique_ptr<Arena> A;  // device buffers: one arena holding this one case
  Placement pl;
  CaseCtl* d_case = nullptr;                // [1]
  unsigned long long* d_dtmin = nullptr;    // [DT_CASE]
  double* d_force = nullptr;                // [2]
  int* d_flag = nullptr;
  int* d_levels = nullptr;
  int* d_walk = nullptr;
  std::vector<void*> allocs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  // the same step captured GRAPH_STEPS times in one graph: one launch per GRAPH_STEPS steps
  hipGraph_t graph_n = nullptr;
  hipGraphExec_t graph_exec_n = nullptr;
  dgm_error err{};
  double cfl = 0.0;
  void* scratch = nullptr;   // this device's address of the scratch array
  CaseCtl* h_ctl = nullptr;  // pinned host copy of the control record (read-back without allocation)
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
  *p = (X*)v;
  return DGM_OK;
}

size_t nstate(const dgm_ctx* c) { return (size_t)c->M.ne * 4 * c->T.np; }
DeviceState& dev(dgm_ctx* c) { return c->A->S; }
const Solver& solver(dgm_ctx* c) { return *c->A->solver; }

// Device address of the scratch array on the current device, looked up once per device.
int scratch_address(void** out, std::string& err) {
  static std::mutex mu;
  static void* addr[256] = {};
  int d = 0;
  hipError_t e = hipGetDevice(&d);
  if (e != hipSuccess || d < 0 || d >= 256) {
    err = std::string("hipGetDevice: ") + hipGetErrorString(e);
    return DGM_HIP_ERROR;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[d]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return DGM_HIP_ERROR;
    }
    addr[d] = p;
  }
  *out = addr[d];
  return DGM_OK;
}

int read_ctl(dgm_ctx* c, CaseCtl& h) {
  if (!c->stream) return fail(c, DGM_BAD_ARGUMENT, "host-only context");
  HIPCHK(c, hipMemcpyAsync(&h, c->d_case, sizeof(CaseCtl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DGM_OK;
}
// writes the control record and empties both dt slots
int write_ctl(dgm_ctx* c, const CaseCtl& h) {
  if (!c->stream) return fail(c, DGM_BAD_ARGUMENT, "host-only context");
  const double big = 1.0e10;
  unsigned long long b;
  std::memcpy(&b, &big, 8);
  const std::vector<unsigned long long> empty(DT_CASE, b);
  HIPCHK(c, hipMemcpyAsync(c->d_case, &h, sizeof(CaseCtl), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_dtmin, empty.data(), DT_CASE * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DGM_OK;
}

CaseCtl fresh_ctl(const dgm_ctx* c) {
  CaseCtl h = fresh_case(c->cfl);
  h.e_off = 0;
  h.e_end = (int)c->M.ne;
  h.ed_off = 0;
  h.ed_end = (int)c->M.nedge;
  return h;
}

// status word of the device: lowest failing edge and its code
int ctl_status(dgm_ctx* c, const CaseCtl& h) {
  const unsigned long long err = case_error(h);
  if (err == ~0ull) return DGM_OK;
  const int code = (int)(err & 0xff);
  const int64_t edge = (int64_t)(err >> 8);
  if (code == 1) return fail(c, DGM_NEG_DENSITY, "Negative density found in HLLC!", edge);
  if (code == 2) return fail(c, DGM_NEG_PRESSURE, "Negative pressure found in HLLC!", edge);
  if (code == 5) return fail(c, DGM_BAD_MESH, "Invalid number of boundary neighbours!", -1, edge == 0xFFFFFF ? -1 : edge);
  return fail(c, DGM_HIP_ERROR, "unknown device status");
}

int stage_in(dgm_ctx* c, const double* U) {
  if (!c->stream) return fail(c, DGM_BAD_ARGUMENT, "host-only context");
  std::string e;
  if (int s = c->A->upload_state(U, c->pl, e)) return fail(c, s, e);
  return DGM_OK;
}
int stage_out(dgm_ctx* c, const double* soa, double* U) {
  std::string e;
  if (int s = c->A->download_state(soa, U, c->pl, e)) return fail(c, s, e);
  return DGM_OK;
}

// The Jacobi sweep's last stage reduces the next step's time step itself; other modes run the
// stand-alone reduction at the start of each step.
bool fused_dt(const dgm_ctx* c) { return c->prm.limiter == DGM_LIMITER_JACOBI; }

// the step control runs inside the stage-0 edge launch (one extra wavefront, beside the edges)
void enqueue_step(dgm_ctx* c) {
  if (!fused_dt(c)) solver(c).time_step(dev(c), c->stream, 0);
  for (int s = 0; s < 3; ++s) solver(c).stage(dev(c), s, c->stream, s == 0 ? 1 : 0);
}

// fresh control block for a resident run, with the first step's time step already reduced
int start_run(dgm_ctx* c, int32_t final_step, double final_time) {
  solver(c).flush(dev(c), c->stream);  // a resident state may still hold rebuilds of the last run
  CaseCtl h = fresh_ctl(c);
  h.final_step = final_step;
  h.final_time = final_time;
  if (int s = write_ctl(c, h)) return s;
  if (fused_dt(c)) solver(c).time_step(dev(c), c->stream, 0);
  HIPCHK(c, hipGetLastError());
  return DGM_OK;
}

// the step counted at the start of the next one (CaseCtl::pending) is already done
void effective_clock(const CaseCtl& h, double& t, int32_t& steps) {
  t = h.pending ? h.t + h.dt : h.t;
  steps = h.step + (h.pending ? 1 : 0);
}

// Steps per launch of the multi-step graph. Every step decides on the device whether it runs, so steps
// past a stop condition are no-ops and any number of steps may be queued.
constexpr int GRAPH_STEPS = 8;

void drop_graphs(dgm_ctx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->graph_exec_n) (void)hipGraphExecDestroy(c->graph_exec_n);
  if (c->graph_n) (void)hipGraphDestroy(c->graph_n);
  c->graph_exec = c->graph_exec_n = nullptr;
  c->graph = c->graph_n = nullptr;
}

int ensure_graph(dgm_ctx* c) {
  if (!c->stream) return fail(c, DGM_BAD_ARGUMENT, "host-only context");
  if (c->graph_exec) return DGM_OK;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  enqueue_step(c);
  HIPCHK(c, hipStreamEndCapture(c->stream, &c->graph));
  HIPCHK(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  for (int i = 0; i < GRAPH_STEPS; ++i) enqueue_step(c);
  HIPCHK(c, hipStreamEndCapture(c->stream, &c->graph_n));
  HIPCHK(c, hipGraphInstantiate(&c->graph_exec_n, c->graph_n, nullptr, nullptr, 0));
  // both graphs go to the device now, so that their first launches carry no one-time cost
  HIPCHK(c, hipGraphUpload(c->graph_exec, c->stream));
  HIPCHK(c, hipGraphUpload(c->graph_exec_n, c->stream));
  return DGM_OK;
}

// n steps from the captured graphs: whole multiples of GRAPH_STEPS from the multi-step graph
int launch_steps(dgm_ctx* c, int32_t n) {
  for (int32_t i = 0; i + GRAPH_STEPS <= n; i += GRAPH_STEPS) HIPCHK(c, hipGraphLaunch(c->graph_exec_n, c->stream));
  for (int32_t i = 0; i < n % GRAPH_STEPS; ++i) HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
  return DGM_OK;
}

int build_device(dgm_ctx* c) {
  const dgm_params& P = c->prm;
  std::string e;
  if (int s = dalloc(c, &c->d_case, 1)) return s;
  if (int s = dalloc(c, &c->d_dtmin, DT_CASE)) return s;
  if (int s = dalloc(c, &c->d_force, 2)) return s;
  if (int s = dalloc(c, &c->d_flag, 1)) return s;
  if (int s = scratch_address(&c->scratch, e)) return fail(c, s, e);
  HIPCHK(c, hipHostMalloc((void**)&c->h_ctl, sizeof(CaseCtl), hipHostMallocDefault));
  c->A.reset(new Arena());
  Arena& A = *c->A;
  if (int s = A.init(P.order, P.n_cell_gp, P.n_edge_gp, P.precision, P.limiter, c->M.ne, c->M.nedge,
                     (int64_t)c->M.patch_type.size(), true, c->d_case, c->d_dtmin, c->stream, e))
    return fail(c, s, e);
  A.S.mesh.hcr = P.hcr;
  {
    const char* v = std::getenv("DGM_ELEMENT_KERNEL");  // "valu" selects the one-lane-per-element kernel
    if (v && std::strcmp(v, "valu") == 0) A.S.use_mfma = 0;
    const char* r = std::getenv("DGM_REMAP");            // "0": plain block order in the gather kernels
    if (r && r[0] == '0') A.S.tune.remap = 0;
  }
  // the case record first (the smoothness build and the kernels read it), then the mesh
  if (int s = write_ctl(c, fresh_ctl(c))) return s;
  const std::vector<double> pq = farfield_states(c->M, P.fs_mach, P.fs_aoa_deg, P.fs_rho, P.fs_p, P.fs_hcr, P.hcr);
  if (int s = A.place(c->M, pq, 0, c->pl, e)) return fail(c, s, e.empty() ? "mesh does not fit its arena" : e);
  A.update_extent();
  A.S.mesh.single_case = 0;  // the one case owns every block of this arena
  std::vector<int> lv(c->M.level_elems.begin(), c->M.level_elems.end());
  if (int s = dalloc(c, &c->d_levels, lv.size())) return s;
  if (!lv.empty()) HIPCHK(c, hipMemcpy(c->d_levels, lv.data(), lv.size() * 4, hipMemcpyHostToDevice));
  A.S.levels = c->d_levels;
  A.S.level_offsets.assign(c->M.level_offsets.begin(), c->M.level_offsets.end());
  const char* wv = std::getenv("DGM_SEQUENTIAL_WALK");  // "0": one launch per level instead of the walk
  if (P.limiter == DGM_LIMITER_SEQUENTIAL && !(wv && wv[0] == '0')) {
    // walk records: {e, lower-numbered neighbours or -1} in level order, each level padded to 64 records
    std::vector<int> wk;
    for (size_t l = 0; l + 1 < c->M.level_offsets.size(); ++l) {
      for (int32_t q = c->M.level_offsets[l]; q < c->M.level_offsets[l + 1]; ++q) {
        const int el = c->M.level_elems[q];
        wk.push_back(el);
        for (int i = 0; i < 3; ++i) {
          const int64_t n = c->M.el_nbr[3 * (size_t)el + i];
          wk.push_back(n >= 0 && n < el ? (int)n : -1);
        }
      }
      while (wk.size() % 256) wk.push_back(-1);
    }
    if (int s = dalloc(c, &c->d_walk, wk.size())) return s;
    if (!wk.empty()) HIPCHK(c, hipMemcpy(c->d_walk, wk.data(), wk.size() * 4, hipMemcpyHostToDevice));
    A.S.walk = c->d_walk;
    A.S.walk_chunks = (int)(wk.size() / 256);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DGM_OK;
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
  if (!make_solver(P.precision, P.order, P.n_cell_gp, P.n_edge_gp)) return DGM_BAD_CONFIG;
  std::string err;
  if (int s = build_mesh(*mesh, c->M, err)) return s;
  if (c->M.ne >= (1 << 29) || c->M.nedge >= (1 << 28)) return DGM_BAD_MESH;
  c->cfl = default_cfl(P.order, P.cfl);
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
  drop_graphs(c);
  c->A.reset();
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->h_ctl) (void)hipHostFree(c->h_ctl);
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
  info->fused_pass = (c->A && c->A->S.fused && c->A->S.use_mfma && c->A->S.tune.element_waves > 0) ? 1 : 0;
  info->cfl = c->cfl;
  info->device_bytes = c->A ? c->A->bytes : 0;
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
  initial_state(c->T, c->M, density, aoa_deg, mach, pressure, c->prm.hcr, U);
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
  if (!c || !U || !c->stream) return DGM_BAD_ARGUMENT;
  return stage_out(c, dev(c).un, U);
}

int dgm_compute_residuals(dgm_ctx* c, double* R, const double* U, int32_t accumulate) {
  if (!c || !R || !U) return DGM_BAD_ARGUMENT;
  std::vector<double> rin;
  if (accumulate) rin.assign(R, R + nstate(c));
  if (int s = stage_in(c, U)) return s;
  CaseCtl h = fresh_ctl(c);
  if (int s = write_ctl(c, h)) return s;
  solver(c).residual(dev(c), c->stream);
  HIPCHK(c, hipGetLastError());
  if (int s = stage_out(c, dev(c).rout, R)) return s;
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

int dgm_compute_residuals_device_noalloc(dgm_ctx* c, const double* U, double* R) {
  if (!c || !U || !R) return DGM_BAD_ARGUMENT;
  if (!c->stream || !c->scratch || !c->h_ctl) return fail(c, DGM_BAD_ARGUMENT, "host-only context");
  // the call's working arrays, carved from the scratch array at 256-byte boundaries
  DeviceState S = dev(c);
  const bool f32 = solver(c).fp32();
  const size_t ns = (size_t)4 * c->T.np * (size_t)S.mesh.nes;
  const size_t nflux = (size_t)S.mesh.ed_stride * c->T.ne * 4;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = (off + bytes + 255) / 256 * 256;
    return static_cast<unsigned char*>(c->scratch) + at;
  };
  S.un = reinterpret_cast<double*>(take(ns * 8));
  S.rout = reinterpret_cast<double*>(take(ns * 8));
  S.sn = f32 ? take(ns * 4) : nullptr;
  S.flux = take(nflux * (f32 ? 4 : 8));
  if (off > kScratchBytes)
    return fail(c, DGM_NO_SPACE,
                "dgm_compute_residuals_device_noalloc: this mesh and order need " + std::to_string(off) +
                    " bytes of working memory; the scratch array holds " + std::to_string(kScratchBytes));
  S.prof_ev[0] = S.prof_ev[1] = nullptr;
  // Nothing in the scratch must start at zero: the state and the point fluxes are written in full
  // before they are read, and the residual kernel writes every element's row. The control record
  // (the error word the edge kernel reports into) is reset here by a kernel, not by a host copy.
  launch_set_case(c->d_case, fresh_ctl(c), c->stream);
  solver(c).upload(S, U, 0, (int)c->M.ne, c->stream);
  solver(c).residual(S, c->stream);
  solver(c).download(S, S.rout, R, 0, (int)c->M.ne, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->h_ctl, c->d_case, sizeof(CaseCtl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ctl_status(c, *c->h_ctl);
}

int dgm_apply_limiter(dgm_ctx* c, double* U, uint8_t* troubled) {
  if (!c || !U) return DGM_BAD_ARGUMENT;
  if (int s = stage_in(c, U)) return s;
  if (int s = write_ctl(c, fresh_ctl(c))) return s;
  if (troubled) HIPCHK(c, hipMemsetAsync(dev(c).lim_flags, 0, c->M.ne, c->stream));
  solver(c).limiter(dev(c), c->stream);
  HIPCHK(c, hipGetLastError());
  if (int s = stage_out(c, dev(c).un, U)) return s;
  if (troubled) {
    HIPCHK(c, hipMemcpyAsync(troubled, dev(c).lim_flags, c->M.ne, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  CaseCtl h;
  if (int s = read_ctl(c, h)) return s;
  return ctl_status(c, h);
}

int dgm_compute_time_step(dgm_ctx* c, const double* U, double* dt) {
  if (!c || !U || !dt) return DGM_BAD_ARGUMENT;
  if (int s = stage_in(c, U)) return s;
  CaseCtl h = fresh_ctl(c);
  if (int s = write_ctl(c, h)) return s;
  solver(c).time_step(dev(c), c->stream, 1);
  HIPCHK(c, hipGetLastError());
  std::vector<unsigned long long> slot(DT_SLOT);
  HIPCHK(c, hipMemcpyAsync(slot.data(), c->d_dtmin, DT_SLOT * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double m = 1.0e10;
  for (int i = 0; i < DT_LANES; ++i) {
    double v;
    std::memcpy(&v, &slot[i * DT_STRIDE], 8);
    m = std::fmin(m, v);
  }
  *dt = m;
  return DGM_OK;
}

int dgm_rk3(dgm_ctx* c, double* U, double dt) {
  if (!c || !U) return DGM_BAD_ARGUMENT;
  if (int s = stage_in(c, U)) return s;
  CaseCtl h = fresh_ctl(c);
  h.dt = dt;
  if (int s = write_ctl(c, h)) return s;
  for (int st = 0; st < 3; ++st) solver(c).stage(dev(c), st, c->stream);
  HIPCHK(c, hipGetLastError());
  if (int s = read_ctl(c, h)) return s;
  if (int s = ctl_status(c, h)) return s;
  return stage_out(c, dev(c).un, U);
}

int dgm_device_time_march(dgm_ctx* c, int32_t final_step, double final_time, dgm_march_stats* stats) {
  if (!c) return DGM_BAD_ARGUMENT;
  if (int s = ensure_graph(c)) return s;
  if (int s = start_run(c, final_step, final_time)) return s;
  const int batch = 2 * GRAPH_STEPS;
  CaseCtl h;
  double t = 0.0;
  int32_t steps = 0;
  while (true) {
    if (int s = launch_steps(c, batch)) return s;
    if (int s = read_ctl(c, h)) return s;
    effective_clock(h, t, steps);
    if (case_error(h) != ~0ull || !h.active || steps >= final_step || !(t < final_time)) break;
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
  if (int s = stage_out(c, dev(c).un, U)) return s;
  return st;
}

int dgm_device_steps_async(dgm_ctx* c, int32_t n) {
  if (!c) return DGM_BAD_ARGUMENT;
  if (int s = ensure_graph(c)) return s;
  return launch_steps(c, n);
}

int dgm_synchronize(dgm_ctx* c) {
  if (!c) return DGM_BAD_ARGUMENT;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  CaseCtl h;
  if (int s = read_ctl(c, h)) return s;
  return ctl_status(c, h);
}

void* dgm_stream(dgm_ctx* c) { return c ? (void*)c->stream : nullptr; }

int dgm_check_state(dgm_ctx* c) {
  if (!c || !c->stream) return DGM_BAD_ARGUMENT;
  HIPCHK(c, hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
  solver(c).flush(dev(c), c->stream);
  solver(c).check_finite(dev(c), c->d_flag, 0, (int)c->M.ne, c->stream);
  int f = 0;
  HIPCHK(c, hipMemcpyAsync(&f, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return f ? fail(c, DGM_NONFINITE, "non-finite state") : DGM_OK;
}

int dgm_device_clock(dgm_ctx* c, double* t, int32_t* steps, double* last_dt) {
  if (!c) return DGM_BAD_ARGUMENT;
  CaseCtl h;
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
  constexpr int NP_ = Solver::N_PARTS;
  // per part of each stage: a start and a stop event; then stream events: step start, after the
  // time-step reduction, step end (ctl_ev[2] is unused)
  std::vector<hipEvent_t> ev(2 * 3 * NP_ + 4);
  for (auto& e : ev) HIPCHK(c, hipEventCreate(&e));
  hipEvent_t* ctl_ev = &ev[2 * 3 * NP_];
  DeviceState& S = dev(c);
  const int lim = c->prm.limiter;
  // parts that launch nothing count zero: the Jacobi march defers its write-back, the element-order
  // sweep is all in its detect part, and without a limiter there are no limiter parts
  auto launches = [lim](int part) {
    if (part <= Solver::PART_ELEMENT) return true;
    if (lim == DGM_LIMITER_JACOBI) return part != Solver::PART_APPLY;
    if (lim == DGM_LIMITER_SEQUENTIAL) return part == Solver::PART_DETECT;
    return false;
  };
  // one launch per part (surface, element, the Jacobi detector and rebuild): the dispatch records its
  // own start and end; the element-order sweep (several launches) is bracketed by stream events
  auto single = [lim](int part) { return part <= Solver::PART_ELEMENT || lim == DGM_LIMITER_JACOBI; };
  double acc[NP_] = {0, 0, 0, 0, 0};
  double t_dt = 0.0, t_step = 0.0;
  for (int it = 0; it < n; ++it) {
    HIPCHK(c, hipEventRecord(ctl_ev[0], c->stream));
    if (!fused_dt(c)) solver(c).time_step(S, c->stream, 0);
    HIPCHK(c, hipEventRecord(ctl_ev[1], c->stream));
    // the step control runs inside the stage-0 edge launch, as in a march: no launch of its own
    for (int st = 0; st < 3; ++st)
      for (int part = 0; part < NP_; ++part) {
        if (!launches(part)) continue;
        hipEvent_t* pe = &ev[2 * (st * NP_ + part)];
        const int nctl = (st == 0 && part == Solver::PART_SURFACE) ? 1 : 0;
        if (single(part)) {
          S.prof_ev[0] = pe[0];
          S.prof_ev[1] = pe[1];
          solver(c).stage_part(S, st, part, c->stream, nctl);
          S.prof_ev[0] = S.prof_ev[1] = nullptr;
        } else {
          HIPCHK(c, hipEventRecord(pe[0], c->stream));
          solver(c).stage_part(S, st, part, c->stream, nctl);
          HIPCHK(c, hipEventRecord(pe[1], c->stream));
        }
      }
    HIPCHK(c, hipEventRecord(ctl_ev[3], c->stream));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    if (!fused_dt(c)) {  // with the Jacobi limiter the reduction is part of the last detector pass
      HIPCHK(c, hipEventElapsedTime(&ms, ctl_ev[0], ctl_ev[1]));
      t_dt += ms;
    }
    for (int st = 0; st < 3; ++st)
      for (int part = 0; part < NP_; ++part) {
        if (!launches(part)) continue;
        hipEvent_t* pe = &ev[2 * (st * NP_ + part)];
        HIPCHK(c, hipEventElapsedTime(&ms, pe[0], pe[1]));
        acc[part] += ms;
      }
    HIPCHK(c, hipEventElapsedTime(&ms, ctl_ev[0], ctl_ev[3]));
    t_step += ms;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  out->time_step_ms = t_dt / n;
  out->control_ms = 0.0;
  out->surface_ms = acc[0] / (3.0 * n);
  out->element_ms = acc[1] / (3.0 * n);
  out->detect_ms = acc[2] / (3.0 * n);
  out->rebuild_ms = acc[3] / (3.0 * n);
  out->apply_ms = acc[4] / (3.0 * n);
  out->step_ms = t_step / n;
  out->steps = n;
  CaseCtl h;
  if (int s = read_ctl(c, h)) return s;
  return ctl_status(c, h);
}

int dgm_troubled_counts(dgm_ctx* c, int32_t counts[4]) {
  if (!c || !counts) return DGM_BAD_ARGUMENT;
  CaseCtl h;
  if (int s = read_ctl(c, h)) return s;
  for (int i = 0; i < 4; ++i) counts[i] = h.count[i];
  return DGM_OK;
}

int dgm_lift_coefficient(dgm_ctx* c, double chord, double* cl) {
  if (!c || !c->stream || !cl || !(chord > 0.0)) return DGM_BAD_ARGUMENT;
  HIPCHK(c, hipMemsetAsync(c->d_force, 0, 2 * sizeof(double), c->stream));
  solver(c).flush(dev(c), c->stream);
  solver(c).wall_force(dev(c), c->d_force, 0, (int)c->M.nedge, c->stream);
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

int dgm_get_tuning(const dgm_ctx* c, dgm_tuning* t) {
  if (!c || !t || !c->A) return DGM_BAD_ARGUMENT;
  *t = from_tuning(c->A->S.tune);
  return DGM_OK;
}

int dgm_set_tuning(dgm_ctx* c, const dgm_tuning* t) {
  if (!c || !t || !c->A || !valid_tuning(*t)) return DGM_BAD_ARGUMENT;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->A->S.tune = to_tuning(*t);
  drop_graphs(c);
  return DGM_OK;
}

int dgm_autotune(dgm_ctx* c, int32_t reps, dgm_tuning* chosen) {
  if (!c || !c->A || reps < 1) return DGM_BAD_ARGUMENT;
  std::string e;
  if (int s = autotune_steps({c->A.get()}, c->d_case, c->d_dtmin, 1, c->stream, [c]() { enqueue_step(c); }, reps, e))
    return fail(c, s, e);
  drop_graphs(c);
  if (chosen) *chosen = from_tuning(c->A->S.tune);
  CaseCtl h;
  if (int s = read_ctl(c, h)) return s;
  return ctl_status(c, h);
}

}  // extern "C"
