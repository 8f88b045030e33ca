// Batches of cases (include/dgm.h, dgm_batch_*): one arena per order, cases added and removed between
// steps, one captured graph per step covering every running case, per-case control records, lift,
// and an autotuner of the launch shapes of each order's kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "arena.h"
#include "dgm_internal.h"
#include "solver.h"

using namespace dgm;

namespace {

struct BatchCase {
  bool live = false;
  int order = 0;
  Placement pl;
  dgm_case_params cp{};
  double cfl = 0.0;
  int64_t footprint = 0;
};

// A removed case's slot and ranges, kept out of use until the batch stream has passed its removal (work
// queued before the removal still touches them): released once a poll queued after it has completed.
struct Retired {
  int64_t after_poll;  // released when the poll of this index (or a later one) has completed
  int slot;            // -1: none
  int order;           // 0: no ranges
  Placement pl;
  hipGraphExec_t exec;
  hipGraph_t graph;
};

}  // namespace

struct dgm_batch {
  dgm_batch_params prm{};
  hipStream_t stream = nullptr;
  CaseCtl* d_cases = nullptr;             // [max_cases]
  unsigned long long* d_dtmin = nullptr;  // [max_cases][DT_CASE]
  double* d_force = nullptr;              // [2 max_cases]
  CaseCtl* h_cases = nullptr;             // pinned mirror
  unsigned long long* h_dt = nullptr;     // pinned, one case's dt slots (all empty)
  std::vector<void*> allocs;
  std::unique_ptr<Arena> arena[4];
  std::vector<BatchCase> cases;
  // one stream per order: with more than one order present, a step runs the arenas' stage sequences
  // as parallel branches (fork and join through events on the batch stream)
  hipStream_t astream[4] = {};
  hipEvent_t fork_ev = nullptr, join_ev[4] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph_n = nullptr;           // the step captured GRAPH_STEPS times: one launch per GRAPH_STEPS steps
  hipGraphExec_t exec_n = nullptr;
  int64_t captured[4][2] = {};            // arena extents the graphs were captured with
  dgm_error err{};
  // Admissions: a new case's mesh, state and first time step go on ustream beside the running steps; its
  // record goes live on the batch stream, between steps, once those are done (admit_ev) and only with
  // launches that cover its ranges.
  hipStream_t ustream = nullptr;
  hipStream_t cstream = nullptr;  // graph capture only
  hipEvent_t admit_ev = nullptr;
  bool admit_pending = false;
  std::vector<int> activate;              // slots whose record the next batch-stream operation writes
  CaseCtl* h_act = nullptr;               // pinned, per slot: the record that goes live
  CaseCtl* h_empty = nullptr;             // pinned: the record of an empty slot
  std::vector<Retired> retired;
  std::vector<char> held;                 // per slot: removed, not yet released
  // serving polls: lift, convergence stop and a copy of every record, into one of two pinned buffers
  double* d_cpar = nullptr;               // [max_cases][4]: cos aoa, sin aoa, 1 / (rho V^2 / 2), unused
  double* d_clprev = nullptr;             // [max_cases] lift coefficient at the last poll
  int* d_conv = nullptr;                  // [max_cases] stopped by the convergence test
  double* d_cl = nullptr;                 // [max_cases]
  CaseCtl* h_poll_cases[2] = {};
  double* h_poll_cl[2] = {};
  int* h_poll_conv[2] = {};
  hipEvent_t poll_ev[2] = {};
  int64_t polls = 0, polls_done = 0;      // polls queued, and the lowest index not yet known complete
  // device time of the step launches: an event pair per dgm_batch_steps_async call
  std::vector<std::pair<hipEvent_t, hipEvent_t>> step_ev;
  double device_ms = 0.0;
};

namespace {

int fail(dgm_batch* b, int status, const std::string& msg, int64_t edge = -1) {
  if (b) {
    b->err.status = status;
    b->err.edge = edge;
    b->err.element = -1;
    std::snprintf(b->err.message, sizeof(b->err.message), "%s", msg.c_str());
  }
  return status;
}

#define BCHK(b, call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return fail((b), DGM_HIP_ERROR, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class X>
int balloc(dgm_batch* b, X** p, size_t n) {
  *p = nullptr;
  void* v = nullptr;
  const hipError_t e = hipMalloc(&v, std::max<size_t>(1, n) * sizeof(X));
  if (e != hipSuccess) return fail(b, DGM_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  b->allocs.push_back(v);
  *p = (X*)v;
  return DGM_OK;
}

Ctl shared_ctl(dgm_batch* b) { return Ctl{b->d_cases, b->d_dtmin, nullptr}; }

// Work on the batch stream that must see every admission so far: it waits for the admission stream, then
// the admitted records go live (from their pinned copies), ahead of whatever is queued next.
int flush_admissions(dgm_batch* b) {
  if (!b->admit_pending) return DGM_OK;
  BCHK(b, hipStreamWaitEvent(b->stream, b->admit_ev, 0));
  for (int id : b->activate)
    BCHK(b, hipMemcpyAsync(b->d_cases + id, &b->h_act[id], sizeof(CaseCtl), hipMemcpyHostToDevice, b->stream));
  b->activate.clear();
  b->admit_pending = false;
  return DGM_OK;
}

// Frees what removals and re-captures left behind once the batch stream has passed them: all of it after
// a full synchronisation (done = true), else what precedes the polls known complete.
void release_retired(dgm_batch* b, bool done) {
  std::vector<Retired> keep;
  for (const Retired& r : b->retired) {
    if (!done && r.after_poll >= b->polls_done) {
      keep.push_back(r);
      continue;
    }
    if (r.slot >= 0) b->held[r.slot] = 0;
    if (r.order > 0) {
      b->arena[r.order - 1]->release_ranges(r.pl);
      b->arena[r.order - 1]->update_extent();
    }
    if (r.exec) (void)hipGraphExecDestroy(r.exec);
    if (r.graph) (void)hipGraphDestroy(r.graph);
  }
  b->retired.swap(keep);
}

// adds the device time of every completed step launch pair to device_ms (all of them when the stream is idle)
void collect_step_times(dgm_batch* b, bool all) {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> keep;
  for (auto& e : b->step_ev) {
    if (!all && hipEventQuery(e.second) != hipSuccess) {
      keep.push_back(e);
      continue;
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) b->device_ms += ms;
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  b->step_ev.swap(keep);
}

int full_sync(dgm_batch* b) {
  if (int s = flush_admissions(b)) return s;
  BCHK(b, hipStreamSynchronize(b->ustream));
  BCHK(b, hipStreamSynchronize(b->stream));
  b->polls_done = b->polls;
  collect_step_times(b, true);
  release_retired(b, true);
  return DGM_OK;
}

// Steps per launch of the multi-step graph. Every step decides on the device, case by case, whether it
// runs, so steps past a case's stop condition are no-ops and any number of steps may be queued.
constexpr int GRAPH_STEPS = 8;

bool graph_current(const dgm_batch* b) {
  if (!b->exec) return false;
  for (int g = 0; g < 4; ++g) {
    const int64_t ne = b->arena[g] ? b->arena[g]->extent_elements() : 0;
    const int64_t ned = b->arena[g] ? b->arena[g]->extent_edges() : 0;
    if (ne != b->captured[g][0] || ned != b->captured[g][1]) return false;
  }
  return true;
}

// The current graph is no longer launched; launches of it may still be queued, so it is destroyed only
// once the batch stream has passed them.
void drop_graph(dgm_batch* b) {
  if (b->exec || b->graph) b->retired.push_back(Retired{b->polls, -1, 0, Placement(), b->exec, b->graph});
  if (b->exec_n || b->graph_n) b->retired.push_back(Retired{b->polls, -1, 0, Placement(), b->exec_n, b->graph_n});
  b->exec = b->exec_n = nullptr;
  b->graph = b->graph_n = nullptr;
}

// one step of every arena on stream st (the capture stream, or the batch stream when timed eagerly)
void enqueue_step(dgm_batch* b, hipStream_t st) {
  if (b->prm.limiter != DGM_LIMITER_JACOBI)
    for (int g = 0; g < 4; ++g)
      if (b->arena[g] && b->arena[g]->extent_elements() > 0) b->arena[g]->solver->time_step(b->arena[g]->S, st, 0);
  int present = 0;
  for (int g = 0; g < 4; ++g) present += (b->arena[g] && b->arena[g]->extent_elements() > 0) ? 1 : 0;
  if (present > 1) {
    // Arenas share nothing but the case records and dt slots, each touching only its own cases': after
    // the step control (its own launch here, since every arena's stage 0 needs it), each arena's three
    // stages run on the arena's stream, and the batch stream waits for all of them.
    launch_step_begin(shared_ctl(b), b->prm.max_cases, st);
    (void)hipEventRecord(b->fork_ev, st);
    for (int g = 0; g < 4; ++g) {
      if (!b->arena[g] || b->arena[g]->extent_elements() <= 0) continue;
      (void)hipStreamWaitEvent(b->astream[g], b->fork_ev, 0);
      for (int s = 0; s < 3; ++s) b->arena[g]->solver->stage(b->arena[g]->S, s, b->astream[g], 0);
      (void)hipEventRecord(b->join_ev[g], b->astream[g]);
    }
    for (int g = 0; g < 4; ++g)
      if (b->arena[g] && b->arena[g]->extent_elements() > 0) (void)hipStreamWaitEvent(st, b->join_ev[g], 0);
    return;
  }
  // one order: the step control of every case slot runs inside its stage-0 edge launch
  int nctl = b->prm.max_cases;
  for (int s = 0; s < 3; ++s)
    for (int g = 0; g < 4; ++g)
      if (b->arena[g] && b->arena[g]->extent_elements() > 0) {
        b->arena[g]->solver->stage(b->arena[g]->S, s, st, s == 0 ? nctl : 0);
        if (s == 0) nctl = 0;
      }
  if (nctl > 0) launch_step_begin(shared_ctl(b), nctl, st);  // no arena holds elements
}

int ensure_graph(dgm_batch* b) {
  if (graph_current(b)) return DGM_OK;
  drop_graph(b);
  // captured on a stream that never runs work, so the steps already queued on the batch stream go on
  BCHK(b, hipStreamBeginCapture(b->cstream, hipStreamCaptureModeThreadLocal));
  enqueue_step(b, b->cstream);
  BCHK(b, hipStreamEndCapture(b->cstream, &b->graph));
  BCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  BCHK(b, hipStreamBeginCapture(b->cstream, hipStreamCaptureModeThreadLocal));
  for (int i = 0; i < GRAPH_STEPS; ++i) enqueue_step(b, b->cstream);
  BCHK(b, hipStreamEndCapture(b->cstream, &b->graph_n));
  BCHK(b, hipGraphInstantiate(&b->exec_n, b->graph_n, nullptr, nullptr, 0));
  // both graphs go to the device now, so that their first launches carry no one-time cost
  BCHK(b, hipGraphUpload(b->exec, b->stream));
  BCHK(b, hipGraphUpload(b->exec_n, b->stream));
  for (int g = 0; g < 4; ++g) {
    b->captured[g][0] = b->arena[g] ? b->arena[g]->extent_elements() : 0;
    b->captured[g][1] = b->arena[g] ? b->arena[g]->extent_edges() : 0;
  }
  return DGM_OK;
}

int case_status_code(const CaseCtl& h, int64_t* edge) {
  *edge = -1;
  const unsigned long long err = case_error(h);
  if (err == ~0ull) return DGM_OK;
  const int code = (int)(err & 0xff);
  const int64_t e = (int64_t)(err >> 8);
  *edge = e == 0xFFFFFF ? -1 : e;
  if (code == 1) return DGM_NEG_DENSITY;
  if (code == 2) return DGM_NEG_PRESSURE;
  if (code == 5) return DGM_BAD_MESH;
  return DGM_HIP_ERROR;
}

}  // namespace

extern "C" {

int dgm_batch_create(const dgm_batch_params* params, dgm_batch** out) {
  if (!params || !out) return DGM_BAD_ARGUMENT;
  *out = nullptr;
  const dgm_batch_params& P = *params;
  if (P.precision != DGM_FP64 && P.precision != DGM_FP32_BULK) return DGM_BAD_CONFIG;
  if (P.limiter != DGM_LIMITER_OFF && P.limiter != DGM_LIMITER_JACOBI) return DGM_BAD_CONFIG;
  if (!(P.hcr > 1.0) || P.max_cases < 1 || P.max_cases > (1 << 20)) return DGM_BAD_CONFIG;
  int64_t total_cap = 0;
  for (int g = 0; g < 4; ++g) {
    if (P.capacity[g] < 0) return DGM_BAD_CONFIG;
    total_cap += P.capacity[g];
  }
  if (total_cap <= 0) return DGM_BAD_CONFIG;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return DGM_HIP_ERROR;
  std::unique_ptr<dgm_batch> b(new dgm_batch());
  b->prm = P;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return DGM_HIP_ERROR;
  const int nc = P.max_cases;
  int s = DGM_OK;
  auto cleanup = [&](int st) {
    dgm_batch_destroy(b.release());
    return st;
  };
  if (hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming) != hipSuccess) return cleanup(DGM_HIP_ERROR);
  for (int g = 0; g < 4; ++g) {
    if (P.capacity[g] <= 0) continue;
    if (hipStreamCreateWithFlags(&b->astream[g], hipStreamNonBlocking) != hipSuccess) return cleanup(DGM_HIP_ERROR);
    if (hipEventCreateWithFlags(&b->join_ev[g], hipEventDisableTiming) != hipSuccess) return cleanup(DGM_HIP_ERROR);
  }
  if ((s = balloc(b.get(), &b->d_cases, nc))) return cleanup(s);
  if ((s = balloc(b.get(), &b->d_dtmin, (size_t)nc * DT_CASE))) return cleanup(s);
  if ((s = balloc(b.get(), &b->d_force, 2 * (size_t)nc))) return cleanup(s);
  if (hipHostMalloc((void**)&b->h_cases, sizeof(CaseCtl) * nc, hipHostMallocDefault) != hipSuccess) return cleanup(DGM_OUT_OF_MEMORY);
  if (hipHostMalloc((void**)&b->h_dt, DT_CASE * 8, hipHostMallocDefault) != hipSuccess) return cleanup(DGM_OUT_OF_MEMORY);
  if (hipStreamCreateWithFlags(&b->ustream, hipStreamNonBlocking) != hipSuccess) return cleanup(DGM_HIP_ERROR);
  if (hipStreamCreateWithFlags(&b->cstream, hipStreamNonBlocking) != hipSuccess) return cleanup(DGM_HIP_ERROR);
  if (hipEventCreateWithFlags(&b->admit_ev, hipEventDisableTiming) != hipSuccess) return cleanup(DGM_HIP_ERROR);
  if (hipHostMalloc((void**)&b->h_act, sizeof(CaseCtl) * nc, hipHostMallocDefault) != hipSuccess) return cleanup(DGM_OUT_OF_MEMORY);
  if (hipHostMalloc((void**)&b->h_empty, sizeof(CaseCtl), hipHostMallocDefault) != hipSuccess) return cleanup(DGM_OUT_OF_MEMORY);
  std::memset(b->h_act, 0, sizeof(CaseCtl) * nc);
  std::memset(b->h_empty, 0, sizeof(CaseCtl));
  b->h_empty->error = b->h_empty->error_s0 = ~0ull;
  for (int f = 0; f < 3; ++f) b->h_empty->error_f[f] = ~0ull;
  if ((s = balloc(b.get(), &b->d_cpar, 4 * (size_t)nc))) return cleanup(s);
  if ((s = balloc(b.get(), &b->d_clprev, nc))) return cleanup(s);
  if ((s = balloc(b.get(), &b->d_conv, nc))) return cleanup(s);
  if ((s = balloc(b.get(), &b->d_cl, nc))) return cleanup(s);
  for (int k = 0; k < 2; ++k) {
    if (hipHostMalloc((void**)&b->h_poll_cases[k], sizeof(CaseCtl) * nc, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->h_poll_cl[k], sizeof(double) * nc, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->h_poll_conv[k], sizeof(int) * nc, hipHostMallocDefault) != hipSuccess)
      return cleanup(DGM_OUT_OF_MEMORY);
    if (hipEventCreateWithFlags(&b->poll_ev[k], hipEventDisableTiming) != hipSuccess) return cleanup(DGM_HIP_ERROR);
  }
  b->held.assign(nc, 0);
  std::memset(b->h_cases, 0, sizeof(CaseCtl) * nc);
  {
    const double big = 1.0e10;
    unsigned long long bits;
    std::memcpy(&bits, &big, 8);
    for (int i = 0; i < DT_CASE; ++i) b->h_dt[i] = bits;
    std::vector<unsigned long long> all((size_t)nc * DT_CASE, bits);
    if (hipMemcpy(b->d_dtmin, all.data(), all.size() * 8, hipMemcpyHostToDevice) != hipSuccess) return cleanup(DGM_HIP_ERROR);
    if (hipMemcpy(b->d_cases, b->h_cases, sizeof(CaseCtl) * nc, hipMemcpyHostToDevice) != hipSuccess) return cleanup(DGM_HIP_ERROR);
  }
  for (int g = 0; g < 4; ++g) {
    if (P.capacity[g] <= 0) continue;
    const int order = g + 1;
    const int64_t cap_e = P.capacity[g];
    // edges: an O-grid or periodic triangulation has 1.5 edges per triangle plus a boundary row; each
    // case also rounds its range up to 64
    const int64_t cap_ed = cap_e * 8 / 5 + 64 * (int64_t)std::min(nc, 4096);
    b->arena[g].reset(new Arena());
    std::string e;
    if ((s = b->arena[g]->init(order, default_cell_rule(order), default_edge_rule(order), P.precision, P.limiter, cap_e,
                               cap_ed, 8 * (int64_t)nc, false, b->d_cases, b->d_dtmin, b->stream, e))) {
      std::fprintf(stderr, "dgm_batch_create: %s\n", e.c_str());
      return cleanup(s);
    }
    b->arena[g]->S.mesh.hcr = P.hcr;
    b->arena[g]->ustream = b->ustream;
  }
  b->cases.resize(nc);
  *out = b.release();
  return DGM_OK;
}

void dgm_batch_destroy(dgm_batch* b) {
  if (!b) return;
  if (b->ustream) (void)hipStreamSynchronize(b->ustream);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto st : b->astream)
    if (st) (void)hipStreamSynchronize(st);
  drop_graph(b);
  b->polls_done = b->polls;
  if (!b->held.empty()) release_retired(b, true);
  collect_step_times(b, true);
  for (auto& a : b->arena) a.reset();
  for (void* p : b->allocs) (void)hipFree(p);
  if (b->h_cases) (void)hipHostFree(b->h_cases);
  if (b->h_dt) (void)hipHostFree(b->h_dt);
  if (b->h_act) (void)hipHostFree(b->h_act);
  if (b->h_empty) (void)hipHostFree(b->h_empty);
  for (int k = 0; k < 2; ++k) {
    if (b->h_poll_cases[k]) (void)hipHostFree(b->h_poll_cases[k]);
    if (b->h_poll_cl[k]) (void)hipHostFree(b->h_poll_cl[k]);
    if (b->h_poll_conv[k]) (void)hipHostFree(b->h_poll_conv[k]);
    if (b->poll_ev[k]) (void)hipEventDestroy(b->poll_ev[k]);
  }
  if (b->admit_ev) (void)hipEventDestroy(b->admit_ev);
  if (b->ustream) (void)hipStreamDestroy(b->ustream);
  if (b->cstream) (void)hipStreamDestroy(b->cstream);
  for (int g = 0; g < 4; ++g) {
    if (b->astream[g]) (void)hipStreamDestroy(b->astream[g]);
    if (b->join_ev[g]) (void)hipEventDestroy(b->join_ev[g]);
  }
  if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int dgm_batch_last_error(const dgm_batch* b, dgm_error* e) {
  if (!b || !e) return DGM_BAD_ARGUMENT;
  *e = b->err;
  return DGM_OK;
}

int dgm_batch_footprint(const dgm_batch* b, int32_t order, int64_t ne, int64_t ned, int64_t npatch, int64_t* bytes) {
  if (!b || !bytes || order < 1 || order > 4) return DGM_BAD_ARGUMENT;
  *bytes = Arena::footprint(order, default_cell_rule(order), default_edge_rule(order), b->prm.precision,
                            b->prm.limiter, false)
               .of(ne, ned, npatch);
  return DGM_OK;
}

int dgm_batch_add(dgm_batch* b, const dgm_mesh* mesh, const dgm_case_params* cp, const double* U0, int32_t* case_id) {
  if (!b || !mesh || !cp || !case_id) return DGM_BAD_ARGUMENT;
  *case_id = -1;
  if (cp->order < 1 || cp->order > 4 || !b->arena[cp->order - 1])
    return fail(b, DGM_BAD_CONFIG, "the batch has no arena for this order");
  if (!(cp->fs_rho > 0.0) || !(cp->fs_p > 0.0) || cp->final_step < 0)
    return fail(b, DGM_BAD_CONFIG, "bad case parameters");
  int slot = -1;
  for (int i = 0; i < b->prm.max_cases; ++i)
    if (!b->cases[i].live && !b->held[i]) {
      slot = i;
      break;
    }
  if (slot < 0) return fail(b, DGM_NO_SPACE, "no free case slot");
  Arena& A = *b->arena[cp->order - 1];
  HostMesh M;
  std::string err;
  if (int s = build_mesh(*mesh, M, err)) return fail(b, s, err);
  std::vector<double> U;
  if (!U0) {
    U.resize((size_t)M.ne * 4 * A.T.np);
    initial_state(A.T, M, cp->fs_rho, cp->fs_aoa_deg, cp->fs_mach, cp->fs_p, b->prm.hcr, U.data());
    U0 = U.data();
  }
  const std::vector<double> pq = farfield_states(M, cp->fs_mach, cp->fs_aoa_deg, cp->fs_rho, cp->fs_p, cp->fs_hcr,
                                                 b->prm.hcr);
  // Everything below runs on the admission stream, beside the steps already queued: the case's ranges are
  // free (ranges of removed cases come back only once the batch stream has passed the removal).
  Placement pl;
  if (int s = A.place(M, pq, slot, pl, err)) return fail(b, s, err.empty() ? "no free range in the arena" : err);
  A.update_extent();  // the kernels below (and every later graph) must cover the new range
  if (int s = A.upload_state(U0, pl, err)) return fail(b, s, err);
  const double cfl = default_cfl(cp->order, cp->cfl);
  CaseCtl h = fresh_case(cfl);
  h.e_off = (int)pl.e_off;
  h.e_end = (int)(pl.e_off + pl.ne);
  h.ed_off = (int)pl.ed_off;
  h.ed_end = (int)(pl.ed_off + pl.ned);
  h.group = cp->order - 1;
  h.final_step = cp->final_step;
  h.final_time = cp->final_time;
  // the record that goes live later, and for now the same record held inactive and not live (the first
  // time step below reads its CFL; kernels already queued skip it, and the step control ignores it)
  b->h_act[slot] = h;
  b->h_cases[slot] = h;
  b->h_cases[slot].live = 0;
  b->h_cases[slot].active = 0;
  BCHK(b, hipMemcpyAsync(b->d_cases + slot, &b->h_cases[slot], sizeof(CaseCtl), hipMemcpyHostToDevice, b->ustream));
  BCHK(b, hipMemcpyAsync(b->d_dtmin + (size_t)slot * DT_CASE, b->h_dt, DT_CASE * 8, hipMemcpyHostToDevice, b->ustream));
  // the first step's time step, of this case only (the other cases' slots hold their fused reductions)
  A.solver->time_step(A.S, b->ustream, 1, slot);
  BCHK(b, hipGetLastError());
  launch_case_poll_init(b->d_cpar, b->d_clprev, b->d_conv, slot, cp->fs_mach, cp->fs_aoa_deg, cp->fs_rho, cp->fs_p,
                        cp->fs_hcr, b->ustream);
  BCHK(b, hipEventRecord(b->admit_ev, b->ustream));
  // the host copies above must not be overwritten before they are made: h_cases is also the target of
  // dgm_batch_case_info, so wait for them (the admission stream holds only this case's work)
  BCHK(b, hipStreamSynchronize(b->ustream));
  b->admit_pending = true;
  b->activate.push_back(slot);
  BatchCase& bc = b->cases[slot];
  bc.live = true;
  bc.order = cp->order;
  bc.pl = pl;
  bc.cp = *cp;
  bc.cfl = cfl;
  int64_t fp = 0;
  dgm_batch_footprint(b, cp->order, pl.ne, pl.ned, pl.npatch, &fp);
  bc.footprint = fp;
  *case_id = slot;
  return DGM_OK;
}

int dgm_batch_remove(dgm_batch* b, int32_t id) {
  if (!b || id < 0 || id >= b->prm.max_cases || !b->cases[id].live) return DGM_BAD_ARGUMENT;
  BatchCase& bc = b->cases[id];
  // stream-ordered after the work already queued (which still sees the case), without waiting for it
  if (int s = flush_admissions(b)) return s;
  BCHK(b, hipMemcpyAsync(b->d_cases + id, b->h_empty, sizeof(CaseCtl), hipMemcpyHostToDevice, b->stream));
  Arena& A = *b->arena[bc.order - 1];
  std::string err;
  if (int s = A.unplace(bc.pl, err, false)) return fail(b, s, err);
  // the slot and the ranges stay out of use until the batch stream has passed this point
  b->retired.push_back(Retired{b->polls, id, bc.order, bc.pl, nullptr, nullptr});
  b->held[id] = 1;
  bc = BatchCase();
  return DGM_OK;
}

int dgm_batch_steps_async(dgm_batch* b, int32_t n) {
  if (!b || n < 0) return DGM_BAD_ARGUMENT;
  bool any = false;
  for (int g = 0; g < 4; ++g) any = any || (b->arena[g] && b->arena[g]->extent_elements() > 0);
  if (int s = flush_admissions(b)) return s;
  if (!any) return DGM_OK;
  if (int s = ensure_graph(b)) return s;
  if (n <= 0) return DGM_OK;  // n = 0: the step graphs are captured, nothing is launched
  hipEvent_t e0, e1;
  BCHK(b, hipEventCreate(&e0));
  BCHK(b, hipEventCreate(&e1));
  BCHK(b, hipEventRecord(e0, b->stream));
  for (int i = 0; i + GRAPH_STEPS <= n; i += GRAPH_STEPS) BCHK(b, hipGraphLaunch(b->exec_n, b->stream));
  for (int i = 0; i < n % GRAPH_STEPS; ++i) BCHK(b, hipGraphLaunch(b->exec, b->stream));
  BCHK(b, hipEventRecord(e1, b->stream));
  b->step_ev.push_back({e0, e1});
  collect_step_times(b, false);
  return DGM_OK;
}

int dgm_batch_synchronize(dgm_batch* b) {
  if (!b) return DGM_BAD_ARGUMENT;
  return full_sync(b);
}

int dgm_batch_device_time(dgm_batch* b, double* ms) {
  if (!b || !ms) return DGM_BAD_ARGUMENT;
  collect_step_times(b, false);
  *ms = b->device_ms;
  return DGM_OK;
}

int dgm_batch_poll_async(dgm_batch* b, double chord, double cl_tol, int64_t* ticket) {
  if (!b || !ticket || !(chord > 0.0) || !(cl_tol >= 0.0)) return DGM_BAD_ARGUMENT;
  if (b->polls - b->polls_done >= 2) return fail(b, DGM_BAD_ARGUMENT, "two polls are outstanding: wait for one first");
  if (int s = flush_admissions(b)) return s;
  const int nc = b->prm.max_cases, k = (int)(b->polls & 1);
  BCHK(b, hipMemsetAsync(b->d_force, 0, 2 * sizeof(double) * nc, b->stream));
  for (int g = 0; g < 4; ++g)
    if (b->arena[g] && b->arena[g]->extent_edges() > 0) {
      b->arena[g]->solver->flush(b->arena[g]->S, b->stream);
      b->arena[g]->solver->wall_force(b->arena[g]->S, b->d_force, 0, (int)b->arena[g]->extent_edges(), b->stream);
    }
  launch_converge(b->d_cases, b->d_force, b->d_cpar, b->d_clprev, b->d_conv, b->d_cl, nc, chord, cl_tol, b->stream);
  BCHK(b, hipGetLastError());
  BCHK(b, hipMemcpyAsync(b->h_poll_cases[k], b->d_cases, sizeof(CaseCtl) * nc, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipMemcpyAsync(b->h_poll_cl[k], b->d_cl, sizeof(double) * nc, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipMemcpyAsync(b->h_poll_conv[k], b->d_conv, sizeof(int) * nc, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipEventRecord(b->poll_ev[k], b->stream));
  *ticket = b->polls++;
  return DGM_OK;
}

int dgm_batch_poll_wait(dgm_batch* b, int64_t ticket, int32_t n, dgm_case_info* out, double* cl, int32_t* converged) {
  if (!b || n < 0 || n > b->prm.max_cases || ticket < b->polls_done || ticket >= b->polls) return DGM_BAD_ARGUMENT;
  if (ticket != b->polls_done) return fail(b, DGM_BAD_ARGUMENT, "polls complete in order: wait for the oldest first");
  const int k = (int)(ticket & 1);
  BCHK(b, hipEventSynchronize(b->poll_ev[k]));
  for (int i = 0; i < n; ++i) {
    const CaseCtl& h = b->h_poll_cases[k][i];
    if (out) {
      dgm_case_info& o = out[i];
      std::memset(&o, 0, sizeof(o));
      o.error_edge = -1;
      if (h.live) {
        const double t = h.pending ? h.t + h.dt : h.t;
        const int32_t steps = h.step + (h.pending ? 1 : 0);
        o.live = 1;
        o.status = case_status_code(h, &o.error_edge);
        o.active = (o.status == DGM_OK && steps < h.final_step && t < h.final_time) ? 1 : 0;
        o.steps = steps;
        o.time = t;
        o.last_dt = h.dt;
        o.troubled_last_stage = h.pending ? h.count[2] : h.last_count;
        o.order = h.group + 1;
        o.n_elements = h.e_end - h.e_off;
      }
    }
    if (cl) cl[i] = h.live ? b->h_poll_cl[k][i] : std::nan("");
    if (converged) converged[i] = h.live ? b->h_poll_conv[k][i] : 0;
  }
  b->polls_done = ticket + 1;
  collect_step_times(b, false);
  release_retired(b, false);
  return DGM_OK;
}

int dgm_batch_case_info(dgm_batch* b, int32_t n, dgm_case_info* out) {
  if (!b || !out || n < 0 || n > b->prm.max_cases) return DGM_BAD_ARGUMENT;
  if (n == 0) return DGM_OK;
  if (int s = flush_admissions(b)) return s;
  BCHK(b, hipMemcpyAsync(b->h_cases, b->d_cases, sizeof(CaseCtl) * n, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  for (int i = 0; i < n; ++i) {
    const CaseCtl& h = b->h_cases[i];
    const BatchCase& bc = b->cases[i];
    dgm_case_info& o = out[i];
    std::memset(&o, 0, sizeof(o));
    o.error_edge = -1;
    if (!bc.live) continue;
    const double t = h.pending ? h.t + h.dt : h.t;
    const int32_t steps = h.step + (h.pending ? 1 : 0);
    o.live = 1;
    o.status = case_status_code(h, &o.error_edge);
    o.active = (o.status == DGM_OK && steps < h.final_step && t < h.final_time) ? 1 : 0;
    o.steps = steps;
    o.time = t;
    o.last_dt = h.dt;
    o.troubled_last_stage = h.pending ? h.count[2] : h.last_count;
    o.order = bc.order;
    o.n_elements = bc.pl.ne;
    o.footprint_bytes = bc.footprint;
  }
  return DGM_OK;
}

int dgm_batch_case_state(dgm_batch* b, int32_t id, double* U) {
  if (!b || !U || id < 0 || id >= b->prm.max_cases || !b->cases[id].live) return DGM_BAD_ARGUMENT;
  if (int s = flush_admissions(b)) return s;
  Arena& A = *b->arena[b->cases[id].order - 1];
  std::string err;
  if (int s = A.download_state(A.S.un, U, b->cases[id].pl, err)) return fail(b, s, err);
  return DGM_OK;
}

int dgm_batch_lift(dgm_batch* b, double chord, int32_t n, double* cl) {
  if (!b || !cl || !(chord > 0.0) || n < 0 || n > b->prm.max_cases) return DGM_BAD_ARGUMENT;
  if (int s = flush_admissions(b)) return s;
  BCHK(b, hipMemsetAsync(b->d_force, 0, 2 * sizeof(double) * b->prm.max_cases, b->stream));
  for (int g = 0; g < 4; ++g)
    if (b->arena[g] && b->arena[g]->extent_edges() > 0) {
      b->arena[g]->solver->flush(b->arena[g]->S, b->stream);
      b->arena[g]->solver->wall_force(b->arena[g]->S, b->d_force, 0, (int)b->arena[g]->extent_edges(), b->stream);
    }
  BCHK(b, hipGetLastError());
  std::vector<double> f(2 * (size_t)std::max(n, 1));
  if (n > 0) BCHK(b, hipMemcpyAsync(f.data(), b->d_force, 2 * sizeof(double) * n, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  for (int i = 0; i < n; ++i) {
    const BatchCase& bc = b->cases[i];
    if (!bc.live) {
      cl[i] = std::nan("");
      continue;
    }
    const dgm_case_params& P = bc.cp;
    const double a = std::sqrt(P.fs_hcr * P.fs_p / P.fs_rho);
    const double V = P.fs_mach * a;
    const double rad = P.fs_aoa_deg * (M_PI / 180.0);
    cl[i] = (f[2 * i + 1] * std::cos(rad) - f[2 * i] * std::sin(rad)) / (0.5 * P.fs_rho * V * V * chord);
  }
  return DGM_OK;
}

int dgm_batch_get_tuning(const dgm_batch* b, int32_t order, dgm_tuning* t) {
  if (!b || !t || order < 1 || order > 4 || !b->arena[order - 1]) return DGM_BAD_ARGUMENT;
  *t = from_tuning(b->arena[order - 1]->S.tune);
  return DGM_OK;
}

int dgm_batch_set_tuning(dgm_batch* b, int32_t order, const dgm_tuning* t) {
  if (!b || !t || order < 1 || order > 4 || !b->arena[order - 1] || !valid_tuning(*t)) return DGM_BAD_ARGUMENT;
  if (int s = full_sync(b)) return s;
  b->arena[order - 1]->S.tune = to_tuning(*t);
  drop_graph(b);
  return DGM_OK;
}

int dgm_batch_autotune(dgm_batch* b, int32_t reps, dgm_tuning chosen[4]) {
  if (!b || reps < 1) return DGM_BAD_ARGUMENT;
  if (int s = full_sync(b)) return s;
  std::vector<Arena*> as;
  for (int g = 0; g < 4; ++g)
    if (b->arena[g]) as.push_back(b->arena[g].get());
  std::string err;
  if (int s = autotune_steps(as, b->d_cases, b->d_dtmin, b->prm.max_cases, b->stream, [b]() { enqueue_step(b, b->stream); }, reps,
                             err))
    return fail(b, s, err);
  for (int g = 0; g < 4; ++g)
    if (b->arena[g] && chosen) chosen[g] = from_tuning(b->arena[g]->S.tune);
  drop_graph(b);
  return DGM_OK;
}

int dgm_batch_profile_steps(dgm_batch* b, int32_t n, dgm_kernel_times out[4]) {
  if (!b || !out || n < 1) return DGM_BAD_ARGUMENT;
  if (int s = full_sync(b)) return s;
  std::vector<Arena*> as;
  std::vector<int> order_of;
  for (int g = 0; g < 4; ++g)
    if (b->arena[g]) {
      as.push_back(b->arena[g].get());
      order_of.push_back(g);
    }
  std::vector<std::array<double, 5>> ms;
  double step_ms = 0.0;
  std::string err;
  if (int s = profile_step_parts(as, b->d_cases, b->d_dtmin, b->prm.max_cases, b->stream, n, ms, step_ms, err))
    return fail(b, s, err);
  for (int g = 0; g < 4; ++g) std::memset(&out[g], 0, sizeof(dgm_kernel_times));
  for (size_t a = 0; a < as.size(); ++a) {
    dgm_kernel_times& o = out[order_of[a]];
    o.surface_ms = ms[a][Solver::PART_SURFACE];
    o.element_ms = ms[a][Solver::PART_ELEMENT];
    o.detect_ms = ms[a][Solver::PART_DETECT];
    o.rebuild_ms = ms[a][Solver::PART_REBUILD];
    o.apply_ms = ms[a][Solver::PART_APPLY];
    o.step_ms = step_ms;
    o.steps = n;
  }
  return DGM_OK;
}

}  // extern "C"
