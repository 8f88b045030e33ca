// Arena: the device buffers of one polynomial order, holding any number of cases as disconnected
// pieces of one structure-of-arrays mesh (see kernels.h). Used with one case by the single-case
// context (dgm.cpp) and with many by the batch / serving context (batch.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "dgm_internal.h"
#include "solver.h"

namespace dgm {

// First-fit allocator of ranges of units (one unit = 64 elements or 64 edges, or one patch).
class RangeAlloc {
 public:
  void reset(int64_t capacity);
  int64_t alloc(int64_t n);  // start, or -1 when no free range is long enough
  void release(int64_t start, int64_t n);
  int64_t used_end() const;  // one past the highest allocated unit
  int64_t free_units() const;
  int64_t largest_free() const;

 private:
  std::map<int64_t, int64_t> free_;  // start -> length
  std::map<int64_t, int64_t> used_;  // start -> length
  int64_t cap_ = 0;
};

// Where one case lives in an arena.
struct Placement {
  int64_t e_off = -1, ed_off = -1, patch_off = -1;  // first element / edge / patch
  int64_t ne = 0, ned = 0, npatch = 0;              // sizes
  int64_t ne_alloc = 0, ned_alloc = 0;              // sizes rounded up to 64
};

// Per-element and per-edge device bytes of an arena of this configuration (the admission footprint).
struct Footprint {
  int64_t per_element = 0, per_edge = 0, per_patch = 0;
  int64_t of(int64_t ne, int64_t ned, int64_t npatch) const {
    return per_element * ((ne + 63) / 64 * 64) + per_edge * ((ned + 63) / 64 * 64) + per_patch * npatch;
  }
};

class Arena {
 public:
  Arena() = default;
  Arena(const Arena&) = delete;
  Arena& operator=(const Arena&) = delete;
  ~Arena();

  static Footprint footprint(int order, int nq, int ne_pts, int precision, int limiter, bool residual_buffer);

  // Tables and all buffers at the given capacities (elements and edges are rounded up to 64).
  // cases / dtmin are the owner's control records; the arena keeps its own list counters.
  int init(int order, int nq, int ne_pts, int precision, int limiter, int64_t cap_elements, int64_t cap_edges,
           int64_t cap_patches, bool residual_buffer, CaseCtl* cases, unsigned long long* dtmin, hipStream_t st,
           std::string& err);
  // Reserves ranges for a case and uploads its mesh, far-field states and smoothness matrices, then
  // marks its 64-blocks as belonging to case_id. Stream-ordered; returns DGM_NO_SPACE if a range does
  // not fit (nothing is changed then).
  int place(const HostMesh& M, const std::vector<double>& patch_q, int case_id, Placement& out, std::string& err);
  // Marks the case's blocks unused, stream-ordered on `stream` without waiting for it, and frees its ranges
  // (release = false: keeps them allocated; release_ranges then frees them once the stream has passed).
  int unplace(const Placement& pl, std::string& err, bool release = true);
  void release_ranges(const Placement& pl);
  // host layout [ne][4][Np] <-> the case's state (stream-ordered; returns when the copy is done)
  int upload_state(const double* U, const Placement& pl, std::string& err);
  int download_state(const double* soa, double* U, const Placement& pl, std::string& err);
  // kernel extents = allocated extents (a captured graph must be re-captured when these change)
  void update_extent();

  int64_t extent_elements() const { return S.mesh.ne; }
  int64_t extent_edges() const { return S.mesh.nedge; }

  Tables T;
  std::unique_ptr<Solver> solver;
  DeviceState S;
  int order = 0, precision = 0, limiter = 0;
  int64_t cap_e = 0, cap_ed = 0, cap_patch = 0;
  int64_t bytes = 0;
  RangeAlloc r_e, r_ed, r_patch;  // units of 64 elements, 64 edges, 1 patch
  hipStream_t stream = nullptr;
  // when set, place() and upload_state() run on this stream instead (a batch admits cases beside its steps)
  hipStream_t ustream = nullptr;

 private:
  template <class X>
  int dalloc(X** p, size_t n, std::string& err);
  int stage_copy(const std::vector<std::pair<void*, std::pair<const void*, size_t>>>& pieces, std::string& err);
  hipStream_t pstream() const { return ustream ? ustream : stream; }
  int grow_staging(size_t n, std::string& err);  // device buffer of n doubles for host-layout states
  std::vector<void*> allocs_;
  void* pinned_ = nullptr;
  size_t pinned_size_ = 0;
  double* hbuf_ = nullptr;  // device staging of a host-layout state
  size_t hbuf_size_ = 0;
};

// Far-field conserved state per patch (src/initialization.rs:124-133, boundary_condition.rs:50-55).
std::vector<double> farfield_states(const HostMesh& M, double mach, double aoa_deg, double rho, double p, double fs_hcr,
                                    double hcr);

// Empty control record of a case.
CaseCtl fresh_case(double cfl);

// Solver::set_initial_solution (src/solver.rs:45-79): L2 projection of a uniform state, host layout.
void initial_state(const Tables& T, const HostMesh& M, double density, double aoa_deg, double mach, double pressure,
                   double hcr, double* U);

// Chooses the launch shapes of every arena's kernels by timing whole steps (`step` enqueues one on st),
// one knob at a time per arena. Each trial starts from a snapshot of what a step reads from earlier
// steps (u^n and its FP32 shadow, the case records and their dt slots), which is restored at the end,
// so every case's state and clock are left as they were.
int autotune_steps(const std::vector<Arena*>& arenas, CaseCtl* cases, unsigned long long* dtmin, int ncases,
                   hipStream_t st, const std::function<void()>& step, int reps, std::string& err);

// Per-kernel timing of n eager steps on st: the step control once per step, then each stage's parts
// arena by arena, each launch timed by events its own dispatch records. ms[a][part] is the mean per
// launch of arenas[a]'s part (0 for a part that launches nothing), step_ms the mean eager step. The
// state, case records and dt slots are restored afterwards (as by autotune_steps).
int profile_step_parts(const std::vector<Arena*>& arenas, CaseCtl* cases, unsigned long long* dtmin, int ncases,
                       hipStream_t st, int n, std::vector<std::array<double, 5>>& ms, double& step_ms,
                       std::string& err);

// launch shapes <-> their C-ABI record
Tuning to_tuning(const dgm_tuning& t);
dgm_tuning from_tuning(const Tuning& u);
bool valid_tuning(const dgm_tuning& t);

// CFL of the reference for order p (src/initialization.rs:29-33), or the given value when > 0.
double default_cfl(int order, double cfl);

}  // namespace dgm
