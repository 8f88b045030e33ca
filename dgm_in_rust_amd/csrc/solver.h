// Launch-sequence interface shared by the C ABI (dgm.cpp, batch.cpp) and the kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "../../include/dgm.h"
#include "kernels.h"

namespace dgm {

// Device buffers of one arena (one polynomial order; one or many cases). Sizes: state arrays
// [4 Np][nes]; flux [edge capacity][Ne][4].
struct DeviceState {
  DevMesh mesh{};
  double* un = nullptr;    // u^n, FP64
  void* us = nullptr;      // FP64 mode: stage state
  void* sn = nullptr;      // FP32-bulk: float shadow of u^n
  void* s = nullptr;       // FP32-bulk: float state of stage 1 (u^(1)), written by stage 0
  void* s2 = nullptr;      // FP32-bulk: float state of stage 2 (u^(2)), written by stage 1
  void* d = nullptr;       // FP32-bulk: stage increment
  double* rout = nullptr;  // residual output
  void* flux = nullptr;
  void* tab = nullptr;     // bulk-precision tables
  double* ltab = nullptr;  // FP64 limiter tables
  double* smat = nullptr;  // per-element packed smoothness matrices [nes][Np(Np+1)/2]
  int* lim_list = nullptr;
  double* lim_w = nullptr;
  uint8_t* lim_flags = nullptr;
  int* lim_pend = nullptr;     // [nes] slot in lim_w of a rebuilt element not yet written back, or -1
  Ctl ctl{};               // case records (shared by the arenas of a batch) and this arena's list counters
  int limiter = 0;
  int use_mfma = 1;        // FP32-bulk element kernel on the matrix cores (p <= 3)
  int fused = 0;           // a march's edge and element passes run as one fused pass (FP32-bulk, p = 3)
  Tuning tune{};
  std::vector<int> level_offsets;
  const int* levels = nullptr;
  // element-order walk records {e, lower-numbered neighbours or -1} in level order, each level padded
  // to 64-record chunks; nullptr: one launch per level instead
  const int* walk = nullptr;
  int walk_chunks = 0;
  // when set, the next single-launch stage part records these at its dispatch's start and end
  hipEvent_t prof_ev[2] = {nullptr, nullptr};
};

class Solver {
 public:
  virtual ~Solver() = default;
  virtual int table_len() const = 0;
  virtual int limiter_table_len() const = 0;
  virtual bool fp32() const = 0;
  virtual bool has_mfma() const = 0;
  // residual of the resident u^n into rout (no time control, always runs)
  virtual void residual(const DeviceState& s, hipStream_t st) const = 0;
  // SpatialDisc::apply_limiter on the resident u^n (always runs)
  virtual void limiter(const DeviceState& s, hipStream_t st) const = 0;
  // RK stage 0, 1, 2 (surface + element + limiter), each case gated by its CaseCtl::active. With
  // nctl > 0 (stage 0), the stage's edge launch also runs the step control of case slots 0 .. nctl - 1,
  // which then must not be launched on its own.
  virtual void stage(const DeviceState& s, int stage_idx, hipStream_t st, int nctl = 0) const = 0;
  // the parts of a stage, in order: surface flux, element update, detector (or the whole
  // element-order sweep), WENO rebuild, apply
  enum { PART_SURFACE = 0, PART_ELEMENT = 1, PART_DETECT = 2, PART_REBUILD = 3, PART_APPLY = 4, N_PARTS = 5 };
  virtual void stage_part(const DeviceState& s, int stage_idx, int part, hipStream_t st, int nctl = 0) const = 0;
  // per-element smoothness matrices of the WENO indicator (set-up), elements e0 .. e0 + n - 1
  virtual int smoothness_len() const = 0;
  virtual void build_smoothness(const DeviceState& s, int e0, int n, hipStream_t st) const = 0;
  // time-step reduction into the dt slot (step + pending) of every live case, or of case `only` when >= 0
  virtual void time_step(const DeviceState& s, hipStream_t st, int force, int only = -1) const = 0;
  // host layout [n][4][Np] (device copy) <-> state arrays, elements e0 .. e0 + n - 1
  virtual void upload(const DeviceState& s, const double* host_layout_dev, int e0, int n, hipStream_t st) const = 0;
  virtual void download(const DeviceState& s, const double* soa, double* host_layout_dev, int e0, int n,
                        hipStream_t st) const = 0;
  virtual void check_finite(const DeviceState& s, int* flag, int e0, int n, hipStream_t st) const = 0;
  // wall force per case into force[2 case + i], edges ed0 .. ed0 + n - 1 (ed0 a multiple of 64)
  virtual void wall_force(const DeviceState& s, double* force, int ed0, int n, hipStream_t st) const = 0;
  // writes back every rebuilt element still pending from the last limiter pass (a last stage's pass:
  // u^n and its FP32 shadow) and clears the marks; needed before the state is read or snapshot
  virtual void flush(const DeviceState& s, hipStream_t st) const = 0;
};

std::unique_ptr<Solver> make_solver(int precision, int p, int nq, int ne);

// per-case step control: one thread per case slot
void launch_step_begin(Ctl ctl, int ncases, hipStream_t st);
// Serving polls. For the case in slot i: its lift normalisation cpar[i] = {cos aoa, sin aoa, 1 / (rho V^2 / 2)},
// its last lift coefficient (0 at admission) and its convergence mark.
void launch_case_poll_init(double* cpar, double* clprev, int* conv, int slot, double mach, double aoa_deg, double rho,
                           double p, double hcr, hipStream_t st);
// Per slot: the lift coefficient from the wall force (force[2 i], force[2 i + 1]) and chord; a running case
// whose coefficient moved by less than tol since the previous poll stops after the step it is at
// (final_step = its step count) and is marked converged.
void launch_converge(CaseCtl* cases, const double* force, const double* cpar, double* clprev, int* conv, double* cl,
                     int ncases, double chord, double tol, hipStream_t st);
// stream-ordered write of one control record by a one-thread kernel
void launch_set_case(CaseCtl* dst, const CaseCtl& value, hipStream_t st);

}  // namespace dgm
