// Device-side data structures and kernel launchers of the explicit SSP-RK3 step (gfx950).
//
// Device layout (structure of arrays, element index fastest so that one lane per element
// gives coalesced loads):
//   state   [4 * Np][nes]  component c = var * Np + mode, nes = element capacity (a multiple of 64)
//   flux    [n_edges][Ne][4]  numerical flux times half the edge length at each edge point
//   element [3][nes] edge id << 1 | side (side 1 = the element is the edge's right neighbour),
//           [3][nes] neighbour id or -1, geometry [8][nes]
//   edge    [2][n_edges] left / right element, info word, geometry [3][n_edges]
//
// One arena holds any number of cases of one polynomial order as disconnected pieces of one mesh.
// Every case's element range and edge range start at a multiple of 64, so each 64-wide wavefront
// of a one-lane-per-element (or per-edge) kernel belongs to exactly one case; el_case64[e >> 6] and
// ed_case64[ed >> 6] give it (-1: unused capacity). A single-case context is an arena with one case.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dgm {

// element geometry rows: J, metric (dxi/dx, dxi/dy, deta/dx, deta/dy), circumradius^1.5 (the KXRCF
// scale), minimum height, 2/J
enum ElemGeo { G_J = 0, G_XX = 1, G_XY = 2, G_EX = 3, G_EY = 4, G_CR15 = 5, G_HMIN = 6, G_2J = 7, G_N = 8 };
enum EdgeBc { BC_INTERIOR = 0, BC_WALL = 1, BC_FARFIELD = 2 };
// element kernel operations: residual only, or residual fused with RK stage 1, 2, 3
enum ElemOp { OP_RESID = 0, OP_S0 = 1, OP_S1 = 2, OP_S2 = 3 };

struct DevMesh {
  int ne, nes, nedge, npatch;  // ne / nedge: extent the kernels cover; nes: element capacity (stride)
  double hcr;
  const int* el_edge;      // [3][nes]
  const int* el_nbr;       // [3][nes]
  const double* el_geo;    // [G_N][nes]
  const float* el_geo32;   // [G_N][nes]
  const int* ed_elem;      // [2][nedge capacity]
  const int* ed_info;      // [nedge]: lic_left | lic_right << 2 | bc << 4 | patch << 8
  const double* ed_geo;    // [3][nedge capacity]: nx, ny, half length (boundary edges: normal oriented by the owner)
  const float* ed_geo32;   // [3][nedge capacity]
  const double* patch_q;   // [npatch][4] far-field conserved state
  const int* el_case64;    // [nes / 64] case of each 64-element block, -1 unused
  const int* ed_case64;    // [edge capacity / 64] case of each 64-edge block, -1 unused
  int ed_stride;           // edge capacity (stride of ed_elem / ed_geo)
  int single_case;         // >= 0: every block belongs to this case (no map lookup); -1: use the maps
  // The fused edge-and-element pass (FP32-bulk, p = 3, DGM_FUSED=1): per group of 16 elements (one wavefront), the
  // records of the distinct edges of its elements in increasing edge order, j < grp_n[g]: grp_rec =
  // {left element, right element or -1 (boundary), info word, edge id}, grp_geo = {nx, ny, half length, 0}
  // (the edge arrays' values, gathered per group at set-up so that a wavefront reads them in one trip);
  // el_gslot[i][e] = (place of face i's edge in its group's list) << 1 | side. nullptr without that pass.
  const int4* grp_rec;     // [nes / 16][GRP_EDGES]
  const float4* grp_geo;   // [nes / 16][GRP_EDGES]
  const int* grp_n;        // [nes / 16]
  const int* el_gslot;     // [3][nes]
};
constexpr int GRP_EDGES = 48;  // at most three distinct edges per element of a 16-element group

// Step control of one case, resident on the device so that whole steps replay from a captured graph.
// A step is counted at the start of the next one (pending): the time-step reduction for step n + 1
// is fused into step n's last limiter pass and lands in the case's dt slot (n + 1) & 1.
struct CaseCtl {
  double t, dt, final_time, cfl;
  unsigned long long error;         // (local edge << 8) | status of the lowest failing edge, ~0ull when clean
  int step, final_step, active, pending;
  int live;                         // the slot holds a case
  int e_off, ed_off;                // first element / edge of the case in its arena
  int e_end, ed_end;                // one past its last element / edge (its 64-blocks may extend further)
  int group;                        // arena (order) of the case
  int count[4];                     // troubled cells per stage of the running step (slot 3: stand-alone limiter)
  int last_count;                   // troubled cells of the last stage of the previous step
  int pad0;
  // first failing edge of the stage-0 edge pass of the running step (same coding as error): that pass
  // runs beside the step control, which reads error, so it records here and the stage-0 element pass
  // moves it into error when the case runs the step
  unsigned long long error_s0;
  // First failing edge of the fused edge-and-element pass of stage s (same coding as error). Every
  // wavefront of such a pass decides at its start whether the case runs, so the pass records into a
  // word that none of its own wavefronts reads; the passes after it stop the case on it (case_error).
  unsigned long long error_f[3];
  int pad[2];
};
static_assert(sizeof(CaseCtl) == 144, "CaseCtl is one 144-byte record");

// the error status of a case: its error word and those of the fused passes
__host__ __device__ inline unsigned long long case_error(const CaseCtl& c) {
  unsigned long long e = c.error;
  for (int s = 0; s < 3; ++s) e = c.error_f[s] < e ? c.error_f[s] : e;
  return e;
}

// The time-step minimum of a case is reduced into DT_LANES words 128 bytes apart: atomics from many
// workgroups to one word serialise in L2, so each wavefront picks a word by its index.
constexpr int DT_LANES = 32;
constexpr int DT_STRIDE = 16;                       // unsigned long longs between two targets (128 bytes)
constexpr int DT_SLOT = DT_LANES * DT_STRIDE;       // one slot
constexpr int DT_CASE = 2 * DT_SLOT;                // two slots (step parity) per case

// Kernel-argument view of the control state of one arena.
struct Ctl {
  CaseCtl* cases;                // [max cases]
  unsigned long long* dtmin;     // [max cases][2][DT_SLOT]
  int* count;                    // [4] troubled-list lengths of this arena per stage slot
};

// Buffers the element kernel reads and writes. FP64 mode uses in/un/us; FP32-bulk uses the
// float shadows sn/s and increments d, with un kept in FP64.
template <class T>
struct StageBufs {
  const T* in;     // stage input (read by the volume integral)
  double* un;      // u^n (FP64)
  T* us;           // FP64 mode: stage state
  T* sn;           // FP32-bulk: float shadow of u^n
  T* s;            // FP32-bulk: float state of stage 1, u^(1) (written by stage 0)
  T* s2;           // FP32-bulk: float state of stage 2, u^(2) (written by stage 1). A stage never writes
                   // the array it reads, so kernels that read the stage input of other elements may run
                   // beside the update of their own.
  T* d;            // FP32-bulk: stage increment u^(s) - u^n
  double* rout;    // OP_RESID output, [4 Np][nes]
  // Deferred limiter output (Jacobi march): pend[e] >= 0 means element e's stage input is the rebuilt
  // polynomial pw[pend[e]][4][Np] of the last limiter pass, not the stored state; nullptr: none.
  const int* pend;
  const double* pw;
};

struct LimiterBufs {
  int* list;       // troubled element ids
  double* w;       // reconstructed polynomials [slot][4][Np]
  uint8_t* flags;  // per-element troubled flags (1 troubled, 0 not); between the element-order limiter's
                   // detector/rebuild pass and its walk, 2 marks a troubled cell whose rebuild failed
  int* pend;       // per element: its slot in w after the last detector pass, or -1
};

// Launch shapes of one arena, chosen by the autotuner (defaults are the measured best at p = 3).
struct Tuning {
  int surface_block = 256;   // threads per block of the edge kernel (64, 128, 256)
  int element_waves = 4;     // FP32 matrix-core element kernel: wavefronts (16 elements each) per block (1, 2, 4);
                             // 0: the one-lane-per-element kernel instead (the faster one at p <= 2)
  int element_block = 128;   // one-lane-per-element kernel (always in FP64): threads per block (64, 128, 256)
  int detect_block = 256;    // detector threads per block (64, 128, 256)
  int remap = 1;             // XCD-contiguous block order for the gather kernels
  int weno_grid = 2048;      // wavefronts of the WENO rebuild
};

}  // namespace dgm
