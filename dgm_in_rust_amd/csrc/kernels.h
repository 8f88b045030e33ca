// Device-side data structures and kernel launchers of the explicit SSP-RK3 step (gfx950).
//
// Device layout (structure of arrays, element index fastest so that one lane per element
// gives coalesced loads):
//   state   [4 * Np][nes]  component c = var * Np + mode, nes = n_elements rounded up to 64
//   flux    [n_edges][Ne][4]  numerical flux times half the edge length at each edge point
//   element [3][nes] edge id << 1 | side (side 1 = the element is the edge's right neighbour),
//           [3][nes] neighbour id or -1, geometry [8][nes]
//   edge    [2][n_edges] left / right element, info word, geometry [3][n_edges]
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dgm {

enum ElemGeo { G_J = 0, G_XX = 1, G_XY = 2, G_EX = 3, G_EY = 4, G_CR = 5, G_HMIN = 6, G_2J = 7, G_N = 8 };
enum EdgeBc { BC_INTERIOR = 0, BC_WALL = 1, BC_FARFIELD = 2 };
// element kernel operations: residual only, or residual fused with RK stage 1, 2, 3
enum ElemOp { OP_RESID = 0, OP_S0 = 1, OP_S1 = 2, OP_S2 = 3 };

struct DevMesh {
  int ne, nes, nedge, npatch;
  const int* el_edge;      // [3][nes]
  const int* el_nbr;       // [3][nes]
  const double* el_geo;    // [G_N][nes]
  const float* el_geo32;   // [G_N][nes]
  const int* ed_elem;      // [2][nedge]
  const int* ed_info;      // [nedge]: lic_left | lic_right << 2 | bc << 4 | patch << 8
  const double* ed_geo;    // [3][nedge]: nx, ny, half length (boundary edges: normal oriented by the owner)
  const float* ed_geo32;   // [3][nedge]
  const double* patch_q;   // [npatch][4] far-field conserved state
};

// Step control, resident on the device so that whole steps replay from a captured graph.
// A step is counted at the start of the next one (pending): the time-step reduction for step n + 1
// is fused into step n's last limiter pass and lands in dtmin_bits[(n + 1) & 1].
// The reduction targets are spread over DT_LANES words 128 bytes apart: atomics from many workgroups
// to one word serialise in L2, so each workgroup / wavefront picks a word by its index.
constexpr int DT_LANES = 32;
constexpr int DT_STRIDE = 16;  // unsigned long longs between two targets (128 bytes)
struct Ctl {
  double t, dt, final_time, cfl, hcr;
  unsigned long long error;         // (edge << 8) | status of the lowest failing edge, ~0ull when clean
  int step, final_step, active, pending;
  int count[4];                     // troubled cells per stage (slot 3: stand-alone limiter call)
  int last_count, pad;
  unsigned long long dtmin_bits[2][DT_LANES * DT_STRIDE];  // atomicMin targets by step parity
};

// Buffers the element kernel reads and writes. FP64 mode uses in/un/us; FP32-bulk uses the
// float shadows sn/s and increments d, with un kept in FP64.
template <class T>
struct StageBufs {
  const T* in;     // stage input (read by the volume integral)
  double* un;      // u^n (FP64)
  T* us;           // FP64 mode: stage state
  T* sn;           // FP32-bulk: float shadow of u^n
  T* s;            // FP32-bulk: float stage state
  T* d;            // FP32-bulk: stage increment u^(s) - u^n
  double* rout;    // OP_RESID output, [4 Np][nes]
};

struct LimiterBufs {
  int* list;       // troubled element ids
  double* w;       // reconstructed polynomials [slot][4][Np]
  uint8_t* flags;  // optional per-element troubled flags
};

}  // namespace dgm
