// Internal declarations shared by the host preprocessing (tables.cpp, mesh.cpp) and the
// device code / C ABI (dgm.hip). Not part of the public ABI (include/dgm.h is).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dgm.h"

namespace dgm {

// Basis / quadrature tables for one order p (DubinerBasis::new src/basis_function.rs:15-38,
// GaussPoints::new src/gauss_point.rs:13-37).
struct Tables {
  int p = 0, np = 0, nq = 0, ne = 0;
  std::vector<double> cell_pts, cell_w;  // [nq][2], [nq]
  std::vector<double> edge_pts, edge_w;  // [3][ne][2], [ne]
  std::vector<double> phi_cell;          // [nq][np]
  std::vector<double> phi_edge;          // [3][ne][np]
  std::vector<double> mass;              // [np]
  int maxd = 0, nkeys = 0;               // reference-element derivative keys (a,b), 1 <= a+b <= maxd
  std::vector<int> key_a, key_b;         // ordered by total degree, then a descending
  std::vector<double> dref;              // [nkeys][nq][np] d^(a+b) phi / dxi^a deta^b at the cell points
  int key(int a, int b) const;
};

int default_cell_rule(int p);
int default_edge_rule(int p);
// Returns false for an unsupported (p, nq, ne) combination.
bool build_tables(Tables& t, int p, int nq, int ne);

// Mesh after process_mesh (src/initialization.rs:112-261) and the geometry of src/mesh.rs:15-147,
// kept in the reference's numbering, structure-of-arrays.
struct HostMesh {
  int64_t nv = 0, ne = 0, nedge = 0, n_boundary_edges = 0;
  std::vector<double> vx, vy;
  std::vector<double> exy;                 // optional [ne][3][2]
  // edges
  std::vector<int64_t> ed_elem;            // [nedge][2], right = -1 on boundary edges
  std::vector<int64_t> ed_vert;            // [nedge][2]
  std::vector<int64_t> ed_lic;             // [nedge][2], right = -1 on boundary edges
  std::vector<int64_t> ed_patch;           // [nedge], -1 interior
  std::vector<double> ed_normal;           // [nedge][2]
  std::vector<double> ed_len;              // [nedge]
  std::vector<double> ed_bc_normal;        // [nedge][2]: boundary normal oriented by the owning element
  // elements
  std::vector<int64_t> el_vert, el_edge, el_nbr;  // [ne][3]
  std::vector<double> el_J, el_cr, el_hmin;       // [ne]
  std::vector<double> el_metric;                  // [ne][4]: dxi/dx, dxi/dy, deta/dx, deta/dy
  std::vector<int64_t> internal_edges, internal_elements, boundary_elements;
  // patches (boundary blocks)
  std::vector<int> patch_type;             // DGM_WALL / DGM_FARFIELD
  // level schedule of the in-place element-order limiter sweep: element i after every lower-numbered neighbour
  std::vector<int32_t> level_offsets;      // [nlevels + 1]
  std::vector<int32_t> level_elems;        // [ne] grouped by level, increasing element index within a level
};

// Builds connectivity, geometry and the level schedule. Returns a dgm_status.
int build_mesh(const dgm_mesh& in, HostMesh& m, std::string& err);

}  // namespace dgm
