/*
 * refcpu — FP64 CPU restatement of the Decem5150/DGM_in_rust explicit time step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the oracle that the HIP product is checked
 * against. It may be loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg, and by nothing else: the product (dgm_in_rust_amd/) never links, loads or calls it.
 *
 * Each function restates the reference file:line named in refcpu.cpp, keeping the
 * reference's operation order, literals and quirks (sum-of-modes time step, altered HLLC
 * p*, KXRCF neighbour-trace index, WENO linear-weight index, in-place sequential limiter).
 * Orders p != 2 and periodic meshes are extensions of the reference (which supports p = 2
 * only); their p = 2 path is bit-identical to the restatement.
 *
 * Status codes match include/dgm.h.
 */
#ifndef DGM_REFCPU_H
#define DGM_REFCPU_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ref_gmsh ref_gmsh;
typedef struct ref_case ref_case;

/* GmshParser::parse (src/io/mesh_parser.rs:28-158). Returns NULL if the file cannot be opened. */
ref_gmsh* ref_gmsh_parse(const char* path);
void ref_gmsh_sizes(const ref_gmsh* g, int64_t* n_vertices, int64_t* n_edge_blocks, int64_t* n_edge_nodes_total,
                    int64_t* n_elem_blocks, int64_t* n_elements_total);
/* vxy: [nv][2]; eb_type: [neb] (0 wall, 1 farfield, -1 other); eb_count: [neb]; eb_nodes: [total][2] (0-based);
 * el_count: [nelb]; el_nodes: [total][3] (0-based). */
void ref_gmsh_get(const ref_gmsh* g, double* vxy, int32_t* eb_type, int64_t* eb_count, int64_t* eb_nodes,
                  int64_t* el_count, int64_t* el_nodes);
void ref_gmsh_free(ref_gmsh* g);

typedef struct {
  int32_t order;        /* p */
  int32_t n_cell_gp;    /* triangle rule: 4, 6, 16, 25 */
  int32_t n_edge_gp;    /* Gauss-Legendre rule: 3, 4, 5 */
  double hcr;           /* ratio of specific heats */
  double cfl;           /* <= 0: the reference formula of src/initialization.rs:29-33 */
  double fs_mach, fs_aoa_deg, fs_rho, fs_p, fs_hcr; /* freestream.json fields (far-field patch state) */
  int32_t limiter;      /* 0 off, 1 sequential in-place (reference order), 2 Jacobi (all cells from the pre-limiter state) */
  int32_t nthreads;     /* residual folds: 1 = deterministic sequential order */
  int32_t gemm_fma;     /* 1: 4x4 gemm as an FMA chain (matrixmultiply's FMA kernel), 0: mul/add chain */
} ref_params;

/* Mesh in GmshParser output form. elem_xy (nullable): [ne][3][2] per-element vertex coordinates,
 * used for periodic meshes whose vertices are identified across the period. */
ref_case* ref_create(int64_t nv, const double* vxy, int64_t neb, const int32_t* eb_type, const int64_t* eb_count,
                     const int64_t* eb_nodes, int64_t nelb, const int64_t* el_count, const int64_t* el_nodes,
                     const double* elem_xy, const ref_params* prm, int32_t* status);
void ref_destroy(ref_case* c);

void ref_sizes(const ref_case* c, int64_t* n_elements, int64_t* n_edges, int64_t* n_internal_edges,
               int64_t* n_internal_elements, int64_t* n_boundary_elements, int32_t* nbasis, int32_t* ncgp, int32_t* negp);
/* connectivity and geometry as built by process_mesh + Mesh::compute_* */
void ref_get_edges(const ref_case* c, int64_t* ielements, int64_t* ivertices, int64_t* in_cell, int64_t* ipatch,
                   double* normal, double* jacob);
void ref_get_elements(const ref_case* c, int64_t* ivertices, int64_t* iedges, int64_t* ineighbours, double* jacob,
                      double* circumradius, double* minimum_height);
void ref_get_lists(const ref_case* c, int64_t* internal_edges, int64_t* internal_elements, int64_t* boundary_elements);
/* basis tables: phi_cell [nq][np], phi_edge [3][ne][np], mass [np], cell points [nq][2], cell weights [nq],
 * edge points [3][ne][2], edge weights [ne], dref [nkeys][nq][np] (keys (a,b), 1<=a+b<=max(p,2), ordered by
 * total degree then a descending: (1,0),(0,1),(2,0),(1,1),(0,2),...) */
int32_t ref_n_deriv_keys(const ref_case* c);
void ref_get_tables(const ref_case* c, double* phi_cell, double* phi_edge, double* mass, double* cell_pts,
                    double* cell_w, double* edge_pts, double* edge_w, double* dref);
double ref_cfl(const ref_case* c);
/* edge (or element) at which the last failing residual fold stopped: the first failing entry of its list */
int64_t ref_last_bad_index(const ref_case* c);
/* physical derivative d^(l1+l2) phi_k / dx^l1 dy^l2 at cell point igp of element e (src/mesh.rs:85-132) */
double ref_phys_deriv(const ref_case* c, int64_t e, int32_t igp, int32_t k, int32_t l1, int32_t l2);

/* Solver::set_initial_solution (src/solver.rs:45-79): L2 projection of a uniform state. */
void ref_initial_solution(const ref_case* c, double density, double aoa_deg, double mach, double pressure, double* U);
/* extension: L2 projection of pointwise conserved values q[ne][nq][4] given at the cell points */
void ref_project(const ref_case* c, const double* q_at_points, double* U);
/* physical coordinates of the cell points, [ne][nq][2] */
void ref_cell_points_xy(const ref_case* c, double* xy);

/* SpatialDisc::compute_residuals (src/spatial_disc.rs:27-39): ACCUMULATES into R then divides by the mass. */
int32_t ref_compute_residuals(ref_case* c, double* R, const double* U, int64_t* bad_edge);
/* component terms (each accumulating into R, no mass division) */
int32_t ref_volume_terms(ref_case* c, double* R, const double* U);
int32_t ref_edge_terms(ref_case* c, double* R, const double* U, int64_t* bad_edge);
int32_t ref_boundary_terms(ref_case* c, double* R, const double* U);
void ref_divide_by_mass(const ref_case* c, double* R);

/* SpatialDisc::apply_limiter (src/spatial_disc/limiter.rs:7-9); mode from ref_params.limiter. troubled (nullable)
 * receives the KXRCF flags that were acted on. */
int32_t ref_apply_limiter(ref_case* c, double* U, uint8_t* troubled);
int32_t ref_kxrcf(const ref_case* c, const double* U, int64_t ielem);
// WENO reconstruction of one element from U (weno.rs:86-269 for one troubled cell), U untouched.
int32_t ref_weno_element(const ref_case* c, const double* U, int64_t ielem, double* out);
double ref_compute_time_step(const ref_case* c, const double* U);
int32_t ref_rk3(ref_case* c, double* R, double* U, double dt);
int32_t ref_time_march(ref_case* c, double* R, double* U, int32_t final_step, double final_time, int32_t* steps_done,
                       double* t_done);

/* pieces for unit checks */
int32_t ref_hllc(const double* ql, const double* qr, const double* n, double hcr, double* f);
void ref_flux(const double* q, double hcr, double* f, double* g);
void ref_eigenmatrix(const double* q, double nx, double ny, double hcr, double* L, double* R);
/* RK3 combination of src/temporal_disc/rk3.rs:7-26 with the residual R(u) = lambda*u, limiter off */
void ref_rk3_linear(double* u, int64_t n, double lambda, double dt);

#ifdef __cplusplus
}
#endif
#endif
