/*
 * dgm.h — C ABI of the MI355X (gfx950) DG-Euler explicit time step.
 *
 * Drop-in for the call surface of the Rust crate Decem5150/DGM_in_rust:
 *   SpatialDisc::compute_residuals  src/spatial_disc.rs:27-39          -> dgm_compute_residuals
 *   SpatialDisc::apply_limiter      src/spatial_disc/limiter.rs:7-9     -> dgm_apply_limiter
 *   TemperalDisc::compute_time_step src/temporal_disc.rs:58-88          -> dgm_compute_time_step
 *   TemperalDisc::rk3               src/temporal_disc/rk3.rs:7-26       -> dgm_rk3
 *   TemperalDisc::time_march        src/temporal_disc.rs:21-57          -> dgm_time_march
 *   TemperalDisc::set_residual_to_zero src/temporal_disc.rs:89-93       -> (no call: dgm_compute_residuals overwrites)
 * and of the set-up that feeds them:
 *   initialize_mesh_basis_gauss_params / process_mesh / Mesh::compute_*  src/initialization.rs:11-261,
 *   src/mesh.rs:15-147, DubinerBasis::new src/basis_function.rs:15-38, GaussPoints::new src/gauss_point.rs:13-37
 *                                                                       -> dgm_ctx_create
 *   Solver::set_initial_solution    src/solver.rs:45-79                 -> dgm_initial_solution
 *
 * Host arrays use the reference layout: solutions / residuals are C-order [n_elements][4][n_basis]
 * f64 (ndarray Array<f64, Ix3>, src/initialization.rs:70-71). All calls return a dgm_status;
 * none aborts. Where the reference panics (negative density or pressure in HLLC,
 * src/spatial_disc.rs:189-201; unsupported rule sizes, src/gauss_point.rs:56-58,90-92; more than
 * two boundary faces in WENO, src/spatial_disc/limiter/weno.rs:107), the call returns the
 * matching status and dgm_last_error names the offending edge or element.
 *
 * A context owns its device buffers and one HIP stream; it is not thread-safe. Different
 * contexts may be used from different threads.
 */
#ifndef DGM_H
#define DGM_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  DGM_OK = 0,
  DGM_NEG_DENSITY = 1,   /* flux::hllc Err("Negative density found in HLLC!") */
  DGM_NEG_PRESSURE = 2,  /* flux::hllc Err("Negative pressure found in HLLC!") */
  DGM_NONFINITE = 3,     /* non-finite state found by dgm_check_state */
  DGM_BAD_CONFIG = 4,    /* unsupported order / rule / option */
  DGM_BAD_MESH = 5,      /* mesh the reference would panic on */
  DGM_HIP_ERROR = 6,
  DGM_OUT_OF_MEMORY = 7,
  DGM_BAD_ARGUMENT = 8,
  DGM_NO_SPACE = 9       /* batch: the case does not fit the free capacity now (admission refused) */
} dgm_status;

enum { DGM_FP64 = 0, DGM_FP32_BULK = 1 };
/* limiter sweep order: off; the reference's in-place element-order sweep (weno.rs:90,263), run on the
 * device as a detector and rebuild pass over every cell, then one wavefront walking the level schedule
 * that redoes the cells with a rebuilt lower-numbered neighbour (one launch per level when the
 * environment sets DGM_SEQUENTIAL_WALK=0 at create, or above 2^18 cells); or Jacobi (every cell tested
 * and rebuilt from the pre-limiter state). */
enum { DGM_LIMITER_OFF = 0, DGM_LIMITER_SEQUENTIAL = 1, DGM_LIMITER_JACOBI = 2 };
enum { DGM_WALL = 0, DGM_FARFIELD = 1 };

/* The mesh as GmshParser::parse returns it (src/io/mesh_parser.rs:28-158). Node ids are 0-based.
 * element_xy (nullable) gives each triangle's own vertex coordinates [ne][3][2]: periodic meshes
 * identify vertices across the period and need it (extension; the reference has no periodic BC). */
typedef struct {
  int64_t n_vertices;
  const double* vertices_xy;            /* [n_vertices][2] */
  int64_t n_boundary_blocks;
  const int32_t* boundary_block_type;   /* DGM_WALL / DGM_FARFIELD, per block */
  const int64_t* boundary_block_count;  /* lines per block */
  const int64_t* boundary_nodes;        /* [sum counts][2] */
  int64_t n_element_blocks;
  const int64_t* element_block_count;   /* triangles per block */
  const int64_t* element_nodes;         /* [sum counts][3], CCW */
  const double* element_xy;             /* nullable */
} dgm_mesh;

/* parameters.json / freestream.json fields (src/io/param_parser.rs:6-19, boundary_condition_parser.rs:6-12). */
typedef struct {
  int32_t order;      /* p in 1..4 (the reference supports p = 2 only) */
  int32_t n_cell_gp;  /* 0 = default: 6 for p <= 2, 16 for p = 3, 25 for p = 4 */
  int32_t n_edge_gp;  /* 0 = default: 3 for p <= 2, 5 for p >= 3 */
  double hcr;         /* gamma */
  double cfl;         /* <= 0: 1/((2p+1)(1+4/(p+2)^2)), src/initialization.rs:29-33 */
  double fs_mach, fs_aoa_deg, fs_rho, fs_p, fs_hcr; /* far-field state, src/initialization.rs:124-133 */
  int32_t precision;  /* DGM_FP64 or DGM_FP32_BULK */
  int32_t limiter;    /* DGM_LIMITER_* */
} dgm_params;

typedef struct {
  int64_t n_elements, n_edges, n_internal_edges, n_internal_elements, n_boundary_elements;
  int32_t n_basis, n_cell_gp, n_edge_gp, order, precision, limiter;
  int32_t n_limiter_levels;  /* level-schedule depth of the sequential sweep */
  int32_t fused_pass;        /* 1: a march evaluates the edge fluxes inside the element pass (FP32-bulk, p = 3,
                                with DGM_FUSED=1 in the environment at context creation; same results) */
  double cfl;
  int64_t device_bytes;      /* device memory held by the context */
} dgm_info;

typedef struct {
  int32_t status;
  int64_t edge;      /* offending edge (reference numbering) or -1 */
  int64_t element;   /* offending element or -1 */
  char message[256];
} dgm_error;

typedef struct {
  int32_t steps;
  double time;
  double last_dt;
  int64_t troubled_cells_last_stage;
} dgm_march_stats;

typedef struct dgm_ctx dgm_ctx;

/* initialize_mesh_basis_gauss_params + initialize_solver (src/initialization.rs:11-111). */
int dgm_ctx_create(const dgm_mesh* mesh, const dgm_params* params, dgm_ctx** out);
/* Same set-up without touching the GPU: connectivity, geometry, tables and the projection helpers work;
 * compute entry points return DGM_BAD_ARGUMENT. For host-side checks on machines without a GPU. */
int dgm_ctx_create_host(const dgm_mesh* mesh, const dgm_params* params, dgm_ctx** out);
void dgm_ctx_destroy(dgm_ctx* ctx);
int dgm_ctx_info(const dgm_ctx* ctx, dgm_info* info);
int dgm_last_error(const dgm_ctx* ctx, dgm_error* err);

/* Connectivity in the reference's numbering (process_mesh, src/initialization.rs:112-261; Mesh::set_neighbours,
 * src/mesh.rs:133-147). Any pointer may be NULL. Sizes: edges [n_edges][2] / [n_edges]; elements [ne][3]. */
int dgm_get_connectivity(const dgm_ctx* ctx, int64_t* edge_elements, int64_t* edge_vertices, int64_t* edge_in_cell,
                         int64_t* edge_patch, int64_t* element_vertices, int64_t* element_edges,
                         int64_t* element_neighbours, int64_t* internal_edges, int64_t* internal_elements,
                         int64_t* boundary_elements);
/* Geometry (src/mesh.rs:15-84): edge normal [n_edges][2], edge length, element jacobian, circumradius, min height. */
int dgm_get_geometry(const dgm_ctx* ctx, double* edge_normal, double* edge_length, double* element_jacob,
                     double* element_circumradius, double* element_minimum_height);
/* Basis and quadrature tables (DubinerBasis::new, GaussPoints::new). Any pointer may be NULL. */
int dgm_get_tables(const dgm_ctx* ctx, double* phi_cell, double* phi_edge, double* mass, double* cell_points,
                   double* cell_weights, double* edge_points, double* edge_weights);

/* Solver::set_initial_solution (src/solver.rs:45-79): L2 projection of a uniform state into U (host). */
int dgm_initial_solution(dgm_ctx* ctx, double density, double aoa_deg, double mach, double pressure, double* U);
/* Extension: L2 projection of a state given at the cell quadrature points, q[ne][n_cell_gp][4]. */
int dgm_project(dgm_ctx* ctx, const double* q_at_points, double* U);
/* Physical coordinates of the cell quadrature points, xy[ne][n_cell_gp][2]. */
int dgm_cell_points_xy(const dgm_ctx* ctx, double* xy);

/* ---- host-buffer entry points (reference layout; each uploads, runs, downloads) ---- */
/* SpatialDisc::compute_residuals. accumulate = 0: R = M^-1 (V - S - B) (overwrite; the caller need not zero R);
 * accumulate = 1: the reference's semantics R = M^-1 (R_in + V - S - B). */
int dgm_compute_residuals(dgm_ctx* ctx, double* residuals, const double* solutions, int32_t accumulate);
/* SpatialDisc::apply_limiter; troubled (nullable, [ne] bytes) receives the cells that were rebuilt. */
int dgm_apply_limiter(dgm_ctx* ctx, double* solutions, uint8_t* troubled);
/* TemperalDisc::compute_time_step (wave speed from the sum of all modes, as the reference). */
int dgm_compute_time_step(dgm_ctx* ctx, const double* solutions, double* dt);
/* TemperalDisc::rk3: one SSP-RK3 step with the limiter after every stage. */
int dgm_rk3(dgm_ctx* ctx, double* solutions, double dt);
/* TemperalDisc::time_march: loop while step < final_step && t < final_time (dt clamped to final_time). */
int dgm_time_march(dgm_ctx* ctx, double* solutions, int32_t final_step, double final_time, dgm_march_stats* stats);

/* ---- device-resident state ---- */
/* Allocation-free residual (extension; no counterpart in the reference). For callers that cannot
 * allocate device memory: harnesses that forbid hipMalloc, tight serving loops. Same result as
 * dgm_compute_residuals with accumulate = 0 (overwrite), bit for bit, but U and R are DEVICE arrays in
 * the reference layout [n_elements][4][n_basis] f64, and the call allocates nothing: its working arrays
 * (the structure-of-arrays state and residual, the edge point fluxes) live in a fixed 128 MiB scratch
 * array that the library holds on each GPU until the process ends. A context whose working set exceeds
 * it (about 150k triangles at p = 3 in FP64) is refused with DGM_NO_SPACE. Runs on the context stream
 * and returns when the result is in R. Calls that overlap on one GPU (contexts used from several
 * threads at once) share the scratch array and race. */
int dgm_compute_residuals_device_noalloc(dgm_ctx* ctx, const double* solutions_dev, double* residuals_dev);
int dgm_state_upload(dgm_ctx* ctx, const double* solutions);
int dgm_state_download(dgm_ctx* ctx, double* solutions);
/* time_march on the resident state: steps run as replays of captured HIP graphs (eight RK3 steps per
 * launch, then single steps; each step tests on the device whether it runs, so steps past the stop
 * condition change nothing). */
int dgm_device_time_march(dgm_ctx* ctx, int32_t final_step, double final_time, dgm_march_stats* stats);
/* enqueue n whole steps (time-step reduction + 3 stages + limiter) without synchronising; n = 0 only
 * captures the step graphs (set-up that a later call would otherwise do first). */
int dgm_device_steps_async(dgm_ctx* ctx, int32_t n_steps);
int dgm_synchronize(dgm_ctx* ctx);
/* the context's HIP stream (hipStream_t), for event timing by the caller */
void* dgm_stream(dgm_ctx* ctx);
/* DGM_NONFINITE if the resident state holds a NaN or Inf */
int dgm_check_state(dgm_ctx* ctx);
/* time and step counters of the resident run */
int dgm_device_clock(dgm_ctx* ctx, double* time, int32_t* steps, double* last_dt);
/* Per-kernel timing: runs n whole steps on the resident state (eagerly, not from the graph) on the
 * context stream and reports the mean milliseconds per launch of each stage part. A part that is one
 * kernel launch is timed by events its dispatch records at its own start and end (hipExtLaunchKernel);
 * a part of several launches (the element-order limiter: its pass, walk and write-back, all in
 * detect_ms) by stream events around it. The step
 * control runs inside the stage-0 edge launch, as in a march: surface_ms includes it, and control_ms
 * (kept for the layout) is 0. */
typedef struct {
  double surface_ms, element_ms, detect_ms, rebuild_ms, apply_ms, time_step_ms, control_ms;
  double step_ms;  /* mean wall time of one step between its first and last event */
  int32_t steps;
} dgm_kernel_times;
int dgm_profile_steps(dgm_ctx* ctx, int32_t n_steps, dgm_kernel_times* out);
/* troubled cells found by the limiter in the last run stage 1, 2, 3 and the stand-alone limiter call */
int dgm_troubled_counts(dgm_ctx* ctx, int32_t counts[4]);
/* Lift coefficient Cl = (Fy cos a - Fx sin a) / (0.5 rho V^2 c) of the wall pressure (extension), chord c. */
int dgm_lift_coefficient(dgm_ctx* ctx, double chord, double* cl);
/* Launch shapes of the kernels (extension; no counterpart in the reference): threads per block of the
 * edge kernel; FP32-bulk element kernel: wavefronts per block of the matrix-core kernel (16 elements
 * each; 1, 2, 4), or 0 for the one-lane-per-element kernel, which the FP64 path always uses; threads per
 * block of the one-lane-per-element kernel and of the detector; XCD-contiguous block order; and
 * wavefronts of the WENO rebuild. */
typedef struct {
  int32_t surface_block, element_waves, element_block, detect_block, remap, weno_grid;
} dgm_tuning;
/* read, set, or choose by timing them on the resident state (dgm_state_upload first; only stage-0
 * passes run, so the state and clock stay as they are) */
int dgm_get_tuning(const dgm_ctx* ctx, dgm_tuning* t);
int dgm_set_tuning(dgm_ctx* ctx, const dgm_tuning* t);
int dgm_autotune(dgm_ctx* ctx, int32_t repeats, dgm_tuning* chosen);

/* ---- batches: many cases of uneven mesh size and order advanced together (extension) ----
 * The reference runs one case per process (src/main.rs, Solver::solve src/solver.rs:80-82); these entry
 * points have no counterpart there. A batch keeps one arena per order p: every case is a disconnected,
 * 64-aligned piece of that arena's element / edge arrays, so each phase of a step is one launch per
 * order over all of its cases (surface, element, detector, rebuild, apply), with per-case time steps,
 * stop conditions and error status. Cases are added and removed between steps (continuous batching);
 * every case gives the same result alone as in any mix. */
typedef struct dgm_batch dgm_batch;
typedef struct {
  int32_t precision;      /* DGM_FP64 / DGM_FP32_BULK */
  int32_t limiter;        /* DGM_LIMITER_OFF or DGM_LIMITER_JACOBI */
  double hcr;             /* ratio of specific heats of every case */
  int32_t max_cases;      /* case slots */
  int32_t pad;
  int64_t capacity[4];    /* element capacity of the arena of order p = index + 1 (0: order not served) */
} dgm_batch_params;
typedef struct {
  int32_t order;          /* 1..4 with the default rules */
  int32_t final_step;     /* stop after this many steps ... */
  double final_time;      /* ... or at this time (TemperalDisc::time_march, src/temporal_disc.rs:21-57) */
  double cfl;             /* <= 0: the reference formula of the order */
  double fs_mach, fs_aoa_deg, fs_rho, fs_p, fs_hcr;  /* far-field and initial state */
} dgm_case_params;
typedef struct {
  int32_t live, active;   /* slot holds a case; the case is still stepping */
  int32_t status;         /* dgm_status of the case (DGM_OK while clean) */
  int32_t steps;
  double time, last_dt;
  int64_t error_edge;     /* case-local edge of a failure, or -1 */
  int32_t troubled_last_stage, order;
  int64_t n_elements, footprint_bytes;
} dgm_case_info;

int dgm_batch_create(const dgm_batch_params* params, dgm_batch** out);
void dgm_batch_destroy(dgm_batch* b);
int dgm_batch_last_error(const dgm_batch* b, dgm_error* err);
/* device bytes a case of this size takes in the batch (its admission footprint) */
int dgm_batch_footprint(const dgm_batch* b, int32_t order, int64_t n_elements, int64_t n_edges, int64_t n_patches,
                        int64_t* bytes);
/* Admits a case: initial state U0 in the reference layout [ne][4][Np] (NULL: the uniform far-field state,
 * Solver::set_initial_solution src/solver.rs:45-79). DGM_NO_SPACE when its arena has no free range now.
 * Its mesh, state and first time step are written beside the steps already queued (which it does not
 * join); it takes its first step with the next dgm_batch_steps_async. dgm_batch_remove is stream-ordered
 * too and does not wait: a removed case's slot and ranges return to use once the steps queued before the
 * removal have run (after a later poll or a synchronisation). */
int dgm_batch_add(dgm_batch* b, const dgm_mesh* mesh, const dgm_case_params* params, const double* U0,
                  int32_t* case_id);
int dgm_batch_remove(dgm_batch* b, int32_t case_id);
/* enqueues n steps of every running case (captured graphs of eight steps, then of one); returns at once;
 * n = 0 only captures the step graphs */
int dgm_batch_steps_async(dgm_batch* b, int32_t n_steps);
int dgm_batch_synchronize(dgm_batch* b);
/* control records of slots 0 .. n-1 (one device read) */
int dgm_batch_case_info(dgm_batch* b, int32_t n, dgm_case_info* out);
int dgm_batch_case_state(dgm_batch* b, int32_t case_id, double* U);
/* lift coefficient of every slot 0 .. n-1 (NaN for empty slots), chord c */
int dgm_batch_lift(dgm_batch* b, double chord, int32_t n, double* cl);
/* Serving support (extension). Enqueues, after the work already queued: every live case's lift coefficient
 * (chord c); for every running case whose coefficient moved by less than cl_tol since the previous poll (0 at
 * admission), a stop after the step it is at, marked converged; and a copy of every slot's record, coefficient
 * and mark into host memory of the batch. Returns at once with a ticket; at most two polls may be outstanding. */
int dgm_batch_poll_async(dgm_batch* b, double chord, double cl_tol, int64_t* ticket);
/* Waits for the oldest outstanding poll (ticket) and reports slots 0 .. n-1 as they were when it ran (any
 * output may be NULL). Cases admitted and removed meanwhile are not reflected; cases admitted while steps
 * run go live between steps, so a poll queued before a case's first step shows it not live. */
int dgm_batch_poll_wait(dgm_batch* b, int64_t ticket, int32_t n, dgm_case_info* info, double* cl, int32_t* converged);
/* device time (ms) of the step launches so far (events around each dgm_batch_steps_async call; completed ones) */
int dgm_batch_device_time(dgm_batch* b, double* ms);
int dgm_batch_get_tuning(const dgm_batch* b, int32_t order, dgm_tuning* t);
int dgm_batch_set_tuning(dgm_batch* b, int32_t order, const dgm_tuning* t);
/* Times the launch shapes of every served order on the resident cases (stage-0 passes only, which
 * leave every case's state and clock unchanged) and keeps the fastest; chosen[p - 1] reports them. */
int dgm_batch_autotune(dgm_batch* b, int32_t repeats, dgm_tuning chosen[4]);
/* Per-kernel timing of the batch (no counterpart in the reference; the batched form of
 * dgm_profile_steps): runs n steps eagerly on the batch stream (the step control once per step, then each
 * stage's parts order by order, each launch timed by events its own dispatch records), then restores
 * every case's state, record and time-step slots. out[p - 1]: mean ms per launch of order p's parts and
 * the mean eager step (all zero for an order the batch does not serve). */
int dgm_batch_profile_steps(dgm_batch* b, int32_t n_steps, dgm_kernel_times out[4]);

#ifdef __cplusplus
}
#endif
#endif
