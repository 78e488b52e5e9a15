/*
 * wasl_hip: MI355X (gfx950) engine for the wasl RKDG hot path.
 *
 * C ABI over the engine. Plain pointers and sizes only; every call returns a
 * status (0 = ok, negative = error) and wasl_hip_last_error() gives the
 * message, which reuses the reference's exception texts where the reference
 * throws (e.g. "simulation: Euler cell average is not physical",
 * cpp/src/time/cfl_timestep.cpp:39-40).
 *
 * Ownership: host buffers belong to the caller; device memory belongs to the
 * engine. One engine drives one HIP stream and is not thread safe; distinct
 * engines may be driven from distinct threads.
 *
 * Layouts are the reference's (docs/NUMERICAL_CONVENTIONS.md):
 *   Euler state   U[variable][mode][element]  (dg/modal_layout.hpp:16-22)
 *   scalar state  U[element][mode]            (dg/scalar_advection/line_residual.hpp:14)
 * Reference paths below are relative to the reference's cpp/ directory
 * unless they start with crates/.
 */
#ifndef WASL_HIP_H
#define WASL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WASL_HIP_OK 0
#define WASL_HIP_ERROR (-1)       /* invalid argument or setup failure */
#define WASL_HIP_SIMULATION (-2)  /* a numerical failure the reference throws for */
#define WASL_HIP_DEVICE (-3)      /* a HIP runtime failure */

/* Boundary kinds, as wasl::BoundaryKind (boundaries/line_boundary_conditions.hpp:24). */
#define WASL_HIP_PERIODIC 0
#define WASL_HIP_INFLOW 1
#define WASL_HIP_OUTFLOW 2

/* Equations. */
#define WASL_HIP_SCALAR_ADVECTION 0
#define WASL_HIP_EULER 1

/* Arithmetic of the stage contractions. */
#define WASL_HIP_FP64 0  /* plain drop-in: everything in FP64 */
#define WASL_HIP_MIXED 1 /* FP32 MFMA contractions, FP64 where the method needs it */
#define WASL_HIP_BF16 2  /* ablation: BF16 MFMA operands, FP32 accumulation */

typedef struct wasl_hip_mesh wasl_hip_mesh;
typedef struct wasl_hip_engine wasl_hip_engine;

/* Message of the last failed call on this thread (never NULL). */
const char *wasl_hip_last_error(void);

/* ------------------------------------------------------------------ mesh --
 * Canonical topology from MSH-style records, replacing
 * wasl::build_triangle_mesh (src/mesh/triangle/build_triangle_mesh.cpp:120-148)
 * and wasl::build_line_mesh (src/mesh/line/build_line_mesh.cpp:69-85):
 * counter-clockwise orientation, shared-edge faces with the smaller element
 * on the left, faces sorted by (left element, left local edge), left outward
 * normals; lines oriented left to right and sorted by position.
 * node_xyz is [node][3]; element_nodes lists node tags per element (3 for
 * triangles, 2 for lines) in record order. */
int wasl_hip_mesh_build(int dimension, int64_t node_count, const int64_t *node_tags,
                        const double *node_xyz, int64_t element_count,
                        const int64_t *element_nodes, wasl_hip_mesh **out);

/* A mesh from the reference's prepared arrays (src/mesh/prepared/
 * mesh_from_prepared.cpp:169-229; PreparedMesh, include/wasl/mesh/prepared/
 * prepared_mesh.hpp:18-48): node_tags/points [node][3], element_nodes are
 * local node indices, face_nodes/face_elements/face_local_edges are [face][2],
 * face_normals [face][dimension]. The faces are taken as given. */
int wasl_hip_mesh_from_prepared(int dimension, int64_t node_count, const int64_t *node_tags,
                                const double *points, int64_t element_count,
                                const int64_t *element_nodes, int64_t face_count,
                                const int64_t *face_nodes, const int32_t *face_elements,
                                const int32_t *face_local_edges, const double *face_normals,
                                wasl_hip_mesh **out);

/* Counts: 0 elements, 1 faces, 2 boundary faces, 3 nodes; 4 gives the dimension. */
int64_t wasl_hip_mesh_count(const wasl_hip_mesh *mesh, int what);

/* Prepared-mesh arrays of the canonical mesh (prepare_mesh, src/mesh/prepared/
 * prepare_mesh.cpp:79-140), node order = first use by elements. Sizes:
 * node_tags [N], points [N][3], element_nodes [E][npe] (local indices),
 * element_metrics [E][5] (triangles: J00 J01 J10 J11 detJ; lines: [E][1]),
 * face_nodes [F][2], face_elements [F][2], face_local_edges [F][2],
 * face_normals [F][dimension], face_measures [F]. Any pointer may be NULL. */
int wasl_hip_mesh_export(const wasl_hip_mesh *mesh, int64_t *node_tags, double *points,
                         int64_t *element_nodes, double *element_metrics, int64_t *face_nodes,
                         int32_t *face_elements, int32_t *face_local_edges, double *face_normals,
                         double *face_measures);

/* Per-face vertex traversal of the triangle topology ([F][4]: left v0, left v1,
 * right v0, right v1; TriangleFace, include/wasl/mesh/triangle/triangle_mesh.hpp:51-68). */
int wasl_hip_mesh_face_vertices(const wasl_hip_mesh *mesh, int64_t *vertices);

void wasl_hip_mesh_destroy(wasl_hip_mesh *mesh);

/* Host-side setup, no device needed.
 * Periodic donor traces of an axis-aligned box, replacing
 * periodic_triangle_trace (src/boundaries/periodic_triangle_trace.cpp:40-67):
 * donors [F][p+1] (-1 on interior faces), basis [F][p+1][Np]. The donor of a
 * point is the lowest-index element containing it (tolerance 1e-10), as the
 * reference's linear search finds it, located here through a bin grid. */
int wasl_hip_mesh_periodic_trace(const wasl_hip_mesh *mesh, int polynomial_order,
                                 int32_t *donors, double *basis);

/* Limiter stencil, replacing make_euler_limiter_stencil for triangles
 * (src/limiting/euler_limiter_stencil_triangle.cpp:97-112). periodic != 0 takes
 * boundary neighbours from the periodic donors; 0 leaves boundary faces
 * without a neighbour. neighbors [E][3], transforms [E][3][Np][Np],
 * normals [E][2], check_basis [C][Np] with C = (p+1)^2 + 3(p+1). */
int wasl_hip_mesh_limiter_stencil(const wasl_hip_mesh *mesh, int polynomial_order, int periodic,
                                  int32_t *neighbors, double *transforms, double *normals,
                                  double *check_basis);

/* L2 projection of an initial condition on the host (see wasl_hip_project
 * for the codes); Euler state U[4][Np][E], scalar U[E][Np]. */
int wasl_hip_mesh_project(const wasl_hip_mesh *mesh, int polynomial_order, double gamma, int ic,
                          const double *params, int64_t param_count, double *state);

/* ---------------------------------------------------------------- engine --
 * Numerics of one case: the fields of wasl::EulerCase / ScalarAdvectionCase
 * (include/wasl/case/euler_case.hpp:28-42, scalar_advection_case.hpp:29-45)
 * that reach the hot path. */
typedef struct {
  int32_t equation;          /* WASL_HIP_EULER: triangles (2D) or lines (1D, euler_line.cpp:14-48,
                                periodic or outflow ends, FP64); WASL_HIP_SCALAR_ADVECTION: lines
                                (1D) or periodic triangles (2D, scalar_advection_triangle.cpp:15-46) */
  int32_t polynomial_order;  /* 0..3 */
  double heat_capacity_ratio;
  double wavespeed;          /* scalar advection a (1D), velocity_x (2D) */
  int32_t boundary_left;     /* 1D: per end. 2D: boundary_left applies to every boundary face
                                (periodic box or transmissive outflow). */
  int32_t boundary_right;
  double boundary_value_left; /* inflow state */
  double boundary_value_right;
  int32_t limiting;
  double sensor_threshold;   /* default 1e-3 */
  double density_floor;      /* default 1e-10 */
  double pressure_floor;     /* default 1e-10 */
  int32_t precision;         /* WASL_HIP_FP64, WASL_HIP_MIXED or WASL_HIP_BF16 */
  int32_t lanes_per_element; /* 0: autotuned tile shape */
  int32_t stage_kernel;      /* 0: autotuned; 1: point split (1, 2 or 4 lanes per element,
                                each lane a share of the quadrature points); 2: variable split
                                (4 or 8 lanes per element, each lane one conserved variable) */
  double velocity_y;         /* 2D scalar advection: y velocity (ScalarAdvectionCase::velocity_y) */
} wasl_hip_numerics;

/* Set fields to the reference defaults (config.rs:85-103). */
void wasl_hip_numerics_default(wasl_hip_numerics *numerics);

/* Builds setup data on the host (periodic donor traces,
 * src/boundaries/periodic_triangle_trace.cpp:40-67; limiter stencil,
 * src/limiting/euler_limiter_stencil_triangle.cpp:97-112; basis and
 * quadrature tables) and uploads it. The mesh may be destroyed afterwards. */
int wasl_hip_engine_create(const wasl_hip_mesh *mesh, const wasl_hip_numerics *numerics,
                           wasl_hip_engine **out);
void wasl_hip_engine_destroy(wasl_hip_engine *engine);

/* Sizes: 0 elements, 1 modes, 2 variables, 3 state doubles, 4 faces,
 * 5 device bytes held, 6 lanes per element in use, 7 limiter check points,
 * 8 stage kernel in use, 9 boundary faces kept as boundary (transmissive, or
 * periodic with per-point donors), 10 conforming periodic face pairs recorded
 * as interior faces. */
int64_t wasl_hip_engine_size(const wasl_hip_engine *engine, int what);

/* Periodic donors [F][p+1] and donor basis [F][p+1][Np]
 * (PeriodicTriangleTrace, include/wasl/boundaries/periodic_triangle_trace.hpp:19-23). */
int wasl_hip_engine_trace(const wasl_hip_engine *engine, int32_t *donors, double *basis);

/* Limiter stencil (EulerLimiterStencil, include/wasl/limiting/euler_limiter_stencil.hpp:19-50):
 * neighbors [E][K], transforms [E][K][Np][Np], normals [E][2] with K = 3 on
 * triangles and 2 on lines (euler_limiter_stencil_line.cpp:79-94),
 * check_basis [C][Np] with C = wasl_hip_engine_size(engine, 7). */
int wasl_hip_engine_stencil(const wasl_hip_engine *engine, int32_t *neighbors, double *transforms,
                            double *normals, double *check_basis);

/* L2 projection of an initial condition (src/case/initial_state.cpp:16-64).
 * ic 0: Euler regions, params k*10 (shape, lower, upper, cx, cy, r, rho, u, v, p)
 *       as packed by the bridge (crates/wasl-bridge/src/bridge.cpp:192-214);
 * ic 1: isentropic vortex (strength, cx, cy; background rho=u=v=p=1);
 * ic 2: radial rarefaction (rho, p, speed);
 * ic 3: smooth field (L, then 4 x (a, m, n, phase));
 * ic 10: scalar pulse sin(2 pi x) + exp(-((x-0.5)/0.05)^2);
 * ic 11: scalar regions (default, sine flag, amplitude, k*3 intervals);
 * ic 12: 2D scalar circles (default, k*4 (cx, cy, r, value)),
 *        src/case/initial_regions.cpp:28-34;
 * ic 13: 2D scalar plane wave (L, kx, ky): sin(2 pi (kx x + ky y) / L).
 * Codes 0-3 need an Euler engine, 10-11 a line engine, 12-13 a 2D scalar one.
 * The result is written to state_out (host) and made the resident state. */
int wasl_hip_project(wasl_hip_engine *engine, int ic, const double *params, int64_t param_count,
                     double *state_out);

int wasl_hip_upload_state(wasl_hip_engine *engine, const double *state);
int wasl_hip_download_state(wasl_hip_engine *engine, double *state);

/* One residual evaluation R(state) with the case's boundary exterior,
 * replacing euler_triangle_residual<hllc_flux> (dg/euler/triangle_residual.hpp:25-107)
 * with periodic_triangle_euler_exterior (src/boundaries/periodic_triangle_trace.cpp:84-105),
 * scalar_advection_residual (dg/scalar_advection/line_residual.hpp:48-111)
 * with make_line_boundary_exterior (boundaries/line_boundary_conditions.hpp:36-57),
 * or triangle_scalar_advection_residual (dg/scalar_advection/triangle_residual.hpp:25-105)
 * with periodic_triangle_exterior (src/boundaries/periodic_triangle_trace.cpp:69-82).
 * Host buffers in and out; the resident state is not touched. */
int wasl_hip_residual(wasl_hip_engine *engine, const double *state, double *rhs);

/* limit_euler_modal on the resident state in place
 * (src/limiting/euler_modal_limiter.cpp:70-98); flags [E] receives the sensor
 * decision (may be NULL). */
int wasl_hip_limit(wasl_hip_engine *engine, uint8_t *flags);

/* wasl_hip_limit on a caller's DEVICE buffer, allocating nothing. No
 * counterpart in the reference: it is for callers that cannot allocate or
 * synchronise (code under HIP graph capture, tight serving loops that drive
 * their own stages), e.g. as the post_stage of a caller's own RK loop.
 *   state   device U[variable][mode][element] of this engine's mesh, limited
 *           in place with the same kernels and results as wasl_hip_limit;
 *   flags   device uint8 [E] receiving the sensor decision, or NULL;
 *   status  device int receiving WASL_HIP_OK, or WASL_HIP_SIMULATION when a
 *           cell average is below a floor (the state then holds the elements
 *           before the first such element limited, as the reference leaves
 *           it), or NULL;
 *   stream  hipStream_t the work is enqueued on (NULL: the engine's stream).
 * The call only enqueues kernels: it calls no hipMalloc/hipFree and does not
 * wait. Its working memory is one static 64 MiB device array per GPU
 * (error word, troubled list, staged values: 256 + 4 E + 32 Np E bytes,
 * e.g. at most about 205,000 elements at P3 or 340,000 at P2); larger meshes
 * are rejected with an error. Every call on a GPU shares that array, so calls
 * that overlap on one GPU (other streams, other threads) race; calls in one
 * stream order are safe. The engine must have been created with limiting on.
 */
int wasl_hip_limit_noalloc(wasl_hip_engine *engine, double *state, uint8_t *flags, int *status,
                           void *stream);

/* euler_cfl_timestep / advection_cfl_timestep / triangle_scalar_cfl_timestep
 * (src/time/cfl_timestep.cpp:53-84) of the resident state. */
int wasl_hip_stable_dt(wasl_hip_engine *engine, double cfl, double *dt);

/* One SSP-RK3 step of the resident state (time/ssp_rk3.hpp:37-61), the
 * limiter after every stage when enabled. */
int wasl_hip_step(wasl_hip_engine *engine, double dt);

/* The stepping loop of run_serial_case (src/case/serial_case_runner.cpp:117-149)
 * on the device: dt from the CFL condition at the start of every step, clipped
 * to land on end_time (:107-113), without a host round trip per step. Starts
 * from (*time, *step); returns the final time and step. history [rows][4]
 * (step, time, dt, mass) receives the initial row and every history_interval-th
 * and the final step when history is not NULL (capacity in rows). Reaching
 * max_steps before end_time fails with the reference's message
 * ("simulation: max_steps reached before end_time"); *time and *step still
 * hold the position reached. */
int wasl_hip_advance(wasl_hip_engine *engine, double cfl, double end_time, uint64_t max_steps,
                     uint64_t history_interval, double *time, uint64_t *step, double *history,
                     int64_t history_capacity, int64_t *history_rows);

/* euler_density_mass / scalar_mass (src/case/euler_diagnostics.cpp:30-36,
 * scalar_diagnostics.cpp:18-25) of the resident state, summed in FP64 with
 * compensation. */
int wasl_hip_mass(wasl_hip_engine *engine, double *mass);

/* Cell averages packed [element][variable] (euler_averages,
 * src/case/euler_diagnostics.cpp:10-21; scalar_cell_averages). */
int wasl_hip_cell_averages(wasl_hip_engine *engine, double *averages);

/* Device time of the last wasl_hip_advance/step call in milliseconds, and
 * the average duration of the fused stage kernel over that call (HIP events
 * on the engine's stream). */
int wasl_hip_last_timing(const wasl_hip_engine *engine, double *wall_ms, double *stage_ms,
                         int64_t *stage_launches);

/* No reference counterpart (measurement only): when on, wasl_hip_advance
 * records a HIP event pair around every stage kernel so
 * wasl_hip_last_timing can report the stage kernel's average duration. */
int wasl_hip_set_timing(wasl_hip_engine *engine, int on);

/* Not part of the reference interface. wasl_hip_advance replays its steps
 * from captured HIP graphs of 1, 2, 4, ..., 64 steps at either step parity,
 * captured when first needed; this captures them all now (setup), so that no
 * capture falls inside a caller's timed or latency-critical loop. */
int wasl_hip_capture_graphs(wasl_hip_engine *engine);

/* Not part of the reference interface: the stage kernel's average duration
 * in microseconds per stage (stage_us[3]), each stage launched reps times
 * back to back from a captured graph with fixed dt and timed with HIP events
 * on the engine's stream (benchmarks). The resident state is not meaningful
 * afterwards. */
int wasl_hip_time_stages(wasl_hip_engine *engine, int reps, double dt, double *stage_us);

/* Not part of the reference interface (measurement only): the FP64 rates in
 * TFLOP/s that the current device sustains on its vector ALU (v_fma_f64
 * chains) and on its matrix cores (v_mfma_f64_16x16x4f64 chains), each the
 * best of three timed launches; the benchmark's roofline reports them. */
int wasl_hip_fp64_ceilings(double *valu_tflops, double *mfma_tflops);

#ifdef __cplusplus
}
#endif

#endif
