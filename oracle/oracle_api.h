/*
 * C API shared by the two CPU oracles of the RKDG hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (wasl_amd/, include/)
 * links or calls this; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg do, and only as the checker.
 *
 * Two libraries implement this header:
 *   oracle/liboracle.so        the project's own single-threaded C++
 *                              restatement of the reference algorithms
 *                              (oracle/cpu_ref.cpp);
 *   oracle/_ref/libwaslref.so  the reference's own sources compiled from
 *                              /root/reference/cpp by oracle/Makefile and
 *                              driven through oracle/ref_bind.cpp.
 * Both are built with -ffp-contract=off and no -march, matching the
 * reference's default CMake build, so their results can be compared bit for
 * bit.
 *
 * Conventions follow the reference:
 *   Euler state  U[variable][mode][element]   (cpp/include/wasl/dg/modal_layout.hpp:1-22)
 *   scalar state U[element][mode]             (dg/scalar_advection/line_residual.hpp:14)
 * Every function returns 0 on success and -1 on failure; orc_error() then
 * returns the exception text the reference throws.
 */
#ifndef WASL_ORACLE_API_H
#define WASL_ORACLE_API_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_case orc_case;

/* Boundary kinds (reference BoundaryKind, line_boundary_conditions.hpp:24). */
enum { ORC_PERIODIC = 0, ORC_INFLOW = 1, ORC_OUTFLOW = 2 };

/* Initial conditions understood by orc_project. */
enum {
  ORC_IC_EULER_REGIONS = 0,  /* params: k*10 doubles (shape,lower,upper,cx,cy,r,rho,u,v,p), bridge.cpp:192 */
  ORC_IC_VORTEX = 1,         /* params: strength, cx, cy (background rho=u=v=p=1) */
  ORC_IC_RAREFACTION = 2,    /* params: rho, p, speed (velocity speed * r_hat, zero at origin) */
  ORC_IC_SMOOTH = 3,         /* params: L, then 4 x (a, m, n, phase): rho = 1 + 0.1 sum a sin(2pi(mx+ny)/L+phase), p=1, u=v=0.5 */
  ORC_IC_SCALAR_PULSE = 10,  /* sin(2 pi x) + exp(-((x-0.5)/0.05)^2) */
  ORC_IC_SCALAR_REGIONS = 11, /* params: default, sine(0/1), amplitude, k*3 (lower, upper, value) */
  ORC_IC_SCALAR_CIRCLES = 12, /* 2D: default, k*4 (cx, cy, r, value), initial_regions.cpp:28-34 */
  ORC_IC_SCALAR_WAVE = 13     /* 2D: L, kx, ky: sin(2 pi (kx x + ky y) / L) */
};

const char *orc_error(void);

/* 2D Euler on straight-sided triangles. Nodes carry tags and (x, y);
 * triangles list three node tags in file order (MSH records). boundary is
 * ORC_PERIODIC (axis-aligned box, periodic donor traces) or ORC_OUTFLOW
 * (transmissive: exterior = own trace, no limiter neighbour across). */
orc_case *orc_tri_euler_create(int64_t node_count, const int64_t *node_tags, const double *node_xy,
                               int64_t triangle_count, const int64_t *triangle_nodes, int order,
                               double gamma, int boundary);

/* 2D scalar advection on periodic triangles with constant velocity
 * (scalar_advection_triangle.cpp:15-46). State U[element][mode]; no limiter. */
orc_case *orc_tri_scalar_create(int64_t node_count, const int64_t *node_tags, const double *node_xy,
                                int64_t triangle_count, const int64_t *triangle_nodes, int order,
                                double velocity_x, double velocity_y);

/* 1D Euler on a line mesh (euler_line.cpp:14-48): periodic or outflow ends
 * (ORC_INFLOW throws at the first residual, as line_euler_exterior does). */
orc_case *orc_line_euler_create(int64_t node_count, const int64_t *node_tags,
                                const double *node_x, int64_t segment_count,
                                const int64_t *segment_nodes, int order, double gamma,
                                int left_kind, int right_kind);

/* 1D scalar advection on a line mesh. Segments list two node tags. */
orc_case *orc_line_scalar_create(int64_t node_count, const int64_t *node_tags,
                                 const double *node_x, int64_t segment_count,
                                 const int64_t *segment_nodes, int order, double wavespeed,
                                 int left_kind, double left_value, int right_kind,
                                 double right_value);

void orc_destroy(orc_case *c);

/* Sizes: 0 elements, 1 faces, 2 modes, 3 variables, 4 edge points (p+1). */
int64_t orc_size(const orc_case *c, int what);

/* Canonical triangle topology (build_triangle_mesh.cpp:38-148).
 * element_nodes [E][3] tags after CCW orientation; element_xy [E][6].
 * face_elements [F][2] (left, right or -1); face_edges [F][2];
 * face_vertices [F][4] (left v0, left v1, right v0, right v1);
 * face_normals [F][2]. */
int orc_tri_topology(const orc_case *c, int64_t *element_nodes, double *element_xy,
                     int32_t *face_elements, int32_t *face_edges, int64_t *face_vertices,
                     double *face_normals);

/* Line topology (build_line_mesh.cpp:15-85): element_x [E][2] (left,right);
 * face_elements [F][2]; face_normals [F]. */
int orc_line_topology(const orc_case *c, double *element_x, int32_t *face_elements,
                      double *face_normals);

/* Periodic donor traces (periodic_triangle_trace.cpp:40-67):
 * donors [F][p+1] (-1 on interior faces), basis [F][p+1][Np]. */
int orc_tri_trace(const orc_case *c, int32_t *donors, double *basis);

/* Limiter stencil (euler_limiter_stencil_*.cpp): neighbors [E][K],
 * transforms [E][K][Np][Np], normals [E][2], check_basis [C][Np].
 * K = 3 (triangles) or 2 (lines); C from orc_stencil_check_count. Line
 * scalar cases have no stencil (C = 0). */
int64_t orc_stencil_check_count(const orc_case *c);
int orc_stencil(const orc_case *c, int32_t *neighbors, double *transforms, double *normals,
                double *check_basis);

/* One residual evaluation with the case's own boundary exterior
 * (euler_triangle_residual / scalar_advection_residual). */
int orc_residual(const orc_case *c, const double *state, double *rhs);

/* Troubled-cell WENO + positivity limiter in place (euler_modal_limiter.cpp:70-98).
 * flags [E] receives the sensor decision (1 troubled). */
int orc_limit(const orc_case *c, double threshold, double density_floor, double pressure_floor,
              double *state, uint8_t *flags);

/* Stable time step (cfl_timestep.cpp). */
int orc_stable_dt(const orc_case *c, const double *state, double cfl, double *dt);

/* One SSP-RK3 step (ssp_rk3.hpp:37-61), limiter after every stage if enabled. */
int orc_step(orc_case *c, double dt, int limiting, double threshold, double density_floor,
             double pressure_floor, double *state);

/* Density mass (Euler, euler_diagnostics.cpp:30-36) or scalar mass
 * (scalar_diagnostics.cpp:18-25). */
int orc_mass(const orc_case *c, const double *state, double *mass);

/* L2 projection of an initial condition (initial_state.cpp:16-64). The case
 * drivers limit the projected Euler state once when limiting is on
 * (euler_triangle.cpp:37-41); callers do that with orc_limit. */
int orc_project(const orc_case *c, int ic, const double *params, int64_t param_count,
                double *state);

/* The serial runner loop (serial_case_runner.cpp:117-149) starting from
 * *state at start_time/start_step; writes the final state back and the
 * history rows (step, time, dt, mass) at history_interval and the final step
 * into history [rows][4] (capacity history_capacity rows). */
int orc_run(orc_case *c, double cfl, int limiting, double threshold, double density_floor,
            double pressure_floor, double start_time, uint64_t start_step, double end_time,
            uint64_t max_steps, uint64_t history_interval, double *state, uint64_t *final_step,
            double *final_time, double *history, int64_t history_capacity, int64_t *history_rows);

/* ---- function-level hooks (known-answer tests of single routines) ---- */

/* Basis values: kind 0 Legendre (xi) with p+1 values, kind 1 Dubiner
 * (xi, eta) with (p+1)(p+2)/2 values; gradient != 0 fills gx (and ge). */
int orc_basis(int kind, int p, double xi, double eta, int gradient, double *values, double *gx,
              double *ge);
/* Rules: kind 0 Gauss-Legendre n points (x, w); kind 1 triangle volume rule
 * of order n ((n+1)^2 points: xi, eta, w); kind 2 edge rule of order n for
 * edge `edge` (xi, eta, w). Returns the point count (or -1). */
int orc_rule(int kind, int n, int edge, double *a, double *b, double *w);
/* Euler physics on conserved states [4]. */
int orc_euler_flux(double gamma, const double *state, double *fx, double *fy);
int orc_hllc(double gamma, const double *left, const double *right, double nx, double ny,
             double *flux);
int orc_conserved(double gamma, const double *primitive, double *state);
/* Characteristic transform of change[4] frozen at average[4] and back. */
int orc_characteristic(double gamma, const double *average, const double *change, double nx,
                       double ny, double *waves, double *back);
/* simple_weno on one field: target [n], candidates [K][n] -> limited [n]. */
int orc_simple_weno(const double *target, const double *candidates, int K, int n,
                    double *limited);
/* euler_characteristic_weno: target [4][n], candidates [K][4][n]. */
int orc_characteristic_weno(double gamma, const double *average, double nx, double ny,
                            const double *target, const double *candidates, int K, int n,
                            double *limited);
/* limit_euler_positivity on packed [4][n] coefficients with check basis
 * [points][n]; returns the total scale in *scale. */
int orc_positivity(double gamma, double *coefficients, int n, const double *basis, int points,
                   double density_floor, double pressure_floor, double *scale);
/* modal_decay_troubled of one contiguous triangle polynomial of order p. */
int orc_modal_decay_troubled(const double *coefficients, int p, double threshold, int *troubled);

#ifdef __cplusplus
}
#endif

#endif
