/*
 * wasl_hip serving: many resident cases advanced together on one GPU.
 *
 * No counterpart in the reference, which runs one case per process
 * (src/case/serial_case_runner.cpp:117-149). A server takes requests (a mesh,
 * numerics as for an engine, an initial condition and an end time), prepares
 * each on a host worker thread, and admits prepared cases while device slots
 * and device memory allow (continuous batching). Each scheduling round
 * (wasl_hip_server_pump) advances every resident case by up to steps_per_batch
 * SSP-RK3 steps with one batched kernel launch per stage and class, each case
 * taking its own CFL dt on the device and stopping at its own end time;
 * finished cases are retired and their slots reused. A captured HIP graph per
 * class serves every round until the class's membership changes.
 *
 * A case computes exactly what it computes alone: tiles of the batched
 * kernels are the blocks of the single-case kernel, so results do not depend
 * on the other resident cases.
 *
 * 2D Euler cases run the stage in the server's precision (WASL_HIP_MIXED by
 * default, or WASL_HIP_BF16); the limiter, the CFL speed and the state stay
 * FP64. Two kinds of 2D Euler case run the FP64 stage instead (each is a
 * config switch): orders up to fp64_max_order, where the matrix-core form
 * has no speed advantage, and cases with limiting, whose positivity
 * guarantee needs FP64 point values near vacuum. 1D advection cases run in
 * FP64. Cases of one (dimension, order, stage precision) form a class;
 * classes run concurrently on streams of their own.
 *
 * Cases with limiting are limited once after their projection, as the
 * reference's case drivers do (src/case/euler_triangle.cpp:37-41).
 *
 * Threading: the server owns its worker threads. Its calls may come from
 * several threads; they are serialised (a submit made while another thread
 * is inside wasl_hip_server_pump waits for that round to end).
 */
#ifndef WASL_HIP_SERVE_H
#define WASL_HIP_SERVE_H

#include <stdint.h>

#include "wasl_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wasl_hip_server wasl_hip_server;

typedef struct {
  int32_t precision;        /* stage arithmetic of 2D Euler cases: WASL_HIP_MIXED or WASL_HIP_BF16 */
  int32_t max_resident;     /* device slots (default 256) */
  uint64_t memory_budget;   /* device bytes for prepared and resident cases; 0: 80% of free */
  int32_t steps_per_batch;  /* steps per scheduling round, even (default 16) */
  int32_t use_graphs;       /* capture one graph per resident set (default 1) */
  int32_t setup_threads;    /* host threads preparing cases (default 4, at most 16) */
  int32_t fp64_max_order;   /* 2D Euler orders up to this run the FP64 stage (default 1; -1 none) */
  int32_t fp64_limited;     /* 2D Euler cases with limiting run the FP64 stage (default 1) */
  /* Tile of the FP64 2D Euler stage by order, {stage_kernel, lanes_per_element}
   * as in wasl_hip_numerics, for served cases without limiting whose numerics
   * leave the tile to the autotuner (limited cases keep the autotuned tile, so
   * that their troubled-cell decisions are those of a single engine). A served class runs one launch over all its cases, so the
   * tile that suits a large mesh can beat the one a small case's autotuner
   * picks. {0, 0}: the engine's autotuned tile. Defaults: P0 {0, 0}, P1 {1, 2},
   * P2 {1, 2}, P3 {2, 4}. */
  int32_t fp64_tile[4][2];
} wasl_hip_server_config;

void wasl_hip_server_config_default(wasl_hip_server_config *config);
int wasl_hip_server_create(const wasl_hip_server_config *config, wasl_hip_server **out);
/* Waits for the worker threads; pending and unclaimed requests are dropped. */
void wasl_hip_server_destroy(wasl_hip_server *server);

/* Queues one case. The mesh is copied (it may be destroyed afterwards); the
 * numerics are those of wasl_hip_engine_create (the server's precision applies
 * to 2D Euler); ic/params as wasl_hip_project. The case runs to end_time, or,
 * when end_steps > 0, to end_steps times the CFL step of its projected (and,
 * with limiting, limited) initial state, at most to end_time when end_time is
 * positive as well; max_steps bounds the step count. */
int wasl_hip_server_submit(wasl_hip_server *server, const wasl_hip_mesh *mesh,
                           const wasl_hip_numerics *numerics, int ic, const double *params,
                           int64_t param_count, double cfl, double end_time, uint64_t end_steps,
                           uint64_t max_steps, uint64_t *id);

/* One scheduling round, pipelined with the device: wait for the round the
 * previous call launched and take its finished cases out of the resident set,
 * start preparing queued cases that fit the budget, admit prepared ones,
 * launch the next round of up to steps_per_batch steps without waiting for it,
 * and retire the finished cases (state copied out, engine released) while it
 * runs. Returns up to capacity finished request ids. When nothing is resident
 * it waits (up to 50 ms) for a preparation to finish. */
int wasl_hip_server_pump(wasl_hip_server *server, uint64_t *completed, int64_t capacity,
                         int64_t *count);

/* Outcome of a finished request: the final state (the engine's layout; pass
 * NULL to skip), time and step, and the case's status (0, or the engine
 * status of its failure, with the message from wasl_hip_last_error()).
 * Releases the request. */
int wasl_hip_server_result(wasl_hip_server *server, uint64_t id, double *state,
                           int64_t state_capacity, double *time, uint64_t *step,
                           int *case_status);

/* Size of a request's state in doubles (-1 if unknown). */
int64_t wasl_hip_server_state_size(const wasl_hip_server *server, uint64_t id);

/* Counters: 0 queued, 1 preparing or prepared, 2 resident, 3 finished and
 * unclaimed, 4 device bytes reserved, 5 graphs captured, 6 rounds run,
 * 7 resident-set changes, 8 microseconds spent rebuilding launch tables and
 * graphs, 9 microseconds spent waiting for rounds to complete, 10
 * microseconds of case preparation (summed over the setup threads),
 * microseconds spent 11 admitting prepared cases, 12 retiring finished ones,
 * 13 waiting for a preparation with nothing resident, 14 in submit, 15
 * releasing retired cases' device memory (part of 12). */
int64_t wasl_hip_server_count(const wasl_hip_server *server, int what);

#ifdef __cplusplus
}
#endif

#endif /* WASL_HIP_SERVE_H */
