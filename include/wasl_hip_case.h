/*
 * wasl_hip case layer: the bridge's coarse calls over the engine of
 * wasl_hip.h, with the reference's on-disk formats.
 *
 * C mirrors of the cxx bridge functions (crates/wasl-bridge/include/bridge.hpp:7-32,
 * crates/wasl-bridge/src/bridge.cpp:82-216). rust::Str arguments become
 * NUL-terminated strings; rust::Slice arguments become a pointer and a count.
 * Argument order and packing are the bridge's. A call that returns a
 * rust::String here returns a status and hands back a malloc'd string through
 * `summary`, which the caller releases with wasl_hip_string_free.
 *
 * Errors: a negative status, with the message from wasl_hip_case_last_error()
 * (the reference's exception text, which the Rust side would prefix with
 * "run: ", crates/wasl-bridge/src/ops.rs:116,181).
 *
 * The prepared-mesh and checkpoint files are the reference's HDF5 layouts, at
 * format version 2 (src/io/prepared_mesh_file.cpp:34-97,
 * src/io/checkpoint_file.cpp:17-71). The run writes history.csv, solution-*.vtu,
 * solution.pvd and rotated checkpoint-*.h5 like run_serial_case
 * (src/case/serial_case_runner.cpp:48-104).
 *
 * Reference paths are relative to the reference's cpp/ directory unless they
 * start with crates/.
 */
#ifndef WASL_HIP_CASE_H
#define WASL_HIP_CASE_H

#include <stddef.h>
#include <stdint.h>

#include "wasl_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed case-layer call on this thread (never NULL). */
const char *wasl_hip_case_last_error(void);

/* Releases a string returned through `summary`. */
void wasl_hip_string_free(char *text);

/* run_euler_case (crates/wasl-bridge/src/bridge.cpp:173-216): regions are
 * packed 10 per region (shape 0 default / 1 interval / 2 circle, lower, upper,
 * center_x, center_y, radius, rho, u, v, p); boundary kinds 0 periodic,
 * 1 inflow, 2 outflow. Summary "step=<n> time=<%.17g> checkpoint=<path>"
 * (bridge.cpp:40-46). 2D runs need periodic boundaries as in the reference
 * (src/case/euler_triangle.cpp:17-18), or, beyond the reference, every boundary
 * outflow. dimensions = 1 runs Euler on lines (src/case/euler_line.cpp:14-48):
 * periodic or outflow ends; inflow fails with the reference's message. */
int wasl_hip_run_euler_case(const char *mesh_path, const char *output_path,
                            const char *restart_path, int limiting, double sensor_threshold,
                            double density_floor, double pressure_floor, int dimensions,
                            int polynomial_order, double heat_capacity_ratio, double cfl,
                            double start_time, double end_time, uint64_t max_steps,
                            uint64_t history_interval, uint64_t visualization_interval,
                            uint64_t checkpoint_interval, uint64_t keep_checkpoints,
                            const double *regions, size_t region_values,
                            const char *const *boundary_names, const uint8_t *boundary_kinds,
                            const double *boundary_values, size_t boundary_count,
                            char **summary);

/* run_scalar_advection_case (crates/wasl-bridge/src/bridge.cpp:132-171):
 * intervals packed 3 per interval (lower, upper, value), circles 4 per circle
 * (center_x, center_y, radius, value). Serial only (mpi != 0 fails). Lines
 * (dimensions = 1) or periodic triangles (dimensions = 2,
 * src/case/scalar_advection_triangle.cpp:15-46: every boundary periodic, the
 * initial state from the circles, velocity (velocity_x, velocity_y)). */
int wasl_hip_run_scalar_advection_case(
    const char *mesh_path, const char *output_path, const char *restart_path, int mpi,
    int dimensions, int polynomial_order, double velocity_x, double velocity_y, double cfl,
    double start_time, double end_time, uint64_t max_steps, uint64_t history_interval,
    uint64_t visualization_interval, uint64_t checkpoint_interval, uint64_t keep_checkpoints,
    double initial_default, int initial_sine, double initial_sine_amplitude,
    const double *intervals, size_t interval_values, const double *circles, size_t circle_values,
    const char *const *boundary_names, const uint8_t *boundary_kinds,
    const double *boundary_values, size_t boundary_count, char **summary);

/* prepare_mesh + write_prepared_mesh (bridge.cpp:82-92) for a mesh already
 * built from MSH-style records (wasl_hip_mesh_build): no physical tags (-1),
 * no physical names, one partition rank. Summary "prepared <n> elements". */
int wasl_hip_prepare_mesh(const wasl_hip_mesh *mesh, const char *output_path, char **summary);

/* The same with physical tags: element_physical_tags [E] in the mesh's element
 * order, face_physical_tags [F] in its face order (either may be NULL for all
 * -1), and name_count physical names (dimension, tag, NUL-terminated name).
 * Names are stored sorted by (dimension, tag), as a std::map orders them. */
int wasl_hip_prepare_mesh_tagged(const wasl_hip_mesh *mesh, const int32_t *element_physical_tags,
                                 const int32_t *face_physical_tags, size_t name_count,
                                 const int32_t *name_dimensions, const int32_t *name_tags,
                                 const char *const *names, const char *output_path,
                                 char **summary);

/* inspect_prepared_mesh (bridge.cpp:105-109): "<d>D mesh with <n> elements". */
int wasl_hip_inspect_prepared_mesh(const char *mesh_path, char **summary);

/* inspect_checkpoint (bridge.cpp:94-103):
 * "step=<n> time=<%.17g> variables=<v> modes=<m> elements=<e>". */
int wasl_hip_inspect_checkpoint(const char *checkpoint_path, char **summary);

/* export_checkpoint (bridge.cpp:111-131): VTU of the checkpoint's cell
 * averages (c0 / sqrt 2) on the prepared mesh. "exported <n> cells". */
int wasl_hip_export_checkpoint(const char *checkpoint_path, const char *mesh_path,
                               const char *output_path, char **summary);

/* The canonical mesh stored in a prepared-mesh file (read_prepared_mesh, then
 * mesh_from_prepared): faces are taken as stored. Destroy with
 * wasl_hip_mesh_destroy. */
int wasl_hip_mesh_read_prepared(const char *mesh_path, wasl_hip_mesh **out);

/* Checkpoint contents (read_checkpoint, src/io/checkpoint_file.cpp:44-71).
 * Call with coefficients = NULL to learn the sizes, then again with a buffer of
 * variable_count * mode_count * element_count doubles ([variable][mode][element]). */
int wasl_hip_checkpoint_read(const char *checkpoint_path, double *time, uint64_t *step,
                             uint64_t *variable_count, uint64_t *mode_count,
                             uint64_t *element_count, double *coefficients);

/* write_checkpoint for a serial checkpoint of the given layout; prepared_mesh
 * is the provenance path stored in the file. */
int wasl_hip_checkpoint_write(const char *checkpoint_path, double time, uint64_t step,
                              uint64_t variable_count, uint64_t mode_count,
                              uint64_t element_count, const char *prepared_mesh,
                              const double *coefficients);

#ifdef __cplusplus
}
#endif

#endif /* WASL_HIP_CASE_H */
