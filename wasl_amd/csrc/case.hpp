// Host-side case layer over the wasl_hip engine: the reference's case types,
// on-disk formats and serial runner, with the stepping done on the device.
// Names and meanings follow the reference's cpp/include/wasl/case and io
// headers so a caller of run_euler_case / run_scalar_advection_case can switch.

#pragma once

#include <cstdint>
#include <filesystem>
#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "../../include/wasl_hip.h"

namespace wasl_hip {

// ------------------------------------------------------------ formats ----

// include/wasl/mesh/prepared/prepared_mesh.hpp:17-48
struct PreparedMesh {
  int dimension = 0;
  int nodes_per_element = 0;
  std::vector<std::int64_t> node_tags;
  std::vector<double> points;
  std::vector<std::int64_t> element_nodes;
  std::vector<std::int32_t> element_physical_tags;
  std::vector<double> element_metrics;
  std::vector<std::int64_t> face_nodes;
  std::vector<std::int32_t> face_elements;
  std::vector<std::int32_t> face_local_edges;
  std::vector<double> face_normals;
  std::vector<double> face_measures;
  std::vector<std::int32_t> face_physical_tags;
  std::vector<std::int32_t> physical_name_dimensions;
  std::vector<std::int32_t> physical_name_tags;
  std::vector<std::uint64_t> physical_name_offsets;
  std::vector<unsigned char> physical_name_bytes;
  int partition_rank_count = 1;
  std::vector<std::int32_t> element_owners;
  std::vector<std::int32_t> face_owners;

  std::size_t element_count() const noexcept { return element_physical_tags.size(); }
  std::size_t face_count() const noexcept { return face_physical_tags.size(); }
  std::size_t metrics_per_element() const noexcept { return dimension == 1 ? 1U : 5U; }
};

// include/wasl/io/checkpoint.hpp:17-29
struct Checkpoint {
  double time = 0.0;
  std::uint64_t step = 0;
  std::uint64_t variable_count = 0;
  std::uint64_t mode_count = 0;
  std::uint64_t element_count = 0;
  std::string prepared_mesh;
  std::vector<double> coefficients;
  int partition_rank = 0;
  int partition_rank_count = 1;
  std::uint64_t global_element_count = 0;
  std::vector<std::int64_t> element_indices;
};

struct HistoryRow {
  std::uint64_t step;
  double time;
  double time_step;
  double mass;
};

void validate_prepared_mesh(const PreparedMesh &mesh);
PreparedMesh read_prepared_mesh(const std::filesystem::path &path);
void write_prepared_mesh(const std::filesystem::path &path, const PreparedMesh &mesh);
void validate_checkpoint(const Checkpoint &checkpoint);
Checkpoint read_checkpoint(const std::filesystem::path &path);
void write_checkpoint(const std::filesystem::path &path, const Checkpoint &checkpoint);
void write_csv_history(const std::filesystem::path &path, const std::vector<HistoryRow> &rows);
void write_vtu(const std::filesystem::path &path, const PreparedMesh &mesh,
               const std::vector<double> &cell_values, std::size_t component_count);
void write_pvd(const std::filesystem::path &path,
               const std::vector<std::pair<double, std::string>> &snapshots);
std::string numbered_step_name(const std::string &prefix, std::uint64_t step,
                               const std::string &extension);

// Prepared arrays of an engine mesh (prepare_mesh, src/mesh/prepared/
// prepare_mesh.cpp:29-90) with the given physical data, one partition rank.
PreparedMesh prepare_mesh(const wasl_hip_mesh *mesh, std::vector<std::int32_t> element_tags,
                          std::vector<std::int32_t> face_tags,
                          const std::vector<std::pair<std::pair<int, int>, std::string>> &names);

// ---------------------------------------------------------- case types ----

// include/wasl/case/case_files.hpp:9-13
struct CaseFiles {
  std::filesystem::path prepared_mesh;
  std::filesystem::path output_directory;
  std::filesystem::path restart_checkpoint;
};

// include/wasl/case/run_control.hpp:11-19
struct RunControl {
  double start_time = 0.0;
  double end_time = 0.0;
  std::uint64_t max_steps = 0;
  std::uint64_t history_interval = 0;
  std::uint64_t visualization_interval = 0;
  std::uint64_t checkpoint_interval = 0;
  std::uint64_t keep_checkpoints = 0;
};

// include/wasl/case/simulation_summary.hpp:10-14
struct SimulationSummary {
  std::uint64_t step = 0;
  double time = 0.0;
  std::filesystem::path checkpoint;
};

enum class BoundaryKind { periodic = 0, inflow = 1, outflow = 2 };

struct BoundaryCondition {
  BoundaryKind kind = BoundaryKind::periodic;
  double value = 0.0;
};

struct NamedBoundary {
  std::string name;
  BoundaryCondition condition;
};
using NamedBoundaries = std::vector<NamedBoundary>;

struct EulerPrimitive {
  double density = 0.0;
  double velocity_x = 0.0;
  double velocity_y = 0.0;
  double pressure = 0.0;
};

// include/wasl/case/euler_case.hpp:17-42
struct EulerRegion {
  enum class Shape { default_region, interval, circle };
  Shape shape = Shape::default_region;
  double lower = 0.0;
  double upper = 0.0;
  double center_x = 0.0;
  double center_y = 0.0;
  double radius = 0.0;
  EulerPrimitive value{};
};

struct EulerCase {
  CaseFiles files;
  RunControl run;
  bool limiting = false;
  double sensor_threshold = 1e-3;
  double density_floor = 1e-10;
  double pressure_floor = 1e-10;
  int dimensions = 1;
  int polynomial_order = 0;
  double heat_capacity_ratio = 1.4;
  double cfl = 0.0;
  std::vector<EulerRegion> initial_regions;
  NamedBoundaries boundaries;
};

// include/wasl/case/scalar_advection_case.hpp
struct ScalarInterval {
  double lower;
  double upper;
  double value;
};
struct ScalarCircle {
  double center_x;
  double center_y;
  double radius;
  double value;
};

struct ScalarAdvectionCase {
  CaseFiles files;
  RunControl run;
  bool mpi = false;
  int dimensions = 1;
  int polynomial_order = 0;
  double wavespeed = 0.0;
  double velocity_y = 0.0;
  double cfl = 0.0;
  double initial_default = 0.0;
  bool initial_sine = false;
  double initial_sine_amplitude = 0.0;
  std::vector<ScalarInterval> initial_intervals;
  std::vector<ScalarCircle> initial_circles;
  NamedBoundaries boundaries;
};

// include/wasl/case/case_mesh.hpp: the prepared mesh and its canonical path
// (the checkpoint provenance).
struct CaseMesh {
  std::filesystem::path path;
  PreparedMesh prepared;
};

CaseMesh load_case_mesh(const std::filesystem::path &prepared_mesh, int dimensions);

// Where a run stands.
struct RunPosition {
  std::uint64_t step = 0;
  double time = 0.0;
};

// The runner's view of one case, with the state resident on the device
// (include/wasl/case/serial_case_runner.hpp:22-34 keeps it in a host vector).
// Persisted layouts are [variable][mode][element].
struct SerialCaseModel {
  using State = std::vector<double>;

  std::size_t variable_count = 0;
  std::size_t mode_count = 0;
  std::function<void()> initial_state;                   // fresh starts: make it resident
  std::function<void(const State &)> from_checkpoint;    // persisted -> resident
  std::function<State()> to_checkpoint;                  // resident -> persisted
  std::function<double()> mass;                          // history diagnostic
  std::function<State()> cell_averages;                  // packed [element][variable]
  // Steps on the device until end_time or step `limit`, with the CFL dt of
  // every step taken at its start (serial_case_runner.cpp:107-113, 128-133).
  // Returns the dt of the last step taken.
  std::function<double(RunPosition &, std::uint64_t limit, double end_time)> advance_until;
};

SimulationSummary run_serial_case(const CaseFiles &files, const RunControl &run,
                                  const CaseMesh &mesh, const SerialCaseModel &model);

void validate_euler_case(const EulerCase &config);
void validate_scalar_advection_case(const ScalarAdvectionCase &config);

SimulationSummary run_euler_case(const EulerCase &config);
SimulationSummary run_scalar_advection_case(const ScalarAdvectionCase &config);

}  // namespace wasl_hip
