// The case layer's drivers: configuration checks, the serial runner with the
// stepping on the device, the Euler-on-triangles and advection-on-lines
// drivers, and the extern "C" mirrors of the bridge's coarse calls.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>

#include "../../include/wasl_hip_case.h"
#include "case.hpp"

namespace wasl_hip {
namespace {

constexpr const char *kMaxStepsMessage = "simulation: max_steps reached before end_time";
constexpr int kMaximumOrder = 3;  // src/case/case_validation.cpp:12

// Engine calls raise the engine's message.
void call(int status) {
  if (status != WASL_HIP_OK)
    throw std::runtime_error(wasl_hip_last_error());
}

struct MeshDeleter {
  void operator()(wasl_hip_mesh *m) const { wasl_hip_mesh_destroy(m); }
};
struct EngineDeleter {
  void operator()(wasl_hip_engine *g) const { wasl_hip_engine_destroy(g); }
};
using MeshPtr = std::unique_ptr<wasl_hip_mesh, MeshDeleter>;
using EnginePtr = std::unique_ptr<wasl_hip_engine, EngineDeleter>;

MeshPtr engine_mesh(const PreparedMesh &p) {
  wasl_hip_mesh *m = nullptr;
  call(wasl_hip_mesh_from_prepared(p.dimension, static_cast<int64_t>(p.node_tags.size()),
                                   p.node_tags.data(), p.points.data(),
                                   static_cast<int64_t>(p.element_count()), p.element_nodes.data(),
                                   static_cast<int64_t>(p.face_count()), p.face_nodes.data(),
                                   p.face_elements.data(), p.face_local_edges.data(),
                                   p.face_normals.data(), &m));
  return MeshPtr(m);
}

// ------------------------------------------------------- validation ----
// src/case/case_validation.cpp:14-58, with its messages.

bool valid_files(const CaseFiles &f) {
  return !f.prepared_mesh.empty() && !f.output_directory.empty();
}
bool valid_order(int p) { return p >= 0 && p <= kMaximumOrder; }
bool valid_run(const RunControl &r) {
  return std::isfinite(r.start_time) && std::isfinite(r.end_time) && r.end_time > r.start_time &&
         r.max_steps != 0 && r.history_interval != 0 && r.visualization_interval != 0 &&
         r.checkpoint_interval != 0 && r.keep_checkpoints != 0;
}
bool positive_finite(double v) { return std::isfinite(v) && v > 0.0; }
bool nonnegative_finite(double v) { return std::isfinite(v) && v >= 0.0; }
bool valid_interval(double a, double b) { return std::isfinite(a) && std::isfinite(b) && a < b; }
bool valid_circle(double x, double y, double r) {
  return std::isfinite(x) && std::isfinite(y) && positive_finite(r);
}

void validate_boundaries(const NamedBoundaries &boundaries) {
  std::set<std::string> names;
  for (const auto &b : boundaries) {
    if (b.name.empty() || !names.insert(b.name).second)
      throw std::runtime_error("simulation: boundary names must be unique and nonempty");
    if (!std::isfinite(b.condition.value))
      throw std::runtime_error("simulation: boundary " + b.name + " value must be finite");
  }
}

bool all_of_kind(const NamedBoundaries &boundaries, BoundaryKind kind) {
  return std::all_of(boundaries.begin(), boundaries.end(),
                     [kind](const NamedBoundary &b) { return b.condition.kind == kind; });
}

// ------------------------------------------------------------- runner ----

// Output events at a step: a history row, a VTU snapshot, a checkpoint.
enum OutputEvent : unsigned { kHistory = 1u, kSnapshot = 2u, kCheckpoint = 4u };

// The output cadence of a run, fixed by its run control: every multiple of
// an interval, and every kind of output at the final step (end_time or
// max_steps), as src/case/serial_case_runner.cpp:132-141 emits them. The
// device loop runs uninterrupted from one stop to the next.
class OutputPlan {
public:
  explicit OutputPlan(const RunControl &run) : run_(run) {}

  bool running(const RunPosition &at) const {
    return at.time < run_.end_time && at.step < run_.max_steps;
  }

  // The first step after `step` at which some output falls due (or
  // max_steps): where the device loop next hands control back.
  std::uint64_t next_stop(std::uint64_t step) const {
    const auto after = [step](std::uint64_t every) { return step - step % every + every; };
    return std::min({run_.max_steps, after(run_.history_interval),
                     after(run_.visualization_interval), after(run_.checkpoint_interval)});
  }

  unsigned events(const RunPosition &at) const {
    const bool last = at.time >= run_.end_time || at.step == run_.max_steps;
    const auto due = [&](std::uint64_t every, unsigned bit) {
      return (last || at.step % every == 0) ? bit : 0u;
    };
    return due(run_.history_interval, kHistory) | due(run_.visualization_interval, kSnapshot) |
           due(run_.checkpoint_interval, kCheckpoint);
  }

private:
  const RunControl &run_;
};

// The files a run leaves in its output directory: history.csv and
// solution.pvd written at the end from the rows and snapshots gathered, the
// solution_<step>.vtu snapshots, and the newest keep checkpoint_<step>.h5
// files, the older ones deleted as new ones arrive (a ring of keep slots).
class CaseOutputs {
public:
  CaseOutputs(const CaseFiles &files, const CaseMesh &mesh, const SerialCaseModel &model,
              std::uint64_t keep)
      : dir_(files.output_directory), mesh_(mesh), model_(model), keep_(keep) {
    std::filesystem::create_directories(dir_);
  }

  void emit(unsigned events, const RunPosition &at, double dt) {
    if (events & kHistory)
      rows_.push_back({at.step, at.time, dt, model_.mass()});
    if (events & kSnapshot) {
      const std::string vtu = numbered_step_name("solution", at.step, ".vtu");
      write_vtu(dir_ / vtu, mesh_.prepared, model_.cell_averages(), model_.variable_count);
      pvd_.emplace_back(at.time, vtu);
    }
    if (events & kCheckpoint)
      checkpoint(at);
  }

  // The run's last checkpoint (one at the final position if the cadence
  // wrote none), then the history and collection files.
  std::filesystem::path close(const RunPosition &at) {
    if (newest_.empty())
      checkpoint(at);
    write_csv_history(dir_ / "history.csv", rows_);
    write_pvd(dir_ / "solution.pvd", pvd_);
    return newest_;
  }

private:
  void checkpoint(const RunPosition &at) {
    Checkpoint c;
    c.time = at.time;
    c.step = at.step;
    c.variable_count = model_.variable_count;
    c.mode_count = model_.mode_count;
    c.element_count = mesh_.prepared.element_count();
    c.prepared_mesh = mesh_.path.string();
    c.coefficients = model_.to_checkpoint();
    newest_ = dir_ / numbered_step_name("checkpoint", at.step, ".h5");
    write_checkpoint(newest_, c);
    if (ring_.size() < keep_) {
      ring_.push_back(newest_);
    } else {
      std::filesystem::remove(ring_[oldest_]);
      ring_[oldest_] = newest_;
      oldest_ = (oldest_ + 1) % ring_.size();
    }
  }

  std::filesystem::path dir_;
  const CaseMesh &mesh_;
  const SerialCaseModel &model_;
  std::uint64_t keep_;
  std::vector<HistoryRow> rows_;
  std::vector<std::pair<double, std::string>> pvd_;
  std::vector<std::filesystem::path> ring_;   // checkpoints on disk
  std::size_t oldest_ = 0;                    // ring slot written longest ago
  std::filesystem::path newest_;
};

// A device engine behind the model's callbacks.
struct DeviceCase {
  EnginePtr engine;
  double cfl = 0.0;
  int V = 1, Np = 1, E = 0;
  bool scalar = false;

  std::vector<double> download() const {
    std::vector<double> U(static_cast<std::size_t>(V) * Np * E);
    call(wasl_hip_download_state(engine.get(), U.data()));
    return U;
  }

  // Scalar solver state is [element][mode]; checkpoints persist
  // [variable][mode][element] (src/case/scalar_coefficient_layout.cpp:5-21).
  std::vector<double> persisted(const std::vector<double> &solver) const {
    if (!scalar)
      return solver;
    std::vector<double> out(solver.size());
    for (int e = 0; e < E; ++e)
      for (int m = 0; m < Np; ++m)
        out[static_cast<std::size_t>(m) * E + e] = solver[static_cast<std::size_t>(e) * Np + m];
    return out;
  }
  std::vector<double> solver(const std::vector<double> &stored) const {
    if (!scalar)
      return stored;
    std::vector<double> out(stored.size());
    for (int e = 0; e < E; ++e)
      for (int m = 0; m < Np; ++m)
        out[static_cast<std::size_t>(e) * Np + m] = stored[static_cast<std::size_t>(m) * E + e];
    return out;
  }

  void bind(SerialCaseModel &model) {
    model.variable_count = static_cast<std::size_t>(V);
    model.mode_count = static_cast<std::size_t>(Np);
    model.from_checkpoint = [this](const SerialCaseModel::State &stored) {
      const auto s = solver(stored);
      call(wasl_hip_upload_state(engine.get(), s.data()));
    };
    model.to_checkpoint = [this] { return persisted(download()); };
    model.mass = [this] {
      double m = 0.0;
      call(wasl_hip_mass(engine.get(), &m));
      return m;
    };
    model.cell_averages = [this] {
      std::vector<double> a(static_cast<std::size_t>(V) * E);
      call(wasl_hip_cell_averages(engine.get(), a.data()));
      return a;
    };
    model.advance_until = [this](RunPosition &at, std::uint64_t limit, double end_time) {
      double history[8] = {0};
      int64_t rows = 0;
      double t = at.time;
      uint64_t s = at.step;
      const int status = wasl_hip_advance(engine.get(), cfl, end_time, limit, 0, &t, &s, history,
                                          2, &rows);
      // Stopping at `limit` short of end_time is the chunk boundary, not an
      // error: the runner decides whether max_steps was reached.
      if (status != WASL_HIP_OK &&
          !(status == WASL_HIP_SIMULATION && s == limit &&
            std::strcmp(wasl_hip_last_error(), kMaxStepsMessage) == 0))
        throw std::runtime_error(wasl_hip_last_error());
      at.time = t;
      at.step = s;
      return rows >= 2 ? history[4 * (rows - 1) + 2] : 0.0;
    };
  }
};

Checkpoint restore(const CaseFiles &files, const CaseMesh &mesh, const SerialCaseModel &model,
                   RunPosition &at) {
  const auto c = read_checkpoint(files.restart_checkpoint);
  if (c.variable_count != model.variable_count || c.mode_count != model.mode_count ||
      c.element_count != mesh.prepared.element_count() || c.partition_rank_count != 1 ||
      std::filesystem::weakly_canonical(c.prepared_mesh) != mesh.path)
    throw std::runtime_error("simulation restart: checkpoint layout or mesh provenance mismatch");
  at.step = c.step;
  at.time = c.time;
  return c;
}

// Line ends from physical names (src/case/line_boundary_resolution.cpp:34-64).
void resolve_line_boundaries(const PreparedMesh &p, const NamedBoundaries &boundaries,
                             BoundaryCondition &left, BoundaryCondition &right) {
  left = right = BoundaryCondition{};
  if (boundaries.empty())
    return;
  std::map<std::pair<int, int>, std::string> names;
  for (std::size_t i = 0; i < p.physical_name_tags.size(); ++i)
    names[{p.physical_name_dimensions[i], p.physical_name_tags[i]}] =
        std::string(p.physical_name_bytes.begin() + static_cast<std::ptrdiff_t>(p.physical_name_offsets[i]),
                    p.physical_name_bytes.begin() + static_cast<std::ptrdiff_t>(p.physical_name_offsets[i + 1]));
  std::set<std::string> used;
  for (std::size_t f = 0; f < p.face_count(); ++f) {
    if (p.face_elements[2 * f + 1] != -1)
      continue;
    const int tag = p.face_physical_tags[f];
    const auto found = names.find({0, tag});
    if (tag == -1 || found == names.end()) {
      // The reference prints the face position with std::to_string.
      const double x = p.points[3 * static_cast<std::size_t>(p.face_nodes[2 * f])];
      throw std::runtime_error("boundaries: the line boundary face at x = " + std::to_string(x) +
                               " has no physical name");
    }
    const std::string &name = found->second;
    used.insert(name);
    const auto cfg = std::find_if(boundaries.begin(), boundaries.end(),
                                  [&](const NamedBoundary &b) { return b.name == name; });
    if (cfg == boundaries.end())
      throw std::runtime_error("boundaries." + name + ": the mesh boundary has no condition");
    (p.face_normals[f] < 0.0 ? left : right) = cfg->condition;
  }
  for (const auto &b : boundaries)
    if (!used.count(b.name))
      throw std::runtime_error("boundaries." + b.name +
                               ": no line boundary face has this physical name");
  if ((left.kind == BoundaryKind::periodic) != (right.kind == BoundaryKind::periodic))
    throw std::runtime_error("boundaries: periodic conditions must be applied to both line ends");
}

}  // namespace

CaseMesh load_case_mesh(const std::filesystem::path &prepared_mesh, int dimensions) {
  CaseMesh mesh;
  mesh.path = std::filesystem::weakly_canonical(prepared_mesh);
  mesh.prepared = read_prepared_mesh(mesh.path);
  if (mesh.prepared.dimension != dimensions)
    throw std::runtime_error("simulation: case dimension does not match prepared mesh");
  return mesh;
}

// run_serial_case (src/case/serial_case_runner.cpp:117-149): the same outputs
// at the same steps, but the device takes every step from one output stop of
// the plan to the next without a host round trip.
SimulationSummary run_serial_case(const CaseFiles &files, const RunControl &run,
                                  const CaseMesh &mesh, const SerialCaseModel &model) {
  RunPosition at{0, run.start_time};
  if (files.restart_checkpoint.empty())
    model.initial_state();
  else
    model.from_checkpoint(restore(files, mesh, model, at).coefficients);
  if (!std::isfinite(at.time) || at.time < run.start_time || at.time > run.end_time ||
      at.step > run.max_steps)
    throw std::runtime_error("simulation restart: checkpoint is outside the configured run range");

  const OutputPlan plan(run);
  CaseOutputs outputs(files, mesh, model, run.keep_checkpoints);
  outputs.emit(kHistory | kSnapshot, at, 0.0);
  while (plan.running(at)) {
    const std::uint64_t from = at.step;
    const double dt = model.advance_until(at, plan.next_stop(at.step), run.end_time);
    if (at.step == from)
      break;
    outputs.emit(plan.events(at), at, dt);
  }
  if (at.time < run.end_time)
    throw std::runtime_error(kMaxStepsMessage);
  const std::filesystem::path last = outputs.close(at);
  return {at.step, at.time, last};
}

void validate_euler_case(const EulerCase &c) {
  if (!valid_files(c.files) || !valid_order(c.polynomial_order) ||
      !std::isfinite(c.heat_capacity_ratio) || c.heat_capacity_ratio <= 1.0 ||
      !nonnegative_finite(c.sensor_threshold) || !nonnegative_finite(c.density_floor) ||
      !nonnegative_finite(c.pressure_floor) || !positive_finite(c.cfl) || !valid_run(c.run) ||
      (c.dimensions != 1 && c.dimensions != 2) || c.initial_regions.empty())
    throw std::runtime_error("simulation: invalid Euler case configuration");
  int defaults = 0;
  for (const auto &r : c.initial_regions) {
    const auto &v = r.value;
    if (!(positive_finite(v.density) && std::isfinite(v.velocity_x) &&
          std::isfinite(v.velocity_y) && positive_finite(v.pressure)))
      throw std::runtime_error("simulation: invalid Euler initial primitive state");
    if (r.shape == EulerRegion::Shape::interval &&
        (!valid_interval(r.lower, r.upper) || c.dimensions != 1))
      throw std::runtime_error("simulation: invalid Euler initial interval");
    if (r.shape == EulerRegion::Shape::circle &&
        (!valid_circle(r.center_x, r.center_y, r.radius) || c.dimensions != 2))
      throw std::runtime_error("simulation: invalid Euler initial circle");
    if (r.shape == EulerRegion::Shape::default_region)
      ++defaults;
  }
  if (defaults != 1)
    throw std::runtime_error("simulation: Euler initial state requires one default region");
  validate_boundaries(c.boundaries);
  for (const auto &b : c.boundaries)
    if (b.condition.kind == BoundaryKind::inflow)
      throw std::runtime_error("simulation: boundary " + b.name +
                               ": Euler inflow is not implemented");
}

void validate_scalar_advection_case(const ScalarAdvectionCase &c) {
  if (!valid_files(c.files) || !valid_order(c.polynomial_order) || !std::isfinite(c.wavespeed) ||
      !positive_finite(c.cfl) || !std::isfinite(c.initial_default) ||
      !std::isfinite(c.initial_sine_amplitude) || !valid_run(c.run))
    throw std::runtime_error("simulation: invalid scalar-advection case configuration");
  if (c.dimensions < 1 || c.dimensions > 2 || !std::isfinite(c.velocity_y))
    throw std::runtime_error("simulation: invalid scalar-advection dimension or velocity");
  for (const auto &r : c.initial_intervals)
    if (!valid_interval(r.lower, r.upper) || !std::isfinite(r.value))
      throw std::runtime_error("simulation: invalid initial interval");
  for (const auto &r : c.initial_circles)
    if (!valid_circle(r.center_x, r.center_y, r.radius) || !std::isfinite(r.value))
      throw std::runtime_error("simulation: invalid initial circle");
  validate_boundaries(c.boundaries);
}

// src/case/euler_case.cpp:8-12 with run_triangle_euler
// (src/case/euler_triangle.cpp:15-47) and run_line_euler
// (src/case/euler_line.cpp:14-48).
SimulationSummary run_euler_case(const EulerCase &config) {
  validate_euler_case(config);
  const auto mesh = load_case_mesh(config.files.prepared_mesh, config.dimensions);
  int boundary = WASL_HIP_PERIODIC, boundary_right = WASL_HIP_PERIODIC;
  if (config.dimensions == 1) {
    BoundaryCondition left, right;
    resolve_line_boundaries(mesh.prepared, config.boundaries, left, right);
    boundary = static_cast<int>(left.kind);
    boundary_right = static_cast<int>(right.kind);
  } else if (!all_of_kind(config.boundaries, BoundaryKind::periodic)) {
    // Beyond the reference: every boundary transmissive (own trace outside).
    if (!config.boundaries.empty() && all_of_kind(config.boundaries, BoundaryKind::outflow))
      boundary = WASL_HIP_OUTFLOW;
    else
      throw std::runtime_error("simulation: 2D Euler boundaries must be periodic");
    boundary_right = boundary;
  }
  auto m = engine_mesh(mesh.prepared);
  wasl_hip_numerics num;
  wasl_hip_numerics_default(&num);
  num.equation = WASL_HIP_EULER;
  num.polynomial_order = config.polynomial_order;
  num.heat_capacity_ratio = config.heat_capacity_ratio;
  num.boundary_left = boundary;
  num.boundary_right = boundary_right;
  num.limiting = config.limiting ? 1 : 0;
  num.sensor_threshold = config.sensor_threshold;
  num.density_floor = config.density_floor;
  num.pressure_floor = config.pressure_floor;
  DeviceCase dev;
  wasl_hip_engine *g = nullptr;
  call(wasl_hip_engine_create(m.get(), &num, &g));
  dev.engine.reset(g);
  dev.cfl = config.cfl;
  dev.V = 4;
  dev.Np = static_cast<int>(wasl_hip_engine_size(g, 1));
  dev.E = static_cast<int>(wasl_hip_engine_size(g, 0));

  SerialCaseModel model;
  dev.bind(model);
  model.initial_state = [&] {
    // initial_coefficients over the regions, then the limiter once
    // (euler_triangle.cpp:37-41).
    std::vector<double> packed;
    for (const auto &r : config.initial_regions)
      packed.insert(packed.end(), {static_cast<double>(static_cast<int>(r.shape)), r.lower, r.upper,
                                   r.center_x, r.center_y, r.radius, r.value.density,
                                   r.value.velocity_x, r.value.velocity_y, r.value.pressure});
    std::vector<double> U(static_cast<std::size_t>(dev.V) * dev.Np * dev.E);
    call(wasl_hip_project(dev.engine.get(), 0, packed.data(), static_cast<int64_t>(packed.size()),
                          U.data()));
    if (config.limiting)
      call(wasl_hip_limit(dev.engine.get(), nullptr));
  };
  return run_serial_case(config.files, config.run, mesh, model);
}

// src/case/scalar_advection_case.cpp:8-16 with run_line_scalar_advection
// (src/case/scalar_advection_line.cpp:12-34) and run_triangle_scalar_advection
// (src/case/scalar_advection_triangle.cpp:15-46).
SimulationSummary run_scalar_advection_case(const ScalarAdvectionCase &config) {
  validate_scalar_advection_case(config);
  const auto mesh = load_case_mesh(config.files.prepared_mesh, config.dimensions);
  if (config.mpi)
    throw std::runtime_error(config.dimensions == 2
                                 ? "simulation: 2D case-level MPI is not implemented"
                                 : "wasl_hip: MPI runs are not part of this engine (one case per GPU)");
  wasl_hip_numerics num;
  wasl_hip_numerics_default(&num);
  num.equation = WASL_HIP_SCALAR_ADVECTION;
  num.polynomial_order = config.polynomial_order;
  num.wavespeed = config.wavespeed;
  if (config.dimensions == 2) {
    if (!all_of_kind(config.boundaries, BoundaryKind::periodic))
      throw std::runtime_error("simulation: 2D scalar-advection boundaries must be periodic");
    num.velocity_y = config.velocity_y;
    num.boundary_left = num.boundary_right = WASL_HIP_PERIODIC;
  } else {
    BoundaryCondition left, right;
    resolve_line_boundaries(mesh.prepared, config.boundaries, left, right);
    num.boundary_left = static_cast<int>(left.kind);
    num.boundary_right = static_cast<int>(right.kind);
    num.boundary_value_left = left.value;
    num.boundary_value_right = right.value;
  }
  auto m = engine_mesh(mesh.prepared);
  DeviceCase dev;
  wasl_hip_engine *g = nullptr;
  call(wasl_hip_engine_create(m.get(), &num, &g));
  dev.engine.reset(g);
  dev.cfl = config.cfl;
  dev.V = 1;
  dev.Np = static_cast<int>(wasl_hip_engine_size(g, 1));
  dev.E = static_cast<int>(wasl_hip_engine_size(g, 0));
  dev.scalar = true;

  SerialCaseModel model;
  dev.bind(model);
  model.initial_state = [&] {
    std::vector<double> U(static_cast<std::size_t>(dev.Np) * dev.E);
    if (config.dimensions == 2) {
      // the first containing circle, else the default (initial_regions.cpp:28-34)
      std::vector<double> packed{config.initial_default};
      for (const auto &c : config.initial_circles)
        packed.insert(packed.end(), {c.center_x, c.center_y, c.radius, c.value});
      call(wasl_hip_project(dev.engine.get(), 12, packed.data(),
                            static_cast<int64_t>(packed.size()), U.data()));
      return;
    }
    std::vector<double> packed{config.initial_default, config.initial_sine ? 1.0 : 0.0,
                               config.initial_sine_amplitude};
    for (const auto &r : config.initial_intervals)
      packed.insert(packed.end(), {r.lower, r.upper, r.value});
    call(wasl_hip_project(dev.engine.get(), 11, packed.data(), static_cast<int64_t>(packed.size()),
                          U.data()));
  };
  return run_serial_case(config.files, config.run, mesh, model);
}

}  // namespace wasl_hip

// ------------------------------------------------------------ C ABI ----

namespace {

thread_local std::string g_case_error;

template <typename F> int guarded(F &&f) {
  try {
    f();
    return WASL_HIP_OK;
  } catch (const std::exception &ex) {
    g_case_error = ex.what();
    return WASL_HIP_ERROR;
  }
}

char *copy_string(const std::string &s) {
  char *out = static_cast<char *>(std::malloc(s.size() + 1));
  if (!out)
    throw std::runtime_error("out of host memory");
  std::memcpy(out, s.c_str(), s.size() + 1);
  return out;
}

std::string str(const char *s) { return s ? std::string(s) : std::string(); }

wasl_hip::CaseFiles case_files(const char *mesh, const char *output, const char *restart) {
  return {str(mesh), str(output), str(restart)};
}

wasl_hip::RunControl run_control(double t0, double t1, uint64_t max_steps, uint64_t hist,
                                 uint64_t vis, uint64_t ckpt, uint64_t keep) {
  return {t0, t1, max_steps, hist, vis, ckpt, keep};
}

// crates/wasl-bridge/src/bridge.cpp:48-63
wasl_hip::NamedBoundaries named_boundaries(const char *const *names, const uint8_t *kinds,
                                           const double *values, size_t count) {
  wasl_hip::NamedBoundaries out;
  if (count && (!names || !kinds || !values))
    throw std::runtime_error("boundaries: invalid boundary data");
  for (size_t i = 0; i < count; ++i) {
    if (kinds[i] > static_cast<uint8_t>(wasl_hip::BoundaryKind::outflow))
      throw std::runtime_error("boundaries: invalid boundary kind");
    out.push_back({str(names[i]), {static_cast<wasl_hip::BoundaryKind>(kinds[i]), values[i]}});
  }
  return out;
}

// crates/wasl-bridge/src/bridge.cpp:40-46
std::string summary_message(const wasl_hip::SimulationSummary &s) {
  std::ostringstream m;
  m.precision(17);
  m << "step=" << s.step << " time=" << s.time << " checkpoint=" << s.checkpoint.string();
  return m.str();
}

void put(char **summary, const std::string &text) {
  if (summary)
    *summary = copy_string(text);
}

}  // namespace

extern "C" {

const char *wasl_hip_case_last_error(void) { return g_case_error.c_str(); }

void wasl_hip_string_free(char *text) { std::free(text); }

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
                            char **summary) {
  return guarded([&] {
    constexpr size_t kWidth = 10;  // bridge.cpp:192
    if (region_values % kWidth != 0 || (region_values && !regions))
      throw std::runtime_error("initial.regions: invalid Euler region data");
    wasl_hip::EulerCase c;
    c.files = case_files(mesh_path, output_path, restart_path);
    c.run = run_control(start_time, end_time, max_steps, history_interval, visualization_interval,
                        checkpoint_interval, keep_checkpoints);
    c.limiting = limiting != 0;
    c.sensor_threshold = sensor_threshold;
    c.density_floor = density_floor;
    c.pressure_floor = pressure_floor;
    c.dimensions = dimensions;
    c.polynomial_order = polynomial_order;
    c.heat_capacity_ratio = heat_capacity_ratio;
    c.cfl = cfl;
    for (size_t i = 0; i < region_values; i += kWidth) {
      wasl_hip::EulerRegion r;
      const int shape = static_cast<int>(regions[i]);
      if (shape < 0 || shape > 2)
        throw std::runtime_error("initial.regions: invalid Euler region shape");
      r.shape = static_cast<wasl_hip::EulerRegion::Shape>(shape);
      r.lower = regions[i + 1];
      r.upper = regions[i + 2];
      r.center_x = regions[i + 3];
      r.center_y = regions[i + 4];
      r.radius = regions[i + 5];
      r.value = {regions[i + 6], regions[i + 7], regions[i + 8], regions[i + 9]};
      c.initial_regions.push_back(r);
    }
    c.boundaries = named_boundaries(boundary_names, boundary_kinds, boundary_values, boundary_count);
    put(summary, summary_message(wasl_hip::run_euler_case(c)));
  });
}

int wasl_hip_run_scalar_advection_case(
    const char *mesh_path, const char *output_path, const char *restart_path, int mpi,
    int dimensions, int polynomial_order, double velocity_x, double velocity_y, double cfl,
    double start_time, double end_time, uint64_t max_steps, uint64_t history_interval,
    uint64_t visualization_interval, uint64_t checkpoint_interval, uint64_t keep_checkpoints,
    double initial_default, int initial_sine, double initial_sine_amplitude,
    const double *intervals, size_t interval_values, const double *circles, size_t circle_values,
    const char *const *boundary_names, const uint8_t *boundary_kinds,
    const double *boundary_values, size_t boundary_count, char **summary) {
  return guarded([&] {
    // bridge.cpp:141-146
    if (interval_values % 3 != 0 || (interval_values && !intervals))
      throw std::runtime_error("initial.regions: invalid interval data");
    if (circle_values % 4 != 0 || (circle_values && !circles))
      throw std::runtime_error("initial.regions: invalid circle data");
    wasl_hip::ScalarAdvectionCase c;
    c.files = case_files(mesh_path, output_path, restart_path);
    c.run = run_control(start_time, end_time, max_steps, history_interval, visualization_interval,
                        checkpoint_interval, keep_checkpoints);
    c.mpi = mpi != 0;
    c.dimensions = dimensions;
    c.polynomial_order = polynomial_order;
    c.wavespeed = velocity_x;
    c.velocity_y = velocity_y;
    c.cfl = cfl;
    c.initial_default = initial_default;
    c.initial_sine = initial_sine != 0;
    c.initial_sine_amplitude = initial_sine_amplitude;
    for (size_t i = 0; i < interval_values; i += 3)
      c.initial_intervals.push_back({intervals[i], intervals[i + 1], intervals[i + 2]});
    for (size_t i = 0; i < circle_values; i += 4)
      c.initial_circles.push_back({circles[i], circles[i + 1], circles[i + 2], circles[i + 3]});
    c.boundaries = named_boundaries(boundary_names, boundary_kinds, boundary_values, boundary_count);
    put(summary, summary_message(wasl_hip::run_scalar_advection_case(c)));
  });
}

int wasl_hip_prepare_mesh(const wasl_hip_mesh *mesh, const char *output_path, char **summary) {
  return wasl_hip_prepare_mesh_tagged(mesh, nullptr, nullptr, 0, nullptr, nullptr, nullptr,
                                      output_path, summary);
}

int wasl_hip_prepare_mesh_tagged(const wasl_hip_mesh *mesh, const int32_t *element_physical_tags,
                                 const int32_t *face_physical_tags, size_t name_count,
                                 const int32_t *name_dimensions, const int32_t *name_tags,
                                 const char *const *names, const char *output_path,
                                 char **summary) {
  return guarded([&] {
    if (!mesh || !output_path)
      throw std::runtime_error("prepare mesh: null argument");
    if (name_count && (!name_dimensions || !name_tags || !names))
      throw std::runtime_error("prepare mesh: invalid physical names");
    const int64_t E = wasl_hip_mesh_count(mesh, 0), F = wasl_hip_mesh_count(mesh, 1);
    std::vector<int32_t> et, ft;
    if (element_physical_tags)
      et.assign(element_physical_tags, element_physical_tags + E);
    if (face_physical_tags)
      ft.assign(face_physical_tags, face_physical_tags + F);
    std::vector<std::pair<std::pair<int, int>, std::string>> named;
    for (size_t i = 0; i < name_count; ++i)
      named.push_back({{name_dimensions[i], name_tags[i]}, str(names[i])});
    const auto prepared = wasl_hip::prepare_mesh(mesh, std::move(et), std::move(ft), named);
    wasl_hip::write_prepared_mesh(output_path, prepared);
    put(summary, "prepared " + std::to_string(prepared.element_count()) + " elements");
  });
}

int wasl_hip_inspect_prepared_mesh(const char *mesh_path, char **summary) {
  return guarded([&] {
    const auto m = wasl_hip::read_prepared_mesh(str(mesh_path));
    put(summary, std::to_string(m.dimension) + "D mesh with " + std::to_string(m.element_count()) +
                     " elements");
  });
}

int wasl_hip_inspect_checkpoint(const char *checkpoint_path, char **summary) {
  return guarded([&] {
    const auto c = wasl_hip::read_checkpoint(str(checkpoint_path));
    std::ostringstream s;
    s.precision(17);
    s << "step=" << c.step << " time=" << c.time << " variables=" << c.variable_count
      << " modes=" << c.mode_count << " elements=" << c.element_count;
    put(summary, s.str());
  });
}

int wasl_hip_export_checkpoint(const char *checkpoint_path, const char *mesh_path,
                               const char *output_path, char **summary) {
  return guarded([&] {
    const auto c = wasl_hip::read_checkpoint(str(checkpoint_path));
    const auto m = wasl_hip::read_prepared_mesh(str(mesh_path));
    if (c.element_count != m.element_count())
      throw std::runtime_error("checkpoint export: mesh element count mismatch");
    const double inverse_sqrt2 = 1.0 / std::sqrt(2.0);
    std::vector<double> averages(c.variable_count * c.element_count);
    for (std::size_t v = 0; v < c.variable_count; ++v)
      for (std::size_t e = 0; e < c.element_count; ++e)
        averages[e * c.variable_count + v] =
            c.coefficients[(v * c.mode_count) * c.element_count + e] * inverse_sqrt2;
    wasl_hip::write_vtu(str(output_path), m, averages, c.variable_count);
    put(summary, "exported " + std::to_string(c.element_count) + " cells");
  });
}

int wasl_hip_mesh_read_prepared(const char *mesh_path, wasl_hip_mesh **out) {
  return guarded([&] {
    if (!out)
      throw std::runtime_error("prepared mesh: null argument");
    const auto p = wasl_hip::read_prepared_mesh(str(mesh_path));
    wasl_hip_mesh *m = nullptr;
    if (wasl_hip_mesh_from_prepared(p.dimension, static_cast<int64_t>(p.node_tags.size()),
                                    p.node_tags.data(), p.points.data(),
                                    static_cast<int64_t>(p.element_count()), p.element_nodes.data(),
                                    static_cast<int64_t>(p.face_count()), p.face_nodes.data(),
                                    p.face_elements.data(), p.face_local_edges.data(),
                                    p.face_normals.data(), &m) != WASL_HIP_OK)
      throw std::runtime_error(wasl_hip_last_error());
    *out = m;
  });
}

int wasl_hip_checkpoint_read(const char *checkpoint_path, double *time, uint64_t *step,
                             uint64_t *variable_count, uint64_t *mode_count,
                             uint64_t *element_count, double *coefficients) {
  return guarded([&] {
    const auto c = wasl_hip::read_checkpoint(str(checkpoint_path));
    if (time)
      *time = c.time;
    if (step)
      *step = c.step;
    if (variable_count)
      *variable_count = c.variable_count;
    if (mode_count)
      *mode_count = c.mode_count;
    if (element_count)
      *element_count = c.element_count;
    if (coefficients)
      std::copy(c.coefficients.begin(), c.coefficients.end(), coefficients);
  });
}

int wasl_hip_checkpoint_write(const char *checkpoint_path, double time, uint64_t step,
                              uint64_t variable_count, uint64_t mode_count,
                              uint64_t element_count, const char *prepared_mesh,
                              const double *coefficients) {
  return guarded([&] {
    wasl_hip::Checkpoint c;
    c.time = time;
    c.step = step;
    c.variable_count = variable_count;
    c.mode_count = mode_count;
    c.element_count = element_count;
    c.prepared_mesh = str(prepared_mesh);
    const std::size_t n = variable_count * mode_count * element_count;
    if (n && !coefficients)
      throw std::runtime_error("checkpoint: null coefficients");
    c.coefficients.assign(coefficients, coefficients + n);
    wasl_hip::write_checkpoint(str(checkpoint_path), c);
  });
}

}  // extern "C"
