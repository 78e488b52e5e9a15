// The reference's on-disk formats driven through a small C API: its own
// prepared-mesh and checkpoint readers and writers (src/io/prepared_mesh_file.cpp:
// 32-97, src/io/checkpoint_file.cpp:17-71), compiled from the reference
// sources where they lie by oracle/Makefile into oracle/_ref/libwaslref_io.so.
//
// TEST INFRASTRUCTURE ONLY: tests/ use it to check that files written by the
// engine's case layer are read by the reference and the other way round.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "wasl/io/checkpoint.hpp"
#include "wasl/io/checkpoint_file.hpp"
#include "wasl/io/history_csv.hpp"
#include "wasl/io/prepared_mesh_file.hpp"
#include "wasl/io/pvd_writer.hpp"
#include "wasl/io/vtu_writer.hpp"
#include "wasl/mesh/msh/msh_document.hpp"
#include "wasl/mesh/prepared/prepare_mesh.hpp"
#include "wasl/mesh/triangle/build_triangle_mesh.hpp"

namespace {

thread_local std::string g_error;

template <typename F> int guarded(F &&f) {
  try {
    f();
    return 0;
  } catch (const std::exception &ex) {
    g_error = ex.what();
    return -1;
  }
}

} // namespace

extern "C" {

const char *orc_io_error(void) { return g_error.c_str(); }

// The reference's prepare_mesh of a triangle mesh given as MSH-style records
// (node tags, x y, triangles of node tags), written with write_prepared_mesh.
int orc_io_prepared_write(int64_t node_count, const int64_t *tags, const double *xy,
                          int64_t triangle_count, const int64_t *triangles, const char *path) {
  return guarded([&] {
    wasl::MshDocument doc;
    for (int64_t i = 0; i < node_count; ++i)
      doc.nodes.push_back({tags[i], 2, 1, xy[2 * i], xy[2 * i + 1], 0.0});
    for (int64_t e = 0; e < triangle_count; ++e)
      doc.elements.push_back({e + 1, 2, 1, wasl::msh_triangle_element,
                              {triangles[3 * e], triangles[3 * e + 1], triangles[3 * e + 2]}});
    wasl::write_prepared_mesh(path, wasl::prepare_mesh(wasl::build_triangle_mesh(doc)));
  });
}

// read_prepared_mesh, then write_prepared_mesh of what it read: a file the
// reference accepts, re-emitted by the reference. counts receives dimension,
// nodes, elements, faces.
int orc_io_prepared_copy(const char *in, const char *out, int64_t *counts) {
  return guarded([&] {
    const wasl::PreparedMesh m = wasl::read_prepared_mesh(in);
    if (counts) {
      counts[0] = m.dimension;
      counts[1] = static_cast<int64_t>(m.node_tags.size());
      counts[2] = static_cast<int64_t>(m.element_count());
      counts[3] = static_cast<int64_t>(m.face_count());
    }
    if (out && *out)
      wasl::write_prepared_mesh(out, m);
  });
}

int orc_io_checkpoint_write(const char *path, double time, uint64_t step,
                            uint64_t variable_count, uint64_t mode_count,
                            uint64_t element_count, const char *prepared_mesh,
                            const double *coefficients) {
  return guarded([&] {
    wasl::Checkpoint c;
    c.time = time;
    c.step = step;
    c.variable_count = variable_count;
    c.mode_count = mode_count;
    c.element_count = element_count;
    c.prepared_mesh = prepared_mesh;
    c.coefficients.assign(coefficients,
                          coefficients + variable_count * mode_count * element_count);
    wasl::write_checkpoint(path, c);
  });
}

// read_checkpoint: the header fields, and the coefficients and mesh path when
// the buffers are given (capacities in doubles and bytes).
int orc_io_checkpoint_read(const char *path, double *time, uint64_t *step, uint64_t *layout,
                           double *coefficients, int64_t capacity, char *mesh_path,
                           int64_t mesh_capacity) {
  return guarded([&] {
    const wasl::Checkpoint c = wasl::read_checkpoint(path);
    if (time)
      *time = c.time;
    if (step)
      *step = c.step;
    if (layout) {
      layout[0] = c.variable_count;
      layout[1] = c.mode_count;
      layout[2] = c.element_count;
      layout[3] = static_cast<uint64_t>(c.partition_rank_count);
    }
    if (coefficients && capacity >= static_cast<int64_t>(c.coefficients.size()))
      std::copy(c.coefficients.begin(), c.coefficients.end(), coefficients);
    if (mesh_path && mesh_capacity > 0) {
      const size_t n = std::min<size_t>(c.prepared_mesh.size(), mesh_capacity - 1);
      std::memcpy(mesh_path, c.prepared_mesh.data(), n);
      mesh_path[n] = '\0';
    }
  });
}

// The bridge's export_checkpoint (crates/wasl-bridge/src/bridge.cpp:107-126):
// cell averages c_0 / sqrt(2) of a checkpoint on its prepared mesh, written by
// the reference's write_vtu.
int orc_io_export(const char *checkpoint_path, const char *mesh_path, const char *out) {
  return guarded([&] {
    const auto c = wasl::read_checkpoint(checkpoint_path);
    const auto mesh = wasl::read_prepared_mesh(mesh_path);
    if (c.element_count != mesh.element_count())
      throw std::runtime_error("checkpoint export: mesh element count mismatch");
    const double inverse_sqrt2 = 1.0 / std::sqrt(2.0);
    std::vector<double> averages(c.variable_count * c.element_count);
    for (std::size_t v = 0; v < c.variable_count; ++v)
      for (std::size_t e = 0; e < c.element_count; ++e)
        averages[e * c.variable_count + v] =
            c.coefficients[(v * c.mode_count) * c.element_count + e] * inverse_sqrt2;
    wasl::write_vtu(out, mesh, averages, c.variable_count);
  });
}

// write_csv_history of rows [n][4] = (step, time, dt, mass).
int orc_io_write_history(const char *path, int64_t n, const double *rows) {
  return guarded([&] {
    std::vector<wasl::HistoryRow> h;
    for (int64_t i = 0; i < n; ++i)
      h.push_back({static_cast<std::uint64_t>(rows[4 * i]), rows[4 * i + 1], rows[4 * i + 2],
                   rows[4 * i + 3]});
    wasl::write_csv_history(path, h);
  });
}

// write_pvd of n snapshots: times[n] and the file names joined by newlines.
int orc_io_write_pvd(const char *path, int64_t n, const double *times, const char *names) {
  return guarded([&] {
    std::vector<std::pair<double, std::string>> snaps;
    std::istringstream in(names);
    std::string name;
    for (int64_t i = 0; i < n && std::getline(in, name); ++i)
      snaps.emplace_back(times[i], name);
    wasl::write_pvd(path, snaps);
  });
}

} // extern "C"
