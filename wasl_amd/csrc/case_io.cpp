// On-disk formats of the case layer: the prepared mesh and the restart
// checkpoint as HDF5 files (scalar attributes and one-dimensional datasets of
// native types, format version 2), and the text outputs history.csv, VTU and
// PVD. Written against the HDF5 C API.

#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <limits>
#include <map>
#include <numeric>
#include <sstream>
#include <stdexcept>

#include "case.hpp"

namespace wasl_hip {
namespace {

constexpr int kPreparedMeshVersion = 2;  // src/io/prepared_mesh_file.cpp:13
constexpr int kCheckpointVersion = 2;    // src/io/checkpoint_file.cpp:13
constexpr int kEarliestVersion = 1;      // src/io/hdf5_io.hpp

void require(bool condition, const char *message) {
  if (!condition)
    throw std::runtime_error(message);
}

// ------------------------------------------------------------- HDF5 ----

// Raised by the wrappers below; the callers turn it into the reference's
// "<what> HDF5 read/write: <detail>" messages.
struct H5Failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class Handle {
public:
  Handle(hid_t id, herr_t (*close)(hid_t), const std::string &what) : id_(id), close_(close) {
    if (id_ < 0)
      throw H5Failure(what);
  }
  Handle(const Handle &) = delete;
  Handle &operator=(const Handle &) = delete;
  ~Handle() {
    if (id_ >= 0)
      close_(id_);
  }
  hid_t get() const { return id_; }

private:
  hid_t id_;
  herr_t (*close_)(hid_t);
};

template <typename T> hid_t native();
template <> hid_t native<double>() { return H5T_NATIVE_DOUBLE; }
template <> hid_t native<std::int64_t>() { return H5T_NATIVE_INT64; }
template <> hid_t native<std::int32_t>() { return H5T_NATIVE_INT32; }
template <> hid_t native<std::uint64_t>() { return H5T_NATIVE_UINT64; }
template <> hid_t native<unsigned char>() { return H5T_NATIVE_UCHAR; }

void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

template <typename T> void write_attribute(hid_t file, const char *name, const T &value) {
  Handle space(H5Screate(H5S_SCALAR), H5Sclose, std::string("dataspace for ") + name);
  Handle attr(H5Acreate2(file, name, native<T>(), space.get(), H5P_DEFAULT, H5P_DEFAULT), H5Aclose,
              std::string("create attribute ") + name);
  if (H5Awrite(attr.get(), native<T>(), &value) < 0)
    throw H5Failure(std::string("write attribute ") + name);
}

template <typename T> T read_attribute(hid_t file, const char *name) {
  Handle attr(H5Aopen(file, name, H5P_DEFAULT), H5Aclose, std::string("open attribute ") + name);
  T value{};
  if (H5Aread(attr.get(), native<T>(), &value) < 0)
    throw H5Failure(std::string("read attribute ") + name);
  return value;
}

template <typename T>
void write_vector(hid_t file, const char *name, const std::vector<T> &values) {
  const hsize_t size = values.size();
  Handle space(H5Screate_simple(1, &size, nullptr), H5Sclose, std::string("dataspace for ") + name);
  Handle set(H5Dcreate2(file, name, native<T>(), space.get(), H5P_DEFAULT, H5P_DEFAULT,
                        H5P_DEFAULT),
             H5Dclose, std::string("create dataset ") + name);
  if (!values.empty() &&
      H5Dwrite(set.get(), native<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, values.data()) < 0)
    throw H5Failure(std::string("write dataset ") + name);
}

template <typename T> std::vector<T> read_vector(hid_t file, const char *name) {
  Handle set(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose, std::string("open dataset ") + name);
  Handle space(H5Dget_space(set.get()), H5Sclose, std::string("dataspace of ") + name);
  if (H5Sget_simple_extent_ndims(space.get()) != 1)
    throw std::runtime_error(std::string("HDF5 I/O: dataset ") + name + " must be one-dimensional");
  hsize_t size = 0;
  H5Sget_simple_extent_dims(space.get(), &size, nullptr);
  std::vector<T> values(static_cast<std::size_t>(size));
  if (!values.empty() &&
      H5Dread(set.get(), native<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, values.data()) < 0)
    throw H5Failure(std::string("read dataset ") + name);
  return values;
}

void create_parent_directories(const std::filesystem::path &path) {
  if (path.has_parent_path())
    std::filesystem::create_directories(path.parent_path());
}

std::ofstream open_text_output(const std::filesystem::path &path) {
  create_parent_directories(path);
  std::ofstream output(path);
  if (!output)
    throw std::runtime_error("file output: cannot open " + path.string());
  output << std::setprecision(17);
  return output;
}

std::string xml_escape(const std::string &value) {
  std::string result;
  for (const char c : value) {
    if (c == '&')
      result += "&amp;";
    else if (c == '<')
      result += "&lt;";
    else if (c == '"')
      result += "&quot;";
    else
      result += c;
  }
  return result;
}

bool all_finite(const std::vector<double> &v) {
  return std::all_of(v.begin(), v.end(), [](double x) { return std::isfinite(x); });
}

bool node_indices_valid(const std::vector<std::int64_t> &v, std::size_t n) {
  return std::all_of(v.begin(), v.end(),
                     [n](std::int64_t i) { return i >= 0 && static_cast<std::size_t>(i) < n; });
}

}  // namespace

// src/mesh/prepared/prepared_mesh_validation.cpp:122-128, with its messages.
void validate_prepared_mesh(const PreparedMesh &m) {
  require(m.dimension >= 1 && m.dimension <= 2, "prepared mesh: dimension must be 1 or 2");
  require(m.nodes_per_element == m.dimension + 1,
          "prepared mesh: nodes_per_element does not match dimension");
  require(m.points.size() == m.node_tags.size() * 3,
          "prepared mesh: points size does not match node tags");
  require(all_finite(m.points), "prepared mesh: point coordinates must be finite");

  require(!m.element_nodes.empty() &&
              m.element_nodes.size() % static_cast<std::size_t>(m.nodes_per_element) == 0,
          "prepared mesh: invalid element connectivity");
  const std::size_t E = m.element_nodes.size() / static_cast<std::size_t>(m.nodes_per_element);
  require(node_indices_valid(m.element_nodes, m.node_tags.size()),
          "prepared mesh: element has invalid local node index");
  require(m.element_physical_tags.size() == E,
          "prepared mesh: element physical-tag count mismatch");
  require(m.element_metrics.size() == E * m.metrics_per_element(),
          "prepared mesh: element metric count mismatch");
  require(all_finite(m.element_metrics), "prepared mesh: element metrics must be finite");

  const std::size_t F = m.face_count();
  require(m.face_nodes.size() == 2 * F && m.face_elements.size() == 2 * F &&
              m.face_local_edges.size() == 2 * F &&
              m.face_normals.size() == F * static_cast<std::size_t>(m.dimension) &&
              m.face_measures.size() == F,
          "prepared mesh: face array size mismatch");
  for (std::size_t f = 0; f < F; ++f) {
    const auto l = m.face_elements[2 * f], r = m.face_elements[2 * f + 1];
    require(l >= 0 && static_cast<std::size_t>(l) < E &&
                (r == -1 || (r >= 0 && static_cast<std::size_t>(r) < E)),
            "prepared mesh: face has invalid element index");
  }
  require(node_indices_valid(m.face_nodes, m.node_tags.size()),
          "prepared mesh: face has invalid local node index");
  require(all_finite(m.face_normals) &&
              std::all_of(m.face_measures.begin(), m.face_measures.end(),
                          [](double v) { return std::isfinite(v) && v > 0.0; }),
          "prepared mesh: face geometry must be finite with positive measures");

  require(!m.physical_name_offsets.empty() &&
              m.physical_name_dimensions.size() == m.physical_name_tags.size() &&
              m.physical_name_offsets.size() == m.physical_name_tags.size() + 1 &&
              m.physical_name_offsets.front() == 0 &&
              m.physical_name_offsets.back() == m.physical_name_bytes.size() &&
              std::is_sorted(m.physical_name_offsets.begin(), m.physical_name_offsets.end()),
          "prepared mesh: physical-name array size mismatch");

  const int R = m.partition_rank_count;
  require(R >= 1, "prepared mesh: partition_rank_count must be positive");
  require(m.element_owners.size() == E, "prepared mesh: element owner count mismatch");
  require(m.face_owners.size() == F, "prepared mesh: face owner count mismatch");
  const auto in_range = [R](std::int32_t o) { return o >= 0 && o < R; };
  require(std::all_of(m.element_owners.begin(), m.element_owners.end(), in_range),
          "prepared mesh: element owner is out of range");
  require(std::all_of(m.face_owners.begin(), m.face_owners.end(), in_range),
          "prepared mesh: face owner is out of range");
  if (E >= static_cast<std::size_t>(R)) {
    std::vector<unsigned char> owns(static_cast<std::size_t>(R), 0);
    for (const auto o : m.element_owners)
      owns[static_cast<std::size_t>(o)] = 1;
    require(std::find(owns.begin(), owns.end(), 0) == owns.end(),
            "prepared mesh: a rank owns no elements");
  }
  for (std::size_t f = 0; f < F; ++f)
    require(m.face_owners[f] == m.element_owners.at(static_cast<std::size_t>(m.face_elements[2 * f])),
            "prepared mesh: face owner does not match its left element owner");
}

// src/io/prepared_mesh_file.cpp:32-61
void write_prepared_mesh(const std::filesystem::path &path, const PreparedMesh &m) {
  validate_prepared_mesh(m);
  create_parent_directories(path);
  quiet();
  try {
    Handle file(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose,
                "cannot create " + path.string());
    const hid_t h = file.get();
    write_attribute(h, "format_version", kPreparedMeshVersion);
    write_attribute(h, "dimension", m.dimension);
    write_attribute(h, "nodes_per_element", m.nodes_per_element);
    write_vector(h, "node_tags", m.node_tags);
    write_vector(h, "points", m.points);
    write_vector(h, "element_nodes", m.element_nodes);
    write_vector(h, "element_physical_tags", m.element_physical_tags);
    write_vector(h, "element_metrics", m.element_metrics);
    write_vector(h, "face_nodes", m.face_nodes);
    write_vector(h, "face_elements", m.face_elements);
    write_vector(h, "face_local_edges", m.face_local_edges);
    write_vector(h, "face_normals", m.face_normals);
    write_vector(h, "face_measures", m.face_measures);
    write_vector(h, "face_physical_tags", m.face_physical_tags);
    write_vector(h, "physical_name_dimensions", m.physical_name_dimensions);
    write_vector(h, "physical_name_tags", m.physical_name_tags);
    write_vector(h, "physical_name_offsets", m.physical_name_offsets);
    write_vector(h, "physical_name_bytes", m.physical_name_bytes);
    write_attribute(h, "partition_rank_count", m.partition_rank_count);
    write_vector(h, "element_owners", m.element_owners);
    write_vector(h, "face_owners", m.face_owners);
  } catch (const H5Failure &e) {
    throw std::runtime_error(std::string("prepared mesh HDF5 write: ") + e.what());
  }
}

// src/io/prepared_mesh_file.cpp:63-97. Version 1 files carry no ownership: a
// single rank owns everything (partition_mesh(mesh, 1)).
PreparedMesh read_prepared_mesh(const std::filesystem::path &path) {
  if (!std::filesystem::is_regular_file(path))
    throw std::runtime_error("prepared mesh HDF5 read: file does not exist: " + path.string());
  quiet();
  try {
    Handle file(H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose,
                "cannot open " + path.string());
    const hid_t h = file.get();
    const int version = read_attribute<std::int32_t>(h, "format_version");
    if (version < kEarliestVersion || version > kPreparedMeshVersion)
      throw std::runtime_error("prepared mesh HDF5 read: unsupported format_version");
    PreparedMesh m;
    m.dimension = read_attribute<std::int32_t>(h, "dimension");
    m.nodes_per_element = read_attribute<std::int32_t>(h, "nodes_per_element");
    m.node_tags = read_vector<std::int64_t>(h, "node_tags");
    m.points = read_vector<double>(h, "points");
    m.element_nodes = read_vector<std::int64_t>(h, "element_nodes");
    m.element_physical_tags = read_vector<std::int32_t>(h, "element_physical_tags");
    m.element_metrics = read_vector<double>(h, "element_metrics");
    m.face_nodes = read_vector<std::int64_t>(h, "face_nodes");
    m.face_elements = read_vector<std::int32_t>(h, "face_elements");
    m.face_local_edges = read_vector<std::int32_t>(h, "face_local_edges");
    m.face_normals = read_vector<double>(h, "face_normals");
    m.face_measures = read_vector<double>(h, "face_measures");
    m.face_physical_tags = read_vector<std::int32_t>(h, "face_physical_tags");
    m.physical_name_dimensions = read_vector<std::int32_t>(h, "physical_name_dimensions");
    m.physical_name_tags = read_vector<std::int32_t>(h, "physical_name_tags");
    m.physical_name_offsets = read_vector<std::uint64_t>(h, "physical_name_offsets");
    m.physical_name_bytes = read_vector<unsigned char>(h, "physical_name_bytes");
    if (version >= 2) {
      m.partition_rank_count = read_attribute<std::int32_t>(h, "partition_rank_count");
      m.element_owners = read_vector<std::int32_t>(h, "element_owners");
      m.face_owners = read_vector<std::int32_t>(h, "face_owners");
    } else {
      if (m.face_elements.size() != m.face_count() * 2)
        throw std::runtime_error("prepared mesh: face array size mismatch");
      m.partition_rank_count = 1;
      m.element_owners.assign(m.element_count(), 0);
      m.face_owners.assign(m.face_count(), 0);
    }
    validate_prepared_mesh(m);
    return m;
  } catch (const H5Failure &e) {
    throw std::runtime_error(std::string("prepared mesh HDF5 read: ") + e.what());
  }
}

// src/io/checkpoint.cpp:25-57
void validate_checkpoint(const Checkpoint &c) {
  bool layout = c.variable_count != 0 && c.mode_count != 0 && c.element_count != 0 &&
                !c.prepared_mesh.empty() &&
                c.variable_count <= std::numeric_limits<std::uint64_t>::max() / c.mode_count;
  if (layout) {
    const auto vm = c.variable_count * c.mode_count;
    layout = vm <= std::numeric_limits<std::uint64_t>::max() / c.element_count &&
             vm * c.element_count == c.coefficients.size();
  }
  if (!layout)
    throw std::runtime_error("checkpoint: coefficient size does not match declared layout");
  if (c.partition_rank_count < 1 || c.partition_rank < 0 ||
      c.partition_rank >= c.partition_rank_count)
    throw std::runtime_error("checkpoint: invalid partition rank");
  const std::uint64_t global = c.global_element_count == 0 ? c.element_count : c.global_element_count;
  std::vector<std::int64_t> indices = c.element_indices;
  if (indices.empty()) {
    indices.resize(c.element_count);
    std::iota(indices.begin(), indices.end(), 0);
  }
  if (indices.size() != c.element_count || global < c.element_count)
    throw std::runtime_error("checkpoint: invalid partition element count");
  std::vector<unsigned char> seen(global, 0);
  for (const auto i : indices) {
    if (i < 0 || static_cast<std::uint64_t>(i) >= global || seen[static_cast<std::size_t>(i)])
      throw std::runtime_error("checkpoint: invalid partition element index");
    seen[static_cast<std::size_t>(i)] = 1;
  }
}

// src/io/checkpoint_file.cpp:17-42
void write_checkpoint(const std::filesystem::path &path, const Checkpoint &c) {
  validate_checkpoint(c);
  create_parent_directories(path);
  quiet();
  try {
    Handle file(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose,
                "cannot create " + path.string());
    const hid_t h = file.get();
    write_attribute(h, "format_version", kCheckpointVersion);
    write_attribute(h, "time", c.time);
    write_attribute(h, "step", c.step);
    write_attribute(h, "variable_count", c.variable_count);
    write_attribute(h, "mode_count", c.mode_count);
    write_attribute(h, "element_count", c.element_count);
    write_vector(h, "prepared_mesh",
                 std::vector<unsigned char>(c.prepared_mesh.begin(), c.prepared_mesh.end()));
    write_vector(h, "coefficients", c.coefficients);
    write_attribute(h, "partition_rank", static_cast<std::int32_t>(c.partition_rank));
    write_attribute(h, "partition_rank_count", static_cast<std::int32_t>(c.partition_rank_count));
    const std::uint64_t global = c.global_element_count == 0 ? c.element_count : c.global_element_count;
    write_attribute(h, "global_element_count", global);
    std::vector<std::int64_t> indices = c.element_indices;
    if (indices.empty()) {
      indices.resize(c.element_count);
      std::iota(indices.begin(), indices.end(), 0);
    }
    write_vector(h, "element_indices", indices);
  } catch (const H5Failure &e) {
    throw std::runtime_error(std::string("checkpoint HDF5 write: ") + e.what());
  }
}

// src/io/checkpoint_file.cpp:44-71
Checkpoint read_checkpoint(const std::filesystem::path &path) {
  if (!std::filesystem::is_regular_file(path))
    throw std::runtime_error("checkpoint HDF5 read: file does not exist: " + path.string());
  quiet();
  try {
    Handle file(H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose,
                "cannot open " + path.string());
    const hid_t h = file.get();
    const int version = read_attribute<std::int32_t>(h, "format_version");
    if (version < kEarliestVersion || version > kCheckpointVersion)
      throw std::runtime_error("checkpoint HDF5 read: unsupported format_version");
    Checkpoint c;
    c.time = read_attribute<double>(h, "time");
    c.step = read_attribute<std::uint64_t>(h, "step");
    c.variable_count = read_attribute<std::uint64_t>(h, "variable_count");
    c.mode_count = read_attribute<std::uint64_t>(h, "mode_count");
    c.element_count = read_attribute<std::uint64_t>(h, "element_count");
    const auto mesh = read_vector<unsigned char>(h, "prepared_mesh");
    c.prepared_mesh.assign(mesh.begin(), mesh.end());
    c.coefficients = read_vector<double>(h, "coefficients");
    if (version >= 2) {
      c.partition_rank = read_attribute<std::int32_t>(h, "partition_rank");
      c.partition_rank_count = read_attribute<std::int32_t>(h, "partition_rank_count");
      c.global_element_count = read_attribute<std::uint64_t>(h, "global_element_count");
      c.element_indices = read_vector<std::int64_t>(h, "element_indices");
    }
    validate_checkpoint(c);
    return c;
  } catch (const H5Failure &e) {
    throw std::runtime_error(std::string("checkpoint HDF5 read: ") + e.what());
  }
}

// src/io/history_csv.cpp:7-12
void write_csv_history(const std::filesystem::path &path, const std::vector<HistoryRow> &rows) {
  auto out = open_text_output(path);
  out << "step,time,dt,mass\n";
  for (const auto &r : rows)
    out << r.step << ',' << r.time << ',' << r.time_step << ',' << r.mass << '\n';
}

// src/io/vtu_writer.cpp:50-65: ASCII unstructured grid of cell values.
void write_vtu(const std::filesystem::path &path, const PreparedMesh &mesh,
               const std::vector<double> &cell_values, std::size_t component_count) {
  validate_prepared_mesh(mesh);
  const std::size_t cells = mesh.element_count();
  if (component_count == 0 || cell_values.size() != cells * component_count)
    throw std::runtime_error("VTU output: cell-value size mismatch");
  auto out = open_text_output(path);
  out << "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
         "byte_order=\"LittleEndian\">\n<UnstructuredGrid><Piece NumberOfPoints=\""
      << mesh.node_tags.size() << "\" NumberOfCells=\"" << cells << "\">\n";
  out << "<Points><DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
  for (const double v : mesh.points)
    out << v << ' ';
  out << "\n</DataArray></Points>\n";
  out << "<Cells>\n<DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
  for (const auto i : mesh.element_nodes)
    out << i << ' ';
  out << "\n</DataArray>\n<DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
  for (std::size_t c = 1; c <= cells; ++c)
    out << c * static_cast<std::size_t>(mesh.nodes_per_element) << ' ';
  out << "\n</DataArray>\n<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
  const int type = mesh.dimension == 1 ? 3 : 5;  // VTK_LINE, VTK_TRIANGLE
  for (std::size_t c = 0; c < cells; ++c)
    out << type << ' ';
  out << "\n</DataArray>\n</Cells>\n";
  out << "<CellData Scalars=\"state\"><DataArray type=\"Float64\" Name=\"state\" "
         "NumberOfComponents=\""
      << component_count << "\" format=\"ascii\">\n";
  for (const double v : cell_values)
    out << v << ' ';
  out << "\n</DataArray></CellData>\n</Piece></UnstructuredGrid></VTKFile>\n";
}

// src/io/pvd_writer.cpp:7-16
void write_pvd(const std::filesystem::path &path,
               const std::vector<std::pair<double, std::string>> &snapshots) {
  auto out = open_text_output(path);
  out << "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\" "
         "byte_order=\"LittleEndian\">\n<Collection>\n";
  for (const auto &[time, file] : snapshots)
    out << "<DataSet timestep=\"" << time << "\" group=\"\" part=\"0\" file=\"" << xml_escape(file)
        << "\"/>\n";
  out << "</Collection></VTKFile>\n";
}

// src/io/output_names.cpp:19-22: prefix-NNNNNN.ext
std::string numbered_step_name(const std::string &prefix, std::uint64_t step,
                               const std::string &extension) {
  std::ostringstream text;
  text << prefix << '-' << std::setfill('0') << std::setw(6) << step << extension;
  return text.str();
}

PreparedMesh prepare_mesh(const wasl_hip_mesh *mesh, std::vector<std::int32_t> element_tags,
                          std::vector<std::int32_t> face_tags,
                          const std::vector<std::pair<std::pair<int, int>, std::string>> &names) {
  if (!mesh)
    throw std::runtime_error("prepare mesh: null mesh");
  const int64_t E = wasl_hip_mesh_count(mesh, 0), F = wasl_hip_mesh_count(mesh, 1),
                N = wasl_hip_mesh_count(mesh, 3);
  const int dim = wasl_hip_mesh_count(mesh, 4) == 1 ? 1 : 2;
  PreparedMesh m;
  m.dimension = dim;
  m.nodes_per_element = dim + 1;
  m.node_tags.resize(N);
  m.points.resize(3 * N);
  m.element_nodes.resize(E * m.nodes_per_element);
  m.element_metrics.resize(E * m.metrics_per_element());
  m.face_nodes.resize(2 * F);
  m.face_elements.resize(2 * F);
  m.face_local_edges.resize(2 * F);
  m.face_normals.resize(F * dim);
  m.face_measures.resize(F);
  if (wasl_hip_mesh_export(mesh, m.node_tags.data(), m.points.data(), m.element_nodes.data(),
                           m.element_metrics.data(), m.face_nodes.data(), m.face_elements.data(),
                           m.face_local_edges.data(), m.face_normals.data(),
                           m.face_measures.data()) != WASL_HIP_OK)
    throw std::runtime_error(wasl_hip_last_error());
  if (element_tags.empty())
    element_tags.assign(E, -1);
  if (face_tags.empty())
    face_tags.assign(F, -1);
  if (static_cast<int64_t>(element_tags.size()) != E || static_cast<int64_t>(face_tags.size()) != F)
    throw std::runtime_error("prepare mesh: physical-tag count mismatch");
  m.element_physical_tags = std::move(element_tags);
  m.face_physical_tags = std::move(face_tags);
  // pack_physical_names (prepare_mesh.cpp:11-19) walks a std::map keyed by
  // (dimension, tag).
  std::map<std::pair<int, int>, std::string> sorted(names.begin(), names.end());
  m.physical_name_offsets.push_back(0);
  for (const auto &[key, name] : sorted) {
    m.physical_name_dimensions.push_back(key.first);
    m.physical_name_tags.push_back(key.second);
    m.physical_name_bytes.insert(m.physical_name_bytes.end(), name.begin(), name.end());
    m.physical_name_offsets.push_back(m.physical_name_bytes.size());
  }
  m.partition_rank_count = 1;
  m.element_owners.assign(E, 0);
  m.face_owners.assign(F, 0);
  return m;
}

}  // namespace wasl_hip
