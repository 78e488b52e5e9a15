// The oracle C API (oracle_api.h) implemented on the reference's own code.
//
// TEST INFRASTRUCTURE ONLY. oracle/Makefile compiles this file together with
// the reference's numeric sources where they lie under /root/reference/cpp
// into oracle/_ref/libwaslref.so. Everything numerical below is a call into
// the reference (mesh builders, periodic trace, limiter stencil, residuals,
// limiter, CFL, SSP-RK3, projection, diagnostics); this file only adapts
// arrays to the reference types and composes the calls the way the
// reference's case drivers do (src/case/euler_triangle.cpp:16-50,
// src/case/euler_line.cpp:14-48,
// src/case/scalar_advection_line.cpp:12-31,
// src/case/scalar_advection_triangle.cpp:15-46, src/case/serial_case_runner.cpp).
//
// Two compositions have no reference counterpart and are marked:
//   - the transmissive 2D exterior (own trace, no limiter neighbour across),
//   - the analytic initial conditions (vortex, rarefaction, smooth field,
//     scalar pulse, 2D scalar plane wave), passed to the reference's
//     projection templates.

#include "oracle_api.h"

#include <algorithm>
#include <cmath>
#include <numbers>
#include <stdexcept>
#include <string>
#include <vector>

#include "wasl/basis/dubiner_triangle.hpp"
#include "wasl/basis/mode_count.hpp"
#include "wasl/boundaries/line_boundary_conditions.hpp"
#include "wasl/boundaries/periodic_triangle_trace.hpp"
#include "wasl/case/euler_case.hpp"
#include "wasl/case/euler_diagnostics.hpp"
#include "wasl/case/initial_state.hpp"
#include "wasl/case/scalar_advection_case.hpp"
#include "wasl/case/scalar_coefficient_layout.hpp"
#include "wasl/case/scalar_diagnostics.hpp"
#include "wasl/boundaries/line_euler_exterior.hpp"
#include "wasl/dg/euler/line_residual.hpp"
#include "wasl/dg/euler/triangle_residual.hpp"
#include "wasl/dg/l2_projection.hpp"
#include "wasl/dg/modal_layout.hpp"
#include "wasl/dg/scalar_advection/line_residual.hpp"
#include "wasl/dg/scalar_advection/triangle_residual.hpp"
#include "wasl/equations/euler/primitive.hpp"
#include "wasl/flux/hllc.hpp"
#include "wasl/limiting/euler_limiter_stencil.hpp"
#include "wasl/limiting/euler_modal_limiter.hpp"
#include "wasl/limiting/positivity/euler_positivity.hpp"
#include "wasl/limiting/troubled_cell/modal_decay.hpp"
#include "wasl/limiting/weno/euler_characteristic.hpp"
#include "wasl/limiting/weno/euler_characteristic_weno.hpp"
#include "wasl/limiting/weno/simple_weno.hpp"
#include "wasl/mesh/line/build_line_mesh.hpp"
#include "wasl/mesh/msh/msh_document.hpp"
#include "wasl/mesh/prepared/mesh_from_prepared.hpp"
#include "wasl/mesh/prepared/prepare_mesh.hpp"
#include "wasl/mesh/triangle/build_triangle_mesh.hpp"
#include "wasl/quadrature/gauss_legendre.hpp"
#include "wasl/quadrature/triangle_edge_rules.hpp"
#include "wasl/quadrature/triangle_volume_rule.hpp"
#include "wasl/time/cfl_timestep.hpp"
#include "wasl/time/line_advection_step.hpp"
#include "wasl/time/ssp_rk3.hpp"

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

}  // namespace

struct orc_case {
  int dim = 2;
  bool tri_scalar = false;  // 2D scalar advection, U[element][mode]
  bool line_euler = false;  // 1D Euler, U[variable][mode][element]
  int p = 0;
  std::size_t n = 0;
  int boundary = ORC_PERIODIC;
  wasl::Euler equation{1.4};
  wasl::TriangleMesh tri;
  wasl::PeriodicTriangleTrace trace;
  wasl::LineMesh line;
  wasl::ScalarAdvection scalar{0.0};
  wasl::ScalarAdvection2D advection{0.0, 0.0};
  wasl::LineBoundaryConditions conditions;
  wasl::ExteriorFunction exterior;
  wasl::EulerLimiterStencil stencil;
  wasl::SspRk3Workspace workspace;

  std::size_t E() const { return dim == 2 ? tri.num_elements() : line.num_elements(); }
  std::size_t F() const { return dim == 2 ? tri.faces.size() : line.faces.size(); }
  std::size_t V() const { return (dim == 2 && !tri_scalar) || line_euler ? 4 : 1; }
};

namespace {

// Composition without a reference counterpart: the transmissive 2D
// exterior is the owner's own trace at each edge point, built from the
// reference's basis and layout helpers.
std::vector<double> own_trace_exterior(const orc_case &c, const std::vector<double> &U) {
  const int P = c.p + 1;
  const auto rules = wasl::triangle_edge_rules(c.p);
  const wasl::ModalLayout layout{c.n, c.E()};
  std::vector<double> ext(4 * c.F() * P, 0.0), basis(c.n);
  for (std::size_t f = 0; f < c.F(); ++f) {
    const auto &face = c.tri.faces[f];
    if (!face.is_boundary())
      continue;
    const auto &rule = rules[static_cast<std::size_t>(face.left_local_edge)];
    for (int q = 0; q < P; ++q) {
      wasl::dubiner_basis(rule.xi[q], rule.eta[q], c.p, basis.data());
      const auto s = wasl::evaluate_euler_state(layout, U.data(),
                                                static_cast<std::size_t>(face.left_element),
                                                basis.data());
      for (int v = 0; v < 4; ++v)
        ext[(static_cast<std::size_t>(v) * c.F() + f) * P + q] = wasl::euler_component(s, v);
    }
  }
  return ext;
}

void residual(const orc_case &c, const std::vector<double> &U, std::vector<double> &rhs) {
  rhs.resize(U.size());
  if (c.line_euler) {
    const auto ext = wasl::line_euler_exterior(c.line, c.p, c.conditions, U);
    wasl::euler_line_residual(c.line, c.equation, wasl::hllc_flux, c.p, U.data(), ext.data(),
                              rhs.data());
  } else if (c.tri_scalar) {
    const auto ext = wasl::periodic_triangle_exterior(c.trace, U, c.n);
    wasl::triangle_scalar_advection_residual(c.tri, c.advection, c.p, U.data(), ext.data(),
                                             rhs.data());
  } else if (c.dim == 2) {
    const auto ext = c.boundary == ORC_PERIODIC
                         ? wasl::periodic_triangle_euler_exterior(c.trace, U, c.E(), c.n, c.F())
                         : own_trace_exterior(c, U);
    wasl::euler_triangle_residual(c.tri, c.equation, wasl::hllc_flux, c.p, U.data(), ext.data(),
                                  rhs.data());
  } else {
    const auto ext = c.exterior(U);
    wasl::scalar_advection_residual(c.line, c.scalar, c.p, U.data(), ext.data(), rhs.data());
  }
}

// The sensor decision of euler_modal_limiter.cpp:18-24, evaluated with the
// public modal_energy helper so flags can be reported.
void sensor_flags(const orc_case &c, const std::vector<double> &U, double thr, uint8_t *flags) {
  const std::size_t E = c.E();
  for (std::size_t e = 0; e < E; ++e) {
    if (c.p == 0) {
      flags[e] = 0;
      continue;
    }
    const auto energy =
        wasl::modal_energy(U.data() + e, E, c.stencil.mode_count, c.stencil.highest_degree_begin);
    flags[e] = !(energy.total == 0.0 || energy.highest_degree / energy.total <= thr) ? 1 : 0;
  }
}

wasl::MshDocument triangle_document(int64_t node_count, const int64_t *tags, const double *xy,
                                    int64_t count, const int64_t *tri) {
  wasl::MshDocument doc;
  for (int64_t i = 0; i < node_count; ++i)
    doc.nodes.push_back({tags[i], 2, 1, xy[2 * i], xy[2 * i + 1], 0.0});
  for (int64_t e = 0; e < count; ++e)
    doc.elements.push_back(
        {e + 1, 2, 1, wasl::msh_triangle_element, {tri[3 * e], tri[3 * e + 1], tri[3 * e + 2]}});
  return doc;
}

// Analytic fields for the benchmark configurations (no reference
// counterpart); evaluated pointwise and handed to project_*_field.
wasl::EulerPrimitive analytic_euler(int ic, const double *prm, double gamma, double x, double y) {
  constexpr double pi = std::numbers::pi;
  if (ic == ORC_IC_VORTEX) {
    const double eps = prm[0];
    const double dx = x - prm[1], dy = y - prm[2];
    const double r2 = dx * dx + dy * dy;
    const double swirl = eps / (2.0 * pi) * std::exp(0.5 * (1.0 - r2));
    const double T = 1.0 - (gamma - 1.0) * eps * eps / (8.0 * gamma * pi * pi) * std::exp(1.0 - r2);
    const double rho = std::pow(T, 1.0 / (gamma - 1.0));
    return {rho, 1.0 - swirl * dy, 1.0 + swirl * dx, std::pow(rho, gamma)};
  }
  if (ic == ORC_IC_RAREFACTION) {
    const double r = std::sqrt(x * x + y * y);
    if (r == 0.0)
      return {prm[0], 0.0, 0.0, prm[1]};
    return {prm[0], prm[2] * x / r, prm[2] * y / r, prm[1]};
  }
  if (ic == ORC_IC_SMOOTH) {
    const double L = prm[0];
    double rho = 0.0;
    for (int k = 0; k < 4; ++k) {
      const double *t = prm + 1 + 4 * k;
      rho += t[0] * std::sin(2.0 * pi * (t[1] * x + t[2] * y) / L + t[3]);
    }
    return {1.0 + 0.1 * rho, 0.5, 0.5, 1.0};
  }
  throw std::runtime_error("oracle: unknown Euler initial condition");
}

void project(const orc_case &c, int ic, const double *prm, int64_t np, std::vector<double> &U) {
  if (c.line_euler) {
    if (ic != ORC_IC_EULER_REGIONS)
      throw std::runtime_error("oracle: 1D Euler takes region initial states");
    wasl::EulerCase config;
    config.dimensions = 1;
    config.polynomial_order = c.p;
    for (int64_t k = 0; k + 10 <= np; k += 10) {
      wasl::EulerRegion region;
      region.shape = static_cast<wasl::EulerRegion::Shape>(static_cast<int>(prm[k]));
      region.lower = prm[k + 1];
      region.upper = prm[k + 2];
      region.center_x = prm[k + 3];
      region.center_y = prm[k + 4];
      region.radius = prm[k + 5];
      region.value = {prm[k + 6], prm[k + 7], prm[k + 8], prm[k + 9]};
      config.initial_regions.push_back(region);
    }
    U = wasl::initial_coefficients(c.line, config, c.equation);
    return;
  }
  if (c.tri_scalar) {
    if (ic == ORC_IC_SCALAR_CIRCLES) {
      wasl::ScalarAdvectionCase config;
      config.polynomial_order = c.p;
      config.initial_default = prm[0];
      for (int64_t k = 1; k + 4 <= np; k += 4)
        config.initial_circles.push_back({prm[k], prm[k + 1], prm[k + 2], prm[k + 3]});
      U = wasl::initial_coefficients(c.tri, config);
      return;
    }
    if (ic != ORC_IC_SCALAR_WAVE)
      throw std::runtime_error("oracle: unknown 2D scalar initial condition");
    const auto projected = wasl::project_triangle_field(
        c.tri, c.p, wasl::triangle_volume_rule(std::max(c.p, 5)), 1,
        [&](double x, double y, double *v) {
          v[0] = std::sin(2.0 * std::numbers::pi * (prm[1] * x + prm[2] * y) / prm[0]);
        });
    U = wasl::solver_coefficients(projected, c.E(), c.n);
    return;
  }
  if (c.dim == 2) {
    if (ic == ORC_IC_EULER_REGIONS) {
      wasl::EulerCase config;
      config.polynomial_order = c.p;
      for (int64_t k = 0; k + 10 <= np; k += 10) {
        wasl::EulerRegion region;
        region.shape = static_cast<wasl::EulerRegion::Shape>(static_cast<int>(prm[k]));
        region.lower = prm[k + 1];
        region.upper = prm[k + 2];
        region.center_x = prm[k + 3];
        region.center_y = prm[k + 4];
        region.radius = prm[k + 5];
        region.value = {prm[k + 6], prm[k + 7], prm[k + 8], prm[k + 9]};
        config.initial_regions.push_back(region);
      }
      U = wasl::initial_coefficients(c.tri, config, c.equation);
      return;
    }
    const double gamma = c.equation.heat_capacity_ratio;
    U = wasl::project_triangle_field(
        c.tri, c.p, wasl::triangle_volume_rule(std::max(c.p, 5)), 4,
        [&](double x, double y, double *values) {
          const auto s = wasl::conserved(c.equation, analytic_euler(ic, prm, gamma, x, y));
          for (int v = 0; v < 4; ++v)
            values[v] = wasl::euler_component(s, v);
        });
    return;
  }
  if (ic == ORC_IC_SCALAR_REGIONS) {
    wasl::ScalarAdvectionCase config;
    config.polynomial_order = c.p;
    config.initial_default = prm[0];
    config.initial_sine = prm[1] != 0.0;
    config.initial_sine_amplitude = prm[2];
    for (int64_t k = 3; k + 3 <= np; k += 3)
      config.initial_intervals.push_back({prm[k], prm[k + 1], prm[k + 2]});
    U = wasl::initial_coefficients(c.line, config);
    return;
  }
  if (ic != ORC_IC_SCALAR_PULSE)
    throw std::runtime_error("oracle: unknown scalar initial condition");
  const auto projected =
      wasl::project_line_field(c.line, c.p, wasl::gauss_legendre(7), 1, [](double x, double *v) {
        const double s = (x - 0.5) / 0.05;
        v[0] = std::sin(2.0 * std::numbers::pi * x) + std::exp(-(s * s));
      });
  U = wasl::solver_coefficients(projected, c.E(), c.n);
}

double mass(const orc_case &c, const std::vector<double> &U) {
  if (c.line_euler)
    return wasl::euler_density_mass(c.line, U);
  if (c.tri_scalar)
    return wasl::scalar_mass(c.tri, U, c.n);
  return c.dim == 2 ? wasl::euler_density_mass(c.tri, U) : wasl::scalar_mass(c.line, U, c.n);
}

double stable_dt(const orc_case &c, const std::vector<double> &U, double cfl) {
  if (c.line_euler)
    return wasl::euler_cfl_timestep(c.line, c.equation, U, c.n, c.p, cfl);
  if (c.tri_scalar)
    return wasl::triangle_scalar_cfl_timestep(c.tri, c.advection.velocity_x,
                                              c.advection.velocity_y, c.p, cfl);
  return c.dim == 2 ? wasl::euler_cfl_timestep(c.tri, c.equation, U, c.n, c.p, cfl)
                    : wasl::advection_cfl_timestep(c.line, c.scalar, c.p, cfl);
}

void step(orc_case &c, double dt, int limiting, double thr, double rf, double pf,
          std::vector<double> &U) {
  if (c.dim == 1 && !c.line_euler) {
    wasl::ssp_rk3_step(c.line, c.scalar, c.p, dt, c.exterior, U, c.workspace);
    return;
  }
  const wasl::ResidualFunction res = [&c](std::vector<double> &stage, std::vector<double> &rhs) {
    residual(c, stage, rhs);
  };
  if (c.tri_scalar) {
    wasl::ssp_rk3_step(dt, res, U, c.workspace);
    return;
  }
  const wasl::StateFunction post = [&](std::vector<double> &s) {
    if (limiting)
      wasl::limit_euler_modal(c.stencil, c.equation, thr, rf, pf, s);
  };
  wasl::ssp_rk3_step(dt, res, post, U, c.workspace);
}

}  // namespace

extern "C" {

const char *orc_error(void) { return g_error.c_str(); }

orc_case *orc_tri_euler_create(int64_t node_count, const int64_t *node_tags, const double *node_xy,
                               int64_t triangle_count, const int64_t *triangle_nodes, int order,
                               double gamma, int boundary) {
  auto *c = new orc_case;
  if (guarded([&] {
        c->dim = 2;
        c->p = order;
        c->n = wasl::triangle_mode_count(order);
        c->boundary = boundary;
        c->equation = wasl::Euler{gamma};
        // The case drivers rebuild the canonical mesh from the prepared mesh
        // (euler_triangle.cpp:19); the round trip is kept here.
        const auto built = wasl::build_triangle_mesh(
            triangle_document(node_count, node_tags, node_xy, triangle_count, triangle_nodes));
        c->tri = wasl::triangle_mesh_from_prepared(wasl::prepare_mesh(built));
        std::vector<std::int32_t> donors;
        if (boundary == ORC_PERIODIC) {
          c->trace = wasl::periodic_triangle_trace(c->tri, order);
          donors = c->trace.elements;
        } else {
          c->trace.elements.assign(c->tri.faces.size() * (order + 1), wasl::invalid_element);
          c->trace.basis.assign(c->tri.faces.size() * (order + 1) * c->n, 0.0);
          c->trace.point_count = order + 1;
        }
        c->stencil = wasl::make_euler_limiter_stencil(c->tri, order, donors, order + 1);
      }) != 0) {
    delete c;
    return nullptr;
  }
  return c;
}

orc_case *orc_tri_scalar_create(int64_t node_count, const int64_t *node_tags, const double *node_xy,
                                int64_t triangle_count, const int64_t *triangle_nodes, int order,
                                double velocity_x, double velocity_y) {
  auto *c = new orc_case;
  if (guarded([&] {
        c->dim = 2;
        c->tri_scalar = true;
        c->p = order;
        c->n = wasl::triangle_mode_count(order);
        c->boundary = ORC_PERIODIC;
        c->advection = wasl::ScalarAdvection2D{velocity_x, velocity_y};
        const auto built = wasl::build_triangle_mesh(
            triangle_document(node_count, node_tags, node_xy, triangle_count, triangle_nodes));
        c->tri = wasl::triangle_mesh_from_prepared(wasl::prepare_mesh(built));
        c->trace = wasl::periodic_triangle_trace(c->tri, order);
      }) != 0) {
    delete c;
    return nullptr;
  }
  return c;
}

orc_case *orc_line_euler_create(int64_t node_count, const int64_t *node_tags,
                                const double *node_x, int64_t segment_count,
                                const int64_t *segment_nodes, int order, double gamma,
                                int left_kind, int right_kind) {
  auto *c = new orc_case;
  if (guarded([&] {
        if ((left_kind == ORC_PERIODIC) != (right_kind == ORC_PERIODIC))
          throw std::runtime_error("boundaries: periodic ends must be paired");
        c->dim = 1;
        c->line_euler = true;
        c->p = order;
        c->n = wasl::line_mode_count(order);
        c->equation = wasl::Euler{gamma};
        wasl::MshDocument doc;
        for (int64_t i = 0; i < node_count; ++i)
          doc.nodes.push_back({node_tags[i], 1, 1, node_x[i], 0.0, 0.0});
        for (int64_t s = 0; s < segment_count; ++s)
          doc.elements.push_back({s + 1, 1, 1, wasl::msh_line_element,
                                  {segment_nodes[2 * s], segment_nodes[2 * s + 1]}});
        c->line = wasl::line_mesh_from_prepared(wasl::prepare_mesh(wasl::build_line_mesh(doc)));
        c->conditions.left = {static_cast<wasl::BoundaryKind>(left_kind), 0.0};
        c->conditions.right = {static_cast<wasl::BoundaryKind>(right_kind), 0.0};
        c->stencil = wasl::make_euler_limiter_stencil(
            c->line, order,
            left_kind == ORC_PERIODIC ? wasl::LineEnds::periodic : wasl::LineEnds::open);
      }) != 0) {
    delete c;
    return nullptr;
  }
  return c;
}

orc_case *orc_line_scalar_create(int64_t node_count, const int64_t *node_tags,
                                 const double *node_x, int64_t segment_count,
                                 const int64_t *segment_nodes, int order, double wavespeed,
                                 int left_kind, double left_value, int right_kind,
                                 double right_value) {
  auto *c = new orc_case;
  if (guarded([&] {
        c->dim = 1;
        c->p = order;
        c->n = wasl::line_mode_count(order);
        c->scalar = wasl::ScalarAdvection{wavespeed};
        wasl::MshDocument doc;
        for (int64_t i = 0; i < node_count; ++i)
          doc.nodes.push_back({node_tags[i], 1, 1, node_x[i], 0.0, 0.0});
        for (int64_t s = 0; s < segment_count; ++s)
          doc.elements.push_back({s + 1, 1, 1, wasl::msh_line_element,
                                  {segment_nodes[2 * s], segment_nodes[2 * s + 1]}});
        c->line = wasl::line_mesh_from_prepared(wasl::prepare_mesh(wasl::build_line_mesh(doc)));
        c->conditions.left = {static_cast<wasl::BoundaryKind>(left_kind), left_value};
        c->conditions.right = {static_cast<wasl::BoundaryKind>(right_kind), right_value};
        c->exterior = wasl::make_line_boundary_exterior(c->line, order, c->conditions);
      }) != 0) {
    delete c;
    return nullptr;
  }
  return c;
}

void orc_destroy(orc_case *c) { delete c; }

int64_t orc_size(const orc_case *c, int what) {
  switch (what) {
  case 0:
    return static_cast<int64_t>(c->E());
  case 1:
    return static_cast<int64_t>(c->F());
  case 2:
    return static_cast<int64_t>(c->n);
  case 3:
    return static_cast<int64_t>(c->V());
  case 4:
    return c->p + 1;
  default:
    return -1;
  }
}

int orc_tri_topology(const orc_case *c, int64_t *element_nodes, double *element_xy,
                     int32_t *face_elements, int32_t *face_edges, int64_t *face_vertices,
                     double *face_normals) {
  return guarded([&] {
    for (std::size_t e = 0; e < c->E(); ++e) {
      const auto &t = c->tri.elements[e];
      element_nodes[3 * e] = t.node0;
      element_nodes[3 * e + 1] = t.node1;
      element_nodes[3 * e + 2] = t.node2;
      const double xy[6] = {t.x0, t.y0, t.x1, t.y1, t.x2, t.y2};
      std::copy(xy, xy + 6, element_xy + 6 * e);
    }
    for (std::size_t f = 0; f < c->F(); ++f) {
      const auto &x = c->tri.faces[f];
      face_elements[2 * f] = x.left_element;
      face_elements[2 * f + 1] = x.right_element;
      face_edges[2 * f] = x.left_local_edge;
      face_edges[2 * f + 1] = x.right_local_edge;
      face_vertices[4 * f] = x.left_vertex0;
      face_vertices[4 * f + 1] = x.left_vertex1;
      face_vertices[4 * f + 2] = x.right_vertex0;
      face_vertices[4 * f + 3] = x.right_vertex1;
      face_normals[2 * f] = x.outward_normal_x;
      face_normals[2 * f + 1] = x.outward_normal_y;
    }
  });
}

int orc_line_topology(const orc_case *c, double *element_x, int32_t *face_elements,
                      double *face_normals) {
  return guarded([&] {
    for (std::size_t e = 0; e < c->E(); ++e) {
      element_x[2 * e] = c->line.elements[e].x_left;
      element_x[2 * e + 1] = c->line.elements[e].x_right;
    }
    for (std::size_t f = 0; f < c->F(); ++f) {
      face_elements[2 * f] = c->line.faces[f].left_element;
      face_elements[2 * f + 1] = c->line.faces[f].right_element;
      face_normals[f] = c->line.faces[f].outward_normal;
    }
  });
}

int orc_tri_trace(const orc_case *c, int32_t *donors, double *basis) {
  return guarded([&] {
    std::copy(c->trace.elements.begin(), c->trace.elements.end(), donors);
    std::copy(c->trace.basis.begin(), c->trace.basis.end(), basis);
  });
}

int64_t orc_stencil_check_count(const orc_case *c) {
  return c->V() == 4 ? static_cast<int64_t>(c->stencil.check_point_count()) : 0;
}

int orc_stencil(const orc_case *c, int32_t *neighbors, double *transforms, double *normals,
                double *check_basis) {
  return guarded([&] {
    const auto &s = c->stencil;
    std::copy(s.neighbors.begin(), s.neighbors.end(), neighbors);
    std::copy(s.transforms.begin(), s.transforms.end(), transforms);
    for (std::size_t e = 0; e < s.normal_x.size(); ++e) {
      normals[2 * e] = s.normal_x[e];
      normals[2 * e + 1] = s.normal_y[e];
    }
    std::copy(s.check_basis.begin(), s.check_basis.end(), check_basis);
  });
}

int orc_residual(const orc_case *c, const double *state, double *rhs) {
  return guarded([&] {
    const std::vector<double> U(state, state + c->V() * c->n * c->E());
    std::vector<double> out;
    residual(*c, U, out);
    std::copy(out.begin(), out.end(), rhs);
  });
}

int orc_limit(const orc_case *c, double threshold, double density_floor, double pressure_floor,
              double *state, uint8_t *flags) {
  return guarded([&] {
    if (c->V() != 4)
      throw std::runtime_error("oracle: the limiter applies to Euler cases");
    std::vector<double> U(state, state + 4 * c->n * c->E());
    if (flags)
      sensor_flags(*c, U, threshold, flags);
    // limit_euler_modal works in place and may throw part way through (a cell
    // average below a floor): the elements limited before the throw are
    // handed back as well, as the restatement's in-place loop does.
    try {
      wasl::limit_euler_modal(c->stencil, c->equation, threshold, density_floor, pressure_floor,
                              U);
    } catch (...) {
      std::copy(U.begin(), U.end(), state);
      throw;
    }
    std::copy(U.begin(), U.end(), state);
  });
}

int orc_stable_dt(const orc_case *c, const double *state, double cfl, double *dt) {
  return guarded([&] {
    const std::vector<double> U(state, state + c->V() * c->n * c->E());
    *dt = stable_dt(*c, U, cfl);
  });
}

int orc_step(orc_case *c, double dt, int limiting, double threshold, double density_floor,
             double pressure_floor, double *state) {
  return guarded([&] {
    std::vector<double> U(state, state + c->V() * c->n * c->E());
    step(*c, dt, limiting, threshold, density_floor, pressure_floor, U);
    std::copy(U.begin(), U.end(), state);
  });
}

int orc_mass(const orc_case *c, const double *state, double *m) {
  return guarded([&] {
    const std::vector<double> U(state, state + c->V() * c->n * c->E());
    *m = mass(*c, U);
  });
}

int orc_project(const orc_case *c, int ic, const double *params, int64_t param_count,
                double *state) {
  return guarded([&] {
    std::vector<double> U;
    project(*c, ic, params, param_count, U);
    std::copy(U.begin(), U.end(), state);
  });
}

// Restates the loop of run_serial_case (serial_case_runner.cpp:117-149)
// around the reference operators; next_time_step (:107-113) is restated as
// it lives in an anonymous namespace.
int orc_run(orc_case *c, double cfl, int limiting, double threshold, double density_floor,
            double pressure_floor, double start_time, uint64_t start_step, double end_time,
            uint64_t max_steps, uint64_t history_interval, double *state, uint64_t *final_step,
            double *final_time, double *history, int64_t history_capacity,
            int64_t *history_rows) {
  return guarded([&] {
    std::vector<double> U(state, state + c->V() * c->n * c->E());
    double time = start_time;
    uint64_t stepno = start_step;
    int64_t rows = 0;
    const auto record = [&](double dt) {
      if (history && rows < history_capacity) {
        history[4 * rows] = static_cast<double>(stepno);
        history[4 * rows + 1] = time;
        history[4 * rows + 2] = dt;
        history[4 * rows + 3] = mass(*c, U);
      }
      ++rows;
    };
    record(0.0);
    while (time < end_time && stepno < max_steps) {
      const double stable = stable_dt(*c, U, cfl);
      const double remaining = end_time - time;
      const double dt = std::min(stable > 0.0 ? stable : remaining, remaining);
      if (dt <= 0.0 || time + dt == time)
        throw std::runtime_error("simulation: timestep cannot advance time");
      step(*c, dt, limiting, threshold, density_floor, pressure_floor, U);
      time += dt;
      ++stepno;
      const bool final = time >= end_time || stepno == max_steps;
      if (stepno % history_interval == 0 || final)
        record(dt);
    }
    if (time < end_time)
      throw std::runtime_error("simulation: max_steps reached before end_time");
    std::copy(U.begin(), U.end(), state);
    *final_step = stepno;
    *final_time = time;
    if (history_rows)
      *history_rows = rows;
  });
}

// ---- function-level hooks: direct calls into the reference routines ----

int orc_basis(int kind, int p, double xi, double eta, int gradient, double *values, double *gx,
              double *ge) {
  return guarded([&] {
    if (kind == 0) {
      wasl::legendre_basis(xi, p, values);
      if (gradient)
        wasl::legendre_basis_gradient(xi, p, gx);
    } else {
      wasl::dubiner_basis(xi, eta, p, values);
      if (gradient)
        wasl::dubiner_basis_gradient(xi, eta, p, gx, ge);
    }
  });
}

int orc_rule(int kind, int n, int edge, double *a, double *b, double *w) {
  int count = -1;
  if (guarded([&] {
        if (kind == 0) {
          const auto r = wasl::gauss_legendre(n);
          count = r.size();
          std::copy(r.points.begin(), r.points.end(), a);
          std::copy(r.weights.begin(), r.weights.end(), w);
        } else if (kind == 1) {
          const auto r = wasl::triangle_volume_rule(n);
          count = r.size();
          std::copy(r.xi.begin(), r.xi.end(), a);
          std::copy(r.eta.begin(), r.eta.end(), b);
          std::copy(r.weights.begin(), r.weights.end(), w);
        } else {
          const auto r = wasl::triangle_edge_rules(n).at(static_cast<std::size_t>(edge));
          count = static_cast<int>(r.xi.size());
          std::copy(r.xi.begin(), r.xi.end(), a);
          std::copy(r.eta.begin(), r.eta.end(), b);
          std::copy(r.weights.begin(), r.weights.end(), w);
        }
      }) != 0)
    return -1;
  return count;
}

namespace {
wasl::EulerConserved cons4(const double *s) { return {s[0], s[1], s[2], s[3]}; }
void store4(const wasl::EulerConserved &s, double *out) {
  for (int v = 0; v < 4; ++v)
    out[v] = wasl::euler_component(s, v);
}
}  // namespace

int orc_euler_flux(double gamma, const double *s, double *fx, double *fy) {
  return guarded([&] {
    const wasl::Euler eq{gamma};
    store4(wasl::euler_flux_x(eq, cons4(s)), fx);
    store4(wasl::euler_flux_y(eq, cons4(s)), fy);
  });
}

int orc_hllc(double gamma, const double *l, const double *r, double nx, double ny,
             double *flux) {
  return guarded(
      [&] { store4(wasl::hllc_flux(wasl::Euler{gamma}, cons4(l), cons4(r), nx, ny), flux); });
}

int orc_conserved(double gamma, const double *w, double *s) {
  return guarded([&] {
    store4(wasl::conserved(wasl::Euler{gamma}, wasl::EulerPrimitive{w[0], w[1], w[2], w[3]}), s);
  });
}

int orc_characteristic(double gamma, const double *avg, const double *d, double nx, double ny,
                       double *waves, double *back) {
  return guarded([&] {
    const wasl::Euler eq{gamma};
    const auto c = wasl::euler_to_characteristic(eq, cons4(avg), cons4(d), nx, ny);
    std::copy(c.begin(), c.end(), waves);
    store4(wasl::euler_from_characteristic(eq, cons4(avg), c, nx, ny), back);
  });
}

int orc_simple_weno(const double *target, const double *candidates, int K, int n,
                    double *limited) {
  return guarded([&] { wasl::simple_weno(target, candidates, K, n, limited); });
}

int orc_characteristic_weno(double gamma, const double *avg, double nx, double ny,
                            const double *target, const double *candidates, int K, int n,
                            double *limited) {
  return guarded([&] {
    wasl::euler_characteristic_weno(wasl::Euler{gamma}, cons4(avg), nx, ny, target, candidates,
                                    K, n, limited);
  });
}

int orc_positivity(double gamma, double *c, int n, const double *basis, int points, double rf,
                   double pf, double *scale) {
  return guarded([&] {
    *scale = wasl::limit_euler_positivity(wasl::Euler{gamma}, c, n, basis, points, rf, pf);
  });
}

int orc_modal_decay_troubled(const double *c, int p, double threshold, int *troubled) {
  return guarded([&] { *troubled = wasl::modal_decay_troubled(c, p, threshold) ? 1 : 0; });
}

}  // extern "C"
