// Host-side setup for the wasl_hip engine (see setup.hpp).
//
// The reference-element evaluations follow the reference's formulas term for
// term (cpp/include/wasl/basis/*.hpp, quadrature/*.hpp, mesh/triangle/
// triangle_geometry.hpp) because the engine's tables must hold the same
// doubles the reference evaluates. The mesh, donor and stencil builders are
// organised for large meshes instead: edge pairing by sorting, a bin grid for
// donor search, per-element parallel stencil and projection loops.

#include "setup.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_map>

namespace whip {

// ------------------------------------------------------------ reference ----

// Three-term recurrence of P_n^(alpha,beta) (basis/jacobi.hpp:12-27).
double jacobi(double x, int n, double alpha, double beta) {
  if (n <= 0)
    return 1.0;
  double p0 = 1.0;
  double p1 = 0.5 * (alpha - beta + (alpha + beta + 2.0) * x);
  for (int k = 1; k < n; ++k) {
    const double s = 2.0 * k + alpha + beta;
    const double a = 2.0 * (k + 1) * (k + alpha + beta + 1.0) * s;
    const double b = (s + 1.0) * (s * (s + 2.0) * x + (alpha * alpha - beta * beta));
    const double c = 2.0 * (k + alpha) * (k + beta) * (s + 2.0);
    const double p2 = (b * p1 - c * p0) / a;
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

double jacobi_derivative(double x, int n, double alpha, double beta) {
  if (n <= 0)
    return 0.0;
  return 0.5 * (n + alpha + beta + 1.0) * jacobi(x, n - 1, alpha + 1.0, beta + 1.0);
}

static double legendre_norm(int j) { return std::sqrt((2.0 * j + 1.0) / 2.0); }

// Orthonormal Legendre values and derivatives (basis/legendre.hpp:23-60).
void legendre(double xi, int p, double *v) {
  v[0] = legendre_norm(0);
  if (p <= 0)
    return;
  double a = 1.0, b = xi;
  v[1] = legendre_norm(1) * b;
  for (int j = 1; j < p; ++j) {
    const double c = ((2.0 * j + 1.0) * xi * b - j * a) / (j + 1);
    a = b;
    b = c;
    v[j + 1] = legendre_norm(j + 1) * b;
  }
}

void legendre_gradient(double xi, int p, double *g) {
  g[0] = 0.0;
  if (p <= 0)
    return;
  double a = 1.0, b = xi, da = 0.0, db = 1.0;
  g[1] = legendre_norm(1) * db;
  for (int j = 1; j < p; ++j) {
    const double c = ((2.0 * j + 1.0) * xi * b - j * a) / (j + 1);
    const double dc = ((2.0 * j + 1.0) * (b + xi * db) - j * da) / (j + 1);
    a = b;
    b = c;
    da = db;
    db = dc;
    g[j + 1] = legendre_norm(j + 1) * db;
  }
}

// Dubiner modes on the collapsed coordinates a = (2+2xi)/(1-eta) - 1, b = eta,
// ordered by degree then first index (basis/dubiner_triangle.hpp:49-105).
void dubiner(double xi, double eta, int p, double *v) {
  const double a = (eta < 1.0) ? (2.0 + 2.0 * xi) / (1.0 - eta) - 1.0 : -1.0;
  const double half = 0.5 * (1.0 - eta);
  int k = 0;
  for (int n = 0; n <= p; ++n)
    for (int i = 0; i <= n; ++i) {
      const double norm = std::sqrt((2.0 * i + 1.0) * (n + 1.0) / 2.0);
      v[k++] = norm * jacobi(a, i, 0.0, 0.0) * std::pow(half, static_cast<double>(i)) *
               jacobi(eta, n - i, 2.0 * i + 1.0, 0.0);
    }
}

void dubiner_gradient(double xi, double eta, int p, double *gxi, double *geta) {
  const double a = (eta < 1.0) ? (2.0 + 2.0 * xi) / (1.0 - eta) - 1.0 : -1.0;
  const double half = 0.5 * (1.0 - eta);
  const double inv = 1.0 / (1.0 - eta);
  const double da_dxi = 2.0 * inv;
  const double da_deta = (a + 1.0) * inv;
  int k = 0;
  for (int n = 0; n <= p; ++n)
    for (int i = 0; i <= n; ++i) {
      const int j = n - i;
      const double norm = std::sqrt((2.0 * i + 1.0) * (n + 1.0) / 2.0);
      const double pa = jacobi(a, i, 0.0, 0.0);
      const double w = std::pow(half, static_cast<double>(i));
      const double qb = jacobi(eta, j, 2.0 * i + 1.0, 0.0);
      const double dpa = jacobi_derivative(a, i, 0.0, 0.0);
      const double dqb = jacobi_derivative(eta, j, 2.0 * i + 1.0, 0.0);
      const double dw =
          i > 0 ? -0.5 * static_cast<double>(i) * std::pow(half, static_cast<double>(i - 1)) : 0.0;
      const double d_a = norm * dpa * w * qb;
      const double d_b = norm * (pa * dw * qb + pa * w * dqb);
      gxi[k] = d_a * da_dxi;
      geta[k] = d_a * da_deta + d_b;
      ++k;
    }
}

// Gauss-Legendre nodes and weights for n = 1..7 (tabulated constants, the
// same values as quadrature/gauss_legendre.hpp:24-42), emitted ascending.
LineRule gauss_legendre(int n) {
  static const double nodes[8][4] = {
      {0, 0, 0, 0},
      {0.0, 0, 0, 0},
      {0.5773502691896257, 0, 0, 0},
      {0.0, 0.7745966692414834, 0, 0},
      {0.33998104358485626, 0.8611363115940526, 0, 0},
      {0.0, 0.5384693101056831, 0.9061798459386640, 0},
      {0.2386191860831969, 0.6612093864662645, 0.9324695142031521, 0},
      {0.0, 0.4058451513773972, 0.7415311855993945, 0.9491079123427585}};
  static const double weights[8][4] = {
      {0, 0, 0, 0},
      {2.0, 0, 0, 0},
      {1.0, 0, 0, 0},
      {0.8888888888888888, 0.5555555555555556, 0, 0},
      {0.6521451548625461, 0.34785484513745385, 0, 0},
      {0.5688888888888889, 0.47862867049936647, 0.2369268850561891, 0},
      {0.4679139345726910, 0.3607615730481386, 0.17132449237917036, 0},
      {0.4179591836734694, 0.3818300505051188, 0.2797053914892766, 0.1294849661688697}};
  if (n < 1 || n > 7)
    throw std::out_of_range("gauss_legendre: unsupported point count");
  const int half = (n + 1) / 2;
  const bool centre = n % 2 == 1;
  LineRule r;
  for (int h = half - 1; h >= 0; --h) {
    if (centre && h == 0)
      continue;
    r.x.push_back(-nodes[n][h]);
    r.w.push_back(weights[n][h]);
  }
  for (int h = 0; h < half; ++h) {
    r.x.push_back(nodes[n][h]);
    r.w.push_back(weights[n][h]);
  }
  return r;
}

// Collapsed tensor rule with Jacobian (1-eta)/2 (quadrature/triangle_volume_rule.hpp:34-51).
TriRule triangle_rule(int p) {
  const LineRule g = gauss_legendre(p + 1);
  TriRule r;
  for (int qb = 0; qb < g.size(); ++qb) {
    const double eta = g.x[qb];
    const double jac = 0.5 * (1.0 - eta);
    for (int qa = 0; qa < g.size(); ++qa) {
      r.xi.push_back(0.5 * (g.x[qa] + 1.0) * (1.0 - eta) - 1.0);
      r.eta.push_back(eta);
      r.w.push_back(g.w[qb] * g.w[qa] * jac);
    }
  }
  return r;
}

// Edge rules: edge 0 eta=-1, edge 1 the hypotenuse (weights scaled by its
// length 2 sqrt2 over 2), edge 2 xi=-1 (quadrature/triangle_edge_rules.hpp:28-58).
std::array<EdgeRule, 3> edge_rules(int p) {
  const LineRule g = gauss_legendre(p + 1);
  const double is2 = 1.0 / std::sqrt(2.0);
  std::array<EdgeRule, 3> e;
  for (int q = 0; q < g.size(); ++q) {
    const double s = g.x[q], w = g.w[q];
    e[0].xi.push_back(s);
    e[0].eta.push_back(-1.0);
    e[0].w.push_back(w);
    e[1].xi.push_back(1.0 + 0.5 * (s + 1.0) * (-2.0));
    e[1].eta.push_back(-1.0 + 0.5 * (s + 1.0) * (2.0));
    e[1].w.push_back(w * 2.0 * is2);
    e[2].xi.push_back(-1.0);
    e[2].eta.push_back(s);
    e[2].w.push_back(w);
  }
  return e;
}

double reference_edge_length(int edge) { return edge == 1 ? 2.0 * std::sqrt(2.0) : 2.0; }

// The left point is reduced to its edge parameter along the left traversal,
// reversed when the right side traverses the edge the other way, then placed
// on the right element's edge (dg/triangle_face_trace.hpp:16-33).
void neighbor_point(int kl, int kr, bool flip, double lxi, double leta, double &rxi,
                    double &reta) {
  double t = kl == 0 ? lxi : kl == 1 ? leta : -leta;
  if (flip)
    t = -t;
  rxi = -1.0;
  reta = -1.0;
  if (kr == 0) {
    rxi = t;
  } else if (kr == 1) {
    rxi = -t;
    reta = t;
  } else {
    reta = -t;
  }
}

// ------------------------------------------------------------- geometry ----

void Triangle::to_physical(double xi, double eta, double &px, double &py) const {
  px = x[0] + 0.5 * (ax() * (xi + 1.0) + bx() * (eta + 1.0));
  py = y[0] + 0.5 * (ay() * (xi + 1.0) + by() * (eta + 1.0));
}

void Triangle::to_reference(double px, double py, double &xi, double &eta) const {
  const double dx = px - x[0], dy = py - y[0];
  const double inv = 1.0 / det_a();
  const double s = (by() * dx - bx() * dy) * inv;
  const double t = (-ay() * dx + ax() * dy) * inv;
  xi = 2.0 * s - 1.0;
  eta = 2.0 * t - 1.0;
}

void Triangle::edge_vector(int k, double &ex, double &ey) const {
  const int a = k, b = (k + 1) % 3;
  ex = x[b] - x[a];
  ey = y[b] - y[a];
}

double Triangle::edge_length(int k) const {
  double ex, ey;
  edge_vector(k, ex, ey);
  return std::hypot(ex, ey);
}

// Outward unit normal: the tangent rotated clockwise for a counter-clockwise
// triangle (mesh/triangle/triangle_geometry.hpp:96-105).
void Triangle::edge_normal(int k, double &nx, double &ny) const {
  double ex, ey;
  edge_vector(k, ex, ey);
  const double len = std::hypot(ex, ey);
  const double tx = ex / len, ty = ey / len;
  const int sign = det_a() > 0.0 ? 1 : -1;
  nx = sign * ty;
  ny = sign * (-tx);
}

size_t Mesh::boundary_faces() const {
  size_t n = 0;
  if (dimension == 2) {
    for (const auto &f : faces)
      n += f.boundary();
  } else {
    for (const auto &f : lfaces)
      n += f.right < 0;
  }
  return n;
}

// ------------------------------------------------------------------ mesh ----

namespace {

struct NodeIndex {
  std::unordered_map<int64_t, int64_t> map;
  std::string context;
  NodeIndex(int64_t n, const int64_t *tags, std::string ctx) : context(std::move(ctx)) {
    map.reserve(static_cast<size_t>(n) * 2);
    for (int64_t i = 0; i < n; ++i)
      if (!map.emplace(tags[i], i).second)
        throw std::runtime_error(context + ": duplicate node tag " + std::to_string(tags[i]));
  }
  int64_t at(int64_t tag) const {
    auto it = map.find(tag);
    if (it == map.end())
      throw std::runtime_error(context + ": element references missing node " +
                               std::to_string(tag));
    return it->second;
  }
};

// Shared-edge faces: each (element, local edge) side becomes a record keyed
// by its sorted vertex pair; sorting the records groups the two sides of an
// edge with the lower element index first, which is the left side. Faces are
// then emitted in (left element, left edge) order, which is the reference's
// face order (build_triangle_mesh.cpp:92-140).
void build_faces(Mesh &m) {
  struct Side {
    int64_t a, b;  // sorted key
    int32_t element;
    int8_t edge;
  };
  const size_t E = m.tris.size();
  std::vector<Side> sides(3 * E);
  for (size_t e = 0; e < E; ++e)
    for (int k = 0; k < 3; ++k) {
      const int64_t v0 = m.tris[e].node[k], v1 = m.tris[e].node[(k + 1) % 3];
      sides[3 * e + k] = {std::min(v0, v1), std::max(v0, v1), static_cast<int32_t>(e),
                          static_cast<int8_t>(k)};
    }
  std::vector<uint32_t> order(sides.size());
  std::iota(order.begin(), order.end(), 0u);
  std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    const Side &s = sides[x], &t = sides[y];
    if (s.a != t.a)
      return s.a < t.a;
    if (s.b != t.b)
      return s.b < t.b;
    return x < y;
  });
  // partner[side] = the other side of the same edge, or -1.
  std::vector<int64_t> partner(sides.size(), -1);
  for (size_t i = 0; i < order.size();) {
    size_t j = i + 1;
    while (j < order.size() && sides[order[j]].a == sides[order[i]].a &&
           sides[order[j]].b == sides[order[i]].b)
      ++j;
    if (j - i > 2)
      throw std::runtime_error("triangle mesh: edge " + std::to_string(sides[order[i]].a) + "-" +
                               std::to_string(sides[order[i]].b) +
                               " has more than two adjacent elements (non-conforming mesh)");
    if (j - i == 2) {
      partner[order[i]] = order[i + 1];
      partner[order[i + 1]] = order[i];
    }
    i = j;
  }
  m.faces.clear();
  m.faces.reserve(E * 3 / 2 + 16);
  for (size_t s = 0; s < sides.size(); ++s) {
    const int64_t other = partner[s];
    if (other >= 0 && static_cast<size_t>(other) < s)
      continue;  // this side is the right side of an earlier face
    const Triangle &t = m.tris[sides[s].element];
    TriFace f;
    f.left = sides[s].element;
    f.left_edge = sides[s].edge;
    f.va = sides[s].a;
    f.vb = sides[s].b;
    f.lv0 = t.node[f.left_edge];
    f.lv1 = t.node[(f.left_edge + 1) % 3];
    if (other >= 0) {
      const Triangle &r = m.tris[sides[other].element];
      f.right = sides[other].element;
      f.right_edge = sides[other].edge;
      f.rv0 = r.node[f.right_edge];
      f.rv1 = r.node[(f.right_edge + 1) % 3];
    }
    t.edge_normal(f.left_edge, f.nx, f.ny);
    m.faces.push_back(f);
  }
}

void build_line_faces(Mesh &m) {
  m.lfaces.clear();
  m.lfaces.push_back({0, -1, m.lines.front().xl, -1.0});
  for (size_t i = 0; i + 1 < m.lines.size(); ++i) {
    if (m.lines[i].right_node != m.lines[i + 1].left_node)
      throw std::runtime_error("line mesh: gap or discontinuity between elements " +
                               std::to_string(i) + " and " + std::to_string(i + 1));
    m.lfaces.push_back(
        {static_cast<int32_t>(i), static_cast<int32_t>(i + 1), m.lines[i].xr, 1.0});
  }
  m.lfaces.push_back({static_cast<int32_t>(m.lines.size() - 1), -1, m.lines.back().xr, 1.0});
}

}  // namespace

Mesh build_mesh(int dimension, int64_t node_count, const int64_t *tags, const double *xyz,
                int64_t element_count, const int64_t *element_nodes) {
  Mesh m;
  m.dimension = dimension;
  if (dimension == 2) {
    const NodeIndex nodes(node_count, tags, "triangle mesh");
    m.tris.resize(static_cast<size_t>(element_count));
    for (int64_t e = 0; e < element_count; ++e) {
      int64_t t[3] = {element_nodes[3 * e], element_nodes[3 * e + 1], element_nodes[3 * e + 2]};
      int64_t v[3] = {nodes.at(t[0]), nodes.at(t[1]), nodes.at(t[2])};
      const double *p0 = xyz + 3 * v[0], *p1 = xyz + 3 * v[1], *p2 = xyz + 3 * v[2];
      const double area2 =
          p0[0] * (p1[1] - p2[1]) + p1[0] * (p2[1] - p0[1]) + p2[0] * (p0[1] - p1[1]);
      if (area2 == 0.0)
        throw std::runtime_error("triangle mesh: zero-area element " + std::to_string(e + 1));
      if (area2 < 0.0) {
        std::swap(v[1], v[2]);
        std::swap(t[1], t[2]);
      }
      Triangle &tri = m.tris[static_cast<size_t>(e)];
      for (int k = 0; k < 3; ++k) {
        tri.node[k] = t[k];
        tri.x[k] = xyz[3 * v[k]];
        tri.y[k] = xyz[3 * v[k] + 1];
      }
    }
    if (m.tris.empty())
      throw std::runtime_error("triangle mesh: document contains no triangle elements");
    build_faces(m);
    return m;
  }
  if (dimension != 1)
    throw std::runtime_error("mesh: dimension must be 1 or 2");
  const NodeIndex nodes(node_count, tags, "line mesh");
  for (int64_t s = 0; s < element_count; ++s) {
    const int64_t ta = element_nodes[2 * s], tb = element_nodes[2 * s + 1];
    const double xa = xyz[3 * nodes.at(ta)], xb = xyz[3 * nodes.at(tb)];
    if (xa == xb)
      throw std::runtime_error("line mesh: zero-length element " + std::to_string(s + 1));
    m.lines.push_back(xa < xb ? Segment{ta, tb, xa, xb} : Segment{tb, ta, xb, xa});
  }
  if (m.lines.empty())
    throw std::runtime_error("line mesh: document contains no line elements");
  std::stable_sort(m.lines.begin(), m.lines.end(),
                   [](const Segment &a, const Segment &b) { return a.xl < b.xl; });
  build_line_faces(m);
  return m;
}

// mesh_from_prepared.cpp:169-229: elements from node indices, faces as stored.
Mesh mesh_from_prepared(int dimension, int64_t node_count, const int64_t *tags,
                        const double *points, int64_t element_count, const int64_t *element_nodes,
                        int64_t face_count, const int64_t *face_nodes,
                        const int32_t *face_elements, const int32_t *face_local_edges,
                        const double *face_normals) {
  Mesh m;
  m.dimension = dimension;
  const auto node = [&](int64_t i) {
    if (i < 0 || i >= node_count)
      throw std::runtime_error("prepared mesh: node index out of range");
    return i;
  };
  if (dimension == 2) {
    m.tris.resize(static_cast<size_t>(element_count));
    for (int64_t e = 0; e < element_count; ++e)
      for (int k = 0; k < 3; ++k) {
        const int64_t n = node(element_nodes[3 * e + k]);
        m.tris[e].node[k] = tags[n];
        m.tris[e].x[k] = points[3 * n];
        m.tris[e].y[k] = points[3 * n + 1];
      }
    m.faces.resize(static_cast<size_t>(face_count));
    for (int64_t f = 0; f < face_count; ++f) {
      TriFace &x = m.faces[f];
      x.left = face_elements[2 * f];
      x.right = face_elements[2 * f + 1];
      x.left_edge = face_local_edges[2 * f];
      x.right_edge = face_local_edges[2 * f + 1];
      if (x.left < 0 || x.left >= element_count || x.right >= element_count ||
          x.left_edge < 0 || x.left_edge > 2 || (x.right >= 0 && (x.right_edge < 0 || x.right_edge > 2)))
        throw std::runtime_error("prepared mesh: invalid face connectivity");
      x.va = tags[node(face_nodes[2 * f])];
      x.vb = tags[node(face_nodes[2 * f + 1])];
      const Triangle &l = m.tris[x.left];
      x.lv0 = l.node[x.left_edge];
      x.lv1 = l.node[(x.left_edge + 1) % 3];
      if (x.right >= 0) {
        const Triangle &r = m.tris[x.right];
        x.rv0 = r.node[x.right_edge];
        x.rv1 = r.node[(x.right_edge + 1) % 3];
      }
      x.nx = face_normals[2 * f];
      x.ny = face_normals[2 * f + 1];
    }
    return m;
  }
  for (int64_t e = 0; e < element_count; ++e) {
    const int64_t a = node(element_nodes[2 * e]), b = node(element_nodes[2 * e + 1]);
    m.lines.push_back({tags[a], tags[b], points[3 * a], points[3 * b]});
  }
  for (int64_t f = 0; f < face_count; ++f)
    m.lfaces.push_back({face_elements[2 * f], face_elements[2 * f + 1],
                        points[3 * node(face_nodes[2 * f])], face_normals[f]});
  return m;
}

// ------------------------------------------------------------ boundaries ----

PeriodicTrace periodic_trace(const Mesh &m, int p) {
  const int P = p + 1, Np = tri_modes(p);
  PeriodicTrace tr;
  tr.points = P;
  tr.donors.assign(m.faces.size() * P, -1);
  tr.basis.assign(m.faces.size() * P * Np, 0.0);
  double x0 = std::numeric_limits<double>::max(), x1 = std::numeric_limits<double>::lowest();
  double y0 = x0, y1 = x1;
  for (const auto &t : m.tris)
    for (int k = 0; k < 3; ++k) {
      x0 = std::min(x0, t.x[k]);
      x1 = std::max(x1, t.x[k]);
      y0 = std::min(y0, t.y[k]);
      y1 = std::max(y1, t.y[k]);
    }
  // Bin grid over the box; each element is listed in every bin its slightly
  // enlarged bounding box touches.
  const size_t E = m.tris.size();
  const int nb = std::max(1, static_cast<int>(std::sqrt(static_cast<double>(E) / 2.0)));
  const double w = x1 - x0, h = y1 - y0;
  const auto bin_of = [&](double v, double lo, double span) {
    int b = static_cast<int>((v - lo) / span * nb);
    return std::min(nb - 1, std::max(0, b));
  };
  std::vector<std::vector<int32_t>> bins(static_cast<size_t>(nb) * nb);
  for (size_t e = 0; e < E; ++e) {
    const Triangle &t = m.tris[e];
    double bx0 = std::min({t.x[0], t.x[1], t.x[2]}), bx1 = std::max({t.x[0], t.x[1], t.x[2]});
    double by0 = std::min({t.y[0], t.y[1], t.y[2]}), by1 = std::max({t.y[0], t.y[1], t.y[2]});
    const double margin = 1e-6 * std::max(bx1 - bx0, by1 - by0) + 1e-12 * std::max(w, h);
    for (int j = bin_of(by0 - margin, y0, h); j <= bin_of(by1 + margin, y0, h); ++j)
      for (int i = bin_of(bx0 - margin, x0, w); i <= bin_of(bx1 + margin, x0, w); ++i)
        bins[static_cast<size_t>(j) * nb + i].push_back(static_cast<int32_t>(e));
  }
  const auto rules = edge_rules(p);
  for (size_t f = 0; f < m.faces.size(); ++f) {
    const TriFace &face = m.faces[f];
    if (!face.boundary())
      continue;
    const Triangle &t = m.tris[face.left];
    const EdgeRule &rule = rules[face.left_edge];
    for (int q = 0; q < P; ++q) {
      double px, py;
      t.to_physical(rule.xi[q], rule.eta[q], px, py);
      // periodic image across an axis-aligned box (periodic_triangle_trace.cpp:17-24)
      if (std::fabs(face.nx) > 0.5 && std::fabs(face.ny) < 1e-10)
        px += face.nx < 0.0 ? x1 - x0 : x0 - x1;
      else if (std::fabs(face.ny) > 0.5 && std::fabs(face.nx) < 1e-10)
        py += face.ny < 0.0 ? y1 - y0 : y0 - y1;
      else
        throw std::runtime_error("simulation: 2D periodic boundaries require an axis-aligned box");
      int32_t best = -1;
      double bxi = 0.0, beta = 0.0;
      for (int32_t e : bins[static_cast<size_t>(bin_of(py, y0, h)) * nb + bin_of(px, x0, w)]) {
        if (best >= 0 && e > best)
          continue;
        double xi, eta;
        m.tris[e].to_reference(px, py, xi, eta);
        if (xi >= -1.0 - 1e-10 && eta >= -1.0 - 1e-10 && xi + eta <= 1e-10) {
          best = e;
          bxi = xi;
          beta = eta;
        }
      }
      if (best < 0)
        throw std::runtime_error("simulation: periodic 2D trace has no donor element");
      tr.donors[f * P + q] = best;
      dubiner(bxi, beta, p, tr.basis.data() + (f * P + q) * Np);
    }
  }
  return tr;
}

// ---------------------------------------------------------------- stencil ----

LimiterStencil line_limiter_stencil(const Mesh &m, int p, bool periodic) {
  LimiterStencil s;
  const size_t E = m.lines.size();
  const int Np = p + 1;
  s.K = 2;
  s.Np = Np;
  s.highest_begin = p > 0 ? Np - 1 : Np;
  s.neighbors.assign(E * 2, -1);
  s.nx.assign(E, 1.0);
  s.ny.assign(E, 0.0);
  const auto add = [&](size_t e, int32_t n) {
    int32_t *v = s.neighbors.data() + 2 * e;
    for (int i = 0; i < 2; ++i) {
      if (v[i] == n)
        return;
      if (v[i] < 0) {
        v[i] = n;
        return;
      }
    }
  };
  int32_t left_end = -1, right_end = -1;
  for (const LineFace &f : m.lfaces) {
    if (f.right >= 0) {
      add(static_cast<size_t>(f.left), f.right);
      add(static_cast<size_t>(f.right), f.left);
    } else if (f.normal < 0.0) {
      left_end = f.left;
    } else {
      right_end = f.left;
    }
  }
  if (periodic && left_end >= 0 && right_end >= 0) {
    add(static_cast<size_t>(left_end), right_end);
    add(static_cast<size_t>(right_end), left_end);
  }
  // entry [m][s] += w_q phi_m(xi_q) phi_s(source xi of q), the neighbour moved
  // by the period when its centre is more than half a period away
  const LineRule rule = gauss_legendre(std::max(1, p + 1));
  const double period = m.lines.back().xr - m.lines.front().xl;
  s.transforms.assign(E * 2 * Np * Np, 0.0);
  std::vector<double> sb(Np), tb(Np);
  for (size_t e = 0; e < E; ++e) {
    const double tc = (m.lines[e].xl + m.lines[e].xr) / 2.0;
    const double tj = (m.lines[e].xr - m.lines[e].xl) / 2.0;
    for (int k = 0; k < 2; ++k) {
      const int32_t nb = s.neighbors[2 * e + k];
      if (nb < 0)
        continue;
      double sl = m.lines[static_cast<size_t>(nb)].xl, sr = m.lines[static_cast<size_t>(nb)].xr;
      const double gap = tc - (sl + sr) / 2.0;
      const double shift = gap > 0.5 * period ? period : gap < -0.5 * period ? -period : 0.0;
      sl += shift;
      sr += shift;
      const double sc = (sl + sr) / 2.0, sj = (sr - sl) / 2.0;
      double *M = s.transforms.data() + (2 * e + k) * Np * Np;
      for (int q = 0; q < rule.size(); ++q) {
        legendre((tc + tj * rule.x[q] - sc) / sj, p, sb.data());
        legendre(rule.x[q], p, tb.data());
        for (int a = 0; a < Np; ++a)
          for (int b = 0; b < Np; ++b)
            M[a * Np + b] += rule.w[q] * tb[a] * sb[b];
      }
    }
  }
  std::vector<double> b(Np);
  const auto append = [&](double xi) {
    legendre(xi, p, b.data());
    s.check.insert(s.check.end(), b.begin(), b.end());
  };
  for (int q = 0; q < rule.size(); ++q)
    append(rule.x[q]);
  append(-1.0);
  append(1.0);
  return s;
}

LimiterStencil limiter_stencil(const Mesh &m, int p, const std::vector<int32_t> &donors,
                               int points) {
  LimiterStencil s;
  const size_t E = m.tris.size();
  const int Np = tri_modes(p);
  s.K = 3;
  s.Np = Np;
  s.highest_begin = p > 0 ? tri_modes(p - 1) : Np;
  s.neighbors.assign(E * 3, -1);
  s.nx.assign(E, 1.0);
  s.ny.assign(E, 0.0);
  std::vector<uint8_t> has(E, 0);
  const auto add = [&](size_t e, int32_t n) {
    int32_t *v = s.neighbors.data() + 3 * e;
    for (int i = 0; i < 3; ++i) {
      if (v[i] == n)
        return;
      if (v[i] < 0) {
        v[i] = n;
        return;
      }
    }
  };
  const auto normal = [&](size_t e, double nx, double ny) {
    if (has[e])
      return;
    s.nx[e] = nx;
    s.ny[e] = ny;
    has[e] = 1;
  };
  // Neighbours in face order; the characteristic normal comes from the first
  // face that gives the element a neighbour (euler_limiter_stencil_triangle.cpp:18-44).
  for (size_t f = 0; f < m.faces.size(); ++f) {
    const TriFace &face = m.faces[f];
    int32_t right = face.right;
    if (face.boundary() && !donors.empty())
      right = donors[f * points];
    if (right < 0)
      continue;
    add(static_cast<size_t>(face.left), right);
    normal(static_cast<size_t>(face.left), face.nx, face.ny);
    if (!face.boundary()) {
      add(static_cast<size_t>(right), face.left);
      normal(static_cast<size_t>(right), -face.nx, -face.ny);
    }
  }
  // Transforms: entry [m][s] = sum_q w_q phi_s(source point of q) phi_m(q), the
  // L2 projection of the neighbour's mode s onto the element basis, with a
  // periodic image brought next to the element (candidate_transform.hpp:18-39,
  // limiter_stencil_assembly.hpp:45-47).
  double x0 = std::numeric_limits<double>::max(), x1 = std::numeric_limits<double>::lowest();
  double y0 = x0, y1 = x1;
  for (const auto &t : m.tris)
    for (int k = 0; k < 3; ++k) {
      x0 = std::min(x0, t.x[k]);
      x1 = std::max(x1, t.x[k]);
      y0 = std::min(y0, t.y[k]);
      y1 = std::max(y1, t.y[k]);
    }
  const double width = x1 - x0, height = y1 - y0;
  const auto offset = [](double gap, double period) {
    return gap > 0.5 * period ? period : gap < -0.5 * period ? -period : 0.0;
  };
  const TriRule rule = triangle_rule(p);
  const int Q = rule.size();
  std::vector<double> tb(static_cast<size_t>(Q) * Np);
  for (int q = 0; q < Q; ++q)
    dubiner(rule.xi[q], rule.eta[q], p, tb.data() + q * Np);
  s.transforms.assign(E * 3 * Np * Np, 0.0);
  parallel_for(E, [&](size_t begin, size_t end) {
    std::vector<double> sb(Np);
    for (size_t e = begin; e < end; ++e) {
      const Triangle &tgt = m.tris[e];
      for (int k = 0; k < 3; ++k) {
        const int32_t nb = s.neighbors[3 * e + k];
        if (nb < 0)
          continue;
        Triangle src = m.tris[nb];
        const double gx =
            (tgt.x[0] + tgt.x[1] + tgt.x[2] - src.x[0] - src.x[1] - src.x[2]) / 3.0;
        const double gy =
            (tgt.y[0] + tgt.y[1] + tgt.y[2] - src.y[0] - src.y[1] - src.y[2]) / 3.0;
        const double ox = offset(gx, width), oy = offset(gy, height);
        for (int c = 0; c < 3; ++c) {
          src.x[c] += ox;
          src.y[c] += oy;
        }
        double *M = s.transforms.data() + (e * 3 + k) * Np * Np;
        for (int q = 0; q < Q; ++q) {
          double px, py, sxi, seta;
          tgt.to_physical(rule.xi[q], rule.eta[q], px, py);
          src.to_reference(px, py, sxi, seta);
          dubiner(sxi, seta, p, sb.data());
          for (int sm = 0; sm < Np; ++sm) {
            const double ws = rule.w[q] * sb[sm];
            for (int tm = 0; tm < Np; ++tm)
              M[tm * Np + sm] += ws * tb[q * Np + tm];
          }
        }
      }
    }
  });
  // Positivity check points: the order-p volume rule, then the edge points
  // (euler_limiter_stencil_triangle.cpp:81-93).
  s.check.assign(tb.begin(), tb.end());
  std::vector<double> b(Np);
  for (const EdgeRule &edge : edge_rules(p))
    for (size_t q = 0; q < edge.xi.size(); ++q) {
      dubiner(edge.xi[q], edge.eta[q], p, b.data());
      s.check.insert(s.check.end(), b.begin(), b.end());
    }
  return s;
}

double minimum_length(const Mesh &m) {
  double r = std::numeric_limits<double>::max();
  if (m.dimension == 1) {
    for (const auto &l : m.lines)
      r = std::min(r, l.xr - l.xl);
    return r;
  }
  for (const auto &t : m.tris) {
    const double longest = std::max({t.edge_length(0), t.edge_length(1), t.edge_length(2)});
    r = std::min(r, 4.0 * t.jacobian_determinant() / longest);
  }
  return r;
}

// ------------------------------------------------------------- projection ----

namespace {

struct Prim {
  double r, u, v, p;
};

// Regions: the first containing interval (1D) or circle (2D) wins, the
// default region fills the rest (src/case/initial_regions.cpp:154-174).
Prim euler_value(int ic, const double *prm, int64_t np, double gamma, double x, double y) {
  switch (ic) {
  case IC_EULER_REGIONS: {
    const double *fallback = nullptr;
    for (int64_t k = 0; k + 10 <= np; k += 10) {
      const double *r = prm + k;
      const int shape = static_cast<int>(r[0]);
      if (shape == 0) {
        fallback = r;
      } else if (shape == 1) {
        if (x >= r[1] && x <= r[2])
          return {r[6], r[7], r[8], r[9]};
      } else {
        const double dx = x - r[3], dy = y - r[4];
        if (dx * dx + dy * dy <= r[5] * r[5])
          return {r[6], r[7], r[8], r[9]};
      }
    }
    if (!fallback)
      throw std::runtime_error("simulation: Euler initial state requires a default region");
    return {fallback[6], fallback[7], fallback[8], fallback[9]};
  }
  case IC_VORTEX: {
    // Isentropic vortex of strength eps about (cx, cy) on a unit background.
    const double eps = prm[0];
    const double dx = x - prm[1], dy = y - prm[2];
    const double r2 = dx * dx + dy * dy;
    const double swirl = eps / (2.0 * kPi) * std::exp(0.5 * (1.0 - r2));
    const double T =
        1.0 - (gamma - 1.0) * eps * eps / (8.0 * gamma * kPi * kPi) * std::exp(1.0 - r2);
    const double rho = std::pow(T, 1.0 / (gamma - 1.0));
    return {rho, 1.0 - swirl * dy, 1.0 + swirl * dx, std::pow(rho, gamma)};
  }
  case IC_RAREFACTION: {
    const double r = std::sqrt(x * x + y * y);
    if (r == 0.0)
      return {prm[0], 0.0, 0.0, prm[1]};
    return {prm[0], prm[2] * x / r, prm[2] * y / r, prm[1]};
  }
  case IC_SMOOTH: {
    const double L = prm[0];
    double s = 0.0;
    for (int k = 0; k < 4; ++k) {
      const double *t = prm + 1 + 4 * k;
      s += t[0] * std::sin(2.0 * kPi * (t[1] * x + t[2] * y) / L + t[3]);
    }
    return {1.0 + 0.1 * s, 0.5, 0.5, 1.0};
  }
  default:
    throw std::runtime_error("initial state: unknown Euler initial condition " +
                             std::to_string(ic));
  }
}

double scalar_value(int ic, const double *prm, int64_t np, double x) {
  if (ic == IC_SCALAR_PULSE) {
    const double s = (x - 0.5) / 0.05;
    return std::sin(2.0 * kPi * x) + std::exp(-(s * s));
  }
  if (ic == IC_SCALAR_REGIONS) {
    if (np < 3)
      throw std::runtime_error("initial state: scalar regions need default, sine, amplitude");
    if (prm[1] != 0.0)
      return prm[2] * std::sin(2.0 * kPi * x);
    for (int64_t k = 3; k + 3 <= np; k += 3)
      if (x >= prm[k] && x <= prm[k + 1])
        return prm[k + 2];
    return prm[0];
  }
  throw std::runtime_error("initial state: unknown scalar initial condition " +
                           std::to_string(ic));
}

// 2D scalar initial values: the first containing circle, else the default
// (src/case/initial_regions.cpp:28-34), or a plane sine wave.
double scalar_value_2d(int ic, const double *prm, int64_t np, double x, double y) {
  if (ic == IC_SCALAR_CIRCLES) {
    if (np < 1)
      throw std::runtime_error("initial state: scalar circles need a default value");
    for (int64_t k = 1; k + 4 <= np; k += 4) {
      const double dx = x - prm[k], dy = y - prm[k + 1];
      if (dx * dx + dy * dy <= prm[k + 2] * prm[k + 2])
        return prm[k + 3];
    }
    return prm[0];
  }
  if (ic == IC_SCALAR_WAVE) {
    if (np < 3)
      throw std::runtime_error("initial state: a plane wave needs L, kx, ky");
    return std::sin(2.0 * kPi * (prm[1] * x + prm[2] * y) / prm[0]);
  }
  throw std::runtime_error("initial state: unknown 2D scalar initial condition " +
                           std::to_string(ic));
}

}  // namespace

// Orthonormal basis: c_m = sum_q w_q f(x_q) phi_m(q), over-integrated with the
// order max(p, 5) triangle rule or 7 Gauss points on lines
// (src/case/initial_state.cpp:16-22, dg/l2_projection.hpp:36-76).
std::vector<double> project(const Mesh &m, int p, double gamma, int ic, const double *prm,
                            int64_t np) {
  const size_t E = m.elements();
  if (m.dimension == 2 && ic >= IC_SCALAR_PULSE) {
    // project_triangle_field of one variable, transposed to U[E][Np]
    // (initial_state.cpp:39-46, scalar_coefficient_layout.cpp:14-21).
    const int Np = tri_modes(p);
    const TriRule rule = triangle_rule(std::max(p, 5));
    std::vector<double> basis(static_cast<size_t>(rule.size()) * Np);
    for (int q = 0; q < rule.size(); ++q)
      dubiner(rule.xi[q], rule.eta[q], p, basis.data() + q * Np);
    scalar_value_2d(ic, prm, np, 0.0, 0.0);  // validate once on this thread
    std::vector<double> U(static_cast<size_t>(Np) * E, 0.0);
    parallel_for(E, [&](size_t begin, size_t end) {
      for (size_t e = begin; e < end; ++e) {
        const Triangle &t = m.tris[e];
        for (int q = 0; q < rule.size(); ++q) {
          double x, y;
          t.to_physical(rule.xi[q], rule.eta[q], x, y);
          const double f = scalar_value_2d(ic, prm, np, x, y);
          for (int k = 0; k < Np; ++k)
            U[e * Np + k] += rule.w[q] * f * basis[q * Np + k];
        }
      }
    });
    return U;
  }
  if (m.dimension == 2) {
    const int Np = tri_modes(p);
    const TriRule rule = triangle_rule(std::max(p, 5));
    std::vector<double> basis(static_cast<size_t>(rule.size()) * Np);
    for (int q = 0; q < rule.size(); ++q)
      dubiner(rule.xi[q], rule.eta[q], p, basis.data() + q * Np);
    std::vector<double> U(4 * Np * E, 0.0);
    euler_value(ic, prm, np, gamma, 0.0, 0.0);  // validate once on this thread
    parallel_for(E, [&](size_t begin, size_t end) {
      for (size_t e = begin; e < end; ++e) {
        const Triangle &t = m.tris[e];
        for (int q = 0; q < rule.size(); ++q) {
          double x, y;
          t.to_physical(rule.xi[q], rule.eta[q], x, y);
          const Prim w = euler_value(ic, prm, np, gamma, x, y);
          const double ke = 0.5 * w.r * (w.u * w.u + w.v * w.v);
          const double s[4] = {w.r, w.r * w.u, w.r * w.v, w.p / (gamma - 1.0) + ke};
          for (int v = 0; v < 4; ++v)
            for (int k = 0; k < Np; ++k)
              U[(static_cast<size_t>(v) * Np + k) * E + e] +=
                  rule.w[q] * s[v] * basis[q * Np + k];
        }
      }
    });
    return U;
  }
  const int Np = p + 1;
  const LineRule rule = gauss_legendre(7);
  std::vector<double> basis(static_cast<size_t>(rule.size()) * Np);
  for (int q = 0; q < rule.size(); ++q)
    legendre(rule.x[q], p, basis.data() + q * Np);
  if (ic < IC_SCALAR_PULSE) {
    // project_line_field of the conserved state, U[4][Np][E]
    // (initial_state.cpp:48-54; euler_initial_value with y = 0)
    std::vector<double> U(4 * Np * E, 0.0);
    euler_value(ic, prm, np, gamma, 0.0, 0.0);
    parallel_for(E, [&](size_t begin, size_t end) {
      for (size_t e = begin; e < end; ++e) {
        const double xl = m.lines[e].xl, xr = m.lines[e].xr;
        for (int q = 0; q < rule.size(); ++q) {
          const double x = (xl + xr) / 2.0 + (xr - xl) / 2.0 * rule.x[q];
          const Prim w = euler_value(ic, prm, np, gamma, x, 0.0);
          const double ke = 0.5 * w.r * (w.u * w.u + w.v * w.v);
          const double s[4] = {w.r, w.r * w.u, w.r * w.v, w.p / (gamma - 1.0) + ke};
          for (int v = 0; v < 4; ++v)
            for (int k = 0; k < Np; ++k)
              U[(static_cast<size_t>(v) * Np + k) * E + e] += rule.w[q] * s[v] * basis[q * Np + k];
        }
      }
    });
    return U;
  }
  scalar_value(ic, prm, np, 0.0);
  std::vector<double> U(Np * E, 0.0);
  parallel_for(E, [&](size_t begin, size_t end) {
    for (size_t e = begin; e < end; ++e) {
      const double xl = m.lines[e].xl, xr = m.lines[e].xr;
      for (int q = 0; q < rule.size(); ++q) {
        const double x = (xl + xr) / 2.0 + (xr - xl) / 2.0 * rule.x[q];
        const double f = scalar_value(ic, prm, np, x);
        for (int k = 0; k < Np; ++k)
          U[e * Np + k] += rule.w[q] * f * basis[q * Np + k];
      }
    }
  });
  return U;
}

}  // namespace whip
