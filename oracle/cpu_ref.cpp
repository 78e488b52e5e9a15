// CPU oracle: a single-threaded C++ restatement of the reference RKDG hot
// path, written against the reference's behaviour file by file.
//
// TEST INFRASTRUCTURE ONLY (see oracle_api.h). It is the checker for the HIP
// engine in wasl_amd/ and the "port" CPU baseline; it is never linked into
// the product.
//
// Arithmetic follows the reference's expression and loop order so that, when
// both are compiled with -ffp-contract=off, results agree bit for bit with
// oracle/_ref (the reference compiled from its own sources). Each block cites
// the reference file:line it restates (paths relative to cpp/).

#include "oracle_api.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace {

thread_local std::string g_error;

using Vec = std::vector<double>;

constexpr double kPi = 3.141592653589793238462643383279502884;

// ---------------------------------------------------------------- basis ----

// include/wasl/basis/jacobi.hpp:12-27
double jacobi(double x, int n, double alpha, double beta) {
  if (n <= 0)
    return 1.0;
  double prev = 1.0;
  double curr = 0.5 * (alpha - beta + (alpha + beta + 2.0) * x);
  for (int k = 1; k < n; ++k) {
    const double s = 2.0 * k + alpha + beta;
    const double a = 2.0 * (k + 1) * (k + alpha + beta + 1.0) * s;
    const double b = (s + 1.0) * (s * (s + 2.0) * x + (alpha * alpha - beta * beta));
    const double c = 2.0 * (k + alpha) * (k + beta) * (s + 2.0);
    const double next = (b * curr - c * prev) / a;
    prev = curr;
    curr = next;
  }
  return curr;
}

// include/wasl/basis/jacobi.hpp:31-35
double jacobi_derivative(double x, int n, double alpha, double beta) {
  if (n <= 0)
    return 0.0;
  return 0.5 * (n + alpha + beta + 1.0) * jacobi(x, n - 1, alpha + 1.0, beta + 1.0);
}

// include/wasl/basis/legendre.hpp:20-38
double legendre_norm(int j) { return std::sqrt((2.0 * j + 1.0) / 2.0); }

void legendre(double xi, int p, double *v) {
  v[0] = legendre_norm(0);
  if (p <= 0)
    return;
  double prev = 1.0, curr = xi;
  v[1] = legendre_norm(1) * curr;
  for (int j = 1; j < p; ++j) {
    const double next = ((2.0 * j + 1.0) * xi * curr - j * prev) / (j + 1);
    prev = curr;
    curr = next;
    v[j + 1] = legendre_norm(j + 1) * curr;
  }
}

// include/wasl/basis/legendre.hpp:41-60
void legendre_gradient(double xi, int p, double *g) {
  g[0] = 0.0;
  if (p <= 0)
    return;
  double prev = 1.0, curr = xi, dprev = 0.0, dcurr = 1.0;
  g[1] = legendre_norm(1) * dcurr;
  for (int j = 1; j < p; ++j) {
    const double next = ((2.0 * j + 1.0) * xi * curr - j * prev) / (j + 1);
    const double dnext = ((2.0 * j + 1.0) * (curr + xi * dcurr) - j * dprev) / (j + 1);
    prev = curr;
    curr = next;
    dprev = dcurr;
    dcurr = dnext;
    g[j + 1] = legendre_norm(j + 1) * dcurr;
  }
}

// include/wasl/basis/dubiner_triangle.hpp:37-43
double collapsed_a(double xi, double eta) {
  return (eta < 1.0) ? (2.0 + 2.0 * xi) / (1.0 - eta) - 1.0 : -1.0;
}
double dubiner_norm(int i, int n) { return std::sqrt((2.0 * i + 1.0) * (n + 1.0) / 2.0); }

// include/wasl/basis/dubiner_triangle.hpp:49-64
void dubiner(double xi, double eta, int p, double *v) {
  const double b = eta;
  const double a = collapsed_a(xi, eta);
  const double half = 0.5 * (1.0 - b);
  int k = 0;
  for (int n = 0; n <= p; ++n)
    for (int i = 0; i <= n; ++i) {
      const int j = n - i;
      const double pi = jacobi(a, i, 0.0, 0.0);
      const double warp = std::pow(half, static_cast<double>(i));
      const double qj = jacobi(b, j, 2.0 * i + 1.0, 0.0);
      v[k++] = dubiner_norm(i, n) * pi * warp * qj;
    }
}

// include/wasl/basis/dubiner_triangle.hpp:70-105
void dubiner_gradient(double xi, double eta, int p, double *gx, double *ge) {
  const double b = eta;
  const double a = collapsed_a(xi, eta);
  const double half = 0.5 * (1.0 - b);
  const double inv = 1.0 / (1.0 - b);
  const double da_dxi = 2.0 * inv;
  const double da_deta = (a + 1.0) * inv;
  int k = 0;
  for (int n = 0; n <= p; ++n)
    for (int i = 0; i <= n; ++i) {
      const int j = n - i;
      const double norm = dubiner_norm(i, n);
      const double pi = jacobi(a, i, 0.0, 0.0);
      const double warp = std::pow(half, static_cast<double>(i));
      const double qj = jacobi(b, j, 2.0 * i + 1.0, 0.0);
      const double dpi = jacobi_derivative(a, i, 0.0, 0.0);
      const double dqj = jacobi_derivative(b, j, 2.0 * i + 1.0, 0.0);
      double dwarp = 0.0;
      if (i > 0)
        dwarp = -0.5 * static_cast<double>(i) * std::pow(half, static_cast<double>(i - 1));
      const double dphi_da = norm * dpi * warp * qj;
      const double dphi_db = norm * (pi * dwarp * qj + pi * warp * dqj);
      gx[k] = dphi_da * da_dxi;
      ge[k] = dphi_da * da_deta + dphi_db;
      ++k;
    }
}

int tri_modes(int p) { return (p + 1) * (p + 2) / 2; }

// ----------------------------------------------------------- quadrature ----

// include/wasl/quadrature/gauss_legendre.hpp:24-65. The node/weight constants
// are data: the same tabulated Gauss-Legendre values.
struct Rule {
  Vec x, w;
  int size() const { return static_cast<int>(x.size()); }
};

Rule gauss(int n) {
  static const std::vector<std::vector<std::pair<double, double>>> halves = {
      {},
      {{0.0, 2.0}},
      {{0.5773502691896257, 1.0}},
      {{0.0, 0.8888888888888888}, {0.7745966692414834, 0.5555555555555556}},
      {{0.33998104358485626, 0.6521451548625461}, {0.8611363115940526, 0.34785484513745385}},
      {{0.0, 0.5688888888888889},
       {0.5384693101056831, 0.47862867049936647},
       {0.9061798459386640, 0.2369268850561891}},
      {{0.2386191860831969, 0.4679139345726910},
       {0.6612093864662645, 0.3607615730481386},
       {0.9324695142031521, 0.17132449237917036}},
      {{0.0, 0.4179591836734694},
       {0.4058451513773972, 0.3818300505051188},
       {0.7415311855993945, 0.2797053914892766},
       {0.9491079123427585, 0.1294849661688697}},
  };
  if (n < 1 || n >= static_cast<int>(halves.size()))
    throw std::out_of_range("gauss_legendre: unsupported point count");
  Rule r;
  const auto &h = halves[n];
  for (auto it = h.rbegin(); it != h.rend(); ++it) {
    if (it->first == 0.0)
      continue;
    r.x.push_back(-it->first);
    r.w.push_back(it->second);
  }
  for (const auto &e : h) {
    r.x.push_back(e.first);
    r.w.push_back(e.second);
  }
  return r;
}

struct TriRule {
  Vec xi, eta, w;
  int size() const { return static_cast<int>(xi.size()); }
};

// include/wasl/quadrature/triangle_volume_rule.hpp:34-51
TriRule tri_rule(int p) {
  const int n = p + 1;
  const Rule b = gauss(n), a = gauss(n);
  TriRule r;
  for (int qb = 0; qb < b.size(); ++qb) {
    const double eta = b.x[qb];
    const double jac = 0.5 * (1.0 - eta);
    for (int qa = 0; qa < a.size(); ++qa) {
      r.xi.push_back(0.5 * (a.x[qa] + 1.0) * (1.0 - eta) - 1.0);
      r.eta.push_back(eta);
      r.w.push_back(b.w[qb] * a.w[qa] * jac);
    }
  }
  return r;
}

// include/wasl/quadrature/triangle_edge_rules.hpp:28-58
struct EdgeRule {
  Vec xi, eta, w;
};

std::vector<EdgeRule> edge_rules(int p) {
  const Rule g = gauss(p + 1);
  const double inv_sqrt2 = 1.0 / std::sqrt(2.0);
  std::vector<EdgeRule> e(3);
  for (int q = 0; q < g.size(); ++q) {
    const double s = g.x[q], w = g.w[q];
    e[0].xi.push_back(s);
    e[0].eta.push_back(-1.0);
    e[0].w.push_back(w);
    e[1].xi.push_back(1.0 + 0.5 * (s + 1.0) * (-2.0));
    e[1].eta.push_back(-1.0 + 0.5 * (s + 1.0) * (2.0));
    e[1].w.push_back(w * 2.0 * inv_sqrt2);
    e[2].xi.push_back(-1.0);
    e[2].eta.push_back(s);
    e[2].w.push_back(w);
  }
  return e;
}

// ------------------------------------------------------------- geometry ----

// include/wasl/mesh/triangle/triangle_geometry.hpp:24-116
double reference_edge_length(int edge) { return edge == 1 ? 2.0 * std::sqrt(2.0) : 2.0; }

struct TriMap {
  double x0, y0, x1, y1, x2, y2;
  double ax() const { return x1 - x0; }
  double ay() const { return y1 - y0; }
  double bx() const { return x2 - x0; }
  double by() const { return y2 - y0; }
  double det_a() const { return ax() * by() - bx() * ay(); }
  double jacobian_determinant() const { return det_a() / 4.0; }
  int orientation() const { return det_a() > 0.0 ? 1 : -1; }
  void to_physical(double xi, double eta, double &x, double &y) const {
    x = x0 + 0.5 * (ax() * (xi + 1.0) + bx() * (eta + 1.0));
    y = y0 + 0.5 * (ay() * (xi + 1.0) + by() * (eta + 1.0));
  }
  void to_reference(double x, double y, double &xi, double &eta) const {
    const double dx = x - x0, dy = y - y0;
    const double inv = 1.0 / det_a();
    const double s = (by() * dx - bx() * dy) * inv;
    const double t = (-ay() * dx + ax() * dy) * inv;
    xi = 2.0 * s - 1.0;
    eta = 2.0 * t - 1.0;
  }
  void scaled_gradient(double gxi, double geta, double &gx, double &gy) const {
    gx = 0.5 * (by() * gxi - ay() * geta);
    gy = 0.5 * (-bx() * gxi + ax() * geta);
  }
  void edge_vector(int edge, double &ex, double &ey) const {
    if (edge == 0) {
      ex = x1 - x0;
      ey = y1 - y0;
    } else if (edge == 1) {
      ex = x2 - x1;
      ey = y2 - y1;
    } else {
      ex = x0 - x2;
      ey = y0 - y2;
    }
  }
  double edge_length(int edge) const {
    double ex, ey;
    edge_vector(edge, ex, ey);
    return std::hypot(ex, ey);
  }
  void edge_normal(int edge, double &nx, double &ny) const {
    double ex, ey;
    edge_vector(edge, ex, ey);
    const double len = std::hypot(ex, ey);
    const double tx = ex / len, ty = ey / len;
    const int sign = orientation();
    nx = sign * ty;
    ny = sign * (-tx);
  }
  void shift(double dx, double dy) {
    x0 += dx;
    x1 += dx;
    x2 += dx;
    y0 += dy;
    y1 += dy;
    y2 += dy;
  }
};

// ------------------------------------------------------------------ mesh ----

struct Tri {
  int64_t node[3];
  double x[3], y[3];
  TriMap map() const { return {x[0], y[0], x[1], y[1], x[2], y[2]}; }
};

struct Face {
  int32_t left = -1, right = -1;
  int left_edge = 0, right_edge = 0;
  int64_t lv0 = 0, lv1 = 0, rv0 = 0, rv1 = 0;
  double nx = 0.0, ny = 0.0;
  bool boundary() const { return right == -1; }
};

constexpr int kEdgeVertices[3][2] = {{0, 1}, {1, 2}, {2, 0}};

// src/mesh/triangle/build_triangle_mesh.cpp:38-148. Orientation by signed
// area, edge sharing keyed on the sorted vertex pair, left side = smaller
// element index (the first element in element order to touch the edge),
// faces sorted by (left element, left local edge).
void build_triangles(int64_t node_count, const int64_t *tags, const double *xy, int64_t count,
                     const int64_t *tri_nodes, std::vector<Tri> &elements,
                     std::vector<Face> &faces) {
  std::map<int64_t, int64_t> lookup;
  for (int64_t i = 0; i < node_count; ++i)
    lookup[tags[i]] = i;
  const auto node = [&](int64_t tag) -> int64_t {
    auto it = lookup.find(tag);
    if (it == lookup.end())
      throw std::runtime_error("triangle mesh: missing node " + std::to_string(tag));
    return it->second;
  };
  elements.clear();
  for (int64_t e = 0; e < count; ++e) {
    int64_t t[3] = {tri_nodes[3 * e], tri_nodes[3 * e + 1], tri_nodes[3 * e + 2]};
    int64_t v[3] = {node(t[0]), node(t[1]), node(t[2])};
    const double area2 = xy[2 * v[0]] * (xy[2 * v[1] + 1] - xy[2 * v[2] + 1]) +
                         xy[2 * v[1]] * (xy[2 * v[2] + 1] - xy[2 * v[0] + 1]) +
                         xy[2 * v[2]] * (xy[2 * v[0] + 1] - xy[2 * v[1] + 1]);
    if (area2 == 0.0)
      throw std::runtime_error("triangle mesh: zero-area element " + std::to_string(e + 1));
    if (area2 < 0.0) {
      std::swap(v[1], v[2]);
      std::swap(t[1], t[2]);
    }
    Tri tri;
    for (int k = 0; k < 3; ++k) {
      tri.node[k] = t[k];
      tri.x[k] = xy[2 * v[k]];
      tri.y[k] = xy[2 * v[k] + 1];
    }
    elements.push_back(tri);
  }
  struct Side {
    int32_t element;
    int edge;
    int64_t v0, v1;
  };
  std::map<std::pair<int64_t, int64_t>, std::vector<Side>> edges;
  for (size_t e = 0; e < elements.size(); ++e)
    for (int k = 0; k < 3; ++k) {
      const int64_t a = elements[e].node[kEdgeVertices[k][0]];
      const int64_t b = elements[e].node[kEdgeVertices[k][1]];
      auto &sides = edges[a < b ? std::make_pair(a, b) : std::make_pair(b, a)];
      if (sides.size() == 2)
        throw std::runtime_error("triangle mesh: edge has more than two adjacent elements "
                                 "(non-conforming mesh)");
      sides.push_back({static_cast<int32_t>(e), k, a, b});
    }
  faces.clear();
  for (const auto &[key, sides] : edges) {
    (void)key;
    const Side &l = sides[0];
    Face f;
    f.left = l.element;
    f.left_edge = l.edge;
    f.lv0 = l.v0;
    f.lv1 = l.v1;
    if (sides.size() == 2) {
      f.right = sides[1].element;
      f.right_edge = sides[1].edge;
      f.rv0 = sides[1].v0;
      f.rv1 = sides[1].v1;
    }
    elements[static_cast<size_t>(l.element)].map().edge_normal(l.edge, f.nx, f.ny);
    faces.push_back(f);
  }
  std::sort(faces.begin(), faces.end(), [](const Face &a, const Face &b) {
    return a.left != b.left ? a.left < b.left : a.left_edge < b.left_edge;
  });
}

// include/wasl/dg/triangle_face_trace.hpp:16-33
void neighbor_edge_point(const Face &f, double lxi, double leta, double &rxi, double &reta) {
  double t = f.left_edge == 0 ? lxi : f.left_edge == 1 ? leta : -leta;
  if (f.lv0 != f.rv0)
    t = -t;
  rxi = -1.0;
  reta = -1.0;
  if (f.right_edge == 0) {
    rxi = t;
  } else if (f.right_edge == 1) {
    rxi = -t;
    reta = t;
  } else {
    reta = -t;
  }
}

struct Box {
  double min_x = std::numeric_limits<double>::max();
  double max_x = std::numeric_limits<double>::lowest();
  double min_y = std::numeric_limits<double>::max();
  double max_y = std::numeric_limits<double>::lowest();
};

// include/wasl/mesh/triangle/bounding_box.hpp:22-31
Box bounding_box(const std::vector<Tri> &elements) {
  Box b;
  for (const auto &t : elements) {
    b.min_x = std::min({b.min_x, t.x[0], t.x[1], t.x[2]});
    b.max_x = std::max({b.max_x, t.x[0], t.x[1], t.x[2]});
    b.min_y = std::min({b.min_y, t.y[0], t.y[1], t.y[2]});
    b.max_y = std::max({b.max_y, t.y[0], t.y[1], t.y[2]});
  }
  return b;
}

// --------------------------------------------------------------- physics ----

struct Cons {
  double r, mx, my, E;
  double &operator[](int v) { return v == 0 ? r : v == 1 ? mx : v == 2 ? my : E; }
  double operator[](int v) const { return v == 0 ? r : v == 1 ? mx : v == 2 ? my : E; }
};
struct Prim {
  double r, u, v, p;
};

// include/wasl/equations/euler/primitive.hpp:11-40
Cons conserved(double g, const Prim &s) {
  const double ke = 0.5 * s.r * (s.u * s.u + s.v * s.v);
  return {s.r, s.r * s.u, s.r * s.v, s.p / (g - 1.0) + ke};
}
Prim primitive(double g, const Cons &s) {
  const double inv = 1.0 / s.r;
  const double u = s.mx * inv, v = s.my * inv;
  const double ke = 0.5 * (s.mx * u + s.my * v);
  return {s.r, u, v, (g - 1.0) * (s.E - ke)};
}
double sound_speed(double g, const Prim &s) { return std::sqrt(g * s.p / s.r); }
bool is_physical(double g, const Cons &s) { return s.r > 0.0 && primitive(g, s).p > 0.0; }

// include/wasl/equations/euler/flux.hpp:11-33
Cons flux_x(double g, const Cons &s) {
  const Prim w = primitive(g, s);
  return {s.mx, s.mx * w.u + w.p, s.my * w.u, (s.E + w.p) * w.u};
}
Cons flux_y(double g, const Cons &s) {
  const Prim w = primitive(g, s);
  return {s.my, s.mx * w.v, s.my * w.v + w.p, (s.E + w.p) * w.v};
}
Cons normal_flux(double g, const Cons &s, double nx, double ny) {
  const Cons fx = flux_x(g, s), fy = flux_y(g, s);
  return {nx * fx.r + ny * fy.r, nx * fx.mx + ny * fy.mx, nx * fx.my + ny * fy.my,
          nx * fx.E + ny * fy.E};
}

// include/wasl/flux/hllc.hpp:12-74
Cons hllc_star(const Cons &s, const Prim &w, double nx, double ny, double S, double Sm) {
  const double un = w.u * nx + w.v * ny;
  const double rs = s.r * (S - un) / (S - Sm);
  const double shift = Sm - un;
  const double es = s.E / s.r + shift * (Sm + w.p / (s.r * (S - un)));
  return {rs, rs * (w.u + shift * nx), rs * (w.v + shift * ny), rs * es};
}
Cons add_jump(const Cons &f, double S, const Cons &star, const Cons &s) {
  return {f.r + S * (star.r - s.r), f.mx + S * (star.mx - s.mx), f.my + S * (star.my - s.my),
          f.E + S * (star.E - s.E)};
}
Cons hllc(double g, const Cons &L, const Cons &R, double nx, double ny) {
  const Prim wl = primitive(g, L), wr = primitive(g, R);
  const double ul = wl.u * nx + wl.v * ny, ur = wr.u * nx + wr.v * ny;
  const double cl = sound_speed(g, wl), cr = sound_speed(g, wr);
  const double SL = std::min(ul - cl, ur - cr);
  const double SR = std::max(ul + cl, ur + cr);
  const Cons FL = normal_flux(g, L, nx, ny);
  if (SL >= 0.0)
    return FL;
  const Cons FR = normal_flux(g, R, nx, ny);
  if (SR <= 0.0)
    return FR;
  const double Sm = (wr.p - wl.p + L.r * ul * (SL - ul) - R.r * ur * (SR - ur)) /
                    (L.r * (SL - ul) - R.r * (SR - ur));
  if (Sm >= 0.0)
    return add_jump(FL, SL, hllc_star(L, wl, nx, ny, SL, Sm), L);
  return add_jump(FR, SR, hllc_star(R, wr, nx, ny, SR, Sm), R);
}

// ------------------------------------------------------------ limiter ----

// include/wasl/limiting/weno/euler_characteristic.hpp:19-67
void to_characteristic(double g, const Cons &avg, const Cons &d, double nx, double ny,
                       double *w) {
  const Prim m = primitive(g, avg);
  const double un = m.u * nx + m.v * ny;
  const double ut = -m.u * ny + m.v * nx;
  const double mn = d.mx * nx + d.my * ny;
  const double mt = -d.mx * ny + d.my * nx;
  const double c2 = g * m.p / m.r;
  const double c = std::sqrt(c2);
  const double ke = 0.5 * (un * un + ut * ut);
  const double factor = (g - 1.0) / c2;
  const double common = factor * (ke * d.r - un * mn - ut * mt + d.E);
  const double diff = (un * d.r - mn) / c;
  w[0] = 0.5 * (common + diff);
  w[1] = d.r - common;
  w[2] = mt - ut * d.r;
  w[3] = 0.5 * (common - diff);
}

Cons from_characteristic(double g, const Cons &avg, const double *w, double nx, double ny) {
  const Prim m = primitive(g, avg);
  const double un = m.u * nx + m.v * ny;
  const double ut = -m.u * ny + m.v * nx;
  const double c = sound_speed(g, m);
  const double H = (avg.E + m.p) / avg.r;
  const double ke = 0.5 * (un * un + ut * ut);
  const double r = w[0] + w[1] + w[3];
  const double mn = (un - c) * w[0] + un * w[1] + (un + c) * w[3];
  const double mt = ut * r + w[2];
  const double E = (H - un * c) * w[0] + ke * w[1] + ut * w[2] + (H + un * c) * w[3];
  return {r, nx * mn - ny * mt, ny * mn + nx * mt, E};
}

// include/wasl/limiting/weno/simple_weno.hpp:11-48
double smoothness(const double *c, int n) {
  double v = 0.0;
  for (int m = 1; m < n; ++m)
    v += c[m] * c[m];
  return v;
}

void simple_weno(const double *target, const double *cand, int K, int n, double *out) {
  const double eps = 1e-12;
  if (K == 0) {
    for (int m = 0; m < n; ++m)
      out[m] = target[m];
    return;
  }
  double w[4];
  const double d0 = eps + smoothness(target, n);
  w[0] = 0.5 / (d0 * d0);
  double sum = w[0];
  const double ideal = 0.5 / static_cast<double>(K);
  for (int k = 0; k < K; ++k) {
    const double d = eps + smoothness(cand + k * n, n);
    w[k + 1] = ideal / (d * d);
    sum += w[k + 1];
  }
  out[0] = target[0];
  for (int m = 1; m < n; ++m) {
    double v = w[0] * target[m];
    for (int k = 0; k < K; ++k)
      v += w[k + 1] * cand[k * n + m];
    out[m] = v / sum;
  }
}

// include/wasl/limiting/weno/euler_characteristic_weno.hpp:17-67
void characteristic_weno(double g, const Cons &avg, double nx, double ny, const double *target,
                         const double *cand, int K, int n, double *out) {
  std::vector<double> tc(4 * n), cc(static_cast<size_t>(K) * 4 * n), lc(4 * n);
  const auto project = [&](const double *src, double *dst) {
    for (int m = 0; m < n; ++m) {
      const Cons d{src[m], src[n + m], src[2 * n + m], src[3 * n + m]};
      double w[4];
      to_characteristic(g, avg, d, nx, ny, w);
      for (int wave = 0; wave < 4; ++wave)
        dst[wave * n + m] = w[wave];
    }
  };
  project(target, tc.data());
  for (int k = 0; k < K; ++k)
    project(cand + k * 4 * n, cc.data() + k * 4 * n);
  std::vector<double> wc(static_cast<size_t>(K) * n);
  for (int wave = 0; wave < 4; ++wave) {
    for (int k = 0; k < K; ++k)
      for (int m = 0; m < n; ++m)
        wc[k * n + m] = cc[k * 4 * n + wave * n + m];
    simple_weno(tc.data() + wave * n, wc.data(), K, n, lc.data() + wave * n);
  }
  for (int m = 0; m < n; ++m) {
    double w[4];
    for (int wave = 0; wave < 4; ++wave)
      w[wave] = lc[wave * n + m];
    const Cons s = from_characteristic(g, avg, w, nx, ny);
    out[m] = s.r;
    out[n + m] = s.mx;
    out[2 * n + m] = s.my;
    out[3 * n + m] = s.E;
  }
  for (int v = 0; v < 4; ++v)
    out[v * n] = target[v * n];
}

// include/wasl/limiting/positivity/euler_positivity.hpp:21-102
Cons eval_packed(const double *c, int n, const double *basis) {
  Cons s{0.0, 0.0, 0.0, 0.0};
  for (int v = 0; v < 4; ++v) {
    double acc = 0.0;
    for (int m = 0; m < n; ++m)
      acc += c[v * n + m] * basis[m];
    s[v] = acc;
  }
  return s;
}

Cons interpolate(const Cons &a, const Cons &p, double t) {
  return {a.r + t * (p.r - a.r), a.mx + t * (p.mx - a.mx), a.my + t * (p.my - a.my),
          a.E + t * (p.E - a.E)};
}

void scale_high(double *c, int n, double s) {
  for (int v = 0; v < 4; ++v)
    for (int m = 1; m < n; ++m)
      c[v * n + m] *= s;
}

double pressure_theta(double g, const Cons &avg, const Cons &pt, double floor) {
  double lo = 0.0, hi = 1.0;
  for (int it = 0; it < 64; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (primitive(g, interpolate(avg, pt, mid)).p >= floor)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

double positivity(double g, double *c, int n, const double *basis, int points, double rf,
                  double pf) {
  if (n == 0 || rf < 0.0 || pf < 0.0)
    throw std::invalid_argument("Euler positivity limiter: invalid modes or floors");
  const double is2 = 1.0 / std::sqrt(2.0);
  const Cons avg{c[0] * is2, c[n] * is2, c[2 * n] * is2, c[3 * n] * is2};
  if (avg.r <= rf || primitive(g, avg).p <= pf)
    throw std::invalid_argument("Euler positivity limiter: cell average is below a floor");
  double ds = 1.0;
  for (int q = 0; q < points; ++q) {
    const Cons s = eval_packed(c, n, basis + q * n);
    if (s.r < rf)
      ds = std::min(ds, (avg.r - rf) / (avg.r - s.r));
  }
  scale_high(c, n, ds);
  double ps = 1.0;
  for (int q = 0; q < points; ++q) {
    const Cons s = eval_packed(c, n, basis + q * n);
    if (primitive(g, s).p >= pf)
      continue;
    ps = std::min(ps, pressure_theta(g, avg, s, pf));
  }
  scale_high(c, n, ps);
  return ds * ps;
}

// src/limiting/limiter_stencil_assembly.hpp:28-47
void append_unique(std::vector<int32_t> &nbrs, int K, size_t e, int32_t n) {
  int32_t *v = nbrs.data() + e * K;
  for (int i = 0; i < K; ++i) {
    if (v[i] == n)
      return;
    if (v[i] == -1) {
      v[i] = n;
      return;
    }
  }
}

double periodic_offset(double gap, double period) {
  return gap > 0.5 * period ? period : gap < -0.5 * period ? -period : 0.0;
}

struct Stencil {
  int K = 0;
  int n = 0;
  int highest_begin = 0;
  std::vector<int32_t> nbrs;
  Vec transforms, nx, ny, check;
  int checks() const { return n ? static_cast<int>(check.size()) / n : 0; }
};

// ------------------------------------------------------------------ case ----

struct LineElem {
  int64_t ln, rn;
  double xl, xr;
};
struct LineFace {
  int32_t left, right;
  double x, normal;
};

}  // namespace

struct orc_case {
  int dim = 2;
  bool scalar = false;  // 2D scalar advection (U[element][mode]) instead of Euler
  int p = 0;
  int n = 0;  // modes
  int V = 4;  // variables
  double gamma = 1.4;
  int boundary = ORC_PERIODIC;
  // triangles
  std::vector<Tri> tris;
  std::vector<Face> faces;
  std::vector<int32_t> donors;
  Vec donor_basis;
  double vx = 0.0, vy = 0.0;  // 2D scalar advection velocity
  // lines
  std::vector<LineElem> lines;
  std::vector<LineFace> lfaces;
  double wavespeed = 0.0;
  int left_kind = ORC_PERIODIC, right_kind = ORC_PERIODIC;
  double left_value = 0.0, right_value = 0.0;
  Stencil stencil;
  bool has_stencil = false;
  Vec stage, rhs, euler;

  size_t E() const { return dim == 2 ? tris.size() : lines.size(); }
  bool is_euler() const { return V == 4; }
  size_t F() const { return dim == 2 ? faces.size() : lfaces.size(); }
  size_t state_size() const { return static_cast<size_t>(V) * n * E(); }
};

namespace {

// src/boundaries/periodic_triangle_trace.cpp:17-67
void build_trace(orc_case &c) {
  const int P = c.p + 1;
  c.donors.assign(c.faces.size() * P, -1);
  c.donor_basis.assign(c.faces.size() * P * c.n, 0.0);
  if (c.boundary != ORC_PERIODIC)
    return;
  const auto rules = edge_rules(c.p);
  const Box box = bounding_box(c.tris);
  for (size_t f = 0; f < c.faces.size(); ++f) {
    const Face &face = c.faces[f];
    if (!face.boundary())
      continue;
    const TriMap map = c.tris[face.left].map();
    const EdgeRule &rule = rules[face.left_edge];
    for (int q = 0; q < P; ++q) {
      double x, y;
      map.to_physical(rule.xi[q], rule.eta[q], x, y);
      if (std::fabs(face.nx) > 0.5 && std::fabs(face.ny) < 1e-10)
        x += face.nx < 0.0 ? box.max_x - box.min_x : box.min_x - box.max_x;
      else if (std::fabs(face.ny) > 0.5 && std::fabs(face.nx) < 1e-10)
        y += face.ny < 0.0 ? box.max_y - box.min_y : box.min_y - box.max_y;
      else
        throw std::runtime_error("simulation: 2D periodic boundaries require an axis-aligned box");
      const size_t idx = f * P + q;
      double xi = 0.0, eta = 0.0;
      int32_t donor = -1;
      for (size_t e = 0; e < c.tris.size(); ++e) {
        c.tris[e].map().to_reference(x, y, xi, eta);
        if (xi >= -1.0 - 1e-10 && eta >= -1.0 - 1e-10 && xi + eta <= 1e-10) {
          donor = static_cast<int32_t>(e);
          break;
        }
      }
      if (donor < 0)
        throw std::runtime_error("simulation: periodic 2D trace has no donor element");
      c.donors[idx] = donor;
      dubiner(xi, eta, c.p, c.donor_basis.data() + idx * c.n);
    }
  }
}

// Exterior per [variable][face][point]: periodic donors
// (periodic_triangle_trace.cpp:84-105) or, for the transmissive boundary, the
// owner's own trace, so HLLC(U, U, n) = F(U).n.
Vec tri_exterior(const orc_case &c, const double *U) {
  const int P = c.p + 1;
  const size_t E = c.E(), F = c.faces.size();
  Vec ext(4 * F * P, 0.0);
  if (c.boundary == ORC_PERIODIC) {
    for (size_t pt = 0; pt < c.donors.size(); ++pt) {
      if (c.donors[pt] < 0)
        continue;
      const size_t d = static_cast<size_t>(c.donors[pt]);
      const size_t f = pt / P, q = pt % P;
      for (int v = 0; v < 4; ++v)
        for (int m = 0; m < c.n; ++m)
          ext[(v * F + f) * P + q] +=
              U[(static_cast<size_t>(v) * c.n + m) * E + d] * c.donor_basis[pt * c.n + m];
    }
  } else {
    const auto rules = edge_rules(c.p);
    Vec b(c.n);
    for (size_t f = 0; f < F; ++f) {
      const Face &face = c.faces[f];
      if (!face.boundary())
        continue;
      const EdgeRule &rule = rules[face.left_edge];
      for (int q = 0; q < P; ++q) {
        dubiner(rule.xi[q], rule.eta[q], c.p, b.data());
        for (int v = 0; v < 4; ++v) {
          double acc = 0.0;
          for (int m = 0; m < c.n; ++m)
            acc += U[(static_cast<size_t>(v) * c.n + m) * E + face.left] * b[m];
          ext[(v * F + f) * P + q] = acc;
        }
      }
    }
  }
  return ext;
}

// include/wasl/dg/modal_layout.hpp:26-36
Cons eval_state(const double *U, int n, size_t E, size_t e, const double *b) {
  Cons s{0.0, 0.0, 0.0, 0.0};
  for (int v = 0; v < 4; ++v) {
    double acc = 0.0;
    for (int m = 0; m < n; ++m)
      acc += U[(static_cast<size_t>(v) * n + m) * E + e] * b[m];
    s[v] = acc;
  }
  return s;
}

// include/wasl/dg/euler/triangle_residual.hpp:25-107
void tri_euler_residual(const orc_case &c, const double *U, const double *ext, double *rhs) {
  const int n = c.n, P = c.p + 1;
  const size_t E = c.E(), F = c.faces.size();
  const TriRule vol = tri_rule(c.p + 2);
  const auto edges = edge_rules(c.p);
  std::fill(rhs, rhs + 4 * n * E, 0.0);
  Vec b(n), gx(n), ge(n);
  for (size_t e = 0; e < E; ++e) {
    const TriMap map = c.tris[e].map();
    for (int q = 0; q < vol.size(); ++q) {
      dubiner(vol.xi[q], vol.eta[q], c.p, b.data());
      dubiner_gradient(vol.xi[q], vol.eta[q], c.p, gx.data(), ge.data());
      const Cons s = eval_state(U, n, E, e, b.data());
      const Cons fx = flux_x(c.gamma, s), fy = flux_y(c.gamma, s);
      for (int m = 0; m < n; ++m) {
        double dx, dy;
        map.scaled_gradient(gx[m], ge[m], dx, dy);
        for (int v = 0; v < 4; ++v)
          rhs[(static_cast<size_t>(v) * n + m) * E + e] += vol.w[q] * (fx[v] * dx + fy[v] * dy);
      }
    }
  }
  Vec lb(n), rb(n);
  for (size_t f = 0; f < F; ++f) {
    const Face &face = c.faces[f];
    const size_t L = static_cast<size_t>(face.left);
    const EdgeRule &rule = edges[face.left_edge];
    const TriMap lmap = c.tris[L].map();
    const double scale = lmap.edge_length(face.left_edge) / reference_edge_length(face.left_edge);
    for (int q = 0; q < P; ++q) {
      dubiner(rule.xi[q], rule.eta[q], c.p, lb.data());
      const Cons ls = eval_state(U, n, E, L, lb.data());
      Cons rs{0.0, 0.0, 0.0, 0.0};
      if (!face.boundary()) {
        double rxi, reta;
        neighbor_edge_point(face, rule.xi[q], rule.eta[q], rxi, reta);
        dubiner(rxi, reta, c.p, rb.data());
        rs = eval_state(U, n, E, static_cast<size_t>(face.right), rb.data());
      } else {
        for (int v = 0; v < 4; ++v)
          rs[v] = ext[(v * F + f) * P + q];
      }
      const Cons flux = hllc(c.gamma, ls, rs, face.nx, face.ny);
      const double w = rule.w[q] * scale;
      for (int v = 0; v < 4; ++v) {
        const double value = flux[v] * w;
        for (int m = 0; m < n; ++m)
          rhs[(static_cast<size_t>(v) * n + m) * E + L] -= value * lb[m];
        if (!face.boundary()) {
          const size_t R = static_cast<size_t>(face.right);
          for (int m = 0; m < n; ++m)
            rhs[(static_cast<size_t>(v) * n + m) * E + R] += value * rb[m];
        }
      }
    }
  }
  for (size_t e = 0; e < E; ++e) {
    const double inv = 1.0 / c.tris[e].map().jacobian_determinant();
    for (int v = 0; v < 4; ++v)
      for (int m = 0; m < n; ++m)
        rhs[(static_cast<size_t>(v) * n + m) * E + e] *= inv;
  }
}

// include/wasl/basis/modal_expansion.hpp:14-20
double modal_sum(const double *c, const double *b, int n) {
  double v = 0.0;
  for (int m = 0; m < n; ++m)
    v += c[m] * b[m];
  return v;
}

double line_trace(const orc_case &c, const double *U, size_t e, double xi) {
  Vec b(c.n);
  legendre(xi, c.p, b.data());
  return modal_sum(U + e * c.n, b.data(), c.n);
}

// include/wasl/boundaries/line_boundary_conditions.hpp:36-57 via
// line_exterior.hpp:28-43
Vec line_exterior(const orc_case &c, const double *U) {
  Vec ext(c.lfaces.size(), 0.0);
  for (size_t f = 0; f < c.lfaces.size(); ++f) {
    const LineFace &face = c.lfaces[f];
    if (face.right != -1) {
      ext[f] = line_trace(c, U, static_cast<size_t>(face.right), -1.0);
      continue;
    }
    const bool left_end = face.normal < 0.0;
    const int kind = left_end ? c.left_kind : c.right_kind;
    if (kind == ORC_INFLOW)
      ext[f] = left_end ? c.left_value : c.right_value;
    else if (kind == ORC_OUTFLOW)
      ext[f] = line_trace(c, U, static_cast<size_t>(face.left), left_end ? -1.0 : 1.0);
    else
      ext[f] = left_end ? line_trace(c, U, c.lines.size() - 1, 1.0) : line_trace(c, U, 0, -1.0);
  }
  return ext;
}

// include/wasl/equations/scalar_advection/scalar_advection.hpp:36-42
double upwind(double a, double l, double r) { return a >= 0.0 ? a * l : a * r; }

// include/wasl/dg/scalar_advection/line_residual.hpp:48-111
void line_scalar_residual(const orc_case &c, const double *U, const double *ext, double *rhs) {
  const int n = c.n;
  const size_t E = c.lines.size();
  const Rule vol = gauss(c.p > 0 ? c.p : 1);
  std::fill(rhs, rhs + E * n, 0.0);
  Vec b(n), g(n);
  for (size_t e = 0; e < E; ++e) {
    const double *modes = U + e * n;
    double *r = rhs + e * n;
    for (int q = 0; q < vol.size(); ++q) {
      legendre(vol.x[q], c.p, b.data());
      legendre_gradient(vol.x[q], c.p, g.data());
      const double flux = c.wavespeed * modal_sum(modes, b.data(), n);
      for (int m = 0; m < n; ++m)
        r[m] += vol.w[q] * flux * g[m];
    }
  }
  Vec bl(n), br(n);
  legendre(-1.0, c.p, bl.data());
  legendre(1.0, c.p, br.data());
  for (size_t f = 0; f < c.lfaces.size(); ++f) {
    const LineFace &face = c.lfaces[f];
    const double *lm = U + static_cast<size_t>(face.left) * n;
    double *lr = rhs + static_cast<size_t>(face.left) * n;
    if (face.normal > 0.0) {
      const double interior = modal_sum(lm, br.data(), n);
      const double flux = upwind(c.wavespeed, interior, ext[f]);
      for (int m = 0; m < n; ++m)
        lr[m] -= br[m] * flux;
      if (face.right != -1) {
        double *rr = rhs + static_cast<size_t>(face.right) * n;
        for (int m = 0; m < n; ++m)
          rr[m] += bl[m] * flux;
      }
    } else {
      const double interior = modal_sum(lm, bl.data(), n);
      const double flux = upwind(c.wavespeed, ext[f], interior);
      for (int m = 0; m < n; ++m)
        lr[m] += bl[m] * flux;
    }
  }
  for (size_t e = 0; e < E; ++e) {
    const double inv = 1.0 / ((c.lines[e].xr - c.lines[e].xl) / 2.0);
    for (int m = 0; m < n; ++m)
      rhs[e * n + m] *= inv;
  }
}

// src/boundaries/periodic_triangle_trace.cpp:69-82: exterior per [face][point]
Vec tri_scalar_exterior(const orc_case &c, const double *U) {
  Vec ext(c.donors.size(), 0.0);
  for (size_t pt = 0; pt < c.donors.size(); ++pt) {
    if (c.donors[pt] < 0)
      continue;
    const size_t d = static_cast<size_t>(c.donors[pt]);
    for (int m = 0; m < c.n; ++m)
      ext[pt] += U[d * c.n + m] * c.donor_basis[pt * c.n + m];
  }
  return ext;
}

// include/wasl/equations/scalar_advection/scalar_advection.hpp:45-50
double upwind_normal(double vx, double vy, double nx, double ny, double l, double r) {
  const double ns = vx * nx + vy * ny;
  return ns * (ns >= 0.0 ? l : r);
}

// include/wasl/dg/scalar_advection/triangle_residual.hpp:25-105 (constant
// velocity). The volume rule has (p+1)^2 points, two fewer orders than the
// Euler residual's.
void tri_scalar_residual(const orc_case &c, const double *U, const double *ext, double *rhs) {
  const int n = c.n, P = c.p + 1;
  const size_t E = c.E();
  const TriRule vol = tri_rule(c.p);
  const auto edges = edge_rules(c.p);
  std::fill(rhs, rhs + E * n, 0.0);
  Vec b(n), gx(n), ge(n);
  for (size_t e = 0; e < E; ++e) {
    const TriMap map = c.tris[e].map();
    const double *modes = U + e * n;
    double *r = rhs + e * n;
    for (int q = 0; q < vol.size(); ++q) {
      dubiner(vol.xi[q], vol.eta[q], c.p, b.data());
      dubiner_gradient(vol.xi[q], vol.eta[q], c.p, gx.data(), ge.data());
      const double state = modal_sum(modes, b.data(), n);
      for (int m = 0; m < n; ++m) {
        double dx, dy;
        map.scaled_gradient(gx[m], ge[m], dx, dy);
        r[m] += vol.w[q] * state * (c.vx * dx + c.vy * dy);
      }
    }
  }
  Vec lb(n), rb(n);
  for (size_t f = 0; f < c.faces.size(); ++f) {
    const Face &face = c.faces[f];
    const size_t L = static_cast<size_t>(face.left);
    const EdgeRule &rule = edges[face.left_edge];
    const TriMap lmap = c.tris[L].map();
    const double scale = lmap.edge_length(face.left_edge) / reference_edge_length(face.left_edge);
    for (int q = 0; q < P; ++q) {
      dubiner(rule.xi[q], rule.eta[q], c.p, lb.data());
      const double ls = modal_sum(U + L * n, lb.data(), n);
      double rs = 0.0;
      if (!face.boundary()) {
        double rxi, reta;
        neighbor_edge_point(face, rule.xi[q], rule.eta[q], rxi, reta);
        dubiner(rxi, reta, c.p, rb.data());
        rs = modal_sum(U + static_cast<size_t>(face.right) * n, rb.data(), n);
      } else {
        rs = ext[f * P + q];
      }
      const double flux = upwind_normal(c.vx, c.vy, face.nx, face.ny, ls, rs);
      const double w = rule.w[q] * scale;
      for (int m = 0; m < n; ++m)
        rhs[L * n + m] -= w * lb[m] * flux;
      if (!face.boundary()) {
        const size_t R = static_cast<size_t>(face.right);
        for (int m = 0; m < n; ++m)
          rhs[R * n + m] += w * rb[m] * flux;
      }
    }
  }
  for (size_t e = 0; e < E; ++e) {
    const double inv = 1.0 / c.tris[e].map().jacobian_determinant();
    for (int m = 0; m < n; ++m)
      rhs[e * n + m] *= inv;
  }
}

// src/boundaries/line_euler_exterior.cpp:13-43: exterior per [variable][face]
Vec line_euler_exterior(const orc_case &c, const double *U) {
  const size_t E = c.E(), F = c.lfaces.size();
  Vec bl(c.n), br(c.n);
  legendre(-1.0, c.p, bl.data());
  legendre(1.0, c.p, br.data());
  Vec ext(4 * F, 0.0);
  for (size_t f = 0; f < F; ++f) {
    const LineFace &face = c.lfaces[f];
    if (face.right != -1)
      continue;
    const bool left_end = face.normal < 0.0;
    const int kind = left_end ? c.left_kind : c.right_kind;
    if (kind == ORC_INFLOW)
      throw std::invalid_argument("line Euler exterior: inflow is not implemented");
    const bool periodic = kind == ORC_PERIODIC;
    const size_t donor = periodic ? (left_end ? E - 1 : 0) : static_cast<size_t>(face.left);
    const bool right_endpoint = periodic ? left_end : !left_end;
    const Vec &b = right_endpoint ? br : bl;
    for (int v = 0; v < 4; ++v)
      for (int m = 0; m < c.n; ++m)
        ext[v * F + f] += U[(static_cast<size_t>(v) * c.n + m) * E + donor] * b[m];
  }
  return ext;
}

// include/wasl/dg/euler/line_residual.hpp:23-86
void line_euler_residual(const orc_case &c, const double *U, const double *ext, double *rhs) {
  const int n = c.n;
  const size_t E = c.E(), F = c.lfaces.size();
  const Rule vol = gauss(std::max(1, c.p + 1));
  std::fill(rhs, rhs + 4 * n * E, 0.0);
  Vec b(n), g(n);
  for (size_t e = 0; e < E; ++e) {
    for (int q = 0; q < vol.size(); ++q) {
      legendre(vol.x[q], c.p, b.data());
      legendre_gradient(vol.x[q], c.p, g.data());
      const Cons fx = flux_x(c.gamma, eval_state(U, n, E, e, b.data()));
      for (int v = 0; v < 4; ++v)
        for (int m = 0; m < n; ++m)
          rhs[(static_cast<size_t>(v) * n + m) * E + e] += vol.w[q] * fx[v] * g[m];
    }
  }
  Vec bl(n), br(n);
  legendre(-1.0, c.p, bl.data());
  legendre(1.0, c.p, br.data());
  for (size_t f = 0; f < F; ++f) {
    const LineFace &face = c.lfaces[f];
    const size_t L = static_cast<size_t>(face.left);
    const bool owned_at_left_end = face.normal < 0.0;
    const Vec &ob = owned_at_left_end ? bl : br;
    const Cons owner = eval_state(U, n, E, L, ob.data());
    Cons exterior{0.0, 0.0, 0.0, 0.0};
    if (face.right != -1) {
      exterior = eval_state(U, n, E, static_cast<size_t>(face.right), bl.data());
    } else {
      for (int v = 0; v < 4; ++v)
        exterior[v] = ext[v * F + f];
    }
    const Cons flux = owned_at_left_end ? hllc(c.gamma, exterior, owner, 1.0, 0.0)
                                        : hllc(c.gamma, owner, exterior, 1.0, 0.0);
    for (int v = 0; v < 4; ++v) {
      const double value = flux[v];
      for (int m = 0; m < n; ++m)
        rhs[(static_cast<size_t>(v) * n + m) * E + L] +=
            (owned_at_left_end ? 1.0 : -1.0) * ob[m] * value;
      if (face.right != -1) {
        const size_t R = static_cast<size_t>(face.right);
        for (int m = 0; m < n; ++m)
          rhs[(static_cast<size_t>(v) * n + m) * E + R] += bl[m] * value;
      }
    }
  }
  for (size_t e = 0; e < E; ++e) {
    const double factor = 1.0 / ((c.lines[e].xr - c.lines[e].xl) / 2.0);
    for (int v = 0; v < 4; ++v)
      for (int m = 0; m < n; ++m)
        rhs[(static_cast<size_t>(v) * n + m) * E + e] *= factor;
  }
}

void residual(const orc_case &c, const double *U, double *rhs) {
  if (c.dim == 1 && c.is_euler()) {
    const Vec ext = line_euler_exterior(c, U);
    line_euler_residual(c, U, ext.data(), rhs);
  } else if (c.scalar) {
    const Vec ext = tri_scalar_exterior(c, U);
    tri_scalar_residual(c, U, ext.data(), rhs);
  } else if (c.dim == 2) {
    const Vec ext = tri_exterior(c, U);
    tri_euler_residual(c, U, ext.data(), rhs);
  } else {
    const Vec ext = line_exterior(c, U);
    line_scalar_residual(c, U, ext.data(), rhs);
  }
}

// src/limiting/euler_limiter_stencil_line.cpp:14-94 (the neighbour slots,
// normals (1, 0) and the high-mode start are set by the caller).
void build_line_stencil(orc_case &c) {
  Stencil &s = c.stencil;
  const size_t E = c.E();
  int32_t left_boundary = -1, right_boundary = -1;
  for (const LineFace &face : c.lfaces) {
    if (face.right != -1) {
      append_unique(s.nbrs, s.K, static_cast<size_t>(face.left), face.right);
      append_unique(s.nbrs, s.K, static_cast<size_t>(face.right), face.left);
    } else if (face.normal < 0.0) {
      left_boundary = face.left;
    } else {
      right_boundary = face.left;
    }
  }
  if (c.left_kind == ORC_PERIODIC && left_boundary != -1 && right_boundary != -1) {
    append_unique(s.nbrs, s.K, static_cast<size_t>(left_boundary), right_boundary);
    append_unique(s.nbrs, s.K, static_cast<size_t>(right_boundary), left_boundary);
  }
  const Rule rule = gauss(std::max(1, c.p + 1));
  const double period = c.lines.back().xr - c.lines.front().xl;
  Vec sb(c.n), tb(c.n);
  for (size_t e = 0; e < E; ++e) {
    const double txl = c.lines[e].xl, txr = c.lines[e].xr;
    const double tcentre = (txl + txr) / 2.0, tjac = (txr - txl) / 2.0;
    for (int k = 0; k < s.K; ++k) {
      const int32_t nb = s.nbrs[e * s.K + k];
      if (nb == -1)
        continue;
      double sxl = c.lines[static_cast<size_t>(nb)].xl, sxr = c.lines[static_cast<size_t>(nb)].xr;
      const double shift = periodic_offset(tcentre - (sxl + sxr) / 2.0, period);
      sxl += shift;
      sxr += shift;
      const double scentre = (sxl + sxr) / 2.0, sjac = (sxr - sxl) / 2.0;
      double *M = s.transforms.data() + (e * s.K + k) * c.n * c.n;
      for (int q = 0; q < rule.size(); ++q) {
        const double sxi = (tcentre + tjac * rule.x[q] - scentre) / sjac;
        legendre(sxi, c.p, sb.data());
        legendre(rule.x[q], c.p, tb.data());
        for (int tm = 0; tm < c.n; ++tm)
          for (int sm = 0; sm < c.n; ++sm)
            M[tm * c.n + sm] += rule.w[q] * tb[tm] * sb[sm];
      }
    }
  }
  Vec b(c.n);
  const auto append = [&](double xi) {
    legendre(xi, c.p, b.data());
    s.check.insert(s.check.end(), b.begin(), b.end());
  };
  for (int q = 0; q < rule.size(); ++q)
    append(rule.x[q]);
  append(-1.0);
  append(1.0);
}

// src/limiting/euler_limiter_stencil_triangle.cpp:18-112
void build_stencil(orc_case &c) {
  Stencil &s = c.stencil;
  const size_t E = c.E();
  s.n = c.n;
  s.K = c.dim == 2 ? 3 : 2;
  s.highest_begin = c.p > 0 ? (c.dim == 2 ? tri_modes(c.p - 1) : c.n - 1) : c.n;
  s.nbrs.assign(E * s.K, -1);
  s.nx.assign(E, 1.0);
  s.ny.assign(E, 0.0);
  s.transforms.assign(E * s.K * c.n * c.n, 0.0);
  s.check.clear();
  if (c.dim != 2) {
    if (c.is_euler())
      build_line_stencil(c);
    c.has_stencil = true;
    return;
  }
  const int P = c.p + 1;
  std::vector<unsigned char> has(E, 0);
  const auto set_normal = [&](size_t e, double nx, double ny) {
    if (has[e])
      return;
    s.nx[e] = nx;
    s.ny[e] = ny;
    has[e] = 1;
  };
  for (size_t f = 0; f < c.faces.size(); ++f) {
    const Face &face = c.faces[f];
    int32_t right = face.right;
    if (face.boundary() && c.boundary == ORC_PERIODIC)
      right = c.donors[f * P];
    if (right == -1)
      continue;
    append_unique(s.nbrs, s.K, static_cast<size_t>(face.left), right);
    set_normal(static_cast<size_t>(face.left), face.nx, face.ny);
    if (!face.boundary()) {
      append_unique(s.nbrs, s.K, static_cast<size_t>(right), face.left);
      set_normal(static_cast<size_t>(right), -face.nx, -face.ny);
    }
  }
  // Transforms: column s of each matrix is the L2 projection of the
  // neighbour's mode s (candidate_transform.hpp:18-39).
  const Box box = bounding_box(c.tris);
  const double width = box.max_x - box.min_x, height = box.max_y - box.min_y;
  const TriRule rule = tri_rule(c.p);
  Vec src(c.n, 0.0), proj(c.n), sb(c.n), tb(c.n);
  for (size_t e = 0; e < E; ++e) {
    const TriMap target = c.tris[e].map();
    for (int k = 0; k < s.K; ++k) {
      const int32_t nb = s.nbrs[e * s.K + k];
      if (nb == -1)
        continue;
      TriMap source = c.tris[static_cast<size_t>(nb)].map();
      const double gx =
          (target.x0 + target.x1 + target.x2 - source.x0 - source.x1 - source.x2) / 3.0;
      const double gy =
          (target.y0 + target.y1 + target.y2 - source.y0 - source.y1 - source.y2) / 3.0;
      source.shift(periodic_offset(gx, width), periodic_offset(gy, height));
      double *M = s.transforms.data() + (e * s.K + k) * c.n * c.n;
      for (int sm = 0; sm < c.n; ++sm) {
        src[sm] = 1.0;
        for (int m = 0; m < c.n; ++m)
          proj[m] = 0.0;
        for (int q = 0; q < rule.size(); ++q) {
          double x, y, sxi, seta;
          target.to_physical(rule.xi[q], rule.eta[q], x, y);
          source.to_reference(x, y, sxi, seta);
          dubiner(sxi, seta, c.p, sb.data());
          dubiner(rule.xi[q], rule.eta[q], c.p, tb.data());
          const double value = modal_sum(src.data(), sb.data(), c.n);
          for (int m = 0; m < c.n; ++m)
            proj[m] += rule.w[q] * value * tb[m];
        }
        src[sm] = 0.0;
        for (int m = 0; m < c.n; ++m)
          M[m * c.n + sm] = proj[m];
      }
    }
  }
  Vec b(c.n);
  for (int q = 0; q < rule.size(); ++q) {
    dubiner(rule.xi[q], rule.eta[q], c.p, b.data());
    s.check.insert(s.check.end(), b.begin(), b.end());
  }
  for (const EdgeRule &edge : edge_rules(c.p))
    for (size_t q = 0; q < edge.xi.size(); ++q) {
      dubiner(edge.xi[q], edge.eta[q], c.p, b.data());
      s.check.insert(s.check.end(), b.begin(), b.end());
    }
  c.has_stencil = true;
}

// src/limiting/euler_modal_limiter.cpp:18-98
void limit(const orc_case &c, double thr, double rf, double pf, double *U, uint8_t *flags) {
  const size_t E = c.E();
  if (flags)
    std::fill(flags, flags + E, 0);
  if (c.p == 0)
    return;
  const Stencil &s = c.stencil;
  const int n = c.n;
  const Vec source(U, U + 4 * n * E);
  Vec target(4 * n), cand(s.K * 4 * n), out(4 * n);
  for (size_t e = 0; e < E; ++e) {
    double total = 0.0, high = 0.0;
    for (int m = 0; m < n; ++m) {
      const double v = source[static_cast<size_t>(m) * E + e];
      total += v * v;
      if (m >= s.highest_begin)
        high += v * v;
    }
    if (total == 0.0 || high / total <= thr)
      continue;
    if (flags)
      flags[e] = 1;
    for (int v = 0; v < 4; ++v)
      for (int m = 0; m < n; ++m)
        target[v * n + m] = source[(static_cast<size_t>(v) * n + m) * E + e];
    int K = 0;
    for (int k = 0; k < s.K; ++k) {
      const int32_t nb = s.nbrs[e * s.K + k];
      if (nb == -1)
        continue;
      const double *M = s.transforms.data() + (e * s.K + k) * n * n;
      for (int v = 0; v < 4; ++v) {
        double *res = cand.data() + (K * 4 + v) * n;
        std::fill(res, res + n, 0.0);
        for (int tm = 0; tm < n; ++tm)
          for (int sm = 0; sm < n; ++sm)
            res[tm] += M[tm * n + sm] * source[(static_cast<size_t>(v) * n + sm) * E + nb];
        res[0] = target[v * n];
      }
      ++K;
    }
    const Cons avg{target[0] / std::sqrt(2.0), target[n] / std::sqrt(2.0),
                   target[2 * n] / std::sqrt(2.0), target[3 * n] / std::sqrt(2.0)};
    characteristic_weno(c.gamma, avg, s.nx[e], s.ny[e], target.data(), cand.data(), K, n,
                        out.data());
    positivity(c.gamma, out.data(), n, s.check.data(), s.checks(), rf, pf);
    for (int v = 0; v < 4; ++v)
      for (int m = 0; m < n; ++m)
        U[(static_cast<size_t>(v) * n + m) * E + e] = out[v * n + m];
  }
}

// src/time/cfl_timestep.cpp:14-84
double stable_dt(const orc_case &c, const double *U, double cfl) {
  if (c.dim == 1 && c.is_euler()) {  // :14-19, :33-53, :70-76 with |u_x|
    const size_t E = c.E();
    double h = std::numeric_limits<double>::max();
    for (const auto &l : c.lines)
      h = std::min(h, l.xr - l.xl);
    const double scale = 1.0 / std::sqrt(2.0);
    double speed = 0.0;
    for (size_t e = 0; e < E; ++e) {
      const Cons avg{scale * U[e], scale * U[static_cast<size_t>(c.n) * E + e],
                     scale * U[2 * static_cast<size_t>(c.n) * E + e],
                     scale * U[3 * static_cast<size_t>(c.n) * E + e]};
      if (!is_physical(c.gamma, avg))
        throw std::runtime_error("simulation: Euler cell average is not physical");
      const Prim w = primitive(c.gamma, avg);
      speed = std::max(speed, std::fabs(w.u) + sound_speed(c.gamma, w));
    }
    return speed > 0.0 ? cfl * h / (speed * (2 * c.p + 1)) : 0.0;
  }
  if (c.dim == 1) {
    if (c.lines.empty() || c.wavespeed == 0.0)
      return 0.0;
    double h = std::numeric_limits<double>::max();
    for (const auto &l : c.lines)
      h = std::min(h, l.xr - l.xl);
    return cfl * h / (std::fabs(c.wavespeed) * (2.0 * c.p + 1.0));
  }
  const size_t E = c.E();
  double h = std::numeric_limits<double>::max();
  for (size_t e = 0; e < E; ++e) {
    const TriMap m = c.tris[e].map();
    const double longest = std::max({m.edge_length(0), m.edge_length(1), m.edge_length(2)});
    h = std::min(h, 4.0 * m.jacobian_determinant() / longest);
  }
  if (c.scalar) {  // cfl_timestep.cpp:62-68: constant for the whole run
    const double speed = std::hypot(c.vx, c.vy);
    if (E == 0 || speed == 0.0)
      return 0.0;
    return cfl * h / (speed * (2.0 * c.p + 1.0));
  }
  const double scale = 1.0 / std::sqrt(2.0);
  double speed = 0.0;
  for (size_t e = 0; e < E; ++e) {
    const Cons avg{scale * U[e], scale * U[static_cast<size_t>(c.n) * E + e],
                   scale * U[2 * static_cast<size_t>(c.n) * E + e],
                   scale * U[3 * static_cast<size_t>(c.n) * E + e]};
    if (!is_physical(c.gamma, avg))
      throw std::runtime_error("simulation: Euler cell average is not physical");
    const Prim w = primitive(c.gamma, avg);
    speed = std::max(speed, std::hypot(w.u, w.v) + sound_speed(c.gamma, w));
  }
  return speed > 0.0 ? cfl * h / (speed * (2 * c.p + 1)) : 0.0;
}

// include/wasl/time/ssp_rk3.hpp:37-61
void step(orc_case &c, double dt, int limiting, double thr, double rf, double pf, double *U) {
  const size_t N = c.state_size();
  c.stage.resize(N);
  c.rhs.resize(N);
  c.euler.resize(N);
  const auto post = [&](double *s) {
    if (limiting && c.is_euler())
      limit(c, thr, rf, pf, s, nullptr);
  };
  residual(c, U, c.rhs.data());
  for (size_t i = 0; i < N; ++i)
    c.euler[i] = U[i] + dt * c.rhs[i];
  post(c.euler.data());
  residual(c, c.euler.data(), c.rhs.data());
  for (size_t i = 0; i < N; ++i)
    c.stage[i] = 0.75 * U[i] + 0.25 * (c.euler[i] + dt * c.rhs[i]);
  post(c.stage.data());
  residual(c, c.stage.data(), c.rhs.data());
  for (size_t i = 0; i < N; ++i)
    U[i] = (1.0 / 3.0) * U[i] + (2.0 / 3.0) * (c.stage[i] + dt * c.rhs[i]);
  post(U);
}

// src/case/euler_diagnostics.cpp:30-36, scalar_diagnostics.cpp:18-25
double mass(const orc_case &c, const double *U) {
  double r = 0.0;
  if (c.scalar) {  // scalar_diagnostics.cpp:27-34
    for (size_t e = 0; e < c.E(); ++e)
      r += std::sqrt(2.0) * c.tris[e].map().jacobian_determinant() * U[e * c.n];
  } else if (c.dim == 2) {
    for (size_t e = 0; e < c.E(); ++e)
      r += std::sqrt(2.0) * c.tris[e].map().jacobian_determinant() * U[e];
  } else if (c.is_euler()) {  // euler_diagnostics.cpp:22-27
    for (size_t e = 0; e < c.E(); ++e)
      r += std::sqrt(2.0) * ((c.lines[e].xr - c.lines[e].xl) / 2.0) * U[e];
  } else {
    for (size_t e = 0; e < c.E(); ++e)
      r += std::sqrt(2.0) * ((c.lines[e].xr - c.lines[e].xl) / 2.0) * U[e * c.n];
  }
  return r;
}

// Initial conditions. Regions follow src/case/initial_regions.cpp:127-174;
// the analytic fields are the benchmark's own definitions.
Prim euler_ic(int ic, const double *prm, int64_t np, double gamma, double x, double y) {
  switch (ic) {
  case ORC_IC_EULER_REGIONS: {
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
  case ORC_IC_VORTEX: {
    const double eps = prm[0];
    const double dx = x - prm[1], dy = y - prm[2];
    const double r2 = dx * dx + dy * dy;
    const double swirl = eps / (2.0 * kPi) * std::exp(0.5 * (1.0 - r2));
    const double T = 1.0 - (gamma - 1.0) * eps * eps / (8.0 * gamma * kPi * kPi) * std::exp(1.0 - r2);
    const double rho = std::pow(T, 1.0 / (gamma - 1.0));
    return {rho, 1.0 - swirl * dy, 1.0 + swirl * dx, std::pow(rho, gamma)};
  }
  case ORC_IC_RAREFACTION: {
    const double r = std::sqrt(x * x + y * y);
    if (r == 0.0)
      return {prm[0], 0.0, 0.0, prm[1]};
    return {prm[0], prm[2] * x / r, prm[2] * y / r, prm[1]};
  }
  case ORC_IC_SMOOTH: {
    const double L = prm[0];
    double rho = 0.0;
    for (int k = 0; k < 4; ++k) {
      const double *t = prm + 1 + 4 * k;
      rho += t[0] * std::sin(2.0 * kPi * (t[1] * x + t[2] * y) / L + t[3]);
    }
    return {1.0 + 0.1 * rho, 0.5, 0.5, 1.0};
  }
  default:
    throw std::runtime_error("oracle: unknown Euler initial condition");
  }
}

// src/case/initial_regions.cpp:28-34 (circles) and a plane sine wave
double scalar_ic_2d(int ic, const double *prm, int64_t np, double x, double y) {
  if (ic == ORC_IC_SCALAR_CIRCLES) {
    for (int64_t k = 1; k + 4 <= np; k += 4) {
      const double dx = x - prm[k], dy = y - prm[k + 1];
      if (dx * dx + dy * dy <= prm[k + 2] * prm[k + 2])
        return prm[k + 3];
    }
    return prm[0];
  }
  if (ic == ORC_IC_SCALAR_WAVE)
    return std::sin(2.0 * kPi * (prm[1] * x + prm[2] * y) / prm[0]);
  throw std::runtime_error("oracle: unknown 2D scalar initial condition");
}

double scalar_ic(int ic, const double *prm, int64_t np, double x) {
  if (ic == ORC_IC_SCALAR_PULSE) {
    const double s = (x - 0.5) / 0.05;
    return std::sin(2.0 * kPi * x) + std::exp(-(s * s));
  }
  if (ic == ORC_IC_SCALAR_REGIONS) {
    if (prm[1] != 0.0)
      return prm[2] * std::sin(2.0 * kPi * x);
    for (int64_t k = 3; k + 3 <= np; k += 3)
      if (x >= prm[k] && x <= prm[k + 1])
        return prm[k + 2];
    return prm[0];
  }
  throw std::runtime_error("oracle: unknown scalar initial condition");
}

// src/case/initial_state.cpp:16-64 with dg/l2_projection.hpp:36-76
void project(const orc_case &c, int ic, const double *prm, int64_t np, double *U) {
  const size_t E = c.E();
  if (c.scalar) {
    // project_triangle_field writes U[0][mode][element]; solver_coefficients
    // transposes it to U[element][mode] (scalar_coefficient_layout.hpp).
    const TriRule rule = tri_rule(std::max(c.p, 5));
    Vec tmp(c.n * E, 0.0), b(c.n);
    for (size_t e = 0; e < E; ++e) {
      const TriMap map = c.tris[e].map();
      for (int q = 0; q < rule.size(); ++q) {
        double x, y;
        map.to_physical(rule.xi[q], rule.eta[q], x, y);
        dubiner(rule.xi[q], rule.eta[q], c.p, b.data());
        const double value = scalar_ic_2d(ic, prm, np, x, y);
        for (int m = 0; m < c.n; ++m)
          tmp[static_cast<size_t>(m) * E + e] += rule.w[q] * value * b[m];
      }
    }
    for (size_t e = 0; e < E; ++e)
      for (int m = 0; m < c.n; ++m)
        U[e * c.n + m] = tmp[static_cast<size_t>(m) * E + e];
    return;
  }
  if (c.dim == 2) {
    const TriRule rule = tri_rule(std::max(c.p, 5));
    std::fill(U, U + 4 * c.n * E, 0.0);
    Vec b(c.n);
    for (size_t e = 0; e < E; ++e) {
      const TriMap map = c.tris[e].map();
      for (int q = 0; q < rule.size(); ++q) {
        double x, y;
        map.to_physical(rule.xi[q], rule.eta[q], x, y);
        dubiner(rule.xi[q], rule.eta[q], c.p, b.data());
        const Cons s = conserved(c.gamma, euler_ic(ic, prm, np, c.gamma, x, y));
        for (int v = 0; v < 4; ++v)
          for (int m = 0; m < c.n; ++m)
            U[(static_cast<size_t>(v) * c.n + m) * E + e] += rule.w[q] * s[v] * b[m];
      }
    }
    return;
  }
  if (c.is_euler()) {  // initial_state.cpp:48-54: U[variable][mode][element]
    const Rule rule = gauss(7);
    std::fill(U, U + 4 * c.n * E, 0.0);
    Vec b(c.n);
    for (size_t e = 0; e < E; ++e) {
      const double xl = c.lines[e].xl, xr = c.lines[e].xr;
      for (int q = 0; q < rule.size(); ++q) {
        legendre(rule.x[q], c.p, b.data());
        const double centre = (xl + xr) / 2.0, jac = (xr - xl) / 2.0;
        const Cons s = conserved(c.gamma, euler_ic(ic, prm, np, c.gamma, centre + jac * rule.x[q],
                                                   0.0));
        for (int v = 0; v < 4; ++v)
          for (int m = 0; m < c.n; ++m)
            U[(static_cast<size_t>(v) * c.n + m) * E + e] += rule.w[q] * s[v] * b[m];
      }
    }
    return;
  }
  // project_line_field writes U[variable][mode][element]; with one variable
  // the transpose to the solver layout is U[element][mode].
  const Rule rule = gauss(7);
  Vec tmp(c.n * E, 0.0), b(c.n);
  for (size_t e = 0; e < E; ++e) {
    const double xl = c.lines[e].xl, xr = c.lines[e].xr;
    for (int q = 0; q < rule.size(); ++q) {
      legendre(rule.x[q], c.p, b.data());
      const double centre = (xl + xr) / 2.0, jac = (xr - xl) / 2.0;
      const double value = scalar_ic(ic, prm, np, centre + jac * rule.x[q]);
      for (int m = 0; m < c.n; ++m)
        tmp[static_cast<size_t>(m) * E + e] += rule.w[q] * value * b[m];
    }
  }
  for (size_t e = 0; e < E; ++e)
    for (int m = 0; m < c.n; ++m)
      U[e * c.n + m] = tmp[static_cast<size_t>(m) * E + e];
}

// src/mesh/line/build_line_mesh.cpp:15-85: elements oriented left to right
// and sorted by their left end, faces from left to right.
void build_lines(orc_case &c, int64_t node_count, const int64_t *node_tags, const double *node_x,
                 int64_t segment_count, const int64_t *segment_nodes) {
  std::map<int64_t, int64_t> lookup;
  for (int64_t i = 0; i < node_count; ++i)
    lookup[node_tags[i]] = i;
  for (int64_t s = 0; s < segment_count; ++s) {
    const int64_t ta = segment_nodes[2 * s], tb = segment_nodes[2 * s + 1];
    const auto ia = lookup.find(ta), ib = lookup.find(tb);
    if (ia == lookup.end() || ib == lookup.end())
      throw std::runtime_error("line mesh: missing node");
    const double xa = node_x[ia->second], xb = node_x[ib->second];
    if (xa == xb)
      throw std::runtime_error("line mesh: zero-length element " + std::to_string(s + 1));
    const bool fwd = xa < xb;
    c.lines.push_back(fwd ? LineElem{ta, tb, xa, xb} : LineElem{tb, ta, xb, xa});
  }
  if (c.lines.empty())
    throw std::runtime_error("line mesh: document contains no line elements");
  std::stable_sort(c.lines.begin(), c.lines.end(),
                   [](const LineElem &a, const LineElem &b) { return a.xl < b.xl; });
  c.lfaces.push_back({0, -1, c.lines.front().xl, -1.0});
  for (size_t i = 0; i + 1 < c.lines.size(); ++i) {
    if (c.lines[i].rn != c.lines[i + 1].ln)
      throw std::runtime_error("line mesh: gap or discontinuity between elements " +
                               std::to_string(i) + " and " + std::to_string(i + 1));
    c.lfaces.push_back(
        {static_cast<int32_t>(i), static_cast<int32_t>(i + 1), c.lines[i].xr, 1.0});
  }
  c.lfaces.push_back({static_cast<int32_t>(c.lines.size() - 1), -1, c.lines.back().xr, 1.0});
}

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

extern "C" {

const char *orc_error(void) { return g_error.c_str(); }

orc_case *orc_tri_euler_create(int64_t node_count, const int64_t *node_tags, const double *node_xy,
                               int64_t triangle_count, const int64_t *triangle_nodes, int order,
                               double gamma, int boundary) {
  orc_case *c = new orc_case;
  if (guarded([&] {
        if (order < 0 || order > 3)
          throw std::runtime_error("oracle: polynomial order must be 0..3");
        c->dim = 2;
        c->p = order;
        c->n = tri_modes(order);
        c->V = 4;
        c->gamma = gamma;
        c->boundary = boundary;
        build_triangles(node_count, node_tags, node_xy, triangle_count, triangle_nodes, c->tris,
                        c->faces);
        build_trace(*c);
        build_stencil(*c);
      }) != 0) {
    delete c;
    return nullptr;
  }
  return c;
}

// src/case/scalar_advection_triangle.cpp:15-46: periodic boundaries only,
// no limiter.
orc_case *orc_tri_scalar_create(int64_t node_count, const int64_t *node_tags, const double *node_xy,
                                int64_t triangle_count, const int64_t *triangle_nodes, int order,
                                double velocity_x, double velocity_y) {
  orc_case *c = new orc_case;
  if (guarded([&] {
        if (order < 0 || order > 3)
          throw std::runtime_error("oracle: polynomial order must be 0..3");
        c->dim = 2;
        c->scalar = true;
        c->p = order;
        c->n = tri_modes(order);
        c->V = 1;
        c->vx = velocity_x;
        c->vy = velocity_y;
        c->boundary = ORC_PERIODIC;
        build_triangles(node_count, node_tags, node_xy, triangle_count, triangle_nodes, c->tris,
                        c->faces);
        build_trace(*c);
      }) != 0) {
    delete c;
    return nullptr;
  }
  return c;
}

// src/mesh/line/build_line_mesh.cpp:15-85
orc_case *orc_line_scalar_create(int64_t node_count, const int64_t *node_tags,
                                 const double *node_x, int64_t segment_count,
                                 const int64_t *segment_nodes, int order, double wavespeed,
                                 int left_kind, double left_value, int right_kind,
                                 double right_value) {
  orc_case *c = new orc_case;
  if (guarded([&] {
        c->dim = 1;
        c->p = order;
        c->n = order + 1;
        c->V = 1;
        c->wavespeed = wavespeed;
        c->left_kind = left_kind;
        c->right_kind = right_kind;
        c->left_value = left_value;
        c->right_value = right_value;
        build_lines(*c, node_count, node_tags, node_x, segment_count, segment_nodes);
        build_stencil(*c);
      }) != 0) {
    delete c;
    return nullptr;
  }
  return c;
}

// src/case/euler_line.cpp:14-48: line limiter stencil with periodic or open
// ends (line_boundary_resolution.cpp pairs periodic ends).
orc_case *orc_line_euler_create(int64_t node_count, const int64_t *node_tags,
                                const double *node_x, int64_t segment_count,
                                const int64_t *segment_nodes, int order, double gamma,
                                int left_kind, int right_kind) {
  orc_case *c = new orc_case;
  if (guarded([&] {
        if (order < 0 || order > 3)
          throw std::runtime_error("oracle: polynomial order must be 0..3");
        if ((left_kind == ORC_PERIODIC) != (right_kind == ORC_PERIODIC))
          throw std::runtime_error("boundaries: periodic ends must be paired");
        c->dim = 1;
        c->p = order;
        c->n = order + 1;
        c->V = 4;
        c->gamma = gamma;
        c->left_kind = left_kind;
        c->right_kind = right_kind;
        build_lines(*c, node_count, node_tags, node_x, segment_count, segment_nodes);
        build_stencil(*c);
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
    return c->n;
  case 3:
    return c->V;
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
    for (size_t e = 0; e < c->tris.size(); ++e)
      for (int k = 0; k < 3; ++k) {
        element_nodes[3 * e + k] = c->tris[e].node[k];
        element_xy[6 * e + 2 * k] = c->tris[e].x[k];
        element_xy[6 * e + 2 * k + 1] = c->tris[e].y[k];
      }
    for (size_t f = 0; f < c->faces.size(); ++f) {
      const Face &x = c->faces[f];
      face_elements[2 * f] = x.left;
      face_elements[2 * f + 1] = x.right;
      face_edges[2 * f] = x.left_edge;
      face_edges[2 * f + 1] = x.right_edge;
      face_vertices[4 * f] = x.lv0;
      face_vertices[4 * f + 1] = x.lv1;
      face_vertices[4 * f + 2] = x.rv0;
      face_vertices[4 * f + 3] = x.rv1;
      face_normals[2 * f] = x.nx;
      face_normals[2 * f + 1] = x.ny;
    }
  });
}

int orc_line_topology(const orc_case *c, double *element_x, int32_t *face_elements,
                      double *face_normals) {
  return guarded([&] {
    for (size_t e = 0; e < c->lines.size(); ++e) {
      element_x[2 * e] = c->lines[e].xl;
      element_x[2 * e + 1] = c->lines[e].xr;
    }
    for (size_t f = 0; f < c->lfaces.size(); ++f) {
      face_elements[2 * f] = c->lfaces[f].left;
      face_elements[2 * f + 1] = c->lfaces[f].right;
      face_normals[f] = c->lfaces[f].normal;
    }
  });
}

int orc_tri_trace(const orc_case *c, int32_t *donors, double *basis) {
  return guarded([&] {
    std::copy(c->donors.begin(), c->donors.end(), donors);
    std::copy(c->donor_basis.begin(), c->donor_basis.end(), basis);
  });
}

int64_t orc_stencil_check_count(const orc_case *c) { return c->stencil.checks(); }

int orc_stencil(const orc_case *c, int32_t *neighbors, double *transforms, double *normals,
                double *check_basis) {
  return guarded([&] {
    const Stencil &s = c->stencil;
    std::copy(s.nbrs.begin(), s.nbrs.end(), neighbors);
    std::copy(s.transforms.begin(), s.transforms.end(), transforms);
    for (size_t e = 0; e < s.nx.size(); ++e) {
      normals[2 * e] = s.nx[e];
      normals[2 * e + 1] = s.ny[e];
    }
    std::copy(s.check.begin(), s.check.end(), check_basis);
  });
}

int orc_residual(const orc_case *c, const double *state, double *rhs) {
  return guarded([&] { residual(*c, state, rhs); });
}

int orc_limit(const orc_case *c, double threshold, double density_floor, double pressure_floor,
              double *state, uint8_t *flags) {
  return guarded([&] {
    if (!c->is_euler())
      throw std::runtime_error("oracle: the limiter applies to Euler cases");
    limit(*c, threshold, density_floor, pressure_floor, state, flags);
  });
}

int orc_stable_dt(const orc_case *c, const double *state, double cfl, double *dt) {
  return guarded([&] { *dt = stable_dt(*c, state, cfl); });
}

int orc_step(orc_case *c, double dt, int limiting, double threshold, double density_floor,
             double pressure_floor, double *state) {
  return guarded(
      [&] { step(*c, dt, limiting, threshold, density_floor, pressure_floor, state); });
}

int orc_mass(const orc_case *c, const double *state, double *m) {
  return guarded([&] { *m = mass(*c, state); });
}

int orc_project(const orc_case *c, int ic, const double *params, int64_t param_count,
                double *state) {
  return guarded([&] { project(*c, ic, params, param_count, state); });
}

// src/case/serial_case_runner.cpp:107-149
int orc_run(orc_case *c, double cfl, int limiting, double threshold, double density_floor,
            double pressure_floor, double start_time, uint64_t start_step, double end_time,
            uint64_t max_steps, uint64_t history_interval, double *state, uint64_t *final_step,
            double *final_time, double *history, int64_t history_capacity,
            int64_t *history_rows) {
  return guarded([&] {
    double time = start_time;
    uint64_t stepno = start_step;
    int64_t rows = 0;
    const auto record = [&](double dt) {
      if (history && rows < history_capacity) {
        history[4 * rows] = static_cast<double>(stepno);
        history[4 * rows + 1] = time;
        history[4 * rows + 2] = dt;
        history[4 * rows + 3] = mass(*c, state);
      }
      ++rows;
    };
    record(0.0);
    while (time < end_time && stepno < max_steps) {
      const double stable = stable_dt(*c, state, cfl);
      const double remaining = end_time - time;
      const double dt = std::min(stable > 0.0 ? stable : remaining, remaining);
      if (dt <= 0.0 || time + dt == time)
        throw std::runtime_error("simulation: timestep cannot advance time");
      step(*c, dt, limiting, threshold, density_floor, pressure_floor, state);
      time += dt;
      ++stepno;
      const bool final = time >= end_time || stepno == max_steps;
      if (stepno % history_interval == 0 || final)
        record(dt);
    }
    if (time < end_time)
      throw std::runtime_error("simulation: max_steps reached before end_time");
    *final_step = stepno;
    *final_time = time;
    if (history_rows)
      *history_rows = rows;
  });
}

// ---- function-level hooks ----

int orc_basis(int kind, int p, double xi, double eta, int gradient, double *values, double *gx,
              double *ge) {
  return guarded([&] {
    if (kind == 0) {
      legendre(xi, p, values);
      if (gradient)
        legendre_gradient(xi, p, gx);
    } else {
      dubiner(xi, eta, p, values);
      if (gradient)
        dubiner_gradient(xi, eta, p, gx, ge);
    }
  });
}

int orc_rule(int kind, int n, int edge, double *a, double *b, double *w) {
  int count = -1;
  if (guarded([&] {
        if (kind == 0) {
          const Rule r = gauss(n);
          count = r.size();
          std::copy(r.x.begin(), r.x.end(), a);
          std::copy(r.w.begin(), r.w.end(), w);
        } else if (kind == 1) {
          const TriRule r = tri_rule(n);
          count = r.size();
          std::copy(r.xi.begin(), r.xi.end(), a);
          std::copy(r.eta.begin(), r.eta.end(), b);
          std::copy(r.w.begin(), r.w.end(), w);
        } else {
          const EdgeRule r = edge_rules(n).at(static_cast<size_t>(edge));
          count = static_cast<int>(r.xi.size());
          std::copy(r.xi.begin(), r.xi.end(), a);
          std::copy(r.eta.begin(), r.eta.end(), b);
          std::copy(r.w.begin(), r.w.end(), w);
        }
      }) != 0)
    return -1;
  return count;
}

int orc_euler_flux(double gamma, const double *s, double *fx, double *fy) {
  return guarded([&] {
    const Cons u{s[0], s[1], s[2], s[3]};
    const Cons a = flux_x(gamma, u), b = flux_y(gamma, u);
    for (int v = 0; v < 4; ++v) {
      fx[v] = a[v];
      fy[v] = b[v];
    }
  });
}

int orc_hllc(double gamma, const double *l, const double *r, double nx, double ny,
             double *flux) {
  return guarded([&] {
    const Cons f = hllc(gamma, {l[0], l[1], l[2], l[3]}, {r[0], r[1], r[2], r[3]}, nx, ny);
    for (int v = 0; v < 4; ++v)
      flux[v] = f[v];
  });
}

int orc_conserved(double gamma, const double *w, double *s) {
  return guarded([&] {
    const Cons u = conserved(gamma, {w[0], w[1], w[2], w[3]});
    for (int v = 0; v < 4; ++v)
      s[v] = u[v];
  });
}

int orc_characteristic(double gamma, const double *avg, const double *d, double nx, double ny,
                       double *waves, double *back) {
  return guarded([&] {
    const Cons a{avg[0], avg[1], avg[2], avg[3]};
    to_characteristic(gamma, a, {d[0], d[1], d[2], d[3]}, nx, ny, waves);
    const Cons b = from_characteristic(gamma, a, waves, nx, ny);
    for (int v = 0; v < 4; ++v)
      back[v] = b[v];
  });
}

int orc_simple_weno(const double *target, const double *candidates, int K, int n,
                    double *limited) {
  return guarded([&] { simple_weno(target, candidates, K, n, limited); });
}

int orc_characteristic_weno(double gamma, const double *avg, double nx, double ny,
                            const double *target, const double *candidates, int K, int n,
                            double *limited) {
  return guarded([&] {
    characteristic_weno(gamma, {avg[0], avg[1], avg[2], avg[3]}, nx, ny, target, candidates, K,
                        n, limited);
  });
}

int orc_positivity(double gamma, double *c, int n, const double *basis, int points, double rf,
                   double pf, double *scale) {
  return guarded([&] { *scale = positivity(gamma, c, n, basis, points, rf, pf); });
}

// include/wasl/limiting/troubled_cell/modal_decay.hpp:36-47
int orc_modal_decay_troubled(const double *c, int p, double threshold, int *troubled) {
  return guarded([&] {
    double ratio = 0.0;
    if (p >= 1) {
      const int n = tri_modes(p), hb = tri_modes(p - 1);
      double total = 0.0, high = 0.0;
      for (int m = 0; m < n; ++m) {
        total += c[m] * c[m];
        if (m >= hb)
          high += c[m] * c[m];
      }
      ratio = total > 0.0 ? high / total : 0.0;
    }
    *troubled = ratio > threshold ? 1 : 0;
  });
}

}  // extern "C"
