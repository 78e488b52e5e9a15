// Host-side setup for the wasl_hip engine: reference-element tables, the
// canonical mesh topology, periodic donor traces, the limiter stencil and
// initial projection. Everything here runs once per case on the CPU, in
// FP64 with the same libm calls as the reference, so the tables the kernels
// read are bit-identical to the reference's own evaluations.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace whip {

constexpr double kPi = 3.141592653589793238462643383279502884;

// ------------------------------------------------------------ reference ----
double jacobi(double x, int n, double alpha, double beta);
double jacobi_derivative(double x, int n, double alpha, double beta);
void legendre(double xi, int p, double *values);
void legendre_gradient(double xi, int p, double *gradients);
void dubiner(double xi, double eta, int p, double *values);
void dubiner_gradient(double xi, double eta, int p, double *gxi, double *geta);
inline int tri_modes(int p) { return (p + 1) * (p + 2) / 2; }

struct LineRule {
  std::vector<double> x, w;
  int size() const { return static_cast<int>(x.size()); }
};
struct TriRule {
  std::vector<double> xi, eta, w;
  int size() const { return static_cast<int>(xi.size()); }
};
struct EdgeRule {
  std::vector<double> xi, eta, w;
};

LineRule gauss_legendre(int n);
TriRule triangle_rule(int p);                // collapsed tensor, (p+1)^2 points
std::array<EdgeRule, 3> edge_rules(int p);   // p+1 points per edge
double reference_edge_length(int edge);

// Right-element reference point of left-edge point (lxi, leta) across a face
// (left edge kl, right edge kr, flip = vertex traversal differs).
void neighbor_point(int kl, int kr, bool flip, double lxi, double leta, double &rxi,
                    double &reta);

// ------------------------------------------------------------------ mesh ----
struct Triangle {
  int64_t node[3];
  double x[3], y[3];
  double ax() const { return x[1] - x[0]; }
  double ay() const { return y[1] - y[0]; }
  double bx() const { return x[2] - x[0]; }
  double by() const { return y[2] - y[0]; }
  double det_a() const { return ax() * by() - bx() * ay(); }
  double jacobian_determinant() const { return det_a() / 4.0; }
  void to_physical(double xi, double eta, double &px, double &py) const;
  void to_reference(double px, double py, double &xi, double &eta) const;
  void edge_vector(int k, double &ex, double &ey) const;
  double edge_length(int k) const;
  void edge_normal(int k, double &nx, double &ny) const;
};

struct TriFace {
  int32_t left = -1, right = -1;
  int left_edge = 0, right_edge = 0;
  int64_t va = 0, vb = 0;              // sorted vertex pair
  int64_t lv0 = 0, lv1 = 0, rv0 = 0, rv1 = 0;
  double nx = 0.0, ny = 0.0;
  bool boundary() const { return right < 0; }
  bool flip() const { return lv0 != rv0; }
};

struct Segment {
  int64_t left_node, right_node;
  double xl, xr;
};
struct LineFace {
  int32_t left, right;
  double x, normal;
};

struct Mesh {
  int dimension = 2;
  std::vector<Triangle> tris;
  std::vector<TriFace> faces;
  std::vector<Segment> lines;
  std::vector<LineFace> lfaces;

  size_t elements() const { return dimension == 2 ? tris.size() : lines.size(); }
  size_t face_count() const { return dimension == 2 ? faces.size() : lfaces.size(); }
  size_t boundary_faces() const;
};

// Canonicalisation from MSH-style records (throws std::runtime_error with the
// reference's messages).
Mesh build_mesh(int dimension, int64_t node_count, const int64_t *tags, const double *xyz,
                int64_t element_count, const int64_t *element_nodes);
Mesh mesh_from_prepared(int dimension, int64_t node_count, const int64_t *tags,
                        const double *points, int64_t element_count, const int64_t *element_nodes,
                        int64_t face_count, const int64_t *face_nodes,
                        const int32_t *face_elements, const int32_t *face_local_edges,
                        const double *face_normals);

// ------------------------------------------------------------ boundaries ----
struct PeriodicTrace {
  int points = 0;                 // p + 1
  std::vector<int32_t> donors;    // [F][points], -1 on interior faces
  std::vector<double> basis;      // [F][points][Np]
};
// Axis-aligned periodic box: every boundary edge point is translated by one
// period and located in the lowest-index element containing it (tolerance
// 1e-10 in reference coordinates), found through a uniform bin grid.
PeriodicTrace periodic_trace(const Mesh &mesh, int p);

// ---------------------------------------------------------------- stencil ----
struct LimiterStencil {
  int K = 3;
  int Np = 0;
  int highest_begin = 0;
  std::vector<int32_t> neighbors;  // [E][K]
  std::vector<double> transforms;  // [E][K][Np][Np]
  std::vector<double> nx, ny;      // [E]
  std::vector<double> check;       // [C][Np]
  int checks() const { return Np ? static_cast<int>(check.size()) / Np : 0; }
};
// donors: periodic donors or empty (no neighbour across a boundary face).
LimiterStencil limiter_stencil(const Mesh &mesh, int p, const std::vector<int32_t> &donors,
                               int points);

// Line limiter stencil (euler_limiter_stencil_line.cpp:14-94): K = 2
// neighbours from the faces in order (the two end elements joined when the
// ends are periodic), normals (1, 0), transforms by the Gauss rule of
// max(1, p+1) points, check points at those Gauss points and the two ends.
LimiterStencil line_limiter_stencil(const Mesh &mesh, int p, bool periodic);

// Smallest altitude 4 detJ / longest edge (triangles) or smallest length.
double minimum_length(const Mesh &mesh);

// ------------------------------------------------------------- projection ----
// Initial-condition codes shared with the C ABI.
enum { IC_EULER_REGIONS = 0, IC_VORTEX = 1, IC_RAREFACTION = 2, IC_SMOOTH = 3,
       IC_SCALAR_PULSE = 10, IC_SCALAR_REGIONS = 11, IC_SCALAR_CIRCLES = 12,
       IC_SCALAR_WAVE = 13 };
// Euler (codes 0-3): U[4][Np][E] (lines: x with y = 0); scalar (10-13): U[E][Np].
std::vector<double> project(const Mesh &mesh, int p, double gamma, int ic, const double *params,
                            int64_t param_count);

// Run fn(begin, end) over [0, n) on up to hardware_concurrency threads.
template <typename F> void parallel_for(size_t n, F fn);

}  // namespace whip

#include "setup_parallel.hpp"
