// wasl_hip engine: device memory, kernel dispatch and the C ABI of
// include/wasl_hip.h. Host setup lives in setup.cpp, device code in
// kernels.hpp.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <tuple>
#include <unordered_map>
#include <string>
#include <vector>

#include "../../include/wasl_hip.h"
#include "kernels.hpp"
#include "setup.hpp"

// A built mesh is immutable: engines and queued serving requests share it
// instead of copying it.
struct wasl_hip_mesh {
  explicit wasl_hip_mesh(whip::Mesh built)
      : shared(std::make_shared<const whip::Mesh>(std::move(built))), m(*shared) {}
  std::shared_ptr<const whip::Mesh> shared;
  const whip::Mesh &m;
};

namespace {

thread_local std::string g_error;

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct SimulationError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw DeviceError(std::string("hip: ") + what + ": " + hipGetErrorString(e));
}
#define HIPCHECK(x) check((x), #x)

template <typename F> int guarded(F &&f) {
  try {
    f();
    return WASL_HIP_OK;
  } catch (const SimulationError &ex) {
    g_error = ex.what();
    return WASL_HIP_SIMULATION;
  } catch (const DeviceError &ex) {
    g_error = ex.what();
    return WASL_HIP_DEVICE;
  } catch (const std::exception &ex) {
    g_error = ex.what();
    return WASL_HIP_ERROR;
  }
}

const char *device_error_text(int code) {
  switch (code) {
  case whip::kErrUnphysical:
    return "simulation: Euler cell average is not physical";
  case whip::kErrPositivity:
    return "Euler positivity limiter: cell average is below a floor";
  case whip::kErrTimestep:
    return "simulation: timestep cannot advance time";
  default:
    return "simulation: device error";
  }
}

// Device allocation owned by the engine. With a stream it is stream-ordered
// (hipMallocAsync / hipFreeAsync from the device's memory pool): releasing it
// neither synchronises the device nor returns memory to the driver, which
// makes creating and destroying many engines (one per served case) cheap.
struct Buffer {
  void *p = nullptr;
  size_t bytes = 0;
  hipStream_t stream = nullptr;
  Buffer() = default;
  Buffer(const Buffer &) = delete;
  Buffer &operator=(const Buffer &) = delete;
  ~Buffer() { release(); }
  void release() {
    if (p) {
      if (stream)
        (void)hipFreeAsync(p, stream);
      else
        (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t n, size_t &total, hipStream_t s = nullptr) {
    release();
    bytes = n;
    stream = s;
    if (n) {
      if (s)
        HIPCHECK(hipMallocAsync(&p, n, s));
      else
        HIPCHECK(hipMalloc(&p, n));
      total += n;
    }
  }
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

template <typename T>
void upload(Buffer &b, const std::vector<T> &v, size_t &total, hipStream_t s) {
  b.alloc(v.size() * sizeof(T), total, s);
  if (!v.empty())
    HIPCHECK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

// Memory freed by stream-ordered releases stays in the device's pool for the
// next engine instead of going back to the driver at each synchronisation.
void keep_pool_memory(int device) {
  static std::mutex mu;
  static std::vector<int> done;
  std::lock_guard<std::mutex> lock(mu);
  if (std::find(done.begin(), done.end(), device) != done.end())
    return;
  hipMemPool_t pool;
  HIPCHECK(hipDeviceGetDefaultMemPool(&pool, device));
  uint64_t keep = ~0ull;
  HIPCHECK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
  done.push_back(device);
}

// Small pinned host slots for step-control read-backs. Pinning and unpinning
// per engine costs about a millisecond each way, so slots come from 64 KiB
// pinned chunks kept for the life of the process.
constexpr size_t kPinnedSlot = 256;

class PinnedSlots {
public:
  void *get() {
    std::lock_guard<std::mutex> lock(mu_);
    if (free_.empty()) {
      char *chunk = nullptr;
      HIPCHECK(hipHostMalloc(reinterpret_cast<void **>(&chunk), kChunk));
      for (size_t i = 0; i < kChunk / kPinnedSlot; ++i)
        free_.push_back(chunk + i * kPinnedSlot);
    }
    void *slot = free_.back();
    free_.pop_back();
    return slot;
  }
  void put(void *slot) {
    std::lock_guard<std::mutex> lock(mu_);
    free_.push_back(slot);
  }

private:
  static constexpr size_t kChunk = size_t(64) << 10;
  std::mutex mu_;
  std::vector<void *> free_;
};

PinnedSlots &pinned_slots() {
  static PinnedSlots *slots = new PinnedSlots;   // outlives every engine
  return *slots;
}

// Engine streams, reused: creating and destroying a HIP stream per engine
// costs about a millisecond. A returned stream may still hold its engine's
// stream-ordered frees; the next engine's work simply queues behind them.
class StreamPool {
public:
  hipStream_t get(int device) {
    {
      std::lock_guard<std::mutex> lock(mu_);
      for (size_t i = 0; i < free_.size(); ++i)
        if (free_[i].first == device) {
          hipStream_t s = free_[i].second;
          free_.erase(free_.begin() + static_cast<std::ptrdiff_t>(i));
          return s;
        }
    }
    hipStream_t s;
    HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
  }
  void put(int device, hipStream_t s) {
    std::lock_guard<std::mutex> lock(mu_);
    free_.emplace_back(device, s);
  }

private:
  std::mutex mu_;
  std::vector<std::pair<int, hipStream_t>> free_;
};

StreamPool &stream_pool() {
  static StreamPool *pool = new StreamPool;   // outlives every engine
  return *pool;
}

// Default tile shape of the Euler stage kernel when the caller names none.
void choose_tile(int order, int &kernel, int &lanes) {
  // Measured on MI355X (32k and 524k triangles): the point split with four
  // lanes per element is fastest up to P2, the variable split with four lanes
  // from P3 on (its per-lane register state grows with one variable, not four).
  kernel = order >= 3 ? 2 : 1;
  lanes = 4;
}

}  // namespace

struct wasl_hip_engine {
  wasl_hip_numerics num{};
  int dim = 2, p = 0, Np = 1, V = 4;
  int E = 0, F = 0, NB = 0, periodic_pairs = 0;
  int lanes = 1;
  int stage_kernel = 1;       // 1 point split, 2 variable split
  double hmin = 0.0;
  hipStream_t stream = nullptr;
  int device = 0;             // the HIP device current when the engine was created
  size_t device_bytes = 0;

  whip::PeriodicTrace trace;      // host copy for export
  whip::LimiterStencil stencil;   // host copy (built when limiting or on request)
  bool have_stencil = false;
  std::shared_ptr<const whip::Mesh> mesh_ptr;   // kept for stencil export and projection

  // state and stage buffers
  Buffer u, w1, w2;
  // geometry and tables
  Buffer geo, nbr, meta, fn, bdonor, bbasis, table, vtable, inv_jac, mxtable;
  // mixed precision: FP32 edge traces of u, w1, w2 ([E][3][4][4] floats each)
  Buffer tr_u, tr_w1, tr_w2;
  // limiter
  Buffer lnbr, T, nrm, check, flags, list, counters, side, err;
  // control and reductions
  Buffer ctl, mass_w, partials, result;
  whip::RunCtl *ctl_host = nullptr;   // pinned
  double *result_host = nullptr;      // pinned

  // timing of the last step/advance call
  std::vector<hipEvent_t> ev;
  hipEvent_t wall[2] = {nullptr, nullptr};   // around wasl_hip_advance
  double last_wall_ms = 0.0, last_stage_ms = 0.0;
  int64_t last_stage_launches = 0;
  bool timing = false;

  // Captured step graphs for wasl_hip_advance: graphs[k][parity] holds 2^k
  // steps starting at that step parity, so a batch of any length replays as
  // at most kGraphLevels graphs.
  static constexpr int kGraphLevels = 7;   // 1 .. 64 steps
  hipGraphExec_t graphs[kGraphLevels][2] = {};
  bool use_graphs = true;

  ~wasl_hip_engine() {
    for (auto &level : graphs)
      for (hipGraphExec_t x : level)
        if (x)
          (void)hipGraphExecDestroy(x);
    for (auto e : ev)
      (void)hipEventDestroy(e);
    for (auto e : wall)
      if (e)
        (void)hipEventDestroy(e);
    if (ctl_host)
      pinned_slots().put(ctl_host);
    if (result_host)
      pinned_slots().put(result_host);
    // stream-ordered releases need the stream: before it goes
    for (Buffer *b : {&u, &w1, &w2, &geo, &nbr, &meta, &fn, &bdonor, &bbasis, &table, &vtable,
                      &inv_jac, &mxtable, &tr_u, &tr_w1, &tr_w2, &lnbr, &T, &nrm, &check,
                      &flags, &list, &counters, &side, &err, &ctl, &mass_w, &partials, &result})
      b->release();
    if (stream)
      stream_pool().put(device, stream);
  }

  size_t state_size() const { return static_cast<size_t>(V) * Np * E; }

  // 2D Euler on triangles; otherwise scalar advection on lines (dim 1) or
  // periodic triangles (dim 2).
  bool euler() const { return num.equation == WASL_HIP_EULER; }

  bool mixed() const { return euler() && num.precision != WASL_HIP_FP64; }

  // The 2D scalar state lives on the device as U[mode][element]; the C ABI
  // speaks the reference's U[element][mode] (triangle_residual.hpp:3-4).
  bool mode_major_scalar() const { return !euler() && dim == 2; }

  // Host layout -> device layout (a copy in `tmp` when they differ).
  const double *to_device(const double *host, std::vector<double> &tmp) const {
    if (!mode_major_scalar())
      return host;
    tmp.resize(state_size());
    for (int e = 0; e < E; ++e)
      for (int m = 0; m < Np; ++m)
        tmp[static_cast<size_t>(m) * E + e] = host[static_cast<size_t>(e) * Np + m];
    return tmp.data();
  }
  // Device layout (in `dev`) -> host layout in `host`.
  void to_host(const double *dev, double *host) const {
    if (!mode_major_scalar()) {
      std::copy(dev, dev + state_size(), host);
      return;
    }
    for (int e = 0; e < E; ++e)
      for (int m = 0; m < Np; ++m)
        host[static_cast<size_t>(e) * Np + m] = dev[static_cast<size_t>(m) * E + e];
  }

  // The trace buffer that belongs to a state buffer (mixed precision only).
  float *traces_of(const double *p) const {
    if (!mixed())
      return nullptr;
    if (p == u.as<double>())
      return tr_u.as<float>();
    if (p == w1.as<double>())
      return tr_w1.as<float>();
    if (p == w2.as<double>())
      return tr_w2.as<float>();
    return nullptr;
  }

  whip::EulerMesh euler_mesh() const {
    whip::EulerMesh m;
    m.E = E;
    m.geo = geo.as<double>();
    m.nbr = nbr.as<int>();
    m.meta = meta.as<int>();
    m.fn = fn.as<double>();
    m.bdonor = bdonor.as<int>();
    m.bbasis = bbasis.as<double>();
    m.table = table.as<double>();
    m.vtable = vtable.as<double>();
    m.gamma = num.heat_capacity_ratio;
    return m;
  }

  whip::AdvectMesh advect_mesh() const {
    whip::AdvectMesh m;
    m.E = E;
    m.vel = geo.as<double>();
    m.nbr = nbr.as<int>();
    m.meta = meta.as<int>();
    m.flux = fn.as<double>();
    m.bdonor = bdonor.as<int>();
    m.bbasis = bbasis.as<double>();
    m.table = table.as<double>();
    return m;
  }

  whip::EulerLineDev euler_line_mesh() const {
    whip::EulerLineDev m;
    m.E = E;
    m.inv_jac = inv_jac.as<double>();
    m.table = table.as<double>();
    m.gamma = num.heat_capacity_ratio;
    m.periodic = num.boundary_left == WASL_HIP_PERIODIC ? 1 : 0;
    return m;
  }

  whip::LineMeshDev line_mesh() const {
    whip::LineMeshDev m;
    m.E = E;
    m.inv_jac = inv_jac.as<double>();
    m.table = table.as<double>();
    m.wavespeed = num.wavespeed;
    m.left_kind = num.boundary_left;
    m.right_kind = num.boundary_right;
    m.left_value = num.boundary_value_left;
    m.right_value = num.boundary_value_right;
    return m;
  }

  whip::LimiterDev limiter() const {
    whip::LimiterDev L;
    L.E = E;
    L.nbr = lnbr.as<int>();
    L.T = T.as<double>();
    L.nrm = nrm.as<double>();
    L.check = check.as<double>();
    L.gamma = num.heat_capacity_ratio;
    L.density_floor = num.density_floor;
    L.pressure_floor = num.pressure_floor;
    return L;
  }

  void build_stencil_host() {
    if (have_stencil || !euler())
      return;
    if (dim == 1) {
      stencil = whip::line_limiter_stencil(*mesh_ptr, p, num.boundary_left == WASL_HIP_PERIODIC);
      have_stencil = true;
      return;
    }
    stencil = whip::limiter_stencil(*mesh_ptr, p, num.boundary_left == WASL_HIP_PERIODIC
                                                 ? trace.donors
                                                 : std::vector<int32_t>{},
                                    p + 1);
    have_stencil = true;
  }
};

namespace {

// ------------------------------------------------------------- tables ----

// Edge tables shared by the triangle stages: own-side basis at the edge
// points bl [3][NE][Np], the neighbour's basis at the matching points
// br [kl][kr][flip][NE][Np], edge weights we [3][NE].
void edge_tables(int p, double *bl, double *br, double *we) {
  const int Np = whip::tri_modes(p), NE = p + 1;
  const auto edges = whip::edge_rules(p);
  for (int k = 0; k < 3; ++k)
    for (int q = 0; q < NE; ++q) {
      whip::dubiner(edges[k].xi[q], edges[k].eta[q], p, bl + (k * NE + q) * Np);
      we[k * NE + q] = edges[k].w[q];
    }
  for (int kl = 0; kl < 3; ++kl)
    for (int kr = 0; kr < 3; ++kr)
      for (int fl = 0; fl < 2; ++fl)
        for (int q = 0; q < NE; ++q) {
          double rxi, reta;
          whip::neighbor_point(kl, kr, fl != 0, edges[kl].xi[q], edges[kl].eta[q], rxi, reta);
          whip::dubiner(rxi, reta, p, br + (((kl * 3 + kr) * 2 + fl) * NE + q) * Np);
        }
}

std::vector<double> euler_tables(int p) {
  const int Np = whip::tri_modes(p), NQ = (p + 3) * (p + 3), NE = p + 1;
  const int kWx = NQ * Np, kWy = 2 * NQ * Np, kBl = 3 * NQ * Np;
  const int kBr = kBl + 3 * NE * Np, kWe = kBr + 18 * NE * Np, kCs = kWe + 3 * NE;
  std::vector<double> t(kCs + 2 * Np, 0.0);
  const whip::TriRule vol = whip::triangle_rule(p + 2);
  std::vector<double> b(Np), gx(Np), ge(Np);
  for (int q = 0; q < NQ; ++q) {
    whip::dubiner(vol.xi[q], vol.eta[q], p, b.data());
    whip::dubiner_gradient(vol.xi[q], vol.eta[q], p, gx.data(), ge.data());
    for (int m = 0; m < Np; ++m) {
      t[q * Np + m] = b[m];
      t[kWx + q * Np + m] = vol.w[q] * gx[m];
      t[kWy + q * Np + m] = vol.w[q] * ge[m];
      // the stage kernels leave these products out
      if ((whip::dubiner_dxi_zero(m) && gx[m] != 0.0) ||
          (whip::dubiner_deta_zero(m) && ge[m] != 0.0))
        throw std::runtime_error("engine: a Dubiner gradient entry taken as zero is not");
    }
  }
  // column sums of the weighted gradients (the volume term of a constant flux)
  for (int m = 0; m < Np; ++m)
    for (int q = 0; q < NQ; ++q) {
      t[kCs + m] += t[kWx + q * Np + m];
      t[kCs + Np + m] += t[kWy + q * Np + m];
    }
  edge_tables(p, t.data() + kBl, t.data() + kBr, t.data() + kWe);
  return t;
}

// 2D scalar advection (whip::TriAdv<P>): the volume term as differentiation
// matrices Sx[m][n] = sum_q w_q dphi_m/dxi phi_n, Sy likewise with d/deta, over
// the order-p rule of the reference's residual (triangle_residual.hpp:31), then
// the edge tables.
std::vector<double> advect_tables(int p) {
  const int Np = whip::tri_modes(p), NE = p + 1;
  const int kSy = Np * Np, kBl = 2 * Np * Np, kBr = kBl + 3 * NE * Np, kWe = kBr + 18 * NE * Np;
  std::vector<double> t(kWe + 3 * NE, 0.0);
  const whip::TriRule vol = whip::triangle_rule(p);
  std::vector<double> b(Np), gx(Np), ge(Np);
  for (int q = 0; q < vol.size(); ++q) {
    whip::dubiner(vol.xi[q], vol.eta[q], p, b.data());
    whip::dubiner_gradient(vol.xi[q], vol.eta[q], p, gx.data(), ge.data());
    for (int m = 0; m < Np; ++m)
      for (int n = 0; n < Np; ++n) {
        t[m * Np + n] += vol.w[q] * gx[m] * b[n];
        t[kSy + m * Np + n] += vol.w[q] * ge[m] * b[n];
      }
  }
  edge_tables(p, t.data() + kBl, t.data() + kBr, t.data() + kWe);
  return t;
}

// Volume tables padded to a multiple of 8 points with zero weights (the
// padded points reuse point 0's basis values so every state stays finite).
std::vector<double> euler_vtables(int p) {
  const int Np = whip::tri_modes(p), NQ = (p + 3) * (p + 3), NQ8 = (NQ + 7) / 8 * 8;
  std::vector<double> t(3 * NQ8 * Np, 0.0);
  const whip::TriRule vol = whip::triangle_rule(p + 2);
  std::vector<double> b(Np), gx(Np), ge(Np);
  for (int q = 0; q < NQ8; ++q) {
    const int src = q < NQ ? q : 0;
    whip::dubiner(vol.xi[src], vol.eta[src], p, b.data());
    whip::dubiner_gradient(vol.xi[src], vol.eta[src], p, gx.data(), ge.data());
    for (int m = 0; m < Np; ++m) {
      t[q * Np + m] = b[m];
      t[NQ8 * Np + q * Np + m] = q < NQ ? vol.w[q] * gx[m] : 0.0;
      t[2 * NQ8 * Np + q * Np + m] = q < NQ ? vol.w[q] * ge[m] : 0.0;
    }
  }
  return t;
}

uint16_t bf16_bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, sizeof u);
  return static_cast<uint16_t>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// Per-lane operand tables of the mixed-precision stage kernel (whip::Mx<P>):
// every entry is the value the lane of that slot feeds to the matrix core, so
// the kernel reads tab[offset + lane].
template <int P> std::vector<float> mixed_tables() {
  using X = whip::Mx<P>;
  constexpr int Np = X::Np, NQ = X::NQ, NE = X::NE, MT = X::MT, KS = X::KS;
  std::vector<float> t(X::kTable, 0.0f);
  const whip::TriRule vol = whip::triangle_rule(P + 2);
  std::vector<double> B(NQ * Np), Wx(NQ * Np), Wy(NQ * Np), b(Np), gx(Np), ge(Np);
  for (int q = 0; q < NQ; ++q) {
    whip::dubiner(vol.xi[q], vol.eta[q], P, b.data());
    whip::dubiner_gradient(vol.xi[q], vol.eta[q], P, gx.data(), ge.data());
    for (int m = 0; m < Np; ++m) {
      B[q * Np + m] = b[m];
      Wx[q * Np + m] = vol.w[q] * gx[m];
      Wy[q * Np + m] = vol.w[q] * ge[m];
    }
  }
  const auto vb = [&](int q, int m) { return q < NQ && m < Np ? B[q * Np + m] : 0.0; };
  const auto wd = [&](int d, int q, int m) {
    return q < NQ && m < Np ? (d == 0 ? Wx : Wy)[q * Np + m] : 0.0;
  };
  const auto edges = whip::edge_rules(P);
  std::vector<double> Be(3 * NE * Np);
  for (int k = 0; k < 3; ++k)
    for (int s = 0; s < NE; ++s)
      whip::dubiner(edges[k].xi[s], edges[k].eta[s], P, Be.data() + (k * NE + s) * Np);
  const auto eb = [&](int k, int s, int m) {
    return k < 3 && s < NE && m < Np ? Be[(k * NE + s) * Np + m] : 0.0;
  };
  for (int l = 0; l < 64; ++l) {
    const int i = l & 15, kk = l >> 4;
    for (int tt = 0; tt < MT; ++tt) {
      // the states at points from the higher modes only (the kernel works
      // with deviations from the cell average)
      for (int s = 0; s < KS; ++s)
        t[X::kAvol + (tt * KS + s) * 64 + l] =
            KS * kk + s == 0 ? 0.0f : static_cast<float>(vb(X::point(tt, i), KS * kk + s));
      // K step r of tile tt takes point X::point(tt, 4 kk + r) from row group kk
      for (int r = 0; r < 4; ++r)
        for (int d = 0; d < 2; ++d)
          t[X::kAgrad + ((tt * 4 + r) * 2 + d) * 64 + l] =
              static_cast<float>(wd(d, X::point(tt, 4 * kk + r), X::row_mode(i)));
      // bf16: A[i][4 kk + c] with four consecutive K entries per lane; in the
      // volume product K entry 4 kk + c is mode KS kk + c (c < KS)
      uint16_t hv[4], hx[4], he[4];
      for (int c = 0; c < 4; ++c) {
        hv[c] = bf16_bits(static_cast<float>(c < KS && KS * kk + c != 0
                                                 ? vb(X::point(tt, i), KS * kk + c) : 0.0));
        hx[c] = bf16_bits(static_cast<float>(wd(0, X::point(tt, 4 * kk + c), X::row_mode(i))));
        he[c] = bf16_bits(static_cast<float>(wd(1, X::point(tt, 4 * kk + c), X::row_mode(i))));
      }
      std::memcpy(&t[X::kBvol16 + (tt * 64 + l) * 2], hv, sizeof hv);
      std::memcpy(&t[X::kBgrad16 + ((tt * 2 + 0) * 64 + l) * 2], hx, sizeof hx);
      std::memcpy(&t[X::kBgrad16 + ((tt * 2 + 1) * 64 + l) * 2], he, sizeof he);
    }
    // Trace rows and lifting steps follow the face pairs: row 4g + t of the
    // trace products and K step t of the lifting (row group g) are pair
    // PPG g + t, that is edge (PPG g + t) / NE, point (PPG g + t) % NE.
    const auto pair_basis = [&](int g, int tt, int m) {
      const int q = X::PPG * g + tt;
      return tt < X::PPG && q < X::PAIRS ? eb(q / NE, q % NE, m) : 0.0;
    };
    for (int s = 0; s < KS; ++s)
      t[X::kAface + s * 64 + l] = static_cast<float>(pair_basis(i >> 2, i & 3, KS * kk + s));
    for (int s = 0; s < X::PPG; ++s)
      t[X::kAlift + s * 64 + l] = static_cast<float>(pair_basis(kk, s, X::row_mode(i)));
  }
  for (int k = 0; k < 3; ++k)
    for (int s = 0; s < NE; ++s) {
      for (int m = 0; m < Np; ++m)
        t[X::kBe + (k * NE + s) * Np + m] = static_cast<float>(Be[(k * NE + s) * Np + m]);
      t[X::kWe + k * NE + s] = static_cast<float>(edges[k].w[s]);
    }
  return t;
}

std::vector<float> mixed_tables(int p) {
  switch (p) {
  case 0: return mixed_tables<0>();
  case 1: return mixed_tables<1>();
  case 2: return mixed_tables<2>();
  default: return mixed_tables<3>();
  }
}

// 1D Euler (whip::LineE<P>): basis and w_q dphi/dxi at the p+1 Gauss points,
// then phi(-1), phi(+1).
std::vector<double> euler_line_tables(int p) {
  const int Np = p + 1, NQ = p + 1;
  std::vector<double> t(2 * NQ * Np + 2 * Np, 0.0);
  const whip::LineRule g = whip::gauss_legendre(NQ);
  std::vector<double> b(Np), d(Np);
  for (int q = 0; q < NQ; ++q) {
    whip::legendre(g.x[q], p, b.data());
    whip::legendre_gradient(g.x[q], p, d.data());
    for (int m = 0; m < Np; ++m) {
      t[q * Np + m] = b[m];
      t[NQ * Np + q * Np + m] = g.w[q] * d[m];
    }
  }
  whip::legendre(-1.0, p, t.data() + 2 * NQ * Np);
  whip::legendre(1.0, p, t.data() + 2 * NQ * Np + Np);
  return t;
}

std::vector<double> line_tables(int p) {
  const int Np = p + 1, NQ = p > 0 ? p : 1;
  std::vector<double> t(2 * NQ * Np + 2 * Np, 0.0);
  const whip::LineRule g = whip::gauss_legendre(NQ);
  std::vector<double> b(Np), d(Np);
  for (int q = 0; q < NQ; ++q) {
    whip::legendre(g.x[q], p, b.data());
    whip::legendre_gradient(g.x[q], p, d.data());
    for (int m = 0; m < Np; ++m) {
      t[q * Np + m] = b[m];
      t[NQ * Np + q * Np + m] = g.w[q] * d[m];
    }
  }
  whip::legendre(-1.0, p, t.data() + 2 * NQ * Np);
  whip::legendre(1.0, p, t.data() + 2 * NQ * Np + Np);
  return t;
}

// ------------------------------------------------------------- launch ----

template <int P>
void launch_euler_p(const wasl_hip_engine &g, const whip::StageArgs &a) {
  const whip::EulerMesh m = g.euler_mesh();
  if (g.num.precision != WASL_HIP_FP64) {
    // 64 elements per 256-thread block (16 per wave)
    const dim3 grid((g.E + 63) / 64);
    const float *t = g.mxtable.as<float>();
    if (g.num.precision == WASL_HIP_BF16)
      hipLaunchKernelGGL((whip::euler_stage_mx<P, true>), grid, dim3(256), 0, g.stream, m, a, t);
    else
      hipLaunchKernelGGL((whip::euler_stage_mx<P, false>), grid, dim3(256), 0, g.stream, m, a, t);
    return;
  }
  const int per = g.stage_kernel == 2 ? 256 / g.lanes : 256 / g.lanes;
  const dim3 grid((g.E + per - 1) / per);
  if (g.stage_kernel == 2) {
    if (g.lanes == 8)
      hipLaunchKernelGGL((whip::euler_stage_vs<P, 2>), grid, dim3(256), 0, g.stream, m, a);
    else
      hipLaunchKernelGGL((whip::euler_stage_vs<P, 1>), grid, dim3(256), 0, g.stream, m, a);
    return;
  }
  switch (g.lanes) {
  case 1:
    hipLaunchKernelGGL((whip::euler_stage<P, 1>), grid, dim3(256), 0, g.stream, m, a);
    break;
  case 2:
    hipLaunchKernelGGL((whip::euler_stage<P, 2>), grid, dim3(256), 0, g.stream, m, a);
    break;
  default:
    hipLaunchKernelGGL((whip::euler_stage<P, 4>), grid, dim3(256), 0, g.stream, m, a);
    break;
  }
}

template <int P> void launch_advect_p(const wasl_hip_engine &g, const whip::StageArgs &a) {
  const dim3 grid((g.E + 255) / 256);
  hipLaunchKernelGGL((whip::advect_stage<P>), grid, dim3(256), 0, g.stream, g.advect_mesh(), a);
}

template <int P> void launch_euler_line_p(const wasl_hip_engine &g, const whip::StageArgs &a) {
  // whip::kLineCells cells per wave, four waves per workgroup
  const int waves = (g.E + whip::kLineCells - 1) / whip::kLineCells;
  const dim3 grid((waves + 3) / 4);
  hipLaunchKernelGGL((whip::euler_line_stage<P>), grid, dim3(256), 0, g.stream,
                     g.euler_line_mesh(), a);
}

template <int P> void launch_line_p(const wasl_hip_engine &g, const whip::StageArgs &a) {
  const dim3 grid((g.E + 255) / 256);
  hipLaunchKernelGGL((whip::line_stage<P>), grid, dim3(256), 0, g.stream, g.line_mesh(), a);
}

void launch_stage(const wasl_hip_engine &g, const whip::StageArgs &a) {
  if (g.dim == 2 && !g.euler()) {
    switch (g.p) {
    case 0: launch_advect_p<0>(g, a); break;
    case 1: launch_advect_p<1>(g, a); break;
    case 2: launch_advect_p<2>(g, a); break;
    default: launch_advect_p<3>(g, a); break;
    }
  } else if (g.dim == 2) {
    switch (g.p) {
    case 0: launch_euler_p<0>(g, a); break;
    case 1: launch_euler_p<1>(g, a); break;
    case 2: launch_euler_p<2>(g, a); break;
    default: launch_euler_p<3>(g, a); break;
    }
  } else if (g.euler()) {
    switch (g.p) {
    case 0: launch_euler_line_p<0>(g, a); break;
    case 1: launch_euler_line_p<1>(g, a); break;
    case 2: launch_euler_line_p<2>(g, a); break;
    default: launch_euler_line_p<3>(g, a); break;
    }
  } else {
    switch (g.p) {
    case 0: launch_line_p<0>(g, a); break;
    case 1: launch_line_p<1>(g, a); break;
    case 2: launch_line_p<2>(g, a); break;
    default: launch_line_p<3>(g, a); break;
    }
  }
  HIPCHECK(hipGetLastError());
}

// FP32 edge traces of a state buffer (mixed precision), on `stream`.
void launch_traces(const wasl_hip_engine &g, const double *U, float *T, const float *mxtab,
                   hipStream_t stream) {
  const dim3 grid((g.E + 63) / 64);
  switch (g.p) {
  case 0:
    hipLaunchKernelGGL(whip::mx_traces<0>, grid, dim3(256), 0, stream, U, g.E, T, mxtab);
    break;
  case 1:
    hipLaunchKernelGGL(whip::mx_traces<1>, grid, dim3(256), 0, stream, U, g.E, T, mxtab);
    break;
  case 2:
    hipLaunchKernelGGL(whip::mx_traces<2>, grid, dim3(256), 0, stream, U, g.E, T, mxtab);
    break;
  default:
    hipLaunchKernelGGL(whip::mx_traces<3>, grid, dim3(256), 0, stream, U, g.E, T, mxtab);
    break;
  }
  HIPCHECK(hipGetLastError());
}

// Working memory of one limiter call: the limited values staged before the
// scatter, the troubled list, its counters and the error word (err[0] code,
// err[1] first element whose average is below a floor).
struct LimitBuffers {
  double *side;
  int *list, *counters, *err;
};

LimitBuffers engine_buffers(const wasl_hip_engine &g) {
  return {g.side.as<double>(), g.list.as<int>(), g.counters.as<int>(), g.err.as<int>()};
}

template <int P>
void launch_limit_p(const wasl_hip_engine &g, double *U, const int *count, whip::RunCtl *ctl,
                    int parity, const LimitBuffers &b, hipStream_t stream) {
  const whip::LimiterDev L = g.limiter();
  const dim3 grid(whip::limit_blocks(g.E)), grid256(whip::kScatterBlocks);
  // Outside the stepping loop, err[1] holds the first element whose average
  // is below a floor (INT_MAX when none); the scatter stops there.
  const int *fail = ctl ? nullptr : b.err + 1;
  if (g.dim == 1) {
    hipLaunchKernelGGL((whip::limit_troubled_line<P>), grid, dim3(64), 0, stream, L, U, b.side,
                       b.list, count, ctl, parity, b.err);
    hipLaunchKernelGGL((whip::scatter_limited_line<P>), grid256, dim3(256), 0, stream, U, g.E,
                       b.side, b.list, count, ctl, parity, fail);
    return;
  }
  hipLaunchKernelGGL((whip::limit_troubled<P>), grid, dim3(64), 0, stream, L, U, b.side, b.list,
                     count, ctl, parity, b.err);
  hipLaunchKernelGGL((whip::scatter_limited<P>), grid256, dim3(256), 0, stream, U, g.E, b.side,
                     b.list, count, ctl, parity, g.traces_of(U), g.mxtable.as<float>(), fail);
}

// Limit the troubled elements listed under counter `slot` of b.
void launch_limit(const wasl_hip_engine &g, double *U, int slot, whip::RunCtl *ctl, int parity,
                  const LimitBuffers &b, hipStream_t stream) {
  const int *count = b.counters + slot;
  switch (g.p) {
  case 0: return;  // P0 is returned unchanged (euler_modal_limiter.cpp:73-74)
  case 1: launch_limit_p<1>(g, U, count, ctl, parity, b, stream); break;
  case 2: launch_limit_p<2>(g, U, count, ctl, parity, b, stream); break;
  default: launch_limit_p<3>(g, U, count, ctl, parity, b, stream); break;
  }
  HIPCHECK(hipGetLastError());
}

void launch_limit(const wasl_hip_engine &g, double *U, int slot, whip::RunCtl *ctl, int parity) {
  launch_limit(g, U, slot, ctl, parity, engine_buffers(g), g.stream);
}

// Troubled-cell sensor of U into b's list (counter 3) and flags (may be null).
template <int P>
void launch_sensor_p(const wasl_hip_engine &g, const double *U, uint8_t *flags,
                     const LimitBuffers &b, hipStream_t stream) {
  constexpr int kTriNp = (P + 1) * (P + 2) / 2;
  const auto kernel = g.dim == 1 ? whip::sensor<P + 1> : whip::sensor<kTriNp>;
  hipLaunchKernelGGL(kernel, dim3((g.E + 255) / 256), dim3(256), 0, stream, U, g.E,
                     g.num.sensor_threshold, g.stencil.highest_begin, flags, b.list,
                     b.counters + 3);
}

void launch_sensor(const wasl_hip_engine &g, const double *U, uint8_t *flags,
                   const LimitBuffers &b, hipStream_t stream) {
  switch (g.p) {
  case 1: launch_sensor_p<1>(g, U, flags, b, stream); break;
  case 2: launch_sensor_p<2>(g, U, flags, b, stream); break;
  default: launch_sensor_p<3>(g, U, flags, b, stream); break;
  }
  HIPCHECK(hipGetLastError());
}

bool limiting(const wasl_hip_engine &g) { return g.euler() && g.num.limiting && g.p > 0; }

// Arguments of Shu-Osher stage s (time/ssp_rk3.hpp:45-60): U1 = U + dt R(U);
// U2 = 3/4 U + 1/4 (U1 + dt R(U1)); U = 1/3 U + 2/3 (U2 + dt R(U2)).
whip::StageArgs stage_args(const wasl_hip_engine &g, int s, double dt, whip::RunCtl *ctl,
                           int parity, bool sensor) {
  double *U = g.u.as<double>(), *A = g.w1.as<double>(), *B = g.w2.as<double>();
  const double *in[3] = {U, A, B};
  double *out[3] = {A, B, U};
  const double ca[3] = {0.0, 0.75, 1.0 / 3.0};
  const double cb[3] = {1.0, 0.25, 2.0 / 3.0};
  whip::StageArgs a{};
  a.uin = in[s];
  a.un = U;
  a.out = out[s];
  a.ca = ca[s];
  a.cb = cb[s];
  a.dt = dt;
  a.ctl = ctl;
  a.parity = parity;
  a.stage = s;
  a.residual_only = 0;
  a.want_lambda = (g.euler() && ctl && s == 2) ? 1 : 0;
  a.sensor = sensor ? 1 : 0;
  a.threshold = g.num.sensor_threshold;
  a.highest_begin = g.stencil.highest_begin;
  a.flags = nullptr;
  a.list = g.list.as<int>();
  a.counters = g.counters.as<int>();
  a.tin = g.traces_of(in[s]);
  a.tout = g.traces_of(out[s]);
  return a;
}

// One SSP-RK3 step: the three fused stages, the limiter after each.
void launch_step(wasl_hip_engine &g, double dt, whip::RunCtl *ctl, int parity, bool timed,
                 int &timed_index) {
  const bool lim = limiting(g);
  for (int s = 0; s < 3; ++s) {
    const whip::StageArgs a = stage_args(g, s, dt, ctl, parity, lim);
    if (timed)
      HIPCHECK(hipEventRecord(g.ev[2 * timed_index], g.stream));
    launch_stage(g, a);
    if (timed) {
      HIPCHECK(hipEventRecord(g.ev[2 * timed_index + 1], g.stream));
      ++timed_index;
    }
    if (lim)
      launch_limit(g, a.out, s, ctl, parity);
  }
}

void ensure_events(wasl_hip_engine &g, size_t pairs) {
  while (g.ev.size() < 2 * pairs) {
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    g.ev.push_back(e);
  }
}

// Sum of the per-stage event intervals recorded by launch_step.
double stage_time_ms(wasl_hip_engine &g, int pairs) {
  double total = 0.0;
  for (int i = 0; i < pairs; ++i) {
    float ms = 0.0f;
    HIPCHECK(hipEventElapsedTime(&ms, g.ev[2 * i], g.ev[2 * i + 1]));
    total += ms;
  }
  return total;
}

// The engine's error word: err[0] the device error code, err[1] the first
// element (ascending) that the standalone limiter left unlimited because its
// cell average is below a floor, INT_MAX when none.
void reset_device_error(wasl_hip_engine &g) {
  const int clear[2] = {0, INT_MAX};
  HIPCHECK(hipMemcpyAsync(g.err.p, clear, sizeof clear, hipMemcpyHostToDevice, g.stream));
  HIPCHECK(hipStreamSynchronize(g.stream));
}

void raise_device_error(wasl_hip_engine &g) {
  // read back through the engine's pinned slot (a pageable copy is slower)
  int *code = reinterpret_cast<int *>(g.result_host);
  HIPCHECK(hipMemcpyAsync(code, g.err.p, 2 * sizeof(int), hipMemcpyDeviceToHost, g.stream));
  HIPCHECK(hipStreamSynchronize(g.stream));
  if (code[0] != 0 || code[1] != INT_MAX) {
    reset_device_error(g);
    HIPCHECK(hipMemsetAsync(g.counters.p, 0, g.counters.bytes, g.stream));
    HIPCHECK(hipStreamSynchronize(g.stream));
    throw SimulationError(device_error_text(code[0]));
  }
}

double mass_of(wasl_hip_engine &g) {
  const int blocks = (g.E + 255) / 256;
  hipLaunchKernelGGL(whip::mass_partials, dim3(blocks), dim3(256), 0, g.stream,
                     g.u.as<double>(), g.euler() || g.dim == 2 ? 1 : g.Np, g.mass_w.as<double>(), g.E,
                     g.partials.as<double>());
  hipLaunchKernelGGL(whip::mass_finish, dim3(1), dim3(64), 0, g.stream, g.partials.as<double>(),
                     blocks, g.result.as<double>());
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(g.result_host, g.result.p, sizeof(double), hipMemcpyDeviceToHost,
                          g.stream));
  HIPCHECK(hipStreamSynchronize(g.stream));
  return *g.result_host;
}

// Largest |u|+c of the resident state's cell averages (throws when one is
// unphysical), or the advection speed |a| (1D) or |v| (2D).
double advection_speed(const wasl_hip_engine &g) {
  return g.dim == 1 ? std::fabs(g.num.wavespeed) : std::hypot(g.num.wavespeed, g.num.velocity_y);
}

double resident_speed(wasl_hip_engine &g) {
  if (!g.euler())
    return advection_speed(g);
  whip::RunCtl *h = g.ctl_host;
  std::memset(h, 0, sizeof(*h));
  HIPCHECK(hipMemcpyAsync(g.ctl.p, h, sizeof(*h), hipMemcpyHostToDevice, g.stream));
  hipLaunchKernelGGL(whip::cfl_speed, dim3((g.E + 255) / 256), dim3(256), 0, g.stream,
                     g.u.as<double>(), g.E, g.Np, g.num.heat_capacity_ratio,
                     g.ctl.as<whip::RunCtl>(), 0, g.dim);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(h, g.ctl.p, sizeof(*h), hipMemcpyDeviceToHost, g.stream));
  HIPCHECK(hipStreamSynchronize(g.stream));
  if (h->s[0].bad)
    throw SimulationError("simulation: Euler cell average is not physical");
  double speed;
  std::memcpy(&speed, &h->s[0].lambda, sizeof(double));
  return speed;
}

// ------------------------------------------------------------- autotuner ----
//
// Tile shape of the FP64 Euler stage kernel (point split with 1, 2 or 4 lanes
// per element, variable split with 4 or 8), keyed on dimension, equation,
// order, precision and a mesh-size class. The first engine of a key times
// every candidate on its own mesh -- second-stage launches (residual, RK
// combination with U^n, CFL speed) of a free-stream state without step
// control, HIP events on its stream, the faster of two rounds -- and the
// fastest serves every later engine of that key in the process.
// WASL_HIP_AUTOTUNE=0 keeps choose_tile's table.

struct TuneKey {
  int dim, equation, order, precision, size_class;
  bool operator<(const TuneKey &o) const {
    return std::tie(dim, equation, order, precision, size_class) <
           std::tie(o.dim, o.equation, o.order, o.precision, o.size_class);
  }
};

std::mutex g_tune_mu;
std::map<TuneKey, std::pair<int, int>> g_tuned;

int size_class(int E) { return E < 8192 ? 0 : E < 131072 ? 1 : 2; }

void launch_stage(const wasl_hip_engine &g, const whip::StageArgs &a);

void autotune(wasl_hip_engine &g) {
  if (const char *env = std::getenv("WASL_HIP_AUTOTUNE"))
    if (std::atoi(env) == 0)
      return;
  const TuneKey key{g.dim, g.num.equation, g.p, g.num.precision, size_class(g.E)};
  // Held through the timing: engines created on several threads (the
  // server's setup threads) tune a key once and then all take its choice.
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = g_tuned.find(key);
  if (it != g_tuned.end()) {
    g.stage_kernel = it->second.first;
    g.lanes = it->second.second;
    return;
  }
  // free stream (rho, u, v, p) = (1, 0.5, 0.25, 1): mode 0 only
  std::vector<double> U(g.state_size(), 0.0);
  const double gm = g.num.heat_capacity_ratio, s2 = std::sqrt(2.0);
  const double cons[4] = {1.0, 0.5, 0.25, 1.0 / (gm - 1.0) + 0.5 * (0.25 + 0.0625)};
  for (int v = 0; v < 4; ++v)
    for (int e = 0; e < g.E; ++e)
      U[static_cast<size_t>(v) * g.Np * g.E + e] = s2 * cons[v];
  HIPCHECK(hipMemcpyAsync(g.w1.p, U.data(), U.size() * sizeof(double), hipMemcpyHostToDevice,
                          g.stream));
  // A second-stage launch without step control: residual, RK combination
  // with U^n and the CFL speed, the paths a stepping stage takes.
  whip::StageArgs a{};
  a.uin = g.w1.as<double>();
  a.un = g.w1.as<double>();
  a.out = g.w2.as<double>();
  a.cb = 0.25;
  a.ca = 0.75;
  a.stage = 1;
  a.want_lambda = 1;
  hipEvent_t t0, t1;
  HIPCHECK(hipEventCreate(&t0));
  HIPCHECK(hipEventCreate(&t1));
  const std::pair<int, int> candidates[] = {{1, 1}, {1, 2}, {1, 4}, {2, 4}, {2, 8}};
  constexpr int kCand = sizeof(candidates) / sizeof(candidates[0]);
  float ms_of[kCand];
  // two rounds over the candidates, each candidate keeping its faster round
  for (int round = 0; round < 2; ++round)
    for (int k = 0; k < kCand; ++k) {
      g.stage_kernel = candidates[k].first;
      g.lanes = candidates[k].second;
      for (int i = 0; i < 2; ++i)
        launch_stage(g, a);
      HIPCHECK(hipEventRecord(t0, g.stream));
      constexpr int kReps = 5;
      for (int i = 0; i < kReps; ++i)
        launch_stage(g, a);
      HIPCHECK(hipEventRecord(t1, g.stream));
      HIPCHECK(hipEventSynchronize(t1));
      float ms = 0.0f;
      HIPCHECK(hipEventElapsedTime(&ms, t0, t1));
      ms_of[k] = round == 0 ? ms : std::min(ms_of[k], ms);
    }
  std::pair<int, int> best = candidates[0];
  float best_ms = ms_of[0];
  for (int k = 1; k < kCand; ++k)
    if (ms_of[k] < best_ms) {
      best_ms = ms_of[k];
      best = candidates[k];
    }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  g.stage_kernel = best.first;
  g.lanes = best.second;
  g_tuned.emplace(key, best);
}

// Edge records of a triangle mesh, gathered per element so a stage kernel
// reads its three faces without atomics: neighbour element (-1 across the
// boundary) and meta bits per [edge][element], the face normal (from the left
// side) and edge scale per [3][edge][element], and for each boundary face (in
// face order, its slot in the meta bits) the periodic donors and donor basis
// of its edge points when a trace is given.
struct EdgeRecords {
  std::vector<int> nbr, meta, bdon;
  std::vector<double> fn, bbas;
  int slots = 0, pairs = 0;
};

// A periodic boundary face whose donor, at each of its Gauss points, is the
// element across the opposite boundary edge, at one of that edge's own Gauss
// points (a conforming periodic mesh): the partner face, the partner's local
// edge, and the two orderings an interior face records.
struct PeriodicPair {
  int partner = -1, kr = 0, flip = 0, rev = 0;
};

std::vector<PeriodicPair> periodic_pairs(const whip::Mesh &mesh, int p,
                                         const whip::PeriodicTrace &trace) {
  const size_t F = mesh.faces.size(), E = mesh.tris.size();
  const int NE = p + 1;
  std::vector<PeriodicPair> out(F);
  if (trace.points != NE || trace.donors.size() != F * NE)
    return out;
  double x0 = std::numeric_limits<double>::max(), x1 = std::numeric_limits<double>::lowest();
  double y0 = x0, y1 = x1;
  for (const auto &t : mesh.tris)
    for (int k = 0; k < 3; ++k) {
      x0 = std::min(x0, t.x[k]);
      x1 = std::max(x1, t.x[k]);
      y0 = std::min(y0, t.y[k]);
      y1 = std::max(y1, t.y[k]);
    }
  const double tol = 1e-9 * std::max(x1 - x0, y1 - y0);
  std::vector<std::array<int, 3>> bface(E, {-1, -1, -1});
  for (size_t f = 0; f < F; ++f)
    if (mesh.faces[f].boundary())
      bface[mesh.faces[f].left][mesh.faces[f].left_edge] = static_cast<int>(f);
  const auto edges = whip::edge_rules(p);
  // the image of the face's Gauss point q across the box, as the donor
  // search places it
  const auto image = [&](const whip::TriFace &x, int q, double &px, double &py) {
    mesh.tris[x.left].to_physical(edges[x.left_edge].xi[q], edges[x.left_edge].eta[q], px, py);
    if (std::fabs(x.nx) > 0.5 && std::fabs(x.ny) < 1e-10)
      px += x.nx < 0.0 ? x1 - x0 : x0 - x1;
    else if (std::fabs(x.ny) > 0.5 && std::fabs(x.nx) < 1e-10)
      py += x.ny < 0.0 ? y1 - y0 : y0 - y1;
    else
      px = std::numeric_limits<double>::quiet_NaN();
  };
  const auto same_donor = [&](size_t f, int d) {
    for (int q = 0; q < NE; ++q)
      if (trace.donors[f * NE + q] != d)
        return false;
    return true;
  };
  for (size_t f = 0; f < F; ++f) {
    const whip::TriFace &x = mesh.faces[f];
    if (!x.boundary())
      continue;
    const int d = trace.donors[f * NE];
    if (d < 0 || d == x.left || !same_donor(f, d))
      continue;
    for (int k = 0; k < 3; ++k) {
      const int g = bface[d][k];
      if (g < 0 || !same_donor(static_cast<size_t>(g), x.left))
        continue;
      const whip::Triangle &dt = mesh.tris[d];
      int found = -1;
      for (int flip = 0; flip < 2 && found < 0; ++flip) {
        bool ok = true;
        for (int q = 0; q < NE && ok; ++q) {
          double px, py, rxi, reta, dx, dy;
          image(x, q, px, py);
          whip::neighbor_point(x.left_edge, k, flip != 0, edges[x.left_edge].xi[q],
                               edges[x.left_edge].eta[q], rxi, reta);
          dt.to_physical(rxi, reta, dx, dy);
          ok = std::hypot(px - dx, py - dy) <= tol;
        }
        if (ok)
          found = flip;
      }
      if (found < 0)
        continue;
      int rev = 0;
      if (NE > 1) {
        double px, py, r0x, r0y, r1x, r1y;
        image(x, 0, px, py);
        dt.to_physical(edges[k].xi[0], edges[k].eta[0], r0x, r0y);
        dt.to_physical(edges[k].xi[NE - 1], edges[k].eta[NE - 1], r1x, r1y);
        rev = std::hypot(px - r0x, py - r0y) > std::hypot(px - r1x, py - r1y) ? 1 : 0;
      }
      out[f] = {g, k, found, rev};
      break;
    }
  }
  // only pairs found from both sides
  for (size_t f = 0; f < F; ++f)
    if (out[f].partner >= 0 && out[static_cast<size_t>(out[f].partner)].partner != static_cast<int>(f))
      out[f].partner = -1;
  return out;
}

EdgeRecords edge_records(const whip::Mesh &mesh, int p, int Np, const whip::PeriodicTrace *trace) {
  const size_t E = mesh.tris.size();
  const int NE = p + 1;
  EdgeRecords r;
  r.nbr.assign(3 * E, -2);
  r.meta.assign(3 * E, 0);
  r.fn.assign(9 * E, 0.0);
  const auto edges = whip::edge_rules(p);
  const std::vector<PeriodicPair> pairs =
      trace ? periodic_pairs(mesh, p, *trace) : std::vector<PeriodicPair>(mesh.faces.size());
  int slot = 0;
  for (size_t f = 0; f < mesh.faces.size(); ++f) {
    const whip::TriFace &x = mesh.faces[f];
    const whip::Triangle &l = mesh.tris[x.left];
    const double scale = l.edge_length(x.left_edge) / whip::reference_edge_length(x.left_edge);
    const int flip = x.boundary() ? 0 : (x.flip() ? 1 : 0);
    const int kr = x.boundary() ? 0 : x.right_edge;
    int m = (x.left_edge << 1) | (kr << 3) | (flip << 5);
    const auto put = [&](int e, int k, int other, int bits) {
      const size_t i = static_cast<size_t>(k) * E + e;
      if (r.nbr[i] != -2)
        throw std::runtime_error("engine: element edge belongs to two faces");
      r.nbr[i] = other;
      r.meta[i] = bits;
      r.fn[(3 * k) * E + e] = x.nx;
      r.fn[(3 * k + 1) * E + e] = x.ny;
      r.fn[(3 * k + 2) * E + e] = scale;
    };
    const PeriodicPair &pp = pairs[f];
    if (x.boundary() && pp.partner >= 0) {
      // A conforming periodic pair becomes an interior face: the face listed
      // first is its left side, and the stage evaluates the partner's trace
      // at the left side's points from the partner's coefficients, with the
      // left normal, as for any interior face.
      if (pp.partner < static_cast<int>(f))
        continue;
      const whip::TriFace &y = mesh.faces[static_cast<size_t>(pp.partner)];
      const int mp = (x.left_edge << 1) | (pp.kr << 3) | (pp.flip << 5);
      const int rb = pp.rev ? whip::kMetaReverse : 0;
      put(x.left, x.left_edge, y.left, mp | whip::kMetaLeft | rb);
      put(y.left, y.left_edge, x.left, mp | rb);
      ++r.pairs;
    } else if (x.boundary()) {
      put(x.left, x.left_edge, -1, m | whip::kMetaLeft | whip::kMetaBoundary | (slot << 8));
      if (trace) {
        for (int q = 0; q < NE; ++q) {
          r.bdon.push_back(trace->donors[f * NE + q]);
          for (int k = 0; k < Np; ++k)
            r.bbas.push_back(trace->basis[(f * NE + q) * Np + k]);
        }
      }
      ++slot;
    } else {
      // Whether the right side lists the face's Gauss points in the reverse
      // order of the left side (each in its own edge rule), by position.
      int rev = 0;
      if (NE > 1) {
        const whip::Triangle &rt = mesh.tris[x.right];
        double lx, ly, r0x, r0y, r1x, r1y;
        l.to_physical(edges[x.left_edge].xi[0], edges[x.left_edge].eta[0], lx, ly);
        rt.to_physical(edges[x.right_edge].xi[0], edges[x.right_edge].eta[0], r0x, r0y);
        rt.to_physical(edges[x.right_edge].xi[NE - 1], edges[x.right_edge].eta[NE - 1], r1x,
                       r1y);
        rev = std::hypot(lx - r0x, ly - r0y) > std::hypot(lx - r1x, ly - r1y) ? 1 : 0;
      }
      const int rb = rev ? whip::kMetaReverse : 0;
      put(x.left, x.left_edge, x.right, m | whip::kMetaLeft | rb);
      put(x.right, x.right_edge, x.left, m | rb);
    }
  }
  if (std::find(r.nbr.begin(), r.nbr.end(), -2) != r.nbr.end())
    throw std::runtime_error("engine: element edge without a face");
  r.slots = slot;
  return r;
}

void create_engine(std::shared_ptr<const whip::Mesh> shared, const wasl_hip_numerics &num,
                   wasl_hip_engine &g) {
  const whip::Mesh &mesh = *shared;
  g.num = num;
  g.p = num.polynomial_order;
  if (g.p < 0 || g.p > 3)
    throw std::runtime_error("engine: polynomial order must be 0..3");
  if (num.precision != WASL_HIP_FP64 && num.precision != WASL_HIP_MIXED &&
      num.precision != WASL_HIP_BF16)
    throw std::runtime_error("engine: precision must be FP64, MIXED or BF16");
  g.mesh_ptr = std::move(shared);
  HIPCHECK(hipGetDevice(&g.device));
  g.stream = stream_pool().get(g.device);
  if (const char *env = std::getenv("WASL_HIP_GRAPHS"))
    g.use_graphs = std::atoi(env) != 0;
  static_assert(sizeof(whip::RunCtl) <= kPinnedSlot, "step control fits a pinned slot");
  g.ctl_host = static_cast<whip::RunCtl *>(pinned_slots().get());
  g.result_host = static_cast<double *>(pinned_slots().get());
  keep_pool_memory(g.device);
  size_t &tot = g.device_bytes;

  if (num.equation == WASL_HIP_EULER && mesh.dimension == 1) {
    // run_line_euler (src/case/euler_line.cpp:14-48): periodic or outflow
    // ends, the line limiter stencil, FP64.
    const int kinds[2] = {num.boundary_left, num.boundary_right};
    for (int k : kinds) {
      if (k < 0 || k > 2)
        throw std::runtime_error("boundaries: invalid boundary kind");
      if (k == WASL_HIP_INFLOW)
        throw std::runtime_error("line Euler exterior: inflow is not implemented");
    }
    if ((num.boundary_left == WASL_HIP_PERIODIC) != (num.boundary_right == WASL_HIP_PERIODIC))
      throw std::runtime_error("boundaries: periodic ends must be paired");
    if (num.precision != WASL_HIP_FP64)
      throw std::runtime_error("engine: 1D Euler runs in FP64");
    g.dim = 1;
    g.V = 4;
    g.Np = g.p + 1;
    g.E = static_cast<int>(mesh.lines.size());
    g.F = static_cast<int>(mesh.lfaces.size());
    g.lanes = 1;
    g.stage_kernel = 1;
    std::vector<double> ij(g.E), mw(g.E);
    for (int e = 0; e < g.E; ++e) {
      const double jac = (mesh.lines[e].xr - mesh.lines[e].xl) / 2.0;
      ij[e] = 1.0 / jac;
      mw[e] = std::sqrt(2.0) * jac;
    }
    upload(g.inv_jac, ij, tot, g.stream);
    upload(g.mass_w, mw, tot, g.stream);
    upload(g.table, euler_line_tables(g.p), tot, g.stream);
    g.hmin = whip::minimum_length(mesh);
    const size_t E = static_cast<size_t>(g.E);
    if (num.limiting && g.p > 0) {
      g.build_stencil_host();
      upload(g.lnbr, g.stencil.neighbors, tot, g.stream);
      upload(g.T, g.stencil.transforms, tot, g.stream);
      std::vector<double> nrm(2 * E);
      for (size_t e = 0; e < E; ++e) {
        nrm[2 * e] = g.stencil.nx[e];
        nrm[2 * e + 1] = g.stencil.ny[e];
      }
      upload(g.nrm, nrm, tot, g.stream);
      upload(g.check, g.stencil.check, tot, g.stream);
      g.side.alloc(E * 4 * g.Np * sizeof(double), tot, g.stream);
    } else {
      g.stencil.highest_begin = g.p > 0 ? g.Np - 1 : g.Np;
    }
    g.flags.alloc(E, tot, g.stream);
    g.list.alloc(E * sizeof(int), tot, g.stream);
  } else if (num.equation == WASL_HIP_EULER) {
    if (mesh.dimension != 2)
      throw std::runtime_error("engine: Euler runs on line or triangle meshes");
    if (num.boundary_left != WASL_HIP_PERIODIC && num.boundary_left != WASL_HIP_OUTFLOW)
      throw std::runtime_error("simulation: 2D Euler boundaries must be periodic or outflow");
    g.dim = 2;
    g.V = 4;
    g.Np = whip::tri_modes(g.p);
    g.E = static_cast<int>(mesh.tris.size());
    g.F = static_cast<int>(mesh.faces.size());
    choose_tile(g.p, g.stage_kernel, g.lanes);
    if (num.stage_kernel != 0 || num.lanes_per_element != 0) {
      g.stage_kernel = num.stage_kernel != 0 ? num.stage_kernel : (num.lanes_per_element > 4 ? 2 : 1);
      g.lanes = num.lanes_per_element != 0 ? num.lanes_per_element : (g.stage_kernel == 2 ? 4 : 4);
    }
    if (g.stage_kernel == 1 && g.lanes != 1 && g.lanes != 2 && g.lanes != 4)
      throw std::runtime_error("engine: the point-split kernel takes 1, 2 or 4 lanes per element");
    if (g.stage_kernel == 2 && g.lanes != 4 && g.lanes != 8)
      throw std::runtime_error("engine: the variable-split kernel takes 4 or 8 lanes per element");
    if (g.stage_kernel != 1 && g.stage_kernel != 2)
      throw std::runtime_error("engine: stage_kernel must be 0, 1 or 2");
    const size_t E = static_cast<size_t>(g.E);

    // element metrics
    std::vector<double> geo(5 * E), mw(E);
    for (size_t e = 0; e < E; ++e) {
      const whip::Triangle &t = mesh.tris[e];
      geo[e] = 0.5 * t.by();
      geo[E + e] = -0.5 * t.bx();
      geo[2 * E + e] = -0.5 * t.ay();
      geo[3 * E + e] = 0.5 * t.ax();
      geo[4 * E + e] = 1.0 / t.jacobian_determinant();
      mw[e] = std::sqrt(2.0) * t.jacobian_determinant();
    }
    // edge records
    if (num.boundary_left == WASL_HIP_PERIODIC)
      g.trace = whip::periodic_trace(mesh, g.p);
    EdgeRecords rec = edge_records(mesh, g.p, g.Np,
                                   num.boundary_left == WASL_HIP_PERIODIC ? &g.trace : nullptr);
    std::vector<int> &nbr = rec.nbr, &meta = rec.meta, &bdon = rec.bdon;
    std::vector<double> &fn = rec.fn, &bbas = rec.bbas;
    g.NB = rec.slots;
    g.periodic_pairs = rec.pairs;
    upload(g.geo, geo, tot, g.stream);
    upload(g.mass_w, mw, tot, g.stream);
    upload(g.nbr, nbr, tot, g.stream);
    upload(g.meta, meta, tot, g.stream);
    upload(g.fn, fn, tot, g.stream);
    upload(g.bdonor, bdon, tot, g.stream);
    upload(g.bbasis, bbas, tot, g.stream);
    upload(g.table, euler_tables(g.p), tot, g.stream);
    if (num.precision != WASL_HIP_FP64)
      upload(g.mxtable, mixed_tables(g.p), tot, g.stream);
    {
      // Reference-element tables of order p into the constant slot of that
      // order; every engine of the same order writes the same values.
      const std::vector<double> vt = euler_vtables(g.p);
      HIPCHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(whip::g_vtable), vt.data(),
                                      vt.size() * sizeof(double),
                                      static_cast<size_t>(g.p) * whip::kVTableMax * sizeof(double),
                                      hipMemcpyHostToDevice, g.stream));
      HIPCHECK(hipStreamSynchronize(g.stream));
    }
    g.hmin = whip::minimum_length(mesh);

    if (num.limiting && g.p > 0) {
      g.build_stencil_host();
      upload(g.lnbr, g.stencil.neighbors, tot, g.stream);
      upload(g.T, g.stencil.transforms, tot, g.stream);
      std::vector<double> nrm(2 * E);
      for (size_t e = 0; e < E; ++e) {
        nrm[2 * e] = g.stencil.nx[e];
        nrm[2 * e + 1] = g.stencil.ny[e];
      }
      upload(g.nrm, nrm, tot, g.stream);
      upload(g.check, g.stencil.check, tot, g.stream);
      g.side.alloc(E * 4 * g.Np * sizeof(double), tot, g.stream);
    } else {
      g.stencil.highest_begin = g.p > 0 ? whip::tri_modes(g.p - 1) : g.Np;
    }
    g.flags.alloc(E, tot, g.stream);
    g.list.alloc(E * sizeof(int), tot, g.stream);
  } else if (num.equation == WASL_HIP_SCALAR_ADVECTION && mesh.dimension == 2) {
    // src/case/scalar_advection_triangle.cpp:15-46: periodic box, constant
    // velocity (wavespeed, velocity_y), no limiter.
    if (num.boundary_left != WASL_HIP_PERIODIC)
      throw std::runtime_error("simulation: 2D scalar-advection boundaries must be periodic");
    if (num.limiting)
      throw std::runtime_error("check: scalar limiting is not supported");
    if (num.precision != WASL_HIP_FP64)
      throw std::runtime_error("engine: 2D scalar advection runs in FP64");
    g.dim = 2;
    g.V = 1;
    g.Np = whip::tri_modes(g.p);
    g.E = static_cast<int>(mesh.tris.size());
    g.F = static_cast<int>(mesh.faces.size());
    g.lanes = 1;
    g.stage_kernel = 1;
    const size_t E = static_cast<size_t>(g.E);
    g.trace = whip::periodic_trace(mesh, g.p);
    const EdgeRecords rec = edge_records(mesh, g.p, g.Np, &g.trace);
    g.NB = rec.slots;
    g.periodic_pairs = rec.pairs;
    const double vx = num.wavespeed, vy = num.velocity_y;
    std::vector<double> vel(2 * E), flux(3 * E), mw(E);
    for (size_t e = 0; e < E; ++e) {
      const whip::Triangle &t = mesh.tris[e];
      const double inv = 1.0 / t.jacobian_determinant();
      const double A = 0.5 * t.by(), B = -0.5 * t.bx(), C = -0.5 * t.ay(), D = 0.5 * t.ax();
      vel[e] = (A * vx + B * vy) * inv;
      vel[E + e] = (C * vx + D * vy) * inv;
      mw[e] = std::sqrt(2.0) * t.jacobian_determinant();
      for (int k = 0; k < 3; ++k) {
        const double nx = rec.fn[(3 * k) * E + e], ny = rec.fn[(3 * k + 1) * E + e];
        flux[k * E + e] = (vx * nx + vy * ny) * rec.fn[(3 * k + 2) * E + e] * inv;
      }
    }
    upload(g.geo, vel, tot, g.stream);
    upload(g.fn, flux, tot, g.stream);
    upload(g.mass_w, mw, tot, g.stream);
    upload(g.nbr, rec.nbr, tot, g.stream);
    upload(g.meta, rec.meta, tot, g.stream);
    upload(g.bdonor, rec.bdon, tot, g.stream);
    upload(g.bbasis, rec.bbas, tot, g.stream);
    const std::vector<double> tables = advect_tables(g.p);
    upload(g.table, tables, tot, g.stream);
    HIPCHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(whip::g_advvol), tables.data(),
                                    2 * g.Np * g.Np * sizeof(double),
                                    static_cast<size_t>(g.p) * whip::kAdvVolMax * sizeof(double),
                                    hipMemcpyHostToDevice, g.stream));
    HIPCHECK(hipStreamSynchronize(g.stream));
    g.hmin = whip::minimum_length(mesh);
    g.stencil.highest_begin = g.Np;
  } else if (num.equation == WASL_HIP_SCALAR_ADVECTION) {
    if (mesh.dimension != 1)
      throw std::runtime_error("engine: scalar advection runs on line or triangle meshes");
    if (num.limiting)
      throw std::runtime_error("check: scalar limiting is not supported");
    const int kinds[2] = {num.boundary_left, num.boundary_right};
    for (int k : kinds)
      if (k < 0 || k > 2)
        throw std::runtime_error("boundaries: invalid boundary kind");
    if ((num.boundary_left == WASL_HIP_PERIODIC) != (num.boundary_right == WASL_HIP_PERIODIC))
      throw std::runtime_error("boundaries: periodic ends must be paired");
    g.dim = 1;
    g.V = 1;
    g.Np = g.p + 1;
    g.E = static_cast<int>(mesh.lines.size());
    g.F = static_cast<int>(mesh.lfaces.size());
    g.lanes = 1;
    std::vector<double> ij(g.E), mw(g.E);
    for (int e = 0; e < g.E; ++e) {
      const double jac = (mesh.lines[e].xr - mesh.lines[e].xl) / 2.0;
      ij[e] = 1.0 / jac;
      mw[e] = std::sqrt(2.0) * jac;
    }
    upload(g.inv_jac, ij, tot, g.stream);
    upload(g.mass_w, mw, tot, g.stream);
    upload(g.table, line_tables(g.p), tot, g.stream);
    g.hmin = whip::minimum_length(mesh);
  } else {
    throw std::runtime_error("engine: unknown equation");
  }
  const size_t n = g.state_size() * sizeof(double);
  g.u.alloc(n, tot, g.stream);
  g.w1.alloc(n, tot, g.stream);
  g.w2.alloc(n, tot, g.stream);
  if (g.mixed()) {
    const size_t tb = static_cast<size_t>(g.E) * 48 * sizeof(float);
    g.tr_u.alloc(tb, tot, g.stream);
    g.tr_w1.alloc(tb, tot, g.stream);
    g.tr_w2.alloc(tb, tot, g.stream);
  }
  HIPCHECK(hipMemsetAsync(g.u.p, 0, n, g.stream));
  g.counters.alloc(4 * sizeof(int), tot, g.stream);
  HIPCHECK(hipMemsetAsync(g.counters.p, 0, g.counters.bytes, g.stream));
  g.err.alloc(2 * sizeof(int), tot, g.stream);
  reset_device_error(g);
  // the step control and, after it, the slots its stages spread the CFL speed over
  g.ctl.alloc(whip::kCtlBytes, tot, g.stream);
  HIPCHECK(hipMemsetAsync(g.ctl.p, 0, g.ctl.bytes, g.stream));
  g.partials.alloc(2 * sizeof(double) * ((g.E + 255) / 256), tot, g.stream);
  g.result.alloc(sizeof(double), tot, g.stream);
  HIPCHECK(hipStreamSynchronize(g.stream));
  if (g.euler() && g.dim == 2 && num.precision == WASL_HIP_FP64 && num.stage_kernel == 0 &&
      num.lanes_per_element == 0)
    autotune(g);
}

// limit_euler_modal on the resident state in place
// (src/limiting/euler_modal_limiter.cpp:70-98); flags (host, may be null)
// receive the sensor decision. Throws the reference's message on a floor
// error.
void limit_resident(wasl_hip_engine &g, uint8_t *flags) {
  if (g.p == 0) {
    if (flags)
      std::fill(flags, flags + g.E, 0);
    return;
  }
  if (!g.lnbr.p)
    throw std::runtime_error("engine: create the engine with limiting on to limit");
  HIPCHECK(hipMemsetAsync(g.counters.as<int>() + 3, 0, sizeof(int), g.stream));
  launch_sensor(g, g.u.as<double>(), g.flags.as<uint8_t>(), engine_buffers(g), g.stream);
  launch_limit(g, g.u.as<double>(), 3, nullptr, 0);
  if (flags)
    HIPCHECK(hipMemcpyAsync(flags, g.flags.p, g.E, hipMemcpyDeviceToHost, g.stream));
  HIPCHECK(hipStreamSynchronize(g.stream));
  raise_device_error(g);
}

// The captured graph of 2^k device-driven steps starting at step parity
// `parity`, captured on first use. Steps change parity one by one, so a
// graph of an even number of steps also ends at `parity`.
hipGraphExec_t step_graph(wasl_hip_engine &g, int k, int parity, whip::RunCtl *dctl) {
  hipGraphExec_t &x = g.graphs[k][parity];
  if (!x) {
    hipGraph_t graph;
    HIPCHECK(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
    int par = parity, unused = 0;
    for (int i = 0; i < (1 << k); ++i) {
      launch_step(g, 0.0, dctl, par, false, unused);
      par ^= 1;
    }
    HIPCHECK(hipStreamEndCapture(g.stream, &graph));
    HIPCHECK(hipGraphInstantiate(&x, graph, nullptr, nullptr, 0));
    HIPCHECK(hipGraphDestroy(graph));
    // the one-time upload to the device happens here, not at the first replay
    HIPCHECK(hipGraphUpload(x, g.stream));
  }
  return x;
}

}  // namespace

// ------------------------------------------------ allocation-free limiter ----
//
// wasl_hip_limit_noalloc works on a caller's device buffer and keeps its
// working memory (error word, troubled counter and list, staged limited
// values) in this file-scope array instead of the engine's buffers, so the
// call allocates nothing and, with device-side status, synchronises nothing.
// Each device holds its own copy for the life of the process. Its device
// address is looked up once per device with hipGetSymbolAddress; host code
// only ever uses that address, never the array's name.
constexpr size_t kScratchBytes = size_t(64) << 20;   // 64 MiB per device
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;   // per device index

void *scratch_on(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0)
    throw std::runtime_error("limit_noalloc: invalid device");
  if (static_cast<size_t>(device) >= g_scratch_ptr.size())
    g_scratch_ptr.resize(device + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    int current = 0;
    HIPCHECK(hipGetDevice(&current));
    if (current != device)
      HIPCHECK(hipSetDevice(device));
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (current != device)
      (void)hipSetDevice(current);
    check(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

// Scratch layout of one call on E elements of Np modes (bytes from the base):
// the error word and counters in the first 256 bytes, then list[E] int32,
// then side[E][4][Np] FP64, each 256-byte aligned.
struct ScratchLayout {
  size_t list, side, total;
};

ScratchLayout scratch_layout(int E, int Np) {
  const auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  ScratchLayout l;
  l.list = 256;
  l.side = up(l.list + sizeof(int) * static_cast<size_t>(E));
  l.total = l.side + sizeof(double) * 4 * static_cast<size_t>(Np) * E;
  return l;
}

// Clears the error word and troubled counter of a call: a kernel rather than
// a memset, so that a caller may capture the call in a graph.
__global__ void noalloc_begin(int *err, int *counters, uint8_t *flags, int E, int clear_flags) {
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i == 0) {
    err[0] = 0;
    err[1] = INT_MAX;
    counters[3] = 0;
  }
  if (clear_flags && flags)
    for (int e = i; e < E; e += static_cast<int>(gridDim.x * blockDim.x))
      flags[e] = 0;
}

// The call's outcome: 0, or WASL_HIP_SIMULATION when a cell average was
// below a floor (the reference's "Euler positivity limiter: cell average is
// below a floor").
__global__ void noalloc_end(const int *err, int *status) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    *status = (err[0] != 0 || err[1] != INT_MAX) ? WASL_HIP_SIMULATION : WASL_HIP_OK;
}

}  // namespace

// =================================================================== C ABI

extern "C" {

const char *wasl_hip_last_error(void) { return g_error.c_str(); }

int wasl_hip_mesh_build(int dimension, int64_t node_count, const int64_t *node_tags,
                        const double *node_xyz, int64_t element_count,
                        const int64_t *element_nodes, wasl_hip_mesh **out) {
  return guarded([&] {
    if (!out || !node_tags || !node_xyz || !element_nodes)
      throw std::runtime_error("mesh: null argument");
    auto m = std::make_unique<wasl_hip_mesh>(whip::build_mesh(
        dimension, node_count, node_tags, node_xyz, element_count, element_nodes));
    *out = m.release();
  });
}

int wasl_hip_mesh_from_prepared(int dimension, int64_t node_count, const int64_t *node_tags,
                                const double *points, int64_t element_count,
                                const int64_t *element_nodes, int64_t face_count,
                                const int64_t *face_nodes, const int32_t *face_elements,
                                const int32_t *face_local_edges, const double *face_normals,
                                wasl_hip_mesh **out) {
  return guarded([&] {
    if (!out)
      throw std::runtime_error("mesh: null argument");
    auto m = std::make_unique<wasl_hip_mesh>(whip::mesh_from_prepared(
        dimension, node_count, node_tags, points, element_count, element_nodes, face_count,
        face_nodes, face_elements, face_local_edges, face_normals));
    *out = m.release();
  });
}

int64_t wasl_hip_mesh_count(const wasl_hip_mesh *mesh, int what) {
  if (!mesh)
    return -1;
  const whip::Mesh &m = mesh->m;
  switch (what) {
  case 0:
    return static_cast<int64_t>(m.elements());
  case 1:
    return static_cast<int64_t>(m.face_count());
  case 2:
    return static_cast<int64_t>(m.boundary_faces());
  case 3: {
    std::vector<int64_t> tags;
    if (m.dimension == 2)
      for (const auto &t : m.tris)
        tags.insert(tags.end(), t.node, t.node + 3);
    else
      for (const auto &l : m.lines) {
        tags.push_back(l.left_node);
        tags.push_back(l.right_node);
      }
    std::sort(tags.begin(), tags.end());
    return static_cast<int64_t>(std::unique(tags.begin(), tags.end()) - tags.begin());
  }
  case 4:
    return m.dimension;
  default:
    return -1;
  }
}

// prepare_mesh (src/mesh/prepared/prepare_mesh.cpp:79-140): nodes numbered in
// order of first use by the elements.
int wasl_hip_mesh_export(const wasl_hip_mesh *mesh, int64_t *node_tags, double *points,
                         int64_t *element_nodes, double *element_metrics, int64_t *face_nodes,
                         int32_t *face_elements, int32_t *face_local_edges, double *face_normals,
                         double *face_measures) {
  return guarded([&] {
    if (!mesh)
      throw std::runtime_error("mesh: null argument");
    const whip::Mesh &m = mesh->m;
    std::unordered_map<int64_t, int64_t> local;
    std::vector<int64_t> tags;
    std::vector<double> pts;
    const auto add = [&](int64_t tag, double x, double y) {
      auto it = local.find(tag);
      if (it != local.end())
        return it->second;
      const int64_t i = static_cast<int64_t>(tags.size());
      local.emplace(tag, i);
      tags.push_back(tag);
      pts.insert(pts.end(), {x, y, 0.0});
      return i;
    };
    if (m.dimension == 2) {
      for (size_t e = 0; e < m.tris.size(); ++e) {
        const whip::Triangle &t = m.tris[e];
        for (int k = 0; k < 3; ++k) {
          const int64_t i = add(t.node[k], t.x[k], t.y[k]);
          if (element_nodes)
            element_nodes[3 * e + k] = i;
        }
        if (element_metrics) {
          const double v[5] = {t.ax() / 2.0, t.bx() / 2.0, t.ay() / 2.0, t.by() / 2.0,
                               t.jacobian_determinant()};
          std::copy(v, v + 5, element_metrics + 5 * e);
        }
      }
      for (size_t f = 0; f < m.faces.size(); ++f) {
        const whip::TriFace &x = m.faces[f];
        if (face_nodes) {
          face_nodes[2 * f] = local.at(x.va);
          face_nodes[2 * f + 1] = local.at(x.vb);
        }
        if (face_elements) {
          face_elements[2 * f] = x.left;
          face_elements[2 * f + 1] = x.right;
        }
        if (face_local_edges) {
          face_local_edges[2 * f] = x.left_edge;
          face_local_edges[2 * f + 1] = x.right_edge;
        }
        if (face_normals) {
          face_normals[2 * f] = x.nx;
          face_normals[2 * f + 1] = x.ny;
        }
        if (face_measures)
          face_measures[f] = m.tris[x.left].edge_length(x.left_edge);
      }
    } else {
      for (size_t e = 0; e < m.lines.size(); ++e) {
        const whip::Segment &s = m.lines[e];
        const int64_t a = add(s.left_node, s.xl, 0.0), b = add(s.right_node, s.xr, 0.0);
        if (element_nodes) {
          element_nodes[2 * e] = a;
          element_nodes[2 * e + 1] = b;
        }
        if (element_metrics)
          element_metrics[e] = (s.xr - s.xl) / 2.0;
      }
      for (size_t f = 0; f < m.lfaces.size(); ++f) {
        const whip::LineFace &x = m.lfaces[f];
        const whip::Segment &s = m.lines[x.left];
        const int64_t node = local.at(x.x == s.xl ? s.left_node : s.right_node);
        if (face_nodes) {
          face_nodes[2 * f] = node;
          face_nodes[2 * f + 1] = node;
        }
        if (face_elements) {
          face_elements[2 * f] = x.left;
          face_elements[2 * f + 1] = x.right;
        }
        if (face_local_edges) {
          face_local_edges[2 * f] = x.normal < 0.0 ? 0 : 1;
          face_local_edges[2 * f + 1] = 0;
        }
        if (face_normals)
          face_normals[f] = x.normal;
        if (face_measures)
          face_measures[f] = 1.0;
      }
    }
    if (node_tags)
      std::copy(tags.begin(), tags.end(), node_tags);
    if (points)
      std::copy(pts.begin(), pts.end(), points);
  });
}

int wasl_hip_mesh_face_vertices(const wasl_hip_mesh *mesh, int64_t *vertices) {
  return guarded([&] {
    if (!mesh || mesh->m.dimension != 2)
      throw std::runtime_error("mesh: face vertices exist for triangle meshes");
    for (size_t f = 0; f < mesh->m.faces.size(); ++f) {
      const whip::TriFace &x = mesh->m.faces[f];
      vertices[4 * f] = x.lv0;
      vertices[4 * f + 1] = x.lv1;
      vertices[4 * f + 2] = x.rv0;
      vertices[4 * f + 3] = x.rv1;
    }
  });
}

void wasl_hip_mesh_destroy(wasl_hip_mesh *mesh) { delete mesh; }

int wasl_hip_mesh_periodic_trace(const wasl_hip_mesh *mesh, int p, int32_t *donors,
                                 double *basis) {
  return guarded([&] {
    if (!mesh || mesh->m.dimension != 2 || p < 0 || p > 3)
      throw std::runtime_error("trace: a triangle mesh and order 0..3 are required");
    const whip::PeriodicTrace t = whip::periodic_trace(mesh->m, p);
    std::copy(t.donors.begin(), t.donors.end(), donors);
    std::copy(t.basis.begin(), t.basis.end(), basis);
  });
}

int wasl_hip_mesh_limiter_stencil(const wasl_hip_mesh *mesh, int p, int periodic,
                                  int32_t *neighbors, double *transforms, double *normals,
                                  double *check_basis) {
  return guarded([&] {
    if (!mesh || mesh->m.dimension != 2 || p < 0 || p > 3)
      throw std::runtime_error("stencil: a triangle mesh and order 0..3 are required");
    std::vector<int32_t> donors;
    if (periodic)
      donors = whip::periodic_trace(mesh->m, p).donors;
    const whip::LimiterStencil s = whip::limiter_stencil(mesh->m, p, donors, p + 1);
    if (neighbors)
      std::copy(s.neighbors.begin(), s.neighbors.end(), neighbors);
    if (transforms)
      std::copy(s.transforms.begin(), s.transforms.end(), transforms);
    if (normals)
      for (size_t e = 0; e < s.nx.size(); ++e) {
        normals[2 * e] = s.nx[e];
        normals[2 * e + 1] = s.ny[e];
      }
    if (check_basis)
      std::copy(s.check.begin(), s.check.end(), check_basis);
  });
}

int wasl_hip_mesh_project(const wasl_hip_mesh *mesh, int p, double gamma, int ic,
                          const double *params, int64_t param_count, double *state) {
  return guarded([&] {
    if (!mesh || !state || p < 0 || p > 3)
      throw std::runtime_error("project: invalid argument");
    const std::vector<double> U = whip::project(mesh->m, p, gamma, ic, params, param_count);
    std::copy(U.begin(), U.end(), state);
  });
}

void wasl_hip_numerics_default(wasl_hip_numerics *n) {
  if (!n)
    return;
  std::memset(n, 0, sizeof(*n));
  n->equation = WASL_HIP_EULER;
  n->polynomial_order = 1;
  n->heat_capacity_ratio = 1.4;
  n->sensor_threshold = 1e-3;
  n->density_floor = 1e-10;
  n->pressure_floor = 1e-10;
  n->precision = WASL_HIP_FP64;
}

int wasl_hip_engine_create(const wasl_hip_mesh *mesh, const wasl_hip_numerics *numerics,
                           wasl_hip_engine **out) {
  return guarded([&] {
    if (!mesh || !numerics || !out)
      throw std::runtime_error("engine: null argument");
    auto g = std::make_unique<wasl_hip_engine>();
    create_engine(mesh->shared, *numerics, *g);
    *out = g.release();
  });
}

void wasl_hip_engine_destroy(wasl_hip_engine *engine) { delete engine; }

int64_t wasl_hip_engine_size(const wasl_hip_engine *g, int what) {
  if (!g)
    return -1;
  switch (what) {
  case 0: return g->E;
  case 1: return g->Np;
  case 2: return g->V;
  case 3: return static_cast<int64_t>(g->state_size());
  case 4: return g->F;
  case 5: return static_cast<int64_t>(g->device_bytes);
  case 6: return g->lanes;
  case 8: return g->stage_kernel;
  case 9: return g->NB;
  case 10: return g->periodic_pairs;
  case 7:
    const_cast<wasl_hip_engine *>(g)->build_stencil_host();
    return g->stencil.checks();
  default: return -1;
  }
}

int wasl_hip_engine_trace(const wasl_hip_engine *g, int32_t *donors, double *basis) {
  return guarded([&] {
    if (!g || g->dim != 2)
      throw std::runtime_error("engine: periodic traces exist for triangle meshes");
    const size_t P = g->p + 1;
    if (g->trace.donors.empty()) {
      std::fill(donors, donors + g->F * P, -1);
      std::fill(basis, basis + g->F * P * g->Np, 0.0);
      return;
    }
    std::copy(g->trace.donors.begin(), g->trace.donors.end(), donors);
    std::copy(g->trace.basis.begin(), g->trace.basis.end(), basis);
  });
}

int wasl_hip_engine_stencil(const wasl_hip_engine *cg, int32_t *neighbors, double *transforms,
                            double *normals, double *check_basis) {
  return guarded([&] {
    if (!cg || !cg->euler())
      throw std::runtime_error("engine: limiter stencils exist for 2D Euler");
    auto *g = const_cast<wasl_hip_engine *>(cg);
    g->build_stencil_host();
    const whip::LimiterStencil &s = g->stencil;
    if (neighbors)
      std::copy(s.neighbors.begin(), s.neighbors.end(), neighbors);
    if (transforms)
      std::copy(s.transforms.begin(), s.transforms.end(), transforms);
    if (normals)
      for (size_t e = 0; e < s.nx.size(); ++e) {
        normals[2 * e] = s.nx[e];
        normals[2 * e + 1] = s.ny[e];
      }
    if (check_basis)
      std::copy(s.check.begin(), s.check.end(), check_basis);
  });
}

int wasl_hip_project(wasl_hip_engine *g, int ic, const double *params, int64_t param_count,
                     double *state_out) {
  return guarded([&] {
    if (!g)
      throw std::runtime_error("engine: null argument");
    const bool fits = g->euler() ? ic < whip::IC_SCALAR_PULSE
                      : g->dim == 1 ? (ic == whip::IC_SCALAR_PULSE || ic == whip::IC_SCALAR_REGIONS)
                                    : ic >= whip::IC_SCALAR_CIRCLES;
    if (!fits)
      throw std::runtime_error("initial state: initial condition " + std::to_string(ic) +
                               " does not belong to this engine's equation");
    const std::vector<double> U =
        whip::project(*g->mesh_ptr, g->p, g->num.heat_capacity_ratio, ic, params, param_count);
    std::vector<double> tmp;
    HIPCHECK(hipMemcpyAsync(g->u.p, g->to_device(U.data(), tmp), U.size() * sizeof(double),
                            hipMemcpyHostToDevice, g->stream));
    HIPCHECK(hipStreamSynchronize(g->stream));
    if (state_out)
      std::copy(U.begin(), U.end(), state_out);
  });
}

int wasl_hip_upload_state(wasl_hip_engine *g, const double *state) {
  return guarded([&] {
    std::vector<double> tmp;
    HIPCHECK(hipMemcpyAsync(g->u.p, g->to_device(state, tmp), g->state_size() * sizeof(double),
                            hipMemcpyHostToDevice, g->stream));
    HIPCHECK(hipStreamSynchronize(g->stream));
  });
}

int wasl_hip_download_state(wasl_hip_engine *g, double *state) {
  return guarded([&] {
    if (!g->mode_major_scalar()) {
      HIPCHECK(hipMemcpyAsync(state, g->u.p, g->state_size() * sizeof(double),
                              hipMemcpyDeviceToHost, g->stream));
      HIPCHECK(hipStreamSynchronize(g->stream));
      return;
    }
    std::vector<double> dev(g->state_size());
    HIPCHECK(hipMemcpyAsync(dev.data(), g->u.p, dev.size() * sizeof(double),
                            hipMemcpyDeviceToHost, g->stream));
    HIPCHECK(hipStreamSynchronize(g->stream));
    g->to_host(dev.data(), state);
  });
}

int wasl_hip_residual(wasl_hip_engine *g, const double *state, double *rhs) {
  return guarded([&] {
    const size_t n = g->state_size() * sizeof(double);
    std::vector<double> tmp;
    HIPCHECK(hipMemcpyAsync(g->w1.p, g->to_device(state, tmp), n, hipMemcpyHostToDevice,
                            g->stream));
    whip::StageArgs a{};
    a.uin = g->w1.as<double>();
    a.un = g->w1.as<double>();
    a.out = g->w2.as<double>();
    a.cb = 1.0;
    a.residual_only = 1;
    a.list = g->list.as<int>();
    a.counters = g->counters.as<int>();
    if (g->mixed()) {
      launch_traces(*g, a.uin, g->tr_w1.as<float>(), g->mxtable.as<float>(), g->stream);
      a.tin = g->tr_w1.as<float>();
    }
    launch_stage(*g, a);
    if (g->mode_major_scalar()) {
      std::vector<double> dev(g->state_size());
      HIPCHECK(hipMemcpyAsync(dev.data(), g->w2.p, n, hipMemcpyDeviceToHost, g->stream));
      HIPCHECK(hipStreamSynchronize(g->stream));
      g->to_host(dev.data(), rhs);
      return;
    }
    HIPCHECK(hipMemcpyAsync(rhs, g->w2.p, n, hipMemcpyDeviceToHost, g->stream));
    HIPCHECK(hipStreamSynchronize(g->stream));
  });
}

int wasl_hip_limit(wasl_hip_engine *g, uint8_t *flags) {
  return guarded([&] {
    if (!g->euler())
      throw std::runtime_error("engine: the limiter applies to 2D Euler");
    limit_resident(*g, flags);
  });
}

int wasl_hip_limit_noalloc(wasl_hip_engine *g, double *state, uint8_t *flags, int *status,
                           void *stream) {
  return guarded([&] {
    if (!g || !state)
      throw std::runtime_error("limit_noalloc: null argument");
    if (!g->euler())
      throw std::runtime_error("engine: the limiter applies to 2D Euler");
    if (g->p > 0 && !g->lnbr.p)
      throw std::runtime_error("engine: create the engine with limiting on to limit");
    const ScratchLayout l = scratch_layout(g->E, g->Np);
    if (l.total > kScratchBytes)
      throw std::runtime_error("limit_noalloc: " + std::to_string(g->E) + " elements of " +
                               std::to_string(g->Np) + " modes need " +
                               std::to_string(l.total) + " bytes of scratch, more than its " +
                               std::to_string(kScratchBytes) + "-byte maximum");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->stream;
    char *base = static_cast<char *>(scratch_on(g->device));
    LimitBuffers b;
    b.err = reinterpret_cast<int *>(base);
    b.counters = reinterpret_cast<int *>(base + 16);
    b.list = reinterpret_cast<int *>(base + l.list);
    b.side = reinterpret_cast<double *>(base + l.side);
    // P0 is returned unchanged (euler_modal_limiter.cpp:73-74) with no flag set
    const int p0 = g->p == 0 ? 1 : 0;
    hipLaunchKernelGGL(noalloc_begin, dim3(p0 ? (g->E + 255) / 256 : 1), dim3(256), 0, s, b.err,
                       b.counters, flags, g->E, p0);
    HIPCHECK(hipGetLastError());
    if (!p0) {
      launch_sensor(*g, state, flags, b, s);
      launch_limit(*g, state, 3, nullptr, 0, b, s);
    }
    if (status) {
      hipLaunchKernelGGL(noalloc_end, dim3(1), dim3(64), 0, s, b.err, status);
      HIPCHECK(hipGetLastError());
    }
  });
}

int wasl_hip_stable_dt(wasl_hip_engine *g, double cfl, double *dt) {
  return guarded([&] {
    const double speed = resident_speed(*g);
    if (!g->euler()) {  // cfl_timestep.cpp:53-68
      *dt = speed == 0.0 ? 0.0 : cfl * g->hmin / (speed * (2.0 * g->p + 1.0));
      return;
    }
    *dt = speed > 0.0 ? cfl * g->hmin / (speed * (2 * g->p + 1)) : 0.0;
  });
}

int wasl_hip_step(wasl_hip_engine *g, double dt) {
  return guarded([&] {
    ensure_events(*g, 3);
    int idx = 0;
    HIPCHECK(hipEventRecord(g->ev[0], g->stream));
    if (g->mixed())
      launch_traces(*g, g->u.as<double>(), g->tr_u.as<float>(), g->mxtable.as<float>(),
                    g->stream);
    launch_step(*g, dt, nullptr, 0, false, idx);
    HIPCHECK(hipStreamSynchronize(g->stream));
    raise_device_error(*g);
  });
}

int wasl_hip_advance(wasl_hip_engine *g, double cfl, double end_time, uint64_t max_steps,
                     uint64_t history_interval, double *time, uint64_t *step, double *history,
                     int64_t history_capacity, int64_t *history_rows) {
  return guarded([&] {
    if (!g || !time || !step)
      throw std::runtime_error("engine: null argument");
    if (history_interval == 0)
      history_interval = ~0ull;
    whip::RunCtl *h = g->ctl_host;
    std::memset(h, 0, sizeof(*h));
    h->end_time = end_time;
    h->cfl_length = cfl * g->hmin;
    h->order_factor = g->dim == 2 ? static_cast<double>(2 * g->p + 1) : (2.0 * g->p + 1.0);
    h->max_steps = max_steps;
    h->fixed_speed = g->euler() ? 0 : 1;
    h->s[0].time = *time;
    h->s[0].step = *step;
    if (!g->euler()) {
      const double a = advection_speed(*g);
      std::memcpy(&h->s[0].lambda, &a, sizeof(double));
    }
    HIPCHECK(hipMemcpyAsync(g->ctl.p, h, sizeof(*h), hipMemcpyHostToDevice, g->stream));
    whip::RunCtl *dctl = g->ctl.as<whip::RunCtl>();
    if (g->euler())
      hipLaunchKernelGGL(whip::cfl_speed, dim3((g->E + 255) / 256), dim3(256), 0, g->stream,
                         g->u.as<double>(), g->E, g->Np, g->num.heat_capacity_ratio, dctl, 0,
                         g->dim);
    if (g->mixed())
      launch_traces(*g, g->u.as<double>(), g->tr_u.as<float>(), g->mxtable.as<float>(),
                    g->stream);
    int64_t rows = 0;
    const auto record = [&](uint64_t st, double t, double dt) {
      if (history && rows < history_capacity) {
        const double m = mass_of(*g);
        history[4 * rows] = static_cast<double>(st);
        history[4 * rows + 1] = t;
        history[4 * rows + 2] = dt;
        history[4 * rows + 3] = m;
      }
      ++rows;
    };
    if (history)
      record(*step, *time, 0.0);

    const bool timed = g->timing;
    constexpr int kBatch = 64;
    if (timed)
      ensure_events(*g, 3 * kBatch);
    int parity = 0;
    uint64_t cur_step = *step;
    double cur_time = *time;
    double last_dt = 0.0;
    double stage_ms = 0.0;
    int64_t launches = 0;
    if (!g->wall[0]) {
      HIPCHECK(hipEventCreate(&g->wall[0]));
      HIPCHECK(hipEventCreate(&g->wall[1]));
    }
    hipEvent_t t0 = g->wall[0], t1 = g->wall[1];
    HIPCHECK(hipEventRecord(t0, g->stream));
    while (true) {
      // Steps beyond max_steps would only be skipped on the device; do not
      // launch them.
      if (cur_step >= max_steps)
        break;
      uint64_t batch = std::min<uint64_t>(kBatch, max_steps - cur_step);
      if (history) {
        const uint64_t to_next = history_interval - (cur_step % history_interval);
        batch = std::min<uint64_t>(batch, to_next);
      }
      int idx = 0;
      if (!timed && g->use_graphs) {
        // the batch's binary digits, largest first: one graph each
        for (int k = wasl_hip_engine::kGraphLevels - 1; k >= 0; --k)
          if (batch & (1ull << k)) {
            HIPCHECK(hipGraphLaunch(step_graph(*g, k, parity, dctl), g->stream));
            parity ^= k == 0 ? 1 : 0;
          }
      } else {
        for (uint64_t i = 0; i < batch; ++i) {
          launch_step(*g, 0.0, dctl, parity, timed, idx);
          parity ^= 1;
        }
      }
      HIPCHECK(hipMemcpyAsync(h, g->ctl.p, sizeof(*h), hipMemcpyDeviceToHost, g->stream));
      HIPCHECK(hipStreamSynchronize(g->stream));
      if (timed) {
        stage_ms += stage_time_ms(*g, idx);
        launches += idx;
      }
      if (h->err)
        break;
      const whip::StepSlot &s = h->s[parity];
      const bool advanced = s.step != cur_step;
      if (advanced)
        last_dt = h->last_dt;
      cur_step = s.step;
      cur_time = s.time;
      const bool done = !(cur_time < end_time && cur_step < max_steps);
      if (history && advanced && (cur_step % history_interval == 0 || done))
        record(cur_step, cur_time, last_dt);
      if (done)
        break;
    }
    HIPCHECK(hipEventRecord(t1, g->stream));
    HIPCHECK(hipEventSynchronize(t1));
    float wall = 0.0f;
    HIPCHECK(hipEventElapsedTime(&wall, t0, t1));
    g->last_wall_ms = wall;
    g->last_stage_ms = stage_ms;
    g->last_stage_launches = launches;
    if (h->err) {
      const int code = h->err;
      HIPCHECK(hipMemsetAsync(g->counters.p, 0, g->counters.bytes, g->stream));
      HIPCHECK(hipStreamSynchronize(g->stream));
      throw SimulationError(device_error_text(code));
    }
    // Inside the loop the kernels report through the step control (h->err,
    // read above); the engine's own error word is only written outside it.
    *time = cur_time;
    *step = cur_step;
    if (history_rows)
      *history_rows = rows;
    // time and step are reported even when the run stops at max_steps
    if (cur_time < end_time)
      throw SimulationError("simulation: max_steps reached before end_time");
  });
}

int wasl_hip_mass(wasl_hip_engine *g, double *mass) {
  return guarded([&] { *mass = mass_of(*g); });
}

int wasl_hip_cell_averages(wasl_hip_engine *g, double *averages) {
  return guarded([&] {
    std::vector<double> U(g->state_size());
    HIPCHECK(hipMemcpyAsync(U.data(), g->u.p, U.size() * sizeof(double), hipMemcpyDeviceToHost,
                            g->stream));
    HIPCHECK(hipStreamSynchronize(g->stream));
    const double s = 1.0 / std::sqrt(2.0);
    const size_t E = g->E;
    if (g->euler()) {
      for (size_t e = 0; e < E; ++e)
        for (int v = 0; v < 4; ++v)
          averages[e * 4 + v] = s * U[(static_cast<size_t>(v) * g->Np) * E + e];
    } else if (g->mode_major_scalar()) {
      for (size_t e = 0; e < E; ++e)
        averages[e] = U[e] / std::sqrt(2.0);
    } else {
      for (size_t e = 0; e < E; ++e)
        averages[e] = U[e * g->Np] / std::sqrt(2.0);
    }
  });
}

int wasl_hip_last_timing(const wasl_hip_engine *g, double *wall_ms, double *stage_ms,
                         int64_t *stage_launches) {
  return guarded([&] {
    if (wall_ms)
      *wall_ms = g->last_wall_ms;
    if (stage_ms)
      *stage_ms = g->last_stage_ms;
    if (stage_launches)
      *stage_launches = g->last_stage_launches;
  });
}

// Not part of the reference interface: the stage kernel's duration. Each
// of the three stages (fixed dt, no step control, no limiter) is launched
// `reps` times back to back from one captured graph and timed with HIP
// events on the engine's stream; stage_us[s] receives the average per
// launch. Stages 1 and 2 repeat exactly; stage 3 overwrites U^n, which it
// also reads, so the resident state is not meaningful afterwards.
int wasl_hip_time_stages(wasl_hip_engine *g, int reps, double dt, double *stage_us) {
  return guarded([&] {
    if (!g || !stage_us || reps < 1 || reps > 4096)
      throw std::runtime_error("engine: time_stages needs an engine, 1..4096 repetitions");
    hipEvent_t t0, t1;
    HIPCHECK(hipEventCreate(&t0));
    HIPCHECK(hipEventCreate(&t1));
    for (int s = 0; s < 3; ++s) {
      const whip::StageArgs a = stage_args(*g, s, dt, nullptr, 0, false);
      if (g->mixed())
        launch_traces(*g, a.uin, const_cast<float *>(a.tin), g->mxtable.as<float>(), g->stream);
      hipGraph_t graph;
      hipGraphExec_t exec;
      HIPCHECK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
      for (int i = 0; i < reps; ++i)
        launch_stage(*g, a);
      HIPCHECK(hipStreamEndCapture(g->stream, &graph));
      HIPCHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      HIPCHECK(hipGraphDestroy(graph));
      HIPCHECK(hipGraphLaunch(exec, g->stream));   // warm
      HIPCHECK(hipEventRecord(t0, g->stream));
      HIPCHECK(hipGraphLaunch(exec, g->stream));
      HIPCHECK(hipEventRecord(t1, g->stream));
      HIPCHECK(hipEventSynchronize(t1));
      float ms = 0.0f;
      HIPCHECK(hipEventElapsedTime(&ms, t0, t1));
      (void)hipGraphExecDestroy(exec);
      stage_us[s] = 1e3 * static_cast<double>(ms) / reps;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
  });
}

// Not part of the reference interface: captures every step graph of
// wasl_hip_advance ahead of time (setup), so that no capture falls inside a
// caller's timed or latency-critical stepping loop.
int wasl_hip_capture_graphs(wasl_hip_engine *g) {
  return guarded([&] {
    if (!g)
      throw std::runtime_error("engine: null argument");
    if (!g->use_graphs)
      return;
    for (int k = 0; k < wasl_hip_engine::kGraphLevels; ++k)
      for (int parity = 0; parity < 2; ++parity)
        step_graph(*g, k, parity, g->ctl.as<whip::RunCtl>());
    // One replay of each graph with a step control whose run is already over
    // (time 0, end time 0): every stage returns at once and the state is not
    // touched, but the first-replay costs are paid here. The next advance
    // writes a fresh step control.
    std::memset(g->ctl_host, 0, sizeof(whip::RunCtl));
    HIPCHECK(hipMemcpyAsync(g->ctl.p, g->ctl_host, sizeof(whip::RunCtl), hipMemcpyHostToDevice,
                            g->stream));
    for (auto &level : g->graphs)
      for (hipGraphExec_t x : level)
        HIPCHECK(hipGraphLaunch(x, g->stream));
    HIPCHECK(hipStreamSynchronize(g->stream));
  });
}

// Not part of the reference interface: the FP64 rates the current device
// sustains on its vector ALU (v_fma_f64) and on its matrix cores
// (v_mfma_f64_16x16x4f64), in TFLOP/s, each the best of three timed launches
// of 4096 blocks of 256 threads (16 waves per SIMD).
int wasl_hip_fp64_ceilings(double *valu_tflops, double *mfma_tflops) {
  return guarded([&] {
    if (!valu_tflops || !mfma_tflops)
      throw std::runtime_error("engine: null argument");
    constexpr int kBlocks = 4096, kValuIters = 2048, kMfmaIters = 256;
    hipStream_t stream = nullptr;
    HIPCHECK(hipStreamCreate(&stream));
    double *out = nullptr;
    HIPCHECK(hipMalloc(&out, sizeof(double) * kBlocks * 256));
    hipEvent_t t0, t1;
    HIPCHECK(hipEventCreate(&t0));
    HIPCHECK(hipEventCreate(&t1));
    const auto best_ms = [&](auto launch) {
      float best = 0.0f;
      launch();   // warm
      for (int r = 0; r < 3; ++r) {
        HIPCHECK(hipEventRecord(t0, stream));
        launch();
        HIPCHECK(hipEventRecord(t1, stream));
        HIPCHECK(hipEventSynchronize(t1));
        float ms = 0.0f;
        HIPCHECK(hipEventElapsedTime(&ms, t0, t1));
        if (r == 0 || ms < best)
          best = ms;
      }
      return static_cast<double>(best);
    };
    const double valu_ms = best_ms([&] {
      hipLaunchKernelGGL(whip::fp64_valu_ceiling, dim3(kBlocks), dim3(256), 0, stream, out,
                         kValuIters);
    });
    const double mfma_ms = best_ms([&] {
      hipLaunchKernelGGL(whip::fp64_mfma_ceiling, dim3(kBlocks), dim3(256), 0, stream, out,
                         kMfmaIters);
    });
    HIPCHECK(hipGetLastError());
    // flops: lanes x iterations x 8 FMAs x 2; waves x iterations x 4 products x 2048
    *valu_tflops = 2.0 * 8.0 * kValuIters * kBlocks * 256.0 / (valu_ms * 1e-3) / 1e12;
    *mfma_tflops = 2048.0 * 4.0 * kMfmaIters * kBlocks * 4.0 / (mfma_ms * 1e-3) / 1e12;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    HIPCHECK(hipFree(out));
    HIPCHECK(hipStreamDestroy(stream));
  });
}

// Not part of the reference interface: switches per-stage HIP event timing
// for wasl_hip_advance (used by the benchmark to time the stage kernel).
int wasl_hip_set_timing(wasl_hip_engine *g, int on) {
  return guarded([&] { g->timing = on != 0; });
}

}  // extern "C"

// The serving layer shares this translation unit (engine internals, kernels).
#include "server.inl"
