// Serving: many resident cases advanced by batched stage kernels
// (include/wasl_hip_serve.h). Part of the engine's translation unit: it is
// included at the end of engine.hip and uses the engine's internals.

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <type_traits>

#include "../../include/wasl_hip_serve.h"

namespace {

enum class Phase { queued, preparing, ready, resident, finishing, done };

struct Request {
  uint64_t id = 0;
  std::shared_ptr<const whip::Mesh> mesh;
  wasl_hip_numerics num{};
  int ic = 0;
  std::vector<double> params;
  double cfl = 0.0, end_time = 0.0;
  uint64_t end_steps = 0, max_steps = 0;
  Phase phase = Phase::queued;
  size_t reserved = 0;                 // device bytes held against the budget
  std::unique_ptr<wasl_hip_engine> eng;
  // outcome
  int status = WASL_HIP_OK;
  std::string message;
  std::vector<double> result;
  double time = 0.0;
  uint64_t step = 0;
};

// A request that finished in a round: its outcome, and where its state sits
// in the server's pinned staging.
struct Finish {
  Request *r;
  int status;
  std::string message;
  size_t offset;
};

// Device bytes of an engine for this case before it exists (state, stage
// buffers, edge records, limiter stencil), for admission.
size_t estimate_bytes(const whip::Mesh &m, const wasl_hip_numerics &n) {
  const size_t E = m.elements(), p = static_cast<size_t>(n.polynomial_order);
  if (m.dimension == 1)
    return E * (3 * (p + 1) * 8 + 8) + (1 << 16);
  const size_t Np = (p + 1) * (p + 2) / 2;
  size_t b = E * (3 * 4 * Np * 8 + 5 * 8 + 3 * 4 * 2 + 9 * 8 + 8);
  if (n.limiting && p > 0)
    b += E * (3 * 4 + 3 * Np * Np * 8 + 2 * 8 + 4 + 1 + 4 * Np * 8);
  return b + (1 << 20);
}

// Resident cases of one batched launch class.
struct ClassKey {
  int dim, p;
  bool fp64;           // 2D Euler: FP64 stage kernels
  int kernel, lanes;   // FP64 2D Euler: the engine's tile (1 point split, 2 variable split)
  bool operator<(const ClassKey &o) const {
    return std::tie(dim, p, fp64, kernel, lanes) <
           std::tie(o.dim, o.p, o.fp64, o.kernel, o.lanes);
  }
  // elements per workgroup of the class's stage kernel
  int block_elements() const {
    if (dim != 2)
      return 256;
    if (!fp64)
      return 64;
    return kernel == 2 ? 64 / (lanes / 4) : 256 / lanes;
  }
};

// One launch class: its resident cases, their launch tables and its own
// stream. Classes hold disjoint cases, so their rounds are independent and
// run concurrently; each round's launches of a class are captured once per
// membership as a hipGraph.
struct ClassRun {
  std::vector<Request *> members;     // admission order
  std::vector<uint64_t> ids;          // their request ids (membership test)
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  Buffer cases, tiles, limit_cases, limit_tiles, scatter_tiles;
  // host images of the tables (alive until the next rebuild of the class)
  std::vector<whip::BatchCase> ec;
  std::vector<whip::BatchLine> lc;
  std::vector<whip::BatchLimit> mc;
  std::vector<int2> et, mt, st;
  int ntiles = 0, nlimit = 0, nscatter = 0;   // live tiles
  int cap = 0, cap_limit = 0, cap_scatter = 0; // launch sizes (tables padded to these)
  hipGraphExec_t graph = nullptr;
  std::vector<const void *> captured;          // launch arguments the graph was captured with

  void release() {
    if (graph)
      (void)hipGraphExecDestroy(graph);
    graph = nullptr;
    if (done)
      (void)hipEventDestroy(done);
    done = nullptr;
    if (stream)
      (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
};

}  // namespace

struct wasl_hip_server {
  wasl_hip_server_config cfg{};
  hipStream_t stream = nullptr;
  size_t budget = 0, reserved = 0;
  uint64_t next_id = 1;
  std::map<uint64_t, std::unique_ptr<Request>> requests;
  std::deque<uint64_t> queue;        // waiting for a preparation slot
  std::vector<uint64_t> resident;    // in admission order
  std::vector<uint64_t> finished;    // done and unclaimed
  bool dirty = true;
  bool in_flight = false;            // a round launched by the previous pump runs
  int64_t graphs = 0, rounds = 0, changes = 0;
  int64_t rebuild_us = 0, wait_us = 0, admit_us = 0, retire_us = 0, idle_us = 0, submit_us = 0;
  int64_t release_us = 0;
  std::atomic<int64_t> prepare_us{0};

  // launch classes of the resident set and the round's step-control gather
  std::map<ClassKey, ClassRun> classes;
  hipEvent_t round_start = nullptr;
  Buffer ctl_ptrs, ctl_out, err_out;
  std::vector<float> mx_host[4];
  Buffer mxtab[4];
  std::vector<whip::RunCtl *> ctl_list;
  std::vector<uint64_t> ctl_ids;          // request of each ctl_list entry
  whip::StepSlot *slots_host = nullptr;   // pinned
  int *errs_host = nullptr;               // pinned
  size_t slots_capacity = 0;
  double *stage_host = nullptr;           // pinned staging of finished states
  size_t stage_capacity = 0;              // doubles

  // Every C ABI call on the server holds api_mu for its duration, so calls
  // from several threads are serialised (a submit from another thread waits
  // for a running pump round). mu guards what the preparation workers share.
  mutable std::mutex api_mu;
  // workers
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_ready;
  std::deque<Request *> work;
  bool stopping = false;

  ~wasl_hip_server() {
    {
      std::lock_guard<std::mutex> lock(mu);
      stopping = true;
    }
    cv_work.notify_all();
    for (auto &t : workers)
      t.join();
    if (stream)
      (void)hipStreamSynchronize(stream);
    for (auto &kv : classes)
      kv.second.release();
    if (round_start)
      (void)hipEventDestroy(round_start);
    if (slots_host)
      (void)hipHostFree(slots_host);
    if (errs_host)
      (void)hipHostFree(errs_host);
    if (stage_host)
      (void)hipHostFree(stage_host);
    if (stream)
      (void)hipStreamDestroy(stream);
  }
};

namespace {

// Worker: build the engine, project the initial condition (and limit it when
// the case limits, as euler_triangle.cpp:37-41 does), fix the end time.
void prepare(Request &r) {
  try {
    auto g = std::make_unique<wasl_hip_engine>();
    create_engine(r.mesh, r.num, *g);
    const std::vector<double> U =
        whip::project(*g->mesh_ptr, g->p, g->num.heat_capacity_ratio, r.ic, r.params.data(),
                      static_cast<int64_t>(r.params.size()));
    HIPCHECK(hipMemcpyAsync(g->u.p, U.data(), U.size() * sizeof(double), hipMemcpyHostToDevice,
                            g->stream));
    if (limiting(*g))
      limit_resident(*g, nullptr);
    if (r.end_steps > 0) {
      const double lambda = resident_speed(*g);
      const double stable = lambda > 0.0
                                ? r.cfl * g->hmin / (lambda * static_cast<double>(2 * g->p + 1))
                                : 0.0;
      if (!(stable > 0.0))
        throw SimulationError("simulation: timestep cannot advance time");
      const double by_steps = static_cast<double>(r.end_steps) * stable;
      r.end_time = r.end_time > 0.0 ? std::min(by_steps, r.end_time) : by_steps;
    }
    HIPCHECK(hipStreamSynchronize(g->stream));
    // A served engine exports nothing: its host copies of the stencil and the
    // traces (the transforms alone reach hundreds of MB) go now, on this
    // worker, rather than when the pump thread retires the case.
    std::vector<double>().swap(g->stencil.transforms);
    std::vector<int32_t>().swap(g->stencil.neighbors);
    std::vector<int32_t>().swap(g->trace.donors);
    std::vector<double>().swap(g->trace.basis);
    r.eng = std::move(g);
  } catch (const SimulationError &ex) {
    r.status = WASL_HIP_SIMULATION;
    r.message = ex.what();
  } catch (const DeviceError &ex) {
    r.status = WASL_HIP_DEVICE;
    r.message = ex.what();
  } catch (const std::exception &ex) {
    r.status = WASL_HIP_ERROR;
    r.message = ex.what();
  }
}

void worker_loop(wasl_hip_server *s) {
  for (;;) {
    Request *r = nullptr;
    {
      std::unique_lock<std::mutex> lock(s->mu);
      s->cv_work.wait(lock, [&] { return s->stopping || !s->work.empty(); });
      if (s->stopping)
        return;
      r = s->work.front();
      s->work.pop_front();
    }
    const auto t0 = std::chrono::steady_clock::now();
    prepare(*r);
    s->prepare_us += std::chrono::duration_cast<std::chrono::microseconds>(
                         std::chrono::steady_clock::now() - t0)
                         .count();
    {
      std::lock_guard<std::mutex> lock(s->mu);
      r->phase = r->status == WASL_HIP_OK ? Phase::ready : Phase::done;
    }
    s->cv_ready.notify_all();
  }
}

// Step control of a newly admitted case, at the server's step parity 0
// (every round is an even number of steps).
void admit_ctl(wasl_hip_server &s, Request &r) {
  wasl_hip_engine &g = *r.eng;
  whip::RunCtl h{};
  h.end_time = r.end_time;
  h.cfl_length = r.cfl * g.hmin;
  h.order_factor = static_cast<double>(2 * g.p + 1);
  h.max_steps = r.max_steps;
  h.fixed_speed = g.dim == 1 ? 1 : 0;
  if (g.dim == 1) {
    const double a = std::fabs(g.num.wavespeed);
    std::memcpy(&h.s[0].lambda, &a, sizeof(double));
  }
  HIPCHECK(hipMemcpyAsync(g.ctl.p, &h, sizeof h, hipMemcpyHostToDevice, s.stream));
  if (g.dim == 2)
    hipLaunchKernelGGL(whip::cfl_speed, dim3((g.E + 255) / 256), dim3(256), 0, s.stream,
                       g.u.as<double>(), g.E, g.Np, g.num.heat_capacity_ratio,
                       g.ctl.as<whip::RunCtl>(), 0, 2);
  if (g.mixed())
    launch_traces(g, g.u.as<double>(), g.tr_u.as<float>(), s.mxtab[g.p].as<float>(), s.stream);
  HIPCHECK(hipGetLastError());
  // the copy above reads the host struct: complete it before it goes out of scope
  HIPCHECK(hipStreamSynchronize(s.stream));
}

// Host table to a device buffer, grown with headroom (its address changes
// only when it grows).
template <typename T> void put_table(Buffer &b, const std::vector<T> &v, hipStream_t st) {
  const size_t n = v.size() * sizeof(T);
  if (b.bytes < n) {
    size_t unused = 0;
    b.alloc(2 * n, unused);
  }
  if (n)
    HIPCHECK(hipMemcpyAsync(b.p, v.data(), n, hipMemcpyHostToDevice, st));
}

// Launch tables of one class from its members (its stream is idle here).
void rebuild_class(wasl_hip_server &s, const ClassKey &key, ClassRun &c) {
  c.ec.clear();
  c.lc.clear();
  c.mc.clear();
  c.et.clear();
  c.mt.clear();
  c.st.clear();
  if (key.dim == 2) {
    for (Request *r : c.members) {
      wasl_hip_engine &g = *r->eng;
      whip::BatchCase b{};
      b.mesh = g.euler_mesh();
      b.u = g.u.as<double>();
      b.w1 = g.w1.as<double>();
      b.w2 = g.w2.as<double>();
      b.tu = g.tr_u.as<float>();
      b.tw1 = g.tr_w1.as<float>();
      b.tw2 = g.tr_w2.as<float>();
      b.ctl = g.ctl.as<whip::RunCtl>();
      const bool lim = limiting(g);
      b.sensor = lim ? 1 : 0;
      b.threshold = g.num.sensor_threshold;
      b.highest_begin = g.stencil.highest_begin;
      b.list = g.list.as<int>();
      b.counters = g.counters.as<int>();
      const int idx = static_cast<int>(c.ec.size());
      c.ec.push_back(b);
      const int per = key.block_elements();
      for (int k = 0; k < (g.E + per - 1) / per; ++k)
        c.et.push_back(make_int2(idx, k));
      if (!lim)
        continue;
      whip::BatchLimit m{};
      m.L = g.limiter();
      m.u = b.u;
      m.w1 = b.w1;
      m.w2 = b.w2;
      m.tu = b.tu;
      m.tw1 = b.tw1;
      m.tw2 = b.tw2;
      m.side = g.side.as<double>();
      m.list = b.list;
      m.counters = b.counters;
      m.ctl = b.ctl;
      m.err = g.err.as<int>();
      const int midx = static_cast<int>(c.mc.size());
      c.mc.push_back(m);
      for (int k = 0; k < whip::limit_blocks(g.E); ++k)
        c.mt.push_back(make_int2(midx, k));
      for (int k = 0; k < (g.E + 63) / 64; ++k)
        c.st.push_back(make_int2(midx, k));
    }
  } else {
    for (Request *r : c.members) {
      wasl_hip_engine &g = *r->eng;
      whip::BatchLine b{};
      b.mesh = g.line_mesh();
      b.u = g.u.as<double>();
      b.w1 = g.w1.as<double>();
      b.w2 = g.w2.as<double>();
      b.ctl = g.ctl.as<whip::RunCtl>();
      const int idx = static_cast<int>(c.lc.size());
      c.lc.push_back(b);
      for (int k = 0; k < (g.E + 255) / 256; ++k)
        c.et.push_back(make_int2(idx, k));
    }
  }
  c.ntiles = static_cast<int>(c.et.size());
  c.nlimit = static_cast<int>(c.mt.size());
  c.nscatter = static_cast<int>(c.st.size());
  // launch sizes grow by doubling and never shrink: padding tiles exit at
  // once, and a graph stays valid while its launch sizes and tables stay
  const auto grow = [](int cap, int n) {
    while (cap < n)
      cap = cap ? 2 * cap : 64;
    return cap;
  };
  c.cap = grow(c.cap, c.ntiles);
  c.cap_limit = c.nlimit ? grow(c.cap_limit, c.nlimit) : 0;
  c.cap_scatter = c.nscatter ? grow(c.cap_scatter, c.nscatter) : 0;
  {
    // stage kernels give each XCD a contiguous range of workgroups: spread
    // the live tiles evenly (in order) over the padded table so every XCD
    // gets its share
    std::vector<int2> spread(c.cap, make_int2(-1, 0));
    const int64_t n = c.ntiles, cap = c.cap;
    for (int64_t j = 0; j < n; ++j)
      spread[static_cast<size_t>(j * cap / n)] = c.et[static_cast<size_t>(j)];
    c.et.swap(spread);
  }
  c.mt.resize(c.cap_limit, make_int2(-1, 0));
  c.st.resize(c.cap_scatter, make_int2(-1, 0));
  if (key.dim == 2)
    put_table(c.cases, c.ec, c.stream);
  else
    put_table(c.cases, c.lc, c.stream);
  put_table(c.tiles, c.et, c.stream);
  put_table(c.limit_cases, c.mc, c.stream);
  put_table(c.limit_tiles, c.mt, c.stream);
  put_table(c.scatter_tiles, c.st, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  const std::vector<const void *> args = {
      c.cases.p, c.tiles.p, c.limit_cases.p, c.limit_tiles.p, c.scatter_tiles.p,
      reinterpret_cast<const void *>(static_cast<intptr_t>(c.cap)),
      reinterpret_cast<const void *>(static_cast<intptr_t>(c.cap_limit)),
      reinterpret_cast<const void *>(static_cast<intptr_t>(c.cap_scatter))};
  if (c.graph && args != c.captured) {
    HIPCHECK(hipGraphExecDestroy(c.graph));
    c.graph = nullptr;
  }
  c.captured = args;
}

// Class membership and tables of the resident set; classes whose members
// changed are rebuilt and lose their graph.
void rebuild(wasl_hip_server &s) {
  std::map<ClassKey, std::vector<Request *>> cls;
  for (uint64_t id : s.resident) {
    Request *r = s.requests.at(id).get();
    const wasl_hip_engine &g = *r->eng;
    const bool fp64 = !g.mixed();
    const bool tiled = g.dim == 2 && fp64;
    cls[{g.dim, g.p, fp64, tiled ? g.stage_kernel : 0, tiled ? g.lanes : 0}].push_back(r);
  }
  for (auto &[key, c] : s.classes)
    if (!cls.count(key))
      cls[key];   // emptied class
  s.ctl_list.clear();
  s.ctl_ids.clear();
  for (auto &[key, list] : cls) {
    ClassRun &c = s.classes[key];
    if (!c.stream) {
      HIPCHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
      HIPCHECK(hipEventCreateWithFlags(&c.done, hipEventDisableTiming));
    }
    std::vector<uint64_t> ids;
    for (Request *r : list)
      ids.push_back(r->id);
    if (ids != c.ids) {
      c.members = list;
      c.ids = ids;
      rebuild_class(s, key, c);
    }
    for (Request *r : c.members) {
      s.ctl_list.push_back(r->eng->ctl.as<whip::RunCtl>());
      s.ctl_ids.push_back(r->id);
    }
  }
  put_table(s.ctl_ptrs, s.ctl_list, s.stream);
  const size_t n = s.ctl_list.size();
  if (s.slots_capacity < n) {
    if (s.slots_host)
      HIPCHECK(hipHostFree(s.slots_host));
    if (s.errs_host)
      HIPCHECK(hipHostFree(s.errs_host));
    s.slots_host = nullptr;
    s.errs_host = nullptr;
    s.slots_capacity = std::max<size_t>(2 * n, 64);
    HIPCHECK(hipHostMalloc(reinterpret_cast<void **>(&s.slots_host),
                           s.slots_capacity * sizeof(whip::StepSlot)));
    HIPCHECK(hipHostMalloc(reinterpret_cast<void **>(&s.errs_host),
                           s.slots_capacity * sizeof(int)));
    size_t unused = 0;
    s.ctl_out.alloc(s.slots_capacity * sizeof(whip::StepSlot), unused);
    s.err_out.alloc(s.slots_capacity * sizeof(int), unused);
  }
  HIPCHECK(hipStreamSynchronize(s.stream));
}

template <int P>
void launch_euler_class(wasl_hip_server &s, const ClassKey &key, ClassRun &c, whip::BatchStage b) {
  const whip::BatchCase *cases = c.cases.as<whip::BatchCase>();
  const int2 *tiles = c.tiles.as<int2>();
  const float *t = s.mxtab[P].as<float>();
  // FP64: the engine's own tile, so a case computes what the engine computes
  if (key.fp64) {
    const dim3 grid(c.cap), block(256);
    if (key.kernel == 2 && key.lanes == 8)
      hipLaunchKernelGGL((whip::euler_stage_vs_batch<P, 2>), grid, block, 0, c.stream, cases,
                         tiles, b);
    else if (key.kernel == 2)
      hipLaunchKernelGGL((whip::euler_stage_vs_batch<P, 1>), grid, block, 0, c.stream, cases,
                         tiles, b);
    else if (key.lanes == 1)
      hipLaunchKernelGGL((whip::euler_stage_batch<P, 1>), grid, block, 0, c.stream, cases, tiles,
                         b);
    else if (key.lanes == 2)
      hipLaunchKernelGGL((whip::euler_stage_batch<P, 2>), grid, block, 0, c.stream, cases, tiles,
                         b);
    else
      hipLaunchKernelGGL((whip::euler_stage_batch<P, 4>), grid, block, 0, c.stream, cases, tiles,
                         b);
  } else if (s.cfg.precision == WASL_HIP_BF16)
    hipLaunchKernelGGL((whip::euler_stage_mx_batch<P, true>), dim3(c.cap), dim3(256), 0,
                       c.stream, cases, tiles, b, t);
  else
    hipLaunchKernelGGL((whip::euler_stage_mx_batch<P, false>), dim3(c.cap), dim3(256), 0,
                       c.stream, cases, tiles, b, t);
}

template <int P> void launch_line_class(ClassRun &c, whip::BatchStage b) {
  hipLaunchKernelGGL((whip::line_stage_batch<P>), dim3(c.cap), dim3(256), 0, c.stream,
                     c.cases.as<whip::BatchLine>(), c.tiles.as<int2>(), b);
}

template <int P> void launch_limit_class(wasl_hip_server &s, ClassRun &c, whip::BatchStage b) {
  const whip::BatchLimit *cases = c.limit_cases.as<whip::BatchLimit>();
  hipLaunchKernelGGL((whip::limit_troubled_batch<P>), dim3(c.cap_limit), dim3(64), 0, c.stream,
                     cases, c.limit_tiles.as<int2>(), b);
  hipLaunchKernelGGL((whip::scatter_limited_batch<P>), dim3(c.cap_scatter), dim3(64), 0,
                     c.stream,
                     cases, c.scatter_tiles.as<int2>(), b, s.mxtab[P].as<float>());
}

template <typename F> void by_order(int p, F &&f) {
  switch (p) {
  case 0: f(std::integral_constant<int, 0>{}); break;
  case 1: f(std::integral_constant<int, 1>{}); break;
  case 2: f(std::integral_constant<int, 2>{}); break;
  default: f(std::integral_constant<int, 3>{}); break;
  }
}

// One round of a class on its stream: steps_per_batch steps of three stages;
// per stage the class's stage launch, then its limiter launches.
void launch_class_round(wasl_hip_server &s, const ClassKey &key, ClassRun &c) {
  int parity = 0;
  for (int k = 0; k < s.cfg.steps_per_batch; ++k) {
    for (int stage = 0; stage < 3; ++stage) {
      const whip::BatchStage b{stage, parity};
      by_order(key.p, [&](auto P) {
        constexpr int kP = decltype(P)::value;
        if (key.dim == 2) {
          launch_euler_class<kP>(s, key, c, b);
          if constexpr (kP > 0)
            if (c.cap_limit > 0)
              launch_limit_class<kP>(s, c, b);
        } else {
          launch_line_class<kP>(c, b);
        }
      });
    }
    parity ^= 1;
  }
  HIPCHECK(hipGetLastError());
}

// A round of every class, each on its own stream after the work queued so
// far on the server stream, which then waits for all of them.
void launch_round(wasl_hip_server &s) {
  HIPCHECK(hipEventRecord(s.round_start, s.stream));
  for (auto &[key, c] : s.classes) {
    if (c.members.empty())
      continue;
    HIPCHECK(hipStreamWaitEvent(c.stream, s.round_start, 0));
    if (s.cfg.use_graphs) {
      if (!c.graph) {
        hipGraph_t gr;
        HIPCHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        launch_class_round(s, key, c);
        HIPCHECK(hipStreamEndCapture(c.stream, &gr));
        HIPCHECK(hipGraphInstantiate(&c.graph, gr, nullptr, nullptr, 0));
        HIPCHECK(hipGraphDestroy(gr));
        ++s.graphs;
      }
      HIPCHECK(hipGraphLaunch(c.graph, c.stream));
    } else {
      launch_class_round(s, key, c);
    }
    HIPCHECK(hipEventRecord(c.done, c.stream));
    HIPCHECK(hipStreamWaitEvent(s.stream, c.done, 0));
  }
}

__global__ void gather_slots(whip::RunCtl *const *ctls, int n, whip::StepSlot *out, int *errs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    out[i] = ctls[i]->s[0];
    errs[i] = ctls[i]->err;
  }
}

// Finishes a request whose final state is already on the host (`state`,
// the engine's state_size() doubles): keeps it and releases the engine.
void retire(wasl_hip_server &s, Request &r, int status, const std::string &message,
            const double *state) {
  wasl_hip_engine &g = *r.eng;
  r.result.assign(state, state + g.state_size());
  r.status = status;
  r.message = message;
  const auto t_release = std::chrono::steady_clock::now();
  r.eng.reset();
  s.release_us += std::chrono::duration_cast<std::chrono::microseconds>(
                      std::chrono::steady_clock::now() - t_release)
                      .count();
  s.reserved -= r.reserved;
  r.reserved = 0;
  r.phase = Phase::done;
}

}  // namespace

extern "C" {

void wasl_hip_server_config_default(wasl_hip_server_config *c) {
  if (!c)
    return;
  *c = wasl_hip_server_config{};
  c->precision = WASL_HIP_MIXED;
  c->max_resident = 256;
  c->memory_budget = 0;
  c->steps_per_batch = 16;
  c->use_graphs = 1;
  c->setup_threads = 4;
  c->fp64_max_order = 1;
  c->fp64_limited = 1;
  // fixed tiles beat each case's autotuned one on the serving mix (P1 point
  // split 2 lanes, P2 point split 2 lanes, P3 variable split 4 lanes); at P1
  // two lanes (166 registers, three waves per SIMD) beat one (240, two) on
  // the 1024-request mix by about 2 %
  const int32_t tiles[4][2] = {{0, 0}, {1, 2}, {1, 2}, {2, 4}};
  for (int p = 0; p < 4; ++p) {
    c->fp64_tile[p][0] = tiles[p][0];
    c->fp64_tile[p][1] = tiles[p][1];
  }
}

int wasl_hip_server_create(const wasl_hip_server_config *config, wasl_hip_server **out) {
  return guarded([&] {
    if (!out)
      throw std::runtime_error("server: null argument");
    auto s = std::make_unique<wasl_hip_server>();
    if (config)
      s->cfg = *config;
    else
      wasl_hip_server_config_default(&s->cfg);
    wasl_hip_server_config &c = s->cfg;
    if (c.precision != WASL_HIP_MIXED && c.precision != WASL_HIP_BF16)
      throw std::runtime_error("server: precision must be MIXED or BF16");
    if (c.steps_per_batch <= 0 || c.steps_per_batch % 2 != 0)
      throw std::runtime_error("server: steps_per_batch must be even and positive");
    if (c.max_resident <= 0)
      throw std::runtime_error("server: max_resident must be positive");
    c.setup_threads = std::max(1, std::min(16, c.setup_threads));
    HIPCHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    HIPCHECK(hipEventCreateWithFlags(&s->round_start, hipEventDisableTiming));
    size_t free_b = 0, total_b = 0;
    HIPCHECK(hipMemGetInfo(&free_b, &total_b));
    s->budget = c.memory_budget ? static_cast<size_t>(c.memory_budget)
                                : static_cast<size_t>(0.8 * static_cast<double>(free_b));
    for (int p = 0; p <= 3; ++p) {
      s->mx_host[p] = mixed_tables(p);
      size_t unused = 0;
      s->mxtab[p].alloc(s->mx_host[p].size() * sizeof(float), unused);
      HIPCHECK(hipMemcpyAsync(s->mxtab[p].p, s->mx_host[p].data(),
                              s->mx_host[p].size() * sizeof(float), hipMemcpyHostToDevice,
                              s->stream));
    }
    HIPCHECK(hipStreamSynchronize(s->stream));
    for (int i = 0; i < c.setup_threads; ++i)
      s->workers.emplace_back(worker_loop, s.get());
    *out = s.release();
  });
}

void wasl_hip_server_destroy(wasl_hip_server *s) { delete s; }

int wasl_hip_server_submit(wasl_hip_server *s, const wasl_hip_mesh *mesh,
                           const wasl_hip_numerics *numerics, int ic, const double *params,
                           int64_t param_count, double cfl, double end_time, uint64_t end_steps,
                           uint64_t max_steps, uint64_t *id) {
  return guarded([&] {
    if (!s || !mesh || !numerics || !id || (param_count > 0 && !params))
      throw std::runtime_error("server: null argument");
    std::lock_guard<std::mutex> api(s->api_mu);
    const auto t_submit = std::chrono::steady_clock::now();
    if (!(cfl > 0.0) || (end_steps == 0 && !(end_time > 0.0)) || max_steps == 0)
      throw std::runtime_error("server: invalid run control");
    const bool euler = numerics->equation == WASL_HIP_EULER;
    if (euler != (mesh->m.dimension == 2))
      throw std::runtime_error(
          "server: serving takes 2D Euler and 1D scalar advection cases; run 1D Euler and 2D "
          "advection on their own engines");
    auto r = std::make_unique<Request>();
    r->mesh = mesh->shared;
    r->num = *numerics;
    const bool fp64 = r->mesh->dimension != 2 ||
                      r->num.polynomial_order <= s->cfg.fp64_max_order ||
                      (s->cfg.fp64_limited && r->num.limiting);
    r->num.precision = fp64 ? WASL_HIP_FP64 : s->cfg.precision;
    const int order = r->num.polynomial_order;
    // Limited cases keep the autotuned tile: troubled-cell decisions near
    // vacuum follow rounding, and the single engine's tile reproduces them.
    if (fp64 && euler && !r->num.limiting && r->num.stage_kernel == 0 && order >= 0 &&
        order <= 3 && s->cfg.fp64_tile[order][0] != 0) {
      r->num.stage_kernel = s->cfg.fp64_tile[order][0];
      r->num.lanes_per_element = s->cfg.fp64_tile[order][1];
    }
    r->ic = ic;
    r->params.assign(params, params + std::max<int64_t>(param_count, 0));
    r->cfl = cfl;
    r->end_time = end_time;
    r->end_steps = end_steps;
    r->max_steps = max_steps;
    std::lock_guard<std::mutex> lock(s->mu);
    r->id = s->next_id++;
    *id = r->id;
    s->queue.push_back(r->id);
    s->requests[r->id] = std::move(r);
    s->submit_us += std::chrono::duration_cast<std::chrono::microseconds>(
                        std::chrono::steady_clock::now() - t_submit)
                        .count();
  });
}

int wasl_hip_server_pump(wasl_hip_server *s, uint64_t *completed, int64_t capacity,
                         int64_t *count) {
  return guarded([&] {
    if (!s)
      throw std::runtime_error("server: null argument");
    std::lock_guard<std::mutex> api(s->api_mu);
    using clock = std::chrono::steady_clock;
    const auto us = [](clock::time_point a, clock::time_point b) {
      return std::chrono::duration_cast<std::chrono::microseconds>(b - a).count();
    };
    std::vector<uint64_t> done_now, ready_now;
    // Requests that finished in the round just completed. They leave the
    // resident set before the next round is launched and are retired (state
    // copied out, engine released) while that round runs.
    std::vector<Finish> finishing;
    if (s->in_flight) {
      const auto t_wait = clock::now();
      HIPCHECK(hipStreamSynchronize(s->stream));
      s->wait_us += us(t_wait, clock::now());
      s->in_flight = false;
      const int n = static_cast<int>(s->ctl_ids.size());
      for (int i = 0; i < n; ++i) {
        Request &r = *s->requests.at(s->ctl_ids[i]);
        const whip::StepSlot &st = s->slots_host[i];
        r.time = st.time;
        r.step = st.step;
        if (s->errs_host[i]) {
          finishing.push_back({&r, WASL_HIP_SIMULATION, device_error_text(s->errs_host[i]), 0});
        } else if (!(st.time < r.end_time) || st.step >= r.max_steps) {
          const bool short_of_end = st.time < r.end_time;
          finishing.push_back({&r, short_of_end ? WASL_HIP_SIMULATION : WASL_HIP_OK,
                               short_of_end ? "simulation: max_steps reached before end_time"
                                            : "",
                               0});
        }
      }
      if (!finishing.empty()) {
        for (Finish &f : finishing)
          f.r->phase = Phase::finishing;
        // keep admission order among the remaining cases
        std::vector<uint64_t> keep;
        for (uint64_t id : s->resident)
          if (s->requests.at(id)->phase == Phase::resident)
            keep.push_back(id);
        s->resident.swap(keep);
        s->dirty = true;
      }
    }
    {
      std::unique_lock<std::mutex> lock(s->mu);
      // start preparations in arrival order while the budget and slots allow
      size_t active = s->resident.size() + finishing.size();
      for (auto &[id, r] : s->requests)
        if (r->phase == Phase::preparing || r->phase == Phase::ready)
          ++active;
      while (!s->queue.empty() && active < static_cast<size_t>(s->cfg.max_resident)) {
        Request &r = *s->requests.at(s->queue.front());
        const size_t need = estimate_bytes(*r.mesh, r.num);
        if (s->reserved + need > s->budget && s->reserved > 0)
          break;
        s->queue.pop_front();
        r.reserved = need;
        s->reserved += need;
        r.phase = Phase::preparing;
        s->work.push_back(&r);
        ++active;
      }
      s->cv_work.notify_all();
      if (s->resident.empty() && finishing.empty()) {
        const auto has_ready = [&] {
          for (auto &[id, r] : s->requests)
            if (r->phase == Phase::ready || (r->phase == Phase::done && r->eng == nullptr &&
                                              r->status != WASL_HIP_OK && r->reserved))
              return true;
          return false;
        };
        const auto t_idle = clock::now();
        s->cv_ready.wait_for(lock, std::chrono::milliseconds(50), has_ready);
        s->idle_us += us(t_idle, clock::now());
      }
      // failed preparations finish here; ready ones are admitted
      for (auto &[id, r] : s->requests) {
        if (r->phase == Phase::done && r->reserved && !r->eng) {
          s->reserved -= r->reserved;
          r->reserved = 0;
          done_now.push_back(id);
        } else if (r->phase == Phase::ready) {
          r->phase = Phase::resident;
          ready_now.push_back(id);
        }
      }
    }
    const auto t_admit = clock::now();
    for (uint64_t id : ready_now) {
      Request *r = s->requests.at(id).get();
      // replace the estimate by what the engine holds
      s->reserved = s->reserved - r->reserved + r->eng->device_bytes;
      r->reserved = r->eng->device_bytes;
      admit_ctl(*s, *r);
      s->resident.push_back(id);
      s->dirty = true;
    }
    s->admit_us += us(t_admit, clock::now());
    if (!s->resident.empty()) {
      const auto t_build = clock::now();
      if (s->dirty) {
        rebuild(*s);
        ++s->changes;
        s->dirty = false;
      }
      launch_round(*s);
      s->rebuild_us += us(t_build, clock::now());
      ++s->rounds;
      // gather the step slots (parity 0 after an even round) and error codes,
      // in the order of the launch tables; read at the next pump
      const int n = static_cast<int>(s->ctl_list.size());
      hipLaunchKernelGGL(gather_slots, dim3((n + 63) / 64), dim3(64), 0, s->stream,
                         s->ctl_ptrs.as<whip::RunCtl *>(), n, s->ctl_out.as<whip::StepSlot>(),
                         s->err_out.as<int>());
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipMemcpyAsync(s->slots_host, s->ctl_out.p, n * sizeof(whip::StepSlot),
                              hipMemcpyDeviceToHost, s->stream));
      HIPCHECK(hipMemcpyAsync(s->errs_host, s->err_out.p, n * sizeof(int), hipMemcpyDeviceToHost,
                              s->stream));
      s->in_flight = true;
    }
    // The finished requests are in no launch table of the running round.
    // Their states come out through pinned staging on each engine's own
    // stream (non-blocking and otherwise idle), beside the round. (Leaving
    // the copies in flight until the next call measured no faster.)
    if (!finishing.empty()) {
      const auto t_retire = clock::now();
      size_t total = 0;
      for (const Finish &f : finishing)
        total += f.r->eng->state_size();
      if (s->stage_capacity < total) {
        if (s->stage_host)
          HIPCHECK(hipHostFree(s->stage_host));
        s->stage_host = nullptr;
        s->stage_capacity = std::max(total, 2 * s->stage_capacity);
        HIPCHECK(hipHostMalloc(reinterpret_cast<void **>(&s->stage_host),
                               s->stage_capacity * sizeof(double)));
      }
      size_t off = 0;
      for (Finish &f : finishing) {
        const wasl_hip_engine &g = *f.r->eng;
        HIPCHECK(hipMemcpyAsync(s->stage_host + off, g.u.p, g.state_size() * sizeof(double),
                                hipMemcpyDeviceToHost, g.stream));
        f.offset = off;
        off += g.state_size();
      }
      for (Finish &f : finishing) {
        HIPCHECK(hipStreamSynchronize(f.r->eng->stream));
        retire(*s, *f.r, f.status, f.message, s->stage_host + f.offset);
        done_now.push_back(f.r->id);
      }
      s->retire_us += us(t_retire, clock::now());
    }
    for (uint64_t id : done_now)
      s->finished.push_back(id);
    int64_t k = 0;
    while (k < capacity && !s->finished.empty()) {
      if (completed)
        completed[k] = s->finished.front();
      s->finished.erase(s->finished.begin());
      ++k;
    }
    if (count)
      *count = k;
  });
}

int wasl_hip_server_result(wasl_hip_server *s, uint64_t id, double *state, int64_t capacity,
                           double *time, uint64_t *step, int *case_status) {
  return guarded([&] {
    if (!s)
      throw std::runtime_error("server: null argument");
    std::lock_guard<std::mutex> api(s->api_mu);
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->requests.find(id);
    if (it == s->requests.end() || it->second->phase != Phase::done)
      throw std::runtime_error("server: request is not finished");
    Request &r = *it->second;
    if (state) {
      if (capacity < static_cast<int64_t>(r.result.size()))
        throw std::runtime_error("server: state buffer too small");
      std::copy(r.result.begin(), r.result.end(), state);
    }
    if (time)
      *time = r.time;
    if (step)
      *step = r.step;
    if (case_status)
      *case_status = r.status;
    if (r.status != WASL_HIP_OK)
      g_error = r.message;
    s->requests.erase(it);
  });
}

int64_t wasl_hip_server_state_size(const wasl_hip_server *s, uint64_t id) {
  if (!s)
    return -1;
  std::lock_guard<std::mutex> api(s->api_mu);
  auto it = s->requests.find(id);
  if (it == s->requests.end())
    return -1;
  const Request &r = *it->second;
  if (!r.result.empty())
    return static_cast<int64_t>(r.result.size());
  if (r.eng)
    return static_cast<int64_t>(r.eng->state_size());
  return -1;
}

int64_t wasl_hip_server_count(const wasl_hip_server *s, int what) {
  if (!s)
    return -1;
  std::lock_guard<std::mutex> api(s->api_mu);
  std::lock_guard<std::mutex> lock(const_cast<wasl_hip_server *>(s)->mu);
  int64_t n = 0;
  switch (what) {
  case 0:
    return static_cast<int64_t>(s->queue.size());
  case 1:
    for (auto &[id, r] : s->requests)
      if (r->phase == Phase::preparing || r->phase == Phase::ready)
        ++n;
    return n;
  case 2:
    return static_cast<int64_t>(s->resident.size());
  case 3:
    return static_cast<int64_t>(s->finished.size());
  case 4:
    return static_cast<int64_t>(s->reserved);
  case 5:
    return s->graphs;
  case 6:
    return s->rounds;
  case 7:
    return s->changes;
  case 8:
    return s->rebuild_us;
  case 9:
    return s->wait_us;
  case 10:
    return s->prepare_us.load();
  case 11:
    return s->admit_us;
  case 12:
    return s->retire_us;
  case 13:
    return s->idle_us;
  case 14:
    return s->submit_us;
  case 15:
    return s->release_us;
  default:
    return -1;
  }
}

}  // extern "C"
