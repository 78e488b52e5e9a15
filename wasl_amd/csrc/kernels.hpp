// Device code of the wasl_hip engine (gfx950, wave64).
//
// Fused SSP-RK3 stage for 2D Euler on triangles: one kernel evaluates the
// modal->quadrature volume integral, the HLLC face fluxes with lifting (faces
// gathered per element through precomputed edge records, no atomics), the
// inverse mass 1/detJ and the Shu-Osher combination, and optionally the
// troubled-cell sensor and the CFL speed of the new state.
//
// Work split: LANES consecutive lanes share one element. Volume points are
// dealt round-robin to the lanes, edges likewise; each lane accumulates a
// partial right-hand side for all variables and modes in registers and the
// partials are summed with quad DPP moves. All reference-element tables sit
// in LDS. State layout is the reference's U[variable][mode][element], so
// every coefficient load is coalesced over the elements of a wave.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <climits>

namespace whip {

// ---------------------------------------------------------------- sizes ----
template <int P> struct Tri {
  static constexpr int Np = (P + 1) * (P + 2) / 2;
  static constexpr int NQ = (P + 3) * (P + 3);          // volume rule of order p+2
  static constexpr int NE = P + 1;                      // points per edge
  static constexpr int NC = (P + 1) * (P + 1) + 3 * (P + 1);  // positivity check points
  // Table offsets (doubles) inside one contiguous block.
  static constexpr int kBvol = 0;
  static constexpr int kWx = kBvol + NQ * Np;
  static constexpr int kWy = kWx + NQ * Np;
  static constexpr int kBl = kWy + NQ * Np;              // [3][NE][Np]
  static constexpr int kBr = kBl + 3 * NE * Np;          // [3][3][2][NE][Np]
  static constexpr int kWe = kBr + 18 * NE * Np;         // [3][NE]
  static constexpr int kTable = kWe + 3 * NE;
  // Column sums of the weighted gradients, sum_q W_xi[q][m] and sum_q
  // W_eta[q][m] ([2][Np]): the volume term of a constant flux, used by the
  // mixed-precision kernel (uploaded after the staged table).
  static constexpr int kCs = kTable;
  static constexpr int kFull = kCs + 2 * Np;
  // Volume tables padded to a multiple of 8 points (zero weights) for the
  // variable-split kernel: B [NQ8][Np], W_xi [NQ8][Np], W_eta [NQ8][Np].
  static constexpr int NQ8 = (NQ + 7) / 8 * 8;
  static constexpr int kVB = 0;
  static constexpr int kVWx = NQ8 * Np;
  static constexpr int kVWy = 2 * NQ8 * Np;
  static constexpr int kVTable = 3 * NQ8 * Np;
  // Edge tables alone (the tail of the full table), staged in LDS.
  static constexpr int kEdge = kBl;
  static constexpr int kEdgeTable = kTable - kBl;
};

// Dubiner modes are ordered by degree n, then by i (basis/dubiner_triangle.hpp:
// 20-23). A mode with i = 0 depends on eta alone, so its xi-derivative is
// identically zero, and mode 0 is constant. The weighted-gradient tables hold
// exact zeros there (the host checks when it builds them), and the volume
// contractions leave those products out.
__host__ __device__ constexpr bool dubiner_dxi_zero(int m) {
  int n = 0;
  while ((n + 1) * (n + 2) / 2 <= m)
    ++n;
  return m == n * (n + 1) / 2;
}
__host__ __device__ constexpr bool dubiner_deta_zero(int m) { return m == 0; }

template <int P> struct Line {
  static constexpr int Np = P + 1;
  static constexpr int NQ = P > 0 ? P : 1;
  static constexpr int kB = 0;                 // [NQ][Np]
  static constexpr int kW = kB + NQ * Np;      // [NQ][Np] weight * derivative
  static constexpr int kL = kW + NQ * Np;      // [Np] phi(-1)
  static constexpr int kR = kL + Np;           // [Np] phi(+1)
  static constexpr int kTable = kR + Np;
};

// Edge-record meta bits.
constexpr int kMetaLeft = 1;        // this element is the face's left side
constexpr int kMetaBoundary = 64;   // boundary face (no right element)
__host__ __device__ inline int meta_kl(int m) { return (m >> 1) & 3; }
__host__ __device__ inline int meta_kr(int m) { return (m >> 3) & 3; }
__host__ __device__ inline int meta_flip(int m) { return (m >> 5) & 1; }
__host__ __device__ inline int meta_slot(int m) { return m >> 8; }

// Error codes raised on the device; the host maps them to the reference texts.
enum : int {
  kErrNone = 0,
  kErrUnphysical = 1,   // simulation: Euler cell average is not physical
  kErrPositivity = 2,   // Euler positivity limiter: cell average is below a floor
  kErrTimestep = 3,     // simulation: timestep cannot advance time
};

// Per-step control, double-buffered by step parity so one kernel can read
// the current step while writing the next.
struct StepSlot {
  double time;
  double dt;
  unsigned long long step;
  unsigned long long lambda;  // max |u|+c over cell averages (bits of a positive double)
  int active;
  int bad;                    // an unphysical cell average was seen
  int spread;                 // 1: the speed is the max of lambda and this parity's spread slots
  int pad_;
};

struct RunCtl {
  double end_time;
  double cfl_length;          // cfl * h_min
  double order_factor;        // 2p + 1
  unsigned long long max_steps;
  int err;
  int fixed_speed;            // 1: lambda is a case constant (scalar advection)
  double last_dt;             // dt of the last step taken
  StepSlot s[2];
};

// The CFL speed of a step's new state is a max over every workgroup of its
// last stage. All of them updating one address serialises the updates at
// the end of the stage, so they are spread over kSpeedSlots slots per step
// parity, one 128-byte line each (workgroup b to slot b % kSpeedSlots), kept
// in the step control's device allocation after the RunCtl itself. The next
// step's first stage takes the max of the slots.
constexpr int kSpeedSlots = 16;
constexpr int kSpeedStride = 16;   // unsigned long longs: 128 bytes
constexpr size_t kSpeedOffset = (sizeof(RunCtl) + 255) / 256 * 256;
constexpr size_t kCtlBytes =
    kSpeedOffset + 2 * kSpeedSlots * kSpeedStride * sizeof(unsigned long long);

__device__ __forceinline__ unsigned long long *speed_slots(RunCtl *ctl, int parity) {
  return reinterpret_cast<unsigned long long *>(reinterpret_cast<char *>(ctl) + kSpeedOffset) +
         parity * kSpeedSlots * kSpeedStride;
}

struct EulerMesh {
  int E;
  const double *geo;          // [5][E]: A B C D 1/detJ with F_xi = A Fx + B Fy, F_eta = C Fx + D Fy
  const int *nbr;             // [3][E] neighbour element or -1
  const int *meta;            // [3][E]
  const double *fn;           // [3][3][E]: nx, ny, edge scale (face values, from the left side)
  const int *bdonor;          // [NB][NE] periodic donors (or null: transmissive)
  const double *bbasis;       // [NB][NE][Np]
  const double *table;        // Tri<P>::kTable doubles
  const double *vtable;       // Tri<P>::kVTable doubles (padded volume tables)
  double gamma;
};

struct StageArgs {
  const double *uin;          // stage input, residual evaluated here
  const double *un;           // step start (RK combination), may alias out
  double *out;                // stage output, or R when residual_only
  double ca, cb;              // out = ca*un + cb*(uin + dt*R)
  double dt;                  // used when ctl == nullptr
  RunCtl *ctl;                // device-driven time stepping, or nullptr
  int parity;
  int stage;                  // 0, 1, 2
  int residual_only;
  int want_lambda;            // accumulate the CFL speed of the output
  // limiter hooks
  int sensor;                 // compute the troubled-cell sensor of the output
  double threshold;
  int highest_begin;
  uint8_t *flags;             // [E] or null
  int *list;                  // troubled elements
  int *counters;              // [3] per-stage troubled counts
  // mixed precision: FP32 edge traces [E][3 edges][4 points][4 variables] of
  // uin (read for neighbours) and of out (written for the next stage, or null)
  const float *tin;
  float *tout;
};

// ------------------------------------------------------------- helpers ----

template <int CTRL> __device__ __forceinline__ double dpp(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
// quad_perm controls: lane i of each quad reads lane perm[i].
constexpr int kQuadXor1 = 0xB1;   // [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;   // [2,3,0,1]
template <int SRC> constexpr int quad_bcast() { return SRC | (SRC << 2) | (SRC << 4) | (SRC << 6); }

template <int LANES> __device__ __forceinline__ double lane_sum(double x) {
  if constexpr (LANES >= 2)
    x += dpp<kQuadXor1>(x);
  if constexpr (LANES >= 4)
    x += dpp<kQuadXor2>(x);
  return x;
}

__device__ __forceinline__ double wave_max(double x) {
  for (int o = 32; o > 0; o >>= 1)
    x = fmax(x, __shfl_xor(x, o, 64));
  return x;
}

__device__ __forceinline__ void raise_error(RunCtl *ctl, int *err, int code) {
  int *slot = ctl ? &ctl->err : err;
  if (slot)
    atomicCAS(slot, 0, code);
}

// Appends `e` to a troubled-element list for every lane whose `want` is set:
// one atomic per wave reserves the wave's entries, which follow the wave's
// base in lane order, so the count is not updated once per element. Entries
// past E are dropped. Every lane must be active.
__device__ __forceinline__ void list_append(int *count, int *list, int E, bool want, int e) {
  const unsigned long long mask = __ballot(want);
  if (mask == 0ull)
    return;
  const int first = __ffsll(static_cast<long long>(mask)) - 1;
  int base = 0;
  if ((static_cast<int>(threadIdx.x) & 63) == first)
    base = atomicAdd(count, __popcll(mask));
  base = __builtin_amdgcn_readlane(base, first);
  const int below = __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                              __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u));
  if (want && base + below < E)
    list[base + below] = e;
}

// Speeds are non-negative doubles, which order like their bit patterns.
__device__ __forceinline__ void speed_update(RunCtl *ctl, int parity, double speed) {
  atomicMax(speed_slots(ctl, parity) + (blockIdx.x % kSpeedSlots) * kSpeedStride,
            static_cast<unsigned long long>(__double_as_longlong(speed)));
}

template <int CTRL> __device__ __forceinline__ unsigned long long dpp_u64(unsigned long long x) {
  const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(x >> 32), CTRL, 0xF, 0xF, false);
  return (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
         static_cast<unsigned>(lo);
}

// Lane l's share of the spread slots of a step parity: slot l % 16.
__device__ __forceinline__ unsigned long long speed_slot_of_lane(RunCtl *ctl, int parity) {
  static_assert(kSpeedSlots == 16, "one slot per lane of a row");
  return speed_slots(ctl, parity)[(static_cast<int>(threadIdx.x) & 15) * kSpeedStride];
}

// The max of the slots the lanes read: each 16-lane row reduces by quad
// permutes and row rotations, and the result is wave-uniform. Every lane
// must be active.
__device__ __forceinline__ unsigned long long speed_slots_max(unsigned long long x) {
  const auto mx = [](unsigned long long a, unsigned long long b) { return a < b ? b : a; };
  x = mx(x, dpp_u64<kQuadXor1>(x));
  x = mx(x, dpp_u64<kQuadXor2>(x));
  x = mx(x, dpp_u64<0x124>(x));   // row_ror:4
  x = mx(x, dpp_u64<0x128>(x));   // row_ror:8
  const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<int>(x));
  const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<int>(x >> 32));
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// Ideal-gas pointwise physics (equations/euler/primitive.hpp, flux.hpp).
struct C4 {
  double r, mx, my, E;
};

// Reciprocal to within an ulp: the hardware estimate refined by two Newton
// steps (5 instructions instead of the ~10 of a correctly rounded divide).
__device__ __forceinline__ double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  return fma(fma(-x, r, 1.0), r, r);
}

struct Side {
  double inv_r, u, v, p, un, c;
};

__device__ __forceinline__ Side side_of(double gamma, const C4 &s, double nx, double ny) {
  Side o;
  o.inv_r = rcp(s.r);
  o.u = s.mx * o.inv_r;
  o.v = s.my * o.inv_r;
  o.p = (gamma - 1.0) * (s.E - 0.5 * (s.mx * o.u + s.my * o.v));
  o.un = o.u * nx + o.v * ny;
  o.c = sqrt(gamma * o.p * o.inv_r);
  return o;
}

__device__ __forceinline__ C4 normal_flux(const C4 &s, const Side &w, double nx, double ny) {
  return {s.r * w.un, s.mx * w.un + w.p * nx, s.my * w.un + w.p * ny, (s.E + w.p) * w.un};
}

// std::min / std::max as the reference calls them: with a NaN operand they
// return the first one (fmin/fmax would return the other), so a NaN sound
// speed (a negative point pressure) propagates exactly as it does there.
__device__ __forceinline__ double first_min(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double first_max(double a, double b) { return a < b ? b : a; }

// HLLC flux with Davis wave-speed bounds, evaluated (left, right, n_left)
// (flux/hllc.hpp:40-74).
__device__ inline C4 hllc(double gamma, const C4 &L, const C4 &R, double nx, double ny) {
  const Side wl = side_of(gamma, L, nx, ny), wr = side_of(gamma, R, nx, ny);
  const double sl = first_min(wl.un - wl.c, wr.un - wr.c);
  const double sr = first_max(wl.un + wl.c, wr.un + wr.c);
  if (sl >= 0.0)
    return normal_flux(L, wl, nx, ny);
  if (sr <= 0.0)
    return normal_flux(R, wr, nx, ny);
  const double sm = (wr.p - wl.p + L.r * wl.un * (sl - wl.un) - R.r * wr.un * (sr - wr.un)) *
                    rcp(L.r * (sl - wl.un) - R.r * (sr - wr.un));
  const bool left = sm >= 0.0;
  const C4 &S = left ? L : R;
  const Side &w = left ? wl : wr;
  const double s = left ? sl : sr;
  const C4 F = normal_flux(S, w, nx, ny);
  const double k = s - w.un;
  const double rs = S.r * k * rcp(s - sm);
  const double shift = sm - w.un;
  const double es = S.E * w.inv_r + shift * (sm + w.p * w.inv_r * rcp(k));
  C4 o;
  o.r = F.r + s * (rs - S.r);
  o.mx = F.mx + s * (rs * (w.u + shift * nx) - S.mx);
  o.my = F.my + s * (rs * (w.v + shift * ny) - S.my);
  o.E = F.E + s * (rs * es - S.E);
  return o;
}

// Device-driven time stepping (serial_case_runner.cpp:107-142): stage 0 of a
// step decides whether the step runs (time < end_time, step < max_steps),
// computes dt = min(cfl h / (lambda (2p+1)), end_time - time) from the speed
// gathered at the end of the previous step, and publishes it for stages 1-2.
// Every block evaluates the same decision from the same data, so the block
// exits are uniform. Returns false when the stage must not run.
__device__ inline bool stage_prologue(RunCtl *ctl, int parity, int stage, double &dt,
                                      bool leader) {
  if (!ctl)
    return true;
  // Every field is read in one batch before any decision: one memory round
  // trip instead of a chain of dependent ones.
  StepSlot &slot = ctl->s[parity];
  const int err = ctl->err;
  const double time = slot.time, slot_dt = slot.dt;
  const unsigned long long step = slot.step, lambda = slot.lambda;
  const int slot_active = slot.active, bad = slot.bad;
  const double end_time = ctl->end_time, cfl_length = ctl->cfl_length;
  const double order_factor = ctl->order_factor;
  const unsigned long long max_steps = ctl->max_steps;
  const int fixed_speed = ctl->fixed_speed;
  const int spread = slot.spread;
  const unsigned long long part = stage == 0 ? speed_slot_of_lane(ctl, parity) : 0ull;
  if (err)
    return false;
  if (stage != 0) {
    if (!slot_active)
      return false;
    dt = slot_dt;
    return true;
  }
  // the speed of the step's start: lambda, with the slots the previous
  // step's last stage filled when it gathered the speed there
  const unsigned long long slots = speed_slots_max(part);
  const unsigned long long lam = spread && slots > lambda ? slots : lambda;
  const bool active = time < end_time && step < max_steps;
  if (!active) {
    if (leader && threadIdx.x == 0) {
      slot.active = 0;
      StepSlot &nxt = ctl->s[parity ^ 1];
      nxt.time = time;
      nxt.step = step;
      nxt.lambda = lam;
      nxt.spread = 0;
      nxt.bad = bad;
      nxt.active = 0;
    }
    return false;
  }
  if (bad) {
    if (threadIdx.x == 0)
      raise_error(ctl, nullptr, kErrUnphysical);
    return false;
  }
  const double speed = __longlong_as_double(static_cast<long long>(lam));
  const double remaining = end_time - time;
  const double stable = speed > 0.0 ? cfl_length / (speed * order_factor) : 0.0;
  dt = fmin(stable > 0.0 ? stable : remaining, remaining);
  if (dt <= 0.0 || time + dt == time) {
    if (threadIdx.x == 0)
      raise_error(ctl, nullptr, kErrTimestep);
    return false;
  }
  if (leader && threadIdx.x == 0) {
    slot.dt = dt;
    slot.active = 1;
    ctl->last_dt = dt;
    StepSlot &nxt = ctl->s[parity ^ 1];
    nxt.lambda = fixed_speed ? lam : 0ull;
    nxt.spread = fixed_speed ? 0 : 1;
    nxt.bad = 0;
  }
  // the next step's slots start empty (its last stage fills them)
  if (leader && !fixed_speed && threadIdx.x < kSpeedSlots)
    speed_slots(ctl, parity ^ 1)[threadIdx.x * kSpeedStride] = 0ull;
  return true;
}

// Workgroups are handed to the 8 XCDs in turn (workgroup b to XCD b % 8),
// each XCD with its own L2. This renumbering gives every XCD one contiguous
// range of workgroups, so the face neighbours of an element, which the mesh
// numbers close to it, are mostly read through the same L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
  const int n = static_cast<int>(gridDim.x), b = static_cast<int>(blockIdx.x);
  const int x = b % kXcds, i = b / kXcds, q = n / kXcds, r = n % kXcds;
  return x * q + (x < r ? x : r) + i;
}

// The last stage of a step advances time and step count for the next step.
__device__ inline void stage_epilogue(RunCtl *ctl, int parity, int stage, double dt,
                                      bool leader) {
  if (ctl && stage == 2 && leader && threadIdx.x == 0) {
    StepSlot &cur = ctl->s[parity];
    StepSlot &nxt = ctl->s[parity ^ 1];
    nxt.time = cur.time + dt;
    nxt.step = cur.step + 1;
  }
}

// ------------------------------------------------------ 2D Euler stage ----

// Trace of one element's coefficients c[4][Np] against basis values b[Np].
template <int Np> __device__ __forceinline__ C4 trace_of(const double (&c)[4][Np],
                                                         const double *b) {
  double t[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    double x = 0.0;
#pragma unroll
    for (int m = 0; m < Np; ++m)
      x += c[v][m] * b[m];
    t[v] = x;
  }
  return {t[0], t[1], t[2], t[3]};
}

// Raw buffer access (gfx9 descriptor, 32-bit offsets): the array base and
// size sit in scalar registers, a lane passes one 32-bit byte offset and a
// row one wave-uniform byte offset, so a load or store needs no address
// arithmetic per row. Loads past `bytes` return zero and stores past it are
// dropped; every access here is in bounds (arrays stay below 4 GiB).
using Rsrc = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ Rsrc rsrc(const void *p, size_t bytes) {
  const unsigned n = bytes < 0xffffffffull ? static_cast<unsigned>(bytes) : 0xffffffffu;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, static_cast<int>(n),
                                           0x00020000);
}
__device__ __forceinline__ double ld_f64(Rsrc r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ int ld_i32(Rsrc r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(int, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ float4 ld_f32x4(Rsrc r, unsigned voff, unsigned soff) {
  const auto q = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  float4 f;
  __builtin_memcpy(&f, &q, sizeof f);
  return f;
}
__device__ __forceinline__ void st_f64(double x, Rsrc r, unsigned voff, unsigned soff) {
  using B = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0u, 0u, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(B, x), r, voff, soff, 0);
}
__device__ __forceinline__ void st_f32x4(float4 x, Rsrc r, unsigned voff, unsigned soff) {
  using B = decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0u, 0u, 0));
  B q;
  __builtin_memcpy(&q, &x, sizeof q);
  __builtin_amdgcn_raw_buffer_store_b128(q, r, voff, soff, 0);
}

// 16 bytes per lane from global memory straight into LDS (gfx950
// global_load_lds_dwordx4): lane l's bytes land at dst + 16 l, with dst
// wave-uniform and 16-byte aligned. The data arrive asynchronously and are
// counted as vector-memory loads of the issuing wave.
__device__ __forceinline__ void lds_copy16(const void *src, void *dst) {
  __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 0);
}

// Element at a 32-bit byte offset from a wave-uniform base: the load takes
// the scalar-base form (base in scalar registers, one 32-bit offset register
// per lane shared by every row), with no per-load address arithmetic.
//
// The access is made through a global-address-space pointer: every array
// read this way is device memory, and a pointer the kernel read from memory
// (a batch's case table) would otherwise compile to flat loads, which wait
// for LDS and vector memory alike.
template <typename T>
__device__ __forceinline__ T at_off(const T *base, unsigned byte_off) {
  using G = const __attribute__((address_space(1))) T;
  return *(G *)(reinterpret_cast<const char *>(base) + byte_off);
}

// Element idx of an array below 4 GiB: one 32-bit offset per lane from a
// scalar base, as at_off.
template <typename T>
__device__ __forceinline__ T at_idx(const T *base, unsigned idx) {
  return at_off(base, idx * static_cast<unsigned>(sizeof(T)));
}

// Row r of an SoA array of E columns at column i (r wave-uniform).
template <typename T>
__device__ __forceinline__ T at_row(const T *base, int E, int r, int i) {
  const T *row = base + static_cast<size_t>(r) * static_cast<unsigned>(E);
  return at_off(row, static_cast<unsigned>(i) * static_cast<unsigned>(sizeof(T)));
}

template <int Np>
__device__ __forceinline__ void load_coefficients(const double *U, int E, int e,
                                                  double (&c)[4][Np]) {
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int m = 0; m < Np; ++m)
      c[v][m] = at_row(U, E, v * Np + m, e);
}

// Two co-resident waves per SIMD: the register budget is capped at 256.
// One workgroup of the point split: elements [block * 256 / LANES, ...) of
// the mesh; the leader workgroup publishes the step control. tab and red are
// the workgroup's LDS.
template <int P, int LANES>
__device__ __forceinline__ void euler_stage_body(const EulerMesh &mesh, const StageArgs &a,
                                                 const int block, const bool leader) {
  using D = Tri<P>;
  constexpr int Np = D::Np, NQ = D::NQ, NE = D::NE;
  constexpr int PER_BLOCK = 256 / LANES;
  __shared__ double tab[D::kTable];
  __shared__ double red[4];
  // From P2 on, U^n of each wave's elements streams into LDS during the
  // volume and face work ([row][element] per wave; rows v Np + m).
  constexpr int EPW = 64 / LANES;   // elements per wave
  constexpr bool kUnLds = P >= 2 && LANES >= 2 && (4 * Np * EPW * 8) % 1024 == 0;
  __shared__ __align__(16) double unl[kUnLds ? 4 : 1][kUnLds ? 4 * Np * EPW : 1];

  const int E = mesh.E;
  const int sub = threadIdx.x % LANES;
  const int e_raw = block * PER_BLOCK + threadIdx.x / LANES;
  const bool valid = e_raw < E;
  const int e = valid ? e_raw : E - 1;
  const double g = mesh.gamma, g1 = g - 1.0;

  // Global requests first, ahead of the step-control read: own coefficients
  // and, with one edge per lane, the edge record and the neighbour's
  // coefficients, so their latency overlaps the prologue, the table staging
  // and the volume integral. Every address is in bounds whether or not the
  // stage runs.
  double c[4][Np];
  load_coefficients<Np>(a.uin, E, e, c);
  constexpr bool kPrefetch = LANES == 4 && P <= 2;
  double nc[4][kPrefetch ? Np : 1];
  int meta0 = 0, nb0 = -1;
  if constexpr (kPrefetch) {
    const int k0 = sub < 3 ? sub : 0;
    meta0 = mesh.meta[k0 * E + e];
    nb0 = mesh.nbr[k0 * E + e];
  }

  // Up to P1 the registers also hold the metrics and 1/detJ from here on.
  // At P2 with 4 lanes they hold the four metrics (the volume then starts
  // without a memory round trip); with fewer lanes, and at P3, there is no
  // room for them before the volume.
  constexpr bool kEarlyMetrics = P <= 1;
  double gm[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  constexpr bool kEarlyGeo = P == 2 && LANES == 4;
  if constexpr (kEarlyMetrics) {
#pragma unroll
    for (int i = 0; i < 5; ++i)
      gm[i] = at_row(mesh.geo, E, i, e);
  } else if constexpr (kEarlyGeo) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      gm[i] = at_row(mesh.geo, E, i, e);
  }
  // this thread's share of the reference tables, staged to LDS below
  constexpr int kPer = (D::kTable + 255) / 256;
  double tpre[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int idx = static_cast<int>(threadIdx.x) + 256 * i;
    tpre[i] = idx < D::kTable ? mesh.table[idx] : 0.0;
  }

  RunCtl *ctl = a.ctl;
  double dt = a.dt;
  if (!stage_prologue(ctl, a.parity, a.stage, dt, leader))
    return;
  if (a.sensor && leader && threadIdx.x == 0)
    a.counters[(a.stage + 1) % 3] = 0;

  // Each wave copies its own slab (16 bytes per lane: two elements of one
  // row), so only the wave itself waits for it. Rows start 16-byte aligned
  // when E is even; lanes past the last element copy nothing.
  const bool un_lds = kUnLds && a.ca != 0.0 && (E % 2) == 0;
  const int wv = static_cast<int>(threadIdx.x) >> 6;
  if (un_lds) {
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const int e0 = block * PER_BLOCK + wv * EPW;
    constexpr int kChunks = 4 * Np * EPW * 8 / 1024;
#pragma unroll
    for (int i = 0; i < kChunks; ++i) {
      const int idx = 64 * i + lane;   // 16-byte unit of the slab
      const int row = idx / (EPW / 2), el = e0 + 2 * (idx % (EPW / 2));
      if (el < E)
        lds_copy16(a.un + static_cast<size_t>(row) * static_cast<unsigned>(E) + el,
                   &unl[wv][128 * i]);
    }
  }

#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int idx = static_cast<int>(threadIdx.x) + 256 * i;
    if (idx < D::kTable)
      tab[idx] = tpre[i];
  }
  __syncthreads();

  double acc[4][Np];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int m = 0; m < Np; ++m)
      acc[v][m] = 0.0;

  // The neighbour's coefficients (its index arrived with the first round
  // trip): requested now, their latency overlapping the volume integral.
  if constexpr (kPrefetch)
    load_coefficients<Np>(a.uin, E, nb0 >= 0 ? nb0 : e, nc);

  // Volume: sum_q (W_xi[q][m] F_xi + W_eta[q][m] F_eta) with the reference
  // fluxes F_xi = A Fx + B Fy, F_eta = C Fx + D Fy (= detJ J^-1 F).
  if constexpr (P > 0) {
    if constexpr (!kEarlyMetrics && !kEarlyGeo) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        gm[i] = at_row(mesh.geo, E, i, e);
    }
    const double gA = gm[0], gB = gm[1], gC = gm[2], gD = gm[3];
#pragma unroll 1
    for (int it = 0; it < (NQ + LANES - 1) / LANES; ++it) {
      const int q = it * LANES + sub;
      if (q < NQ) {
        const C4 s = trace_of<Np>(c, tab + D::kBvol + q * Np);
        const double inv = rcp(s.r);
        const double u = s.mx * inv, w = s.my * inv;
        const double p = g1 * (s.E - 0.5 * (s.mx * u + s.my * w));
        // reference-frame fluxes through the contravariant velocities
        // U_xi = A u + B v, U_eta = C u + D v
        const double uxi = gA * u + gB * w, ueta = gC * u + gD * w, enp = s.E + p;
        const double fxi[4] = {s.r * uxi, s.mx * uxi + gA * p, s.my * uxi + gB * p, enp * uxi};
        const double feta[4] = {s.r * ueta, s.mx * ueta + gC * p, s.my * ueta + gD * p,
                                enp * ueta};
        const double *wx = tab + D::kWx + q * Np;
        const double *wy = tab + D::kWy + q * Np;
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int m = 0; m < Np; ++m) {
            if (!dubiner_dxi_zero(m))
              acc[v][m] = fma(wx[m], fxi[v], acc[v][m]);
            if (!dubiner_deta_zero(m))
              acc[v][m] = fma(wy[m], feta[v], acc[v][m]);
          }
      }
    }
  }

  // Lane `sub` finalises variables sub*VPL .. sub*VPL + VPL - 1. Their U^n
  // (stages 2 and 3): up to P1 it is requested here, its latency overlapping
  // the faces; at P2 and P3 the registers go to the faces and it is read in
  // the epilogue.
  constexpr int VPL = 4 / LANES;
  double uo[VPL][Np];
#pragma unroll
  for (int j = 0; j < VPL; ++j)
#pragma unroll
    for (int m = 0; m < Np; ++m)
      uo[j][m] = 0.0;
  if constexpr (kEarlyMetrics) {
    if (a.ca != 0.0 && !a.residual_only) {
#pragma unroll
      for (int j = 0; j < VPL; ++j)
#pragma unroll
        for (int m = 0; m < Np; ++m)
          uo[j][m] = at_idx(a.un, static_cast<unsigned>((sub * VPL + j) * Np + m) *
                                      static_cast<unsigned>(E) + static_cast<unsigned>(e));
    }
  }

  // 1/detJ is requested here (from P2 on), its latency overlapping the faces.
  const double invdet = kEarlyMetrics ? gm[4] : at_row(mesh.geo, E, 4, e);

  // Faces: each lane takes whole edges; traces on both sides use the left
  // side's quadrature point and HLLC is always evaluated (left, right,
  // n_left), so the two elements of a face see identical flux bits (exact
  // conservation).
  for (int k = sub; k < 3; k += LANES) {
    int meta, nb;
    if constexpr (kPrefetch) {
      meta = meta0;
      nb = nb0;
    } else {
      meta = mesh.meta[k * E + e];
      nb = mesh.nbr[k * E + e];
    }
    const double nx = mesh.fn[(3 * k) * E + e];
    const double ny = mesh.fn[(3 * k + 1) * E + e];
    const double scale = mesh.fn[(3 * k + 2) * E + e];
    const bool is_left = meta & kMetaLeft;
    const bool boundary = meta & kMetaBoundary;
    const int kl = meta_kl(meta), kr = meta_kr(meta), fl = meta_flip(meta);
    const double *bl = tab + D::kBl + kl * NE * Np;
    const double *br = tab + D::kBr + ((kl * 3 + kr) * 2 + fl) * NE * Np;
    const double *own = is_left ? bl : br;
    const double *oth = is_left ? br : bl;
    // Neighbour traces at all edge points, streaming its coefficients once.
    double so[NE][4];
#pragma unroll
    for (int q = 0; q < NE; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        so[q][v] = 0.0;
    if (!boundary) {
      const int src = nb;
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int m = 0; m < Np; ++m) {
          double x;
          if constexpr (kPrefetch)
            x = nc[v][m];
          else
            x = at_row(a.uin, E, v * Np + m, src);
#pragma unroll
          for (int q = 0; q < NE; ++q)
            so[q][v] += x * oth[q * Np + m];
        }
    }
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      const C4 si = trace_of<Np>(c, own + q * Np);
      C4 so4 = {so[q][0], so[q][1], so[q][2], so[q][3]};
      if (boundary) {
        if (mesh.bdonor) {
          const int slot = meta_slot(meta);
          const int d = mesh.bdonor[slot * NE + q];
          const double *db = mesh.bbasis + (slot * NE + q) * Np;
          double t[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            double x = 0.0;
#pragma unroll
            for (int m = 0; m < Np; ++m)
              x += at_row(a.uin, E, v * Np + m, d) * db[m];
            t[v] = x;
          }
          so4 = {t[0], t[1], t[2], t[3]};
        } else {
          so4 = si;  // transmissive: exterior = own trace
        }
      }
      const C4 L = is_left ? si : so4;
      const C4 R = is_left ? so4 : si;
      const C4 f = hllc(g, L, R, nx, ny);
      const double wq = tab[D::kWe + kl * NE + q] * scale;
      const double sw = is_left ? -wq : wq;
      const double val[4] = {f.r * sw, f.mx * sw, f.my * sw, f.E * sw};
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int m = 0; m < Np; ++m)
          acc[v][m] += val[v] * own[q * Np + m];
    }
  }

  // Reduce-scatter over the element's lanes: each lane keeps the full sums of
  // the variables it finalises. Partners one lane apart combine first, then
  // two apart, so every sum is (x0 + x1) + (x2 + x3) on whichever lane holds
  // it.
  double res[VPL][Np];
#pragma unroll
  for (int m = 0; m < Np; ++m) {
    if constexpr (LANES == 1) {
#pragma unroll
      for (int v = 0; v < 4; ++v)
        res[v][m] = acc[v][m];
    } else if constexpr (LANES == 2) {
      const bool odd = sub & 1;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const double send = odd ? acc[k][m] : acc[2 + k][m];
        const double keep = odd ? acc[2 + k][m] : acc[k][m];
        res[k][m] = keep + dpp<kQuadXor1>(send);
      }
    } else {
      const bool odd = sub & 1, hi = sub & 2;
      double kp[2];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const double send = odd ? acc[2 * k][m] : acc[2 * k + 1][m];
        const double keep = odd ? acc[2 * k + 1][m] : acc[2 * k][m];
        kp[k] = keep + dpp<kQuadXor1>(send);
      }
      const double send = hi ? kp[0] : kp[1];
      const double keep = hi ? kp[1] : kp[0];
      res[0][m] = keep + dpp<kQuadXor2>(send);
    }
  }
  // this lane's share of the element's own coefficients
  const auto own_c = [&](int j, int m) -> double {
    if constexpr (LANES == 1) {
      return c[j][m];
    } else if constexpr (LANES == 2) {
      return (sub & 1) ? c[2 + j][m] : c[j][m];
    } else {
      const double lo = (sub & 1) ? c[1][m] : c[0][m];
      const double hi = (sub & 1) ? c[3][m] : c[2][m];
      return (sub & 2) ? hi : lo;
    }
  };

  if (un_lds) {
    __builtin_amdgcn_s_waitcnt(0);   // this wave's slab has landed in LDS
    __asm__ volatile("" ::: "memory");
  }
  const int el_w = static_cast<int>(threadIdx.x) / LANES - wv * EPW;   // element within the wave

  // Finalise. Every U^n read comes before the first store: the slab and
  // global reads would otherwise wait for the stores issued ahead of them.
  // One loop per source: a single loop choosing between the slab and global
  // memory compiles to flat loads, which wait for LDS and vector memory alike.
  if constexpr (!kEarlyMetrics) {
    if (a.ca != 0.0 && !a.residual_only) {
      bool from_slab = false;
      if constexpr (kUnLds)
        from_slab = un_lds;
      if (from_slab) {
#pragma unroll
        for (int j = 0; j < VPL; ++j)
#pragma unroll
          for (int m = 0; m < Np; ++m)
            uo[j][m] = unl[wv][((sub * VPL + j) * Np + m) * EPW + el_w];
      } else {
#pragma unroll
        for (int j = 0; j < VPL; ++j)
#pragma unroll
          for (int m = 0; m < Np; ++m)
            uo[j][m] = at_idx(a.un, static_cast<unsigned>((sub * VPL + j) * Np + m) *
                                            static_cast<unsigned>(E) +
                                        static_cast<unsigned>(e));
      }
    }
  }
  double ys[VPL][Np];
#pragma unroll
  for (int j = 0; j < VPL; ++j)
#pragma unroll
    for (int m = 0; m < Np; ++m) {
      const int row = (sub * VPL + j) * Np + m;
      const double r = res[j][m] * invdet;
      double y;
      if (a.residual_only) {
        y = r;
      } else {
        y = a.cb * (own_c(j, m) + dt * r);
        if (a.ca != 0.0)
          y += a.ca * uo[j][m];
      }
      if (valid)
        a.out[row * E + e] = y;
      ys[j][m] = y;
    }

  // The new cell averages of all four variables on the element's first lane,
  // and its density modes for the sensor (the first lane owns variable 0).
  double out0[4];
  if constexpr (LANES == 1) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
      out0[v] = ys[v][0];
  } else if constexpr (LANES == 2) {
    constexpr int kPairLead = 0xF5;   // quad_perm [1,1,3,3]
    out0[0] = ys[0][0];
    out0[1] = ys[1][0];
    out0[2] = dpp<kPairLead>(ys[0][0]);
    out0[3] = dpp<kPairLead>(ys[1][0]);
  } else {
    out0[0] = ys[0][0];
    out0[1] = dpp<quad_bcast<1>()>(ys[0][0]);
    out0[2] = dpp<quad_bcast<2>()>(ys[0][0]);
    out0[3] = dpp<quad_bcast<3>()>(ys[0][0]);
  }
  const double (&dens)[Np] = ys[0];

  if (a.sensor && !a.residual_only) {
    bool troubled = false;
    if (sub == 0 && valid) {
      // Modal-decay sensor on density (euler_modal_limiter.cpp:18-24).
      double total = 0.0, high = 0.0;
#pragma unroll
      for (int m = 0; m < Np; ++m) {
        total += dens[m] * dens[m];
        if (m >= a.highest_begin)
          high += dens[m] * dens[m];
      }
      troubled = !(total == 0.0 || high / total <= a.threshold);
      if (a.flags)
        a.flags[e] = troubled ? 1 : 0;
    }
    // at most one entry per element and stage (the leader clears the count)
    list_append(&a.counters[a.stage], a.list, E, troubled, e);
  }

  if (a.want_lambda && !a.residual_only) {
    // |u| + c of the new cell average (cfl_timestep.cpp:33-47).
    double speed = 0.0;
    if (sub == 0 && valid) {
      const double is2 = 1.0 / sqrt(2.0);
      const double r = is2 * out0[0], mx = is2 * out0[1], my = is2 * out0[2], en = is2 * out0[3];
      const double inv = 1.0 / r;
      const double u = mx * inv, w = my * inv;
      const double p = g1 * (en - 0.5 * (mx * u + my * w));
      if (!(r > 0.0 && p > 0.0)) {
        if (ctl)
          ctl->s[a.parity ^ 1].bad = 1;
        speed = 0.0;
      } else {
        speed = hypot(u, w) + sqrt(g * p / r);
      }
    }
    speed = wave_max(speed);
    if ((threadIdx.x & 63) == 0)
      red[threadIdx.x >> 6] = speed;
    __syncthreads();
    if (threadIdx.x == 0) {
      const double mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
      if (ctl)
        speed_update(ctl, a.parity ^ 1, mx);
    }
  }
  stage_epilogue(ctl, a.parity, a.stage, dt, leader);
}

template <int P, int LANES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void euler_stage(EulerMesh mesh, StageArgs a) {
  euler_stage_body<P, LANES>(mesh, a, xcd_block(), blockIdx.x == 0);
}

// ------------------------------------------- 2D Euler, variable split ----

// Padded volume tables per order in the constant address space: the values
// depend only on the reference element, and wave-uniform reads from here are
// served by the scalar cache into SGPR operands.
constexpr int kVTableMax = 3 * 40 * 10;   // Tri<3>::kVTable
__constant__ double g_vtable[4][kVTableMax];

// 4x4 transpose across the four lanes of each quad: lane v enters with row v
// (x[j] = M[v][j]) and leaves with column v (x[j] = M[j][v]). Partners two
// lanes apart swap the off-diagonal 2x2 blocks, then partners one lane apart
// swap within them. Every lane of the quad must be active.
__device__ __forceinline__ void quad_transpose(double (&x)[4]) {
  const int v = static_cast<int>(threadIdx.x) & 3;
  const bool b1 = (v & 2) != 0, b0 = (v & 1) != 0;
  double a[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const double t = dpp<kQuadXor2>(x[j ^ 2]);
    a[j] = (((j & 2) != 0) == b1) ? x[j] : t;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const double t = dpp<kQuadXor1>(a[j ^ 1]);
    x[j] = (((j & 1) != 0) == b0) ? a[j] : t;
  }
}

// The same fused stage with a different work split: each element is served
// by H quads of lanes; lane v of a quad owns conserved variable v -- its Np
// coefficients and Np right-hand-side accumulators -- and the quads of an
// element take interleaved groups of four volume points and whole edges.
// States at points and point fluxes move between the four lanes of a quad
// by 4x4 transposes in DPP quad permutes (no LDS round trip), so no lane
// holds another variable's modes and no cross-lane reduction of the
// accumulators is needed (a single one across the H quads when H = 2).
// Scheduling barriers at the exchange points keep the compiler from
// stretching live ranges across them (without them P3 spilled). Volume
// tables are wave-uniform and read through the scalar cache; edge tables
// depend on each element's face configuration and are staged in LDS.
// LDS of one variable-split workgroup.
template <int P> struct alignas(16) VsShared {
  alignas(16) double tab[Tri<P>::kEdgeTable];
  double red[4];
};

// One workgroup of the variable split: elements [block * 64 / H, ...); the
// leader workgroup publishes the step control.
template <int P, int H>
__device__ __forceinline__ void euler_stage_vs_body(const EulerMesh mesh, const StageArgs a,
                                                    const int block, const bool leader,
                                                    VsShared<P> &sh) {
  using D = Tri<P>;
  constexpr int Np = D::Np, NE = D::NE;
  constexpr int PER_BLOCK = 64 / H;
  constexpr int NIT = (D::NQ + 4 * H - 1) / (4 * H);
  double *tab = sh.tab;
  double *red = sh.red;

  const int E = mesh.E;
  const int v = threadIdx.x & 3;              // owned variable
  const int quad = threadIdx.x >> 2;          // quad slot in the block
  const int h = quad % H;                     // quad index within the element
  const int e_raw = block * PER_BLOCK + quad / H;
  const bool valid = e_raw < E;
  const int e = valid ? e_raw : E - 1;
  const double g = mesh.gamma, g1 = g - 1.0;

  // The global reads of the stage are issued here, ahead of the step-control
  // read (all addresses in bounds): own coefficients, metrics, the records
  // of this quad's edges, U^n, 1/detJ and this thread's share of the edge
  // tables. The neighbours' coefficients follow once their indices are in,
  // so the second round trip overlaps the volume integral.
  double c[Np], acc[Np];
#pragma unroll
  for (int m = 0; m < Np; ++m) {
    c[m] = at_idx(a.uin, static_cast<unsigned>((v * Np + m) * E + e));
    acc[m] = 0.0;
  }
  const double gA = mesh.geo[e], gB = mesh.geo[E + e], gC = mesh.geo[2 * E + e],
               gD = mesh.geo[3 * E + e];
  constexpr int KE = (3 + H - 1) / H;   // edges per quad
  // Edges whose neighbour coefficients are fetched here; with one quad per
  // element at P3 the rest are fetched one edge ahead inside the face loop,
  // which keeps the kernel within its register budget.
  constexpr int KP = (H == 1 && Np > 6) ? 1 : KE;
  int emeta[KE], esrc[KE];
  double enx[KE], eny[KE], escale[KE], env[KP][Np];
#pragma unroll
  for (int i = 0; i < KE; ++i) {
    const int k = h + i * H < 3 ? h + i * H : 0;
    emeta[i] = mesh.meta[k * E + e];
    const int nb = mesh.nbr[k * E + e];
    enx[i] = mesh.fn[(3 * k) * E + e];
    eny[i] = mesh.fn[(3 * k + 1) * E + e];
    escale[i] = mesh.fn[(3 * k + 2) * E + e];
    esrc[i] = (emeta[i] & kMetaBoundary) ? e : nb;
  }
  double un[Np];
#pragma unroll
  for (int m = 0; m < Np; ++m)
    un[m] = a.ca != 0.0 ? a.un[(v * Np + m) * E + e] : 0.0;
  const double invdet = mesh.geo[4 * E + e];
  constexpr int kPer = (D::kEdgeTable + 255) / 256;
  double tpre[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int idx = static_cast<int>(threadIdx.x) + 256 * i;
    tpre[i] = idx < D::kEdgeTable ? mesh.table[D::kEdge + idx] : 0.0;
  }

  RunCtl *ctl = a.ctl;
  double dt = a.dt;
  if (!stage_prologue(ctl, a.parity, a.stage, dt, leader))
    return;
  if (a.sensor && leader && threadIdx.x == 0)
    a.counters[(a.stage + 1) % 3] = 0;

#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int idx = static_cast<int>(threadIdx.x) + 256 * i;
    if (idx < D::kEdgeTable)
      tab[idx] = tpre[i];
  }
  __syncthreads();

  // The neighbours' coefficients (their indices arrived with the first round
  // trip): requested now, their latency overlapping the volume integral.
#pragma unroll
  for (int i = 0; i < KP; ++i)
#pragma unroll
    for (int m = 0; m < Np; ++m)
      env[i][m] = at_row(a.uin, E, v * Np + m, esrc[i]);

  static_assert(D::kVTable <= kVTableMax, "volume table slot too small");
  const double *vt = g_vtable[P];

  // Volume integral, four points per quad and iteration.
  if constexpr (P > 0) {
#pragma unroll 1
    for (int it = 0; it < NIT; ++it) {
      const int q0 = (it * H + h) * 4;
      // x[j]: variable v at point q0 + j; after the transpose, x[k]: variable k
      // at point q0 + v
      double x[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double *b = vt + D::kVB + (q0 + j) * Np;
        double t = 0.0;
#pragma unroll
        for (int m = 0; m < Np; ++m)
          t += c[m] * b[m];
        x[j] = t;
      }
      quad_transpose(x);
      __builtin_amdgcn_sched_barrier(0);
      double fxi4[4], feta4[4];
      {
        // lane v evaluates the flux at point q0 + v
        const double s0 = x[0], s1 = x[1], s2 = x[2], s3 = x[3];
        const double inv = rcp(s0);
        const double u = s1 * inv, w = s2 * inv;
        const double p = g1 * (s3 - 0.5 * (s1 * u + s2 * w));
        const double fx[4] = {s1, s1 * u + p, s2 * u, (s3 + p) * u};
        const double fy[4] = {s2, s1 * w, s2 * w + p, (s3 + p) * w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          fxi4[k] = gA * fx[k] + gB * fy[k];
          feta4[k] = gC * fx[k] + gD * fy[k];
        }
      }
      // fxi4[j], feta4[j]: variable v's reference fluxes at point q0 + j
      quad_transpose(fxi4);
      quad_transpose(feta4);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double fxi = fxi4[j], feta = feta4[j];
        const double *wx = vt + D::kVWx + (q0 + j) * Np;
        const double *wy = vt + D::kVWy + (q0 + j) * Np;
#pragma unroll
        for (int m = 0; m < Np; ++m) {
          if (!dubiner_dxi_zero(m))
            acc[m] = fma(wx[m], fxi, acc[m]);
          if (!dubiner_deta_zero(m))
            acc[m] = fma(wy[m], feta, acc[m]);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }

  // Faces: quad h takes edges h, h + H, ...; lane v builds both traces of
  // variable v at the edge points, lane q < NE evaluates HLLC at point q.
  // nv holds the neighbour coefficients of the edge; when next >= 0 the
  // coefficients of element next are fetched into it once the traces are
  // built, overlapping the flux and lifting work of this edge.
  auto face = [&](const int meta, const double nx, const double ny, const double scale,
                  double *nv, const int next) {
    const bool is_left = meta & kMetaLeft;
    const bool boundary = meta & kMetaBoundary;
    const int kl = meta_kl(meta), kr = meta_kr(meta), fl = meta_flip(meta);
    const double *bl = tab + (kl * NE * Np);
    const double *br = tab + (D::kBr - D::kEdge) + ((kl * 3 + kr) * 2 + fl) * NE * Np;
    const double *own = is_left ? bl : br;
    const double *oth = is_left ? br : bl;
    double si4[4] = {0.0, 0.0, 0.0, 0.0}, so4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      double si = 0.0;
#pragma unroll
      for (int m = 0; m < Np; ++m)
        si += c[m] * own[q * Np + m];
      // The interior trace always runs and a boundary face overwrites it: a
      // branch between the LDS tables and the global donor basis would
      // compile to flat loads.
      double so = 0.0;
#pragma unroll
      for (int m = 0; m < Np; ++m)
        so += nv[m] * oth[q * Np + m];
      if (boundary) {
        if (mesh.bdonor) {
          const int slot = meta_slot(meta);
          const int d = mesh.bdonor[slot * NE + q];
          const double *db = mesh.bbasis + (slot * NE + q) * Np;
          double sd = 0.0;
#pragma unroll
          for (int m = 0; m < Np; ++m)
            sd += at_idx(a.uin, static_cast<unsigned>((v * Np + m) * E + d)) * db[m];
          so = sd;
        } else {
          so = si;  // transmissive: exterior = own trace
        }
      }
      si4[q] = si;
      so4[q] = so;
    }
    if (next >= 0) {
#pragma unroll
      for (int m = 0; m < Np; ++m)
        nv[m] = at_idx(a.uin, static_cast<unsigned>((v * Np + m) * E + next));
    }
    // after the transposes lane v holds all four variables at edge point v
    quad_transpose(si4);
    quad_transpose(so4);
    __builtin_amdgcn_sched_barrier(0);
    double fv[4] = {0.0, 0.0, 0.0, 0.0};
    if (v < NE) {
      const C4 si = {si4[0], si4[1], si4[2], si4[3]};
      const C4 so = {so4[0], so4[1], so4[2], so4[3]};
      const C4 L = is_left ? si : so;
      const C4 R = is_left ? so : si;
      const C4 f = hllc(g, L, R, nx, ny);
      const double wq = tab[(D::kWe - D::kEdge) + kl * NE + v] * scale;
      const double sw = is_left ? -wq : wq;
      fv[0] = f.r * sw;
      fv[1] = f.mx * sw;
      fv[2] = f.my * sw;
      fv[3] = f.E * sw;
    }
    // fv[q]: variable v's weighted flux at edge point q
    quad_transpose(fv);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      const double val = fv[q];
#pragma unroll
      for (int m = 0; m < Np; ++m)
        acc[m] += val * own[q * Np + m];
    }
  };
  if constexpr (KP == KE) {
#pragma unroll
    for (int i = 0; i < KE; ++i)
      if (h + i * H < 3)
        face(emeta[i], enx[i], eny[i], escale[i], env[i], -1);
  } else {
    static_assert(KP == 1 && KE == 3, "pipelined faces assume one quad per element");
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {
      const int meta = i == 0 ? emeta[0] : i == 1 ? emeta[1] : emeta[2];
      const double nx = i == 0 ? enx[0] : i == 1 ? enx[1] : enx[2];
      const double ny = i == 0 ? eny[0] : i == 1 ? eny[1] : eny[2];
      const double sc = i == 0 ? escale[0] : i == 1 ? escale[1] : escale[2];
      face(meta, nx, ny, sc, env[0], i == 0 ? esrc[1] : i == 1 ? esrc[2] : -1);
    }
  }

  if constexpr (H == 2) {
#pragma unroll
    for (int m = 0; m < Np; ++m)
      acc[m] += __shfl_xor(acc[m], 4, 64);
  }

  // Finalise variable v (quad 0 of the element stores).
  double y[Np];
#pragma unroll
  for (int m = 0; m < Np; ++m) {
    const double r = acc[m] * invdet;
    const int idx = (v * Np + m) * E + e;
    if (a.residual_only) {
      y[m] = r;
    } else {
      y[m] = a.cb * (c[m] + dt * r);
      if (a.ca != 0.0)
        y[m] += a.ca * un[m];
    }
    if (valid && h == 0)
      a.out[idx] = y[m];
  }

  // New cell averages to every lane of the quad.
  const double o0 = dpp<quad_bcast<0>()>(y[0]);
  const double o1 = dpp<quad_bcast<1>()>(y[0]);
  const double o2 = dpp<quad_bcast<2>()>(y[0]);
  const double o3 = dpp<quad_bcast<3>()>(y[0]);

  if (a.sensor && !a.residual_only) {
    bool troubled = false;
    if (v == 0 && h == 0 && valid) {
      // Modal-decay sensor on density (euler_modal_limiter.cpp:18-24).
      double total = 0.0, high = 0.0;
#pragma unroll
      for (int m = 0; m < Np; ++m) {
        total += y[m] * y[m];
        if (m >= a.highest_begin)
          high += y[m] * y[m];
      }
      troubled = !(total == 0.0 || high / total <= a.threshold);
      if (a.flags)
        a.flags[e] = troubled ? 1 : 0;
    }
    // at most one entry per element and stage (the leader clears the count)
    list_append(&a.counters[a.stage], a.list, E, troubled, e);
  }

  if (a.want_lambda && !a.residual_only) {
    // |u| + c of the new cell average (cfl_timestep.cpp:33-47).
    double speed = 0.0;
    if (v == 0 && h == 0 && valid) {
      const double is2 = 1.0 / sqrt(2.0);
      const double r = is2 * o0, mx = is2 * o1, my = is2 * o2, en = is2 * o3;
      const double inv = 1.0 / r;
      const double u = mx * inv, w = my * inv;
      const double p = g1 * (en - 0.5 * (mx * u + my * w));
      if (!(r > 0.0 && p > 0.0)) {
        if (ctl)
          ctl->s[a.parity ^ 1].bad = 1;
      } else {
        speed = hypot(u, w) + sqrt(g * p / r);
      }
    }
    speed = wave_max(speed);
    if ((threadIdx.x & 63) == 0)
      red[threadIdx.x >> 6] = speed;
    __syncthreads();
    if (threadIdx.x == 0 && ctl) {
      const double mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
      speed_update(ctl, a.parity ^ 1, mx);
    }
  }
  stage_epilogue(ctl, a.parity, a.stage, dt, leader);
}

template <int P, int H>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void euler_stage_vs(EulerMesh mesh, StageArgs a) {
  __shared__ VsShared<P> sh;
  euler_stage_vs_body<P, H>(mesh, a, xcd_block(), blockIdx.x == 0, sh);
}

// -------------------------------------- 2D Euler, mixed precision (MFMA) ----
//
// The same fused stage with the contractions in FP32 on the matrix cores.
// A wave serves 16 consecutive elements, one per matrix column: lane l owns
// element j = l & 15 of the wave and row group g = l >> 4, and the four
// conserved variables are four products sharing one A operand. Operand
// layouts of v_mfma_f32_16x16x4f32 on gfx950:
//   A (16 x 4):  lane l holds A[l & 15][l >> 4]
//   B (4 x 16):  lane l holds B[l >> 4][l & 15]
//   D (16 x 16): lane l holds D[4 (l >> 4) + r][l & 15], r = 0..3
// and of v_mfma_f32_16x16x16bf16_1k: A[l & 15][4 (l >> 4) + i], B[4 (l >> 4) + i][l & 15]
// for i = 0..3, D as above.
//
//   states at points  S_v[q][j]  = sum_m B_vol[q][m] c_v[m][j]
//   volume term       R_v[m][j] += sum_q W_xi[q][m] F_xi,v[q][j] + W_eta[q][m] F_eta,v[q][j]
//   own edge traces   T_v[4k+s][j] = sum_m B_edge[k][s][m] c_v[m][j]
//   lifting           R_v[m][j] += sum_(k,s) B_edge[k][s][m] Fw_v[k][s][j]
// A point tile of S comes out of the first product in exactly the layout the
// second product takes as its B operand (K index = point 16t + 4g + r), and a
// lane holds all four variables of its points, so the pointwise flux is
// computed once per point with no exchange. Faces are evaluated in each
// element's own edge frame: Gauss points are symmetric, so the two sides of a
// face list the same physical points in the same or reversed order (meta bit
// 7). Row group k serves edge k: neighbour traces, HLLC and the edge weights
// run on the vector ALU in FP32, once per face point and side. The state stays
// FP64 in HBM; the RK combination, sensor, CFL speed and limiter are FP64.

constexpr int kMetaReverse = 128;   // the two sides order the face points oppositely

template <int P> struct Mx {
  static constexpr int Np = Tri<P>::Np, NQ = Tri<P>::NQ, NE = Tri<P>::NE;
  static constexpr int MT = (NQ + 15) / 16;   // point tiles of 16
  // Row 4g + r of point tile t is point 16t + 4r + g: the real points of the
  // last tile fill whole K steps r of the flux product, and the steps with no
  // real point (16t + 4r >= NQ) are left out.
  static constexpr int point(int t, int row) { return 16 * t + 4 * (row & 3) + (row >> 2); }
  static constexpr bool step_real(int t, int r) { return 16 * t + 4 * r < NQ; }
  // K steps over modes: row group g owns modes KS g .. KS g + KS - 1 and step
  // r takes mode KS g + r from it, so a lane feeds its own modes. Rows of the
  // residual products follow the same map (row 4g + r is mode KS g + r, rows
  // with r >= KS are padding), so the lane that holds a mode's operand also
  // receives its residual, the layout the epilogue needs.
  static constexpr int KS = (Np + 3) / 4;
  // mode of residual row `row`, Np for a padding row
  static constexpr int row_mode(int row) {
    return (row & 3) < KS && KS * (row >> 2) + (row & 3) < Np ? KS * (row >> 2) + (row & 3) : Np;
  }
  // Face points as (edge, point) pairs q = NE k + s; row group g serves
  // pairs PPG g .. PPG g + PPG - 1 (at P3 the twelve pairs fill all four row
  // groups; below P3 the last row group has none). Row 4g + t of the trace
  // products is pair PPG g + t; the pairs of a row group span at most two
  // edges.
  static constexpr int PAIRS = 3 * NE;
  static constexpr int PPG = (PAIRS + 3) / 4;
  // per-lane operand tables, floats, each [..][64]
  static constexpr int kAvol = 0;                         // [MT][KS]
  static constexpr int kAgrad = kAvol + MT * KS * 64;     // [MT][4 r][2 d]
  static constexpr int kAface = kAgrad + MT * 8 * 64;     // [KS]
  static constexpr int kAlift = kAface + KS * 64;         // [PPG]
  static constexpr int kBe = kAlift + PPG * 64;           // [3][NE][Np] own-edge basis
  static constexpr int kWe = kBe + 3 * NE * Np;           // [3][NE] edge weights
  // bf16 operands (pairs of bf16 per float slot), each lane 4 bf16 = 2 slots
  static constexpr int kBvol16 = (kWe + 3 * NE + 3) / 4 * 4;   // [MT][64][2], 16-byte aligned
  static constexpr int kBgrad16 = kBvol16 + MT * 128;     // [MT][2 d][64][2]
  static constexpr int kTable = kBgrad16 + MT * 256;
  static_assert(NE <= 4 && Np <= 16, "one edge per row group, modes in one tile");
};

typedef float mx_f4 __attribute__((ext_vector_type(4)));

// Waves per SIMD the mixed stage kernels are built for (register budget 512 /
// waves): every global read sits in registers from the start of the kernel.
// The deviation-form volume holds the cell average's state through the
// volume products; P1 fits 128 registers (four waves) and P2 and P3 168
// (three waves). Measured on MI355X at P2: three and four waves run alike
// (146.3 against 146.1 us at 524288 triangles), and the batched kernel
// spills at four.
template <int P, bool BF16> constexpr int mx_waves() {
  return P == 0 ? 5 : P == 1 ? 4 : 3;
}
typedef short mx_s4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mx_f4 mfma4(float a, float b, mx_f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Round to nearest even into bfloat16 (finite inputs).
__device__ __forceinline__ short to_bf16(float x) {
  const uint32_t u = __float_as_uint(x);
  return static_cast<short>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

template <int CTRL> __device__ __forceinline__ float dppf(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, false));
}

struct F4 {
  float r, mx, my, E;
};

__device__ __forceinline__ float rcpf(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return fmaf(fmaf(-x, r, 1.0f), r, r);
}

struct SideF {
  float inv_r, u, v, p, un, c;
};

__device__ __forceinline__ SideF side_of_f(float gamma, const F4 &s, float nx, float ny) {
  SideF o;
  o.inv_r = rcpf(s.r);
  o.u = s.mx * o.inv_r;
  o.v = s.my * o.inv_r;
  o.p = (gamma - 1.0f) * (s.E - 0.5f * (s.mx * o.u + s.my * o.v));
  o.un = o.u * nx + o.v * ny;
  // FP32 rounding can push a positive pressure just below zero near vacuum.
  // The hardware square root (within 1 ulp) is enough for a wave-speed bound
  // and skips the correctly rounded expansion (denormal scaling, corrections).
  o.c = __builtin_amdgcn_sqrtf(fmaxf(gamma * o.p * o.inv_r, 0.0f));
  return o;
}

__device__ __forceinline__ F4 normal_flux_f(const F4 &s, const SideF &w, float nx, float ny) {
  return {s.r * w.un, s.mx * w.un + w.p * nx, s.my * w.un + w.p * ny, (s.E + w.p) * w.un};
}

// HLLC (flux/hllc.hpp:40-74) in FP32.
__device__ inline F4 hllc_f(float gamma, const F4 &L, const F4 &R, float nx, float ny) {
  const SideF wl = side_of_f(gamma, L, nx, ny), wr = side_of_f(gamma, R, nx, ny);
  const float sl = fminf(wl.un - wl.c, wr.un - wr.c);
  const float sr = fmaxf(wl.un + wl.c, wr.un + wr.c);
  if (sl >= 0.0f)
    return normal_flux_f(L, wl, nx, ny);
  if (sr <= 0.0f)
    return normal_flux_f(R, wr, nx, ny);
  const float sm = (wr.p - wl.p + L.r * wl.un * (sl - wl.un) - R.r * wr.un * (sr - wr.un)) *
                   rcpf(L.r * (sl - wl.un) - R.r * (sr - wr.un));
  const bool left = sm >= 0.0f;
  const F4 &S = left ? L : R;
  const SideF &w = left ? wl : wr;
  const float s = left ? sl : sr;
  const F4 F = normal_flux_f(S, w, nx, ny);
  const float k = s - w.un;
  const float rs = S.r * k * rcpf(s - sm);
  const float shift = sm - w.un;
  const float es = S.E * w.inv_r + shift * (sm + w.p * w.inv_r * rcpf(k));
  return {F.r + s * (rs - S.r), F.mx + s * (rs * (w.u + shift * nx) - S.mx),
          F.my + s * (rs * (w.v + shift * ny) - S.my), F.E + s * (rs * es - S.E)};
}

__device__ __forceinline__ mx_s4 bf16x4(const mx_f4 &x) {
  return {to_bf16(x[0]), to_bf16(x[1]), to_bf16(x[2]), to_bf16(x[3])};
}

// Floats of the operand table a block stages in LDS.
template <int P, bool BF16> constexpr int mx_staged() {
  return BF16 ? Mx<P>::kTable : Mx<P>::kBvol16;
}

// One block (64 elements starting at block * 64) of the mixed stage; `leader`
// marks the block that advances the case's step control.
template <int P, bool BF16>
__device__ __forceinline__ void mx_stage(const EulerMesh mesh, const StageArgs a,
                                         const float *mxtab, int block, bool leader, float *tab,
                                         double *red) {
  using X = Mx<P>;
  constexpr int Np = X::Np, NQ = X::NQ, NE = X::NE, MT = X::MT, KS = X::KS;
  constexpr int PPG = X::PPG, PAIRS = X::PAIRS;
  constexpr int kStaged = mx_staged<P, BF16>();

  const int E = mesh.E;
  const int lane = threadIdx.x & 63;
  const int j = lane & 15, grp = lane >> 4;
  const int e_raw = block * 64 + (threadIdx.x >> 6) * 16 + j;
  const bool valid = e_raw < E;
  const int e = valid ? e_raw : E - 1;
  const float g = static_cast<float>(mesh.gamma), g1 = g - 1.0f;

  // The global reads are issued here, ahead of the step-control read; all
  // addresses are in bounds. The element's own modes 4 grp .. 4 grp + 3 of
  // each variable (FP64, kept for the RK combination), its metrics, the
  // record of edge `grp` and this thread's share of the operand table. The
  // neighbour's trace follows once its index is in, so the second round
  // trip overlaps the volume products.
  // State rows (v Np + r) are wave-uniform byte offsets; a lane adds its
  // row group's first mode (4 grp) and element (padded modes read row r of
  // the element and are zeroed).
  const unsigned Eu = static_cast<unsigned>(E);
  const size_t state_bytes = static_cast<size_t>(4 * Np) * Eu * 8;
  const Rsrc rin = rsrc(a.uin, state_bytes);
  const unsigned eoff = static_cast<unsigned>(e) * 8u;
  const unsigned goff = static_cast<unsigned>(KS * grp) * Eu * 8u + eoff;
  double c64[4][KS], un64[4][KS];
  float cf[4][KS];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int r = 0; r < KS; ++r) {
      const bool real = KS * grp + r < Np;
      const double x = ld_f64(rin, real ? goff : eoff, static_cast<unsigned>(v * Np + r) * Eu * 8u);
      c64[v][r] = real ? x : 0.0;
      cf[v][r] = static_cast<float>(c64[v][r]);
    }
  // the operand table in 16-byte pieces (its length is a multiple of 4)
  static_assert(kStaged % 4 == 0, "operand table in 16-byte pieces");
  constexpr int kPer = (kStaged / 4 + 255) / 256;
  float4 tpre[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int idx = static_cast<int>(threadIdx.x) + 256 * i;
    tpre[i] = idx < kStaged / 4 ? reinterpret_cast<const float4 *>(mxtab)[idx]
                                : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  const Rsrc rgeo = rsrc(mesh.geo, static_cast<size_t>(5) * Eu * 8);
  const double invdet = ld_f64(rgeo, eoff, 4u * Eu * 8u);
  const float gA = static_cast<float>(ld_f64(rgeo, eoff, 0u)),
              gB = static_cast<float>(ld_f64(rgeo, eoff, Eu * 8u)),
              gC = static_cast<float>(ld_f64(rgeo, eoff, 2u * Eu * 8u)),
              gD = static_cast<float>(ld_f64(rgeo, eoff, 3u * Eu * 8u));
  // The records of the (at most two) edges of this row group's face pairs;
  // a row group without pairs idles on pair 0.
  const int pfirst = PPG * grp < PAIRS ? PPG * grp : 0;
  const int plast = PPG * grp + PPG - 1 < PAIRS ? PPG * grp + PPG - 1 : pfirst;
  const int kedge[2] = {pfirst / NE, plast / NE};
  const Rsrc rmeta = rsrc(mesh.meta, static_cast<size_t>(3) * Eu * 4);
  const Rsrc rnbr = rsrc(mesh.nbr, static_cast<size_t>(3) * Eu * 4);
  const Rsrc rfn = rsrc(mesh.fn, static_cast<size_t>(9) * Eu * 8);
  // Up to P2 a row group's pairs are exactly one edge's points.
  constexpr bool kOneEdge = PPG == NE;
  int meta2[2], nbr2[2];
  float nx2[2], ny2[2], scale2[2];
#pragma unroll
  for (int h = 0; h < (kOneEdge ? 1 : 2); ++h) {
    const unsigned koff = static_cast<unsigned>(kedge[h]) * Eu + static_cast<unsigned>(e);
    meta2[h] = ld_i32(rmeta, koff * 4u, 0u);
    nbr2[h] = ld_i32(rnbr, koff * 4u, 0u);
    const unsigned foff = (static_cast<unsigned>(3 * kedge[h]) * Eu + static_cast<unsigned>(e)) * 8u;
    nx2[h] = static_cast<float>(ld_f64(rfn, foff, 0u));
    ny2[h] = static_cast<float>(ld_f64(rfn, foff, Eu * 8u));
    scale2[h] = static_cast<float>(ld_f64(rfn, foff, 2u * Eu * 8u));
  }

  RunCtl *ctl = a.ctl;
  double dt = a.dt;
  if (!stage_prologue(ctl, a.parity, a.stage, dt, leader))
    return;
  if (a.sensor && leader && threadIdx.x == 0)
    a.counters[(a.stage + 1) % 3] = 0;

#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int idx = static_cast<int>(threadIdx.x) + 256 * i;
    if (idx < kStaged / 4)
      reinterpret_cast<float4 *>(tab)[idx] = tpre[i];
  }
  __syncthreads();

  // The neighbour's trace at each of this row group's face pairs (16 bytes
  // each; the indices arrived with the first round trip): requested now, their
  // latency overlapping the volume products.
  float4 ntr[PPG];
  {
    const Rsrc rt = rsrc(a.tin, static_cast<size_t>(Eu) * 3 * 64);
#pragma unroll
    for (int t = 0; t < PPG; ++t) {
      const int q = PPG * grp + t < PAIRS ? PPG * grp + t : pfirst;
      const int h = kOneEdge || q / NE == kedge[0] ? 0 : 1;
      const int mt = meta2[h], sp = kOneEdge ? t : q - NE * (q / NE);
      const bool bnd = mt & kMetaBoundary, lft = mt & kMetaLeft;
      const int kn = lft ? meta_kr(mt) : meta_kl(mt);   // the neighbour's local edge
      const int sn = (mt & kMetaReverse) ? NE - 1 - sp : sp;
      const unsigned toff =
          (static_cast<unsigned>(bnd ? e : nbr2[h]) * 3u + static_cast<unsigned>(bnd ? 0 : kn)) *
              64u +
          16u * static_cast<unsigned>(bnd ? 0 : sn);
      ntr[t] = ld_f32x4(rt, toff, 0u);
    }
  }

  mx_f4 R[4];
#pragma unroll
  for (int v = 0; v < 4; ++v)
    R[v] = mx_f4{0.0f, 0.0f, 0.0f, 0.0f};

  // ---- volume: states at points (one product per variable, shared A),
  // pointwise fluxes once per point, flux contraction. The products run on
  // deviations from the cell average: the states at points come from the
  // higher modes only (the operand table leaves mode 0 out), and a point's
  // flux enters the contraction as its deviation from the flux of the
  // average, formed from the deviations so that FP32 rounds small numbers.
  // The constant flux's own volume term is added back in FP64 in the
  // epilogue (column sums of the weighted gradients).
  constexpr double kPhi0 = 0.70710678118654752440;   // mode 0 on the reference triangle
  if constexpr (P > 0) {
    float br, bmx, bmy, bu, bv, bh, buxi, bueta;
    {
      double m[4];
#pragma unroll
      for (int v = 0; v < 4; ++v)
        m[v] = __shfl(c64[v][0], j, 64) * kPhi0;
      const double inv = rcp(m[0]);
      const double u = m[1] * inv, w = m[2] * inv;
      const double pb = (mesh.gamma - 1.0) * (m[3] - 0.5 * (m[1] * u + m[2] * w));
      br = static_cast<float>(m[0]);
      bmx = static_cast<float>(m[1]);
      bmy = static_cast<float>(m[2]);
      bu = static_cast<float>(u);
      bv = static_cast<float>(w);
      bh = static_cast<float>(m[3] + pb);
      buxi = gA * bu + gB * bv;
      bueta = gC * bu + gD * bv;
    }
    mx_s4 cb16[4];
    if constexpr (BF16) {
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int c = 0; c < 4; ++c)   // K entries 4 grp + c; c >= KS is padding
          cb16[v][c] = c < KS ? to_bf16(cf[v][c < KS ? c : 0]) : short(0);
    }
#pragma unroll
    for (int t = 0; t < MT; ++t) {
      mx_f4 S[4];
#pragma unroll
      for (int v = 0; v < 4; ++v)
        S[v] = mx_f4{0.0f, 0.0f, 0.0f, 0.0f};
      if constexpr (BF16) {
        const mx_s4 av = *reinterpret_cast<const mx_s4 *>(tab + X::kBvol16 + (t * 64 + lane) * 2);
#pragma unroll
        for (int v = 0; v < 4; ++v)
          S[v] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, cb16[v], S[v], 0, 0, 0);
      } else {
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          const float av = tab[X::kAvol + (t * KS + s) * 64 + lane];
#pragma unroll
          for (int v = 0; v < 4; ++v)
            S[v] = mfma4(av, cf[v][s], S[v]);
        }
      }
      mx_f4 Fxi[4], Feta[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (!BF16 && !X::step_real(t, r))
          continue;   // no real point in this K step: the step is not issued
        const bool real = X::point(t, 4 * grp + r) < NQ;   // padded points: zero fluxes
        // deviations (r1, x1, y1, e1) of density, momenta and energy
        const float r1 = S[0][r], x1 = S[1][r], y1 = S[2][r], e1 = S[3][r];
        const float inv = rcpf(real ? br + r1 : 1.0f);
        const float du = (x1 - bu * r1) * inv, dv = (y1 - bv * r1) * inv;
        const float u = bu + du, w = bv + dv;
        const float dk = 0.5f * ((bmx * du + x1 * u) + (bmy * dv + y1 * w));
        const float dp = g1 * (e1 - dk), de = e1 + dp;
        // reference-frame flux deviations through the contravariant
        // velocities U_xi = A u + B v, U_eta = C u + D v and their deviations
        const float duxi = gA * du + gB * dv, dueta = gC * du + gD * dv;
        const float uxi = buxi + duxi, ueta = bueta + dueta;
        const float fxi[4] = {br * duxi + r1 * uxi, bmx * duxi + x1 * uxi + gA * dp,
                              bmy * duxi + y1 * uxi + gB * dp, bh * duxi + de * uxi};
        const float feta[4] = {br * dueta + r1 * ueta, bmx * dueta + x1 * ueta + gC * dp,
                               bmy * dueta + y1 * ueta + gD * dp, bh * dueta + de * ueta};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          Fxi[v][r] = real ? fxi[v] : 0.0f;
          Feta[v][r] = real ? feta[v] : 0.0f;
        }
      }
      if constexpr (BF16) {
        const mx_s4 ax =
            *reinterpret_cast<const mx_s4 *>(tab + X::kBgrad16 + ((t * 2 + 0) * 64 + lane) * 2);
        const mx_s4 ae =
            *reinterpret_cast<const mx_s4 *>(tab + X::kBgrad16 + ((t * 2 + 1) * 64 + lane) * 2);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          R[v] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ax, bf16x4(Fxi[v]), R[v], 0, 0, 0);
          R[v] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ae, bf16x4(Feta[v]), R[v], 0, 0, 0);
        }
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (!X::step_real(t, r))
            continue;
          const float ax = tab[X::kAgrad + ((t * 4 + r) * 2 + 0) * 64 + lane];
          const float ae = tab[X::kAgrad + ((t * 4 + r) * 2 + 1) * 64 + lane];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            R[v] = mfma4(ax, Fxi[v][r], R[v]);
            R[v] = mfma4(ae, Feta[v][r], R[v]);
          }
        }
      }
    }
  }

  // U^n of the owned outputs (stages 2 and 3): requested here, while the
  // faces are evaluated, rather than held through the volume phase, and used
  // only in the epilogue, so that its latency overlaps the face work.
  if (a.ca != 0.0) {
    const Rsrc run = rsrc(a.un, state_bytes);
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int r = 0; r < KS; ++r) {
        const bool real = KS * grp + r < Np;
        const double x =
            ld_f64(run, real ? goff : eoff, static_cast<unsigned>(v * Np + r) * Eu * 8u);
        un64[v][r] = real ? x : 0.0;
      }
  } else {
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int r = 0; r < KS; ++r)
        un64[v][r] = 0.0;
  }

  // ---- faces: own traces (rows 4k + s) on the matrix cores, bit-identical to
  // the traces this element's neighbours read; HLLC once per face point and
  // side on the vector ALU; lifting on the matrix cores.
  mx_f4 T[4];
#pragma unroll
  for (int v = 0; v < 4; ++v)
    T[v] = mx_f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const float af = tab[X::kAface + s * 64 + lane];
#pragma unroll
    for (int v = 0; v < 4; ++v)
      T[v] = mfma4(af, cf[v][s], T[v]);
  }

#pragma unroll
  for (int t = 0; t < PPG; ++t) {
    const bool pair = PPG * grp + t < PAIRS;
    const int q = pair ? PPG * grp + t : pfirst;
    const int h = kOneEdge || q / NE == kedge[0] ? 0 : 1;
    const int meta = meta2[h], sp = kOneEdge ? t : q - NE * (q / NE);
    const bool boundary = meta & kMetaBoundary;
    const bool is_left = meta & kMetaLeft;
    const int kl = meta_kl(meta);
    const F4 own = {T[0][t], T[1][t], T[2][t], T[3][t]};
    F4 oth = own;   // transmissive: exterior = own trace
    if (!boundary) {
      oth = {ntr[t].x, ntr[t].y, ntr[t].z, ntr[t].w};
    } else if (mesh.bdonor) {
      const int slot = meta_slot(meta);
      const int d = mesh.bdonor[slot * NE + sp];
      const double *db = mesh.bbasis + (slot * NE + sp) * Np;
      const unsigned doff = static_cast<unsigned>(d) * 8u;
      // The donor basis and a variable's modes (half of them from P3) are
      // requested together, so a wave waits once per group, not per load.
      float bm[Np];
#pragma unroll
      for (int m = 0; m < Np; ++m)
        bm[m] = static_cast<float>(db[m]);
      float x[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      constexpr int kChunk = Np > 6 ? (Np + 1) / 2 : Np;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
#pragma unroll
        for (int m0 = 0; m0 < Np; m0 += kChunk) {
          double cv[kChunk];
#pragma unroll
          for (int i = 0; i < kChunk; ++i)
            cv[i] = m0 + i < Np
                        ? ld_f64(rin, doff, static_cast<unsigned>(v * Np + m0 + i) * Eu * 8u)
                        : 0.0;
#pragma unroll
          for (int i = 0; i < kChunk; ++i)
            if (m0 + i < Np)
              x[v] += static_cast<float>(cv[i]) * bm[m0 + i];
        }
      }
      oth = {x[0], x[1], x[2], x[3]};
    }
    const F4 f = hllc_f(g, is_left ? own : oth, is_left ? oth : own, nx2[h], ny2[h]);
    const float wq = tab[X::kWe + kl * NE + sp] * scale2[h];
    const float sw = is_left ? -wq : wq;
    // A row group without this pair: its (meaningless) flux is selected away,
    // not multiplied by zero, since a NaN would survive the product and
    // spread through the matrix core to every mode of the element.
    const float fw[4] = {pair ? f.r * sw : 0.0f, pair ? f.mx * sw : 0.0f,
                         pair ? f.my * sw : 0.0f, pair ? f.E * sw : 0.0f};
    const float al = tab[X::kAlift + t * 64 + lane];
#pragma unroll
    for (int v = 0; v < 4; ++v)
      R[v] = mfma4(al, fw[v], R[v]);
  }

  // The volume term of the cell average's flux, FP64: F_xi(avg) sum_q W_xi[q][m]
  // + F_eta(avg) sum_q W_eta[q][m] for this lane's modes.
  double vmean[4][KS];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int r = 0; r < KS; ++r)
      vmean[v][r] = 0.0;
  if constexpr (P > 0) {
    double m[4];
#pragma unroll
    for (int v = 0; v < 4; ++v)
      m[v] = __shfl(c64[v][0], j, 64) * kPhi0;
    const double dA = ld_f64(rgeo, eoff, 0u), dB = ld_f64(rgeo, eoff, Eu * 8u),
                 dC = ld_f64(rgeo, eoff, 2u * Eu * 8u), dD = ld_f64(rgeo, eoff, 3u * Eu * 8u);
    const double *cs = mesh.table + Tri<P>::kCs;
    double csx[KS], csy[KS];
#pragma unroll
    for (int r = 0; r < KS; ++r) {
      const int mm = KS * grp + r < Np ? KS * grp + r : 0;
      csx[r] = cs[mm];
      csy[r] = cs[Np + mm];
    }
    const double inv = rcp(m[0]);
    const double u = m[1] * inv, w = m[2] * inv;
    const double pb = (mesh.gamma - 1.0) * (m[3] - 0.5 * (m[1] * u + m[2] * w));
    const double uxi = dA * u + dB * w, ueta = dC * u + dD * w, h = m[3] + pb;
    const double fxi[4] = {m[0] * uxi, m[1] * uxi + dA * pb, m[2] * uxi + dB * pb, h * uxi};
    const double feta[4] = {m[0] * ueta, m[1] * ueta + dC * pb, m[2] * ueta + dD * pb, h * ueta};
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int r = 0; r < KS; ++r)
        vmean[v][r] = fxi[v] * csx[r] + feta[v] * csy[r];
  }

  // ---- finalise: lane holds R[v][row 4 grp + r] of element j, mode KS grp + r
  // for r < KS; FP64 from here
  // Every output value is formed before the first store: a store issued
  // between the waits for U^n would make the next wait include it.
  float yf[4][KS];
  double xo[4][KS];
  double avg[4] = {0.0, 0.0, 0.0, 0.0};
  double total = 0.0, high = 0.0;
  const Rsrc rout = rsrc(a.out, state_bytes);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
#pragma unroll
    for (int r = 0; r < KS; ++r) {
      const int m = KS * grp + r;
      yf[v][r] = 0.0f;
      xo[v][r] = 0.0;
      if (m < Np) {
        const double res = (static_cast<double>(R[v][r]) + vmean[v][r]) * invdet;
        double x;
        if (a.residual_only) {
          x = res;
        } else {
          x = a.cb * (c64[v][r] + dt * res);
          if (a.ca != 0.0)
            x += a.ca * un64[v][r];
        }
        xo[v][r] = x;
        yf[v][r] = static_cast<float>(x);
        if (r == 0)
          avg[v] = x;
        if (v == 0) {
          total += x * x;
          if (m >= a.highest_begin)
            high += x * x;
        }
      }
    }
  }
  if (valid) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int r = 0; r < KS; ++r)
        if (KS * grp + r < Np)
          st_f64(xo[v][r], rout, goff, static_cast<unsigned>(v * Np + r) * Eu * 8u);
  }

  if (a.tout) {
    // FP32 edge traces of the new state for the next stage's neighbours: the
    // same product, operands and order as the own traces above.
    mx_f4 Tn[4];
#pragma unroll
    for (int v = 0; v < 4; ++v)
      Tn[v] = mx_f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const float af = tab[X::kAface + s * 64 + lane];
#pragma unroll
      for (int v = 0; v < 4; ++v)
        Tn[v] = mfma4(af, yf[v][s], Tn[v]);
    }
    if (valid) {
      const Rsrc rt = rsrc(a.tout, static_cast<size_t>(Eu) * 3 * 64);
#pragma unroll
      for (int t = 0; t < PPG; ++t) {
        const int q = PPG * grp + t;   // pair: edge q / NE, point q % NE
        if (q < PAIRS)
          st_f32x4(make_float4(Tn[0][t], Tn[1][t], Tn[2][t], Tn[3][t]), rt,
                   (static_cast<unsigned>(e) * 3u + static_cast<unsigned>(q / NE)) * 64u +
                       16u * static_cast<unsigned>(q % NE),
                   0u);
      }
    }
  }

  if (a.sensor && !a.residual_only) {
    // Modal-decay sensor on density (euler_modal_limiter.cpp:18-24): the
    // element's density modes sit in its four row groups.
    total += __shfl_xor(total, 16, 64);
    high += __shfl_xor(high, 16, 64);
    total += __shfl_xor(total, 32, 64);
    high += __shfl_xor(high, 32, 64);
    bool troubled = false;
    if (grp == 0 && valid) {
      troubled = !(total == 0.0 || high / total <= a.threshold);
      if (a.flags)
        a.flags[e] = troubled ? 1 : 0;
    }
    // at most one entry per element and stage (the leader clears the count)
    list_append(&a.counters[a.stage], a.list, E, troubled, e);
  }

  if (a.want_lambda && !a.residual_only) {
    // |u| + c of the new cell average (cfl_timestep.cpp:33-47), FP64.
    double speed = 0.0;
    if (grp == 0 && valid) {
      const double is2 = 1.0 / sqrt(2.0);
      const double r = is2 * avg[0], mx = is2 * avg[1], my = is2 * avg[2], en = is2 * avg[3];
      const double inv = 1.0 / r;
      const double u = mx * inv, w = my * inv;
      const double p = (mesh.gamma - 1.0) * (en - 0.5 * (mx * u + my * w));
      if (!(r > 0.0 && p > 0.0)) {
        if (ctl)
          ctl->s[a.parity ^ 1].bad = 1;
      } else {
        speed = hypot(u, w) + sqrt(mesh.gamma * p / r);
      }
    }
    speed = wave_max(speed);
    if ((threadIdx.x & 63) == 0)
      red[threadIdx.x >> 6] = speed;
    __syncthreads();
    if (threadIdx.x == 0 && ctl) {
      const double mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
      speed_update(ctl, a.parity ^ 1, mx);
    }
  }
  stage_epilogue(ctl, a.parity, a.stage, dt, leader);
}

// FP32 edge traces [E][3][4][4] of a state, by the product the stage epilogue
// uses (bit-identical to it): at the start of a run and for a state that
// arrives from the host.
template <int P>
__device__ __forceinline__ void mx_traces_body(const double *U, int E, float *T,
                                               const float *mxtab, int block) {
  using X = Mx<P>;
  constexpr int Np = X::Np, KS = X::KS;
  const int lane = threadIdx.x & 63;
  const int j = lane & 15, grp = lane >> 4;
  const int e_raw = block * 64 + (threadIdx.x >> 6) * 16 + j;
  const int e = e_raw < E ? e_raw : E - 1;
  mx_f4 Tn[4];
#pragma unroll
  for (int v = 0; v < 4; ++v)
    Tn[v] = mx_f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const float af = mxtab[X::kAface + s * 64 + lane];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = KS * grp + s;
      const float c = m < Np ? static_cast<float>(U[(v * Np + m) * E + e]) : 0.0f;
      Tn[v] = mfma4(af, c, Tn[v]);
    }
  }
  if (e_raw < E) {
#pragma unroll
    for (int t = 0; t < X::PPG; ++t) {
      const int q = X::PPG * grp + t;   // pair: edge q / NE, point q % NE
      if (q < X::PAIRS)
        reinterpret_cast<float4 *>(T + (static_cast<size_t>(e) * 3 + q / X::NE) * 16)[q % X::NE] =
            make_float4(Tn[0][t], Tn[1][t], Tn[2][t], Tn[3][t]);
    }
  }
}

template <int P>
__global__ __launch_bounds__(256) void mx_traces(const double *U, int E, float *T,
                                                 const float *mxtab) {
  mx_traces_body<P>(U, E, T, mxtab, blockIdx.x);
}

template <int P, bool BF16>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(mx_waves<P, BF16>())))
void euler_stage_mx(EulerMesh mesh, StageArgs a, const float *mxtab) {
  __shared__ __align__(16) float tab[mx_staged<P, BF16>()];
  __shared__ double red[4];
  mx_stage<P, BF16>(mesh, a, mxtab, xcd_block(), blockIdx.x == 0, tab, red);
}

// ------------------------------------------------------------ batching ----
//
// Several resident cases of one class (equation, order, precision) advance in
// one launch per stage: a tile table lists (case, block) pairs (case -1:
// padding, the workgroup exits; a table padded to a capacity keeps its launch
// shape while cases come and go), each case
// carries its own buffers and step control, so every case takes its own CFL
// dt and stops at its own end time. A tile computes exactly what the
// single-case kernel computes for that block, so a case's result does not
// depend on which other cases share the launch.
struct BatchCase {
  EulerMesh mesh;              // 2D Euler geometry and edge records
  double *u, *w1, *w2;         // state and the two stage buffers
  float *tu, *tw1, *tw2;       // their FP32 edge traces (mixed precision)
  RunCtl *ctl;
  int *list, *counters;        // troubled elements (limiting)
  int sensor;
  double threshold;
  int highest_begin;
};

struct BatchStage {
  int stage, parity;
};

// A pointer read from memory (a batch's case table) is generic to the
// compiler: every access through it becomes a flat instruction, and a flat
// load waits for LDS and vector memory alike (stores included). These casts
// state that the case buffers are global memory.
template <typename T> __device__ __forceinline__ T *global_ptr(T *p) {
  using G = __attribute__((address_space(1))) T;
  return (T *)((G *)reinterpret_cast<uintptr_t>(p));
}

__device__ __forceinline__ EulerMesh global_mesh(EulerMesh m) {
  m.geo = global_ptr(m.geo);
  m.nbr = global_ptr(m.nbr);
  m.meta = global_ptr(m.meta);
  m.fn = global_ptr(m.fn);
  m.bdonor = global_ptr(m.bdonor);
  m.bbasis = global_ptr(m.bbasis);
  m.table = global_ptr(m.table);
  m.vtable = global_ptr(m.vtable);
  return m;
}

// Stage buffers and SSP-RK3 (Shu-Osher) weights of one stage: stage s reads
// {u, w1, w2}[s] and writes {w1, w2, u}[s], out = ca U^n + cb (in + dt R).
__device__ __forceinline__ StageArgs stage_args(double *u, double *w1, double *w2, RunCtl *ctl,
                                                BatchStage b) {
  StageArgs a{};
  u = global_ptr(u);
  w1 = global_ptr(w1);
  w2 = global_ptr(w2);
  a.uin = b.stage == 0 ? u : b.stage == 1 ? w1 : w2;
  a.out = b.stage == 0 ? w1 : b.stage == 1 ? w2 : u;
  a.un = u;
  a.ca = b.stage == 0 ? 0.0 : b.stage == 1 ? 0.75 : 1.0 / 3.0;
  a.cb = b.stage == 0 ? 1.0 : b.stage == 1 ? 0.25 : 2.0 / 3.0;
  a.ctl = global_ptr(ctl);
  a.parity = b.parity;
  a.stage = b.stage;
  return a;
}

__device__ __forceinline__ StageArgs batch_args(const BatchCase &c, BatchStage b) {
  StageArgs a = stage_args(c.u, c.w1, c.w2, c.ctl, b);
  a.want_lambda = b.stage == 2 ? 1 : 0;
  a.sensor = c.sensor;
  a.threshold = c.threshold;
  a.highest_begin = c.highest_begin;
  a.list = global_ptr(c.list);
  a.counters = global_ptr(c.counters);
  a.tin = global_ptr(b.stage == 0 ? c.tu : b.stage == 1 ? c.tw1 : c.tw2);
  a.tout = global_ptr(b.stage == 0 ? c.tw1 : b.stage == 1 ? c.tw2 : c.tu);
  return a;
}

template <int P, bool BF16>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(mx_waves<P, BF16>())))
void euler_stage_mx_batch(const BatchCase *cases, const int2 *tiles, BatchStage b,
                          const float *mxtab) {
  __shared__ __align__(16) float tab[mx_staged<P, BF16>()];
  __shared__ double red[4];
  const int2 t = tiles[xcd_block()];
  if (t.x < 0)
    return;   // padding tile
  const BatchCase &c = cases[t.x];
  mx_stage<P, BF16>(global_mesh(c.mesh), batch_args(c, b), mxtab, t.y, t.y == 0, tab, red);
}

// FP64 point split over a batch: tiles index workgroups of 256 / LANES
// elements, the blocks of the single-case kernel with the same tile.
template <int P, int LANES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void euler_stage_batch(const BatchCase *cases, const int2 *tiles, BatchStage b) {
  const int2 t = tiles[xcd_block()];
  if (t.x < 0)
    return;   // padding tile
  const BatchCase &c = cases[t.x];
  euler_stage_body<P, LANES>(global_mesh(c.mesh), batch_args(c, b), t.y, t.y == 0);
}

// FP64 variable split over a batch (H quads per element, 64 / H elements per
// workgroup).
template <int P, int H>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void euler_stage_vs_batch(const BatchCase *cases, const int2 *tiles, BatchStage b) {
  __shared__ VsShared<P> sh;
  const int2 t = tiles[xcd_block()];
  if (t.x < 0)
    return;   // padding tile
  const BatchCase &c = cases[t.x];
  euler_stage_vs_body<P, H>(global_mesh(c.mesh), batch_args(c, b), t.y, t.y == 0, sh);
}

// ------------------------------------------------- 1D scalar advection ----

struct LineMeshDev {
  int E;
  const double *inv_jac;   // [E] 1 / (h/2)
  const double *table;     // Line<P>::kTable
  double wavespeed;
  int left_kind, right_kind;  // periodic / inflow / outflow
  double left_value, right_value;
};

__device__ __forceinline__ double upwind(double a, double l, double r) {
  return a >= 0.0 ? a * l : a * r;
}

// One lane per element, state U[element][mode]. Face fluxes are those of the
// reference's face loop (dg/scalar_advection/line_residual.hpp:76-103) with
// exteriors from make_line_boundary_exterior (line_boundary_conditions.hpp:36-57).
template <int P>
__device__ __forceinline__ void line_body(const LineMeshDev mesh, const StageArgs a, int block,
                                          bool leader, double *tab) {
  using D = Line<P>;
  constexpr int Np = D::Np, NQ = D::NQ;
  const int E = mesh.E;
  const int e_raw = block * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
  const bool valid = e_raw < E;
  const int e = valid ? e_raw : E - 1;
  // Own coefficients first, ahead of the step-control read. The neighbours'
  // traces come from the adjacent lanes (the cells run left to right, one
  // per lane); only a wave's end lanes and the end cells read a neighbour
  // from memory, requested here too.
  double c[Np];
#pragma unroll
  for (int m = 0; m < Np; ++m)
    c[m] = a.uin[e * Np + m];
  const int lane = static_cast<int>(threadIdx.x) & 63;
  const bool far_l = lane == 0 || e == 0;
  const bool far_r = lane == 63 || e_raw >= E - 1;
  double cl[Np], cr[Np];
  if (far_l) {
    const int en = e > 0 ? e - 1 : E - 1;
#pragma unroll
    for (int m = 0; m < Np; ++m)
      cl[m] = a.uin[en * Np + m];
  }
  if (far_r) {
    const int ep = e < E - 1 ? e + 1 : 0;
#pragma unroll
    for (int m = 0; m < Np; ++m)
      cr[m] = a.uin[ep * Np + m];
  }
  double dt = a.dt;
  if (!stage_prologue(a.ctl, a.parity, a.stage, dt, leader))
    return;
  for (int i = threadIdx.x; i < D::kTable; i += blockDim.x)
    tab[i] = mesh.table[i];
  __syncthreads();
  const double sa = mesh.wavespeed;
  double acc[Np];
#pragma unroll
  for (int m = 0; m < Np; ++m)
    acc[m] = 0.0;
  const double *bl = tab + D::kL, *br = tab + D::kR;
  if constexpr (P > 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double u = 0.0;
#pragma unroll
      for (int m = 0; m < Np; ++m)
        u += c[m] * tab[D::kB + q * Np + m];
      const double f = sa * u;
#pragma unroll
      for (int m = 0; m < Np; ++m)
        acc[m] += f * tab[D::kW + q * Np + m];
    }
  }
  const auto trace = [&](const double (&x)[Np], const double *b) {
    double t = 0.0;
#pragma unroll
    for (int m = 0; m < Np; ++m)
      t += x[m] * b[m];
    return t;
  };
  const double own_l = trace(c, bl), own_r = trace(c, br);
  // the left neighbour's right trace and the right neighbour's left trace:
  // the same bits the neighbour evaluates for its own faces
  double ext_l = __shfl_up(own_r, 1, 64), ext_r = __shfl_down(own_l, 1, 64);
  if (far_l)
    ext_l = e > 0 ? trace(cl, br)
                  : mesh.left_kind == 1   ? mesh.left_value
                  : mesh.left_kind == 2 ? own_l
                                        : trace(cl, br);
  if (far_r)
    ext_r = e < E - 1 ? trace(cr, bl)
                      : mesh.right_kind == 1   ? mesh.right_value
                      : mesh.right_kind == 2 ? own_r
                                             : trace(cr, bl);
  const double fl = upwind(sa, ext_l, own_l);   // left face of e
  const double fr = upwind(sa, own_r, ext_r);   // right face of e
  if (valid) {
    const double ij = mesh.inv_jac[e];
    // U^n first, all of it: a read issued after a store would wait for that
    // store as well.
    double un[Np];
#pragma unroll
    for (int m = 0; m < Np; ++m)
      un[m] = (!a.residual_only && a.ca != 0.0) ? a.un[e * Np + m] : 0.0;
#pragma unroll
    for (int m = 0; m < Np; ++m) {
      const double r = (acc[m] + bl[m] * fl - br[m] * fr) * ij;
      double y;
      if (a.residual_only) {
        y = r;
      } else {
        y = a.cb * (c[m] + dt * r);
        if (a.ca != 0.0)
          y += a.ca * un[m];
      }
      a.out[e * Np + m] = y;
    }
  }
  stage_epilogue(a.ctl, a.parity, a.stage, dt, leader);
}

template <int P> __global__ __launch_bounds__(256) void line_stage(LineMeshDev mesh, StageArgs a) {
  __shared__ double tab[Line<P>::kTable];
  line_body<P>(mesh, a, blockIdx.x, blockIdx.x == 0, tab);
}

// Batched 1D advection: (case, block) tiles as for the Euler classes.
struct BatchLine {
  LineMeshDev mesh;
  double *u, *w1, *w2;
  RunCtl *ctl;
};

template <int P>
__global__ __launch_bounds__(256) void line_stage_batch(const BatchLine *cases, const int2 *tiles,
                                                        BatchStage b) {
  __shared__ double tab[Line<P>::kTable];
  const int2 t = tiles[blockIdx.x];
  if (t.x < 0)
    return;   // padding tile
  const BatchLine &c = cases[t.x];
  LineMeshDev m = c.mesh;
  m.inv_jac = global_ptr(m.inv_jac);
  m.table = global_ptr(m.table);
  line_body<P>(m, stage_args(c.u, c.w1, c.w2, c.ctl, b), t.y, t.y == 0, tab);
}

// ------------------------------------------- 2D scalar advection stage ----

// Reference-element tables of the triangle advection stage (doubles):
// differentiation matrices Sx[m][n] = sum_q w_q dphi_m/dxi(q) phi_n(q) and
// Sy (d/deta) over the order-p volume rule, then the edge tables laid out as
// in Tri<P> (own-side basis, neighbour-side basis, edge weights).
template <int P> struct TriAdv {
  static constexpr int Np = (P + 1) * (P + 2) / 2;
  static constexpr int NE = P + 1;
  static constexpr int kSx = 0;
  static constexpr int kSy = Np * Np;
  static constexpr int kBl = 2 * Np * Np;            // [3][NE][Np]
  static constexpr int kBr = kBl + 3 * NE * Np;      // [3][3][2][NE][Np]
  static constexpr int kWe = kBr + 18 * NE * Np;     // [3][NE]
  static constexpr int kTable = kWe + 3 * NE;
  // The edge tables alone (staged in LDS; the matrices are read from the
  // constant slot g_advvol[P] through the scalar cache).
  static constexpr int kEdge = kBl;
  static constexpr int kEdgeTable = kTable - kBl;
};

// Differentiation matrices Sx, Sy of order p in slot p (every engine of that
// order writes the same values at creation).
constexpr int kAdvVolMax = 2 * 10 * 10;
__constant__ double g_advvol[4][kAdvVolMax];

struct AdvectMesh {
  int E;
  const double *vel;      // [2][E]: (A vx + B vy) / detJ, (C vx + D vy) / detJ
  const int *nbr;         // [3][E] neighbour element or -1
  const int *meta;        // [3][E] edge records (as for Euler)
  const double *flux;     // [3][E]: (v . n) * edge scale / detJ, n the face normal (left side)
  const int *bdonor;      // [NB][NE] periodic donors
  const double *bbasis;   // [NB][NE][Np]
  const double *table;    // TriAdv<P>::kTable doubles
};

// One lane per element (dg/scalar_advection/triangle_residual.hpp:25-105 for
// a constant velocity). The device keeps the state as U[mode][element] so
// every coefficient load and store is coalesced over the elements of a wave;
// the engine converts from and to the reference's U[element][mode] at the
// host boundary.
// The volume term sum_q w_q u_q (v . grad phi_m) is linear in the
// coefficients: with the contravariant velocity (a_xi, a_eta) = detJ J^-1 v
// it is a_xi (Sx c)_m + a_eta (Sy c)_m. Faces: the upwind flux
// (v . n) u_upwind at the left side's edge points, lifted onto this
// element's own basis; 1/detJ is folded into the velocity and edge factors.
template <int P>
__device__ __forceinline__ void advect_body(const AdvectMesh &mesh, const StageArgs &a,
                                            const int block, const bool leader, double *tab) {
  using D = TriAdv<P>;
  constexpr int Np = D::Np, NE = D::NE;
  const int E = mesh.E;
  const int e_raw = block * 256 + static_cast<int>(threadIdx.x);
  const bool valid = e_raw < E;
  const int e = valid ? e_raw : E - 1;

  // Global requests ahead of the step-control read: own coefficients, the
  // contravariant velocity and the three edge records.
  double c[Np];
#pragma unroll
  for (int m = 0; m < Np; ++m)
    c[m] = at_row(a.uin, E, m, e);
  const double ax = at_row(mesh.vel, E, 0, e), ay = at_row(mesh.vel, E, 1, e);
  int meta[3], nb[3];
  double fe[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    meta[k] = at_row(mesh.meta, E, k, e);
    nb[k] = at_row(mesh.nbr, E, k, e);
    fe[k] = at_row(mesh.flux, E, k, e);
  }
  constexpr int kPer = (D::kEdgeTable + 255) / 256;
  double tpre[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int idx = static_cast<int>(threadIdx.x) + 256 * i;
    tpre[i] = idx < D::kEdgeTable ? mesh.table[D::kEdge + idx] : 0.0;
  }

  double dt = a.dt;
  if (!stage_prologue(a.ctl, a.parity, a.stage, dt, leader))
    return;

  // The neighbour's coefficients (own again across a boundary face, where
  // they go unused), requested one edge ahead of their use.
  const auto fetch = [&](int k, double (&dst)[Np]) {
    const int src = nb[k] >= 0 ? nb[k] : e;
#pragma unroll
    for (int m = 0; m < Np; ++m)
      dst[m] = at_row(a.uin, E, m, src);
  };
  double na[Np], nbk[Np];
  fetch(0, na);

#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int idx = static_cast<int>(threadIdx.x) + 256 * i;
    if (idx < D::kEdgeTable)
      tab[idx] = tpre[i];
  }
  __syncthreads();

  // Volume: matrix entries are wave-uniform, read into scalar registers.
  static_assert(2 * Np * Np <= kAdvVolMax, "advection matrix slot too small");
  const double *S = g_advvol[P];
  double acc[Np];
#pragma unroll
  for (int m = 0; m < Np; ++m) {
    double sx = 0.0, sy = 0.0;
    if constexpr (P > 0) {
#pragma unroll
      for (int n = 0; n < Np; ++n) {
        sx += S[D::kSx + m * Np + n] * c[n];
        sy += S[D::kSy + m * Np + n] * c[n];
      }
    }
    acc[m] = ax * sx + ay * sy;
  }

  // Upwind flux at the left side's edge points, lifted onto the own basis.
  const auto edge = [&](int mk, double fk, const double (&nc)[Np]) {
    const bool is_left = mk & kMetaLeft;
    const bool boundary = mk & kMetaBoundary;
    const int kl = meta_kl(mk), kr = meta_kr(mk), fl = meta_flip(mk);
    const double *bl = tab + (D::kBl - D::kEdge) + kl * NE * Np;
    const double *br = tab + (D::kBr - D::kEdge) + ((kl * 3 + kr) * 2 + fl) * NE * Np;
    const double *own = is_left ? bl : br;
    const double *oth = is_left ? br : bl;
    // one edge point at a time keeps the register budget spill-free
#pragma unroll 1
    for (int q = 0; q < NE; ++q) {
      double si = 0.0, so = 0.0;
#pragma unroll
      for (int m = 0; m < Np; ++m) {
        si += c[m] * own[q * Np + m];
        so += nc[m] * oth[q * Np + m];
      }
      if (boundary) {
        const int slot = meta_slot(mk);
        const int d = mesh.bdonor[slot * NE + q];
        const double *db = mesh.bbasis + (slot * NE + q) * Np;
        so = 0.0;
#pragma unroll
        for (int m = 0; m < Np; ++m)
          so += at_row(a.uin, E, m, d) * db[m];
      }
      const double L = is_left ? si : so, R = is_left ? so : si;
      const double f = fk * (fk >= 0.0 ? L : R);
      const double wq = tab[(D::kWe - D::kEdge) + kl * NE + q];
      const double sw = is_left ? -wq : wq;
#pragma unroll
      for (int m = 0; m < Np; ++m)
        acc[m] += sw * own[q * Np + m] * f;
    }
  };
  fetch(1, nbk);
  edge(meta[0], fe[0], na);
  fetch(2, na);
  edge(meta[1], fe[1], nbk);
  edge(meta[2], fe[2], na);

  // U^n first, all of it: a read issued after a store would wait for that
  // store as well.
  double un[Np];
#pragma unroll
  for (int m = 0; m < Np; ++m)
    un[m] = 0.0;
  if (!a.residual_only && a.ca != 0.0) {
#pragma unroll
    for (int m = 0; m < Np; ++m)
      un[m] = a.un[static_cast<size_t>(m) * static_cast<unsigned>(E) + e];
  }
#pragma unroll
  for (int m = 0; m < Np; ++m) {
    const size_t idx = static_cast<size_t>(m) * static_cast<unsigned>(E) + e;
    double y;
    if (a.residual_only) {
      y = acc[m];
    } else {
      y = a.cb * (c[m] + dt * acc[m]);
      if (a.ca != 0.0)
        y += a.ca * un[m];
    }
    if (valid)
      a.out[idx] = y;
  }
  stage_epilogue(a.ctl, a.parity, a.stage, dt, leader);
}

// Up to P2 four waves per SIMD fit spill-free in 128 registers; P3 takes three.
template <int P> constexpr int advect_waves() { return P >= 3 ? 3 : 4; }

template <int P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(advect_waves<P>())))
void advect_stage(AdvectMesh mesh, StageArgs a) {
  __shared__ double tab[TriAdv<P>::kEdgeTable];
  advect_body<P>(mesh, a, xcd_block(), blockIdx.x == 0, tab);
}

// --------------------------------------------------- 1D Euler stage ----

// Tables of the line Euler stage (doubles): basis and weighted derivative at
// the max(1, p+1) = p+1 Gauss points of the reference's volume rule
// (dg/euler/line_residual.hpp:29), and the end values phi(-1), phi(+1).
template <int P> struct LineE {
  static constexpr int Np = P + 1;
  static constexpr int NQ = P + 1;
  static constexpr int kB = 0;                 // [NQ][Np]
  static constexpr int kW = kB + NQ * Np;      // [NQ][Np] w_q dphi/dxi
  static constexpr int kL = kW + NQ * Np;      // [Np]
  static constexpr int kR = kL + Np;           // [Np]
  static constexpr int kTable = kR + Np;
};

struct EulerLineDev {
  int E;
  const double *inv_jac;      // [E] 1 / (h/2)
  const double *table;        // LineE<P>::kTable doubles
  double gamma;
  int periodic;               // 1: periodic ends; 0: outflow (own trace outside)
};

__device__ __forceinline__ C4 flux_x_of(double g1, const C4 &s) {
  const double inv = rcp(s.r);
  const double u = s.mx * inv, w = s.my * inv;
  const double p = g1 * (s.E - 0.5 * (s.mx * u + s.my * w));
  return {s.mx, s.mx * u + p, s.my * u, (s.E + p) * u};
}

// Cells a wave finalises: lanes 1-62. Lanes 0 and 63 hold the cells either
// side (wrapped at periodic ends, clamped at outflow ends), so every face an
// output cell needs is evaluated inside the wave.
constexpr int kLineCells = 62;

// One lane per element, state U[4][Np][E] (euler_line_residual,
// dg/euler/line_residual.hpp:23-86, with line_euler_exterior,
// src/boundaries/line_euler_exterior.cpp:13-43). The elements run left to
// right, so the faces of element e are shared with e - 1 and e + 1. Each lane
// evaluates the HLLC flux of its cell's left face, as (left, right, +x) from
// the left neighbour's right trace (the previous lane's) and its own left
// trace; its right face's flux is the next lane's. A wave covers kLineCells
// cells and the halo lane at either end. Epilogue as in the triangle stage:
// RK combination, the density sensor and the CFL speed |u_x| + c of the new
// cell average.
template <int P>
__device__ __forceinline__ void euler_line_body(const EulerLineDev mesh, const StageArgs a,
                                                int block, bool leader, double *tab, double *red) {
  using D = LineE<P>;
  constexpr int Np = D::Np, NQ = D::NQ;
  const int E = mesh.E;
  const int lane = static_cast<int>(threadIdx.x) & 63;
  const int wave = block * 4 + (static_cast<int>(threadIdx.x) >> 6);
  const int e_raw = wave * kLineCells + lane - 1;
  const bool valid = lane >= 1 && lane <= kLineCells && e_raw < E;
  // the lane's cell: wrapped (periodic) or clamped (outflow) into [0, E)
  const int e = mesh.periodic ? ((e_raw % E) + E) % E : (e_raw < 0 ? 0 : (e_raw < E ? e_raw : E - 1));
  const double g = mesh.gamma, g1 = g - 1.0;
  double c[4][Np];
  load_coefficients<Np>(a.uin, E, e, c);
  double dt = a.dt;
  if (!stage_prologue(a.ctl, a.parity, a.stage, dt, leader))
    return;
  if (a.sensor && leader && threadIdx.x == 0)
    a.counters[(a.stage + 1) % 3] = 0;
  for (int i = threadIdx.x; i < D::kTable; i += blockDim.x)
    tab[i] = mesh.table[i];
  __syncthreads();

  double acc[4][Np];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int m = 0; m < Np; ++m)
      acc[v][m] = 0.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const C4 f = flux_x_of(g1, trace_of<Np>(c, tab + D::kB + q * Np));
    const double fv[4] = {f.r, f.mx, f.my, f.E};
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int m = 0; m < Np; ++m)
        acc[v][m] += fv[v] * tab[D::kW + q * Np + m];
  }
  const double *bl = tab + D::kL, *br = tab + D::kR;
  const C4 own_l = trace_of<Np>(c, bl), own_r = trace_of<Np>(c, br);
  // The left face: the previous lane's right trace against this lane's left
  // trace. At outflow ends the exterior is the owner's own trace: cell 0's
  // left face is (own_l, own_l), and the face right of cell E - 1, evaluated
  // by the lane after it, is (its right trace, the same).
  C4 L = {__shfl_up(own_r.r, 1, 64), __shfl_up(own_r.mx, 1, 64), __shfl_up(own_r.my, 1, 64),
          __shfl_up(own_r.E, 1, 64)};
  C4 R = own_l;
  if (!mesh.periodic && e_raw == 0)
    L = own_l;
  if (!mesh.periodic && e_raw == E)
    R = L;
  const C4 fl = hllc(g, L, R, 1.0, 0.0);
  const C4 fr = {__shfl_down(fl.r, 1, 64), __shfl_down(fl.mx, 1, 64), __shfl_down(fl.my, 1, 64),
                 __shfl_down(fl.E, 1, 64)};
  const double vl[4] = {fl.r, fl.mx, fl.my, fl.E}, vr[4] = {fr.r, fr.mx, fr.my, fr.E};
  const double ij = mesh.inv_jac[e];
  // U^n first, all of it: a read issued after a store would wait for that
  // store as well.
  double un[4][Np];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int m = 0; m < Np; ++m)
      un[v][m] = (!a.residual_only && a.ca != 0.0) ? a.un[(v * Np + m) * E + e] : 0.0;
  double out0[4], dens[Np];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int m = 0; m < Np; ++m) {
      const double r = (acc[v][m] + bl[m] * vl[v] - br[m] * vr[v]) * ij;
      const int idx = (v * Np + m) * E + e;
      double y;
      if (a.residual_only) {
        y = r;
      } else {
        y = a.cb * (c[v][m] + dt * r);
        if (a.ca != 0.0)
          y += a.ca * un[v][m];
      }
      if (valid)
        a.out[idx] = y;
      if (m == 0)
        out0[v] = y;
      if (v == 0)
        dens[m] = y;
    }
  if (a.sensor && !a.residual_only) {
    bool troubled = false;
    if (valid) {
      // modal-decay sensor on density (euler_modal_limiter.cpp:18-24)
      double total = 0.0, high = 0.0;
#pragma unroll
      for (int m = 0; m < Np; ++m) {
        total += dens[m] * dens[m];
        if (m >= a.highest_begin)
          high += dens[m] * dens[m];
      }
      troubled = !(total == 0.0 || high / total <= a.threshold);
      if (a.flags)
        a.flags[e] = troubled ? 1 : 0;
    }
    list_append(&a.counters[a.stage], a.list, E, troubled, e);
  }
  if (a.want_lambda && !a.residual_only) {
    // |u_x| + c of the new cell average (cfl_timestep.cpp:33-47, 1D)
    double speed = 0.0;
    if (valid) {
      const double is2 = 1.0 / sqrt(2.0);
      const double r = is2 * out0[0], mx = is2 * out0[1], my = is2 * out0[2], en2 = is2 * out0[3];
      const double inv = 1.0 / r;
      const double u = mx * inv, w = my * inv;
      const double p = g1 * (en2 - 0.5 * (mx * u + my * w));
      if (!(r > 0.0 && p > 0.0)) {
        if (a.ctl)
          a.ctl->s[a.parity ^ 1].bad = 1;
      } else {
        speed = fabs(u) + sqrt(g * p / r);
      }
    }
    speed = wave_max(speed);
    if ((threadIdx.x & 63) == 0)
      red[threadIdx.x >> 6] = speed;
    __syncthreads();
    if (threadIdx.x == 0 && a.ctl) {
      const double mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
      speed_update(a.ctl, a.parity ^ 1, mx);
    }
  }
  stage_epilogue(a.ctl, a.parity, a.stage, dt, leader);
}

template <int P>
__global__ __launch_bounds__(256) void euler_line_stage(EulerLineDev mesh, StageArgs a) {
  __shared__ double tab[LineE<P>::kTable];
  __shared__ double red[4];
  euler_line_body<P>(mesh, a, blockIdx.x, blockIdx.x == 0, tab, red);
}

// ------------------------------------------------------- reductions ----

// CFL speed of the cell averages of U[4][Np][E] into ctl->s[parity].lambda
// (cfl_timestep.cpp:33-47); an unphysical average sets the slot's bad flag.
// dims 1: |u_x| + c (lines); dims 2: |u| + c (triangles).
__global__ __launch_bounds__(256) void cfl_speed(const double *U, int E, int Np, double gamma,
                                                 RunCtl *ctl, int parity, int dims) {
  __shared__ double red[4];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double speed = 0.0;
  if (e < E) {
    const double is2 = 1.0 / sqrt(2.0);
    const double r = is2 * U[e], mx = is2 * U[Np * E + e], my = is2 * U[2 * Np * E + e],
                 en = is2 * U[3 * Np * E + e];
    const double inv = 1.0 / r;
    const double u = mx * inv, w = my * inv;
    const double p = (gamma - 1.0) * (en - 0.5 * (mx * u + my * w));
    if (!(r > 0.0 && p > 0.0))
      ctl->s[parity].bad = 1;
    else
      speed = (dims == 1 ? fabs(u) : hypot(u, w)) + sqrt(gamma * p / r);
  }
  speed = wave_max(speed);
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = speed;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    atomicMax(&ctl->s[parity].lambda, static_cast<unsigned long long>(__double_as_longlong(mx)));
  }
}

// Neumaier-compensated sum of weight[e] * U[e * stride] in fixed order: each
// block sums a contiguous chunk, a second pass sums the block partials.
struct KSum {
  double s, c;
};
__device__ __forceinline__ void ksum_add(KSum &k, double x) {
  const double t = k.s + x;
  k.c += fabs(k.s) >= fabs(x) ? (k.s - t) + x : (x - t) + k.s;
  k.s = t;
}

__global__ __launch_bounds__(256) void mass_partials(const double *U, int stride,
                                                     const double *weight, int E,
                                                     double *partials) {
  __shared__ double ss[256], sc[256];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  KSum k{0.0, 0.0};
  if (e < E)
    ksum_add(k, weight[e] * U[static_cast<size_t>(e) * stride]);
  ss[threadIdx.x] = k.s;
  sc[threadIdx.x] = k.c;
  __syncthreads();
  if (threadIdx.x == 0) {
    KSum b{0.0, 0.0};
    for (int i = 0; i < blockDim.x; ++i) {
      ksum_add(b, ss[i]);
      ksum_add(b, sc[i]);
    }
    partials[2 * blockIdx.x] = b.s;
    partials[2 * blockIdx.x + 1] = b.c;
  }
}

__global__ void mass_finish(const double *partials, int blocks, double *result) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    KSum b{0.0, 0.0};
    for (int i = 0; i < blocks; ++i) {
      ksum_add(b, partials[2 * i]);
      ksum_add(b, partials[2 * i + 1]);
    }
    *result = b.s + b.c;
  }
}

// --------------------------------------------------------------- limiter ----

struct LimiterDev {
  int E;
  const int *nbr;        // [E][K]: K = 3 (triangles) or 2 (lines)
  const double *T;       // [E][K][Np][Np]
  const double *nrm;     // [E][2]
  const double *check;   // [NC][Np]
  double gamma, density_floor, pressure_floor;
};

// Standalone troubled-cell sensor over the Np density modes of an element
// (triangles or lines; one mode means P0, which is never troubled).
template <int Np>
__global__ __launch_bounds__(256) void sensor(const double *U, int E, double threshold,
                                              int highest_begin, uint8_t *flags, int *list,
                                              int *count) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E)
    return;
  double total = 0.0, high = 0.0;
#pragma unroll
  for (int m = 0; m < Np; ++m) {
    const double v = U[m * E + e];
    total += v * v;
    if (m >= highest_begin)
      high += v * v;
  }
  const bool troubled = Np > 1 && !(total == 0.0 || high / total <= threshold);
  if (flags)
    flags[e] = troubled ? 1 : 0;
  if (troubled)
    list[atomicAdd(count, 1)] = e;
}

// Characteristic frame frozen at a cell average (limiting/weno/euler_characteristic.hpp).
struct CharFrame {
  double nx, ny, un, ut, c, ke, factor, H;
};

__device__ inline CharFrame char_frame(double g, const double avg[4], double nx, double ny) {
  CharFrame f;
  const double inv = 1.0 / avg[0];
  const double u = avg[1] * inv, v = avg[2] * inv;
  const double p = (g - 1.0) * (avg[3] - 0.5 * (avg[1] * u + avg[2] * v));
  f.nx = nx;
  f.ny = ny;
  f.un = u * nx + v * ny;
  f.ut = -u * ny + v * nx;
  const double c2 = g * p / avg[0];
  f.c = sqrt(c2);
  f.ke = 0.5 * (f.un * f.un + f.ut * f.ut);
  f.factor = (g - 1.0) / c2;
  f.H = (avg[3] + p) / avg[0];
  return f;
}

__device__ __forceinline__ void to_char(const CharFrame &f, double r, double mx, double my,
                                        double E, double w[4]) {
  const double mn = mx * f.nx + my * f.ny;
  const double mt = -mx * f.ny + my * f.nx;
  const double common = f.factor * (f.ke * r - f.un * mn - f.ut * mt + E);
  const double diff = (f.un * r - mn) / f.c;
  w[0] = 0.5 * (common + diff);
  w[1] = r - common;
  w[2] = mt - f.ut * r;
  w[3] = 0.5 * (common - diff);
}

__device__ __forceinline__ void from_char(const CharFrame &f, const double w[4], double out[4]) {
  const double r = w[0] + w[1] + w[3];
  const double mn = (f.un - f.c) * w[0] + f.un * w[1] + (f.un + f.c) * w[3];
  const double mt = f.ut * r + w[2];
  out[0] = r;
  out[1] = f.nx * mn - f.ny * mt;
  out[2] = f.ny * mn + f.nx * mt;
  out[3] = (f.H - f.un * f.c) * w[0] + f.ke * w[1] + f.ut * w[2] + (f.H + f.un * f.c) * w[3];
}

__device__ __forceinline__ double pressure_of(double g, double r, double mx, double my, double E) {
  const double inv = 1.0 / r;
  const double u = mx * inv, v = my * inv;
  return (g - 1.0) * (E - 0.5 * (mx * u + my * v));
}

// w[v] for a run-time v without an indexed (scratch) access.
__device__ __forceinline__ double pick4(const double (&w)[4], int v) {
  return v == 0 ? w[0] : v == 1 ? w[1] : v == 2 ? w[2] : w[3];
}

__device__ __forceinline__ double wave_min(double x) {
  for (int o = 32; o > 0; o >>= 1)
    x = fmin(x, __shfl_xor(x, o, 64));
  return x;
}

// Troubled elements from the list: characteristic simple WENO against the
// transformed neighbours (all read from the unlimited state), then the
// Zhang-Shu positivity limiter in FP64 (euler_modal_limiter.cpp:81-97,
// euler_characteristic_weno.hpp:17-67, simple_weno.hpp:21-48,
// euler_positivity.hpp:63-102). Results go to side[idx] for scatter().
//
// One wave per troubled element: lane (v, m) owns coefficient m of variable
// v, lane q < NC one positivity check point. Every sum keeps the reference's
// sequential order (one lane walks it through LDS) and minima are exact; the
// modes, the candidates and the 64-step bisections of the check points
// proceed side by side. Blocks of 64 threads take list entries first,
// first + stride, ...
template <int Np, int NC, int KMAX>
__device__ __forceinline__ void limit_wave_core(const LimiterDev L, const double *U, double *side,
                                                const int *list, const int *count, RunCtl *ctl,
                                                int parity, int *err, int first, int stride) {
  constexpr int NV = 4 * Np;
  static_assert(NV <= 64 && NC <= 64, "one wave per element");
  __shared__ double t[4][Np], tc[4][Np], cv[4][Np], cc[KMAX][4][Np], lim[4][Np], o[4][Np];
  __shared__ double w0[4], wk[KMAX][4], wsum[4];
  if (ctl && (ctl->err || !ctl->s[parity].active))
    return;
  const int n = *count;
  const int lane = threadIdx.x;
  const bool vm = lane < NV;
  const int v = vm ? lane / Np : 0, m = vm ? lane % Np : 0;
  const int E = L.E;
  const double g = L.gamma;
  for (int idx = first; idx < n; idx += stride) {
    const int e = list[idx];
    if (vm)
      t[v][m] = U[(v * Np + m) * E + e];
    __syncthreads();
    const double s2 = sqrt(2.0);
    const double avg[4] = {t[0][0] / s2, t[1][0] / s2, t[2][0] / s2, t[3][0] / s2};
    const CharFrame f = char_frame(g, avg, L.nrm[2 * e], L.nrm[2 * e + 1]);
    if (vm) {
      double w[4];
      to_char(f, t[0][m], t[1][m], t[2][m], t[3][m], w);
      tc[v][m] = pick4(w, v);
    }
    int K = 0;
    for (int k = 0; k < KMAX; ++k)
      if (L.nbr[KMAX * e + k] >= 0)
        ++K;
    __syncthreads();
    if (K == 0) {
      if (vm)
        lim[v][m] = tc[v][m];
    } else {
      // candidate j in characteristic variables: T_j U_nb with the own mean
      int j = 0;
      for (int k = 0; k < KMAX; ++k) {
        const int nb = L.nbr[KMAX * e + k];
        if (nb < 0)
          continue;
        const double *M = L.T + (static_cast<size_t>(e) * KMAX + k) * Np * Np;
        if (vm) {
          double x = 0.0;
#pragma unroll
          for (int sm = 0; sm < Np; ++sm)
            x += M[m * Np + sm] * U[(v * Np + sm) * E + nb];
          cv[v][m] = m == 0 ? t[v][0] : x;
        }
        __syncthreads();
        if (vm) {
          double w[4];
          to_char(f, cv[0][m], cv[1][m], cv[2][m], cv[3][m], w);
          cc[j][v][m] = pick4(w, v);
        }
        __syncthreads();
        ++j;
      }
      if (lane < 4) {
        constexpr double eps = 1e-12;
        const int w = lane;
        double b = 0.0;
        for (int mm = 1; mm < Np; ++mm)
          b += tc[w][mm] * tc[w][mm];
        double d = eps + b;
        w0[w] = 0.5 / (d * d);
        double sum = w0[w];
        const double ideal = 0.5 / static_cast<double>(K);
        for (int j = 0; j < K; ++j) {
          b = 0.0;
          for (int mm = 1; mm < Np; ++mm)
            b += cc[j][w][mm] * cc[j][w][mm];
          d = eps + b;
          wk[j][w] = ideal / (d * d);
          sum += wk[j][w];
        }
        wsum[w] = sum;
      }
      __syncthreads();
      if (vm) {
        if (m == 0) {
          lim[v][0] = tc[v][0];
        } else {
          double x = w0[v] * tc[v][m];
          for (int j = 0; j < K; ++j)
            x += wk[j][v] * cc[j][v][m];
          lim[v][m] = x / wsum[v];
        }
      }
    }
    __syncthreads();
    // back to conserved variables, restoring every mean
    if (vm) {
      const double w[4] = {lim[0][m], lim[1][m], lim[2][m], lim[3][m]};
      double out[4];
      from_char(f, w, out);
      o[v][m] = m == 0 ? t[v][0] : pick4(out, v);
    }
    __syncthreads();
    // positivity: density, then pressure by 64-step bisection
    const double is2 = 1.0 / sqrt(2.0);
    const double a[4] = {o[0][0] * is2, o[1][0] * is2, o[2][0] * is2, o[3][0] * is2};
    if (a[0] <= L.density_floor || pressure_of(g, a[0], a[1], a[2], a[3]) <= L.pressure_floor) {
      // The reference throws here (euler_positivity.hpp:74-75): this element
      // keeps its unlimited coefficients, and outside the stepping loop the
      // smallest such element is recorded in err[1] so the scatter leaves it
      // and every later element (ascending order, euler_modal_limiter.cpp:81)
      // as they were.
      if (vm)
        side[static_cast<size_t>(idx) * 4 * Np + v * Np + m] = t[v][m];
      if (lane == 0) {
        raise_error(ctl, err, kErrPositivity);
        if (!ctl && err)
          atomicMin(err + 1, e);
      }
      __syncthreads();
      continue;
    }
    double ds = 1.0;
    if (lane < NC) {
      const double *b = L.check + lane * Np;
      double r = 0.0;
#pragma unroll
      for (int mm = 0; mm < Np; ++mm)
        r += o[0][mm] * b[mm];
      if (r < L.density_floor)
        ds = (a[0] - L.density_floor) / (a[0] - r);
    }
    ds = wave_min(ds);
    __syncthreads();
    if (vm && m > 0)
      o[v][m] *= ds;
    __syncthreads();
    double lo = 1.0;
    if (lane < NC) {
      const double *b = L.check + lane * Np;
      double sv[4];
#pragma unroll
      for (int vv = 0; vv < 4; ++vv) {
        double x = 0.0;
#pragma unroll
        for (int mm = 0; mm < Np; ++mm)
          x += o[vv][mm] * b[mm];
        sv[vv] = x;
      }
      if (!(pressure_of(g, sv[0], sv[1], sv[2], sv[3]) >= L.pressure_floor)) {
        double l = 0.0, h = 1.0;
        for (int it = 0; it < 64; ++it) {
          const double mid = 0.5 * (l + h);
          const double r = a[0] + mid * (sv[0] - a[0]);
          const double mx = a[1] + mid * (sv[1] - a[1]);
          const double my = a[2] + mid * (sv[2] - a[2]);
          const double en = a[3] + mid * (sv[3] - a[3]);
          if (pressure_of(g, r, mx, my, en) >= L.pressure_floor)
            l = mid;
          else
            h = mid;
        }
        lo = l;
      }
    }
    const double ps = wave_min(lo);
    if (vm) {
      double *dst = side + static_cast<size_t>(idx) * 4 * Np;
      dst[v * Np + m] = m == 0 ? o[v][m] : o[v][m] * ps;
    }
    __syncthreads();
  }
}

template <int P>
__device__ __forceinline__ void limit_wave_body(const LimiterDev L, const double *U, double *side,
                                                const int *list, const int *count, RunCtl *ctl,
                                                int parity, int *err, int first, int stride) {
  limit_wave_core<Tri<P>::Np, Tri<P>::NC, 3>(L, U, side, list, count, ctl, parity, err, first,
                                             stride);
}

// Writes the limited coefficients back; with a trace buffer (mixed
// precision) also the FP32 edge traces of the limited element.
template <int P>
__device__ __forceinline__ void scatter_body(double *U, int E, const double *side, const int *list,
                                             const int *count, RunCtl *ctl, int parity, int idx,
                                             float *tr = nullptr, const float *mxtab = nullptr,
                                             const int *fail = nullptr) {
  constexpr int Np = Tri<P>::Np;
  if (ctl && (ctl->err || !ctl->s[parity].active))
    return;
  if (idx >= *count)
    return;
  const int e = list[idx];
  if (fail && e >= *fail)
    return;   // at or after the first element whose average is below a floor
  // Every value is read before the first store: the compiler cannot tell U
  // from side apart, and a read after a store to U would wait for it.
  const double *src = side + static_cast<size_t>(idx) * 4 * Np;
  double val[4 * Np];
#pragma unroll
  for (int i = 0; i < 4 * Np; ++i)
    val[i] = src[i];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int m = 0; m < Np; ++m)
      U[(v * Np + m) * E + e] = val[v * Np + m];
  if (tr) {
    using X = Mx<P>;
    constexpr int NE = X::NE;
    for (int k = 0; k < 3; ++k)
      for (int q = 0; q < 4; ++q) {
        float x[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (q < NE) {
          const float *b = mxtab + X::kBe + (k * NE + q) * Np;
#pragma unroll
          for (int m = 0; m < Np; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v)
              x[v] += static_cast<float>(val[v * Np + m]) * b[m];
        }
        reinterpret_cast<float4 *>(tr + (static_cast<size_t>(e) * 3 + k) * 16)[q] =
            make_float4(x[0], x[1], x[2], x[3]);
      }
  }
}

template <int P>
__global__ __launch_bounds__(64) void limit_troubled(LimiterDev L, const double *U, double *side,
                                                     const int *list, const int *count,
                                                     RunCtl *ctl, int parity, int *err) {
  limit_wave_body<P>(L, U, side, list, count, ctl, parity, err, blockIdx.x, gridDim.x);
}

// The single-case scatter runs on a small fixed grid that strides over the
// troubled list: the list is short next to E, and one thread per element
// would cost more in launch than the copy itself.
constexpr int kScatterBlocks = 32;

// Flattened over (slot, variable, mode): side[slot][4][Np] goes to
// U[variable][mode][list[slot]].
template <int Np>
__device__ __forceinline__ void scatter_strided(double *U, int E, const double *side,
                                                const int *list, const int *count, RunCtl *ctl,
                                                int parity, const int *fail) {
  if (ctl && (ctl->err || !ctl->s[parity].active))
    return;
  const int n = *count * 4 * Np;
  const int first_bad = fail ? *fail : INT_MAX;
  const int stride = static_cast<int>(gridDim.x * blockDim.x);
  // Four values per thread and pass, all read before any is stored: the
  // compiler cannot tell U from side apart, and a read after a store to U
  // would wait for it.
  constexpr int kPer = 4;
  for (int i0 = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); i0 < n;
       i0 += kPer * stride) {
    int row[kPer], el[kPer];
    double x[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = i0 + k * stride;
      el[k] = -1;
      row[k] = 0;
      x[k] = 0.0;
      if (i < n) {
        const int slot = i / (4 * Np);
        row[k] = i - slot * 4 * Np;
        el[k] = list[slot];
        x[k] = side[i];
      }
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k)
      if (el[k] >= 0 && el[k] < first_bad)
        U[static_cast<size_t>(row[k]) * static_cast<unsigned>(E) + el[k]] = x[k];
  }
}

template <int P>
__global__ __launch_bounds__(256) void scatter_limited(double *U, int E, const double *side,
                                                       const int *list, const int *count,
                                                       RunCtl *ctl, int parity, float *tr,
                                                       const float *mxtab, const int *fail) {
  if (!tr) {
    scatter_strided<Tri<P>::Np>(U, E, side, list, count, ctl, parity, fail);
    return;
  }
  // mixed precision: one element per thread (its FP32 traces are refreshed too)
  const int stride = static_cast<int>(gridDim.x * blockDim.x);
  for (int idx = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); idx < *count;
       idx += stride)
    scatter_body<P>(U, E, side, list, count, ctl, parity, idx, tr, mxtab, fail);
}

// Lines: two neighbours, (P+1) Gauss points and the two ends as check points.
template <int P>
__global__ __launch_bounds__(64) void limit_troubled_line(LimiterDev L, const double *U,
                                                          double *side, const int *list,
                                                          const int *count, RunCtl *ctl,
                                                          int parity, int *err) {
  limit_wave_core<P + 1, P + 3, 2>(L, U, side, list, count, ctl, parity, err, blockIdx.x,
                                   gridDim.x);
}

template <int P>
__global__ __launch_bounds__(256) void scatter_limited_line(double *U, int E, const double *side,
                                                            const int *list, const int *count,
                                                            RunCtl *ctl, int parity,
                                                            const int *fail) {
  scatter_strided<P + 1>(U, E, side, list, count, ctl, parity, fail);
}

// Batched limiter: tiles of 64 troubled-list slots per (case, chunk); the
// stage's output buffer is limited in place as in the single-case path.
struct BatchLimit {
  LimiterDev L;
  double *u, *w1, *w2;
  float *tu, *tw1, *tw2;       // edge traces of u, w1, w2 (mixed precision)
  double *side;
  int *list, *counters;
  RunCtl *ctl;
  int *err;
};

__device__ __forceinline__ double *batch_out(const BatchLimit &c, int stage) {
  return global_ptr(stage == 0 ? c.w1 : stage == 1 ? c.w2 : c.u);
}

__device__ __forceinline__ float *batch_out_traces(const BatchLimit &c, int stage) {
  return global_ptr(stage == 0 ? c.tw1 : stage == 1 ? c.tw2 : c.tu);
}

__device__ __forceinline__ LimiterDev global_limiter(LimiterDev L) {
  L.nbr = global_ptr(L.nbr);
  L.T = global_ptr(L.T);
  L.nrm = global_ptr(L.nrm);
  L.check = global_ptr(L.check);
  return L;
}

// Wave-per-element limiter blocks of one case in a batch (tiles (case, b),
// b < limit_blocks(E)).
__host__ __device__ inline int limit_blocks(int E) { return E < 1024 ? E : 1024; }

template <int P>
__global__ __launch_bounds__(64) void limit_troubled_batch(const BatchLimit *cases,
                                                           const int2 *tiles, BatchStage b) {
  const int2 t = tiles[blockIdx.x];
  if (t.x < 0)
    return;   // padding tile
  const BatchLimit &c = cases[t.x];
  limit_wave_body<P>(global_limiter(c.L), batch_out(c, b.stage), global_ptr(c.side),
                     global_ptr(c.list), global_ptr(c.counters) + b.stage, global_ptr(c.ctl),
                     b.parity, global_ptr(c.err), t.y, limit_blocks(c.L.E));
}

template <int P>
__global__ __launch_bounds__(64) void scatter_limited_batch(const BatchLimit *cases,
                                                            const int2 *tiles, BatchStage b,
                                                            const float *mxtab) {
  const int2 t = tiles[blockIdx.x];
  if (t.x < 0)
    return;   // padding tile
  const BatchLimit &c = cases[t.x];
  scatter_body<P>(batch_out(c, b.stage), c.L.E, global_ptr(c.side), global_ptr(c.list),
                  global_ptr(c.counters) + b.stage, global_ptr(c.ctl), b.parity,
                  t.y * 64 + threadIdx.x, batch_out_traces(c, b.stage), mxtab);
}


// ------------------------------------------------- FP64 issue ceilings ----
//
// The two FP64 units a stage kernel could use, measured on the device: eight
// independent v_fma_f64 chains per lane, and four independent
// v_mfma_f64_16x16x4f64 chains per wave (one 16x16x4 product is 2048 flops).
// Every lane writes one value, so neither loop is dead code.
typedef double mf_d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void fp64_valu_ceiling(double *out, int n) {
  double x[8];
#pragma unroll
  for (int k = 0; k < 8; ++k)
    x[k] = 1.0 + 1e-3 * static_cast<double>(threadIdx.x + k);
  const double a = 0.999999, b = 1e-9;
  for (int i = 0; i < n; ++i)
#pragma unroll
    for (int k = 0; k < 8; ++k)
      x[k] = fma(x[k], a, b);
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    s += x[k];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void fp64_mfma_ceiling(double *out, int n) {
  const int l = threadIdx.x & 63;
  const double a = 1.0 + 1e-3 * l, b = 1.0 - 1e-3 * l;
  mf_d4 c0 = {0.0, 0.0, 0.0, 0.0}, c1 = c0, c2 = c0, c3 = c0;
  for (int i = 0; i < n; ++i) {
    c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c0, 0, 0, 0);
    c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c1, 0, 0, 0);
    c2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c2, 0, 0, 0);
    c3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c3, 0, 0, 0);
  }
  out[blockIdx.x * 256 + threadIdx.x] = c0[0] + c1[1] + c2[2] + c3[3];
}

}  // namespace whip
