// Device engine: one workgroup runs the interior-point method for one slot.
//
// All per-iteration work of src/solver.jl:265-393 happens inside one launch:
// residuals, termination, cone scaling, KKT value update, the level-scheduled
// supernodal LDL' (one wavefront per supernode, supernodes of one level in
// parallel), triangular solves, iterative refinement, step length and the
// update.  Scalars never leave the device; the host only polls the phase.
//
// Lane mode (CLH_LANES): the same source runs one slot per lane.  Every lane
// of a 64-lane workgroup executes the method serially for its own problem;
// the lanes of one batch share the structure (uniform index arrays) and keep
// their values lane-interleaved, so a value access by a uniform index is one
// coalesced load for the whole wavefront.
//
// The file also compiles as plain serial C++ (CLH_EMULATE: one thread, one
// one-lane wave; with CLH_LANES the lanes of a batch run one after another)
// for the CPU unit tests of the engine logic; the package only ever loads the
// HIP build.
#pragma once
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include "clh_types.h"

#if defined(CLH_EMULATE) || defined(CLH_LANES)
#define CLH_SERIAL 1  // one thread per slot: no barriers, no cross-lane reductions
#endif

#if defined(CLH_EMULATE)
#include <algorithm>
#include <chrono>
#define CLH_D inline
#else
#include <hip/hip_runtime.h>
#define CLH_D __device__
#endif

#if defined(CLH_SERIAL)
#define CLH_SYNC() \
  do {             \
  } while (0)
#define CLH_WSYNC() \
  do {              \
  } while (0)
#define CLH_WSYNC_LDS() \
  do {                  \
  } while (0)
#define CLH_WS 1
#define CLH_BSYNC() \
  do {              \
  } while (0)
#else
// a barrier over the workgroup (memory writes of its threads made visible)
#define CLH_BSYNC() __syncthreads()
#if defined(CLH_GRID)
// the grid engine: one barrier over every workgroup of a cooperative launch
#define CLH_SYNC() ::clh::grid_barrier()
#else
#define CLH_SYNC() __syncthreads()
#endif
// make one wavefront's memory writes visible to its other lanes
#define CLH_WSYNC()                                        \
  do {                                                     \
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); \
    __builtin_amdgcn_wave_barrier();                       \
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront"); \
  } while (0)
// the same for LDS writes only (waits on LDS operations, not on memory stores)
#define CLH_WSYNC_LDS()                                             \
  do {                                                              \
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local"); \
    __builtin_amdgcn_wave_barrier();                                \
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local"); \
  } while (0)
#define CLH_WS 64
#endif

// Address spaces of the hot loops' pointers.  A load through a generic
// pointer is a FLAT access, and the compiler must drain every outstanding
// FLAT access before using any result; typed as global (or LDS) the loads are
// counted in order, so the loads of an unrolled loop are all in flight before
// the first wait.
#if defined(CLH_EMULATE)
#define CLH_GLOBAL
#define CLH_LDS
#else
#define CLH_GLOBAL __attribute__((address_space(1)))
#define CLH_LDS __attribute__((address_space(3)))
#endif

// the workgroup, lane and grid builds define the same names: keep them apart
#if defined(CLH_LANES)
#define CLH_ENGINE_NS lanes
#elif defined(CLH_GRID)
#define CLH_ENGINE_NS grid
#else
#define CLH_ENGINE_NS wg
#endif

// CLH_PROBE(B, id): a point of the factorization where an instrumented build
// (one that defines CLH_PROBE_HOOK) takes a timestamp; nothing otherwise
#if defined(CLH_PROBE_HOOK)
#define CLH_PROBE(B, id) CLH_PROBE_HOOK(B, id)
#else
#define CLH_PROBE(B, id) \
  do {                   \
  } while (0)
#endif

namespace clh {

#if defined(CLH_GRID) && !defined(CLH_SERIAL)
// Barrier state of the running grid launch (set by the kernel before its
// first barrier): 64-bit counters in device memory, zeroed at allocation, that
// only grow (no reset to race with the next barrier's arrivals).  The
// workgroups are split into kGridBarGroups groups (workgroup w in group
// w % groups); group g keeps its copy of the generation at [16 g] and its
// arrivals at [16 g + 8] (their own 64-byte lines), and [16 kGridBarGroups]
// counts the groups that have arrived.
__shared__ unsigned long long* g_grid_bar;
#if !defined(CLH_GRID_BAR_GROUPS)
#define CLH_GRID_BAR_GROUPS 16
#endif
constexpr int kGridBarGroups = CLH_GRID_BAR_GROUPS;
static_assert(kGridBarGroups >= 1 && 16 * kGridBarGroups + 1 <= 512, "the groups' lines fit the 4 KB barrier area");
// the launch's abort word (host-mapped memory): once it is set, barriers no
// longer wait and every workgroup leaves at the end of its pass
__shared__ int32_t* g_grid_abort;
// set by the workgroup once it has seen the abort word: its later barriers
// return at once without reading host memory again (an entry check over PCIe
// at every barrier cost ~10% of a large factorization)
__shared__ int g_grid_abort_seen;
// a barrier that has waited this long (100 MHz ticks: 30 s) for the other
// workgroups sets the abort word: a workgroup that never arrives ends the
// launch with an error instead of hanging it
constexpr uint64_t kGridBarrierTimeout = 3000000000ull;
CLH_D bool grid_aborted() {
  return g_grid_abort && __hip_atomic_load(g_grid_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0;
}
// A barrier over the whole grid of a cooperative launch.  The workgroup's
// threads meet first; then its first thread publishes the workgroup's writes
// with one agent-scope release fence (the L2 caches of the XCDs are not
// coherent with each other) and arrives with a relaxed atomic on its group's
// counter.  The group's last arrival takes an acquire-release fence (the
// fences before the group's arrivals synchronize with it through the chain
// of read-modify-writes) and arrives on the groups' counter; the last group
// to arrive stores the next generation into every group's copy.  Each
// workgroup polls its group's copy and takes one agent-scope acquire fence.
// Arrivals on one counter are serialized where the atomics execute: one
// counter for 1024 workgroups measured 34 us per barrier on MI355X, against
// 14 us for 512; sixteen groups divide that chain by sixteen and spread the
// polling over sixteen lines (18 us; eight groups were 3% slower on the
// large solves, thirty-two the same as sixteen).
CLH_D void grid_barrier() {
  __syncthreads();
  if (threadIdx.x == 0 && !g_grid_abort_seen) {
    unsigned long long* const bar = g_grid_bar;
    const unsigned nb = gridDim.x;
    const unsigned ng = nb < (unsigned)kGridBarGroups ? nb : (unsigned)kGridBarGroups;
    const unsigned g = blockIdx.x % ng;
    const unsigned long long members = nb / ng + (g < nb % ng ? 1u : 0u);
    unsigned long long* const gen_p = bar + 16 * g;
    // the generation is read before arriving (the release fence waits for
    // the load): the last arrival may advance it as soon as this one arrives
    const unsigned long long gen = __hip_atomic_load(gen_p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned long long a = __hip_atomic_fetch_add(gen_p + 8, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((a + 1) % members == 0) {  // the group's last arrival
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
      const unsigned long long t =
          __hip_atomic_fetch_add(bar + 16 * kGridBarGroups, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((t + 1) % ng == 0) {  // the last group: release every group
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
        for (unsigned k = 0; k < ng; ++k) __hip_atomic_store(bar + 16 * k, gen + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    // every 1024 polls: the abort word, and the time waited so far
    uint64_t t0 = 0;
    for (uint32_t k = 1; __hip_atomic_load(gen_p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen; ++k) {
      __builtin_amdgcn_s_sleep(1);
      if ((k & 1023) != 0) continue;
      const uint64_t t = wall_clock64();
      if (t0 == 0) t0 = t;
      if (grid_aborted()) { g_grid_abort_seen = 1; break; }
      if (t - t0 > kGridBarrierTimeout && g_grid_abort) {
        __hip_atomic_store(g_grid_abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        g_grid_abort_seen = 1;
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}
#endif

#if defined(CLH_LANES) && defined(CLH_EMULATE)
extern int g_emu_lane;  // lane being emulated
CLH_D int slot_lane() { return g_emu_lane; }
#elif defined(CLH_LANES)
CLH_D int slot_lane() { return (int)threadIdx.x; }  // one 64-lane workgroup per batch
#endif

namespace CLH_ENGINE_NS {

#if defined(CLH_LANES)
// lane-interleaved array: element i of this lane is p[i * kBatchLanes + lane]
template <class T>
struct LV {
  T* p;
  LV() = default;
  CLH_D LV(decltype(nullptr)) : p(nullptr) {}
  CLH_D explicit LV(T* q) : p(q) {}
  template <class U>
  CLH_D LV(const LV<U>& o) : p(o.p) {}
  CLH_D T& operator[](ptrdiff_t i) const { return p[i * kBatchLanes + slot_lane()]; }
  CLH_D LV operator+(ptrdiff_t o) const { return LV(p + o * kBatchLanes); }
  CLH_D explicit operator bool() const { return p != nullptr; }
};
template <class T> using VP = LV<T>;
using SD = SlotDescT<VP>;  // same layout as SlotDesc
#else
CLH_D int slot_lane() { return 0; }
template <class T> using VP = T*;
using SD = SlotDesc;
#endif
using VD = VP<double>;
using VCD = VP<const double>;
template <class FT>
CLH_D VP<FT> vcast(void* p) {
  return VP<FT>((FT*)p);
}
// device memory seen as global memory (no-op in the CPU build and for the
// lane engine's interleaved arrays)
template <class T>
CLH_D CLH_GLOBAL T* gmem(T* p) {
  return (CLH_GLOBAL T*)p;
}
#if defined(CLH_LANES)
template <class T>
CLH_D LV<T> gmem(LV<T> p) {
  return p;
}
#endif
// a packed record (a multiple of 16 bytes) read through a global pointer
template <class T>
CLH_D T gload(const CLH_GLOBAL T* p) {
#if defined(CLH_EMULATE)
  return *p;
#else
  static_assert(sizeof(T) % 16 == 0, "records are read in 16-byte pieces");
  typedef int v4i __attribute__((ext_vector_type(4)));
  const CLH_GLOBAL v4i* q = (const CLH_GLOBAL v4i*)p;
  v4i tmp[sizeof(T) / 16];
  for (int i = 0; i < (int)(sizeof(T) / 16); ++i) tmp[i] = q[i];
  T r;
  __builtin_memcpy(&r, tmp, sizeof(T));
  return r;
#endif
}
// this lane's phase timers
CLH_D uint64_t* lane_prof(const SD& D) { return D.prof ? D.prof + 8 * slot_lane() : nullptr; }

constexpr double kDMax = 1.7976931348623157e308;
constexpr double kDEps = 2.220446049250313e-16;

CLH_D double dinf() { return __builtin_huge_val(); }
CLH_D double dnan() { return __builtin_nan(""); }
CLH_D double dmin(double a, double b) { return a < b ? a : b; }
CLH_D double dmax(double a, double b) { return a > b ? a : b; }
CLH_D int imin(int a, int b) { return a < b ? a : b; }
CLH_D int imax(int a, int b) { return a > b ? a : b; }
// src/utils/mathutils.jl:4-18
CLH_D double clipd(double s, double lo, double hi) { return s < lo ? lo : (s > hi ? hi : s); }
CLH_D double logsafe(double v) { return v < 0 ? -dinf() : log(v); }
CLH_D bool finite(double v) { return v - v == 0.0; }
CLH_D int tri(int k) { return k * (k + 1) / 2; }

// The pivot that dynamic regularization puts in place of a pivot d whose
// sign disagrees with the expected one sg (d * sg <= eps; the caller tests
// that).  The reference's rule (QDLDL, reached from
// src/kktsolvers/direct-ldl/directldl_qdldl.jl:62-86) sets sg * delta.  With
// dyn_keep_magnitude the pivot keeps its magnitude when that is above delta:
// a pivot that cancellation has pushed to the wrong side by much more than
// delta (near the end of a solve, when s/z spans 24 orders of magnitude)
// then scales its column by 1/|d| instead of 1/delta, and the columns
// coupled to it are not driven across zero one after the other until the
// factorization overflows.  Both rules perturb K by about |d| on that
// diagonal; iterative refinement answers for it.
CLH_D double dyn_reg_pivot(const DevSettings& st, double d, int sg) {
  const double a = fabs(d);
  return sg * (st.dyn_keep_magnitude && a > st.dynamic_regularization_delta ? a : st.dynamic_regularization_delta);
}

// ---------------------------------------------------------------------------
// workgroup context and deterministic reductions
// ---------------------------------------------------------------------------
struct Blk {
  int tid, nthr, lane, wave, nwaves;
  double* red;  // LDS scratch, >= nwaves doubles
  // supernodes are processed by groups of G lanes inside a wavefront
  int G, gl, grp, ngrp;
  // LDS work area of the workgroup (null in the lane engine): supernode
  // panels are assembled and factored there, the triangular solves keep their
  // vector there, when they fit
  char* lds;
  int lds_bytes;
  // the grid engine: this thread within its workgroup, the group's index
  // among the workgroup's groups (its share of the LDS), the reduction
  // partials in memory (two buffers of nwg), and each thread's buffer parity
  int ltid = -1, lnthr = -1, lgrp = 0, lngrp = 1;
  double* gred = nullptr;
  int wg = 0, nwg = 1;
  unsigned char* epoch = nullptr;
};

// phase timers: 0 setup, 1 residuals+info, 2 cone scaling, 3 factorization,
// 4 triangular solves, 5 refinement products, 6 step length, 7 whole pass
CLH_D uint64_t clh_clock() {  // 100 MHz ticks
#if defined(CLH_EMULATE)
  return (uint64_t)(std::chrono::duration_cast<std::chrono::nanoseconds>(
                        std::chrono::steady_clock::now().time_since_epoch()).count() / 10);
#else
  return wall_clock64();
#endif
}
struct PhaseTimer {
  const Blk& B;
  uint64_t* slot;
  uint64_t t0;
  CLH_D PhaseTimer(const Blk& b, uint64_t* prof, int k) : B(b), slot(prof ? prof + k : nullptr), t0(clh_clock()) {}
  CLH_D ~PhaseTimer() {
    if (slot && B.tid == 0) *slot += clh_clock() - t0;
  }
};

#if defined(CLH_SERIAL)
// one thread per slot: a reduction over the slot is the thread's own value
CLH_D double blk_sum(const Blk&, double v) { return v; }
CLH_D double blk_max(const Blk&, double v) { return v; }
CLH_D double blk_min(const Blk&, double v) { return v; }
#else
CLH_D double wave_sum(double v) {
  for (int o = CLH_WS / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
CLH_D double wave_max(double v) {
  for (int o = CLH_WS / 2; o > 0; o >>= 1) v = dmax(v, __shfl_xor(v, o));
  return v;
}
CLH_D double wave_min(double v) {
  for (int o = CLH_WS / 2; o > 0; o >>= 1) v = dmin(v, __shfl_xor(v, o));
  return v;
}
#if defined(CLH_GRID)
// Grid reductions: a workgroup reduction, one partial per workgroup in
// memory, a grid barrier, then every workgroup combines the partials in
// workgroup order (thread t takes partials t, t + 256, ...; the workgroup
// combines those in its fixed shape), so every thread of the grid gets the
// same, reproducible value.  Two partial buffers alternate: a buffer is
// rewritten two reductions later, after every thread has passed the barrier
// of the reduction in between.
enum { RED_SUM, RED_MAX, RED_MIN };
CLH_D double red_op(int op, double a, double b) { return op == RED_SUM ? a + b : (op == RED_MAX ? dmax(a, b) : dmin(a, b)); }
CLH_D double wave_red(int op, double v) {
  for (int o = CLH_WS / 2; o > 0; o >>= 1) v = red_op(op, v, __shfl_xor(v, o));
  return v;
}
CLH_D double wg_reduce(const Blk& B, int op, double v) {
  v = wave_red(op, v);
  __syncthreads();
  if (B.lane == 0) B.red[B.wave] = v;
  __syncthreads();
  double t = B.red[0];
  for (int w = 1; w < B.nwaves; ++w) t = red_op(op, t, B.red[w]);
  return t;
}
CLH_D double grid_reduce(const Blk& B, int op, double v) {
  v = wg_reduce(B, op, v);
  unsigned char& ep = B.epoch[B.ltid];
  double* buf = B.gred + (size_t)(ep & 1) * B.nwg;
  ep ^= 1;
  if (B.ltid == 0) buf[B.wg] = v;
  grid_barrier();
  const double id = op == RED_SUM ? 0.0 : (op == RED_MAX ? -dinf() : dinf());
  double t = id;
  for (int w = B.ltid; w < B.nwg; w += B.lnthr) t = red_op(op, t, buf[w]);
  return wg_reduce(B, op, t);
}
CLH_D double blk_sum(const Blk& B, double v) { return grid_reduce(B, RED_SUM, v); }
CLH_D double blk_max(const Blk& B, double v) { return grid_reduce(B, RED_MAX, v); }
CLH_D double blk_min(const Blk& B, double v) { return grid_reduce(B, RED_MIN, v); }
#else
// Each returns the same value to every thread; the combination order is
// fixed by the workgroup shape, so results are bitwise reproducible.
CLH_D double blk_sum(const Blk& B, double v) {
  v = wave_sum(v);
  CLH_SYNC();
  if (B.lane == 0) B.red[B.wave] = v;
  CLH_SYNC();
  double t = 0;
  for (int w = 0; w < B.nwaves; ++w) t += B.red[w];
  return t;
}
CLH_D double blk_max(const Blk& B, double v) {
  v = wave_max(v);
  CLH_SYNC();
  if (B.lane == 0) B.red[B.wave] = v;
  CLH_SYNC();
  double t = B.red[0];
  for (int w = 1; w < B.nwaves; ++w) t = dmax(t, B.red[w]);
  return t;
}
CLH_D double blk_min(const Blk& B, double v) {
  v = wave_min(v);
  CLH_SYNC();
  if (B.lane == 0) B.red[B.wave] = v;
  CLH_SYNC();
  double t = B.red[0];
  for (int w = 1; w < B.nwaves; ++w) t = dmin(t, B.red[w]);
  return t;
}
#endif  // CLH_GRID
#endif
CLH_D double blk_dot(const Blk& B, VCD a, VCD b, int n) {
  double t = 0;
  for (int i = B.tid; i < n; i += B.nthr) t += a[i] * b[i];
  return blk_sum(B, t);
}
CLH_D double blk_norminf(const Blk& B, VCD a, int n) {
  double t = 0;
  for (int i = B.tid; i < n; i += B.nthr) t = dmax(t, fabs(a[i]));
  return blk_max(B, t);
}
// 2-norm of m .* v (src/utils/mathutils.jl:62-69)
CLH_D double blk_norm_scaled(const Blk& B, VCD m, VCD v, int n) {
  double t = 0;
  for (int i = B.tid; i < n; i += B.nthr) {
    double p = m[i] * v[i];
    t += p * p;
  }
  return sqrt(blk_sum(B, t));
}
CLH_D double blk_norminf_scaled(const Blk& B, VCD m, VCD v, int n) {
  double t = 0;
  for (int i = B.tid; i < n; i += B.nthr) t = dmax(t, fabs(m[i] * v[i]));
  return blk_max(B, t);
}
#if defined(CLH_GRID) && !defined(CLH_SERIAL)
// the grid engine sums CSR rows longer than this with a whole wavefront (a
// row of A' in a PSD relaxation holds thousands of entries; one thread per
// row would walk it alone)
constexpr int kLongRow = 128;
#define CLH_LONG_ROWS 1
#endif
// y = Sym(P) x with P stored as full symmetric CSR
CLH_D void spmv_csr(const Blk& B, int nrow, const int32_t* rp_, const int32_t* ci_, VCD v_, VCD x_, VD y,
                    double alpha, double beta) {
  const auto rp = gmem(rp_);
  const auto ci = gmem(ci_);
  const auto v = gmem(v_);
  const auto x = gmem(x_);
  for (int i = B.tid; i < nrow; i += B.nthr) {
    const int k1 = rp[i + 1];
#if defined(CLH_LONG_ROWS)
    if (k1 - rp[i] > kLongRow) continue;
#endif
    double t = 0;
#pragma unroll 8
    for (int k = rp[i]; k < k1; ++k) t += v[k] * x[ci[k]];
    y[i] = alpha * t + (beta == 0 ? 0.0 : beta * y[i]);
  }
#if defined(CLH_LONG_ROWS)
  const int lane = B.tid & 63;
  for (int i = B.tid >> 6; i < nrow; i += B.nthr >> 6) {
    const int k0 = rp[i], k1 = rp[i + 1];
    if (k1 - k0 <= kLongRow) continue;  // the same on every lane of the wavefront
    double t = 0;
    for (int k = k0 + lane; k < k1; k += 64) t += v[k] * x[ci[k]];
    t = wave_sum(t);
    if (lane == 0) y[i] = alpha * t + (beta == 0 ? 0.0 : beta * y[i]);
  }
#endif
}
// x' P y with full symmetric CSR P
CLH_D double blk_quad(const Blk& B, const SD& D, VCD x, VCD y_) {
  const auto Pp = gmem(D.Pp);
  const auto Pi = gmem(D.Pi);
  const auto Px = gmem(D.Px);
  const auto y = gmem(y_);
  double t = 0;
  for (int i = B.tid; i < D.n; i += B.nthr) {
    double r = 0;
    const int k1 = Pp[i + 1];
#pragma unroll 8
    for (int k = Pp[i]; k < k1; ++k) r += Px[k] * y[Pi[k]];
    t += x[i] * r;
  }
  return blk_sum(B, t);
}

// ---------------------------------------------------------------------------
// the slot's scalar state while a launch runs
// ---------------------------------------------------------------------------
// Every thread of a slot computes the same scalars (they come from reductions
// that hand every thread the same value), so instead of one copy per lane in
// registers -- about 400 bytes, which the register budget of four resident
// workgroups per CU cannot hold and the compiler spilled to scratch memory --
// each wavefront keeps one copy in LDS.  The lanes of a wavefront execute in
// lockstep and write the same values, read-modify-writes included, so the copy
// evolves as one lane's would; no wavefront touches another's copy.  The lane
// engine (one slot per lane) and the serial CPU build keep a private copy.
#if defined(CLH_SERIAL)
typedef SlotScalars Scal;
#else
typedef CLH_LDS SlotScalars Scal;
#if defined(CLH_BLOCK_THREADS) && !defined(CLH_GRID)
constexpr int kMaxWaves = (CLH_BLOCK_THREADS + 63) / 64;  // the workgroup engine's block size
#else
constexpr int kMaxWaves = 4;  // 256-thread workgroups
#endif
#endif

// ---------------------------------------------------------------------------
// x[k] *= r[i] * c[col[k]] over the entries k of each CSR row i (a missing
// r or c counts as ones); the grid engine takes the long rows by a
// wavefront each.  Each entry's product is the same whichever thread
// forms it.
CLH_D void scale_rows(const Blk& B, int nrow, const int32_t* rp, const int32_t* ci, VD x, VD r, VD c) {
  for (int i = B.tid; i < nrow; i += B.nthr) {
#if defined(CLH_LONG_ROWS)
    if (rp[i + 1] - rp[i] > kLongRow) continue;
#endif
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      if (r && c) x[k] *= r[i] * c[ci[k]];
      else if (r) x[k] *= r[i];
      else x[k] *= c[ci[k]];
    }
  }
#if defined(CLH_LONG_ROWS)
  const int lane = B.tid & 63;
  for (int i = B.tid >> 6; i < nrow; i += B.nthr >> 6) {
    if (rp[i + 1] - rp[i] <= kLongRow) continue;  // the same on every lane
    for (int k = rp[i] + lane; k < rp[i + 1]; k += 64) {
      if (r && c) x[k] *= r[i] * c[ci[k]];
      else if (r) x[k] *= r[i];
      else x[k] *= c[ci[k]];
    }
  }
#endif
}
// x[k] *= c[j] * r[row[k]] over the entries k of each CSC column j
CLH_D void scale_cols(const Blk& B, int ncol, const int32_t* cp, const int32_t* ri, VD x, VD c, VD r) {
  for (int j = B.tid; j < ncol; j += B.nthr) {
#if defined(CLH_LONG_ROWS)
    if (cp[j + 1] - cp[j] > kLongRow) continue;
#endif
    for (int k = cp[j]; k < cp[j + 1]; ++k) {
      if (r && c) x[k] *= r[ri[k]] * c[j];
      else if (r) x[k] *= r[ri[k]];
      else x[k] *= c[j];
    }
  }
#if defined(CLH_LONG_ROWS)
  const int lane = B.tid & 63;
  for (int j = B.tid >> 6; j < ncol; j += B.nthr >> 6) {
    if (cp[j + 1] - cp[j] <= kLongRow) continue;  // the same on every lane
    for (int k = cp[j] + lane; k < cp[j + 1]; k += 64) {
      if (r && c) x[k] *= r[ri[k]] * c[j];
      else if (r) x[k] *= r[ri[k]];
      else x[k] *= c[j];
    }
  }
#endif
}

// setup phase: unscaled norms, Ruiz equilibration, KKT values
// (src/problemdata.jl:47-51, :196-284)
// ---------------------------------------------------------------------------
// rectify_equilibration! (src/cones/coneops_defaults.jl:32-44,
// src/gpucones/coneops_compositecone_gpu.jl:21-39): delta = 1 on the
// elementwise cones and mean(e)/e over each other cone; returns whether there
// is any non-elementwise cone.  Each row is written by one thread only.
CLH_D bool cones_rectify(const Blk& B, const SD& D, VD delta, VCD e) {
  for (int i = B.tid; i < D.m; i += B.nthr)
    if (D.etype[i] == CT_ZERO || D.etype[i] == CT_NONNEG) delta[i] = 1.0;
  for (int k = B.tid; k < D.nsoc + D.nexp + D.npow + D.npsd; k += B.nthr) {
    int off, dim;
    if (k < D.nsoc) { off = D.soc_off[k]; dim = D.soc_dim[k]; }
    else if (k < D.nsoc + D.nexp) { off = D.exp_off[k - D.nsoc]; dim = 3; }
    else if (k < D.nsoc + D.nexp + D.npow) { off = D.pow_off[k - D.nsoc - D.nexp]; dim = 3; }
    else { const int q = k - D.nsoc - D.nexp - D.npow; off = D.psd_off[q]; dim = tri(D.psd_n[q]); }
    double mean = 0;
    for (int i = 0; i < dim; ++i) mean += e[off + i];
    mean /= dim;
    for (int i = 0; i < dim; ++i) delta[off + i] = mean / e[off + i];
  }
  CLH_SYNC();
  return (D.nsoc + D.nexp + D.npow + D.npsd) > 0;
}

CLH_D void setup_phase(const Blk& B, const SD& D, Scal& S) {
  const DevSettings& st = *D.st;
  const int n = D.n, m = D.m;
#if defined(CLH_LANES)
  {  // this lane's uploaded values into the interleaved arrays
    const double* in = D.inbox + (size_t)slot_lane() * D.inbox_lstride;
    for (int k = 0; k < D.nnzP; ++k) D.Px[k] = *in++;
    for (int k = 0; k < D.nnzA; ++k) D.Ax[k] = *in++;
    for (int k = 0; k < D.nnzA; ++k) D.Atx[k] = *in++;
    for (int j = 0; j < n; ++j) D.q[j] = *in++;
    for (int i = 0; i < m; ++i) D.b[i] = *in++;
  }
#endif
  for (int i = B.tid; i < m; i += B.nthr) D.b[i] = dmin(D.b[i], 1e20);
  CLH_SYNC();
  S.normq = blk_norminf(B, D.q, n);
  S.normb = blk_norminf(B, D.b, m);
  if (S.reuse_equil) {
    // new data scaled with the equilibration of the first solve
    // (src/data_updating.jl:70-85, :183-202)
    const double c = S.c;
    for (int i = B.tid; i < n; i += B.nthr)
      for (int k = D.Pp[i]; k < D.Pp[i + 1]; ++k) D.Px[k] = D.d[i] * D.d[D.Pi[k]] * c * D.Px[k];
    for (int i = B.tid; i < m; i += B.nthr)
      for (int k = D.Ap[i]; k < D.Ap[i + 1]; ++k) D.Ax[k] = D.e[i] * D.d[D.Ai[k]] * D.Ax[k];
    for (int j = B.tid; j < n; j += B.nthr) {
      for (int k = D.Atp[j]; k < D.Atp[j + 1]; ++k) D.Atx[k] = D.e[D.Ati[k]] * D.d[j] * D.Atx[k];
      D.q[j] = D.q[j] * D.d[j] * c;
    }
    for (int i = B.tid; i < m; i += B.nthr) D.b[i] = D.b[i] * D.e[i];
    CLH_SYNC();
    for (int k = B.tid; k < D.nnzK; k += B.nthr) D.Kx[k] = 0.0;
    CLH_SYNC();
    for (int k = B.tid; k < D.nnzP; k += B.nthr)
      if (D.Pk[k] >= 0) D.Kx[D.Pk[k]] = D.Px[k];
    for (int k = B.tid; k < D.nnzA; k += B.nthr) D.Kx[D.Ak[k]] = D.Ax[k];
    CLH_SYNC();
    return;
  }
  for (int j = B.tid; j < n; j += B.nthr) D.d[j] = 1.0;
  for (int i = B.tid; i < m; i += B.nthr) D.e[i] = 1.0;
  CLH_SYNC();
  S.c = 1.0;
  VD dw = D.dinv;  // work vectors, as in the reference
  VD ew = D.einv;
  if (st.equilibrate_enable) {
    const double smin = st.equilibrate_min_scaling, smax = st.equilibrate_max_scaling;
    for (int it = 0; it < st.equilibrate_max_iter; ++it) {
      // kkt_col_norms! (src/utils/mathutils.jl:83-95): full-P row = sym column
      // (the maxima are exact in any order: the grid engine takes the long
      // rows of A and A' by a wavefront each, as its SpMV does)
      for (int j = B.tid; j < n; j += B.nthr) {
#if defined(CLH_LONG_ROWS)
        if (D.Atp[j + 1] - D.Atp[j] > kLongRow) continue;
#endif
        double t = 0;
        for (int k = D.Pp[j]; k < D.Pp[j + 1]; ++k) t = dmax(t, fabs(D.Px[k]));
        for (int k = D.Atp[j]; k < D.Atp[j + 1]; ++k) t = dmax(t, fabs(D.Atx[k]));
        if (t == 0) t = 1;
        t = 1 / sqrt(t);
        dw[j] = clipd(t, smin / D.d[j], smax / D.d[j]);
      }
      for (int i = B.tid; i < m; i += B.nthr) {
#if defined(CLH_LONG_ROWS)
        if (D.Ap[i + 1] - D.Ap[i] > kLongRow) continue;
#endif
        double t = 0;
        for (int k = D.Ap[i]; k < D.Ap[i + 1]; ++k) t = dmax(t, fabs(D.Ax[k]));
        if (t == 0) t = 1;
        t = 1 / sqrt(t);
        ew[i] = clipd(t, smin / D.e[i], smax / D.e[i]);
      }
#if defined(CLH_LONG_ROWS)
      {
        const int lane = B.tid & 63;
        for (int j = B.tid >> 6; j < n; j += B.nthr >> 6) {
          if (D.Atp[j + 1] - D.Atp[j] <= kLongRow) continue;  // the same on every lane
          double t = 0;
          for (int k = D.Pp[j] + lane; k < D.Pp[j + 1]; k += 64) t = dmax(t, fabs(D.Px[k]));
          for (int k = D.Atp[j] + lane; k < D.Atp[j + 1]; k += 64) t = dmax(t, fabs(D.Atx[k]));
          t = wave_max(t);
          if (t == 0) t = 1;
          t = 1 / sqrt(t);
          if (lane == 0) dw[j] = clipd(t, smin / D.d[j], smax / D.d[j]);
        }
        for (int i = B.tid >> 6; i < m; i += B.nthr >> 6) {
          if (D.Ap[i + 1] - D.Ap[i] <= kLongRow) continue;
          double t = 0;
          for (int k = D.Ap[i] + lane; k < D.Ap[i + 1]; k += 64) t = dmax(t, fabs(D.Ax[k]));
          t = wave_max(t);
          if (t == 0) t = 1;
          t = 1 / sqrt(t);
          if (lane == 0) ew[i] = clipd(t, smin / D.e[i], smax / D.e[i]);
        }
      }
#endif
      CLH_SYNC();
      // scale_data! (src/problemdata.jl:377-396)
      for (int i = B.tid; i < n; i += B.nthr)
        for (int k = D.Pp[i]; k < D.Pp[i + 1]; ++k) D.Px[k] *= dw[i] * dw[D.Pi[k]];
      scale_rows(B, m, D.Ap, D.Ai, D.Ax, ew, dw);
      scale_cols(B, n, D.Atp, D.Ati, D.Atx, dw, ew);
      for (int j = B.tid; j < n; j += B.nthr) { D.q[j] *= dw[j]; D.d[j] *= dw[j]; }
      for (int i = B.tid; i < m; i += B.nthr) { D.b[i] *= ew[i]; D.e[i] *= ew[i]; }
      CLH_SYNC();
      // column norms of the upper triangle of P, then the cost scaling
      double tsum = 0;
      for (int j = B.tid; j < n; j += B.nthr) {
        double t = 0;
        for (int k = D.Pp[j]; k < D.Pp[j + 1]; ++k)
          if (D.Pi[k] <= j) t = dmax(t, fabs(D.Px[k]));
        tsum += t;
      }
      double mean = blk_sum(B, tsum) / n;
      double infq = blk_norminf(B, D.q, n);
      if (mean != 0 && infq != 0) {
        double ctmp = 1 / dmax(infq, mean);
        ctmp = clipd(ctmp, smin / S.c, smax / S.c);
        for (int k = B.tid; k < D.nnzP; k += B.nthr) D.Px[k] *= ctmp;
        for (int j = B.tid; j < n; j += B.nthr) D.q[j] *= ctmp;
        S.c *= ctmp;
      }
      CLH_SYNC();
    }
    if (cones_rectify(B, D, ew, D.e)) {
      scale_rows(B, m, D.Ap, D.Ai, D.Ax, ew, VD(nullptr));
      scale_cols(B, n, D.Atp, D.Ati, D.Atx, VD(nullptr), ew);
      CLH_SYNC();
      for (int i = B.tid; i < m; i += B.nthr) { D.b[i] *= ew[i]; D.e[i] *= ew[i]; }
    }
  }
  CLH_SYNC();
  for (int j = B.tid; j < n; j += B.nthr) D.dinv[j] = 1 / D.d[j];
  for (int i = B.tid; i < m; i += B.nthr) D.einv[i] = 1 / D.e[i];
  // KKT values: P (upper triangle) and A' blocks; cone blocks are written
  // at every scaling update
  for (int k = B.tid; k < D.nnzK; k += B.nthr) D.Kx[k] = 0.0;
  CLH_SYNC();
  for (int k = B.tid; k < D.nnzP; k += B.nthr)
    if (D.Pk[k] >= 0) D.Kx[D.Pk[k]] = D.Px[k];
  for (int k = B.tid; k < D.nnzA; k += B.nthr) D.Kx[D.Ak[k]] = D.Ax[k];
  CLH_SYNC();
}

// ---------------------------------------------------------------------------
// cones
// ---------------------------------------------------------------------------
// ---- second-order cone helpers (src/cones/coneops_socone.jl)
template <class P>
CLH_D double soc_res(P z, int d) {  // (z0 - |z1|)(z0 + |z1|)
  double t = 0;
  for (int i = 1; i < d; ++i) t += z[i] * z[i];
  double z1 = sqrt(t);
  return (z[0] - z1) * (z[0] + z1);
}
template <class P>
CLH_D double soc_sqrt_res(P z, int d) {
  double r = soc_res(z, d);
  return r > 0 ? sqrt(r) : 0.0;
}
// coneops_socone.jl:78-154
template <class PC, class P>
CLH_D bool soc_scale(PC s, PC z, int d, P w, P lam, double& eta, int sp, P sd, P su, P sv) {
  double zscale = soc_sqrt_res(z, d), sscale = soc_sqrt_res(s, d);
  if (zscale == 0 || sscale == 0) return false;
  eta = sqrt(sscale / zscale);
  for (int i = 0; i < d; ++i) w[i] = s[i] / sscale;
  w[0] += z[0] / zscale;
  for (int i = 1; i < d; ++i) w[i] -= z[i] / zscale;
  double wscale = soc_sqrt_res(w, d);
  if (wscale == 0) return false;
  for (int i = 0; i < d; ++i) w[i] /= wscale;
  double w1sq = 0;
  for (int i = 1; i < d; ++i) w1sq += w[i] * w[i];
  w[0] = sqrt(1 + w1sq);
  double gamma = 0.5 * wscale;
  lam[0] = gamma;
  double c1 = (gamma + z[0] / zscale) / sscale, c2 = (gamma + s[0] / sscale) / zscale;
  for (int i = 1; i < d; ++i) lam[i] = c1 * s[i] + c2 * z[i];
  double inv = 1.0 / (s[0] / sscale + z[0] / zscale + 2 * gamma);
  for (int i = 1; i < d; ++i) lam[i] *= inv;
  double sc = sqrt(sscale * zscale);
  for (int i = 0; i < d; ++i) lam[i] *= sc;
  if (sp >= 0) {
    double al = 2 * w[0];
    double wsq = w[0] * w[0] + w1sq;
    double wsqinv = 1 / wsq;
    sd[0] = wsqinv / 2;
    double u0 = sqrt(wsq - sd[0]);
    double u1 = al / u0;
    double v1 = sqrt(2 * (2 + wsqinv) / (2 * wsq - wsqinv));
    su[0] = u0;
    for (int i = 1; i < d; ++i) su[i] = u1 * w[i];
    sv[0] = 0;
    for (int i = 1; i < d; ++i) sv[i] = v1 * w[i];
  }
  return true;
}
// coneops_socone.jl:470-516
template <class P>
CLH_D double soc_step(P x, P y, double amax, int d) {
  if (x[0] >= 0 && y[0] < 0) amax = dmin(amax, -x[0] / y[0]);
  double a = soc_res(y, d);
  double xy = 0;
  for (int i = 1; i < d; ++i) xy += x[i] * y[i];
  double b = 2 * (x[0] * y[0] - xy);
  double c = dmax(0.0, soc_res(x, d));
  double dd = b * b - 4 * a * c;
  if ((a > 0 && b > 0) || dd < 0) return amax;
  if (a == 0) return amax;
  if (c == 0) return a >= 0 ? amax : 0.0;
  double t = (b >= 0) ? (-b - sqrt(dd)) : (-b + sqrt(dd));
  double r1 = (2 * c) / t, r2 = t / (2 * a);
  if (r1 < 0) r1 = kDMax;
  if (r2 < 0) r2 = kDMax;
  return dmin(amax, dmin(r1, r2));
}
template <class P>
CLH_D double soc_res_shifted(P z, P dz, double a, int d) {
  double x0 = z[0] + a * dz[0];
  double t = 0;
  for (int i = 1; i < d; ++i) {
    double v = z[i] + a * dz[i];
    t += v * v;
  }
  double x1 = sqrt(t);
  return (x0 - x1) * (x0 + x1);
}

// ---- exponential cone (src/cones/coneops_expcone.jl)
// Wright omega, :470-518 (the discarded `q * zinv` at :502 is kept as is)
CLH_D double wright_omega(double z) {
  double w;
  if (z < 0) return dnan();
  if (z < 1 + 3.14159265358979323846) {
    double zm1 = z - 1, p = zm1;
    w = 1 + 0.5 * p;
    p *= zm1; w += (1 / 16.0) * p;
    p *= zm1; w -= (1 / 192.0) * p;
    p *= zm1; w -= (1 / 3072.0) * p;
    p *= zm1; w += (13 / 61440.0) * p;
  } else {
    double logz = logsafe(z), zinv = 1 / z;
    w = z - logz;
    double q = logz * zinv;
    w += q;
    q *= zinv;
    w += q * (logz / 2 - 1);
    w += q * (logz * logz / 3. - (3 / 2.) * logz + 1);
  }
  double r = z - w - logsafe(w);
  for (int i = 0; i < 2; ++i) {
    double wp1 = w + 1;
    double t = wp1 * (wp1 + (2. * r) / 3.0);
    w *= 1 + (r / wp1) * (t - 0.5 * r) / (t - r);
    double wp6 = wp1 * wp1 * wp1 * wp1 * wp1 * wp1;
    r = (2 * w * w - 8 * w - 1) / (72.0 * wp6) * r * r * r * r;
  }
  return w;
}
CLH_D void exp_grad_primal(const double* s, double* g) {  // :318-330
  double om = wright_omega(1 - s[0] / s[1] - logsafe(s[1] / s[2]));
  double g1 = 1 / ((om - 1) * s[1]);
  g[0] = g1;
  g[1] = g1 + g1 * logsafe(om * s[1] / s[2]) - 1 / s[1];
  g[2] = om / ((1 - om) * s[2]);
}
CLH_D void exp_dual_grad_H(const double* z, double* grad, double* H) {  // :431-461
  double l = logsafe(-z[2] / z[0]);
  double r = -z[0] * l - z[0] + z[1];
  double c2 = 1 / r;
  grad[0] = c2 * l - 1 / z[0];
  grad[1] = -c2;
  grad[2] = (c2 * z[0] - 1) / z[2];
  H[0] = ((r * r - z[0] * r + l * l * z[0] * z[0]) / (r * z[0] * z[0] * r));
  H[3] = (-l / (r * r)); H[1] = H[3];
  H[4] = (1 / (r * r));
  H[6] = ((z[1] - z[0]) / (r * r * z[2])); H[2] = H[6];
  H[7] = (-z[0] / (r * r * z[2])); H[5] = H[7];
  H[8] = ((r * r - z[0] * r + z[0] * z[0]) / (r * r * z[2] * z[2]));
}
CLH_D bool exp_primal_feas(const double* s) {
  if (s[2] > 0 && s[1] > 0) return s[1] * logsafe(s[2] / s[1]) - s[0] > 0;
  return false;
}
CLH_D bool exp_dual_feas(const double* z) {
  if (z[2] > 0 && z[0] < 0) return z[1] - z[0] - z[0] * logsafe(-z[2] / z[0]) > 0;
  return false;
}
CLH_D double exp_barrier(const double* z, const double* s) {  // :258-272
  double l = logsafe(-z[2] / z[0]);
  double bd = -logsafe(-z[2] * z[0]) - logsafe(z[1] - z[0] - z[0] * l);
  double om = wright_omega(1 - s[0] / s[1] - logsafe(s[1] / s[2]));
  om = (om - 1) * (om - 1) / om;
  double bp = -logsafe(om) - 2 * logsafe(s[1]) - logsafe(s[2]) - 3;
  return bd + bp;
}

// ---- power cone (src/cones/coneops_powcone.jl)
CLH_D double pow_newton(double s3, double phi, double a) {  // :449-478 with the one-sided Newton of
  // coneops_nonsymmetric_common.jl:221-244
  double x = -1 / s3 + (2 * s3 + sqrt(phi * phi / s3 / s3 + 3 * phi)) / (phi - s3 * s3);
  double t0 = -2 * a * logsafe(a) - 2 * (1 - a) * logsafe(1 - a);
  for (int iter = 0; iter < 100; ++iter) {
    double t1 = x * x, t2 = 2 * x / s3;
    double f0 = 2 * a * logsafe(2 * a * t1 + (1 + a) * t2) + 2 * (1 - a) * logsafe(2 * (1 - a) * t1 + (2 - a) * t2) -
                logsafe(phi) - logsafe(t1 + t2) - 2 * logsafe(t2) + t0;
    double f1 = 2 * a * a / (a * x + (1 + a) / s3) + 2 * (1 - a) * (1 - a) / ((1 - a) * x + (2 - a) / s3) -
                2 * (x + 1 / s3) / (t1 + t2);
    double dx = -f0 / f1;
    if ((dx < kDEps) || (fabs(dx / x) < 1.4901161193847656e-08) || (fabs(f1) < kDEps)) break;
    x += dx;
  }
  return x;
}
CLH_D void pow_grad_primal(const double* s, double a, double* g) {  // :333-356
  double phi = pow(s[0], 2 * a) * pow(s[1], 2 - 2 * a);
  double abs_s = fabs(s[2]);
  if (abs_s > kDEps) {
    g[2] = pow_newton(abs_s, phi, a);
    if (s[2] < 0) g[2] = -g[2];
    g[0] = -(a * g[2] * s[2] + 1 + a) / s[0];
    g[1] = -((1 - a) * g[2] * s[2] + 2 - a) / s[1];
  } else {
    g[2] = 0;
    g[0] = -(1 + a) / s[0];
    g[1] = -(2 - a) / s[1];
  }
}
CLH_D void pow_dual_grad_H(const double* z, double a, double* grad, double* H) {  // :417-445
  double phi = pow(z[0] / a, 2 * a) * pow(z[1] / (1 - a), 2 - 2 * a);
  double psi = phi - z[2] * z[2];
  double g0 = 2 * a * phi / (z[0] * psi), g1 = 2 * (1 - a) * phi / (z[1] * psi), g2 = -2 * z[2] / psi;
  H[0] = g0 * g0 - 2 * a * (2 * a - 1) * phi / (z[0] * z[0] * psi) + (1 - a) / (z[0] * z[0]);
  H[3] = g0 * g1 - 4 * a * (1 - a) * phi / (z[0] * z[1] * psi); H[1] = H[3];
  H[4] = g1 * g1 - 2 * (1 - a) * (1 - 2 * a) * phi / (z[1] * z[1] * psi) + a / (z[1] * z[1]);
  H[6] = g0 * g2; H[2] = H[6];
  H[7] = g1 * g2; H[5] = H[7];
  H[8] = g2 * g2 + 2 / psi;
  grad[0] = -2 * a * phi / (z[0] * psi) - (1 - a) / z[0];
  grad[1] = -2 * (1 - a) * phi / (z[1] * psi) - a / z[1];
  grad[2] = 2 * z[2] / psi;
}
CLH_D bool pow_primal_feas(const double* s, double a) {
  if (s[0] > 0 && s[1] > 0) return exp(2 * a * logsafe(s[0]) + 2 * (1 - a) * logsafe(s[1])) - s[2] * s[2] > 0;
  return false;
}
CLH_D bool pow_dual_feas(const double* z, double a) {
  if (z[0] > 0 && z[1] > 0)
    return exp(2 * a * logsafe(z[0] / a) + 2 * (1 - a) * logsafe(z[1] / (1 - a))) - z[2] * z[2] > 0;
  return false;
}
CLH_D double pow_barrier(const double* z, const double* s, double a) {  // :283-299
  double bd = -logsafe(pow(z[0] / a, 2 * a) * pow(z[1] / (1 - a), 2 - 2 * a) - z[2] * z[2]) - (1 - a) * logsafe(z[0]) -
              a * logsafe(z[1]);
  double g[3];
  pow_grad_primal(s, a, g);
  double bp = logsafe(pow(-g[0] / a, 2 * a) * pow(-g[1] / (1 - a), 2 - 2 * a) - g[2] * g[2]) +
              (1 - a) * logsafe(-g[0]) + a * logsafe(-g[1]) - 3;
  return bd + bp;
}

// 3x3 explicit Cholesky (src/utils/mathutils.jl:381-431), column-major
CLH_D bool chol3(double* L, const double* A) {
  for (int i = 0; i < 9; ++i) L[i] = 0;
  double t = A[0];
  if (t <= 0) return false;
  L[0] = sqrt(t);
  L[1] = A[1] / L[0];
  t = A[4] - L[1] * L[1];
  if (t <= 0) return false;
  L[4] = sqrt(t);
  L[2] = A[2] / L[0];
  L[5] = (A[5] - L[1] * L[2]) / L[4];
  t = A[8] - L[2] * L[2] - L[5] * L[5];
  if (t <= 0) return false;
  L[8] = sqrt(t);
  return true;
}
CLH_D void chol3_solve(const double* L, const double* b, double* x) {
  double L11 = L[0], L21 = L[1], L31 = L[2], L22 = L[4], L32 = L[5], L33 = L[8];
  double c1 = b[0] / L11;
  double c2 = (b[1] * L11 - b[0] * L21) / (L11 * L22);
  double c3 = (b[2] * L11 * L22 - b[1] * L11 * L32 + b[0] * L21 * L32 - b[0] * L22 * L31) / (L11 * L22 * L33);
  x[0] = (c1 * L22 * L33 - c2 * L21 * L33 + c3 * L21 * L32 - c3 * L22 * L31) / (L11 * L22 * L33);
  x[1] = (c2 * L33 - c3 * L32) / (L22 * L33);
  x[2] = c3 / L33;
}
// third-order correction, exp (coneops_expcone.jl:382-428) and pow (coneops_powcone.jl:359-414)
CLH_D void higher_correction(bool isexp, double a, const double* H, const double* z, double* eta, const double* ds,
                             const double* v) {
  double L[9], u[3];
  if (!chol3(L, H)) { eta[0] = eta[1] = eta[2] = 0; return; }
  chol3_solve(L, ds, u);
  if (isexp) {
    eta[1] = 1;
    eta[2] = -z[0] / z[2];
    eta[0] = logsafe(eta[2]);
    double psi = z[0] * eta[0] - z[0] + z[1];
    double du = eta[0] * u[0] + eta[1] * u[1] + eta[2] * u[2];
    double dv = eta[0] * v[0] + eta[1] * v[1] + eta[2] * v[2];
    double coef = ((u[0] * (v[0] / z[0] - v[2] / z[2]) + u[2] * (z[0] * v[2] / z[2] - v[0]) / z[2]) * psi - 2 * du * dv) /
                  (psi * psi * psi);
    for (int i = 0; i < 3; ++i) eta[i] *= coef;
    double ip2 = 1 / psi / psi;
    eta[0] += (1 / psi - 2 / z[0]) * u[0] * v[0] / (z[0] * z[0]) - u[2] * v[2] / (z[2] * z[2]) / psi +
              du * ip2 * (v[0] / z[0] - v[2] / z[2]) + dv * ip2 * (u[0] / z[0] - u[2] / z[2]);
    eta[2] += 2 * (z[0] / psi - 1) * u[2] * v[2] / (z[2] * z[2] * z[2]) - (u[2] * v[0] + u[0] * v[2]) / (z[2] * z[2]) / psi +
              du * ip2 * (z[0] * v[2] / (z[2] * z[2]) - v[0] / z[2]) + dv * ip2 * (z[0] * u[2] / (z[2] * z[2]) - u[0] / z[2]);
    for (int i = 0; i < 3; ++i) eta[i] /= 2;
  } else {
    double phi = pow(z[0] / a, 2 * a) * pow(z[1] / (1 - a), 2 - 2 * a);
    double psi = phi - z[2] * z[2];
    eta[0] = 2 * a * phi / z[0];
    eta[1] = 2 * (1 - a) * phi / z[1];
    eta[2] = -2 * z[2];
    double H11 = 2 * a * (2 * a - 1) * phi / (z[0] * z[0]);
    double H12 = 4 * a * (1 - a) * phi / (z[0] * z[1]);
    double H22 = 2 * (1 - a) * (1 - 2 * a) * phi / (z[1] * z[1]);
    double du = eta[0] * u[0] + eta[1] * u[1] + eta[2] * u[2];
    double dv = eta[0] * v[0] + eta[1] * v[1] + eta[2] * v[2];
    double Hv[3] = {H11 * v[0] + H12 * v[1], H12 * v[0] + H22 * v[1], -2 * v[2]};
    double coef = ((u[0] * Hv[0] + u[1] * Hv[1] + u[2] * Hv[2]) * psi - 2 * du * dv) / (psi * psi * psi);
    double coef2 = 4 * a * (2 * a - 1) * (1 - a) * phi * (u[0] / z[0] - u[1] / z[1]) * (v[0] / z[0] - v[1] / z[1]) / psi;
    double ip2 = 1 / psi / psi;
    eta[0] = coef * eta[0] - 2 * (1 - a) * u[0] * v[0] / (z[0] * z[0] * z[0]) + coef2 / z[0] + Hv[0] * du * ip2;
    eta[1] = coef * eta[1] - 2 * a * u[1] * v[1] / (z[1] * z[1] * z[1]) - coef2 / z[1] + Hv[1] * du * ip2;
    eta[2] = coef * eta[2] + Hv[2] * du * ip2;
    double Hu[3] = {H11 * u[0] + H12 * u[1], H12 * u[0] + H22 * u[1], -2 * u[2]};
    for (int i = 0; i < 3; ++i) eta[i] = (eta[i] + Hu[i] * dv * ip2) / 2;
  }
}
// update_Hs (src/cones/coneops_nonsymmetric_common.jl:64-218); ns = Hd[9] Hs[9] grad[3] zc[3]
CLH_D void nonsym_scale(bool isexp, double a, double* ns, const double* s, const double* z, double mu, int dual) {
  double* Hd = ns;
  double* Hs = ns + 9;
  double* grad = ns + 18;
  if (isexp) exp_dual_grad_H(z, grad, Hd);
  else pow_dual_grad_H(z, a, grad, Hd);
  for (int i = 0; i < 3; ++i) ns[21 + i] = z[i];
  if (dual) {
    for (int i = 0; i < 9; ++i) Hs[i] = mu * Hd[i];
    return;
  }
  double st[3] = {grad[0], grad[1], grad[2]};
  double zt[3];
  if (isexp) exp_grad_primal(s, zt);
  else pow_grad_primal(s, a, zt);
  double dot_sz = z[0] * s[0] + z[1] * s[1] + z[2] * s[2];
  double m = dot_sz / 3;
  double mut = (zt[0] * st[0] + zt[1] * st[1] + zt[2] * st[2]) / 3;
  double dS[3], dZ[3];
  for (int i = 0; i < 3; ++i) { dS[i] = s[i] + m * st[i]; dZ[i] = z[i] + m * zt[i]; }
  double dot_dsz = dS[0] * dZ[0] + dS[1] * dZ[1] + dS[2] * dZ[2];
  double de1 = m * mut - 1;
  double Hzt[3];
  for (int i = 0; i < 3; ++i) Hzt[i] = Hd[i] * zt[0] + Hd[i + 3] * zt[1] + Hd[i + 6] * zt[2];
  double de2 = zt[0] * Hzt[0] + zt[1] * Hzt[1] + zt[2] * Hzt[2] - 3 * mut * mut;
  if (fabs(de1) > 1.4901161193847656e-08 && fabs(de2) > kDEps && dot_sz > 0 && dot_dsz > 0) {
    double tmp[3];
    for (int i = 0; i < 3; ++i) tmp[i] = mut * st[i] - Hd[i] * zt[0] - Hd[i + 3] * zt[1] - Hd[i + 6] * zt[2];
    double W[9];
    for (int i = 0; i < 9; ++i) W[i] = Hd[i];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) W[i + 3 * j] -= st[i] * st[j] / 3 + tmp[i] * tmp[j] / de2;
    double fro = 0;
    for (int i = 0; i < 9; ++i) fro += W[i] * W[i];
    double t = m * sqrt(fro);
    double ax[3] = {z[1] * zt[2] - z[2] * zt[1], z[2] * zt[0] - z[0] * zt[2], z[0] * zt[1] - z[1] * zt[0]};
    double an = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    for (int i = 0; i < 3; ++i) ax[i] /= an;
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j) W[i + 3 * j] = s[i] * s[j] / dot_sz + dS[i] * dS[j] / dot_dsz + t * ax[i] * ax[j];
    W[1] = W[3]; W[2] = W[6]; W[5] = W[7];
    for (int i = 0; i < 9; ++i) Hs[i] = W[i];
  } else {
    for (int i = 0; i < 9; ++i) Hs[i] = m * Hd[i];
  }
}
// backtrack_search (coneops_nonsymmetric_common.jl:9-45); the trip count is
// bounded so that a NaN step length cannot loop forever
template <class P, class F>
CLH_D double backtrack(P dq, P q, double a, double amin, double step, F feas) {
  double wq[3];
  for (int it = 0;; ++it) {
    for (int i = 0; i < 3; ++i) wq[i] = q[i] + a * dq[i];
    if (feas(wq)) break;
    if ((a *= step) < amin || it >= 200) { a = 0; break; }
  }
  return a;
}

// ---- positive semidefinite cone (src/cones/coneops_psdtrianglecone.jl).  One
// wavefront per cone (one thread in the serial builds): a team of lanes works
// on the cone's dense n x n matrices (column-major).  H = R R' (kept from the
// scaling update to the KKT write) lives in the cone's work area of the slot
// arena; the six scratch matrices M1..M6 live in the wavefront's share of the
// LDS area when they fit there, else in the work area after H.  Every entry
// is computed by one lane with the operations, in the order, of the serial
// loop that defines it, so the results do not depend on the team's size;
// CLH_WSYNC() makes one step's writes visible to the next step's readers.
struct Team {
  int l, nl;  // this lane, the team's lanes
};
CLH_D Team psd_team(const Blk& B) {
#if defined(CLH_SERIAL)
  (void)B;
  return Team{0, 1};
#else
  return Team{B.lane, CLH_WS};
#endif
}
// for (k ...) over the PSD cones of this team: one team per wavefront
#if defined(CLH_SERIAL)
#define CLH_PSD_FOR(B, k, cnt) for (int k = (B).tid; k < (cnt); k += (B).nthr)
#else
#define CLH_PSD_FOR(B, k, cnt) for (int k = (B).tid / CLH_WS; k < (cnt); k += (B).nthr / CLH_WS)
#endif

template <class P, class Q>
CLH_D void psd_svec_to_mat(const Team& tm, int n, P x, Q M) {  // :468-482
  const double is2 = 1 / sqrt(2.0);
  for (int e = tm.l; e < n * n; e += tm.nl) {
    const int r = e % n, c = e / n;
    if (r == c) M[e] = x[tri(c) + r];
    else M[e] = x[r < c ? tri(c) + r : tri(r) + c] * is2;
  }
  CLH_WSYNC();
}
template <class P, class Q>
CLH_D void psd_mat_to_svec(const Team& tm, int n, P M, Q x) {  // :485-494
  const double is2 = 1 / sqrt(2.0);
  for (int e = tm.l; e < n * n; e += tm.nl) {
    const int r = e % n, c = e / n;
    if (r <= c) x[tri(c) + r] = (r == c) ? M[e] : (M[e] + M[c + r * n]) * is2;
  }
  CLH_WSYNC();
}
// Cholesky A = L L' (lower, zero above); false on a non-positive pivot (the
// same pivot on every lane)
template <class P, class Q>
CLH_D bool psd_chol(const Team& tm, int n, P A, Q L) {
  for (int e = tm.l; e < n * n; e += tm.nl) L[e] = 0;
  CLH_WSYNC();
  for (int j = 0; j < n; ++j) {
    double s = A[j + j * n];
    for (int k = 0; k < j; ++k) s -= L[j + k * n] * L[j + k * n];
    if (!(s > 0)) return false;
    const double d = sqrt(s);
    if (tm.l == 0) L[j + j * n] = d;
    for (int i = j + 1 + tm.l; i < n; i += tm.nl) {
      double t = A[i + j * n];
      for (int k = 0; k < j; ++k) t -= L[i + k * n] * L[j + k * n];
      L[i + j * n] = t / d;
    }
    CLH_WSYNC();
  }
  return true;
}
// C = alpha op(A) op(B) + beta C, each entry's sum over k in order; C must
// not alias A or B
template <class PA, class PB, class PC>
CLH_D void psd_gemm(const Team& tm, int n, bool ta, bool tb, double alpha, PA A, PB B, double beta, PC C) {
  for (int e = tm.l; e < n * n; e += tm.nl) {
    const int i = e % n, j = e / n;
    double s = 0;
    for (int k = 0; k < n; ++k) {
      const double a = ta ? A[k + i * n] : A[i + k * n];
      const double b = tb ? B[j + k * n] : B[k + j * n];
      s += a * b;
    }
    C[e] = alpha * s + (beta == 0 ? 0.0 : beta * C[e]);
  }
  CLH_WSYNC();
}
// eigenvalues (ascending) of the symmetric A by cyclic Jacobi; A is overwritten.
// Each rotation sets its pair to zero; from the fifth sweep on, an
// off-diagonal entry that a hundredfold would not change either of its two
// diagonal entries is set to zero without a rotation, and the sweeps end when
// every off-diagonal entry is zero (the round-off of the rotations keeps a
// relative test on the off-diagonal norm from ever being met on dense
// matrices).  A rotation of the pair (p, q) multiplies columns p, q and then
// rows p, q by [c s; -s c]: outside the 2 x 2 block the column and the row
// steps touch disjoint entries, so lane k does both for its k, and lane 0
// does the block in the order of the two steps.
template <class P, class Q>
CLH_D void psd_sym_eig(const Team& tm, int n, P A, Q ev) {
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int j = 1; j < n; ++j)
      for (int i = 0; i < j; ++i) off += fabs(A[i + j * n]);
    if (off == 0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[p + q * n];
        const double app = A[p + p * n], aqq = A[q + q * n];
        const double g = 100 * fabs(apq);
        if (sweep > 3 && fabs(app) + g == fabs(app) && fabs(aqq) + g == fabs(aqq)) {
          CLH_WSYNC();  // every lane has read the pair
          if (tm.l == 0) {
            A[p + q * n] = 0;
            A[q + p * n] = 0;
          }
          CLH_WSYNC();
          continue;
        }
        if (apq == 0) continue;
        const double theta = (aqq - app) / (2 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
        const double c = 1 / sqrt(t * t + 1), s = t * c;
        const double aqp = A[q + p * n];
        CLH_WSYNC();  // every lane has read the block before it changes
        for (int k = tm.l; k < n; k += tm.nl) {
          if (k == p || k == q) continue;
          const double akp = A[k + p * n], akq = A[k + q * n];
          A[k + p * n] = c * akp - s * akq;
          A[k + q * n] = s * akp + c * akq;
          const double apk = A[p + k * n], aqk = A[q + k * n];
          A[p + k * n] = c * apk - s * aqk;
          A[q + k * n] = s * apk + c * aqk;
        }
        if (tm.l == 0) {
          const double cpp = c * app - s * apq, cpq = s * app + c * apq;  // columns, k = p
          const double cqp = c * aqp - s * aqq, cqq = s * aqp + c * aqq;  // columns, k = q
          A[p + p * n] = c * cpp - s * cqp;                               // rows, k = p
          A[q + q * n] = s * cpq + c * cqq;                               // rows, k = q
          A[p + q * n] = 0;
          A[q + p * n] = 0;
        }
        CLH_WSYNC();
      }
  }
  if (tm.l == 0) {
    for (int i = 0; i < n; ++i) ev[i] = A[i + i * n];
    for (int i = 1; i < n; ++i) {  // insertion sort, ascending
      const double v = ev[i];
      int j = i - 1;
      while (j >= 0 && ev[j] > v) { ev[j + 1] = ev[j]; --j; }
      ev[j + 1] = v;
    }
  }
  CLH_WSYNC();
}
// one-sided (Hestenes) Jacobi SVD A = U diag(S) V', S descending; U holds A on
// entry.  Every lane forms the pair's three sums in order (the same rotation
// on every lane); lane k rotates row k of U and V.
template <class P, class Q>
CLH_D void psd_svd(const Team& tm, int n, P U, Q S, P V) {
  for (int e = tm.l; e < n * n; e += tm.nl) V[e] = (e % n == e / n) ? 1.0 : 0.0;
  CLH_WSYNC();
  for (int sweep = 0; sweep < 100; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double al = 0, be = 0, ga = 0;
        for (int k = 0; k < n; ++k) {
          al += U[k + p * n] * U[k + p * n];
          be += U[k + q * n] * U[k + q * n];
          ga += U[k + p * n] * U[k + q * n];
        }
        if (fabs(ga) <= 1e-15 * sqrt(al * be) || ga == 0) continue;
        rotated = true;
        const double zeta = (be - al) / (2 * ga);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
        const double c = 1 / sqrt(1 + t * t), s = c * t;
        CLH_WSYNC();  // every lane has read columns p, q
        for (int k = tm.l; k < n; k += tm.nl) {
          const double up = U[k + p * n], uq = U[k + q * n];
          U[k + p * n] = c * up - s * uq;
          U[k + q * n] = s * up + c * uq;
          const double vp = V[k + p * n], vq = V[k + q * n];
          V[k + p * n] = c * vp - s * vq;
          V[k + q * n] = s * vp + c * vq;
        }
        CLH_WSYNC();
      }
    if (!rotated) break;
  }
  for (int j = tm.l; j < n; j += tm.nl) {
    double s = 0;
    for (int k = 0; k < n; ++k) s += U[k + j * n] * U[k + j * n];
    S[j] = sqrt(s);
  }
  CLH_WSYNC();
  for (int e = tm.l; e < n * n; e += tm.nl) {
    const double s = S[e / n];
    if (s > 0) U[e] /= s;
  }
  CLH_WSYNC();
  // descending order, equal values keeping their order: select the first
  // largest remaining value and swap it into place
  for (int j = 0; j < n; ++j) {
    int b = j;
    for (int i = j + 1; i < n; ++i)
      if (S[i] > S[b]) b = i;
    if (b != j) {
      CLH_WSYNC();  // every lane has read S
      if (tm.l == 0) {
        const double t = S[j]; S[j] = S[b]; S[b] = t;
      }
      for (int k = tm.l; k < n; k += tm.nl) {
        double u = U[k + j * n]; U[k + j * n] = U[k + b * n]; U[k + b * n] = u;
        double v = V[k + j * n]; V[k + j * n] = V[k + b * n]; V[k + b * n] = v;
      }
      CLH_WSYNC();
    }
  }
}

// the state and work of PSD cone k
struct PsdCone {
  int n, o, t;
  VD R, Rinv, lam, lisq;  // state
  VD H;                   // R R'
  VD M1, M2, M3, M4, M5, M6;  // scratch (LDS or the work area)
};
// the scratch in this wavefront's share of the LDS area when six n x n
// matrices fit there
CLH_D PsdCone psd_cone(const Blk& B, const SD& D, int k) {
  PsdCone c;
  c.n = D.psd_n[k];
  c.o = D.psd_off[k];
  c.t = tri(c.n);
  const int n2 = c.n * c.n;
  VD st = D.psd_st + D.psd_soff[k];
  c.R = st; c.Rinv = st + n2; c.lam = st + 2 * n2; c.lisq = st + 2 * n2 + c.n;
  VD wk = D.psd_wk + D.psd_woff[k];
  c.H = wk;
  VD m = wk + n2;
#if !defined(CLH_LANES)
  const int share = B.lds ? B.lds_bytes / B.nwaves / (int)sizeof(double) : 0;
  if (6 * n2 <= share) m = (double*)B.lds + (size_t)B.wave * share;
#else
  (void)B;
#endif
  c.M1 = m; c.M2 = m + n2; c.M3 = m + 2 * n2; c.M4 = m + 3 * n2; c.M5 = m + 4 * n2; c.M6 = m + 5 * n2;
  return c;
}
// y = svec(R' mat(x) R) or, transposed, svec(R mat(x) R') (:425-460); uses M1..M3
template <class PY, class PX>
CLH_D void psd_mul_W(const Team& tm, const PsdCone& c, bool transpose, PY y, PX x, VD Rx) {
  const int n = c.n;
  psd_svec_to_mat(tm, n, x, c.M1);
  if (transpose) {
    psd_gemm(tm, n, false, true, 1.0, c.M1, Rx, 0.0, c.M2);  // X R'
    psd_gemm(tm, n, false, false, 1.0, Rx, c.M2, 0.0, c.M3);  // R X R'
  } else {
    psd_gemm(tm, n, true, false, 1.0, Rx, c.M1, 0.0, c.M2);   // R' X
    psd_gemm(tm, n, false, false, 1.0, c.M2, Rx, 0.0, c.M3);  // R' X R
  }
  psd_mat_to_svec(tm, n, c.M3, y);
}
// update_scaling (:109-160)
CLH_D bool psd_update_scaling(const Team& tm, const PsdCone& c, VCD s, VCD z) {
  const int n = c.n;
  VD S = c.M1, Z = c.M2, L1 = c.M3, L2 = c.M4, U = c.M5, V = c.M6;
  psd_svec_to_mat(tm, n, s + c.o, S);
  psd_svec_to_mat(tm, n, z + c.o, Z);
  if (!psd_chol(tm, n, S, L1) || !psd_chol(tm, n, Z, L2)) return false;
  psd_gemm(tm, n, true, false, 1.0, L2, L1, 0.0, U);  // L2' L1, decomposed in place
  psd_svd(tm, n, U, c.lam, V);
  for (int i = tm.l; i < n; i += tm.nl) c.lisq[i] = 1 / sqrt(c.lam[i]);
  CLH_WSYNC();
  psd_gemm(tm, n, false, false, 1.0, L1, V, 0.0, c.R);  // R = L1 V diag(lambda)^{-1/2}
  for (int e = tm.l; e < n * n; e += tm.nl) c.R[e] *= c.lisq[e / n];
  psd_gemm(tm, n, true, true, 1.0, U, L2, 0.0, c.Rinv);  // Rinv = diag(lambda)^{-1/2} U' L2'
  for (int e = tm.l; e < n * n; e += tm.nl) c.Rinv[e] *= c.lisq[e % n];
  CLH_WSYNC();
  psd_gemm(tm, n, false, true, 1.0, c.R, c.R, 0.0, c.H);  // R R', symmetric from its upper triangle
  for (int e = tm.l; e < n * n; e += tm.nl) {
    const int i = e % n, j = e / n;
    if (i > j) c.H[e] = c.H[j + i * n];
  }
  CLH_WSYNC();
  return true;
}
// entry (row, col) of skron(A), A = R R' symmetric, row <= col: the packed
// upper triangle by columns, column col = tri(l) + kk (kk <= l) and row
// = tri(j) + i (i <= j)
CLH_D void psd_skron_column(int col, int& l, int& kk) {
  l = 0;
  while (tri(l + 1) <= col) ++l;
  kk = col - tri(l);
}
// -Hs = -skron(R R') (:499-546) into the KKT values, packed upper triangle by
// columns; lane l writes columns l, l + nl, ...
CLH_D void psd_write_Hs(const Team& tm, const SD& D, const PsdCone& c, int k) {
  const int n = c.n;
  const double sqrt2 = sqrt(2.0);
  const int32_t* ki = D.psd_kidx + D.psd_kblk[k];
  VCD A = c.H;
  for (int col = tm.l; col < c.t; col += tm.nl) {
    int l, kk;
    psd_skron_column(col, l, kk);
    int h = tri(col), row = 0;
    const bool kl_eq = (kk == l);
    for (int j = 0; j < n && row <= col; ++j) {
      const double Ajl = A[j + l * n], Ajk = A[j + kk * n];
      for (int i = 0; i <= j && row <= col; ++i) {
        const bool ij_eq = (i == j);
        double v;
        if (!ij_eq && !kl_eq) v = A[i + kk * n] * Ajl + A[i + l * n] * Ajk;
        else if (ij_eq && !kl_eq) v = sqrt2 * Ajl * Ajk;
        else if (!ij_eq && kl_eq) v = sqrt2 * A[i + l * n] * Ajk;
        else v = Ajl * Ajl;
        D.Kx[ki[h++]] = -v;
        ++row;
      }
    }
  }
}
// minimum eigenvalue of lambda^{-1/2} mat(d) lambda^{-1/2} step bound (:497-521)
template <class P>
CLH_D double psd_step(const Team& tm, const PsdCone& c, P dsvec, double amax) {
  const int n = c.n;
  if (n == 0) return amax;  // an empty cone bounds nothing (and has no eigenvalue to read)
  psd_svec_to_mat(tm, n, dsvec, c.M5);
  for (int e = tm.l; e < n * n; e += tm.nl) c.M5[e] *= c.lisq[e % n] * c.lisq[e / n];
  CLH_WSYNC();
  psd_sym_eig(tm, n, c.M5, c.M6);
  const double g = c.M6[0];
  return g < 0 ? dmin(1 / (-g), amax) : amax;
}
// log det of mat(x + a dx) (2 log of the Cholesky diagonal), +inf when not
// positive definite; the same value on every lane
CLH_D double psd_logdet(const Team& tm, const PsdCone& c, VCD x, VCD dx, double a) {
  const int n = c.n;
  for (int i = tm.l; i < c.t; i += tm.nl) c.M6[i] = x[c.o + i] + a * dx[c.o + i];
  CLH_WSYNC();
  psd_svec_to_mat(tm, n, c.M6, c.M4);
  if (!psd_chol(tm, n, c.M4, c.M5)) return dinf();
  double ld = 0;
  for (int i = 0; i < n; ++i) ld += 2 * log(c.M5[i + i * n]);
  return ld;
}

// ---------------------------------------------------------------------------
// composite cone operations (src/cones/coneops_compositecone.jl)
// ---------------------------------------------------------------------------
// update_scaling! over this thread's share of the cones; 1 when all of them
// succeeded
CLH_D double cones_update_scaling_part(const Blk& B, const SD& D, VCD s, VCD z, double mu, int dual) {
  for (int i = B.tid; i < D.m; i += B.nthr)
    if (D.etype[i] == CT_NONNEG) {
      D.lam[i] = sqrt(s[i] * z[i]);
      D.w[i] = sqrt(s[i] / z[i]);
    }
  double ok = 1;
  for (int k = B.tid; k < D.nsoc; k += B.nthr) {
    int o = D.soc_off[k], d = D.soc_dim[k], sp = D.soc_sp[k];
    double eta;
    VD none = nullptr;
    bool r = soc_scale(s + o, z + o, d, D.w + o, D.lam + o, eta, sp, sp >= 0 ? D.soc_d + k : none,
                       sp >= 0 ? D.sp_u + D.sp_off[sp] : none, sp >= 0 ? D.sp_v + D.sp_off[sp] : none);
    D.soc_eta[k] = eta;
    if (!r) ok = 0;
  }
  for (int k = B.tid; k < D.nexp + D.npow; k += B.nthr) {
    const bool isexp = k < D.nexp;
    const int o = isexp ? D.exp_off[k] : D.pow_off[k - D.nexp];
    double sl[3], zl[3], ns[24];
    for (int i = 0; i < 3; ++i) { sl[i] = s[o + i]; zl[i] = z[o + i]; }
    nonsym_scale(isexp, isexp ? 0.0 : D.pow_alpha[k - D.nexp], ns, sl, zl, mu, dual);
    VD out = D.nsym + 24 * k;
    for (int i = 0; i < 24; ++i) out[i] = ns[i];
  }
  CLH_PSD_FOR(B, k, D.npsd)
    if (!psd_update_scaling(psd_team(B), psd_cone(B, D, k), s, z)) ok = 0;
  return ok;
}
// update_scaling! ; returns success on every thread
CLH_D bool cones_update_scaling(const Blk& B, const SD& D, VCD s, VCD z, double mu, int dual) {
  PhaseTimer pt(B, lane_prof(D), 2);
  return blk_min(B, cones_update_scaling_part(B, D, s, z, mu, dual)) > 0;
}

// set_identity_scaling! for symmetric cones
CLH_D void cones_identity_scaling(const Blk& B, const SD& D) {
  for (int i = B.tid; i < D.m; i += B.nthr)
    if (D.etype[i] == CT_NONNEG) D.w[i] = 1.0;
  for (int k = B.tid; k < D.nsoc; k += B.nthr) {
    int o = D.soc_off[k], d = D.soc_dim[k], sp = D.soc_sp[k];
    for (int i = 0; i < d; ++i) D.w[o + i] = 0;
    D.w[o] = 1;
    D.soc_eta[k] = 1;
    if (sp >= 0) {
      D.soc_d[k] = 0.5;
      VD u = D.sp_u + D.sp_off[sp];
      VD v = D.sp_v + D.sp_off[sp];
      for (int i = 0; i < d; ++i) { u[i] = 0; v[i] = 0; }
      u[0] = sqrt(0.5);
    }
  }
  CLH_PSD_FOR(B, k, D.npsd) {  // R = Rinv = I, so Hs = skron(I) = I
    const Team tm = psd_team(B);
    const PsdCone c = psd_cone(B, D, k);
    for (int e = tm.l; e < c.n * c.n; e += tm.nl) {
      const double v = (e % c.n == e / c.n) ? 1.0 : 0.0;
      c.R[e] = v; c.Rinv[e] = v; c.H[e] = v;
    }
  }
  CLH_SYNC();
}

// get_Hs! + the scatter of -Hs into the KKT values
// (kktsolver_directldl.jl:195-236, directldl_datamaps.jl:64-80)
CLH_D void kkt_write_Hs(const Blk& B, const SD& D) {
  for (int i = B.tid; i < D.m; i += B.nthr) {
    int8_t t = D.etype[i];
    if (t == CT_ZERO) D.Kx[D.lin_kdiag[i]] = -0.0;
    else if (t == CT_NONNEG) D.Kx[D.lin_kdiag[i]] = -(D.w[i] * D.w[i]);
  }
  for (int k = B.tid; k < D.nsoc; k += B.nthr) {
    int o = D.soc_off[k], d = D.soc_dim[k], sp = D.soc_sp[k], kb = D.soc_kblk[k];
    double eta2 = D.soc_eta[k] * D.soc_eta[k];
    VCD w = D.w + o;
    if (sp < 0) {
      // packed triu of eta^2 (2 w w' - J) (coneops_socone.jl:160-178)
      const double s2 = 1.4142135623730951;
      D.Kx[D.soc_kidx[kb]] = -((s2 * w[0] - 1) * (s2 * w[0] + 1) * eta2);
      int h = 1;
      for (int col = 1; col < d; ++col)
        for (int row = 0; row <= col; ++row) {
          double v = 2 * w[row] * w[col];
          if (row == col) v += 1;
          D.Kx[D.soc_kidx[kb + h]] = -(v * eta2);
          ++h;
        }
    } else {
      for (int i = 0; i < d; ++i) D.Kx[D.soc_kidx[kb + i]] = -eta2;
      D.Kx[D.soc_kidx[kb]] = -(eta2 * D.soc_d[k]);
      int so = D.sp_off[sp];
      for (int i = 0; i < d; ++i) {
        D.Kx[D.sp_ku[so + i]] = D.sp_u[so + i] * -eta2;
        D.Kx[D.sp_kv[so + i]] = D.sp_v[so + i] * -eta2;
      }
      D.Kx[D.sp_kD[2 * sp]] = -eta2;
      D.Kx[D.sp_kD[2 * sp + 1]] = +eta2;
    }
  }
  for (int k = B.tid; k < D.nexp + D.npow; k += B.nthr) {
    VCD Hs = D.nsym + 24 * k + 9;
    const int32_t* ki = (k < D.nexp) ? D.exp_kidx + 6 * k : D.pow_kidx + 6 * (k - D.nexp);
    int h = 0;
    for (int col = 0; col < 3; ++col)
      for (int row = 0; row <= col; ++row) D.Kx[ki[h++]] = -Hs[row + 3 * col];
  }
  CLH_PSD_FOR(B, k, D.npsd) psd_write_Hs(psd_team(B), D, psd_cone(B, D, k), k);
  CLH_SYNC();
}

// get_Hs! (:84-109; src/gpucones/coneops_compositecone_gpu.jl:252-309): the
// Hs blocks packed in cone order, a block of d entries (its diagonal) for the
// zero, nonnegative and sparse-expanded second-order cones and d x d
// column-major otherwise.  hs_row gives the block position of each zero or
// nonnegative row, hs_soc / hs_nsym / hs_psd the block start of each
// second-order, exponential-or-power and PSD cone.
CLH_D void cones_get_Hs(const Blk& B, const SD& D, VD out, const int64_t* hs_row, const int64_t* hs_soc,
                        const int64_t* hs_nsym, const int64_t* hs_psd) {
  for (int i = B.tid; i < D.m; i += B.nthr) {
    int8_t t = D.etype[i];
    if (t == CT_ZERO) out[hs_row[i]] = 0.0;
    else if (t == CT_NONNEG) out[hs_row[i]] = D.w[i] * D.w[i];
  }
  for (int k = B.tid; k < D.nsoc; k += B.nthr) {
    const int o = D.soc_off[k], d = D.soc_dim[k], sp = D.soc_sp[k];
    const int64_t h0 = hs_soc[k];
    const double eta2 = D.soc_eta[k] * D.soc_eta[k];
    VCD w = D.w + o;
    if (sp < 0) {  // eta^2 (2 w w' - J)
      const double s2 = 1.4142135623730951;
      for (int col = 0; col < d; ++col)
        for (int row = 0; row < d; ++row) {
          double v;
          if (row == 0 && col == 0) v = (s2 * w[0] - 1) * (s2 * w[0] + 1) * eta2;
          else {
            v = 2 * w[row] * w[col];
            if (row == col) v += 1;
            v *= eta2;
          }
          out[h0 + row + (int64_t)col * d] = v;
        }
    } else {
      out[h0] = eta2 * D.soc_d[k];
      for (int i = 1; i < d; ++i) out[h0 + i] = eta2;
    }
  }
  for (int k = B.tid; k < D.nexp + D.npow; k += B.nthr) {
    VCD Hs = D.nsym + 24 * k + 9;
    for (int i = 0; i < 9; ++i) out[hs_nsym[k] + i] = Hs[i];
  }
  CLH_PSD_FOR(B, k, D.npsd) {  // skron(R R'), the entries psd_write_Hs writes
    const Team tm = psd_team(B);
    const PsdCone c = psd_cone(B, D, k);
    const int n = c.n, t = c.t;
    const int64_t h0 = hs_psd[k];
    const double sqrt2 = sqrt(2.0);
    VCD A = c.H;
    for (int col = tm.l; col < t; col += tm.nl) {
      int l, kk;
      psd_skron_column(col, l, kk);
      {
        int row = 0;
        const bool kl_eq = (kk == l);
        for (int j = 0; j < n && row <= col; ++j) {
          const double Ajl = A[j + l * n], Ajk = A[j + kk * n];
          for (int i = 0; i <= j && row <= col; ++i) {
            const bool ij_eq = (i == j);
            double v;
            if (!ij_eq && !kl_eq) v = A[i + kk * n] * Ajl + A[i + l * n] * Ajk;
            else if (ij_eq && !kl_eq) v = sqrt2 * Ajl * Ajk;
            else if (!ij_eq && kl_eq) v = sqrt2 * A[i + l * n] * Ajk;
            else v = Ajl * Ajl;
            out[h0 + row + (int64_t)col * t] = v;
            out[h0 + col + (int64_t)row * t] = v;
            ++row;
          }
        }
      }
    }
  }
  CLH_SYNC();
}

// y = Hs x (mul_Hs!, :112-124)
CLH_D void cones_mul_Hs(const Blk& B, const SD& D, VD y, VCD x) {
  for (int i = B.tid; i < D.m; i += B.nthr) {
    int8_t t = D.etype[i];
    if (t == CT_ZERO) y[i] = 0;
    else if (t == CT_NONNEG) y[i] = D.w[i] * (D.w[i] * x[i]);
  }
  for (int k = B.tid; k < D.nsoc; k += B.nthr) {
    int o = D.soc_off[k], d = D.soc_dim[k];
    VCD w = D.w + o;
    double c = 0;
    for (int i = 0; i < d; ++i) c += w[i] * x[o + i];
    c *= 2;
    double e2 = D.soc_eta[k] * D.soc_eta[k];
    y[o] = (-x[o] + c * w[0]) * e2;
    for (int i = 1; i < d; ++i) y[o + i] = (x[o + i] + c * w[i]) * e2;
  }
  for (int k = B.tid; k < D.nexp + D.npow; k += B.nthr) {
    int o = (k < D.nexp) ? D.exp_off[k] : D.pow_off[k - D.nexp];
    VCD H = D.nsym + 24 * k + 9;
    double t0 = H[0] * x[o] + H[3] * x[o + 1] + H[6] * x[o + 2];
    double t1 = H[1] * x[o] + H[4] * x[o + 1] + H[7] * x[o + 2];
    double t2 = H[2] * x[o] + H[5] * x[o + 1] + H[8] * x[o + 2];
    y[o] = t0; y[o + 1] = t1; y[o + 2] = t2;
  }
  CLH_PSD_FOR(B, k, D.npsd) {  // W' W x
    const Team tm = psd_team(B);
    const PsdCone c = psd_cone(B, D, k);
    psd_mul_W(tm, c, false, c.M4, x + c.o, c.R);
    psd_mul_W(tm, c, true, y + c.o, c.M4, c.R);
  }
  CLH_SYNC();
}

// affine_ds! (:127-141)
CLH_D void cones_affine_ds(const Blk& B, const SD& D, VD ds, VCD s) {
  for (int i = B.tid; i < D.m; i += B.nthr) {
    int8_t t = D.etype[i];
    if (t == CT_ZERO) ds[i] = 0;
    else if (t == CT_NONNEG) ds[i] = D.lam[i] * D.lam[i];
  }
  for (int k = B.tid; k < D.nsoc; k += B.nthr) {
    int o = D.soc_off[k], d = D.soc_dim[k];
    VCD l = D.lam + o;
    double x0 = 0;
    for (int i = 0; i < d; ++i) x0 += l[i] * l[i];
    for (int i = 1; i < d; ++i) ds[o + i] = 2 * l[0] * l[i];
    ds[o] = x0;
  }
  for (int k = B.tid; k < D.nexp + D.npow; k += B.nthr) {
    int o = (k < D.nexp) ? D.exp_off[k] : D.pow_off[k - D.nexp];
    for (int i = 0; i < 3; ++i) ds[o + i] = s[o + i];
  }
  CLH_PSD_FOR(B, k, D.npsd) {  // svec(diag(lambda^2)), lane l writing columns l, l + nl, ...
    const Team tm = psd_team(B);
    const PsdCone c = psd_cone(B, D, k);
    for (int j = tm.l; j < c.n; j += tm.nl) {
      for (int i = 0; i < j; ++i) ds[c.o + tri(j) + i] = 0;
      ds[c.o + tri(j) + j] = c.lam[j] * c.lam[j];
    }
  }
  CLH_SYNC();
}

// combined_ds_shift! (:144-160, coneops_symmetric_common.jl:6-35); updates step_z, step_s
// of symmetric cones in place as the reference does
CLH_D void cones_combined_ds_shift(const Blk& B, const SD& D, VD shift, VD sz, VD ss, double sigmu) {
  for (int i = B.tid; i < D.m; i += B.nthr) {
    int8_t t = D.etype[i];
    if (t == CT_ZERO) shift[i] = 0;
    else if (t == CT_NONNEG) {
      sz[i] = sz[i] * D.w[i];
      ss[i] = ss[i] / D.w[i];
      shift[i] = ss[i] * sz[i] + -sigmu;
    }
  }
  for (int k = B.tid; k < D.nsoc; k += B.nthr) {
    int o = D.soc_off[k], d = D.soc_dim[k];
    VCD w = D.w + o;
    double eta = D.soc_eta[k];
    // step_z <- W step_z
    double zeta = 0;
    for (int i = 1; i < d; ++i) zeta += w[i] * sz[o + i];
    double c = sz[o] + zeta / (1 + w[0]);
    double y0 = eta * (w[0] * sz[o] + zeta);
    for (int i = 1; i < d; ++i) sz[o + i] = eta * (sz[o + i] + c * w[i]);
    sz[o] = y0;
    // step_s <- W^{-1} step_s
    zeta = 0;
    for (int i = 1; i < d; ++i) zeta += w[i] * ss[o + i];
    c = -ss[o] + zeta / (1 + w[0]);
    y0 = (1 / eta) * (w[0] * ss[o] - zeta);
    for (int i = 1; i < d; ++i) ss[o + i] = (1 / eta) * (ss[o + i] + c * w[i]);
    ss[o] = y0;
    // shift = step_s o step_z - sigmu e
    double x0 = 0;
    for (int i = 0; i < d; ++i) x0 += ss[o + i] * sz[o + i];
    for (int i = 1; i < d; ++i) shift[o + i] = ss[o] * sz[o + i] + sz[o] * ss[o + i];
    shift[o] = x0 + -sigmu;
  }
  for (int k = B.tid; k < D.nexp + D.npow; k += B.nthr) {
    bool isexp = k < D.nexp;
    int o = isexp ? D.exp_off[k] : D.pow_off[k - D.nexp];
    double a = isexp ? 0.0 : D.pow_alpha[k - D.nexp];
    VCD ns = D.nsym + 24 * k;
    double Hd[9], zc[3], dsl[3], dzl[3], eta[3];
    for (int i = 0; i < 9; ++i) Hd[i] = ns[i];
    for (int i = 0; i < 3; ++i) { zc[i] = ns[21 + i]; dsl[i] = ss[o + i]; dzl[i] = sz[o + i]; }
    higher_correction(isexp, a, Hd, zc, eta, dsl, dzl);
    for (int i = 0; i < 3; ++i) shift[o + i] = ns[18 + i] * sigmu - eta[i];
  }
  CLH_PSD_FOR(B, k, D.npsd) {  // coneops_psdtrianglecone.jl:307-353
    const Team tm = psd_team(B);
    const PsdCone c = psd_cone(B, D, k);
    const int n = c.n;
    for (int i = tm.l; i < c.t; i += tm.nl) c.M4[i] = sz[c.o + i];
    CLH_WSYNC();
    psd_mul_W(tm, c, false, sz + c.o, c.M4, c.R);  // step_z <- W step_z
    for (int i = tm.l; i < c.t; i += tm.nl) c.M4[i] = ss[c.o + i];
    CLH_WSYNC();
    psd_mul_W(tm, c, true, ss + c.o, c.M4, c.Rinv);  // step_s <- W^{-T} step_s
    psd_svec_to_mat(tm, n, ss + c.o, c.M5);
    psd_svec_to_mat(tm, n, sz + c.o, c.M6);
    psd_gemm(tm, n, false, false, 0.5, c.M5, c.M6, 0.0, c.M1);  // (Y Z + Z Y) / 2
    psd_gemm(tm, n, false, false, 0.5, c.M6, c.M5, 1.0, c.M1);
    for (int e = tm.l; e < n * n; e += tm.nl) {
      const int i = e % n, j = e / n;
      if (i > j) c.M1[e] = c.M1[j + i * n];
    }
    CLH_WSYNC();
    psd_mat_to_svec(tm, n, c.M1, shift + c.o);
    for (int j = tm.l; j < n; j += tm.nl) shift[c.o + tri(j + 1) - 1] += -sigmu;
  }
  CLH_SYNC();
}

// Δs_from_Δz_offset! (:163-177)
CLH_D void cones_ds_from_dz_offset(const Blk& B, const SD& D, VD out, VCD ds, VCD z) {
  for (int i = B.tid; i < D.m; i += B.nthr) {
    int8_t t = D.etype[i];
    if (t == CT_ZERO) out[i] = 0;
    else if (t == CT_NONNEG) out[i] = ds[i] / z[i];
  }
  for (int k = B.tid; k < D.nsoc; k += B.nthr) {  // coneops_socone.jl:244-270
    int o = D.soc_off[k], d = D.soc_dim[k];
    VCD w = D.w + o, l = D.lam + o, zz = z + o, dd = ds + o;
    double eta = D.soc_eta[k];
    double resz = soc_res(zz, d);
    double l1ds1 = 0, w1ds1 = 0;
    for (int i = 1; i < d; ++i) { l1ds1 += l[i] * dd[i]; w1ds1 += w[i] * dd[i]; }
    double c = l[0] * dd[0] - l1ds1;
    double f = c / resz;
    out[o] = zz[0] * f + eta * w1ds1;
    for (int i = 1; i < d; ++i) out[o + i] = -zz[i] * f + eta * (dd[i] + w1ds1 / (1 + w[0]) * w[i]);
    for (int i = 0; i < d; ++i) out[o + i] *= (1 / l[0]);
  }
  for (int k = B.tid; k < D.nexp + D.npow; k += B.nthr) {
    int o = (k < D.nexp) ? D.exp_off[k] : D.pow_off[k - D.nexp];
    for (int i = 0; i < 3; ++i) out[o + i] = ds[o + i];
  }
  CLH_PSD_FOR(B, k, D.npsd) {  // coneops_psdtrianglecone.jl:396-421
    const Team tm = psd_team(B);
    const PsdCone c = psd_cone(B, D, k);
    const int n = c.n;
    psd_svec_to_mat(tm, n, ds + c.o, c.M5);
    for (int e = tm.l; e < n * n; e += tm.nl) c.M6[e] = 2 * c.M5[e] / (c.lam[e % n] + c.lam[e / n]);
    CLH_WSYNC();
    psd_mat_to_svec(tm, n, c.M6, c.M4);
    psd_mul_W(tm, c, true, out + c.o, c.M4, c.R);
  }
  CLH_SYNC();
}

// step_length (:180-213) with the family kernels: the symmetric cones' bound
// over this thread's share, from amax
CLH_D double cones_step_symmetric_part(const Blk& B, const SD& D, VCD dz, VCD ds, VCD z, VCD s, double amax) {
  double a = amax;
  for (int i = B.tid; i < D.m; i += B.nthr)
    if (D.etype[i] == CT_NONNEG) {
      if (dz[i] < 0) a = dmin(a, -z[i] / dz[i]);
      if (ds[i] < 0) a = dmin(a, -s[i] / ds[i]);
    }
  for (int k = B.tid; k < D.nsoc; k += B.nthr) {
    int o = D.soc_off[k], d = D.soc_dim[k];
    double az = soc_step(z + o, dz + o, a, d);
    double as = soc_step(s + o, ds + o, a, d);
    a = dmin(a, dmin(az, as));
  }
  CLH_PSD_FOR(B, k, D.npsd) {
    const Team tm = psd_team(B);
    const PsdCone c = psd_cone(B, D, k);
    psd_mul_W(tm, c, false, c.M4, dz + c.o, c.R);
    const double az = psd_step(tm, c, c.M4, a);
    psd_mul_W(tm, c, true, c.M4, ds + c.o, c.Rinv);
    const double as = psd_step(tm, c, c.M4, a);
    a = dmin(a, dmin(az, as));
  }
  return a;
}
// the backtracking line search of the exponential and power cones over this
// thread's share, from a (the symmetric bound capped at max_step_fraction)
CLH_D double cones_step_nonsym_part(const Blk& B, const SD& D, VCD dz, VCD ds, VCD z, VCD s, double a) {
  const DevSettings& st = *D.st;
  double step = st.linesearch_backtrack_step, amin = st.min_terminate_step_length;
  double al = a;
  {
    for (int k = B.tid; k < D.nexp; k += B.nthr) {
      int o = D.exp_off[k];
      double az = backtrack(dz + o, z + o, al, amin, step, [](const double* v) { return exp_dual_feas(v); });
      double as = backtrack(ds + o, s + o, al, amin, step, [](const double* v) { return exp_primal_feas(v); });
      al = dmin(al, dmin(az, as));
    }
    for (int k = B.tid; k < D.npow; k += B.nthr) {
      int o = D.pow_off[k];
      double pa = D.pow_alpha[k];
      double az = backtrack(dz + o, z + o, al, amin, step, [pa](const double* v) { return pow_dual_feas(v, pa); });
      double as = backtrack(ds + o, s + o, al, amin, step, [pa](const double* v) { return pow_primal_feas(v, pa); });
      al = dmin(al, dmin(az, as));
    }
  }
  return al;
}
CLH_D double cones_step_length(const Blk& B, const SD& D, VCD dz, VCD ds, VCD z, VCD s, double amax) {
  double a = blk_min(B, cones_step_symmetric_part(B, D, dz, ds, z, s, amax));
  if (!D.symmetric) a = blk_min(B, cones_step_nonsym_part(B, D, dz, ds, z, s, dmin(a, D.st->max_step_fraction)));
  return a;
}

// compute_barrier (:216-233) plus the central terms of variables_barrier,
// over this thread's share
CLH_D double cones_barrier_part(const Blk& B, const SD& D, VCD z, VCD s, VCD dz, VCD ds, double a) {
  double bar = 0;
  for (int i = B.tid; i < D.m; i += B.nthr)
    if (D.etype[i] == CT_NONNEG) bar -= logsafe((s[i] + a * ds[i]) * (z[i] + a * dz[i]));
  for (int k = B.tid; k < D.nsoc; k += B.nthr) {
    int o = D.soc_off[k], d = D.soc_dim[k];
    double rs = soc_res_shifted(s + o, ds + o, a, d), rz = soc_res_shifted(z + o, dz + o, a, d);
    if (rs > 0 && rz > 0) bar += -logsafe(rs * rz) / 2;
    else bar += dinf();
  }
  for (int k = B.tid; k < D.nexp + D.npow; k += B.nthr) {
    bool isexp = k < D.nexp;
    int o = isexp ? D.exp_off[k] : D.pow_off[k - D.nexp];
    double cz[3], cs[3];
    for (int i = 0; i < 3; ++i) { cz[i] = z[o + i] + a * dz[o + i]; cs[i] = s[o + i] + a * ds[o + i]; }
    bar += isexp ? exp_barrier(cz, cs) : pow_barrier(cz, cs, D.pow_alpha[k - D.nexp]);
  }
  CLH_PSD_FOR(B, k, D.npsd) {  // counted once per team
    const Team tm = psd_team(B);
    const PsdCone c = psd_cone(B, D, k);
    const double lz = psd_logdet(tm, c, z, dz, a);
    const double ls = psd_logdet(tm, c, s, ds, a);
    if (tm.l == 0) {
      bar -= lz;
      bar -= ls;
    }
  }
  return bar;
}
CLH_D double cones_barrier(const Blk& B, const SD& D, VCD z, VCD s, VCD dz, VCD ds, double a) {
  return blk_sum(B, cones_barrier_part(B, D, z, s, dz, ds, a));
}

// margins (:37-53) -> (min margin, sum of positive margins) over this thread's share
CLH_D void cones_margins_part(const Blk& B, const SD& D, VCD z, double& a, double& b) {
  a = dinf();
  b = 0;
  for (int i = B.tid; i < D.m; i += B.nthr) {
    int8_t t = D.etype[i];
    if (t == CT_ZERO) a = dmin(a, kDMax);
    else if (t == CT_NONNEG) { a = dmin(a, z[i]); b += z[i] > 0 ? z[i] : 0; }
  }
  for (int k = B.tid; k < D.nsoc; k += B.nthr) {
    int o = D.soc_off[k], d = D.soc_dim[k];
    double t = 0;
    for (int i = 1; i < d; ++i) t += z[o + i] * z[o + i];
    double al = z[o] - sqrt(t);
    a = dmin(a, al);
    b += dmax(0.0, al);
  }
  CLH_PSD_FOR(B, k, D.npsd) {  // eigenvalues of mat(z), summed once per team
    const Team tm = psd_team(B);
    const PsdCone c = psd_cone(B, D, k);
    psd_svec_to_mat(tm, c.n, z + c.o, c.M5);
    psd_sym_eig(tm, c.n, c.M5, c.M6);
    a = dmin(a, c.n > 0 ? c.M6[0] : kDMax);  // an empty cone: floatmax (:16-18)
    if (tm.l == 0)
      for (int i = 0; i < c.n; ++i)
        if (c.M6[i] > 0) b += c.M6[i];
  }
}
CLH_D void cones_margins(const Blk& B, const SD& D, VCD z, double& amin, double& bsum) {
  double a, b;
  cones_margins_part(B, D, z, a, b);
  amin = blk_min(B, a);
  bsum = blk_sum(B, b);
}
// scaled_unit_shift! (:56-67)
CLH_D void cones_unit_shift(const Blk& B, const SD& D, VD z, double a, bool primal) {
  for (int i = B.tid; i < D.m; i += B.nthr) {
    int8_t t = D.etype[i];
    if (t == CT_ZERO) { if (primal) z[i] = 0; }
    else if (t == CT_NONNEG) z[i] += a;
  }
  for (int k = B.tid; k < D.nsoc; k += B.nthr) z[D.soc_off[k]] += a;
  CLH_PSD_FOR(B, k, D.npsd) {
    const Team tm = psd_team(B);
    for (int j = tm.l; j < D.psd_n[k]; j += tm.nl) z[D.psd_off[k] + tri(j + 1) - 1] += a;
  }
  CLH_SYNC();
}
// unit_initialization! (:70-81)
CLH_D void cones_unit_init(const Blk& B, const SD& D, VD z, VD s) {
  for (int i = B.tid; i < D.m; i += B.nthr) {
    int8_t t = D.etype[i];
    if (t == CT_ZERO) { z[i] = 0; s[i] = 0; }
    else if (t == CT_NONNEG) { z[i] = 1; s[i] = 1; }
  }
  for (int k = B.tid; k < D.nsoc; k += B.nthr) {
    int o = D.soc_off[k], d = D.soc_dim[k];
    for (int i = 0; i < d; ++i) { z[o + i] = 0; s[o + i] = 0; }
    z[o] = 1; s[o] = 1;
  }
  for (int k = B.tid; k < D.nexp; k += B.nthr) {
    int o = D.exp_off[k];
    s[o] = -1.051383945322714; s[o + 1] = 0.556409619469370; s[o + 2] = 1.258967884768947;
    for (int i = 0; i < 3; ++i) z[o + i] = s[o + i];
  }
  for (int k = B.tid; k < D.npow; k += B.nthr) {
    int o = D.pow_off[k];
    double a = D.pow_alpha[k];
    s[o] = sqrt(1 + a); s[o + 1] = sqrt(1 + (1 - a)); s[o + 2] = 0;
    for (int i = 0; i < 3; ++i) z[o + i] = s[o + i];
  }
  for (int k = B.tid; k < D.npsd; k += B.nthr) {  // identity matrices
    const int o = D.psd_off[k], n = D.psd_n[k];
    for (int i = 0; i < tri(n); ++i) { z[o + i] = 0; s[o + i] = 0; }
    for (int j = 0; j < n; ++j) { s[o + tri(j + 1) - 1] = 1; z[o + tri(j + 1) - 1] = 1; }
  }
  CLH_SYNC();
}

// ---------------------------------------------------------------------------
// supernodal LDL' : numeric factorization and solves (FT = float or double)
// ---------------------------------------------------------------------------
// One supernode by one lane group.  WP is the working panel's pointer type:
// this group's share of the LDS area when the panel fits, else the panel in
// memory (inlds false).
template <class FT, bool inlds, class WP>
CLH_D void factor_panel_dense(const Blk& B, const SD& D, const SnDesc& sd, WP W, int& ndyn, bool& bad);

// The left-looking updates of panel row p from every descendant whose rows
// meet this supernode's column range (the descendants' panels are final:
// earlier levels), restricted to the panel columns [c0, c1).  The
// contributions to each entry are subtracted in update order, each summed
// over k in order, whatever the column range: splitting a row's columns
// among threads gives the same values.
// Entry (p, column c) is W[p + c * ld]; optionally only the records
// [elo, ehi) of the row's list, with entry (p, c) at T[toff + c * tstride].
template <class FT, class WP>
CLH_D void apply_row_updates_range(const SD& D, const SnDesc& sd, WP T, size_t toff, int tstride, int p, int c0,
                                   int c1, int elo, int ehi) {
  const auto L = gmem(vcast<FT>(D.L));
  const auto Dv = gmem(vcast<FT>(D.Dv));
  const auto rowupd = gmem(D.rowupd);
  const auto upd_rel = gmem(D.upd_rel);
  const bool whole = c0 == 0 && c1 >= sd.ns;
  const int e0 = elo, e1 = ehi;
  const auto W = T + toff;
  const int ld = tstride;
  RowUpd rn = gload(rowupd + e0);  // one record ahead (the record past the end is readable)
  for (int e = e0; e < e1; ++e) {
    const RowUpd r = rn;
    rn = gload(rowupd + e + 1);
    const auto La = L + r.base;
    const auto Dd = Dv + r.fd;
    const auto rel = upd_rel + r.rel;  // the update's rows in this panel: the columns for bc < ncol
    const int nsd = r.nsd, nrd = r.nrd, a = r.a;
    int bc = 0, bmax = r.bmax;
    if (!whole) {  // rel is increasing: the bc range whose columns fall in [c0, c1)
      int lo = 0, hi = bmax;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (rel[mid] < c0) lo = mid + 1;
        else hi = mid;
      }
      bc = lo;
      hi = bmax;
      int l2 = bc;
      while (l2 < hi) {
        const int mid = (l2 + hi) >> 1;
        if (rel[mid] < c1) l2 = mid + 1;
        else hi = mid;
      }
      bmax = l2;
    }
    // four target columns per pass share the loads of row a and of the
    // pivots; each column still sums over k in order
    for (; bc + 4 <= bmax; bc += 4) {
      FT acc0 = FT(0), acc1 = FT(0), acc2 = FT(0), acc3 = FT(0);
#pragma unroll 4
      for (int k = 0; k < nsd; ++k) {
        const FT ak = La[a + k * nrd], dk = Dd[k];
        const auto Lk = La + bc + k * nrd;
        acc0 += ak * (Lk[0] * dk);
        acc1 += ak * (Lk[1] * dk);
        acc2 += ak * (Lk[2] * dk);
        acc3 += ak * (Lk[3] * dk);
      }
      W[rel[bc] * ld] -= acc0;
      W[rel[bc + 1] * ld] -= acc1;
      W[rel[bc + 2] * ld] -= acc2;
      W[rel[bc + 3] * ld] -= acc3;
    }
    for (; bc < bmax; ++bc) {
      FT acc = FT(0);
#pragma unroll 8
      for (int k = 0; k < nsd; ++k) acc += La[a + k * nrd] * (La[bc + k * nrd] * Dd[k]);
      W[rel[bc] * ld] -= acc;
    }
  }
}
template <class FT, class WP>
CLH_D void apply_row_updates(const SD& D, const SnDesc& sd, WP W, int p, int c0, int c1) {
  const auto rp = gmem(D.rowupd_ptr) + sd.r0;
  apply_row_updates_range<FT>(D, sd, W, (size_t)p, sd.nr, p, c0, c1, rp[p], rp[p + 1]);
}

template <class FT, bool inlds, class WP>
CLH_D void factor_supernode_body(const Blk& B, const SD& D, const SnDesc& sd, WP W, double reg_eps, int& ndyn,
                                 bool& bad) {
  const auto Kx = gmem(D.Kx);
  const auto asm_src = gmem(D.asm_src);
  const auto asm_dst = gmem(D.asm_dst);
  const auto asm_sgn = gmem(D.asm_sgn);
  const int ns = sd.ns, nr = sd.nr;
#define CLH_WSYNC_W()                      \
  do {                                     \
    if constexpr (inlds) CLH_WSYNC_LDS();  \
    else CLH_WSYNC();                      \
  } while (0)
  const int gl = B.gl, G = B.G;
  for (int t = gl; t < nr * ns; t += G) W[t] = FT(0);
  CLH_WSYNC_W();
  for (int t = sd.asm0 + gl; t < sd.asm1; t += G) {
    double v = Kx[asm_src[t]];
    int8_t sg = asm_sgn[t];
    if (sg != 0) v = v + (sg > 0 ? reg_eps : -reg_eps);
    W[asm_dst[t]] = FT(v);
  }
  CLH_WSYNC_W();
  // left-looking updates: each lane owns panel rows and applies the
  // contributions landing on them in update order, so no two lanes write one
  // entry and no fence is needed between updates
  for (int p = gl; p < nr; p += G) apply_row_updates<FT>(D, sd, W, p, 0, ns);
  CLH_WSYNC_W();
#undef CLH_WSYNC_W
  factor_panel_dense<FT, inlds>(B, D, sd, W, ndyn, bad);
}

// The dense part of a supernode's factorization once its panel W holds the
// assembled and updated values: LDL' of the panel (right-looking inside the
// supernode; the pivot goes to the panel's diagonal, unused by the solves,
// and to Dv), the scaling of the columns, the inverse of the unit-lower
// diagonal block, and the panel written back when it was factored in LDS.
template <class FT, bool inlds, class WP>
CLH_D void factor_panel_dense(const Blk& B, const SD& D, const SnDesc& sd, WP W, int& ndyn, bool& bad) {
  const DevSettings& st = *D.st;
  const auto L = gmem(vcast<FT>(D.L));
  const auto Dv = gmem(vcast<FT>(D.Dv));
  const int f = sd.f, ns = sd.ns, nr = sd.nr;
  const int ld = nr;
  const int gl = B.gl, G = B.G;
#define CLH_WSYNC_W()                      \
  do {                                     \
    if constexpr (inlds) CLH_WSYNC_LDS();  \
    else CLH_WSYNC();                      \
  } while (0)
  for (int j = 0; j < ns; ++j) {
    FT dj = W[j + j * ld];
    if (st.dynamic_regularization_enable) {
      const int sg = D.dsign[D.perm[f + j]];
      if (double(dj) * sg <= st.dynamic_regularization_eps) {
        dj = FT(dyn_reg_pivot(st, double(dj), sg));
        ++ndyn;
      }
    }
    if (dj == FT(0) || !finite(double(dj))) bad = true;
    const FT djinv = FT(1) / dj;
    for (int i = j + 1 + gl; i < nr; i += G) {
      const FT lij = W[i + j * ld] * djinv;
      const int cmax = imin(i, ns - 1);
      for (int c = j + 1; c <= cmax; ++c) W[i + c * ld] -= lij * W[c + j * ld];
    }
    CLH_WSYNC_W();
    if (gl == 0) {
      W[j + j * ld] = dj;
      Dv[f + j] = dj;
    }
  }
  CLH_WSYNC_W();
  for (int i = gl; i < nr; i += G) {
    const int jm = imin(i, ns);
    for (int j = 0; j < jm; ++j) W[i + j * ld] = W[i + j * ld] / W[j + j * ld];
  }
  CLH_WSYNC_W();
  // inverse of the unit lower-triangular diagonal block (column-major ns x
  // ns): the triangular solves apply it as one small dense product.  Column
  // j is a forward substitution by lane j; with the panel in LDS it runs in
  // the LDS area behind the panel and is copied out.
  if (ns > 1) {
    const auto Lig = gmem(vcast<FT>(D.Linv)) + sd.iptr;
    auto invert = [&](auto Li) {
      for (int j = gl; j < ns; j += G) {
        for (int i = 0; i < j; ++i) Li[i + j * ns] = FT(0);
        Li[j + j * ns] = FT(1);
        for (int i = j + 1; i < ns; ++i) {
          FT acc = W[i + j * ld];
          for (int k = j + 1; k < i; ++k) acc += W[i + k * ld] * Li[k + j * ns];
          Li[i + j * ns] = -acc;
        }
      }
    };
    if constexpr (inlds) {
      const WP Li = W + nr * ns;
      invert(Li);
      CLH_WSYNC_LDS();
      for (int t = gl; t < ns * ns; t += G) Lig[t] = Li[t];
    } else {
      invert(Lig);
    }
    // the level barrier publishes Li to the solves
  }
  if constexpr (inlds) {  // write the factored panel back (made visible by the level barrier)
    const auto Ls = L + sd.lptr;
    for (int t = gl; t < nr * ns; t += G) Ls[t] = W[t];
  }
#undef CLH_WSYNC_W
}

template <class FT>
CLH_D void factor_supernode(const Blk& B, const SD& D, const SnDesc& sd, double reg_eps, int& ndyn, bool& bad) {
#if !defined(CLH_LANES)
  const int cap = B.lds_bytes / (int)sizeof(FT) / B.lngrp;
  if (B.lds && sd.nr * sd.ns + sd.ns * sd.ns <= cap) {
    CLH_LDS FT* W = (CLH_LDS FT*)B.lds + (size_t)B.lgrp * cap;
    factor_supernode_body<FT, true>(B, D, sd, W, reg_eps, ndyn, bad);
    return;
  }
#endif
  factor_supernode_body<FT, false>(B, D, sd, gmem(vcast<FT>(D.L) + sd.lptr), reg_eps, ndyn, bad);
}

#if defined(CLH_GRID)
// A level with at most this many supernodes is factored one supernode at a
// time by the whole grid.
constexpr int kWideLevel = 4;
// panel columns per work item of a wide supernode's updates
constexpr int kWideCols = 16;
constexpr int kTile = 64;  // output tile of the tiled products: 64 x 64 panel entries per workgroup

// Supernodes with at least this many columns take their updates as tiles
// (on the 80 800-row exp/pow logistic problem the factorization takes 0.47 s
// with 128 or 32, 0.80 s with 512); a serial build may lower it to exercise
// the path on small problems
#if !defined(CLH_TILE_MIN_COLS)
#define CLH_TILE_MIN_COLS 128
#endif
constexpr int kTileMinCols = CLH_TILE_MIN_COLS;
constexpr int kTileK = 32;       // descendant columns staged per pass
constexpr int kTileBatch = 128;  // descendants whose rows in the tile are located at once
constexpr size_t kTiledLds(size_t fts) { return fts * 2 * kTileK * kTile + sizeof(int) * (5 * kTileBatch + 4); }
#if defined(CLH_SERIAL)
constexpr int kTileAcc = 256;  // one thread holds all 256 4 x 4 blocks of a tile
#else
constexpr int kTileAcc = 1;  // 256 threads, one 4 x 4 block each
#endif

// The updates of a wide supernode as dense tiles (a GEMM over the
// descendants): workgroup w computes 64 x 64 output tiles of the panel's
// lower part.  For each tile the update list is walked in batches of 128
// descendants: one thread per descendant locates the descendant's rows that
// land in the tile's rows and columns (four interleaved binary searches in
// its sorted row positions), a scan packs the columns of the descendants that
// touch the tile, and the packed columns are gathered into LDS 32 at a time
// (zeros elsewhere, the row side scaled by the pivots, 8 threads per column);
// each thread accumulates a 4 x 4 block of the tile in registers, and the
// tile is subtracted from the panel once, after every descendant.
template <class FT, class WP>
CLH_D void tiled_wide_updates(const Blk& B, const SD& D, const SnDesc& sd, WP W) {
  const auto L = gmem(vcast<FT>(D.L));
  const auto Dv = gmem(vcast<FT>(D.Dv));
  const auto upd_rel = gmem(D.upd_rel);
  const int ns = sd.ns, nr = sd.nr;
  const int lt = B.ltid >= 0 ? B.ltid : B.tid, ln = B.lnthr >= 0 ? B.lnthr : B.nthr;
  CLH_LDS FT* As = (CLH_LDS FT*)B.lds;      // [kTileK][kTile]: packed column k of tile row r at k * kTile + r
  CLH_LDS FT* Bs = As + kTileK * kTile;     // the same for the tile's columns
  CLH_LDS int* bnd = (CLH_LDS int*)(Bs + kTileK * kTile);  // [4][kTileBatch]: rows a0..a1, columns b0..b1
  CLH_LDS int* cst = bnd + 4 * kTileBatch;  // [kTileBatch + 1]: first packed column of each descendant
  const int ntr = (nr + kTile - 1) / kTile, ntc = (ns + kTile - 1) / kTile;
  // tiles (tr, tc) with tr >= tc: the panel's lower part
  int ntiles = 0;
  for (int tc = 0; tc < ntc; ++tc) ntiles += ntr - tc;
  // A panel with few tiles and a long update list (a narrow root over
  // thousands of descendants) leaves most workgroups idle: when at most a
  // quarter of them would have a tile, its list is split into K chunks of
  // batches, tile and chunk one item, each chunk's sums stored apart and
  // added to the panel in chunk order afterwards (the 200-column root of
  // the 5000 x 200 logistic regression, 10 tiles over 12 000 descendants:
  // factorization 0.38 s to 0.09 s over the solve).  Panels with more tiles
  // keep one item per tile and its sums in list order.
  const int nbatch = (sd.upd1 - sd.upd0 + kTileBatch - 1) / kTileBatch;
  int K = 1;
  if (ntiles * 4 <= B.nwg && nbatch > 1) {
    const size_t cap = (size_t)D.ngroups * D.maxns;  // the scratch, in entries
    K = imax(1, imin(imin(B.nwg / ntiles, nbatch), (int)(cap / ((size_t)ntiles * kTile * kTile))));
  }
  const int bpk = (nbatch + K - 1) / K;  // batches per chunk
  const auto Pt = gmem(vcast<FT>(D.gscr));
  for (int item = B.wg; item < ntiles * K; item += B.nwg) {
    const int tile = item % ntiles, kc = item / ntiles;
    const int u0 = sd.upd0 + kc * bpk * kTileBatch, u1 = imin(sd.upd1, u0 + bpk * kTileBatch);
    int tc = 0, rem = tile;
    while (rem >= ntr - tc) { rem -= ntr - tc; ++tc; }
    const int tr = tc + rem;
    const int p0 = tr * kTile, c0 = tc * kTile;
    FT acc[kTileAcc][4][4];
    for (int q = 0; q < kTileAcc; ++q)
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[q][i][j] = FT(0);
    for (int ub = u0; ub < u1; ub += kTileBatch) {
      const int nb = imin(kTileBatch, u1 - ub);
      CLH_BSYNC();
      for (int d = lt; d < kTileBatch; d += ln) {
        int lo[4] = {0, 0, 0, 0}, hi[4] = {0, 0, 0, 0}, cnt = 0;
        if (d < nb) {
          const UpdDesc r = D.upd[ub + d];
          const auto rel = upd_rel + r.rel;
          const int nrem = r.nrd - r.roff;
          const int v[4] = {p0, p0 + kTile, c0, c0 + kTile};
          hi[0] = hi[1] = nrem;
          hi[2] = hi[3] = r.ncol;
          // first position with rel >= v[s], the four searches in step
          for (bool any = true; any;) {
            any = false;
            for (int sI = 0; sI < 4; ++sI)
              if (lo[sI] < hi[sI]) {
                any = true;
                const int m = (lo[sI] + hi[sI]) >> 1;
                if (rel[m] < v[sI]) lo[sI] = m + 1;
                else hi[sI] = m;
              }
          }
          if (lo[0] < lo[1] && lo[2] < lo[3]) cnt = r.nsd;
        }
        for (int sI = 0; sI < 4; ++sI) bnd[sI * kTileBatch + d] = lo[sI];
        cst[d + 1] = cnt;
      }
      CLH_BSYNC();
#if defined(CLH_SERIAL)
      cst[0] = 0;
      for (int d = 0; d < kTileBatch; ++d) cst[d + 1] += cst[d];
#else
      if (lt < 64) {  // exclusive scan of the column counts by one wavefront, two per lane
        const int e0 = cst[2 * lt + 1], e1 = cst[2 * lt + 2];
        int v = e0 + e1;
        for (int o = 1; o < 64; o <<= 1) {
          const int y = __shfl_up(v, o, 64);
          if (lt >= o) v += y;
        }
        cst[2 * lt + 2] = v;
        cst[2 * lt + 1] = v - e1;
        if (lt == 0) cst[0] = 0;
      }
#endif
      CLH_BSYNC();
      const int total = cst[nb];
      for (int q0 = 0; q0 < total; q0 += kTileK) {
        const int kn = imin(kTileK, total - q0);
        CLH_BSYNC();
        for (int t = lt; t < kn * kTile; t += ln) { As[t] = FT(0); Bs[t] = FT(0); }
        CLH_BSYNC();
        for (int t = lt; t < kn * 8; t += ln) {
          const int k = t >> 3, g = t & 7, q = q0 + k;
          int dlo = 0, dhi = nb;  // the descendant with cst[d] <= q < cst[d + 1]
          while (dhi - dlo > 1) {
            const int m = (dlo + dhi) >> 1;
            if (cst[m] <= q) dlo = m;
            else dhi = m;
          }
          const int col = q - cst[dlo];
          const UpdDesc r = D.upd[ub + dlo];
          const auto La = L + r.lptr + r.roff + (size_t)col * r.nrd;
          const auto rel = upd_rel + r.rel;
          const FT dk = Dv[r.fd + col];
          const int a1 = bnd[kTileBatch + dlo], b1 = bnd[3 * kTileBatch + dlo];
          for (int a = bnd[dlo] + g; a < a1; a += 8) As[k * kTile + (rel[a] - p0)] = La[a] * dk;
          for (int b = bnd[2 * kTileBatch + dlo] + g; b < b1; b += 8) Bs[k * kTile + (rel[b] - c0)] = La[b];
        }
        CLH_BSYNC();
        for (int q = lt, qi = 0; q < 256; q += ln, ++qi) {
          const int tx = q & 15, ty = q >> 4;  // rows 4 ty.., columns 4 tx.. of the tile
          for (int k = 0; k < kn; ++k) {
            FT av[4], bv[4];
            for (int i = 0; i < 4; ++i) av[i] = As[k * kTile + ty * 4 + i];
            for (int j = 0; j < 4; ++j) bv[j] = Bs[k * kTile + tx * 4 + j];
            for (int i = 0; i < 4; ++i)
              for (int j = 0; j < 4; ++j) acc[qi][i][j] += av[i] * bv[j];
          }
        }
      }
    }
    for (int q = lt, qi = 0; q < 256; q += ln, ++qi) {
      const int tx = q & 15, ty = q >> 4;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
          const int p = p0 + ty * 4 + i, c = c0 + tx * 4 + j;
          if (K == 1) {
            if (p < nr && c < ns && p >= c) W[p + (size_t)c * nr] -= acc[qi][i][j];
          } else {
            Pt[((size_t)kc * ntiles + tile) * (kTile * kTile) + (ty * 4 + i) + (size_t)(tx * 4 + j) * kTile] = acc[qi][i][j];
          }
        }
    }
  }
  if (K > 1) {
    CLH_SYNC();
    for (size_t t = B.tid; t < (size_t)ntiles * kTile * kTile; t += B.nthr) {
      const int tile = (int)(t / (kTile * kTile)), e = (int)(t % (kTile * kTile));
      int tc = 0, rem = tile;
      while (rem >= ntr - tc) { rem -= ntr - tc; ++tc; }
      const int p = (tc + rem) * kTile + e % kTile, c = tc * kTile + e / kTile;
      if (p < nr && c < ns && p >= c) {
        FT sum = FT(0);
        for (int k = 0; k < K; ++k) sum += Pt[((size_t)k * ntiles + tile) * (kTile * kTile) + e];
        W[p + (size_t)c * nr] -= sum;
      }
    }
  }
  CLH_SYNC();
}

// panels with at least this many columns are factored block by block over
// the whole grid (a serial build may lower it to exercise the path on small
// problems)
#if !defined(CLH_DENSE_GRID_MIN_COLS)
#define CLH_DENSE_GRID_MIN_COLS 128
#endif
constexpr int kDenseGridMinCols = CLH_DENSE_GRID_MIN_COLS;
constexpr size_t kDenseGridMinEntries = 65536;
// columns per block of the blocked dense factorization, and per super-block:
// inside a super-block each block updates only the super-block's later
// columns (and X's rows up to its end); the rest of the trailing part takes
// one rank-kDenseSB update when the super-block is done, so the trailing
// matrix is read and written once per super-block instead of once per block
// (a serial build may lower both to exercise the path on small problems).
// Super-blocks of 512 rather than 256 columns: on the random sparse SOCP
// with n = 20 000 (an 18 754-column root) the trailing updates stream half
// as much, the solves take half as many super-block steps, and the solve
// went from 8.4 s to 8.1 s (FP64) and 10.3 s to 9.5 s (FP32 factors).
// 512 is also the widest that fits: the dense path needs
// (2 kDenseNB 64 + kDenseSB) FP64 words of the workgroup's 36 KB of LDS,
// and without them wide panels fall back to one lane group
#if !defined(CLH_DENSE_NB)
#define CLH_DENSE_NB 32
#endif
#if !defined(CLH_DENSE_SB)
#define CLH_DENSE_SB 512
#endif
constexpr int kDenseNB = CLH_DENSE_NB;
constexpr int kDenseSB = CLH_DENSE_SB;
constexpr size_t kDenseGridLds(size_t fts) {
  return fts * (2 * kDenseNB * kTile + kDenseSB > 4 * kDenseNB * (kDenseNB + 1) ? 2 * kDenseNB * kTile + kDenseSB
                                                                                : 4 * kDenseNB * (kDenseNB + 1));
}

// L^{-1} of a unit lower-triangular block of side n <= 32 held in LDS
// (column-major, leading dimension ldl), into Li (leading dimension ldi).
// Column c is forward substitution, x_c = 1 and
//   x_r = -(L[r,c] + sum_{c<k<r} L[r,k] x_k),
// its sums run right-looking: at step k every entry (r, c) with r > k > c
// adds L[r,k] x_k (x_k final by then), so each entry takes the operations of
// the column-by-column loop in the same order while all threads work at once
// (the column-by-column loop kept one thread per column busy for n^2/2
// dependent steps).  Li holds the running sums until the last pass negates
// them.  Callers synchronize before reading Li.
template <class FT>
CLH_D void unit_lower_inverse_lds(int lt, int ln, int n, const CLH_LDS FT* Lb, int ldl, CLH_LDS FT* Li, int ldi) {
  for (int t = lt; t < n * n; t += ln) {
    const int r = t % n, c = t / n;
    Li[r + c * ldi] = r > c ? Lb[r + c * ldl] : FT(r == c);
  }
  for (int k = 1; k < n - 1; ++k) {
    CLH_BSYNC();
    const int nr = n - 1 - k;  // entries (r, c): k < r < n, c < k
    for (int t = lt; t < nr * k; t += ln) {
      const int r = k + 1 + t % nr, c = t / nr;
      Li[r + c * ldi] += Lb[r + k * ldl] * (-Li[k + c * ldi]);
    }
  }
  CLH_BSYNC();
  for (int t = lt; t < n * n; t += ln) {
    const int r = t % n, c = t / n;
    if (r > c) Li[r + c * ldi] = -Li[r + c * ldi];
  }
}

#if !defined(CLH_SERIAL)
// 16 x 16 x 4 matrix products on the matrix cores: for lane l, A holds
// A[l % 16][l / 16] and B holds B[l / 16][l % 16]; output v of lane l is
// D[drow(l, v)][l % 16], where the FP64 product interleaves the rows
// (l / 16 + 4 v) and the FP32 product keeps 4 consecutive rows per lane
// (4 (l / 16) + v) — both read off the hardware with known inputs
template <class FT>
struct Mfma16;
template <>
struct Mfma16<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static CLH_D v4 op(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static CLH_D int drow(int l, int v) { return (l >> 4) + 4 * v; }
};
template <>
struct Mfma16<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static CLH_D v4 op(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static CLH_D int drow(int l, int v) { return 4 * (l >> 4) + v; }
};

// One 64 x 64 tile of dense_grid_update on the matrix cores (256 threads):
// wave w takes the 32 x 32 quarter (rows 32 (w & 1).., columns 32 (w >> 1)..)
// as 2 x 2 products of 16 x 16, over the pivots kDenseNB at a time staged
// in LDS; each thread fetches its share of the next columns into registers
// while the current ones are multiplied.  The products run transposed
// (columns on the A side, rows on the B side), so a lane's outputs lie in
// one row and 16 lanes write 16 consecutive rows of a column.
template <class FT, class WP, class XP>
CLH_D void dense_tile_mfma(int lt, WP W, XP Lig, int ld, int nr, int ns, int k0, int k1, int cw1, int rx1, int cx1,
                           bool xt, int p0, int c0, CLH_LDS FT* As, CLH_LDS FT* Bs, const CLH_LDS FT* piv) {
  constexpr int NB = kDenseNB;
  constexpr int PT = NB * kTile / 256;  // staged entries of each side per thread
  static_assert(NB % 4 == 0 && PT * 256 == NB * kTile, "a block of columns is staged by 256 threads");
  typedef typename Mfma16<FT>::v4 v4;
  const int l = lt & 63, w = lt >> 6, wr = w & 1, wc = w >> 1;
  FT ra[PT], rb[PT];
  // thread lt stages entries t = lt + 256 u: As[k][r] (r = t % 64, k = t / 64);
  // Bs[k][r] likewise for the panel's tiles, Bs[kk][c] (kk = t % NB, c = t / NB)
  // for X's tiles, so consecutive threads read consecutive rows of X
  auto fetch = [&](int kc) {
    const int kb = imin(NB, k1 - kc);
    for (int u = 0; u < PT; ++u) {
      const int t = lt + 256 * u, r = t % kTile, k = t / kTile;
      const int ip = p0 + r;
      ra[u] = k < kb && ip < (xt ? rx1 : nr) ? W[ip + (size_t)(kc + k) * ld] : FT(0);
      if (!xt) {
        const int ic = c0 + r;
        rb[u] = k < kb && ic < cw1 ? W[ic + (size_t)(kc + k) * ld] : FT(0);
      } else {
        const int kk = t % NB, c = t / NB;
        rb[u] = kk < kb && c0 + c < cx1 ? Lig[kc + kk + (size_t)(c0 + c) * ns] : FT(0);
      }
    }
  };
  v4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = v4{FT(0), FT(0), FT(0), FT(0)};
  fetch(k0);
  for (int kc = k0; kc < k1; kc += NB) {
    CLH_BSYNC();
    for (int u = 0; u < PT; ++u) {
      const int t = lt + 256 * u, r = t % kTile, k = t / kTile;
      const FT pk = kc - k0 + k < k1 - k0 ? piv[kc - k0 + k] : FT(0);
      if (!xt) {
        As[k * kTile + r] = ra[u];
        Bs[k * kTile + r] = rb[u] * pk;
      } else {
        As[k * kTile + r] = ra[u] * pk;
        Bs[(t % NB) * kTile + t / NB] = rb[u];
      }
    }
    CLH_BSYNC();
    if (kc + NB < k1) fetch(kc + NB);
    for (int k4 = 0; k4 < NB; k4 += 4) {
      const int kk = (k4 + (l >> 4)) * kTile + (l & 15);
      const FT b0 = As[kk + 32 * wr], b1 = As[kk + 32 * wr + 16];
      const FT a0 = Bs[kk + 32 * wc], a1 = Bs[kk + 32 * wc + 16];
      acc[0][0] = Mfma16<FT>::op(a0, b0, acc[0][0]);
      acc[0][1] = Mfma16<FT>::op(a1, b0, acc[0][1]);
      acc[1][0] = Mfma16<FT>::op(a0, b1, acc[1][0]);
      acc[1][1] = Mfma16<FT>::op(a1, b1, acc[1][1]);
    }
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int v = 0; v < 4; ++v) {
        const int p = p0 + 32 * wr + 16 * i + (l & 15), c = c0 + 32 * wc + 16 * j + Mfma16<FT>::drow(l, v);
        if (!xt) {
          if (p < nr && c < cw1 && p >= c) W[p + (size_t)c * ld] -= acc[i][j][v];
        } else if (p < rx1 && c < cx1) {
          Lig[p + (size_t)c * ns] -= acc[i][j][v];
        }
      }
}
#endif

// C -= A B over 64 x 64 tiles of a dense panel W (rows nr, leading dimension
// ld, ns columns, L D below the diagonal in the factored columns) and of X
// (ns x ns, column-major: the inverses of the unit-lower diagonal blocks of
// the panel's super-blocks), for the pivots k0 <= k < k1, whose columns are
// staged in LDS kDenseNB at a time:
//   W[p, c] -= sum_k W[p, k] W[c, k] / d_k   for p >= c, r0 <= c < cw1, p < nr
//   X[p, c] -= sum_k (W[p, k] / d_k) X[k, c] for r0 <= p < rx1, cx0 <= c < cx1
// (d_k = W[k, k]).  Each thread holds a 4 x 4 block of the tile over every k
// and subtracts it once.
template <class FT, class WP, class XP>
CLH_D void dense_grid_update(const Blk& B, WP W, XP Lig, int ld, int nr, int ns, int k0, int k1, int r0, int cw1,
                             int rx1, int cx0, int cx1, CLH_LDS FT* S0) {
  constexpr int NB = kDenseNB;
  const int lt = B.ltid >= 0 ? B.ltid : B.tid, ln = B.lnthr >= 0 ? B.lnthr : B.nthr;
  CLH_LDS FT* As = S0;                // [NB][kTile]
  CLH_LDS FT* Bs = As + NB * kTile;   // [NB][kTile]
  CLH_LDS FT* piv = Bs + NB * kTile;  // 1 / d_k, k0 <= k < k1
  const int ntr = imax(0, (nr - r0 + kTile - 1) / kTile), ntc = imax(0, (cw1 - r0 + kTile - 1) / kTile);
  const int nxr = imax(0, (rx1 - r0 + kTile - 1) / kTile), nxc = imax(0, (cx1 - cx0 + kTile - 1) / kTile);
  // The rank-kDenseSB updates walk the tiles in groups of 8 x 8: workgroup w
  // takes slot (w % 8) (nwg / 8) + w / 8 of each round, so the 64 workgroups
  // that one XCD holds (workgroups are dealt to the 8 XCDs in turn) take one
  // group and read its 8 row panels and 8 column panels through their
  // shared L2 instead of each fetching its own from HBM.  Slots outside the
  // region (the groups along the diagonal and the edges) are skipped.
  // (a single workgroup, as in the serial build, walks the same groups)
  const int g = k1 - k0 > NB && (B.nwg % 8 == 0 || B.nwg == 1) ? 8 : 1;
  const int NTR = (ntr + g - 1) / g, NTC = (ntc + g - 1) / g, NXR = (nxr + g - 1) / g, NXC = (nxc + g - 1) / g;
  int nsw = 0;  // groups (TR, TC) with TR >= TC hold the panel's lower tiles
  for (int TC = 0; TC < NTC; ++TC) nsw += NTR - TC;
  const int nslots = (nsw + NXR * NXC) * g * g;
  const int wslot = g > 1 && B.nwg % 8 == 0 ? (B.wg % 8) * (B.nwg / 8) + B.wg / 8 : B.wg;
  CLH_BSYNC();
  for (int k = lt; k < k1 - k0; k += ln) piv[k] = FT(1) / W[k0 + k + (size_t)(k0 + k) * ld];
  for (int q = wslot; q < nslots; q += B.nwg) {
    const int grp = q / (g * g), di = q % g, dj = (q / g) % g;
    const bool xt = grp >= nsw;  // the same on every thread of the workgroup
    int p0, c0;
    if (!xt) {
      int TC = 0, rem = grp;
      while (rem >= NTR - TC) { rem -= NTR - TC; ++TC; }
      const int tr = (TC + rem) * g + di, tc = TC * g + dj;
      if (tr >= ntr || tc >= ntc || tr < tc) continue;
      p0 = r0 + tr * kTile;
      c0 = r0 + tc * kTile;
    } else {
      const int tr = ((grp - nsw) % NXR) * g + di, tc = ((grp - nsw) / NXR) * g + dj;
      if (tr >= nxr || tc >= nxc) continue;
      p0 = r0 + tr * kTile;
      c0 = cx0 + tc * kTile;
    }
#if defined(CLH_SERIAL)
    FT acc[kTileAcc][4][4];
    for (int q = 0; q < kTileAcc; ++q)
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[q][i][j] = FT(0);
    for (int kc = k0; kc < k1; kc += NB) {
      const int kb = imin(NB, k1 - kc);
      CLH_BSYNC();
      if (!xt) {
        for (int t = lt; t < kb * kTile; t += ln) {
          const int r = t % kTile, k = t / kTile;
          const int ip = p0 + r, ic = c0 + r;
          As[k * kTile + r] = ip < nr ? W[ip + (size_t)(kc + k) * ld] : FT(0);
          Bs[k * kTile + r] = ic < cw1 ? W[ic + (size_t)(kc + k) * ld] * piv[kc - k0 + k] : FT(0);
        }
      } else {
        for (int t = lt; t < kb * kTile; t += ln) {
          const int r = t % kTile, k = t / kTile;
          const int ip = p0 + r;
          As[k * kTile + r] = ip < rx1 ? W[ip + (size_t)(kc + k) * ld] * piv[kc - k0 + k] : FT(0);
          const int kk = t % kb, c = t / kb;  // X's column c, its rows kc + kk in order
          Bs[kk * kTile + c] = c0 + c < cx1 ? Lig[kc + kk + (size_t)(c0 + c) * ns] : FT(0);
        }
      }
      CLH_BSYNC();
      for (int q = lt, qi = 0; q < 256; q += ln, ++qi) {
        const int tx = q & 15, ty = q >> 4;  // rows 4 ty.., columns 4 tx.. of the tile
        for (int k = 0; k < kb; ++k) {
          FT av[4], bv[4];
          for (int i = 0; i < 4; ++i) av[i] = As[k * kTile + ty * 4 + i];
          for (int j = 0; j < 4; ++j) bv[j] = Bs[k * kTile + tx * 4 + j];
          for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) acc[qi][i][j] += av[i] * bv[j];
        }
      }
    }
    for (int q = lt, qi = 0; q < 256; q += ln, ++qi) {
      const int tx = q & 15, ty = q >> 4;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
          const int p = p0 + ty * 4 + i, c = c0 + tx * 4 + j;
          if (!xt) {
            if (p < nr && c < cw1 && p >= c) W[p + (size_t)c * ld] -= acc[qi][i][j];
          } else if (p < rx1 && c < cx1) {
            Lig[p + (size_t)c * ns] -= acc[qi][i][j];
          }
        }
    }
#else
    dense_tile_mfma<FT>(lt, W, Lig, ld, nr, ns, k0, k1, cw1, rx1, cx1, xt, p0, c0, As, Bs, piv);
#endif
  }
}

// The dense part of a wide supernode's factorization over the whole grid,
// right-looking by blocks of 32 columns.  Per block: workgroup 0 eliminates
// the block's diagonal part in LDS (pivots with the dynamic regularization
// rule); the rows below it are multiplied by the transposed inverse of the
// block's unit-lower part (LD = A L^-T, 64-row tiles staged in LDS); the
// later columns of the block's super-block (kDenseSB columns) are updated as
// 64 x 64 tiles (the block's columns staged in LDS, a 4 x 4 block per
// thread), and the trailing part beyond the super-block takes one
// rank-kDenseSB update of the same kind once the super-block is factored
// (right-looking by super-blocks, each super-block right-looking by blocks,
// so the trailing matrix is streamed once per super-block).  The panel holds L D below the
// diagonal until the columns are scaled at the end.  The inverses X of the
// unit-lower diagonal blocks of the super-blocks (used by the solves, which
// substitute block by block across super-blocks, dense_solve_forward/
// _backward) are built along, right-looking as well: with the block's rows
// of X holding -sum L_bk X_k over the super-block's earlier blocks, they are
// solved with L_bb by forward substitution (and L_bb^-1 goes on the
// diagonal) beside the rows-below tiles, and the super-block's later rows of
// X receive -L_ib X_b in tiles beside its own columns' update.  (An inverse
// of the whole diagonal block doubled the factorization's work and, applied
// in the solves, took more refinement steps than substitution.)
template <class FT, class WP>
CLH_D void factor_panel_dense_grid(const Blk& B, const SD& D, const SnDesc& sd, WP W, int& ndyn, bool& bad) {
  constexpr int NB = kDenseNB;
  const DevSettings& st = *D.st;
  const auto Dv = gmem(vcast<FT>(D.Dv));
  const int f = sd.f, ns = sd.ns, nr = sd.nr, ld = nr;
  const int lt = B.ltid >= 0 ? B.ltid : B.tid, ln = B.lnthr >= 0 ? B.lnthr : B.nthr;
  CLH_LDS FT* const S0 = (CLH_LDS FT*)B.lds;
  const auto Lig = gmem(vcast<FT>(D.Linv)) + sd.iptr;  // X, [r + c * ns]
  {  // zero the super-blocks' diagonal blocks of X (ordered by the first block's barrier)
    const int nsb = (ns + kDenseSB - 1) / kDenseSB;
    for (size_t t = B.tid; t < (size_t)nsb * kDenseSB * kDenseSB; t += B.nthr) {
      const int q = (int)(t / ((size_t)kDenseSB * kDenseSB)), e = (int)(t % ((size_t)kDenseSB * kDenseSB));
      const int r = q * kDenseSB + e % kDenseSB, c = q * kDenseSB + e / kDenseSB;
      if (r < ns && c < ns) Lig[r + (size_t)c * ns] = FT(0);
    }
  }
  for (int sb = 0; sb < ns; sb += kDenseSB) {
  const int se = imin(ns, sb + kDenseSB);
  for (int jb = sb; jb < se; jb += NB) {
    const int je = imin(se, jb + NB), kb = je - jb;
    // items with L_bb^-1: the rows below the block (LD = A L_bb^-T, tiles of
    // 64 rows), the block's rows of X (L_bb^-1 X_b, tiles of 64 columns
    // left of the block within its super-block) and X's diagonal block.
    // Every workgroup with an
    // item eliminates the block's diagonal part itself, in LDS (the same
    // operations on the same values in each, so the same pivots); workgroup 0
    // writes it back.
    const int nrt = (nr - je + kTile - 1) / kTile, nxt = (jb - sb + kTile - 1) / kTile, nit = nrt + nxt + 1;
    if (B.wg < nit) {
      CLH_LDS FT* Lb = S0;                // the diagonal part, then L_bb, [r + c * NB]
      CLH_LDS FT* Li = S0 + NB * NB;      // its inverse
      CLH_LDS FT* Wt = S0 + 2 * NB * NB;  // row tile [r + j * kTile], or X tile [k + c * NB]
      CLH_BSYNC();
      for (int t = lt; t < kb * kb; t += ln) {
        const int r = t % kb, c = t / kb;
        Lb[r + c * NB] = r >= c ? W[jb + r + (size_t)(jb + c) * ld] : FT(0);
      }
      for (int j = 0; j < kb; ++j) {
        CLH_BSYNC();
        FT dj = Lb[j + j * NB];
        if (st.dynamic_regularization_enable) {
          const int sg = D.dsign[D.perm[f + jb + j]];
          if (double(dj) * sg <= st.dynamic_regularization_eps) {
            dj = FT(dyn_reg_pivot(st, double(dj), sg));
            if (lt == 0 && B.wg == 0) ++ndyn;
          }
        }
        if (dj == FT(0) || !finite(double(dj))) bad = true;
        const FT djinv = FT(1) / dj;
        CLH_BSYNC();
        // rank-1 update of the rows below j, one entry (i, c), j < c <= i,
        // per thread: each entry takes the same operation as in a row loop
        const int m = kb - 1 - j;
        for (int t = lt; t < m * m; t += ln) {
          const int i = j + 1 + t % m, c = j + 1 + t / m;
          if (c <= i) {
            const FT lij = Lb[i + j * NB] * djinv;
            Lb[i + c * NB] -= lij * Lb[c + j * NB];
          }
        }
        if (lt == 0) {
          Lb[j + j * NB] = dj;
          if (B.wg == 0) Dv[f + jb + j] = dj;
        }
      }
      CLH_BSYNC();
      if (B.wg == 0)
        for (int t = lt; t < kb * kb; t += ln) {
          const int r = t % kb, c = t / kb;
          if (r >= c) W[jb + r + (size_t)(jb + c) * ld] = Lb[r + c * NB];
        }
      for (int t = lt; t < kb * kb; t += ln) {  // L = (L D) D^-1 below the diagonal
        const int r = t % kb, c = t / kb;
        if (r > c) Li[r + c * NB] = Lb[r + c * NB] / Lb[c + c * NB];
      }
      CLH_BSYNC();
      for (int t = lt; t < kb * kb; t += ln) {
        const int r = t % kb, c = t / kb;
        Lb[r + c * NB] = r > c ? Li[r + c * NB] : FT(r == c);
      }
      CLH_BSYNC();
      unit_lower_inverse_lds<FT>(lt, ln, kb, Lb, NB, Li, NB);
      for (int it = B.wg; it < nit; it += B.nwg) {
        CLH_BSYNC();
        if (it < nrt) {
          const int p0 = je + it * kTile;
          for (int t = lt; t < kTile * kb; t += ln) {
            const int r = t % kTile, j = t / kTile;
            Wt[r + j * kTile] = p0 + r < nr ? W[p0 + r + (size_t)(jb + j) * ld] : FT(0);
          }
          CLH_BSYNC();
          for (int t = lt; t < kTile * kb; t += ln) {
            const int r = t % kTile, c = t / kTile;
            if (p0 + r < nr) {
              FT a = FT(0);
              for (int j = 0; j <= c; ++j) a += Wt[r + j * kTile] * Li[c + j * NB];
              W[p0 + r + (size_t)(jb + c) * ld] = a;
            }
          }
        } else if (it < nrt + nxt) {
          const int c0 = sb + (it - nrt) * kTile, cn = imin(kTile, jb - c0);
          for (int t = lt; t < kb * cn; t += ln) {
            const int k = t % kb, c = t / kb;
            Wt[k + c * NB] = Lig[jb + k + (size_t)(c0 + c) * ns];
          }
          // forward substitution with L_bb, right-looking: at step k every
          // entry (r, c), r > k, subtracts L[r,k] x_k (final by then), the
          // operations of a column-by-column substitution in its order
          for (int k = 0; k < kb - 1; ++k) {
            CLH_BSYNC();
            const int nrk = kb - 1 - k;
            for (int t = lt; t < nrk * cn; t += ln) {
              const int r = k + 1 + t % nrk, c = t / nrk;
              Wt[r + c * NB] -= Lb[r + k * NB] * Wt[k + c * NB];
            }
          }
          CLH_BSYNC();
          for (int t = lt; t < kb * cn; t += ln) {
            const int r = t % kb, c = t / kb;
            Lig[jb + r + (size_t)(c0 + c) * ns] = Wt[r + c * NB];
          }
        } else {
          for (int t = lt; t < kb * kb; t += ln) {
            const int r = t % kb, c = t / kb;
            if (r >= c) Lig[jb + r + (size_t)(jb + c) * ns] = Li[r + c * NB];
          }
        }
      }
    }
    CLH_SYNC();
    CLH_PROBE(B, 106);
    // the super-block's later columns (every row below them) and X's rows
    // up to the super-block's end receive the block's update now
    if (je < se) dense_grid_update<FT>(B, W, Lig, ld, nr, ns, jb, je, je, se, se, sb, je, S0);
    CLH_SYNC();
    CLH_PROBE(B, 107);
  }
  // the rest of the trailing part and X's later rows: the super-block's
  // columns as one rank-(se - sb) update
  if (se < ns) dense_grid_update<FT>(B, W, Lig, ld, nr, ns, sb, se, se, ns, se, 0, 0, S0);
  CLH_SYNC();
  CLH_PROBE(B, 108);
  }
  CLH_PROBE(B, 103);
  // scale the columns: L = (L D) D^-1
  for (size_t t = B.tid; t < (size_t)nr * ns; t += B.nthr) {
    const int i = (int)(t % nr), j = (int)(t / nr);
    if (i > j) W[t] = W[t] / W[j + (size_t)j * ld];
  }
  // X_sb' into the strict upper triangle of each super-block's diagonal
  // block of X (unused there): the forward solve reads the rows of X_sb as
  // contiguous columns
  {
    const int nsb = (ns + kDenseSB - 1) / kDenseSB;
    for (size_t t = B.tid; t < (size_t)nsb * kDenseSB * kDenseSB; t += B.nthr) {
      const int q = (int)(t / ((size_t)kDenseSB * kDenseSB)), e = (int)(t % ((size_t)kDenseSB * kDenseSB));
      const int r = q * kDenseSB + e % kDenseSB, c = q * kDenseSB + e / kDenseSB;
      if (r < c && c < ns) Lig[r + (size_t)c * ns] = Lig[c + (size_t)r * ns];
    }
  }
  CLH_SYNC();
  CLH_PROBE(B, 104);
}

#if !defined(CLH_SERIAL)
// the forward solve stages 64 x 64 blocks of the panel in the same LDS
static_assert(kDenseGridLds(sizeof(float)) >= 64 * 64 * sizeof(float) &&
                  kDenseGridLds(sizeof(double)) >= 64 * 64 * sizeof(double),
              "the dense panels' LDS holds a 64 x 64 block");
#endif

// whether a wide supernode's panel is factored by factor_panel_dense_grid
// (its X then holds the super-blocks' diagonal-block inverses only)
template <class FT>
CLH_D bool dense_grid_panel(const Blk& B, const SD& D, int ns, int nr) {
  const int dense_min = D.st->dense_min_cols > 0 ? D.st->dense_min_cols : kDenseGridMinCols;
  return (ns >= dense_min || (size_t)nr * ns >= kDenseGridMinEntries) && B.lds &&
         (size_t)B.lds_bytes >= kDenseGridLds(sizeof(FT));
}

// y <- L_11^{-1} y for a panel of factor_panel_dense_grid (unit-lower
// diagonal block L_11 in the panel Ls, leading dimension nr; y the
// supernode's ns entries), by super-blocks: y_sb <- X_sb y_sb with the
// super-block's diagonal-block inverse, then the later rows of the block
// subtract L[i, sb] y_sb.  scr holds kDenseSB entries.  Two grid barriers
// per super-block.
template <class FT, class LP, class XP, class YP>
CLH_D void dense_solve_forward(const Blk& B, LP Ls, int nr, XP Li, int ns, YP y, YP scr) {
  for (int sb = 0; sb < ns; sb += kDenseSB) {
    const int se = imin(ns, sb + kDenseSB), w = se - sb;
#if defined(CLH_SERIAL)
    for (int i = B.tid; i < w; i += B.nthr) {  // row i of X_sb, its sum in column order
      FT acc = FT(0);
      for (int k = 0; k <= i; ++k) acc += Li[(sb + i) + (size_t)(sb + k) * ns] * y[sb + k];
      scr[i] = acc;
    }
#else
    // one wavefront per row of X_sb, read from its transposed copy above
    // the diagonal (the row's entries contiguous); the partial sums meet in
    // a fixed shuffle order
    {
      const int gw = B.tid >> 6, nw = B.nthr >> 6, lane = B.tid & 63;
      for (int i = gw; i < w; i += nw) {
        double acc = 0.0;
#pragma unroll 8
        for (int k = lane; k <= i; k += 64) acc += double(Li[(sb + k) + (size_t)(sb + i) * ns] * y[sb + k]);
        acc = wave_sum(acc);
        if (lane == 0) scr[i] = FT(acc);
      }
    }
#endif
    CLH_SYNC();
    CLH_PROBE(B, 130);
    // the block's entries take X_sb y_sb; each later row subtracts
    // L[i, sb:se] y_sb, its products summed in column order by one thread
    // (column chunks summed apart and then in order were faster, but their
    // other rounding flipped the last iterations of the borderline random
    // SOCPs to ALMOST_SOLVED)
#if defined(CLH_SERIAL)
    for (int i = B.tid; i < ns - sb; i += B.nthr) {
      if (i < w) {
        y[sb + i] = scr[i];
      } else {
        const auto Lr = Ls + (sb + i) + (size_t)sb * nr;
        FT acc = FT(0);
        for (int k = 0; k < w; ++k) acc += Lr[(size_t)k * nr] * scr[k];
        y[sb + i] -= acc;
      }
    }
#else
    // A workgroup per 32 later rows: its 256 threads stage 32 x 64 blocks of
    // L in LDS (each column's 32 rows one coalesced 256-byte read), fetching
    // the next block into registers while the first wavefront's lanes (one
    // per row) add the current block's products to their sums in column
    // order: the same sums in the same order as one thread reading L
    // directly, which left a few waves per step waiting on one strided load
    // after another (0.24 ms per step on the 18 754-column root, 0.12 ms
    // with 64-row blocks staged without the prefetch).
    for (int i = B.tid; i < w; i += B.nthr) y[sb + i] = scr[i];
    {
      constexpr int RT = 32, PV = RT * 64 / 256;
      const int lt = B.ltid >= 0 ? B.ltid : B.tid;
      CLH_LDS FT* const T = (CLH_LDS FT*)B.lds;  // [r + RT k]
      const int nrb = (ns - se + RT - 1) / RT;
      for (int rb = B.wg; rb < nrb; rb += B.nwg) {
        const int r0 = se + rb * RT;
        FT v[PV];
        auto fetch = [&](int kc) {
          const int kb = imin(64, w - kc);
          for (int u = 0; u < PV; ++u) {
            const int t = lt + 256 * u, r = t % RT, k = t / RT;
            v[u] = k < kb && r0 + r < ns ? Ls[(r0 + r) + (size_t)(sb + kc + k) * nr] : FT(0);
          }
        };
        fetch(0);
        FT acc = FT(0);
        for (int kc = 0; kc < w; kc += 64) {
          const int kb = imin(64, w - kc);
          CLH_BSYNC();
          for (int u = 0; u < PV; ++u) T[lt + 256 * u] = v[u];
          CLH_BSYNC();
          if (kc + 64 < w) fetch(kc + 64);
          if (lt < RT)
            for (int k = 0; k < kb; ++k) acc += T[lt + RT * k] * scr[kc + k];
        }
        if (lt < RT && r0 + lt < ns) y[r0 + lt] -= acc;
      }
    }
#endif
    CLH_SYNC();
    CLH_PROBE(B, 131);
  }
}

// y <- L_11^{-T} y for the same panels, super-blocks from the last: y_sb <-
// X_sb' y_sb, then every earlier entry c subtracts L[sb:se, c]' y_sb (a
// right-looking order: each earlier entry receives the super-blocks'
// contributions from the last to the first).  Two grid barriers per
// super-block.
template <class FT, class LP, class XP, class YP>
CLH_D void dense_solve_backward(const Blk& B, LP Ls, int nr, XP Li, int ns, YP y, YP scr) {
  const int nsb = (ns + kDenseSB - 1) / kDenseSB;
  for (int q = nsb - 1; q >= 0; --q) {
    const int sb = q * kDenseSB, se = imin(ns, sb + kDenseSB), w = se - sb;
#if defined(CLH_SERIAL)
    for (int i = B.tid; i < w; i += B.nthr) {
      FT acc = FT(0);
      for (int k = i; k < w; ++k) acc += Li[(sb + k) + (size_t)(sb + i) * ns] * y[sb + k];
      scr[i] = acc;
    }
    CLH_SYNC();
    for (int c = B.tid; c < se; c += B.nthr) {
      if (c >= sb) {
        y[c] = scr[c - sb];
      } else {
        const auto Lc = Ls + sb + (size_t)c * nr;
        FT acc = FT(0);
        for (int k = 0; k < w; ++k) acc += Lc[k] * scr[k];
        y[c] -= acc;
      }
    }
#else
    // one wavefront per column: the lanes read it in 64-entry stretches, the
    // partial sums meet in a fixed shuffle order
    const int gw = B.tid >> 6, nw = B.nthr >> 6, lane = B.tid & 63;
    for (int i = gw; i < w; i += nw) {
      double acc = 0.0;
#pragma unroll 8
      for (int k = i + lane; k < w; k += 64) acc += double(Li[(sb + k) + (size_t)(sb + i) * ns] * y[sb + k]);
      acc = wave_sum(acc);
      if (lane == 0) scr[i] = FT(acc);
    }
    CLH_SYNC();
    CLH_PROBE(B, 132);
    // the block's entries take X_sb' y_sb; every earlier column c (one
    // wavefront, its w entries of rows sb:se contiguous) subtracts its dot
    // with them
    for (int c = gw; c < se; c += nw) {
      if (c >= sb) {
        if (lane == 0) y[c] = scr[c - sb];
      } else {
        const auto Lc = Ls + sb + (size_t)c * nr;
        double acc = 0.0;
#pragma unroll 8
        for (int k = lane; k < w; k += 64) acc += double(Lc[k] * scr[k]);
        acc = wave_sum(acc);
        if (lane == 0) y[c] -= FT(acc);
      }
    }
#endif
    CLH_SYNC();
    CLH_PROBE(B, 133);
  }
}

// One supernode by the whole grid: the panel (in memory) is zeroed and
// assembled by every thread, its updates are split into (row, column range)
// items over every thread, then one 64-lane group of workgroup 0 factors the
// dense panel.  Each entry receives the same contributions in the same order
// as in the one-group factorization.
template <class FT>
CLH_D void factor_supernode_wide(const Blk& B, const SD& D, const SnDesc& sd, double reg_eps, int& ndyn,
                                 bool& bad) {
  const auto W = gmem(vcast<FT>(D.L) + sd.lptr);
  const auto Kx = gmem(D.Kx);
  const auto asm_src = gmem(D.asm_src);
  const auto asm_dst = gmem(D.asm_dst);
  const auto asm_sgn = gmem(D.asm_sgn);
  const int ns = sd.ns, nr = sd.nr;
  for (int t = B.tid; t < nr * ns; t += B.nthr) W[t] = FT(0);
  CLH_SYNC();
  for (int t = sd.asm0 + B.tid; t < sd.asm1; t += B.nthr) {
    double v = Kx[asm_src[t]];
    int8_t sg = asm_sgn[t];
    if (sg != 0) v = v + (sg > 0 ? reg_eps : -reg_eps);
    W[asm_dst[t]] = FT(v);
  }
  CLH_SYNC();
  CLH_PROBE(B, 100);
  const int lnt = B.lnthr >= 0 ? B.lnthr : B.nthr;
  const int tile_min = D.st->tile_min_cols > 0 ? D.st->tile_min_cols : kTileMinCols;
  if (ns >= tile_min && B.lds && (size_t)B.lds_bytes >= kTiledLds(sizeof(FT)) && 256 % lnt == 0 &&
      256 / lnt <= kTileAcc && D.st->grid_rowwise == 0) {
    tiled_wide_updates<FT>(B, D, sd, W);
  } else {
  const int nch = (ns + kWideCols - 1) / kWideCols;
  // the rows' update lists are split into K chunks when the group scratch
  // holds K copies of the panel: chunk k of row p sums its records in order
  // into copy k, then every entry adds its K partial sums in order
  const auto rp = gmem(D.rowupd_ptr) + sd.r0;
  int mx = 0;
  for (int p = B.tid; p < nr; p += B.nthr) mx = imax(mx, rp[p + 1] - rp[p]);
  const int maxc = (int)blk_max(B, double(mx));
  const size_t panel = (size_t)nr * ns;
  const int K = imax(1, imin((maxc + 31) / 32, (int)(((size_t)D.ngroups * D.maxns) / panel)));
  if (K == 1) {
    for (int it = B.tid; it < nr * nch; it += B.nthr) {
      const int p = it / nch, c0 = (it % nch) * kWideCols;
      apply_row_updates<FT>(D, sd, W, p, c0, imin(ns, c0 + kWideCols));
    }
    CLH_SYNC();
  } else {
    const auto Pt = gmem(vcast<FT>(D.gscr));  // K copies of the panel, row-major (p, c) at p * ns + c
    const int C = (maxc + K - 1) / K;
    for (size_t t = B.tid; t < panel * K; t += B.nthr) Pt[t] = FT(0);
    CLH_SYNC();
    for (int it = B.tid; it < nr * nch * K; it += B.nthr) {
      const int p = it / (nch * K), rem = it % (nch * K), cc = rem / K, k = rem % K;
      const int c0 = cc * kWideCols;
      const int e0 = rp[p] + k * C, e1 = imin(rp[p + 1], e0 + C);
      if (e0 < e1)
        apply_row_updates_range<FT>(D, sd, Pt, (size_t)k * panel + (size_t)p * ns, 1, p, c0,
                                    imin(ns, c0 + kWideCols), e0, e1);
    }
    CLH_SYNC();
    for (size_t t = B.tid; t < panel; t += B.nthr) {
      const int p = (int)(t / ns), c = (int)(t % ns);
      FT w = W[p + (size_t)c * nr];
      for (int k = 0; k < K; ++k) w += Pt[(size_t)k * panel + t];
      W[p + (size_t)c * nr] = w;
    }
    CLH_SYNC();
  }
  }
  CLH_PROBE(B, 101);
  // the whole grid also takes narrow panels when they are tall (one lane
  // group would walk thousands of rows per column)
  if (dense_grid_panel<FT>(B, D, ns, nr)) {
    factor_panel_dense_grid<FT>(B, D, sd, W, ndyn, bad);
    CLH_PROBE(B, 102);
    return;
  }
  const int lt = B.ltid >= 0 ? B.ltid : B.tid, ln = B.lnthr >= 0 ? B.lnthr : B.nthr;
  if (B.wg == 0 && lt < 64) {
    Blk g = B;
    g.G = imin(64, ln); g.gl = lt; g.grp = 0; g.ngrp = 1; g.lgrp = 0; g.lngrp = 1;
    factor_panel_dense<FT, false>(g, D, sd, W, ndyn, bad);
    if (g.gl != 0) ndyn = 0;  // lane 0 reports the group's count
  }
  CLH_SYNC();
  CLH_PROBE(B, 102);
}
#endif

// the block regrouped into lane groups of G for one level
CLH_D Blk level_blk(const Blk& B, int G) {
  Blk b = B;
  if (G > B.nthr) G = B.nthr;
  b.G = G;
  b.gl = B.tid % G;
  b.grp = B.tid / G;
  b.ngrp = B.nthr / G;
  const int lt = B.ltid >= 0 ? B.ltid : B.tid, ln = B.lnthr >= 0 ? B.lnthr : B.nthr;
  b.lgrp = lt / G;
  b.lngrp = imax(1, ln / G);
  return b;
}

#if defined(CLH_GRID)
// levels of up to this many supernodes are factored by phases over the grid
// (every thread walks the level's supernodes in each phase); more numerous
// supernodes are each factored by one lane group
constexpr int kLevelWideMax = 512;

// The dense part of one panel by one workgroup, the diagonal block held in
// LDS: elimination with the dynamic regularization rule, columns scaled, the
// block written back, then its unit-lower part inverted in place (column j of
// the inverse is -X[j+1:, j+1:] L[j+1:, j], right to left) and written to the
// solves' inverse; the rows below are then LD = A L^-T, divided by the
// pivots, by tiles of up to 64 rows staged in LDS.  Needs room in LDS for
// ns^2 + 2 ns + 8 ns values (panels_dense_wg_fits).
template <class FT>
CLH_D bool panel_dense_wg_fits(const Blk& B, int ns) {
  return B.lds && ns * ns + 10 * ns <= B.lds_bytes / (int)sizeof(FT);
}
template <class FT, class WP>
CLH_D void factor_panel_dense_wg(const Blk& B, const SD& D, const SnDesc& sd, WP W, int& ndyn, bool& bad) {
  const DevSettings& st = *D.st;
  const auto Dv = gmem(vcast<FT>(D.Dv));
  const int f = sd.f, ns = sd.ns, nr = sd.nr, ld = nr;
  const int lt = B.ltid >= 0 ? B.ltid : B.tid, ln = B.lnthr >= 0 ? B.lnthr : B.nthr;
  const int cap = B.lds_bytes / (int)sizeof(FT);
  const int TR = imin(64, (cap - ns * ns - 2 * ns) / ns);  // rows per tile below the block
  CLH_LDS FT* X = (CLH_LDS FT*)B.lds;  // the diagonal block, [r + c * ns]
  CLH_LDS FT* P = X + ns * ns;         // pivots
  CLH_LDS FT* tmp = P + ns;            // a column of L
  CLH_LDS FT* T = tmp + ns;            // row tile, [r + c * TR]
  CLH_BSYNC();
  for (int t = lt; t < ns * ns; t += ln) {
    const int r = t % ns, c = t / ns;
    X[t] = r >= c ? W[r + (size_t)c * ld] : FT(0);
  }
  for (int j = 0; j < ns; ++j) {
    CLH_BSYNC();
    FT dj = X[j + j * ns];
    if (st.dynamic_regularization_enable) {
      const int sg = D.dsign[D.perm[f + j]];
      if (double(dj) * sg <= st.dynamic_regularization_eps) {
        dj = FT(dyn_reg_pivot(st, double(dj), sg));
        if (lt == 0) ++ndyn;
      }
    }
    if (dj == FT(0) || !finite(double(dj))) bad = true;
    const FT djinv = FT(1) / dj;
    CLH_BSYNC();
    for (int i = j + 1 + lt; i < ns; i += ln) {
      const FT lij = X[i + j * ns] * djinv;
      for (int c = j + 1; c <= i; ++c) X[i + c * ns] -= lij * X[c + j * ns];
    }
    if (lt == 0) { X[j + j * ns] = dj; Dv[f + j] = dj; }
  }
  CLH_BSYNC();
  for (int t = lt; t < ns; t += ln) P[t] = X[t + t * ns];
  for (int t = lt; t < ns * ns; t += ln) {
    const int r = t % ns, c = t / ns;
    if (r > c) X[t] = X[t] / X[c + c * ns];
  }
  CLH_BSYNC();
  for (int t = lt; t < ns * ns; t += ln) {
    const int r = t % ns, c = t / ns;
    if (r >= c) W[r + (size_t)c * ld] = X[t];
  }
  for (int j = ns - 2; j >= 0; --j) {  // X := L^-1 below the diagonal, in place
    CLH_BSYNC();
    for (int i = j + 1 + lt; i < ns; i += ln) tmp[i] = X[i + j * ns];
    CLH_BSYNC();
    for (int i = j + 1 + lt; i < ns; i += ln) {
      FT a = tmp[i];
      for (int k = j + 1; k < i; ++k) a += X[i + k * ns] * tmp[k];
      X[i + j * ns] = -a;
    }
  }
  CLH_BSYNC();
  if (ns > 1) {
    const auto Lig = gmem(vcast<FT>(D.Linv)) + sd.iptr;
    for (int t = lt; t < ns * ns; t += ln) {
      const int r = t % ns, c = t / ns;
      Lig[t] = r > c ? X[t] : FT(r == c);
    }
  }
  for (int p0 = ns; p0 < nr; p0 += TR) {
    const int tr = imin(TR, nr - p0);
    CLH_BSYNC();
    for (int t = lt; t < tr * ns; t += ln) {
      const int r = t % tr, c = t / tr;
      T[r + c * TR] = W[p0 + r + (size_t)c * ld];
    }
    CLH_BSYNC();
    for (int t = lt; t < tr * ns; t += ln) {
      const int r = t % tr, c = t / tr;
      FT a = T[r + c * TR];
      for (int j = 0; j < c; ++j) a += T[r + j * TR] * X[c + j * ns];
      W[p0 + r + (size_t)c * ld] = a / P[c];
    }
  }
  CLH_BSYNC();
}

// Every supernode of a level at once, by phases over the whole grid: the
// panels zeroed and assembled, then their updates as (supernode, row, column
// range) items, then each panel's dense part by a lane group.  A level of a
// few hundred tall supernodes keeps every workgroup busy this way, where one
// group per supernode would leave most of the grid idle.
template <class FT>
CLH_D void factor_level_wide(const Blk& B, const SD& D, int t0, int t1, int G, double reg_eps, int& ndyn,
                             bool& bad) {
  const auto Lg = gmem(vcast<FT>(D.L));
  const auto Kx = gmem(D.Kx);
  const auto asm_src = gmem(D.asm_src);
  const auto asm_dst = gmem(D.asm_dst);
  const auto asm_sgn = gmem(D.asm_sgn);
  const auto sn_lev = gmem(D.sn_lev);
  for (int t = t0; t < t1; ++t) {
    const SnDesc sd = gload(sn_lev + t);
    const auto W = Lg + sd.lptr;
    for (int x = B.tid; x < sd.nr * sd.ns; x += B.nthr) W[x] = FT(0);
  }
  CLH_SYNC();
  for (int t = t0; t < t1; ++t) {
    const SnDesc sd = gload(sn_lev + t);
    const auto W = Lg + sd.lptr;
    for (int x = sd.asm0 + B.tid; x < sd.asm1; x += B.nthr) {
      double v = Kx[asm_src[x]];
      const int8_t sg = asm_sgn[x];
      if (sg != 0) v = v + (sg > 0 ? reg_eps : -reg_eps);
      W[asm_dst[x]] = FT(v);
    }
  }
  CLH_SYNC();
  CLH_PROBE(B, 110);
  // the level's update items numbered supernode after supernode
  long base = 0;
  for (int t = t0; t < t1; ++t) {
    const SnDesc sd = gload(sn_lev + t);
    const int nch = (sd.ns + kWideCols - 1) / kWideCols;
    const long items = (long)sd.nr * nch;
    long g = base + (((long)B.tid - base) % B.nthr + B.nthr) % B.nthr;
    for (; g < base + items; g += B.nthr) {
      const int local = (int)(g - base), p = local / nch, c0 = (local % nch) * kWideCols;
      apply_row_updates<FT>(D, sd, Lg + sd.lptr, p, c0, imin(sd.ns, c0 + kWideCols));
    }
    base += items;
  }
  CLH_SYNC();
  CLH_PROBE(B, 111);
  // dense parts: a workgroup per panel whose diagonal block fits in LDS, a
  // lane group per other panel
  for (int t = t0 + B.wg; t < t1; t += B.nwg) {
    const SnDesc sd = gload(sn_lev + t);
    if (!panel_dense_wg_fits<FT>(B, sd.ns)) continue;
    int nd = 0;
    factor_panel_dense_wg<FT>(B, D, sd, Lg + sd.lptr, nd, bad);
    ndyn += nd;
  }
  const Blk Bl = level_blk(B, G);
  for (int t = t0 + Bl.grp; t < t1; t += Bl.ngrp) {
    const SnDesc sd = gload(sn_lev + t);
    if (panel_dense_wg_fits<FT>(B, sd.ns)) continue;
    int nd = 0;
    factor_panel_dense<FT, false>(Bl, D, sd, Lg + sd.lptr, nd, bad);
    if (Bl.gl == 0) ndyn += nd;
  }
  CLH_SYNC();
  CLH_PROBE(B, 112);
}
#endif

// a single-column supernode without descendants, by one thread
template <class FT>
CLH_D void factor_leaf(const SD& D, const SnDesc& sd, double reg_eps, int& ndyn, bool& bad) {
  const DevSettings& st = *D.st;
  const auto Dv = gmem(vcast<FT>(D.Dv));
  const auto Kx = gmem(D.Kx);
  const auto asm_src = gmem(D.asm_src);
  const auto asm_dst = gmem(D.asm_dst);
  const auto asm_sgn = gmem(D.asm_sgn);
  const int f = sd.f, nr = sd.nr;
  const auto Ls = gmem(vcast<FT>(D.L)) + sd.lptr;
  for (int t = 0; t < nr; ++t) Ls[t] = FT(0);
  for (int t = sd.asm0; t < sd.asm1; ++t) {
    double v = Kx[asm_src[t]];
    int8_t sg = asm_sgn[t];
    if (sg != 0) v = v + (sg > 0 ? reg_eps : -reg_eps);
    Ls[asm_dst[t]] = FT(v);
  }
  FT dj = Ls[0];
  if (st.dynamic_regularization_enable) {
    const int sg = D.dsign[D.perm[f]];
    if (double(dj) * sg <= st.dynamic_regularization_eps) {
      dj = FT(dyn_reg_pivot(st, double(dj), sg));
      ++ndyn;
    }
  }
  if (dj == FT(0) || !finite(double(dj))) bad = true;
  Dv[f] = dj;
  for (int i = 1; i < nr; ++i) Ls[i] = Ls[i] / dj;
}

// the workgroup regrouped into lane groups of G (at most the workgroup size)

template <class FT>
CLH_D bool ldl_factor(const Blk& B, const SD& D, double reg_eps, int& ndyn_total) {
  int ndyn = 0, nleafdyn = 0;
  bool bad = false;
  for (int lev = 0; lev < D.nlev; ++lev) {
    CLH_PROBE(B, 200 + lev);
    int t0 = D.lev_ptr[lev];
    if (lev == 0) {
      for (int t = t0 + B.tid; t < t0 + D.nleaf; t += B.nthr) {
        const SnDesc sd = gload(gmem(D.sn_lev) + t);
        factor_leaf<FT>(D, sd, reg_eps, nleafdyn, bad);
      }
      t0 += D.nleaf;
    }
#if defined(CLH_GRID)
    if (D.lev_ptr[lev + 1] - t0 <= kWideLevel) {
      // few supernodes (the top of the tree): the whole grid takes them one
      // at a time
      for (int t = t0; t < D.lev_ptr[lev + 1]; ++t) {
        const SnDesc sd = gload(gmem(D.sn_lev) + t);
        int nd = 0;
        factor_supernode_wide<FT>(B, D, sd, reg_eps, nd, bad);
        ndyn += nd;
      }
      CLH_SYNC();  // also the barrier after the level's leaves when it has no other supernode
      continue;
    }
    if (D.lev_ptr[lev + 1] - t0 <= kLevelWideMax) {
      factor_level_wide<FT>(B, D, t0, D.lev_ptr[lev + 1], D.lev_group[lev], reg_eps, ndyn, bad);
      continue;
    }
#endif
    const Blk Bl = level_blk(B, D.lev_group[lev]);
    for (int t = t0 + Bl.grp; t < D.lev_ptr[lev + 1]; t += Bl.ngrp) {
      const SnDesc sd = gload(gmem(D.sn_lev) + t);
      int nd = 0;  // every lane of the group counts the pivots; lane 0 reports
      factor_supernode<FT>(Bl, D, sd, reg_eps, nd, bad);
      if (Bl.gl == 0) ndyn += nd;
    }
    CLH_SYNC();
  }
  CLH_PROBE(B, 200 + D.nlev);
  double nd = blk_sum(B, double(ndyn) + double(nleafdyn));
  ndyn_total = int(nd);
  return blk_max(B, bad ? 1.0 : 0.0) == 0.0;
}

// x = K^{-1} b using the factors: permute, L, D, L', unpermute.  YP is the
// work vector's pointer type: the workgroup's LDS area when the vector fits
// (inlds), else memory.
template <class FT, class YP>
CLH_D void ldl_solve_body(const Blk& B, const SD& D, VCD b, VD x, YP y, YP scr0, YP cbuf0, int cbufcap, bool inlds) {
  const auto L = gmem(vcast<FT>(D.L));
  const auto Dv = gmem(vcast<FT>(D.Dv));
  const auto Linv = gmem(vcast<FT>(D.Linv));
  const auto perm = gmem(D.perm);
  const auto lev_ptr = gmem(D.lev_ptr);
  const auto lev_group = gmem(D.lev_group);
  const auto sn_lev = gmem(D.sn_lev);
  const auto rowupd_ptr = gmem(D.rowupd_ptr);
  const auto rowupd = gmem(D.rowupd);
  const int N = D.N;
#define CLH_WSYNC_Y()            \
  do {                           \
    if (inlds) CLH_WSYNC_LDS();  \
    else CLH_WSYNC();            \
  } while (0)
  for (int k = B.tid; k < N; k += B.nthr) y[k] = FT(b[perm[k]]);
  CLH_SYNC();
  for (int lev = 0; lev < D.nlev; ++lev) {
    CLH_PROBE(B, 300 + lev);
    const Blk Bl = level_blk(B, lev_group[lev]);
    const int gl = Bl.gl, G = Bl.G;
    const YP scr = scr0 + (size_t)Bl.grp * D.maxns;
    // single-column leaves have nothing to eliminate in the forward sweep
    const int t0 = lev_ptr[lev] + (lev == 0 ? D.nleaf : 0), t1 = lev_ptr[lev + 1];
#if defined(CLH_GRID)
    if (t1 - t0 <= kWideLevel) {
      // few supernodes: the grid takes them one at a time.  The
      // contributions to each diagonal row are split into chunks of its
      // update list, each chunk summed in order by one thread into the
      // scratch, then each row subtracts its chunks in order.
      const size_t capacity = (size_t)D.ngroups * D.maxns;
      for (int t = t0; t < t1; ++t) {
        const SnDesc sd = gload(sn_lev + t);
        const int f = sd.f, ns = sd.ns;
        const auto rp = rowupd_ptr + sd.r0;
        int mx = 0;
        for (int j = B.tid; j < ns; j += B.nthr) mx = imax(mx, rp[j + 1] - rp[j]);
        const int maxc = (int)blk_max(B, double(mx));
        int K = imax(1, (maxc + 63) / 64);
        if ((size_t)ns * K > capacity) K = imax(1, (int)(capacity / ns));
        const int C = imax(1, (maxc + K - 1) / K);
        for (int it = B.tid; it < ns * K; it += B.nthr) {
          const int j = it / K, k = it % K;
          const int e0 = rp[j] + k * C, e1 = imin(rp[j + 1], e0 + C);
          FT acc = FT(0);
          for (int e = e0; e < e1; ++e) {
            const RowUpd r = gload(rowupd + e);
            const auto La = L + r.base + r.a;
            const int nsd = r.nsd, nrd = r.nrd, fd = r.fd;
#pragma unroll 8
            for (int kk = 0; kk < nsd; ++kk) acc += La[kk * nrd] * y[fd + kk];
          }
          scr0[it] = acc;
        }
        CLH_SYNC();
        CLH_PROBE(B, 120);
        for (int j = B.tid; j < ns; j += B.nthr) {
          FT yj = y[f + j];
          for (int k = 0; k < K; ++k) yj -= scr0[(size_t)j * K + k];
          y[f + j] = yj;
        }
        CLH_SYNC();
        CLH_PROBE(B, 121);
        if (ns > 1 && dense_grid_panel<FT>(B, D, ns, sd.nr)) {
          dense_solve_forward<FT>(B, L + sd.lptr, sd.nr, Linv + sd.iptr, ns, y + f, scr0);
        } else if (ns > 1) {
          const auto Li = Linv + sd.iptr;
#if defined(CLH_SERIAL)
          for (int i = B.tid; i < ns; i += B.nthr) {
            FT acc = FT(0);
            for (int k = 0; k <= i; ++k) acc += Li[i + k * ns] * y[f + k];
            scr0[i] = acc;
          }
          CLH_SYNC();
          for (int i = B.tid; i < ns; i += B.nthr) y[f + i] = scr0[i];
          CLH_SYNC();
#else
          // y <- X y (X lower triangular, column-major): a wavefront takes 64
          // rows (one per lane, loads coalesced along X's columns) over a
          // chunk of CW columns into the scratch; each row then sums its
          // chunks in order.  CW keeps the chunks' partials within the scratch.
          const int maxch = (int)(capacity / ns);  // >= 1: the scratch holds maxns entries per group
          const int CW = imax(256, ((ns + maxch - 1) / maxch + 63) / 64 * 64);
          const int nrb = (ns + 63) / 64, nch = (ns + CW - 1) / CW;  // nch <= maxch
          const int gw = B.tid >> 6, nw = B.nthr >> 6, lane = B.tid & 63;
          for (int it = gw; it < nrb * nch; it += nw) {
            const int rb = it % nrb, c = it / nrb, i = rb * 64 + lane, k0 = c * CW;
            if (k0 > rb * 64 + 63) continue;  // above the diagonal: the same on every lane
            if (i < ns) {
              const int ke = imin(imin(ns, k0 + CW), i + 1);
              FT acc = FT(0);
#pragma unroll 8
              for (int k = k0; k < ke; ++k) acc += Li[i + (size_t)k * ns] * y[f + k];
              scr0[(size_t)c * ns + i] = acc;
            }
          }
          CLH_SYNC();
          for (int i = B.tid; i < ns; i += B.nthr) {
            FT acc = FT(0);
            for (int c = 0; c <= i / CW; ++c) acc += scr0[(size_t)c * ns + i];
            y[f + i] = acc;
          }
          CLH_SYNC();
#endif
        }
      }
      CLH_SYNC();
      CLH_PROBE(B, 122);
      continue;
    }
#endif
    // this group's share of the contribution buffer: split among the groups
    // that have a supernode at this level
    const int cshare = cbufcap / imin(Bl.ngrp, imax(1, t1 - t0));
    const YP cbuf = cbuf0 + (size_t)Bl.grp * cshare;
    for (int t = t0 + Bl.grp; t < t1; t += Bl.ngrp) {
      const SnDesc sd = gload(sn_lev + t);
      const int f = sd.f, ns = sd.ns;
      const auto rp = rowupd_ptr + sd.r0;
      const int c0 = rp[0], c1 = rp[ns];  // contributions to the diagonal rows, row by row
      if (c1 - c0 <= cshare) {
        // every contribution computed in parallel into the buffer, then each
        // row owner subtracts its own in update order
        for (int e = c0 + gl; e < c1; e += G) {
          const RowUpd r = gload(rowupd + e);
          const auto La = L + r.base + r.a;
          const int nsd = r.nsd, nrd = r.nrd, fd = r.fd;
          FT acc = FT(0);
#pragma unroll 8
          for (int k = 0; k < nsd; ++k) acc += La[k * nrd] * y[fd + k];
          cbuf[e - c0] = acc;
        }
        CLH_WSYNC_Y();
        for (int j = gl; j < ns; j += G) {
          FT yj = y[f + j];
          const int e1 = rp[j + 1];
          for (int e = rp[j]; e < e1; ++e) yj -= cbuf[e - c0];
          y[f + j] = yj;
        }
        CLH_WSYNC_Y();
      } else {  // row owners: lane j applies every descendant's contribution to y[f + j] in update order
        for (int j = gl; j < ns; j += G) {
          FT yj = y[f + j];
          const int e0 = rp[j], e1 = rp[j + 1];
          RowUpd r = gload(rowupd + e0);  // the list's records: one ahead (the record past the end is readable)
          for (int e = e0; e < e1; ++e) {
            const RowUpd rn = gload(rowupd + e + 1);
            const auto La = L + r.base + r.a;
            const int nsd = r.nsd, nrd = r.nrd, fd = r.fd;
            FT acc = FT(0);
#pragma unroll 8
            for (int k = 0; k < nsd; ++k) acc += La[k * nrd] * y[fd + k];
            yj -= acc;
            r = rn;
          }
          y[f + j] = yj;
        }
        CLH_WSYNC_Y();
      }
      if (ns > 1) {
        const auto Li = Linv + sd.iptr;
        for (int i = gl; i < ns; i += G) {
          FT acc = FT(0);
#pragma unroll 8
          for (int k = 0; k <= i; ++k) acc += Li[i + k * ns] * y[f + k];
          scr[i] = acc;
        }
        CLH_WSYNC_Y();
        for (int i = gl; i < ns; i += G) y[f + i] = scr[i];
        CLH_WSYNC_Y();
      }
    }
    CLH_SYNC();
    CLH_PROBE(B, 123);
  }
  CLH_PROBE(B, 300 + D.nlev);
  for (int k = B.tid; k < N; k += B.nthr) y[k] = y[k] / Dv[k];
  CLH_SYNC();
  const auto sn_rows = gmem(D.sn_rows);
  for (int lev = D.nlev - 1; lev >= 0; --lev) {
    CLH_PROBE(B, 400 + lev);
    const Blk Bl = level_blk(B, lev_group[lev]);
    const int gl = Bl.gl, G = Bl.G;
    const YP scr = scr0 + (size_t)Bl.grp * D.maxns;
    int t0 = lev_ptr[lev];
    const int t1 = lev_ptr[lev + 1];
    if (lev == 0) {
      for (int t = t0 + B.tid; t < t0 + D.nleaf; t += B.nthr) {
        const SnDesc sd = gload(sn_lev + t);
        const auto Ls = L + sd.lptr;
        const auto rows = sn_rows + sd.r0;
        FT acc = FT(0);
#pragma unroll 8
        for (int t2 = 1; t2 < sd.nr; ++t2) acc += Ls[t2] * y[rows[t2]];
        y[sd.f] -= acc;
      }
      t0 += D.nleaf;
    }
#if defined(CLH_GRID)
    if (t1 - t0 <= kWideLevel) {  // few supernodes: the grid takes them one at a time
      for (int t = t0; t < t1; ++t) {
        const SnDesc sd = gload(sn_lev + t);
        const int f = sd.f, ns = sd.ns, nr = sd.nr;
        const auto Ls = L + sd.lptr;
        const auto rows = sn_rows + sd.r0;
#if defined(CLH_SERIAL)
        for (int jj = B.tid; jj < ns; jj += B.nthr) {
          FT acc = FT(0);
#pragma unroll 8
          for (int t2 = ns; t2 < nr; ++t2) acc += Ls[t2 + jj * nr] * y[rows[t2]];
          y[f + jj] -= acc;
        }
        CLH_SYNC();
        if (ns > 1 && dense_grid_panel<FT>(B, D, ns, nr)) {
          dense_solve_backward<FT>(B, Ls, nr, Linv + sd.iptr, ns, y + f, scr0);
        } else if (ns > 1) {
          const auto Li = Linv + sd.iptr;
          for (int i = B.tid; i < ns; i += B.nthr) {
            FT acc = FT(0);
#pragma unroll 8
            for (int k = i; k < ns; ++k) acc += Li[k + i * ns] * y[f + k];
            scr0[i] = acc;
          }
          CLH_SYNC();
          for (int i = B.tid; i < ns; i += B.nthr) y[f + i] = scr0[i];
          CLH_SYNC();
        }
#else
        // the column products (L's rows below the panel, then X^T y) by one
        // wavefront per column: its lanes read the column in 64-entry
        // stretches and the partial sums meet in a fixed shuffle order
        const int gw = B.tid >> 6, nw = B.nthr >> 6, lane = B.tid & 63;
        if (nr > ns) {
          for (int jj = gw; jj < ns; jj += nw) {
            double acc = 0.0;
            for (int t2 = ns + lane; t2 < nr; t2 += 64) acc += double(Ls[t2 + (size_t)jj * nr] * y[rows[t2]]);
            acc = wave_sum(acc);
            if (lane == 0) y[f + jj] -= FT(acc);
          }
          CLH_SYNC();
        }
        if (ns > 1 && dense_grid_panel<FT>(B, D, ns, nr)) {
          dense_solve_backward<FT>(B, Ls, nr, Linv + sd.iptr, ns, y + f, scr0);
        } else if (ns > 1) {
          const auto Li = Linv + sd.iptr;
          for (int i = gw; i < ns; i += nw) {
            double acc = 0.0;
            for (int k = i + lane; k < ns; k += 64) acc += double(Li[k + (size_t)i * ns] * y[f + k]);
            acc = wave_sum(acc);
            if (lane == 0) scr0[i] = FT(acc);
          }
          CLH_SYNC();
          for (int i = B.tid; i < ns; i += B.nthr) y[f + i] = scr0[i];
          CLH_SYNC();
        }
#endif
      }
      CLH_SYNC();
      CLH_PROBE(B, 124);
      continue;
    }
#endif
    for (int t = t0 + Bl.grp; t < t1; t += Bl.ngrp) {
      const SnDesc sd = gload(sn_lev + t);
      const int f = sd.f, ns = sd.ns, nr = sd.nr;
      const auto Ls = L + sd.lptr;
      const auto rows = sn_rows + sd.r0;
      if (!inlds && nr - ns <= D.maxns) {
        // y in memory: the rows' entries of y gathered once by the group into
        // its scratch, so that a lane's sum over its column reads two
        // contiguous streams instead of an index and then the entry it names
        // (the same products summed in the same order)
        for (int t2 = ns + gl; t2 < nr; t2 += G) scr[t2 - ns] = y[rows[t2]];
        CLH_WSYNC_Y();
        for (int jj = gl; jj < ns; jj += G) {
          FT acc = FT(0);
#pragma unroll 8
          for (int t2 = ns; t2 < nr; ++t2) acc += Ls[t2 + jj * nr] * scr[t2 - ns];
          y[f + jj] -= acc;
        }
      } else {
        for (int jj = gl; jj < ns; jj += G) {
          FT acc = FT(0);
#pragma unroll 8
          for (int t2 = ns; t2 < nr; ++t2) acc += Ls[t2 + jj * nr] * y[rows[t2]];
          y[f + jj] -= acc;
        }
      }
      CLH_WSYNC_Y();
      if (ns > 1) {
        const auto Li = Linv + sd.iptr;
        for (int i = gl; i < ns; i += G) {
          FT acc = FT(0);
#pragma unroll 8
          for (int k = i; k < ns; ++k) acc += Li[k + i * ns] * y[f + k];
          scr[i] = acc;
        }
        CLH_WSYNC_Y();
        for (int i = gl; i < ns; i += G) y[f + i] = scr[i];
        CLH_WSYNC_Y();
      }
    }
    CLH_SYNC();
    CLH_PROBE(B, 125);
  }
  CLH_PROBE(B, 399);
  for (int k = B.tid; k < N; k += B.nthr) x[perm[k]] = double(y[k]);
  CLH_SYNC();
#undef CLH_WSYNC_Y
}

template <class FT>
CLH_D void ldl_solve(const Blk& B, const SD& D, VCD b, VD x) {
  // LDS: the vector, scratch for up to B.ngrp groups (the smallest group size
  // used), and what is left as the forward sweep's contribution buffer
#if !defined(CLH_LANES) && !defined(CLH_GRID)
  const size_t used = (size_t)D.N + (size_t)B.ngrp * D.maxns;
  if (B.lds && used * sizeof(FT) <= (size_t)B.lds_bytes) {
    CLH_LDS FT* y = (CLH_LDS FT*)B.lds;
    const int cap = (int)(B.lds_bytes / sizeof(FT) - used);
    ldl_solve_body<FT>(B, D, b, x, y, y + D.N, y + used, cap, true);
    return;
  }
#endif
  const auto y = gmem(vcast<FT>(D.ly));
  ldl_solve_body<FT>(B, D, b, x, y, gmem(vcast<FT>(D.gscr)), y, 0, false);
}

// e = b - K xi with the unregularized FP64 KKT; returns ||e||_inf
CLH_D double refine_error(const Blk& B, const SD& D, VD e, VCD b, VCD xi_) {
  PhaseTimer pt(B, lane_prof(D), 5);
  const auto Kfp = gmem(D.Kfp);
  const auto Kfk = gmem(D.Kfk);
  const auto Kfi = gmem(D.Kfi);
  const auto Kx = gmem(D.Kx);
  const auto xi = gmem(xi_);
  double t = 0;
  for (int i = B.tid; i < D.N; i += B.nthr) {
    const int k1 = Kfp[i + 1];
#if defined(CLH_LONG_ROWS)
    if (k1 - Kfp[i] > kLongRow) continue;
#endif
    double acc = 0;
#pragma unroll 8
    for (int k = Kfp[i]; k < k1; ++k) acc += Kx[Kfk[k]] * xi[Kfi[k]];
    double v = b[i] - acc;
    e[i] = v;
    t = dmax(t, fabs(v));
  }
#if defined(CLH_LONG_ROWS)
  const int lane = B.tid & 63;
  for (int i = B.tid >> 6; i < D.N; i += B.nthr >> 6) {
    const int k0 = Kfp[i], k1 = Kfp[i + 1];
    if (k1 - k0 <= kLongRow) continue;  // the same on every lane of the wavefront
    double acc = 0;
    for (int k = k0 + lane; k < k1; k += 64) acc += Kx[Kfk[k]] * xi[Kfi[k]];
    acc = wave_sum(acc);
    if (lane == 0) {
      const double v = b[i] - acc;
      e[i] = v;
      t = dmax(t, fabs(v));
    }
  }
#endif
  double r = blk_max(B, t);
  return r;
}

// factor / solve in the slot's current precision
CLH_D bool use_fp32(const SD& D, const Scal& S) { return D.factor_fp32 && !S.prec64; }
CLH_D void ldl_solve_p(const Blk& B, const SD& D, const Scal& S, VCD b, VD x) {
  PhaseTimer pt(B, lane_prof(D), 4);
  if (use_fp32(D, S)) ldl_solve<float>(B, D, b, x);
  else ldl_solve<double>(B, D, b, x);
}

#if !defined(CLH_LANES)
// ---------------------------------------------------------------------------
// LDL' plugin operations on a standalone matrix: refactor!/solve! of the
// reference's LDL solver protocol (directldl_cudss.jl:33-57) and the
// refinement of kktsolver_directldl_gpu.jl:349-428.  `vals` is the caller's
// full symmetric CSR, `map` the CSR position of each triu CSC entry.
// ---------------------------------------------------------------------------
struct LdlOut {
  double ok, ndyn, steps, err;
};
CLH_D void ldl_gather(const Blk& B, const SD& D, const double* vals, const int32_t* map) {
  const auto v = gmem(vals);
  const auto mp = gmem(map);
  for (int k = B.tid; k < D.nnzK; k += B.nthr) D.Kx[k] = v[mp[k]];
  CLH_SYNC();
}
CLH_D void ldl_op_refactor(const Blk& B, const SD& D, const double* vals, const int32_t* map, double static_eps,
                           LdlOut* out) {
  ldl_gather(B, D, vals, map);
  int nd = 0;
  const bool ok = D.factor_fp32 ? ldl_factor<float>(B, D, static_eps, nd) : ldl_factor<double>(B, D, static_eps, nd);
  if (B.tid == 0) { out->ok = ok ? 1 : 0; out->ndyn = nd; out->steps = 0; out->err = 0; }
}
CLH_D void ldl_op_solve(const Blk& B, const SD& D, double* x, const double* b) {
  if (D.factor_fp32) ldl_solve<float>(B, D, b, x);
  else ldl_solve<double>(B, D, b, x);
}
// x = K^{-1} b refined against the (unregularized) values given
CLH_D void ldl_op_solve_refine(const Blk& B, const SD& D, const double* vals, const int32_t* map, double* xout,
                               const double* b, double reltol, double abstol, int maxiter, double stopratio,
                               LdlOut* out) {
  ldl_gather(B, D, vals, map);
  VD x = D.kx, dx = D.kw2, e = D.kw1;
  ldl_op_solve(B, D, x, b);
  const double normb = blk_norminf(B, b, D.N);
  double norme = refine_error(B, D, e, b, x);
  int ok = finite(norme) ? 1 : 0, steps = 0;
  if (ok) {
    for (int it = 0; it < maxiter; ++it) {
      if (norme <= abstol + reltol * normb) break;
      const double last = norme;
      ldl_op_solve(B, D, dx, e);
      ++steps;
      for (int i = B.tid; i < D.N; i += B.nthr) dx[i] += x[i];
      CLH_SYNC();
      norme = refine_error(B, D, e, b, dx);
      if (!finite(norme)) { ok = 0; break; }
      const double ratio = last / norme;
      if (ratio < stopratio) {
        if (ratio > 1) { VD t = x; x = dx; dx = t; }
        else norme = last;
        break;
      }
      VD t = x; x = dx; dx = t;
    }
  }
  for (int i = B.tid; i < D.N; i += B.nthr) xout[i] = x[i];
  CLH_SYNC();
  if (B.tid == 0) { out->ok = ok; out->ndyn = 0; out->steps = steps; out->err = norme; }
}
#endif  // !CLH_LANES

// static regularization + numeric factorization
// (kktsolver_directldl.jl:239-296, _compute_regularizer :299-312)
CLH_D bool kkt_refactor(const Blk& B, const SD& D, Scal& S) {
  PhaseTimer pt(B, lane_prof(D), 3);
  const DevSettings& st = *D.st;
  double eps = 0;
  if (st.static_regularization_enable) {
    double t = 0;
    for (int i = B.tid; i < D.N; i += B.nthr) t = dmax(t, fabs(D.Kx[D.Kdiag[i]]));
    double maxdiag = blk_max(B, t);
    double c0 = (D.factor_fp32 && S.prec64) ? st.mixed_fp64_static_reg : st.static_regularization_constant;
    eps = c0 + st.static_regularization_proportional * maxdiag;
  }
  S.static_eps = eps;
  int ndyn = 0;
  bool ok;
  if (use_fp32(D, S)) {
    ok = ldl_factor<float>(B, D, eps, ndyn);
    S.n_fp32_factors++;
  } else {
    ok = ldl_factor<double>(B, D, eps, ndyn);
  }
  S.n_dynreg = ndyn;
  return ok;
}

// kktsolver_solve! with _iterative_refinement (kktsolver_directldl.jl:352-454)
// solves for D.kb; the answer is returned through *xout (D.kx or D.kw2).
// Returns 0 on failure, 1 on success, 2 when FP32 factors left the refinement
// stalled above the mixed-precision tolerance.
CLH_D int kkt_solve_ir_once(const Blk& B, const SD& D, Scal& S, VD* xout) {
  const DevSettings& st = *D.st;
  VD x = D.kx;
  VD dx = D.kw2;
  VD e = D.kw1;
  ldl_solve_p(B, D, S, D.kb, x);
  S.kkt_solves++;
  int ok = 1;
  if (st.iterative_refinement_enable) {
    double normb = blk_norminf(B, D.kb, D.N);
    double norme = refine_error(B, D, e, D.kb, x);
    double enorm = norme;
    if (!finite(norme)) ok = 0;
    else {
      for (int it = 0; it < st.iterative_refinement_max_iter; ++it) {
        if (norme <= st.iterative_refinement_abstol + st.iterative_refinement_reltol * normb) break;
        double last = norme;
        ldl_solve_p(B, D, S, e, dx);
        S.ir_steps++;
        for (int i = B.tid; i < D.N; i += B.nthr) dx[i] += x[i];
        CLH_SYNC();
        norme = refine_error(B, D, e, D.kb, dx);
        if (!finite(norme)) { ok = 0; break; }
        double ratio = last / norme;
        if (ratio < st.iterative_refinement_stop_ratio) {
          if (ratio > 1) { VD t = x; x = dx; dx = t; enorm = norme; }
          else enorm = last;
          break;
        }
        VD t = x; x = dx; dx = t;
        enorm = norme;
      }
      if (ok && use_fp32(D, S) && enorm > st.mixed_switch_reltol * dmax(1.0, normb)) ok = 2;
    }
  } else {
    double bad = 0;
    for (int i = B.tid; i < D.N; i += B.nthr) if (!finite(x[i])) bad = 1;
    ok = blk_max(B, bad) == 0 ? 1 : 0;
  }
  *xout = x;
  CLH_SYNC();
  return ok;
}

// the solve, moving the slot to FP64 factors when refinement with the FP32
// factors stalls or breaks down
CLH_D bool kkt_solve_ir(const Blk& B, const SD& D, Scal& S, VD* xout) {
  const bool fp32 = use_fp32(D, S);
  int r = kkt_solve_ir_once(B, D, S, xout);
  if (fp32 && r != 1) {
    S.prec64 = 1;
    if (!kkt_refactor(B, D, S)) return false;
    r = kkt_solve_ir_once(B, D, S, xout);
  }
  return r == 1;
}

CLH_D void set_rhs(const Blk& B, const SD& D, VCD rx, VCD rz) {
  for (int i = B.tid; i < D.N; i += B.nthr)
    D.kb[i] = i < D.n ? rx[i] : (i < D.n + D.m ? rz[i - D.n] : 0.0);
  CLH_SYNC();
}
CLH_D void get_lhs(const Blk& B, const SD& D, VCD xs, VD lx, VD lz) {
  if (lx) for (int i = B.tid; i < D.n; i += B.nthr) lx[i] = xs[i];
  if (lz) for (int i = B.tid; i < D.m; i += B.nthr) lz[i] = xs[D.n + i];
  CLH_SYNC();
}

// kktsolver_update! + regularize + refactor + constant-rhs solve
// (src/kktsystem.jl:57-82, kktsolver_directldl.jl:195-296)
CLH_D bool kkt_update(const Blk& B, const SD& D, Scal& S) {
  kkt_write_Hs(B, D);
  if (!kkt_refactor(B, D, S)) {
    if (!use_fp32(D, S)) return false;
    S.prec64 = 1;  // an FP32 factorization that breaks down is redone in FP64
    if (!kkt_refactor(B, D, S)) return false;
  }
  for (int i = B.tid; i < D.n; i += B.nthr) D.workx[i] = -D.q[i];
  CLH_SYNC();
  set_rhs(B, D, D.workx, D.b);
  VD xs;
  bool ok = kkt_solve_ir(B, D, S, &xs);
  if (ok) get_lhs(B, D, xs, D.x2, D.z2);
  return ok;
}

// kkt_solve! (src/kktsystem.jl:130-196); lhs = (dx,ds,dz,dtau,dkappa), rhs = (hx,hs,hz,rtau_rhs,rkappa_rhs)
CLH_D bool kkt_solve(const Blk& B, const SD& D, Scal& S, bool affine) {
  const int n = D.n, m = D.m;
  VD dsc = D.wconic;
  if (affine) {
    for (int i = B.tid; i < m; i += B.nthr) dsc[i] = D.s[i];
    CLH_SYNC();
  } else {
    cones_ds_from_dz_offset(B, D, dsc, D.hs, D.z);
  }
  for (int i = B.tid; i < m; i += B.nthr) D.workz[i] = dsc[i] - D.hz[i];
  CLH_SYNC();
  set_rhs(B, D, D.hx, D.workz);
  VD xs;
  if (!kkt_solve_ir(B, D, S, &xs)) return false;
  get_lhs(B, D, xs, D.x1, D.z1);
  // Δτ
  for (int i = B.tid; i < n; i += B.nthr) D.workx[i] = D.x[i] / S.tau;
  CLH_SYNC();
  double qx1 = blk_dot(B, D.q, D.x1, n), bz1 = blk_dot(B, D.b, D.z1, m);
  double xiPx1 = blk_quad(B, D, D.workx, D.x1);
  double tau_num = S.rtau_rhs - S.rkappa_rhs / S.tau + qx1 + bz1 + 2 * xiPx1;
  for (int i = B.tid; i < n; i += B.nthr) D.workx[i] -= D.x2[i];
  CLH_SYNC();
  double qx2 = blk_dot(B, D.q, D.x2, n), bz2 = blk_dot(B, D.b, D.z2, m);
  double tau_den = S.kappa / S.tau - qx2 - bz2;
  tau_den += blk_quad(B, D, D.workx, D.workx) - blk_quad(B, D, D.x2, D.x2);
  double dtau = tau_num / tau_den;
  S.dtau = dtau;
  for (int i = B.tid; i < n; i += B.nthr) D.dx[i] = D.x1[i] + dtau * D.x2[i];
  for (int i = B.tid; i < m; i += B.nthr) D.dz[i] = D.z1[i] + dtau * D.z2[i];
  CLH_SYNC();
  cones_mul_Hs(B, D, D.ds, D.dz);
  for (int i = B.tid; i < m; i += B.nthr) D.ds[i] = -(D.ds[i] + dsc[i]);
  CLH_SYNC();
  S.dkappa = -(S.rkappa_rhs + S.kappa * dtau) / S.tau;
  return true;
}

// ---------------------------------------------------------------------------
// residuals, info, termination
// ---------------------------------------------------------------------------
CLH_D void residuals_update(const Blk& B, const SD& D, Scal& S) {  // src/residuals.jl:1-37
  PhaseTimer pt(B, lane_prof(D), 1);
  const int n = D.n, m = D.m;
  double qx = blk_dot(B, D.q, D.x, n);
  double bz = blk_dot(B, D.b, D.z, m);
  double sz = blk_dot(B, D.s, D.z, m);
  spmv_csr(B, n, D.Pp, D.Pi, D.Px, D.x, D.Pxv, 1.0, 0.0);
  spmv_csr(B, n, D.Atp, D.Ati, D.Atx, D.z, D.rx_inf, -1.0, 0.0);
  {
    const auto Ap = gmem(D.Ap);
    const auto Ai = gmem(D.Ai);
    const auto Ax = gmem(D.Ax);
    const auto xv = gmem(D.x);
    for (int i = B.tid; i < m; i += B.nthr) {
      double t = 0;
      const int k1 = Ap[i + 1];
#pragma unroll 8
      for (int k = Ap[i]; k < k1; ++k) t += Ax[k] * xv[Ai[k]];
      D.rz_inf[i] = D.s[i] + t;
    }
  }
  CLH_SYNC();
  double xPx = blk_dot(B, D.x, D.Pxv, n);
  for (int i = B.tid; i < n; i += B.nthr) D.rx[i] = D.rx_inf[i] - D.Pxv[i] - D.q[i] * S.tau;
  for (int i = B.tid; i < m; i += B.nthr) D.rz[i] = D.rz_inf[i] - D.b[i] * S.tau;
  CLH_SYNC();
  S.rtau = qx + bz + S.kappa + xPx / S.tau;
  S.dot_qx = qx; S.dot_bz = bz; S.dot_sz = sz; S.dot_xPx = xPx;
}

CLH_D void info_update(const Blk& B, const SD& D, Scal& S) {  // src/info.jl:1-64
  PhaseTimer pt(B, lane_prof(D), 1);
  const int n = D.n, m = D.m;
  double tinv = 1 / S.tau;
  double cinv = 1 / S.c;
  double xPx2 = S.dot_xPx * tinv * tinv / 2;
  S.cost_primal = (+S.dot_qx * tinv + xPx2) * cinv;
  S.cost_dual = (-S.dot_bz * tinv - xPx2) * cinv;
  double normx = blk_norm_scaled(B, D.d, D.x, n);
  double normz = blk_norm_scaled(B, D.e, D.z, m) * cinv;
  double norms = blk_norm_scaled(B, D.einv, D.s, m);
  S.res_primal_inf = (blk_norm_scaled(B, D.dinv, D.rx_inf, n) * cinv) / dmax(1.0, normz);
  double t1 = blk_norm_scaled(B, D.dinv, D.Pxv, n) / dmax(1.0, normx);
  double t2 = blk_norm_scaled(B, D.einv, D.rz_inf, m) / dmax(1.0, normx + norms);
  S.res_dual_inf = dmax(t1, t2);
  normx *= tinv; normz *= tinv; norms *= tinv;
  S.res_primal = blk_norm_scaled(B, D.einv, D.rz, m) * tinv / dmax(1.0, S.normb + normx + norms);
  S.res_dual = blk_norm_scaled(B, D.dinv, D.rx, n) * tinv * cinv / dmax(1.0, S.normq + normx + normz);
  S.gap_abs = fabs(S.cost_primal - S.cost_dual);
  S.gap_rel = S.gap_abs / dmax(1.0, dmin(fabs(S.cost_primal), fabs(S.cost_dual)));
  S.ktratio = S.kappa * tinv;
}

CLH_D void check_convergence(Scal& S, double tga, double tgr, double tf, double tia, double tir, double tkt,
                             int ss, int ps, int ds) {  // src/info.jl:328-408
  if (S.ktratio <= 1 && ((S.gap_abs < tga) || (S.gap_rel < tgr)) && (S.res_primal < tf) && (S.res_dual < tf))
    S.status = ss;
  else if (S.ktratio > 1000. / tkt) {
    if ((S.dot_bz < -tia) && (S.res_primal_inf < -tir * S.dot_bz)) S.status = ps;
    else if ((S.dot_qx < -tia) && (S.res_dual_inf < -tir * S.dot_qx)) S.status = ds;
  }
}
// `late`: the solve has run past its time limit (info.total_time >
// settings.time_limit)
CLH_D bool check_termination(const SD& D, Scal& S, bool late) {  // src/info.jl:135-190
  const DevSettings& st = *D.st;
  S.status = ST_UNSOLVED;
  check_convergence(S, st.tol_gap_abs, st.tol_gap_rel, st.tol_feas, st.tol_infeas_abs, st.tol_infeas_rel,
                    st.tol_ktratio, ST_SOLVED, ST_PRIMAL_INFEASIBLE, ST_DUAL_INFEASIBLE);
  if (S.status == ST_UNSOLVED && S.iter > 1 && (S.res_dual > S.prev_res_dual || S.res_primal > S.prev_res_primal)) {
    if (S.ktratio < 100 * kDEps && (S.prev_gap_abs < st.tol_gap_abs || S.prev_gap_rel < st.tol_gap_rel))
      S.status = ST_INSUFFICIENT_PROGRESS;
    if (S.ktratio < 1) {
      if ((S.res_dual > 100 * st.tol_feas && S.res_dual > 100 * S.prev_res_dual) ||
          (S.res_primal > 100 * st.tol_feas && S.res_primal > 100 * S.prev_res_primal))
        S.status = ST_INSUFFICIENT_PROGRESS;
    }
  }
  if (S.status == ST_UNSOLVED) {
    if (st.max_iter == S.info_iterations) S.status = ST_MAX_ITERATIONS;
    else if (late) S.status = ST_MAX_TIME;
  }
  return S.status != ST_UNSOLVED;
}
// seconds since the solve's first pass, the same on every thread (a
// reduction: the workgroups of a grid launch read their clocks apart)
CLH_D double slot_seconds(const Blk& B, const Scal& S) {
  return blk_max(B, (double)(clh_clock() - S.clock0)) * 1e-8;
}

CLH_D void record_trace(const Blk& B, const SD& D, Scal& S) {
  if (D.trace && B.tid == 0 && S.trace_len < D.st->max_iter + 2) {
    double* r = D.trace + (size_t)slot_lane() * D.trace_lstride + 8 * S.trace_len;
    r[0] = S.cost_primal; r[1] = S.cost_dual; r[2] = dmin(S.gap_abs, S.gap_rel);
    r[3] = S.res_primal; r[4] = S.res_dual; r[5] = S.ktratio; r[6] = S.mu;
    r[7] = S.info_iterations > 0 ? S.alpha : dnan();
  }
  S.trace_len++;
}

CLH_D void copy_vars(const Blk& B, const SD& D, VD dx, VD ds, VD dz, VCD sx, VCD ss, VCD sz) {
  for (int i = B.tid; i < D.n; i += B.nthr) dx[i] = sx[i];
  for (int i = B.tid; i < D.m; i += B.nthr) { ds[i] = ss[i]; dz[i] = sz[i]; }
  CLH_SYNC();
}

// whether the step (dx, ds, dz, dtau, dkappa) is finite
CLH_D bool step_finite(const Blk& B, const SD& D, const Scal& S) {
  double bad = finite(S.dtau) && finite(S.dkappa) ? 0.0 : 1.0;
  for (int i = B.tid; i < D.n; i += B.nthr)
    if (!finite(D.dx[i])) bad = 1;
  for (int i = B.tid; i < D.m; i += B.nthr)
    if (!finite(D.ds[i]) || !finite(D.dz[i])) bad = 1;
  return blk_max(B, bad) == 0.0;
}

// step length (src/solver.jl:449-511, src/variables.jl:14-43)
CLH_D double get_step_length(const Blk& B, const SD& D, Scal& S, bool combined) {
  PhaseTimer pt(B, lane_prof(D), 6);
  const DevSettings& st = *D.st;
  double at = S.dtau < 0 ? -S.tau / S.dtau : kDMax;
  double ak = S.dkappa < 0 ? -S.kappa / S.dkappa : kDMax;
  double a = dmin(dmin(at, ak), 1.0);
  a = cones_step_length(B, D, D.dz, D.ds, D.z, D.s, a);
  if (combined) a *= st.max_step_fraction;
  if (!D.symmetric && combined && S.scaling == 1) {
    const double cc = D.degree + 1;
    for (int j = 0; j < 50; ++j) {
      double ct = S.tau + a * S.dtau, ck = S.kappa + a * S.dkappa;
      double t = 0;
      for (int i = B.tid; i < D.m; i += B.nthr) t += (D.z[i] + a * D.dz[i]) * (D.s[i] + a * D.ds[i]);
      double sz = blk_sum(B, t);
      double mu = (sz + ct * ck) / cc;
      double bar = cc * logsafe(mu) - logsafe(ct) - logsafe(ck);
      bar += cones_barrier(B, D, D.z, D.s, D.dz, D.ds, a);
      if (bar < 1) return a;
      a = st.linesearch_backtrack_step * a;
    }
  }
  return a;
}

CLH_D void shift_to_interior(const Blk& B, const SD& D, VD z, bool primal) {  // src/variables.jl:276-304
  double mm, pm;
  cones_margins(B, D, z, mm, pm);
  double target = dmax(1.0, 0.1 * pm / D.degree);
  if (mm <= 0) {
    cones_unit_shift(B, D, z, -mm, primal);
    cones_unit_shift(B, D, z, target, primal);
  } else if (mm < target) {
    cones_unit_shift(B, D, z, target - mm, primal);
  } else {
    cones_unit_shift(B, D, z, 0.0, primal);
  }
}

// solver_default_start! (src/solver.jl:425-446), kkt_solve_initial_point! (src/kktsystem.jl:84-128)
CLH_D void default_start(const Blk& B, const SD& D, Scal& S) {
  const int n = D.n, m = D.m;
  if (D.symmetric) {
    cones_identity_scaling(B, D);
    kkt_update(B, D, S);
    VD xs;
    if (D.nnzP == 0) {
      for (int i = B.tid; i < n; i += B.nthr) D.workx[i] = 0;
      CLH_SYNC();
      set_rhs(B, D, D.workx, D.b);
      bool ok = kkt_solve_ir(B, D, S, &xs);
      if (ok) get_lhs(B, D, xs, D.x, D.s);
      for (int i = B.tid; i < m; i += B.nthr) D.s[i] = -D.s[i];
      for (int i = B.tid; i < n; i += B.nthr) D.workx[i] = -D.q[i];
      for (int i = B.tid; i < m; i += B.nthr) D.workz[i] = 0;
      CLH_SYNC();
      if (ok) {
        set_rhs(B, D, D.workx, D.workz);
        if (kkt_solve_ir(B, D, S, &xs)) get_lhs(B, D, xs, nullptr, D.z);
      }
    } else {
      for (int i = B.tid; i < n; i += B.nthr) D.workx[i] = -D.q[i];
      CLH_SYNC();
      set_rhs(B, D, D.workx, D.b);
      if (kkt_solve_ir(B, D, S, &xs)) get_lhs(B, D, xs, D.x, D.z);
      for (int i = B.tid; i < m; i += B.nthr) D.s[i] = -D.z[i];
      CLH_SYNC();
    }
    shift_to_interior(B, D, D.s, true);
    shift_to_interior(B, D, D.z, false);
  } else {
    cones_unit_init(B, D, D.z, D.s);
    for (int i = B.tid; i < n; i += B.nthr) D.x[i] = 0;
    CLH_SYNC();
  }
  S.tau = 1;
  S.kappa = 1;
}

// info_post_process! + solution_post_process! + variables_unscale!
CLH_D void finish(const Blk& B, const SD& D, Scal& S) {
  const DevSettings& st = *D.st;
  if (S.status == ST_NUMERICAL_ERROR || S.status == ST_INSUFFICIENT_PROGRESS || S.status == ST_MAX_ITERATIONS ||
      S.status == ST_MAX_TIME)
    check_convergence(S, st.reduced_tol_gap_abs, st.reduced_tol_gap_rel, st.reduced_tol_feas,
                      st.reduced_tol_infeas_abs, st.reduced_tol_infeas_rel, st.reduced_tol_ktratio, ST_ALMOST_SOLVED,
                      ST_ALMOST_PRIMAL_INFEASIBLE, ST_ALMOST_DUAL_INFEASIBLE);
  bool infeas = S.status == ST_PRIMAL_INFEASIBLE || S.status == ST_DUAL_INFEASIBLE ||
                S.status == ST_ALMOST_PRIMAL_INFEASIBLE || S.status == ST_ALMOST_DUAL_INFEASIBLE;
  S.obj_val = infeas ? dnan() : S.cost_primal;
  S.obj_val_dual = infeas ? dnan() : S.cost_dual;
  double scaleinv = infeas ? 1 / S.kappa : 1 / S.tau;
  double cinv = 1 / S.c;
  const size_t lo = (size_t)slot_lane() * D.out_lstride;
  double *ox = D.out_x + lo, *os = D.out_s + lo, *oz = D.out_z + lo;
  for (int i = B.tid; i < D.n; i += B.nthr) ox[i] = D.x[i] * D.d[i] * scaleinv;
  for (int i = B.tid; i < D.m; i += B.nthr) {
    oz[i] = D.z[i] * D.e[i] * (scaleinv * cinv);
    os[i] = D.s[i] * D.einv[i] * scaleinv;
  }
  S.phase = PH_DONE;
  CLH_SYNC();
}

// One pass of the `while true` body of solve! (src/solver.jl:265-393).
CLH_D void ipm_pass(const Blk& B, const SD& D, Scal& S) {
  PhaseTimer pt(B, lane_prof(D), 7);
  const DevSettings& st = *D.st;
  const int n = D.n, m = D.m;
  residuals_update(B, D, S);
  double mu = (S.dot_sz + S.tau * S.kappa) / (D.degree + 1);
  S.mu = mu;
  S.info_iterations = S.iter;
  info_update(B, D, S);
  bool late = S.time_left < dinf() && slot_seconds(B, S) > S.time_left;
#if !defined(CLH_LANES)
  if (D.cancel_flag) {  // cancelled by the host: ends as a time limit does
    double c = 0.0;
    if (B.tid == 0) c = *(volatile int32_t*)D.cancel_flag != 0 ? 1.0 : 0.0;
    late = blk_max(B, c) > 0.0 || late;
  }
#endif
  bool isdone = check_termination(D, S, late);
  record_trace(B, D, S);
  if (isdone) {
    if (S.status != ST_INSUFFICIENT_PROGRESS) { finish(B, D, S); return; }
    // _strategy_checkpoint_insufficient_progress: roll back
    S.cost_primal = S.prev_cost_primal; S.cost_dual = S.prev_cost_dual;
    S.res_primal = S.prev_res_primal; S.res_dual = S.prev_res_dual;
    S.gap_abs = S.prev_gap_abs; S.gap_rel = S.prev_gap_rel;
    copy_vars(B, D, D.x, D.s, D.z, D.px, D.ps, D.pz);
    S.tau = S.ptau; S.kappa = S.pkappa;
    if (!D.symmetric && S.scaling == 0) { S.status = ST_UNSOLVED; S.scaling = 1; return; }
    finish(B, D, S);
    return;
  }
  if (!cones_update_scaling(B, D, D.s, D.z, mu, S.scaling)) {
    S.status = ST_NUMERICAL_ERROR;
    finish(B, D, S);
    return;
  }
  CLH_SYNC();
  S.iter++;
  bool ok;
  double alpha = 0;
  for (int attempt = 0;; ++attempt) {
  ok = kkt_update(B, D, S);
  // affine rhs (src/variables.jl:161-175)
  for (int i = B.tid; i < n; i += B.nthr) D.hx[i] = D.rx[i];
  for (int i = B.tid; i < m; i += B.nthr) D.hz[i] = D.rz[i];
  CLH_SYNC();
  cones_affine_ds(B, D, D.hs, D.s);
  S.rtau_rhs = S.rtau;
  S.rkappa_rhs = S.tau * S.kappa;
  ok = ok && kkt_solve(B, D, S, true);
  if (ok) {
    const double alpha_aff = get_step_length(B, D, S, false);
    double sigma = (1 - alpha_aff) * (1 - alpha_aff) * (1 - alpha_aff);
    S.sigma = sigma;
    double mm = S.iter > 1 ? 1.0 : alpha_aff;
    double sigmu = sigma * mu;
    for (int i = B.tid; i < n; i += B.nthr) D.hx[i] = (1 - sigma) * D.rx[i];
    S.rtau_rhs = (1 - sigma) * S.rtau;
    S.rkappa_rhs = -sigmu + mm * S.dtau * S.dkappa + S.tau * S.kappa;
    if (mm != 1)
      for (int i = B.tid; i < m; i += B.nthr) D.dz[i] *= mm;
    CLH_SYNC();
    cones_combined_ds_shift(B, D, D.hz, D.dz, D.ds, sigmu);
    for (int i = B.tid; i < m; i += B.nthr) {
      D.hs[i] += D.hz[i];
      D.hz[i] = (1 - sigma) * D.rz[i];
    }
    CLH_SYNC();
    ok = kkt_solve(B, D, S, false);
  }
  // with FP32 factors, a direction that is not finite (refinement can reach
  // its tolerance on an operator too ill-conditioned for FP32) is redone once
  // with FP64 factors; nothing before the update depends on the first attempt
  if (ok) alpha = get_step_length(B, D, S, true);
  if (!(ok && use_fp32(D, S) && attempt == 0 && (!step_finite(B, D, S) || !finite(alpha)))) break;
  S.prec64 = 1;
  }
  if (!ok) {  // _strategy_checkpoint_numerical_error
    S.alpha = 0;
    if (!D.symmetric && S.scaling == 0) { S.scaling = 1; return; }
    S.status = ST_NUMERICAL_ERROR;
    finish(B, D, S);
    return;
  }
  // _strategy_checkpoint_small_step
  if (!D.symmetric && S.scaling == 0 && alpha < st.min_switch_step_length) {
    S.scaling = 1; S.alpha = 0; return;
  } else if (alpha <= dmax(0.0, st.min_terminate_step_length)) {
    S.status = ST_INSUFFICIENT_PROGRESS;
    S.alpha = 0;
    finish(B, D, S);
    return;
  }
  S.alpha = alpha;
  // info_save_prev_iterate + variables_add_step!
  S.prev_cost_primal = S.cost_primal; S.prev_cost_dual = S.cost_dual;
  S.prev_res_primal = S.res_primal; S.prev_res_dual = S.res_dual;
  S.prev_gap_abs = S.gap_abs; S.prev_gap_rel = S.gap_rel;
  for (int i = B.tid; i < n; i += B.nthr) { D.px[i] = D.x[i]; D.x[i] += alpha * D.dx[i]; }
  for (int i = B.tid; i < m; i += B.nthr) {
    D.ps[i] = D.s[i]; D.pz[i] = D.z[i];
    D.s[i] += alpha * D.ds[i];
    D.z[i] += alpha * D.dz[i];
  }
  S.ptau = S.tau; S.pkappa = S.kappa;
  S.tau += alpha * S.dtau;
  S.kappa += alpha * S.dkappa;
  CLH_SYNC();
}

// the grid engine's abort word (the serial build reads it too, so the host's
// abort path can be tested without a GPU)
CLH_D bool launch_aborted(const SD& D) {
#if defined(CLH_GRID) && !defined(CLH_SERIAL)
  // one read of host memory per workgroup, shared through LDS (called by
  // every thread of the workgroup at once)
  __syncthreads();
  if (threadIdx.x == 0 && !g_grid_abort_seen && grid_aborted()) g_grid_abort_seen = 1;
  __syncthreads();
  return g_grid_abort_seen != 0;
#elif defined(CLH_GRID)
  return D.abort_flag && *(volatile int32_t*)D.abort_flag != 0;
#else
  return false;
#endif
}

// Advance one slot by at most `passes` loop passes; returns its phase.
// the slot's scalars from memory into this wavefront's copy (or the lane's own)
CLH_D void scal_load(const Blk& B, Scal& S, const SlotScalars* src) {
#if defined(CLH_SERIAL)
  S = *src;
#else
  static_assert(sizeof(SlotScalars) % 4 == 0, "copied in 32-bit words");
  const int32_t* w = reinterpret_cast<const int32_t*>(src);
  CLH_LDS int32_t* d = reinterpret_cast<CLH_LDS int32_t*>(&S);
  for (int k = B.lane; k < (int)(sizeof(SlotScalars) / 4); k += 64) d[k] = w[k];
  CLH_WSYNC_LDS();
#endif
}
CLH_D void scal_store(SlotScalars* dst, const Scal& S) {
#if defined(CLH_SERIAL)
  *dst = S;
#else
  const CLH_LDS int32_t* w = reinterpret_cast<const CLH_LDS int32_t*>(&S);
  int32_t* d = reinterpret_cast<int32_t*>(dst);
  for (int k = 0; k < (int)(sizeof(SlotScalars) / 4); ++k) d[k] = w[k];
#endif
}

CLH_D int run_slot(const Blk& B, const SD& D, int passes) {
  // an aborted launch does nothing (the host resets the grid's state)
  if (launch_aborted(D)) return PH_SETUP;
#if defined(CLH_SERIAL)
  SlotScalars S_own;
  Scal& S = S_own;
#else
  __shared__ SlotScalars S_wave[kMaxWaves];
  Scal& S = *(Scal*)&S_wave[B.wave];
#endif
#if !defined(CLH_LANES)
  if (D.inbox) {
    // a freshly admitted slot: its scalars and values come from the inbox
    scal_load(B, S, reinterpret_cast<const SlotScalars*>(D.inbox));
    const double* in = D.inbox + kInboxHeaderDoubles;
    for (int k = B.tid; k < D.nnzP; k += B.nthr) D.Px[k] = in[k];
    in += D.nnzP;
    for (int k = B.tid; k < D.nnzA; k += B.nthr) D.Ax[k] = in[k];
    in += D.nnzA;
    for (int k = B.tid; k < D.nnzA; k += B.nthr) D.Atx[k] = in[k];
    in += D.nnzA;
    for (int j = B.tid; j < D.n; j += B.nthr) D.q[j] = in[j];
    in += D.n;
    for (int i = B.tid; i < D.m; i += B.nthr) D.b[i] = in[i];
  } else
#endif
  {
    scal_load(B, S, D.sc + slot_lane());
  }
  CLH_SYNC();
  if (S.phase == PH_SETUP) {
    // the time limit counts from here: the device part of setup! (the
    // equilibration), then solve!
    if (S.time_left < dinf()) S.clock0 = (uint64_t)blk_min(B, (double)clh_clock());
    PhaseTimer pt(B, lane_prof(D), 0);
    setup_phase(B, D, S);
    S.phase = PH_START;
  }
  if (S.phase == PH_START) {
    S.iter = 0; S.sigma = 1; S.alpha = 0; S.mu = dinf();
    S.scaling = 0; S.status = ST_UNSOLVED; S.trace_len = 0;
    S.prev_cost_primal = S.prev_cost_dual = S.prev_res_primal = S.prev_res_dual = dinf();
    S.prev_gap_abs = S.prev_gap_rel = dinf();
    S.ir_steps = 0; S.kkt_solves = 0; S.prec64 = 0; S.n_fp32_factors = 0;
    default_start(B, D, S);
    S.phase = PH_ITER;
  }
  for (int p = 0; p < passes && S.phase == PH_ITER && !launch_aborted(D); ++p) ipm_pass(B, D, S);
  CLH_SYNC();
  if (B.tid == 0) {
    scal_store(D.sc + slot_lane(), S);
    if (S.phase == PH_DONE && D.out_sc) scal_store(D.out_sc + slot_lane(), S);
  }
  return S.phase;
}

}  // namespace CLH_ENGINE_NS
}  // namespace clh
