// HIP backend: the gfx950 kernels and the memory/stream operations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include "clh_backend.h"
#include "clh_engine.h"

namespace clh {

#ifndef CLH_BLOCK_THREADS
#define CLH_BLOCK_THREADS 256
#endif
constexpr int kBlockThreads = CLH_BLOCK_THREADS;  // wavefronts of 64 per slot x 64

// One workgroup per slot.  The slot descriptor is staged in LDS, so the
// thousands of pointer reads of an iteration hit LDS instead of memory; a
// dynamic LDS area of kLdsBytes holds the supernode panels being factored
// and the triangular-solve vector.
#ifndef CLH_LDS_BYTES
#define CLH_LDS_BYTES (36 * 1024)  // with the static LDS, four workgroups per CU
#endif
constexpr int kLdsBytes = CLH_LDS_BYTES;
// four workgroups of four wavefronts per CU (128 VGPRs): measured on the MPC
// stream, more resident slots hide more of each slot's memory latency
#ifndef CLH_MIN_WAVES_PER_EU
#define CLH_MIN_WAVES_PER_EU 4
#endif
__global__ __launch_bounds__(kBlockThreads, CLH_MIN_WAVES_PER_EU) void ipm_kernel(const SlotDesc* __restrict__ descs,
                                                                                  int passes,
                                                                                  int32_t* __restrict__ phase_out) {
  using namespace wg;
  extern __shared__ double lds_area[];
  __shared__ SlotDesc sd;
  __shared__ double red[kBlockThreads / 64];
  const int tid = threadIdx.x;
  {
    const int32_t* src = reinterpret_cast<const int32_t*>(descs + blockIdx.x);
    int32_t* dst = reinterpret_cast<int32_t*>(&sd);
    for (int i = tid; i < (int)(sizeof(SlotDesc) / 4); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  if (sd.n < 0) {
    if (tid == 0) phase_out[blockIdx.x] = -1;
    return;
  }
  Blk B;
  B.tid = tid;
  B.nthr = blockDim.x;
  B.lane = tid & 63;
  B.wave = tid >> 6;
  B.nwaves = blockDim.x >> 6;
  B.red = red;
  B.G = sd.group;
  B.gl = tid % sd.group;
  B.grp = tid / sd.group;
  B.ngrp = blockDim.x / sd.group;
  B.lds = reinterpret_cast<char*>(lds_area);
  B.lds_bytes = kLdsBytes;
  int ph = run_slot(B, sd, passes);
  if (tid == 0) phase_out[blockIdx.x] = ph;
}

// The device-resident request loop (backend::ResidentArgs): each workgroup
// takes published requests one after another and runs every pass of each in
// this launch, with the slot code of ipm_kernel.
__global__ __launch_bounds__(kBlockThreads, CLH_MIN_WAVES_PER_EU) void resident_kernel(backend::ResidentArgs a) {
  using namespace wg;
  extern __shared__ double lds_area[];
  __shared__ SlotDesc sd;
  __shared__ double red[kBlockThreads / 64];
  __shared__ uint64_t s_entry;
  __shared__ int s_go;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) {
      int go = 0;
      uint64_t e = 0, t0 = 0;
      for (;;) {
        if (__hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) break;
        const uint64_t h = __hip_atomic_load(a.head, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
        uint64_t t = __hip_atomic_load(a.tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t < h) {
          if (__hip_atomic_compare_exchange_strong(a.tail, &t, t + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)) {
            e = t;
            go = 1;
            break;
          }
          continue;  // another workgroup claimed it first
        }
        const uint64_t now = wall_clock64();
        if (t0 == 0) t0 = now;
        else if (now - t0 > a.idle_ticks) break;
        __builtin_amdgcn_s_sleep(8);
      }
      s_go = go;
      s_entry = e;
    }
    __syncthreads();
    if (!s_go) return;
    const uint64_t e = s_entry;
    {  // the entry's descriptor, published before head, word by word from host memory
      const uint32_t* src = reinterpret_cast<const uint32_t*>(a.ring + (e % (uint64_t)a.ring_size));
      uint32_t* dst = reinterpret_cast<uint32_t*>(&sd);
      for (int i = tid; i < (int)(sizeof(SlotDesc) / 4); i += blockDim.x)
        dst[i] = __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    __syncthreads();
    Blk B;
    B.tid = tid;
    B.nthr = blockDim.x;
    B.lane = tid & 63;
    B.wave = tid >> 6;
    B.nwaves = blockDim.x >> 6;
    B.red = red;
    B.G = sd.group;
    B.gl = tid % sd.group;
    B.grp = tid / sd.group;
    B.ngrp = blockDim.x / sd.group;
    B.lds = reinterpret_cast<char*>(lds_area);
    B.lds_bytes = kLdsBytes;
    // the first call reads the inbox; a solve needing more passes than one
    // call allows continues from the slot's own scalars
    for (int call = 0; call < 4; ++call) {
      const int ph = run_slot(B, sd, a.passes);
      __syncthreads();
      if (ph == PH_DONE) break;
      if (tid == 0) sd.inbox = nullptr;
      __syncthreads();
    }
    __threadfence_system();  // every thread's outputs reach host memory before the flag
    __syncthreads();
    if (tid == 0)
      __hip_atomic_store(a.done + (e % (uint64_t)a.ring_size), e + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    __syncthreads();  // s_go, s_entry and sd are rewritten by the next claim
  }
}

// the LDL' plugin operations on one standalone matrix
__global__ __launch_bounds__(kBlockThreads, CLH_MIN_WAVES_PER_EU) void ldl_kernel(const SlotDesc* __restrict__ desc,
                                                                                  backend::LdlArgs a) {
  using namespace wg;
  extern __shared__ double lds_area[];
  __shared__ SlotDesc sd;
  __shared__ double red[kBlockThreads / 64];
  const int tid = threadIdx.x;
  {
    const int32_t* src = reinterpret_cast<const int32_t*>(desc);
    int32_t* dst = reinterpret_cast<int32_t*>(&sd);
    for (int i = tid; i < (int)(sizeof(SlotDesc) / 4); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  Blk B;
  B.tid = tid;
  B.nthr = blockDim.x;
  B.lane = tid & 63;
  B.wave = tid >> 6;
  B.nwaves = blockDim.x >> 6;
  B.red = red;
  B.G = sd.group;
  B.gl = tid % sd.group;
  B.grp = tid / sd.group;
  B.ngrp = blockDim.x / sd.group;
  B.lds = reinterpret_cast<char*>(lds_area);
  B.lds_bytes = kLdsBytes;
  LdlOut* out = reinterpret_cast<LdlOut*>(a.out);
  if (a.op == 0) ldl_op_refactor(B, sd, a.vals, a.map, a.static_eps, out);
  else if (a.op == 1) ldl_op_solve(B, sd, a.x, a.b);
  else ldl_op_solve_refine(B, sd, a.vals, a.map, a.x, a.b, a.reltol, a.abstol, a.maxiter, a.stopratio, out);
}

// lane engine kernel (clh_lanes_hip.hip)
void launch_lane_kernel(const SlotDesc* batches, int nbatches, int passes, int32_t* phase_out, hipStream_t stream);

namespace backend {

static void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

const char* name() { return "hip-gfx950"; }
int block_threads() { return kBlockThreads; }

void* dev_alloc(size_t bytes) {
  void* p = nullptr;
  check(hipMalloc(&p, bytes ? bytes : 8), "hipMalloc");
  return p;
}
void dev_free(void* p) {
  if (p) (void)hipFree(p);
}
void* host_alloc(size_t bytes) {
  void* p = nullptr;
  check(hipHostMalloc(&p, bytes ? bytes : 8, hipHostMallocDefault), "hipHostMalloc");
  return p;
}
void* host_alloc_coherent(size_t bytes) {
  void* p = nullptr;
  check(hipHostMalloc(&p, bytes ? bytes : 8, hipHostMallocCoherent | hipHostMallocMapped), "hipHostMalloc");
  return p;
}
bool stream_idle(void* stream) {
  const hipError_t e = hipStreamQuery((hipStream_t)stream);
  if (e == hipErrorNotReady) return false;
  check(e, "hipStreamQuery");
  return true;
}
void host_free(void* p) {
  if (p) (void)hipHostFree(p);
}
void h2d(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes) check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream), "hipMemcpyAsync H2D");
}
void d2d(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes) check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream), "hipMemcpyAsync D2D");
}
void d2h(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes) check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream), "hipMemcpyAsync D2H");
}
void sync(void* stream) { check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize"); }
void* stream_create() {
  hipStream_t s;
  check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  return (void*)s;
}
void stream_destroy(void* s) {
  if (s) (void)hipStreamDestroy((hipStream_t)s);
}

void launch_ipm(const LaunchSet& ls, int passes, void* stream) {
  if (ls.nslots > 0) {
    hipLaunchKernelGGL(ipm_kernel, dim3(ls.nslots), dim3(kBlockThreads), kLdsBytes, (hipStream_t)stream, ls.slots,
                       passes, ls.slot_phase);
    check(hipGetLastError(), "ipm_kernel launch");
  }
  if (ls.nbatches > 0) {
    launch_lane_kernel(ls.batches, ls.nbatches, passes, ls.lane_phase, (hipStream_t)stream);
    check(hipGetLastError(), "lane_kernel launch");
  }
}

int resident_blocks() {
  int dev = 0, cus = 0, per_cu = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
  check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, resident_kernel, kBlockThreads, kLdsBytes),
        "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  return std::max(1, cus * std::max(1, per_cu));
}
struct LaunchTimer {
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};
void launch_resident(const ResidentArgs& a, int nblocks, void* stream, void** timer) {
  LaunchTimer* t = (LaunchTimer*)*timer;
  if (!t) {
    t = new LaunchTimer;
    check(hipEventCreate(&t->ev0), "hipEventCreate");
    check(hipEventCreate(&t->ev1), "hipEventCreate");
    *timer = t;
  }
  check(hipEventRecord(t->ev0, (hipStream_t)stream), "hipEventRecord");
  hipLaunchKernelGGL(resident_kernel, dim3(nblocks), dim3(kBlockThreads), kLdsBytes, (hipStream_t)stream, a);
  check(hipGetLastError(), "resident_kernel launch");
  check(hipEventRecord(t->ev1, (hipStream_t)stream), "hipEventRecord");
}
double resident_seconds(void* timer) {
  LaunchTimer* t = (LaunchTimer*)timer;
  float ms = 0;
  if (!t || hipEventElapsedTime(&ms, t->ev0, t->ev1) != hipSuccess) return 0.0;
  return ms * 1e-3;
}
void timer_destroy(void* timer) {
  LaunchTimer* t = (LaunchTimer*)timer;
  if (!t) return;
  (void)hipEventDestroy(t->ev0);
  (void)hipEventDestroy(t->ev1);
  delete t;
}

void launch_ldl(const SlotDesc* desc_dev, const LdlArgs& a, void* stream) {
  hipLaunchKernelGGL(ldl_kernel, dim3(1), dim3(kBlockThreads), kLdsBytes, (hipStream_t)stream, desc_dev, a);
  check(hipGetLastError(), "ldl_kernel launch");
}

struct GraphCache {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
};

// kernels + phase read-back captured once, then replayed; the arguments are
// fixed and the tables they point to are rewritten between replays
void launch_ipm_graph(void** cache, const LaunchSet& ls, int passes, const int32_t* phase_dev, int32_t* phase_host,
                      size_t nphase, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  GraphCache* g = (GraphCache*)*cache;
  if (!g) {
    g = new GraphCache;
    check(hipEventCreate(&g->ev0), "hipEventCreate");
    check(hipEventCreate(&g->ev1), "hipEventCreate");
    *cache = g;
  }
  if (s == nullptr) {  // capture needs a non-default stream
    check(hipEventRecord(g->ev0, s), "hipEventRecord");
    launch_ipm(ls, passes, stream);
    check(hipEventRecord(g->ev1, s), "hipEventRecord");
    d2h(phase_host, phase_dev, sizeof(int32_t) * nphase, stream);
    g->timed = true;
    return;
  }
  if (!g->exec) {
    check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    launch_ipm(ls, passes, stream);
    d2h(phase_host, phase_dev, sizeof(int32_t) * nphase, stream);
    check(hipStreamEndCapture(s, &g->graph), "hipStreamEndCapture");
    check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  }
  check(hipEventRecord(g->ev0, s), "hipEventRecord");
  check(hipGraphLaunch(g->exec, s), "hipGraphLaunch");
  check(hipEventRecord(g->ev1, s), "hipEventRecord");
  g->timed = true;
}

double graph_last_seconds(void* cache) {
  GraphCache* g = (GraphCache*)cache;
  if (!g || !g->timed) return 0.0;
  float ms = 0;
  if (hipEventElapsedTime(&ms, g->ev0, g->ev1) != hipSuccess) return 0.0;
  return ms * 1e-3;
}

void graph_destroy(void* cache) {
  GraphCache* g = (GraphCache*)cache;
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->ev0) (void)hipEventDestroy(g->ev0);
  if (g->ev1) (void)hipEventDestroy(g->ev1);
  delete g;
}

int device_info(char* nm, size_t len, int32_t* num_cu) {
  hipDeviceProp_t p;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) return 1;
  if (nm && len) {
    std::strncpy(nm, p.gcnArchName, len - 1);
    nm[len - 1] = 0;
  }
  if (num_cu) *num_cu = p.multiProcessorCount;
  return 0;
}

}  // namespace backend
}  // namespace clh
