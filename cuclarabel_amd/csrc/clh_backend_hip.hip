// HIP backend: the gfx950 kernels and the memory/stream operations.
#include <hip/hip_runtime.h>

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

// the LDL' plugin operations on one standalone matrix
// one composite-cone operation: thread g of the grid takes rows and cones
// g, g + G, ... (G the grid's threads); per-workgroup partials of reductions
constexpr int kConeThreads = 256;
__global__ __launch_bounds__(kConeThreads) void cone_kernel(const SlotDesc* __restrict__ desc, backend::ConeArgs a) {
  using namespace wg;
  __shared__ SlotDesc sd;
  __shared__ double red[kConeThreads / 64];
  {
    const int32_t* src = reinterpret_cast<const int32_t*>(desc);
    int32_t* dst = reinterpret_cast<int32_t*>(&sd);
    for (int i = threadIdx.x; i < (int)(sizeof(SlotDesc) / 4); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  Blk B;
  B.tid = blockIdx.x * blockDim.x + threadIdx.x;
  B.nthr = gridDim.x * blockDim.x;
  B.lane = threadIdx.x & 63;
  B.wave = threadIdx.x >> 6;
  B.nwaves = blockDim.x >> 6;
  B.red = red;
  B.G = 1; B.gl = 0; B.grp = B.tid; B.ngrp = B.nthr;
  B.lds = nullptr;
  B.lds_bytes = 0;
  double p0 = 0, p1 = 0;
  bool reduce = true;
  switch (a.op) {
    case backend::CO_UPDATE_SCALING: p0 = cones_update_scaling_part(B, sd, a.a, a.b, a.s0, a.flag); break;
    case backend::CO_STEP_SYMMETRIC: p0 = cones_step_symmetric_part(B, sd, a.a, a.b, a.c, a.d, a.s0); break;
    case backend::CO_STEP_NONSYMMETRIC:
      p0 = cones_step_nonsym_part(B, sd, a.a, a.b, a.c, a.d, dmin(a.result[0], sd.st->max_step_fraction));
      break;
    case backend::CO_BARRIER: p0 = cones_barrier_part(B, sd, a.a, a.b, a.c, a.d, a.s0); break;
    case backend::CO_MARGINS: cones_margins_part(B, sd, a.a, p0, p1); break;
    default:
      reduce = false;
      switch (a.op) {
        case backend::CO_IDENTITY_SCALING: cones_identity_scaling(B, sd); break;
        case backend::CO_GET_HS: cones_get_Hs(B, sd, a.x, a.hs_row, a.hs_soc, a.hs_nsym, a.hs_psd); break;
        case backend::CO_MUL_HS: cones_mul_Hs(B, sd, a.x, a.a); break;
        case backend::CO_AFFINE_DS: cones_affine_ds(B, sd, a.x, a.a); break;
        case backend::CO_COMBINED_DS_SHIFT: cones_combined_ds_shift(B, sd, a.x, a.y, a.z, a.s0); break;
        case backend::CO_DS_FROM_DZ_OFFSET: cones_ds_from_dz_offset(B, sd, a.x, a.a, a.b); break;
        case backend::CO_UNIT_SHIFT: cones_unit_shift(B, sd, a.x, a.s0, a.flag != 0); break;
        case backend::CO_UNIT_INIT: cones_unit_init(B, sd, a.x, a.y); break;
        case backend::CO_RECTIFY: cones_rectify(B, sd, a.x, a.a); break;
        default: break;
      }
  }
  if (reduce) {
    const bool mins = a.op != backend::CO_BARRIER;
    const double r0 = mins ? blk_min(B, p0) : blk_sum(B, p0);
    const double r1 = a.op == backend::CO_MARGINS ? blk_sum(B, p1) : 0.0;
    if (threadIdx.x == 0) {
      a.partial[2 * blockIdx.x] = r0;
      a.partial[2 * blockIdx.x + 1] = r1;
    }
  }
}

// result[0] (and result[1]) from the partials of nb workgroups, in workgroup order
__global__ __launch_bounds__(kConeThreads) void cone_reduce_kernel(const double* __restrict__ partial, int nb,
                                                                    int mode, double* __restrict__ result) {
  using namespace wg;
  __shared__ double red[kConeThreads / 64];
  Blk B;
  B.tid = threadIdx.x;
  B.nthr = blockDim.x;
  B.lane = threadIdx.x & 63;
  B.wave = threadIdx.x >> 6;
  B.nwaves = blockDim.x >> 6;
  B.red = red;
  double a0 = mode == 1 ? 0.0 : dinf(), a1 = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    a0 = mode == 1 ? a0 + partial[2 * i] : dmin(a0, partial[2 * i]);
    a1 += partial[2 * i + 1];
  }
  const double r0 = mode == 1 ? blk_sum(B, a0) : blk_min(B, a0);
  const double r1 = blk_sum(B, a1);
  if (threadIdx.x == 0) {
    result[0] = r0;
    if (mode == 2) result[1] = r1;
  }
}

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

void launch_ldl(const SlotDesc* desc_dev, const LdlArgs& a, void* stream) {
  hipLaunchKernelGGL(ldl_kernel, dim3(1), dim3(kBlockThreads), kLdsBytes, (hipStream_t)stream, desc_dev, a);
  check(hipGetLastError(), "ldl_kernel launch");
}

int cone_blocks(int work) { return std::max(1, std::min(4096, (work + kConeThreads - 1) / kConeThreads)); }

void launch_cones(const SlotDesc* desc_dev, const ConeArgs& a, int nblocks, void* stream) {
  hipLaunchKernelGGL(cone_kernel, dim3(nblocks), dim3(kConeThreads), 0, (hipStream_t)stream, desc_dev, a);
  check(hipGetLastError(), "cone_kernel launch");
}

void launch_cones_reduce(const ConeArgs& a, int nblocks, int mode, void* stream) {
  hipLaunchKernelGGL(cone_reduce_kernel, dim3(1), dim3(kConeThreads), 0, (hipStream_t)stream, a.partial, nblocks,
                     mode, a.result);
  check(hipGetLastError(), "cone_reduce_kernel launch");
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
