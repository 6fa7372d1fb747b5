// HIP grid engine: one large problem on the whole GPU.  The interior-point
// engine compiled with CLH_GRID runs over a cooperative grid of 256-thread
// workgroups: every loop over rows, cones and supernodes of a level is spread
// over all the grid's threads (and lane groups), the barriers and reductions
// span the grid.  Used by the single-problem solver for problems whose
// factorization would leave one workgroup far behind the chip.
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#define CLH_GRID 1
#include "clh_backend.h"
#include "clh_engine.h"

namespace clh {

constexpr int kGridThreads = 256;
#ifndef CLH_LDS_BYTES
#define CLH_LDS_BYTES (36 * 1024)
#endif
constexpr int kGridLds = CLH_LDS_BYTES;

__global__ __launch_bounds__(kGridThreads, 4) void grid_ipm_kernel(const SlotDesc* __restrict__ desc, int passes,
                                                                    int32_t* __restrict__ phase_out,
                                                                    double* __restrict__ gred) {
  using namespace grid;
  extern __shared__ double lds_area[];
  __shared__ SlotDesc sd;
  __shared__ double red[kGridThreads / 64];
  __shared__ unsigned char epoch[kGridThreads];
  {
    const int32_t* src = reinterpret_cast<const int32_t*>(desc);
    int32_t* dst = reinterpret_cast<int32_t*>(&sd);
    for (int i = threadIdx.x; i < (int)(sizeof(SlotDesc) / 4); i += blockDim.x) dst[i] = src[i];
  }
  epoch[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    g_grid_bar = reinterpret_cast<unsigned long long*>(gred + 2 * gridDim.x);
    g_grid_abort = desc->abort_flag;  // from memory: sd is still being filled
    g_grid_abort_seen = 0;
  }
  __syncthreads();
  Blk B;
  B.tid = blockIdx.x * blockDim.x + threadIdx.x;
  B.nthr = gridDim.x * blockDim.x;
  B.lane = threadIdx.x & 63;
  B.wave = threadIdx.x >> 6;
  B.nwaves = blockDim.x >> 6;
  B.red = red;
  B.G = sd.group;
  B.gl = B.tid % sd.group;
  B.grp = B.tid / sd.group;
  B.ngrp = B.nthr / sd.group;
  B.lds = reinterpret_cast<char*>(lds_area);
  B.lds_bytes = kGridLds;
  B.ltid = threadIdx.x;
  B.lnthr = blockDim.x;
  B.lgrp = threadIdx.x / sd.group;
  B.lngrp = blockDim.x / sd.group;
  B.gred = gred;
  B.wg = blockIdx.x;
  B.nwg = gridDim.x;
  B.epoch = epoch;
  const int ph = run_slot(B, sd, passes);
  if (B.tid == 0) phase_out[0] = ph;
}

// The LDL' plugin's operations on one large matrix over the same
// cooperative grid (the seam of the reference's cuDSS solver, which drives
// the whole GPU): refactor, solve, solve with refinement.
__global__ __launch_bounds__(kGridThreads, 4) void grid_ldl_kernel(const SlotDesc* __restrict__ desc,
                                                                    backend::LdlArgs a, double* __restrict__ gred) {
  using namespace grid;
  extern __shared__ double lds_area[];
  __shared__ SlotDesc sd;
  __shared__ double red[kGridThreads / 64];
  __shared__ unsigned char epoch[kGridThreads];
  {
    const int32_t* src = reinterpret_cast<const int32_t*>(desc);
    int32_t* dst = reinterpret_cast<int32_t*>(&sd);
    for (int i = threadIdx.x; i < (int)(sizeof(SlotDesc) / 4); i += blockDim.x) dst[i] = src[i];
  }
  epoch[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    g_grid_bar = reinterpret_cast<unsigned long long*>(gred + 2 * gridDim.x);
    g_grid_abort = desc->abort_flag;
    g_grid_abort_seen = 0;
  }
  __syncthreads();
  if (threadIdx.x == 0 && grid_aborted()) g_grid_abort_seen = 1;
  __syncthreads();
  if (g_grid_abort_seen) return;
  Blk B;
  B.tid = blockIdx.x * blockDim.x + threadIdx.x;
  B.nthr = gridDim.x * blockDim.x;
  B.lane = threadIdx.x & 63;
  B.wave = threadIdx.x >> 6;
  B.nwaves = blockDim.x >> 6;
  B.red = red;
  B.G = sd.group;
  B.gl = B.tid % sd.group;
  B.grp = B.tid / sd.group;
  B.ngrp = B.nthr / sd.group;
  B.lds = reinterpret_cast<char*>(lds_area);
  B.lds_bytes = kGridLds;
  B.ltid = threadIdx.x;
  B.lnthr = blockDim.x;
  B.lgrp = threadIdx.x / sd.group;
  B.lngrp = blockDim.x / sd.group;
  B.gred = gred;
  B.wg = blockIdx.x;
  B.nwg = gridDim.x;
  B.epoch = epoch;
  LdlOut* out = reinterpret_cast<LdlOut*>(a.out);
  if (a.op == 0) ldl_op_refactor(B, sd, a.vals, a.map, a.static_eps, out);
  else if (a.op == 1) ldl_op_solve(B, sd, a.x, a.b);
  else ldl_op_solve_refine(B, sd, a.vals, a.map, a.x, a.b, a.reltol, a.abstol, a.maxiter, a.stopratio, out);
}

namespace backend {
namespace {
void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace

// workgroups one cooperative launch can keep resident, at most `cap`
int grid_blocks(int cap) {
  int dev = 0, cus = 0, per_cu = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
  check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, grid_ipm_kernel, kGridThreads, kGridLds),
        "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  return std::max(1, std::min(cap, cus * std::max(1, per_cu)));
}
int grid_threads() { return kGridThreads; }
int grid_ldl_blocks(int cap) {
  int dev = 0, cus = 0, per_cu = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
  check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, grid_ldl_kernel, kGridThreads, kGridLds),
        "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  return std::max(1, std::min(cap, cus * std::max(1, per_cu)));
}
void launch_grid_ldl(const SlotDesc* desc_dev, const LdlArgs& la, double* gred, int nblocks, void* stream) {
  LdlArgs a = la;
  void* args[] = {(void*)&desc_dev, (void*)&a, (void*)&gred};
  check(hipLaunchCooperativeKernel((const void*)grid_ldl_kernel, dim3(nblocks), dim3(kGridThreads), args, kGridLds,
                                   (hipStream_t)stream),
        "grid_ldl_kernel cooperative launch");
}

void launch_grid(const SlotDesc* desc_dev, int passes, int32_t* phase_dev, double* gred, int nblocks, void* stream) {
  void* args[] = {(void*)&desc_dev, (void*)&passes, (void*)&phase_dev, (void*)&gred};
  check(hipLaunchCooperativeKernel((const void*)grid_ipm_kernel, dim3(nblocks), dim3(kGridThreads), args, kGridLds,
                                   (hipStream_t)stream),
        "grid_ipm_kernel cooperative launch");
}

}  // namespace backend
}  // namespace clh
