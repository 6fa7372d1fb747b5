// HIP composite-cone operations: each operation one grid of workgroups over
// the rows and cones of one composite cone (the clh_cones_* entry points).
// A translation unit of its own: the cone functions it calls are shared with
// the interior-point kernel, and compiled next to it they would be emitted
// out of line for both, with this kernel's register budget.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "clh_backend.h"
#include "clh_engine.h"

namespace clh {

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

// Scratch of the allocation-free cone operation: the reduction partials of
// up to kConeScratchBlocks workgroups, then the reduced scalars.  Each GPU
// holds its own copy until the process ends; its device address is looked up
// once per device (cone_scratch) and only that address is handed to kernels
// or copies.  Every call on one GPU shares it: calls that overlap race.
constexpr int kConeScratchBlocks = 2048;
static __device__ __align__(256) double scratch[2 * kConeScratchBlocks + 8];

namespace backend {
namespace {
void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace

double* cone_scratch(int* max_blocks) {
  static std::mutex mu;
  static std::vector<double*> ptr;  // by device
  int dev = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lock(mu);
  if ((int)ptr.size() <= dev) ptr.resize(dev + 1, nullptr);
  if (!ptr[dev]) {
    void* p = nullptr;
    check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    ptr[dev] = (double*)p;
  }
  if (max_blocks) *max_blocks = kConeScratchBlocks;
  return ptr[dev];
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

}  // namespace backend
}  // namespace clh
