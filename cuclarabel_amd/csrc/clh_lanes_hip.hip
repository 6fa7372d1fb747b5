// HIP lane engine: one 64-lane workgroup per batch of same-shape slots, one
// slot per lane.  Each lane runs the interior-point method serially for its
// own problem; the batch's structure arrays are shared (uniform indices) and
// its values are lane-interleaved (coalesced accesses).
#define CLH_LANES 1
#include <hip/hip_runtime.h>

#include "clh_engine.h"

namespace clh {

__global__ __launch_bounds__(kBatchLanes) void lane_kernel(const SlotDesc* __restrict__ descs, int passes,
                                                           int32_t* __restrict__ phase_out) {
  using namespace lanes;
  const SD& D = *reinterpret_cast<const SD*>(descs + blockIdx.x);
  int32_t* ph = phase_out + (size_t)blockIdx.x * kBatchLanes + threadIdx.x;
  if (D.n < 0) {
    *ph = -1;
    return;
  }
  const Blk B{0, 1, 0, 0, 1, nullptr, 1, 0, 0, 1, nullptr, 0};
  *ph = run_slot(B, D, passes);
}

void launch_lane_kernel(const SlotDesc* batches, int nbatches, int passes, int32_t* phase_out, hipStream_t stream) {
  hipLaunchKernelGGL(lane_kernel, dim3(nbatches), dim3(kBatchLanes), 0, stream, batches, passes, phase_out);
}

}  // namespace clh
