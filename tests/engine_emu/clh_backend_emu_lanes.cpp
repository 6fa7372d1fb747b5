// Serial CPU build of the lane engine: the lanes of each batch run one after
// another, each with its lane index set, over the same lane-interleaved
// memory layout the HIP lane kernel uses.
// TEST INFRASTRUCTURE ONLY: the cuclarabel_amd package never loads it.
#define CLH_EMULATE 1
#define CLH_LANES 1
#include "clh_engine.h"

namespace clh {

int g_emu_lane = 0;

void emu_run_lanes(const SlotDesc* batches, int nbatches, int passes, int32_t* lane_phase) {
  for (int b = 0; b < nbatches; ++b) {
    const lanes::SD& D = *reinterpret_cast<const lanes::SD*>(batches + b);
    for (int l = 0; l < kBatchLanes; ++l) {
      g_emu_lane = l;
      int32_t& ph = lane_phase[b * kBatchLanes + l];
      if (D.n < 0) { ph = -1; continue; }
      const lanes::Blk B{0, 1, 0, 0, 1, nullptr, 1, 0, 0, 1, nullptr, 0};
      ph = lanes::run_slot(B, D, passes);
    }
  }
}

}  // namespace clh
