// Serial CPU build of the grid engine: the grid engine's own code paths (the
// top levels of the tree factored and solved one supernode at a time by the
// whole grid) run with one thread, so their arithmetic can be checked
// against the oracle without a GPU.
// TEST INFRASTRUCTURE ONLY: the cuclarabel_amd package never loads it.
#define CLH_EMULATE 1
#define CLH_GRID 1
#include "clh_engine.h"

namespace clh {

int emu_run_grid(const SlotDesc* desc, int passes, int lds_bytes, char* lds) {
  double red[4];
  SlotDesc sd = *desc;
  grid::Blk B{0, 1, 0, 0, 1, red, 1, 0, 0, 1, lds, lds_bytes};
  return grid::run_slot(B, sd, passes);
}

}  // namespace clh
