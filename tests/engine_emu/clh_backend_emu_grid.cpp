// Serial CPU build of the grid engine: the grid engine's own code paths (the
// top levels of the tree factored and solved one supernode at a time by the
// whole grid) run with one thread, so their arithmetic can be checked
// against the oracle without a GPU.
// TEST INFRASTRUCTURE ONLY: the cuclarabel_amd package never loads it.
#define CLH_EMULATE 1
#define CLH_GRID 1
// wide panels from 2 columns up take the blocked whole-grid dense
// factorization (128 on the GPU), so the small test problems reach it
#define CLH_DENSE_GRID_MIN_COLS 2
// in blocks of 4 columns and super-blocks of 12 (32 and 256 on the GPU), so
// the in-super-block and the trailing rank-12 updates both run
#define CLH_DENSE_NB 4
#define CLH_DENSE_SB 12
// and from 8 columns up their updates as tiles (512 on the GPU)
#define CLH_TILE_MIN_COLS 100000
#include "clh_backend.h"
#include "clh_engine.h"

namespace clh {

int emu_run_grid(const SlotDesc* desc, int passes, int lds_bytes, char* lds) {
  double red[4];
  SlotDesc sd = *desc;
  grid::Blk B{0, 1, 0, 0, 1, red, 1, 0, 0, 1, lds, lds_bytes};
  return grid::run_slot(B, sd, passes);
}

int emu_run_grid_ldl(const SlotDesc* desc, const backend::LdlArgs& a, int lds_bytes, char* lds) {
  double red[4];
  SlotDesc sd = *desc;
  grid::Blk B{0, 1, 0, 0, 1, red, 1, 0, 0, 1, lds, lds_bytes};
  if (sd.abort_flag && *sd.abort_flag) return 1;
  grid::LdlOut* out = reinterpret_cast<grid::LdlOut*>(a.out);
  if (a.op == 0) grid::ldl_op_refactor(B, sd, a.vals, a.map, a.static_eps, out);
  else if (a.op == 1) grid::ldl_op_solve(B, sd, a.x, a.b);
  else grid::ldl_op_solve_refine(B, sd, a.vals, a.map, a.x, a.b, a.reltol, a.abstol, a.maxiter, a.stopratio, out);
  return 0;
}

}  // namespace clh
