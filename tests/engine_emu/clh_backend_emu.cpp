// Serial CPU backend for the unit tests of the engine logic: the device
// engine compiled as plain C++ (one thread per workgroup, one-lane waves).
// TEST INFRASTRUCTURE ONLY: the cuclarabel_amd package never loads it.
#define CLH_EMULATE 1
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "clh_backend.h"
#include "clh_engine.h"

namespace clh {
// lane engine, serial (clh_backend_emu_lanes.cpp)
void emu_run_lanes(const SlotDesc* batches, int nbatches, int passes, int32_t* lane_phase);

namespace backend {

const char* name() { return "cpu-emulation"; }
int block_threads() { return 1; }
void* dev_alloc(size_t bytes) {
  void* p = std::calloc(1, bytes ? bytes : 8);
  if (!p) throw std::runtime_error("out of memory");
  return p;
}
void dev_free(void* p) { std::free(p); }
void* host_alloc(size_t bytes) { return dev_alloc(bytes); }
void host_free(void* p) { std::free(p); }
void h2d(void* dst, const void* src, size_t bytes, void*) { std::memcpy(dst, src, bytes); }
void d2h(void* dst, const void* src, size_t bytes, void*) { std::memcpy(dst, src, bytes); }
void sync(void*) {}
void* stream_create() { return nullptr; }
void stream_destroy(void*) {}

// the workgroup's LDS area (CLH_EMU_LDS_BYTES, default the HIP kernel's 36 KiB;
// 0 runs every panel and solve vector in memory)
static int emu_lds_bytes() {
  static int v = [] {
    const char* e = std::getenv("CLH_EMU_LDS_BYTES");
    return e ? std::atoi(e) : 36 * 1024;
  }();
  return v;
}

void launch_ipm(const LaunchSet& ls, int passes, void*) {
  double red[4];
  static double lds_area[60 * 1024 / 8];
  const int lds_bytes = std::min<int>(emu_lds_bytes(), (int)sizeof(lds_area));
  for (int i = 0; i < ls.nslots; ++i) {
    SlotDesc sd = ls.slots[i];
    if (sd.n < 0) { ls.slot_phase[i] = -1; continue; }
    wg::Blk B{0, 1, 0, 0, 1, red, 1, 0, 0, 1, lds_bytes ? (char*)lds_area : nullptr, lds_bytes};
    ls.slot_phase[i] = wg::run_slot(B, sd, passes);
  }
  if (ls.nbatches > 0) emu_run_lanes(ls.batches, ls.nbatches, passes, ls.lane_phase);
}
void launch_ipm_graph(void**, const LaunchSet& ls, int passes, const int32_t* phase_dev, int32_t* phase_host,
                      size_t nphase, void* stream) {
  launch_ipm(ls, passes, stream);
  std::memcpy(phase_host, phase_dev, sizeof(int32_t) * nphase);
}
void launch_ldl(const SlotDesc* desc, const LdlArgs& a, void*) {
  double red[4];
  static double lds_area[60 * 1024 / 8];
  const int lds_bytes = std::min<int>(emu_lds_bytes(), (int)sizeof(lds_area));
  SlotDesc sd = *desc;
  wg::Blk B{0, 1, 0, 0, 1, red, 1, 0, 0, 1, lds_bytes ? (char*)lds_area : nullptr, lds_bytes};
  wg::LdlOut* out = reinterpret_cast<wg::LdlOut*>(a.out);
  if (a.op == 0) wg::ldl_op_refactor(B, sd, a.vals, a.map, a.static_eps, out);
  else if (a.op == 1) wg::ldl_op_solve(B, sd, a.x, a.b);
  else wg::ldl_op_solve_refine(B, sd, a.vals, a.map, a.x, a.b, a.reltol, a.abstol, a.maxiter, a.stopratio, out);
}
void graph_destroy(void*) {}
double graph_last_seconds(void*) { return 0.0; }
int device_info(char* nm, size_t len, int32_t* num_cu) {
  if (nm && len) { std::strncpy(nm, "cpu-emulation", len - 1); nm[len - 1] = 0; }
  if (num_cu) *num_cu = 1;
  return 0;
}

}  // namespace backend
}  // namespace clh
