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

int emu_run_grid(const SlotDesc* desc, int passes, int lds_bytes, char* lds);
int emu_run_grid_ldl(const SlotDesc* desc, const backend::LdlArgs& a, int lds_bytes, char* lds);

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
void* host_alloc_coherent(size_t bytes) { return dev_alloc(bytes); }
void host_free(void* p) { std::free(p); }
bool stream_idle(void*) { return true; }
void h2d(void* dst, const void* src, size_t bytes, void*) { std::memcpy(dst, src, bytes); }
void d2h(void* dst, const void* src, size_t bytes, void*) { std::memcpy(dst, src, bytes); }
void d2d(void* dst, const void* src, size_t bytes, void*) { std::memcpy(dst, src, bytes); }
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
int cone_blocks(int) { return 1; }
// the grid engine's code run by one thread; the arena is planned for a grid
// of 64 workgroups of 256 threads, so its group scratch is as large as on a
// small GPU grid (the wide supernodes then split their update lists)
int grid_blocks(int cap) { return std::min(cap, 64); }
int grid_threads() { return 256; }
void launch_grid(const SlotDesc* desc, int passes, int32_t* phase, double*, int, void*) {
  static double lds_area[60 * 1024 / 8];
  const int lds_bytes = std::min<int>(emu_lds_bytes(), (int)sizeof(lds_area));
  *phase = emu_run_grid(desc, passes, lds_bytes, lds_bytes ? (char*)lds_area : nullptr);
}
// the device-resident loop, serially: every published entry is taken in
// order and solved to completion, then the "launch" ends
int resident_blocks() { return 1; }
double resident_seconds(void*) { return 0.0; }
void timer_destroy(void*) {}
void launch_resident(const ResidentArgs& a, int, void*, void**) {
  static double lds_area[60 * 1024 / 8];
  const int lds_bytes = std::min<int>(emu_lds_bytes(), (int)sizeof(lds_area));
  double red[4];
  while (!*(volatile const uint32_t*)a.stop && *a.tail < *(volatile const uint64_t*)a.head) {
    const uint64_t e = (*a.tail)++;
    SlotDesc sd = a.ring[e % (uint64_t)a.ring_size];
    wg::Blk B{0, 1, 0, 0, 1, red, 1, 0, 0, 1, lds_bytes ? (char*)lds_area : nullptr, lds_bytes};
    for (int call = 0; call < 4; ++call) {
      if (wg::run_slot(B, sd, a.passes) == PH_DONE) break;
      sd.inbox = nullptr;
    }
    a.done[e % (uint64_t)a.ring_size] = e + 1;
  }
}
int grid_ldl_blocks(int cap) { return std::min(cap, 64); }
void launch_grid_ldl(const SlotDesc* desc, const LdlArgs& a, double*, int, void*) {
  static double lds_area[60 * 1024 / 8];
  const int lds_bytes = std::min<int>(emu_lds_bytes(), (int)sizeof(lds_area));
  emu_run_grid_ldl(desc, a, lds_bytes, lds_bytes ? (char*)lds_area : nullptr);
}
void launch_cones(const SlotDesc* desc, const ConeArgs& a, int, void*) {
  double red[4];
  SlotDesc sd = *desc;
  wg::Blk B{0, 1, 0, 0, 1, red, 1, 0, 0, 1, nullptr, 0};
  double p0 = 0, p1 = 0;
  switch (a.op) {
    case CO_UPDATE_SCALING: p0 = wg::cones_update_scaling_part(B, sd, a.a, a.b, a.s0, a.flag); break;
    case CO_STEP_SYMMETRIC: p0 = wg::cones_step_symmetric_part(B, sd, a.a, a.b, a.c, a.d, a.s0); break;
    case CO_STEP_NONSYMMETRIC:
      p0 = wg::cones_step_nonsym_part(B, sd, a.a, a.b, a.c, a.d, std::min(a.result[0], sd.st->max_step_fraction));
      break;
    case CO_BARRIER: p0 = wg::cones_barrier_part(B, sd, a.a, a.b, a.c, a.d, a.s0); break;
    case CO_MARGINS: wg::cones_margins_part(B, sd, a.a, p0, p1); break;
    case CO_IDENTITY_SCALING: wg::cones_identity_scaling(B, sd); break;
    case CO_GET_HS: wg::cones_get_Hs(B, sd, a.x, a.hs_row, a.hs_soc, a.hs_nsym, a.hs_psd); break;
    case CO_MUL_HS: wg::cones_mul_Hs(B, sd, a.x, a.a); break;
    case CO_AFFINE_DS: wg::cones_affine_ds(B, sd, a.x, a.a); break;
    case CO_COMBINED_DS_SHIFT: wg::cones_combined_ds_shift(B, sd, a.x, a.y, a.z, a.s0); break;
    case CO_DS_FROM_DZ_OFFSET: wg::cones_ds_from_dz_offset(B, sd, a.x, a.a, a.b); break;
    case CO_UNIT_SHIFT: wg::cones_unit_shift(B, sd, a.x, a.s0, a.flag != 0); break;
    case CO_UNIT_INIT: wg::cones_unit_init(B, sd, a.x, a.y); break;
    case CO_RECTIFY: wg::cones_rectify(B, sd, a.x, a.a); break;
    default: break;
  }
  if (a.partial) { a.partial[0] = p0; a.partial[1] = p1; }
}
double* cone_scratch(int* max_blocks) {
  static double scratch[2 * 2048 + 8];
  if (max_blocks) *max_blocks = 2048;
  return scratch;
}
void launch_cones_reduce(const ConeArgs& a, int, int mode, void*) {
  a.result[0] = a.partial[0];
  if (mode == 2) a.result[1] = a.partial[1];
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
