// The small set of device operations the runtime needs.  The product
// library implements it with HIP (clh_backend_hip.hip); the CPU unit tests of
// the engine logic link a serial implementation instead.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "clh_types.h"

namespace clh {
namespace backend {

const char* name();
// allocation failures throw std::runtime_error
void* dev_alloc(size_t bytes);
void dev_free(void* p);
void* host_alloc(size_t bytes);  // page-locked where the backend has it
// page-locked and coherent: device atomics on it are seen by the host while a
// kernel runs, and host stores by the kernel (a launch's abort word)
void* host_alloc_coherent(size_t bytes);
void host_free(void* p);
// true when every operation queued on the stream has finished
bool stream_idle(void* stream);
void h2d(void* dst, const void* src, size_t bytes, void* stream);
void d2h(void* dst, const void* src, size_t bytes, void* stream);
void d2d(void* dst, const void* src, size_t bytes, void* stream);
void sync(void* stream);
void* stream_create();
void stream_destroy(void* stream);
// One launch set: the workgroup engine over slots[0..nslots) (one slot per
// workgroup) and the lane engine over batches[0..nbatches) (one batch of up to
// kBatchLanes same-shape slots per workgroup, one slot per lane).  A slot or
// batch with n < 0 is empty.  slot_phase[i] receives the phase of slot i and
// lane_phase[b * kBatchLanes + l] the phase of lane l of batch b (-1 if empty).
struct LaunchSet {
  const SlotDesc* slots;
  int nslots;
  int32_t* slot_phase;  // device
  const SlotDesc* batches;
  int nbatches;
  int32_t* lane_phase;  // device
};
// run `passes` IPM loop passes on every slot and lane of the set
void launch_ipm(const LaunchSet& ls, int passes, void* stream);
// the same followed by the read-back of phase_dev[0..nphase) into phase_host,
// captured once into a graph and replayed (fixed arguments per cache)
void launch_ipm_graph(void** graph_cache, const LaunchSet& ls, int passes, const int32_t* phase_dev,
                      int32_t* phase_host, size_t nphase, void* stream);
// The device-resident request loop: every workgroup of a persistent launch
// claims published ring entries (a compare-and-swap on the device's claim
// counter), solves each to completion and publishes its sequence number + 1
// in done[]; a workgroup that finds no work for idle_ticks (100 MHz) or sees
// stop set leaves, so the launch always ends.  head, stop, ring and done live
// in coherent host memory, written by the host between entries.
struct ResidentArgs {
  const uint64_t* head;   // entries published so far
  const uint32_t* stop;
  const SlotDesc* ring;   // ring_size descriptors, entry e at e % ring_size
  uint64_t* done;         // per ring position: the sequence number + 1 of its last finished entry
  uint64_t* tail;         // device memory: entries claimed so far
  int ring_size;
  int passes;             // loop passes per call of the slot (each call ends at PH_DONE or after them)
  uint64_t idle_ticks;
};
int resident_blocks();
// the launch, timed by events on its stream (*timer is created on first use;
// resident_seconds gives the last launch's device time once it has ended,
// timer_destroy releases it)
void launch_resident(const ResidentArgs& a, int nblocks, void* stream, void** timer);
double resident_seconds(void* timer);
void timer_destroy(void* timer);

// one LDL' plugin operation (op: 0 refactor, 1 solve, 2 solve with
// refinement) on the matrix of desc_dev, one workgroup; out_dev receives
// (ok, dynamically regularized pivots, refinement steps, final error)
struct LdlArgs {
  int op;
  const double* vals;    // device: the caller's CSR values
  const int32_t* map;    // device: CSR position of each triu entry
  double* x;             // device
  const double* b;       // device
  double static_eps, reltol, abstol, stopratio;
  int maxiter;
  double* out;           // device: 4 doubles
};
void launch_ldl(const SlotDesc* desc_dev, const LdlArgs& a, void* stream);

// one composite-cone operation over a grid of workgroups, each thread
// taking every (grid size)-th row and cone; reductions leave two partials per
// workgroup, combined in workgroup order by launch_cones_reduce
enum ConeOp {
  CO_UPDATE_SCALING, CO_IDENTITY_SCALING, CO_GET_HS, CO_MUL_HS, CO_AFFINE_DS, CO_COMBINED_DS_SHIFT,
  CO_DS_FROM_DZ_OFFSET, CO_STEP_SYMMETRIC, CO_STEP_NONSYMMETRIC, CO_BARRIER, CO_MARGINS, CO_UNIT_SHIFT,
  CO_UNIT_INIT, CO_RECTIFY
};
struct ConeArgs {
  int op;
  const double *a, *b, *c, *d;   // device inputs (meaning per op)
  double *x, *y, *z;             // device outputs
  double s0;                     // mu, sigma mu, alpha or alpha_max
  int flag;                      // dual scaling (update_scaling), primal cone (unit shift)
  double* partial;               // device: 2 per workgroup
  double* result;                // device: the reduced scalars (a nonsymmetric step reads its start here)
  const int64_t *hs_row, *hs_soc, *hs_nsym, *hs_psd;  // device: Hs block positions (get_Hs)
};
// reduce modes: 0 min of the first partials, 1 their sum, 2 min of the first
// and sum of the second
void launch_cones(const SlotDesc* desc_dev, const ConeArgs& a, int nblocks, void* stream);
void launch_cones_reduce(const ConeArgs& a, int nblocks, int mode, void* stream);
// the current device's static scratch for one cone operation's partials
// (2 per workgroup, up to *max_blocks workgroups) followed by 8 reduced
// scalars; the first call on a device looks its address up
double* cone_scratch(int* max_blocks);
int cone_blocks(int work);

// the grid engine (one problem over a cooperative grid): resident
// workgroups per launch (at most cap), threads per workgroup, and the launch;
// gred holds 2 * nblocks doubles of reduction partials
int grid_blocks(int cap);
int grid_threads();
// gred: 2 * nblocks doubles of reduction partials followed by kGridBarBytes
// of barrier counters, zeroed when allocated
void launch_grid(const SlotDesc* desc_dev, int passes, int32_t* phase_dev, double* gred, int nblocks, void* stream);
constexpr size_t kGridBarBytes = 4096;
// the LDL' plugin's operations on one matrix over a cooperative grid (same
// gred layout); resident workgroups of that kernel, at most cap
int grid_ldl_blocks(int cap);
void launch_grid_ldl(const SlotDesc* desc_dev, const LdlArgs& a, double* gred, int nblocks, void* stream);
void graph_destroy(void* graph_cache);
// device duration of the last replay (events around the kernel)
double graph_last_seconds(void* graph_cache);
int device_info(char* name, size_t len, int32_t* num_cu);
int block_threads();

}  // namespace backend
}  // namespace clh
