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
void host_free(void* p);
void h2d(void* dst, const void* src, size_t bytes, void* stream);
void d2h(void* dst, const void* src, size_t bytes, void* stream);
void sync(void* stream);
void* stream_create();
void stream_destroy(void* stream);
// run `passes` IPM loop passes on every slot of descs[0..nslots); a slot with
// n < 0 is empty.  phase_out[i] receives the slot's phase afterwards (-1 if empty).
void launch_ipm(const SlotDesc* descs_dev, int nslots, int passes, bool fp32, int32_t* phase_out_dev, void* stream);
// the same, captured once into a graph and replayed (fixed arguments)
void launch_ipm_graph(void** graph_cache, const SlotDesc* descs_dev, int nslots, int passes, bool fp32,
                      int32_t* phase_out_dev, int32_t* phase_out_host, void* stream);
void graph_destroy(void* graph_cache);
// device duration of the last replay (events around the kernel)
double graph_last_seconds(void* graph_cache);
int device_info(char* name, size_t len, int32_t* num_cu);
int block_threads();

}  // namespace backend
}  // namespace clh
