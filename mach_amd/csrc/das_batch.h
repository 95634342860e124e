// Internal interface of the batched (persistent, graph-replayed) delay-and-sum path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "das_kernel.h"

namespace mach {

struct BatchState;

// Allocates two descriptor slots of `max_items` requests (pinned staging + device table + counter).
hipError_t batch_create(uint32_t max_items, bool use_graph, BatchState **out);
// Waits for pending work of both slots, then frees everything.
void batch_destroy(BatchState *b);
uint32_t batch_capacity(const BatchState *b);
// Layout the batched kernel runs for a request whose plain-kernel layout is `want`.
DasLayout batch_layout(const DasParams &p, int store, DasLayout want);
bool batch_layout_supported(DasLayout l);
// Enqueues one persistent-kernel launch covering all `n` requests on `stream`.
hipError_t batch_submit(BatchState *b, const DasParams *params, const int *stores, const int *interps,
                        const DasLayout *layouts, uint32_t n, hipStream_t stream);

// Requests one allocation-free submission takes (the size of its static device table).
constexpr uint32_t kNoAllocMaxRequests = 64;
// batch_submit without a BatchState: the table and counters are the device's static scratch array, the
// table is written by kernels from their arguments, and nothing is allocated (n <= kNoAllocMaxRequests).
hipError_t batch_submit_noalloc(const DasParams *params, const int *stores, const int *interps,
                                const DasLayout *layouts, uint32_t n, hipStream_t stream);

}  // namespace mach
