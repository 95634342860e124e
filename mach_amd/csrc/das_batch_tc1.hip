// Persistent batched kernel of section 0 (single-transmit requests), the counter reset kernel and the
// occupancy query.
#include "das_batch_impl.h"

namespace mach {
namespace batch {
namespace {

__global__ void reset_counters(uint32_t *counters) {
    if (threadIdx.x < kSections) counters[threadIdx.x] = 0u;
}

}  // namespace

hipError_t launch_section0(unsigned grid, hipStream_t stream, const BatchHeader *header, const BatchItem *items,
                           uint32_t *counters) {
    hipLaunchKernelGGL(das_batch_kernel<1>, dim3(grid), dim3(256), 0, stream, header, items, counters);
    return hipGetLastError();
}

hipError_t launch_reset_counters(hipStream_t stream, uint32_t *counters) {
    hipLaunchKernelGGL(reset_counters, dim3(1), dim3(64), 0, stream, counters);
    return hipGetLastError();
}

hipError_t section_occupancy0(int *blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, das_batch_kernel<1>, 256, 0);
}

}  // namespace batch
#ifdef MACH_DEBUG
MACH_DEBUG_READER(debug_violations_batch0)
#endif

}  // namespace mach
