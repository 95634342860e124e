// Persistent batched kernel of section 1 (multi-transmit requests, transmit chunks of kTxChunk) and
// the occupancy query of both sections.
#include "das_batch_impl.h"

namespace mach {
namespace batch {

hipError_t section_occupancy0(int *blocks_per_cu);

hipError_t launch_section1(unsigned grid, hipStream_t stream, const BatchHeader *header, const BatchItem *items,
                           uint32_t *counters) {
    hipLaunchKernelGGL(das_batch_kernel<kTxChunk>, dim3(grid), dim3(256), 0, stream, header, items, counters);
    return hipGetLastError();
}

hipError_t section_occupancy(int section, int *blocks_per_cu) {
    if (section == 0) return section_occupancy0(blocks_per_cu);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, das_batch_kernel<kTxChunk>, 256, 0);
}

}  // namespace batch
#ifdef MACH_DEBUG
MACH_DEBUG_READER(debug_violations_batch1)
#endif

}  // namespace mach
