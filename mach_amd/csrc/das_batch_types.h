// Descriptor types of the batched delay-and-sum path, shared by its host code (das_batch.hip) and
// the persistent kernels (das_batch_tc1.hip, das_batch_tcn.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "das_kernel.h"

namespace mach {
namespace batch {

// (lanes per voxel, frames per lane) variants compiled into the batched kernels: every layout the
// plain kernel's heuristic picks for ensembles (16x4, 32x2, 64x1) and for short frame counts, the
// autotuner's winners for the benchmark workloads (32x2 and 16x4 ensembles, 2x4 volumetric, 4x1 single
// frame, 1x1 compounding), and fallbacks for every frame count.
constexpr int kBatchLayouts[][2] = {{16, 4}, {32, 2}, {64, 1}, {8, 4}, {4, 4}, {2, 4}, {1, 4},
                                    {16, 2}, {4, 2}, {1, 2}, {16, 1}, {4, 1}, {1, 1}};
constexpr int kNumBatchLayouts = sizeof(kBatchLayouts) / sizeof(kBatchLayouts[0]);

// A batch has two sections, drained by two persistent kernels launched back to back: single-transmit
// requests (section 0, no transmit chunking) and multi-transmit requests (section 1, chunks of
// kTxChunk transmits). Each kernel's register budget, and so its occupancy, is set by its own
// variants only: the single-transmit variants need about half the registers of the chunked ones.
constexpr int kSections = 2;

struct BatchHeader {
    uint32_t n_items[kSections];
    uint32_t total_blocks[kSections];
    uint32_t first_item[kSections];
    uint32_t pad[2];
};

struct BatchItem {
    DasParams p;
    int32_t store;
    int32_t interp;
    int32_t layout;  // index into kBatchLayouts
    uint32_t blocks_x;
    uint32_t blocks_y;
    uint32_t first_block;  // within the item's section
    uint32_t pad[2];
};

// Persistent kernels, one translation unit each (so they compile in parallel). Section 0 holds the
// single-transmit requests, section 1 the multi-transmit ones.
hipError_t launch_section0(unsigned grid, hipStream_t stream, const BatchHeader *header, const BatchItem *items,
                           uint32_t *counters);
hipError_t launch_section1(unsigned grid, hipStream_t stream, const BatchHeader *header, const BatchItem *items,
                           uint32_t *counters);
// Resident 256-thread workgroups per CU of each section's kernel.
hipError_t section_occupancy(int section, int *blocks_per_cu);
// Sets both sections' work counters to 0 (a kernel, so that it can be captured into a graph).
hipError_t launch_reset_counters(hipStream_t stream, uint32_t *counters);

}  // namespace batch
}  // namespace mach
