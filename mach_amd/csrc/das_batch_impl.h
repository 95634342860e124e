// Device code of the persistent batched kernels, included by das_batch_tc1.hip and das_batch_tcn.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "das_batch_types.h"
#include "das_device.h"

namespace mach {
namespace batch {
namespace {

using namespace dev;

// Single-frame requests with enough transmits (tx_split_waves): the block's 4 waves split one voxel
// group's transmits, as in the plain kernel. Such requests run as 1 x 1 (batch_layout), which sums in
// the same order as each of their plain layouts.
template <int ST, int INTERP, int LF, int TC>
__device__ __forceinline__ void run_fpl1(const BatchItem &it, uint32_t bx, uint32_t by, float *lds) {
    if constexpr (TC > 1 && LF == 1) {
        if (tx_split_waves(it.p.n_tx, it.p.n_frames) == kTxSplitWaves) {
            das_block<ST, INTERP, LF, 1, TC, 1, false, kTxSplitWaves>(it.p, bx, by, lds);
            return;
        }
    }
    das_block<ST, INTERP, LF, 1, TC>(it.p, bx, by, lds);
}

template <int ST, int INTERP, int TC>
__device__ __forceinline__ void run_layout(const BatchItem &it, uint32_t bx, uint32_t by, float *lds) {
    switch (it.layout) {
        case 0: das_block<ST, INTERP, 16, 4, TC>(it.p, bx, by, lds); break;
        case 1: das_block<ST, INTERP, 32, 2, TC>(it.p, bx, by, lds); break;
        case 2: run_fpl1<ST, INTERP, 64, TC>(it, bx, by, lds); break;
        case 3: das_block<ST, INTERP, 8, 4, TC>(it.p, bx, by, lds); break;
        case 4: das_block<ST, INTERP, 4, 4, TC>(it.p, bx, by, lds); break;
        case 5: das_block<ST, INTERP, 2, 4, TC>(it.p, bx, by, lds); break;
        case 6: das_block<ST, INTERP, 1, 4, TC>(it.p, bx, by, lds); break;
        case 7: das_block<ST, INTERP, 16, 2, TC>(it.p, bx, by, lds); break;
        case 8: das_block<ST, INTERP, 4, 2, TC>(it.p, bx, by, lds); break;
        case 9: das_block<ST, INTERP, 1, 2, TC>(it.p, bx, by, lds); break;
        case 10: run_fpl1<ST, INTERP, 16, TC>(it, bx, by, lds); break;
        case 11: run_fpl1<ST, INTERP, 4, TC>(it, bx, by, lds); break;
        default: run_fpl1<ST, INTERP, 1, TC>(it, bx, by, lds); break;
    }
}

template <int ST, int TC>
__device__ __forceinline__ void run_interp(const BatchItem &it, uint32_t bx, uint32_t by, float *lds) {
    switch (it.interp) {
        case 0: run_layout<ST, 0, TC>(it, bx, by, lds); break;
        case 2: run_layout<ST, 2, TC>(it, bx, by, lds); break;
        default: run_layout<ST, 1, TC>(it, bx, by, lds); break;
    }
}

// Persistent kernel for one section. TC is the transmit chunk the plain kernel uses for the same
// requests (1 for single-transmit requests, kTxChunk otherwise), so a request's result is identical
// in both paths.
template <int TC>
__global__ __launch_bounds__(256) void das_batch_kernel(const BatchHeader *__restrict__ header,
                                                        const BatchItem *__restrict__ all_items,
                                                        uint32_t *counters) {
    constexpr int kSec = TC > 1 ? 1 : 0;
    __shared__ uint32_t s_work;
    // the LDS of the largest variant above (CH = 1): transmit factors of multi-frame lanes (FPL <= 4, no
    // split), or the partial sums and shared pair geometry of a transmit split (FPL = 1)
    constexpr int kLds = TC > 1 ? (das_lds_floats<TC, 1, 4, 1>() > das_lds_floats<TC, 1, 1, kTxSplitWaves>()
                                       ? das_lds_floats<TC, 1, 4, 1>()
                                       : das_lds_floats<TC, 1, 1, kTxSplitWaves>())
                                : 1;
    __shared__ float lds[kLds];
    const uint32_t n_items = header->n_items[kSec];
    const uint32_t total = header->total_blocks[kSec];
    const BatchItem *__restrict__ items = all_items + header->first_item[kSec];
    uint32_t *counter = counters + kSec;
    for (;;) {
        __syncthreads();  // everyone has read the previous s_work
        if (threadIdx.x == 0) s_work = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t w = s_work;
        if (w >= total) break;  // block-uniform: the whole workgroup leaves together
        uint32_t lo = 0, hi = n_items - 1;  // last item with first_block <= w (n_items >= 1 here)
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (items[mid].first_block <= w) lo = mid; else hi = mid - 1;
        }
        const BatchItem &it = items[lo];
        const uint32_t local = w - it.first_block;
        const uint32_t bx = local % it.blocks_x, by = local / it.blocks_x;
        switch (it.store) {
            case kStoreF32: run_interp<kStoreF32, TC>(it, bx, by, lds); break;
            case kStoreC32H: run_interp<kStoreC32H, TC>(it, bx, by, lds); break;
            default: run_interp<kStoreC64, TC>(it, bx, by, lds); break;
        }
    }
}

}  // namespace
}  // namespace batch
}  // namespace mach
