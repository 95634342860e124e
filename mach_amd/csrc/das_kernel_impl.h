// Plain (one request per launch) delay-and-sum kernel templates and their launch dispatch, included by
// one translation unit per storage type (das_kernel_f32.hip, das_kernel_c64.hip, das_kernel_c32h.hip)
// so the variants compile in parallel. The device code is in das_device.h; the lateral-group kernels
// have their own translation unit (das_lateral.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "das_device.h"
#include "das_kernel.h"

namespace mach {
namespace {

using namespace dev;

// Waves per SIMD asked of the compiler. The FP16, linear, multi-transmit 1 x 4 x 2 variant (the tuned
// layout of the 8-frame volume) fits 6 in 80 VGPRs with a few spills, mostly outside its loops: 18.42 ->
// 18.21 ms against 5 waves at 96 VGPRs. The FP16 quadratic 64 x 4 variant (the FP16 Doppler ensemble)
// fits 6 in 78 VGPRs without spills: power image 0.210 -> 0.207 ms, signal unchanged. (The headline's
// lateral 4-voxel kernel at 6 waves, 80 VGPRs with spills, measured 0.363 ms against 0.348 at 5.)
constexpr int das_waves_per_eu(int st, int interp, int lf, int fpl, int ch, int tc, int tsx) {
    if (tsx > 1) return 4;
    return ((st == kStoreC32H && interp == kLinear && lf == 1 && fpl == 4 && ch == 2 && tc > 1) ||
            (st == kStoreC32H && interp == kQuadratic && lf == 64 && fpl == 4 && ch == 1 && tc == 1))
               ? 6
               : 1;
}

template <int ST, int INTERP, int LF, int FPL, int TC, int CH, bool WIDE = false, int TSX = 1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(das_waves_per_eu(ST, INTERP, LF, FPL, CH, TC, TSX))))
void das_kernel(const DasParams p) {
    __shared__ float lds[das_lds_floats<TC, CH, FPL, TSX>()];
    const uint32_t lb = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y, (uint32_t)p.xcd_group);
    const uint32_t bx = lb % gridDim.x;
    if ((int64_t)bx * (4 / TSX) * (64 / LF) >= p.n_vox) return;  // row padding (padded_blocks_x)
    das_block<ST, INTERP, LF, FPL, TC, CH, WIDE, TSX>(p, bx, lb / gridDim.x, lds);
}

template <int ST, int INTERP, int LF, int FPL, int CH, bool WIDE = false>
hipError_t launch_one(const DasParams &p, hipStream_t stream) {
    constexpr int LV = 64 / LF;
    const int64_t n_vgroups = (p.n_vox + LV - 1) / LV;
    // the power epilogue walks all frame blocks inside each block
    const int64_t per = (int64_t)LF * FPL * CH;
    const int64_t blocks_y = p.output == kOutPower ? 1 : (p.n_frames + per - 1) / per;
    DasParams q = p;
    q.xcd_group = (int32_t)xcd_group_for(p, (n_vgroups + 3) / 4, blocks_y);
    const int64_t blocks_x = padded_blocks_x((n_vgroups + 3) / 4, blocks_y, (uint32_t)q.xcd_group);
    if (blocks_x * blocks_y > 0x7fffffff || blocks_y > 65535) return hipErrorInvalidConfiguration;
    const dim3 grid((unsigned)blocks_x, (unsigned)blocks_y, 1);
    if (p.n_tx > 1) {  // amortise the pair geometry over chunks of transmits
        // single-frame requests with enough transmits: the workgroup's 4 waves split one voxel group's
        // transmits (every layout of theirs has one frame per lane)
        if constexpr (FPL == 1) {
            if (tx_split_waves(p.n_tx, p.n_frames) == kTxSplitWaves) {
                const dim3 grid4((unsigned)n_vgroups, 1, 1);
                if (n_vgroups > 0x7fffffff) return hipErrorInvalidConfiguration;
                q.xcd_group = (int32_t)xcd_group_for(p, n_vgroups, 1);
                hipLaunchKernelGGL((das_kernel<ST, INTERP, LF, FPL, kTxChunk, CH, WIDE, kTxSplitWaves>), grid4,
                                   dim3(256, 1, 1), 0, stream, q);
                return hipGetLastError();
            }
        }
        hipLaunchKernelGGL((das_kernel<ST, INTERP, LF, FPL, kTxChunk, CH, WIDE>), grid, dim3(256, 1, 1), 0, stream, q);
    } else {
        hipLaunchKernelGGL((das_kernel<ST, INTERP, LF, FPL, 1, CH, WIDE>), grid, dim3(256, 1, 1), 0, stream, q);
    }
    return hipGetLastError();
}

// Several frame groups per lane (CH > 1) are compiled for the ensemble layouts only
// (multi_group_layout).
template <int ST, int INTERP, int LF, int FPL>
hipError_t dispatch_ch(const DasParams &p, int ch, hipStream_t stream) {
    if constexpr (multi_group_layout(LF, FPL)) {
        if (ch == 4) return launch_one<ST, INTERP, LF, FPL, 4>(p, stream);
        if (ch == 2) return launch_one<ST, INTERP, LF, FPL, 2>(p, stream);
    }
    if (ch != 1) return hipErrorInvalidValue;
    return launch_one<ST, INTERP, LF, FPL, 1>(p, stream);
}

template <int ST, int INTERP, int LF>
hipError_t dispatch_fpl(const DasParams &p, int fpl, int ch, hipStream_t stream) {
    switch (fpl) {
        case 4: return dispatch_ch<ST, INTERP, LF, 4>(p, ch, stream);
        case 2: return dispatch_ch<ST, INTERP, LF, 2>(p, ch, stream);
        default: return dispatch_ch<ST, INTERP, LF, 1>(p, ch, stream);
    }
}

template <int ST, int INTERP>
hipError_t dispatch_lf(const DasParams &p, DasLayout l, hipStream_t stream) {
    const int k = l.frames_per_lane, ch = l.chunks_per_lane;
    if (l.voxels_per_lane != 1) return launch_das_lateral(p, ST, INTERP, l, stream);  // das_lateral.hip
    switch (l.lanes_per_voxel) {
        case 64: return dispatch_fpl<ST, INTERP, 64>(p, k, ch, stream);
        case 32: return dispatch_fpl<ST, INTERP, 32>(p, k, ch, stream);
        case 16: return dispatch_fpl<ST, INTERP, 16>(p, k, ch, stream);
        case 8: return dispatch_fpl<ST, INTERP, 8>(p, k, ch, stream);
        case 4: return dispatch_fpl<ST, INTERP, 4>(p, k, ch, stream);
        case 2: return dispatch_fpl<ST, INTERP, 2>(p, k, ch, stream);
        default: return dispatch_fpl<ST, INTERP, 1>(p, k, ch, stream);
    }
}

template <int ST>
hipError_t dispatch_interp(const DasParams &p, int interp, DasLayout l, hipStream_t stream) {
    switch (interp) {
        case kNearest: return dispatch_lf<ST, kNearest>(p, l, stream);
        case kQuadratic: return dispatch_lf<ST, kQuadratic>(p, l, stream);
        default: return dispatch_lf<ST, kLinear>(p, l, stream);
    }
}

// Channel blocks of 4 GiB or more (64-bit offsets inside a channel) run one layout, 64 lanes x 1
// frame, which is valid for any frame count and alignment and gives the same sums as every other.
template <int ST>
hipError_t launch_wide(const DasParams &p, int interp, hipStream_t stream) {
    switch (interp) {
        case kNearest: return launch_one<ST, kNearest, 64, 1, 1, true>(p, stream);
        case kQuadratic: return launch_one<ST, kQuadratic, 64, 1, 1, true>(p, stream);
        default: return launch_one<ST, kLinear, 64, 1, 1, true>(p, stream);
    }
}

}  // namespace

// Entry of one storage type's translation unit.
template <int ST>
hipError_t launch_das_store(const DasParams &p, int interp, DasLayout l, hipStream_t stream) {
    if (!p.bytes32) return launch_wide<ST>(p, interp, stream);
    return dispatch_interp<ST>(p, interp, l, stream);
}

}  // namespace mach
