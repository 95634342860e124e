// Plain (one request per launch) delay-and-sum kernel for gfx950 and its host-side dispatch. The
// device code is in das_device.h.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "das_device.h"
#include "das_kernel.h"

namespace mach {
namespace {

using namespace dev;

template <int ST, int INTERP, int LF, int FPL, int TC>
__global__ __launch_bounds__(256) void das_kernel(const DasParams p) {
    const uint32_t lb = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    das_block<ST, INTERP, LF, FPL, TC>(p, lb % gridDim.x, lb / gridDim.x);
}

template <int ST, int INTERP, int LF, int FPL>
hipError_t launch_one(const DasParams &p, hipStream_t stream) {
    constexpr int LV = 64 / LF;
    const int64_t n_vgroups = (p.n_vox + LV - 1) / LV;
    const int64_t blocks_x = (n_vgroups + 3) / 4;
    const int64_t blocks_y = (p.n_frames + LF * FPL - 1) / (LF * FPL);
    if (blocks_x * blocks_y > 0x7fffffff || blocks_y > 65535) return hipErrorInvalidConfiguration;
    const dim3 grid((unsigned)blocks_x, (unsigned)blocks_y, 1);
    if (p.n_tx > 1) {  // amortise the pair geometry over chunks of transmits
        hipLaunchKernelGGL((das_kernel<ST, INTERP, LF, FPL, kTxChunk>), grid, dim3(256, 1, 1), 0, stream, p);
    } else {
        hipLaunchKernelGGL((das_kernel<ST, INTERP, LF, FPL, 1>), grid, dim3(256, 1, 1), 0, stream, p);
    }
    return hipGetLastError();
}

template <int ST, int INTERP, int LF>
hipError_t dispatch_fpl(const DasParams &p, int fpl, hipStream_t stream) {
    switch (fpl) {
        case 4: return launch_one<ST, INTERP, LF, 4>(p, stream);
        case 2: return launch_one<ST, INTERP, LF, 2>(p, stream);
        default: return launch_one<ST, INTERP, LF, 1>(p, stream);
    }
}

template <int ST, int INTERP>
hipError_t dispatch_lf(const DasParams &p, DasLayout l, hipStream_t stream) {
    switch (l.lanes_per_voxel) {
        case 64: return dispatch_fpl<ST, INTERP, 64>(p, l.frames_per_lane, stream);
        case 32: return dispatch_fpl<ST, INTERP, 32>(p, l.frames_per_lane, stream);
        case 16: return dispatch_fpl<ST, INTERP, 16>(p, l.frames_per_lane, stream);
        case 8: return dispatch_fpl<ST, INTERP, 8>(p, l.frames_per_lane, stream);
        case 4: return dispatch_fpl<ST, INTERP, 4>(p, l.frames_per_lane, stream);
        case 2: return dispatch_fpl<ST, INTERP, 2>(p, l.frames_per_lane, stream);
        default: return dispatch_fpl<ST, INTERP, 1>(p, l.frames_per_lane, stream);
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

int elem_bytes(int store) { return store == kStoreC64 ? 8 : 4; }
int out_elem_bytes(int store) { return store == kStoreF32 ? 4 : 8; }

}  // namespace

bool layout_valid(const DasParams &p, int store, DasLayout l) {
    const int lf = l.lanes_per_voxel, k = l.frames_per_lane;
    if (lf < 1 || lf > 64 || (lf & (lf - 1)) != 0) return false;
    if (k != 1 && k != 2 && k != 4) return false;
    return (p.n_frames % k) == 0 && ((uintptr_t)p.data % (uintptr_t)(k * elem_bytes(store))) == 0 &&
           ((uintptr_t)p.out % (uintptr_t)(k * out_elem_bytes(store))) == 0;
}

DasLayout choose_layout(const DasParams &p, int store) {
    const int64_t F = p.n_frames;
    auto aligned = [&](int k) {
        return (F % k) == 0 && ((uintptr_t)p.data % (uintptr_t)(k * elem_bytes(store))) == 0 &&
               ((uintptr_t)p.out % (uintptr_t)(k * out_elem_bytes(store))) == 0;
    };
    // Ensembles: 64 frames per voxel group, as 16 lanes x 4 frames (32 x 2 or 64 x 1 when the rows are
    // less aligned). Measured on gfx950 for 256-frame Doppler ensembles: 16x4 and 32x2 are within 4% of
    // each other and ~1.8x faster than 64x4 (c64 and FP16 storage).
    if (F >= 64) {
        if (aligned(4)) return DasLayout{16, 4};
        if (aligned(2)) return DasLayout{32, 2};
        return DasLayout{64, 1};
    }
    // Short ensembles: as many frames per lane as alignment allows, and just enough lanes to cover the
    // frames (8 frames -> 2 x 4, the measured optimum for 8-frame volumes).
    DasLayout l{1, 1};
    l.frames_per_lane = (F >= 4 && aligned(4)) ? 4 : ((F >= 2 && aligned(2)) ? 2 : 1);
    const int64_t lanes = (F + l.frames_per_lane - 1) / l.frames_per_lane;
    int lf = 1;
    while (lf < lanes && lf < 64) lf <<= 1;
    // A single frame of a single transmit is setup-bound: 4 lanes per voxel measured 1.6x faster than
    // 1 on a 65536-voxel plane-wave image. With several transmits the setup is shared by a transmit
    // chunk and 1 lane per voxel is fastest.
    if (F == 1 && p.n_tx == 1) lf = 4;
    l.lanes_per_voxel = lf;
    return l;
}

hipError_t launch_das(const DasParams &p, int store, int interp, DasLayout layout, hipStream_t stream) {
    switch (store) {
        case kStoreF32: return dispatch_interp<kStoreF32>(p, interp, layout, stream);
        case kStoreC32H: return dispatch_interp<kStoreC32H>(p, interp, layout, stream);
        default: return dispatch_interp<kStoreC64>(p, interp, layout, stream);
    }
}

}  // namespace mach
