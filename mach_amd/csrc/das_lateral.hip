// Lateral-group delay-and-sum kernels (das_lateral.h) for every storage type, in a translation unit of
// their own: it is compiled with -structurizecfg-skip-uniform-regions (see CMakeLists.txt), which keeps
// their branches on wave-uniform values (the group's window position, which voxels an element reaches)
// as scalar branches, so the accumulators update in place. The general kernels of FP16 storage ran
// slower with that option and are compiled without it (das_kernel_c32h.hip).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "das_device.h"
#include "das_kernel.h"
#include "das_lateral.h"

namespace mach {
namespace {

using namespace dev;

template <int ST, int INTERP, int VPL, int FPL>
__global__ __launch_bounds__(256) void das_lateral_kernel(const DasParams p) {
    const uint32_t lb = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y, (uint32_t)p.xcd_group);
    const uint32_t bx = lb % gridDim.x;
    if ((int64_t)bx * 4 * VPL >= p.n_vox) return;  // row padding (padded_blocks_x)
    das_lateral_block<ST, INTERP, VPL, FPL>(p, bx, lb / gridDim.x);
}

// Lateral groups of VPL voxels per wave, 64 * FPL frames per block (single transmit).
template <int ST, int INTERP, int VPL, int FPL>
hipError_t launch_lateral(const DasParams &p, hipStream_t stream) {
    const int64_t n_groups = (p.n_vox + VPL - 1) / VPL;
    const int64_t per = 64 * FPL;
    const int64_t blocks_y = p.output == kOutPower ? 1 : (p.n_frames + per - 1) / per;
    DasParams q = p;
    q.xcd_group = (int32_t)xcd_group_for(p, (n_groups + 3) / 4, blocks_y);
    const int64_t blocks_x = padded_blocks_x((n_groups + 3) / 4, blocks_y, (uint32_t)q.xcd_group);
    if (blocks_x * blocks_y > 0x7fffffff || blocks_y > 65535) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((das_lateral_kernel<ST, INTERP, VPL, FPL>), dim3((unsigned)blocks_x, (unsigned)blocks_y, 1),
                       dim3(256, 1, 1), 0, stream, q);
    return hipGetLastError();
}

template <int ST, int INTERP>
hipError_t dispatch_lateral(const DasParams &p, DasLayout l, hipStream_t stream) {
    const bool four = l.frames_per_lane == 4;
    if (l.voxels_per_lane == 4)
        return four ? launch_lateral<ST, INTERP, 4, 4>(p, stream) : launch_lateral<ST, INTERP, 4, 2>(p, stream);
    return four ? launch_lateral<ST, INTERP, 2, 4>(p, stream) : launch_lateral<ST, INTERP, 2, 2>(p, stream);
}

template <int ST>
hipError_t dispatch_lateral_interp(const DasParams &p, int interp, DasLayout l, hipStream_t stream) {
    switch (interp) {
        case kNearest: return dispatch_lateral<ST, kNearest>(p, l, stream);
        case kQuadratic: return dispatch_lateral<ST, kQuadratic>(p, l, stream);
        default: return dispatch_lateral<ST, kLinear>(p, l, stream);
    }
}

}  // namespace

hipError_t launch_das_lateral(const DasParams &p, int store, int interp, DasLayout l, hipStream_t stream) {
    switch (store) {
        case kStoreF32: return dispatch_lateral_interp<kStoreF32>(p, interp, l, stream);
        case kStoreC32H: return dispatch_lateral_interp<kStoreC32H>(p, interp, l, stream);
        default: return dispatch_lateral_interp<kStoreC64>(p, interp, l, stream);
    }
}

#ifdef MACH_DEBUG
MACH_DEBUG_READER(debug_violations_lateral)
#endif

}  // namespace mach
