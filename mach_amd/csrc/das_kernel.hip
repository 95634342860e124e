// Host-side policy of the plain delay-and-sum kernel: layout validity, the layout heuristic and the
// launch by storage type. The kernels are in das_kernel_impl.h (one translation unit per storage type).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "das_kernel.h"
#include "das_post.h"

namespace mach {

hipError_t launch_das_f32(const DasParams &p, int interp, DasLayout l, hipStream_t stream);
hipError_t launch_das_c64(const DasParams &p, int interp, DasLayout l, hipStream_t stream);
hipError_t launch_das_c32h(const DasParams &p, int interp, DasLayout l, hipStream_t stream);

namespace {

int elem_bytes(int store) { return store == kStoreC64 ? 8 : 4; }
// Alignment the epilogue's stores need for `k` frames per lane.
uintptr_t out_align(const DasParams &p, int store, int k) {
    if (p.output != kOutSignal) return 4;  // float32 scalar stores
    return (uintptr_t)k * (store == kStoreF32 ? 4 : 8);
}
// True if `k` frames per lane divide the frames and keep the vector loads and stores aligned.
bool aligned_for(const DasParams &p, int store, int k) {
    return (p.n_frames % k) == 0 && ((uintptr_t)p.data % (uintptr_t)(k * elem_bytes(store))) == 0 &&
           ((uintptr_t)p.out % out_align(p, store, k)) == 0;
}

}  // namespace

bool layout_valid(const DasParams &p, int store, DasLayout l) {
    const int lf = l.lanes_per_voxel, k = l.frames_per_lane, ch = l.chunks_per_lane;
    if (l.voxels_per_lane != 1) {  // lateral groups: single transmit, 64 lanes, 2 or 4 frames per lane
        if ((l.voxels_per_lane != 2 && l.voxels_per_lane != 4) || lf != 64 || (k != 2 && k != 4) || ch != 1 ||
            p.n_tx != 1)
            return false;
    }
    if (lf < 1 || lf > 64 || (lf & (lf - 1)) != 0) return false;
    if (k != 1 && k != 2 && k != 4) return false;
    if (ch != 1 && !((ch == 2 || ch == 4) && multi_group_layout(lf, k))) return false;
    return aligned_for(p, store, k);
}

DasLayout choose_layout(const DasParams &p, int store) {
    // Single-transmit complex64 ensembles of 128 frames or more, in a processing order: lateral groups,
    // whose lateral neighbours share one window of sample rows per element. Measured on gfx950 for the
    // 256-frame Doppler ensemble in depth-major order (the order the Python front end passes with an
    // automatic layout): 4 voxels x 2 frames 0.30 ms for the signal and envelope, 2 voxels x 2 frames
    // 0.30 ms for the power image, against 0.39-0.41 ms for the best general layout. Without an order the
    // caller's voxel order need not put lateral neighbours together (grids stored depth-fastest do not),
    // so the general layouts are kept.
    if (store == kStoreC64 && p.n_tx == 1 && p.n_frames >= 128 && p.order != nullptr && aligned_for(p, store, 2))
        return DasLayout{64, 2, 1, p.output == kOutPower ? 2 : 4};
    return choose_general_layout(p, store);
}

DasLayout choose_general_layout(const DasParams &p, int store) {
    const int64_t F = p.n_frames;
    auto aligned = [&](int k) { return aligned_for(p, store, k); };
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
    if (p.output != kOutSignal && p.peak != nullptr) {  // the epilogue raises it from 0
        const hipError_t e = launch_fill(p.peak, 1, 0.f, stream);
        if (e != hipSuccess) return e;
    }
    switch (store) {
        case kStoreF32: return launch_das_f32(p, interp, layout, stream);
        case kStoreC32H: return launch_das_c32h(p, interp, layout, stream);
        default: return launch_das_c64(p, interp, layout, stream);
    }
}

}  // namespace mach
