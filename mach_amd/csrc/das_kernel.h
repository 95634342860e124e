// Internal interface between the C ABI (mach_capi.hip) and the delay-and-sum kernels (das_kernel.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace mach {

enum StoreKind : int { kStoreF32 = 0, kStoreC64 = 1, kStoreC32H = 2 };

// Transmits whose pair geometry is computed together (compounding / multi-transmit requests).
constexpr int kTxChunk = 8;

// Everything a launch needs, already validated and converted. Geometry scalars are float64 values
// of the float32 API scalars.
struct DasParams {
    const void *data;        // [n_tx][n_rx][n_samples][n_frames] of the storage type
    const float *rx;         // [n_rx][3]
    const float *vox;        // [n_vox][3]
    const float *arrivals;   // [n_tx][n_vox]
    void *out;               // [n_vox][n_frames]
    const int32_t *order;    // optional [n_vox] processing order (nullptr = identity)
    int64_t n_tx, n_rx, n_samples, n_frames, n_vox;
    int64_t tap_step;        // elements between interpolation taps: n_frames (0 when n_samples == 1)
    double two_f_number;     // 2 * F
    double inv_c;            // 1 / c
    double t0;               // rx_start_s
    double fs;               // sampling frequency
    double fmod;             // modulation frequency (used only when rotate != 0)
    double alpha;            // Tukey alpha
    double half_pi_over_alpha;
    int32_t rotate;          // IQ data with fmod != 0
    int32_t accumulate;      // out += result (1) or out = result (0)
    int32_t apod;            // alpha > 0
    int32_t rows32;          // n_samples * n_frames < 2^32: sample offsets within a channel fit 32 bits
};

// Frame-lane layout of a launch: LF lanes of a wave share one voxel and split its frames, each lane
// handling FPL consecutive frames; a wave therefore covers 64 / LF voxels.
struct DasLayout {
    int lanes_per_voxel;   // LF in {1, 2, 4, 8, 16, 32, 64}
    int frames_per_lane;   // FPL in {1, 2, 4}
};

// Picks the layout for a problem (frame count and pointer alignment decide vector widths).
DasLayout choose_layout(const DasParams &p, int store);
// True if `l` is a compiled layout whose frames-per-lane divides n_frames and keeps vector loads and
// stores aligned for these pointers.
bool layout_valid(const DasParams &p, int store, DasLayout l);

// Enqueues the delay-and-sum kernel. Returns hipSuccess or the launch error.
hipError_t launch_das(const DasParams &p, int store, int interp, DasLayout layout, hipStream_t stream);

}  // namespace mach
