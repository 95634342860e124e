// Internal interface between the C ABI (mach_capi.hip) and the delay-and-sum kernels (das_kernel.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace mach {

enum StoreKind : int { kStoreF32 = 0, kStoreC64 = 1, kStoreC32H = 2 };

// Transmits whose pair geometry is computed together (compounding / multi-transmit requests).
constexpr int kTxChunk = 8;

// Waves that split one voxel group's transmits (kTxSplitWaves when a request has one frame and at least
// 4 chunks of transmits, else 1). Each of the workgroup's 4 waves takes a contiguous quarter of the
// transmits; their partial sums are added in transmit order. A property of the request alone, so every
// layout, and the batched kernel, sums in the same order. Why: a single-frame compound has one wave
// per 64 voxels and, on a 236k-voxel grid, too few waves to fill the GPU (3.3 ms; the same kernel on
// 4x the voxels costs 29 % less per voxel).
constexpr int kTxSplitWaves = 4;
__host__ __device__ constexpr int tx_split_waves(int64_t n_tx, int64_t n_frames) {
    return (n_frames == 1 && n_tx >= kTxSplitWaves * kTxChunk) ? kTxSplitWaves : 1;
}

// Everything a launch needs, already validated and converted. Geometry scalars are float64 values
// of the float32 API scalars.
// One transmit's wavefront (layout of mach_wave in include/mach_hip.h).
enum WaveKind { kWavePlane = 0, kWaveSpherical = 1 };
struct WaveDesc {
    int32_t kind;
    int32_t reserved;
    double origin[3];
    double vec[3];  // plane: unit direction; spherical: focus
    double c;
    double offset;
};

// What the kernel writes (mach_output_kind): the beamformed signal [n_vox][n_frames], the power
// Doppler image sum_f |y|^2 [n_vox] (float32), or the envelope |y| [n_vox][n_frames] (float32).
enum OutputKind { kOutSignal = 0, kOutPower = 1, kOutEnvelope = 2 };

struct DasParams {
    const void *data;        // [n_tx][n_rx][n_samples][n_frames] of the storage type
    const float *rx;         // [n_rx][3]
    const float *vox;        // [n_vox][3]
    const float *arrivals;   // [n_tx][n_vox], or nullptr when `waves` is given
    const WaveDesc *waves;   // optional [n_tx] wavefronts: arrivals computed per voxel in the kernel
    void *out;               // [n_vox][n_frames]
    const int32_t *order;    // optional [n_vox] processing order (nullptr = identity)
    int64_t n_tx, n_rx, n_samples, n_frames, n_vox;
    int64_t tap_step;        // elements between interpolation taps: n_frames (0 when n_samples == 1)
    double two_f_number;     // 2 * F
    double inv_c;            // 1 / c
    double t0;               // rx_start_s
    double fs;               // sampling frequency
    double fmod;             // modulation frequency (used only when rotate != 0)
    double flat_r2;          // (1 - alpha)^2: the Tukey window's flat top is h^2 <= flat_r2 * R^2
    double lin_tap_max;      // max(n_samples - 2, 0): the largest first tap of linear interpolation
    float quarter_over_alpha;  // 1 / (4 alpha): taper argument in revolutions per unit of (1 - r)
    int32_t rotate;          // IQ data with fmod != 0
    int32_t accumulate;      // out += result (1) or out = result (0)
    int32_t apod;            // alpha > 0
    int32_t bytes32;         // n_samples * n_frames * element bytes < 2^32: byte offsets within a
                             // channel's block fit 32 bits
    int32_t output;          // OutputKind
    int32_t xcd_group;       // block-to-XCD remap (xcd_remap): 0 = contiguous eighths, > 0 = runs of this
                             // many blocks; < 0 = chosen per launch (xcd_group_for)
    float *peak;             // optional: maximum of the written POWER / ENVELOPE values (zeroed first)
};

// Frame-lane layout of a launch: LF lanes of a wave share one voxel and split its frames, each lane
// handling FPL consecutive frames; a wave therefore covers 64 / LF voxels. With voxels_per_lane
// VPL > 1 (lateral groups, LF = 64) every lane instead holds its FPL frames of VPL voxels that are
// consecutive in the processing order, and the wave reads their taps from shared sample windows.
struct DasLayout {
    int lanes_per_voxel;      // LF in {1, 2, 4, 8, 16, 32, 64}
    int frames_per_lane;      // FPL in {1, 2, 4}
    int chunks_per_lane = 1;  // CH in {1, 2, 4}: frame groups of LF * FPL frames a lane walks per setup
    int voxels_per_lane = 1;  // VPL in {1, 2, 4}; VPL > 1 only with LF = 64, FPL in {2, 4}, CH = 1
};

// (LF, FPL) layouts compiled with CH > 1 (frame ensembles, where the per-pair setup is worth sharing;
// 1 x 4 x 2 gives one lane all 8 frames of a voxel of the volumetric set, with one setup per pair).
constexpr bool multi_group_layout(int lf, int fpl) {
    return (lf == 16 && fpl == 4) || (lf == 32 && fpl == 2) || (lf == 64 && fpl == 1) || (lf == 8 && fpl == 4) ||
           (lf == 1 && fpl == 4);
}

// Layout code of the C ABI: (log2(CH) << 12) | (LF << 4) | FPL; LF == 0 selects the band kernel.
// Lateral groups: 0x8000 | (log2(VPL) << 12) | (64 << 4) | FPL.
// Bits 16-19 optionally fix the block-to-XCD remap of the launch (xcd_remap): 0 = chosen per launch
// (xcd_group_for), 1 = contiguous eighths, 2 <= v <= 10 = runs of 2^(v + 1) blocks (8 .. 4096),
// kXcdWedgeCode = the hardware order itself (runs of 1 block: block bx of every block row on XCD
// bx % 8, the rows padded to a multiple of 8 blocks), for a voxel order that already assigns each XCD
// its own voxels (the host's XCD wedges of a volume: every XCD one angular sector of every depth band).
constexpr int kXcdWedgeCode = 15;
// The XCD remap fixed by a layout code's bits 16-19 (the xcd_group of DasParams), or -1 for none.
inline int32_t layout_xcd_group(int32_t code) {
    const int v = (code >> 16) & 0xf;
    if (v == kXcdWedgeCode) return 1;
    return v == 0 ? -1 : (v == 1 ? 0 : (v > 10 ? -1 : 1 << (v + 1)));
}

// Blocks per row of a launch under remap `group`: with the hardware order (runs of 1 block) and
// several rows, each row is padded to a multiple of the 8 XCDs, so that block bx of every row runs on
// XCD bx % 8 (the padding blocks exit at once).
inline int64_t padded_blocks_x(int64_t blocks_x, int64_t blocks_y, uint32_t group) {
    return (group == 1u && blocks_y > 1) ? (blocks_x + 7) / 8 * 8 : blocks_x;
}

inline DasLayout decode_layout(int32_t code) {
    if (code & 0x8000) {
        const int vpl_log2 = (code >> 12) & 0x7;
        return DasLayout{(code >> 4) & 0xff, code & 0xf, 1, vpl_log2 <= 2 ? 1 << vpl_log2 : 0};
    }
    const int ch_log2 = (code >> 12) & 0xf;
    return DasLayout{(code >> 4) & 0xff, code & 0xf, ch_log2 <= 2 ? 1 << ch_log2 : 0};
}

// Block-to-XCD remap of a plain launch of blocks_x x blocks_y blocks (xcd_remap). Work grows with
// depth (the f-number aperture widens), so with a depth-major voxel order and contiguous eighths the
// deepest XCD finishes last unless several block rows mix the depths. Measured on gfx950 (ms):
//   launch                                   contiguous  runs of 8  runs of 32-128
//   Doppler c64 signal, 32x2x2 (2 rows)        0.639       0.440      0.425 (64)
//   Doppler c64 power image, 32x4 (1 row)      0.640       0.401      0.389 (64)
//   Doppler fp16 power image, 64x4 (1 row)     0.416       0.271      0.257 (128)
//   volume, 2x4, 5 transmits (1 row)          50.5        39.5       37.5 (128)
//   single plane wave, 4x1 (1 row)             0.057       0.054
//   Doppler c64 signal, 32x2 (4 rows)          0.493       0.515
//   75-angle compounding, 1x1 (1 row)          4.50        4.01
// Launches of at most two rows take runs of blocks_x / 128 blocks (a power of two in [8, 128]): each
// XCD then sweeps the depth range in ~16 bands, balanced, while a run keeps neighbouring voxels, which
// share channel-data rows, in one L2. With more rows the L2 reuse of contiguous eighths wins.
inline uint32_t xcd_group_for(const DasParams &p, int64_t blocks_x, int64_t blocks_y) {
    if (p.xcd_group >= 0) return (uint32_t)p.xcd_group;
    if (blocks_y > 2) return 0u;
    uint32_t g = 8u;
    while (g < 128u && (int64_t)(2 * g) * 128 <= blocks_x) g *= 2u;
    return g;
}

// Picks the layout for a problem (frame count and pointer alignment decide vector widths).
DasLayout choose_layout(const DasParams &p, int store);
// The same among the general kernel's layouts (voxels_per_lane 1): for the batched kernels and the band
// kernel's plan, which have no lateral groups.
DasLayout choose_general_layout(const DasParams &p, int store);
// True if `l` is a compiled layout whose frames-per-lane divides n_frames and keeps vector loads and
// stores aligned for these pointers.
bool layout_valid(const DasParams &p, int store, DasLayout l);

// Enqueues the delay-and-sum kernel. Returns hipSuccess or the launch error.
hipError_t launch_das(const DasParams &p, int store, int interp, DasLayout layout, hipStream_t stream);
// The lateral-group kernels (layout.voxels_per_lane 2 or 4; das_lateral.hip), reached through launch_das.
hipError_t launch_das_lateral(const DasParams &p, int store, int interp, DasLayout layout, hipStream_t stream);

#ifdef MACH_DEBUG
// Debug builds: failed kernel bounds checks since the last read, per translation unit (read and clear).
hipError_t debug_violations_f32(uint32_t *count);
hipError_t debug_violations_c64(uint32_t *count);
hipError_t debug_violations_c32h(uint32_t *count);
hipError_t debug_violations_batch0(uint32_t *count);
hipError_t debug_violations_batch1(uint32_t *count);
hipError_t debug_violations_band(uint32_t *count);
hipError_t debug_violations_lateral(uint32_t *count);
#endif

}  // namespace mach
