// LDS-staged delay-and-sum for ensembles (many frames): the "band" kernel.
//
// A workgroup owns a tile of TV = 4 waves x VPW voxels that are consecutive in the depth-major voxel
// order, i.e. laterally adjacent voxels at (almost) the same depth, and one chunk of FC = 64 * FPL
// frames. For such a tile the interpolation taps of one receive element all fall in a short window
// of channel-data rows (a few samples of lateral delay spread + the tap count). Per element the
// workgroup:
//   1. stages that window [lo, lo + W) x FC frames from HBM into LDS with coalesced 16-byte loads
//      (each row's FC frames are contiguous in the [channel][sample][frame] layout), double-buffered
//      so that the loads for the next element are in flight while this one is computed
//      (one barrier per element);
//   2. every wave interpolates its voxels' taps from LDS (lanes over frames: conflict-free 8/16-byte
//      reads), applies the complex apodised coefficient and accumulates in registers.
// Per (voxel, element) geometry is computed once per tile in float64 exactly as in das_kernel (lane =
// element, VPW voxels per wave, kept in VGPRs and read back with v_readlane), so results are
// bit-identical to the plain kernel. The window bounds come from a cross-wave LDS min/max of the
// first taps. Elements whose window exceeds the LDS buffer read their taps from global memory.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>

#include "das_band.h"
#include "das_device.h"

namespace mach {
namespace {

using namespace dev;

constexpr int kWaves = 4;
constexpr int kVoxelsPerWave = 4;
constexpr int kTileVoxels = kWaves * kVoxelsPerWave;
constexpr int kWindowBytes = 24 * 1024;  // one of the two staging buffers

template <int ST>
constexpr int elem_size() {
    return ST == kStoreC64 ? 8 : 4;
}

template <int INTERP>
constexpr int n_taps() {
    return INTERP == 0 ? 1 : (INTERP == 1 ? 2 : 3);
}

template <int ST, int FPL>
constexpr int window_rows() {
    return kWindowBytes / (64 * FPL * elem_size<ST>());
}

template <int ST, int FPL>
constexpr int chunks_per_thread() {  // 16-byte chunks a thread stages for a full window
    return (window_rows<ST, FPL>() * 64 * FPL * elem_size<ST>() / 16 + 255) / 256;
}

template <int ST, int INTERP, int FPL>
__global__ __launch_bounds__(256) void das_band_kernel(const DasParams p) {
    constexpr int FC = 64 * FPL;                  // frames per chunk (one row in LDS)
    constexpr int ROW_BYTES = FC * elem_size<ST>();
    constexpr int WMAX = window_rows<ST, FPL>();
    constexpr int CPT = chunks_per_thread<ST, FPL>();
    constexpr int NT = n_taps<INTERP>();
    __shared__ __attribute__((aligned(16))) unsigned char s_win[2][kWindowBytes];
    __shared__ int s_lo[64];
    __shared__ int s_hi[64];

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t lb = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y,
                                  p.xcd_group > 0 ? (uint32_t)p.xcd_group : 0u);
    const uint32_t tile = lb % gridDim.x, chunk = lb / gridDim.x;
    const int64_t f_base = (int64_t)chunk * FC;
    const int64_t f0 = f_base + (int64_t)lane * FPL;

    int64_t vox[kVoxelsPerWave];
    bool vok[kVoxelsPerWave];
    double xv[kVoxelsPerWave], yv[kVoxelsPerWave], zv[kVoxelsPerWave], R2[kVoxelsPerWave], invR[kVoxelsPerWave];
    float acc_re[kVoxelsPerWave][FPL], acc_im[kVoxelsPerWave][FPL];
#pragma unroll
    for (int j = 0; j < kVoxelsPerWave; ++j) {
        const int64_t slot = (int64_t)tile * kTileVoxels + wave * kVoxelsPerWave + j;
        vok[j] = slot < p.n_vox;
        vox[j] = vok[j] ? voxel_at(p, slot) : slot;
        xv[j] = vok[j] ? (double)p.vox[3 * vox[j] + 0] : 0.0;
        yv[j] = vok[j] ? (double)p.vox[3 * vox[j] + 1] : 0.0;
        zv[j] = vok[j] ? (double)p.vox[3 * vox[j] + 2] : 1.0;
        const double R = zv[j] / p.two_f_number;
        R2[j] = R * R;
        invR[j] = 1.0 / R;
#pragma unroll
        for (int k = 0; k < FPL; ++k) acc_re[j][k] = acc_im[j][k] = 0.f;
    }

    // Staging: the window of one element is W rows x ROW_BYTES, copied in 16-byte chunks.
    auto stage_load = [&](int64_t row0_off_bytes, int W, unsigned int (&regs)[CPT][4]) {
        const int chunks = W * (ROW_BYTES / 16);
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            const int idx = c * 256 + threadIdx.x;
            if (idx < chunks) {
                const int row = idx / (ROW_BYTES / 16), col = idx % (ROW_BYTES / 16);
                const unsigned char *g = (const unsigned char *)p.data + row0_off_bytes +
                                         (int64_t)row * p.n_frames * elem_size<ST>() + col * 16;
                const uint4 v = *(const uint4 *)g;
                regs[c][0] = v.x; regs[c][1] = v.y; regs[c][2] = v.z; regs[c][3] = v.w;
            }
        }
    };
    auto stage_store = [&](int buf, int W, const unsigned int (&regs)[CPT][4]) {
        const int chunks = W * (ROW_BYTES / 16);
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            const int idx = c * 256 + threadIdx.x;
            if (idx < chunks) *(uint4 *)(&s_win[buf][idx * 16]) = make_uint4(regs[c][0], regs[c][1], regs[c][2], regs[c][3]);
        }
    };

    for (int64_t t = 0; t < p.n_tx; ++t) {
        double Tv[kVoxelsPerWave];
#pragma unroll
        for (int j = 0; j < kVoxelsPerWave; ++j) Tv[j] = vok[j] ? (double)p.arrivals[t * p.n_vox + vox[j]] : 0.0;
        for (int64_t eb = 0; eb < p.n_rx; eb += 64) {
            // float64 setup of (voxel j, element eb + lane)
            const int64_t e = eb + lane;
            PairSetup su[kVoxelsPerWave];
            int lo = INT_MAX, hi = -1;
#pragma unroll
            for (int j = 0; j < kVoxelsPerWave; ++j) {
                su[j] = PairSetup{-1, 0.f, 0.f, 0.f};
                if (vok[j] && e < p.n_rx) su[j] = pair_setup<INTERP>(p, xv[j], yv[j], zv[j], R2[j], invR[j], Tv[j], e);
                if (su[j].i0 >= 0) {
                    lo = min(lo, su[j].i0);
                    hi = max(hi, su[j].i0);
                }
            }
            __syncthreads();  // previous block's readers of s_lo/s_hi and windows are done
            if (threadIdx.x < 64) {
                s_lo[threadIdx.x] = INT_MAX;
                s_hi[threadIdx.x] = -1;
            }
            __syncthreads();
            if (hi >= 0) {
                atomicMin(&s_lo[lane], lo);
                atomicMax(&s_hi[lane], hi);
            }
            __syncthreads();
            const int nk = (int)min<int64_t>(64, p.n_rx - eb);
            // window of element k: rows [s_lo[k], s_hi[k] + NT), staged only if it fits. The bounds are
            // workgroup-uniform; readfirstlane tells the compiler so (scalar loop control and lane index).
            auto lds_lo = [&](int k) { return __builtin_amdgcn_readfirstlane(s_lo[k]); };
            auto win_rows = [&](int k) {
                const int h = __builtin_amdgcn_readfirstlane(s_hi[k]);
                return h < 0 ? 0 : h + NT - lds_lo(k);
            };
            auto row_off = [&](int k) {
                return (((t * p.n_rx + eb + k) * p.n_samples + lds_lo(k)) * p.n_frames + f_base) * (int64_t)elem_size<ST>();
            };
            unsigned int regs[CPT][4];
            int k_next = 0;
            while (k_next < nk && win_rows(k_next) == 0) ++k_next;
            if (k_next < nk && win_rows(k_next) <= WMAX) stage_load(row_off(k_next), win_rows(k_next), regs);
            int buf = 0;
            for (int k = k_next; k < nk;) {
                const int W = win_rows(k);
                const bool staged = W <= WMAX;
                if (staged) stage_store(buf, W, regs);
                int kn = k + 1;
                while (kn < nk && win_rows(kn) == 0) ++kn;
                __syncthreads();  // window k visible; everyone is done with the other buffer
                if (kn < nk && win_rows(kn) <= WMAX) stage_load(row_off(kn), win_rows(kn), regs);
                const int lo_k = lds_lo(k);
#pragma unroll
                for (int j = 0; j < kVoxelsPerWave; ++j) {
                    const int i0 = __builtin_amdgcn_readlane(su[j].i0, k);
                    if (i0 < 0) continue;
                    const float frac = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(su[j].frac), k));
                    const float wc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(su[j].wc), k));
                    const float ws = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(su[j].ws), k));
                    if (staged) {
                        accumulate_pair<ST, INTERP, FPL>(s_win[buf], (int64_t)(i0 - lo_k) * FC + lane * FPL, FC, frac,
                                                         wc, ws, acc_re[j], acc_im[j]);
                    } else {
                        const int64_t off = ((t * p.n_rx + eb + k) * p.n_samples + i0) * p.n_frames + f0;
                        accumulate_pair<ST, INTERP, FPL>(p.data, off, p.tap_step, frac, wc, ws, acc_re[j], acc_im[j]);
                    }
                }
                if (staged) buf ^= 1;
                k = kn;
            }
        }
    }

#pragma unroll
    for (int j = 0; j < kVoxelsPerWave; ++j) {
        if (!vok[j] || f0 >= p.n_frames) continue;
        const int64_t o = vox[j] * p.n_frames + f0;
        if constexpr (ST == kStoreF32) {
            float *q = (float *)p.out + o;
#pragma unroll
            for (int k = 0; k < FPL; ++k) q[k] = p.accumulate ? q[k] + acc_re[j][k] : acc_re[j][k];
        } else {
            float *q = (float *)p.out + 2 * o;
#pragma unroll
            for (int k = 0; k < FPL; ++k) {
                q[2 * k + 0] = p.accumulate ? q[2 * k + 0] + acc_re[j][k] : acc_re[j][k];
                q[2 * k + 1] = p.accumulate ? q[2 * k + 1] + acc_im[j][k] : acc_im[j][k];
            }
        }
    }
}

template <int ST, int INTERP, int FPL>
hipError_t launch_band(const DasParams &p, hipStream_t stream) {
    const int64_t tiles = (p.n_vox + kTileVoxels - 1) / kTileVoxels;
    const int64_t chunks = p.n_frames / (64 * FPL);
    if (tiles * chunks > 0x7fffffff || chunks > 65535) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((das_band_kernel<ST, INTERP, FPL>), dim3((unsigned)tiles, (unsigned)chunks, 1), dim3(256, 1, 1),
                       0, stream, p);
    return hipGetLastError();
}

template <int ST, int INTERP>
hipError_t band_fpl(const DasParams &p, int fpl, hipStream_t stream) {
    switch (fpl) {
        case 4: return launch_band<ST, INTERP, 4>(p, stream);
        case 2: return launch_band<ST, INTERP, 2>(p, stream);
        default: return launch_band<ST, INTERP, 1>(p, stream);
    }
}

template <int ST>
hipError_t band_interp(const DasParams &p, int interp, int fpl, hipStream_t stream) {
    switch (interp) {
        case 0: return band_fpl<ST, 0>(p, fpl, stream);
        case 2: return band_fpl<ST, 2>(p, fpl, stream);
        default: return band_fpl<ST, 1>(p, fpl, stream);
    }
}

}  // namespace

int band_frames_per_lane(const DasParams &p, int store) {
    const int elem = store == kStoreC64 ? 8 : 4;
    if (p.order == nullptr || p.n_samples < 2 || p.n_vox < kTileVoxels * 64) return 0;
    // widest lane vector whose 64-lane row is <= 1 KiB, divides the frame count and stays aligned
    for (int k : {4, 2, 1}) {
        if (64 * k * elem > 1024) continue;
        if (p.n_frames % (64 * k) != 0) continue;
        if ((uintptr_t)p.data % 16 != 0 || (uintptr_t)p.out % (uintptr_t)(k * (store == kStoreF32 ? 4 : 8)) != 0)
            continue;
        return k;
    }
    return 0;
}

hipError_t launch_das_band(const DasParams &p, int store, int interp, int fpl, hipStream_t stream) {
    switch (store) {
        case kStoreF32: return band_interp<kStoreF32>(p, interp, fpl, stream);
        case kStoreC32H: return band_interp<kStoreC32H>(p, interp, fpl, stream);
        default: return band_interp<kStoreC64>(p, interp, fpl, stream);
    }
}

#ifdef MACH_DEBUG
MACH_DEBUG_READER(debug_violations_band)
#endif

}  // namespace mach
