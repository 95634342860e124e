// Lateral-group delay-and-sum block: one wave beamforms VPL voxels that are consecutive in the
// processing order (with the depth-major order: lateral neighbours at one depth), for 64 * FPL frames
// (lane l holds frames l * FPL .. l * FPL + FPL - 1 of the block for every voxel of the group).
//
// Why: on the 256-frame Doppler ensemble the general kernel keeps the vector-memory address unit (TA,
// which feeds the L1) 80 % busy, because every voxel reads its own kN interpolation rows per element.
// Neighbouring voxels at one depth have nearly the same delay to an element (their first taps differ
// by at most a sample or two), so the group's taps fall in one window of kN + kSpread rows. The wave
// loads that window once and each voxel interpolates from its own rows of it: for quadratic
// interpolation 4 rows serve 2 voxels and 5 rows serve 4, where the general kernel reads 6 and 12.
// Elements whose tap starts spread wider than the window are read voxel by voxel.
//
// Everything per voxel is wave-uniform here: each voxel's tap index is read with v_readlane into a
// scalar register, so the window position and each voxel's offset in it are scalar branches (its
// weights arrive by ds_bpermute, off the VALU).
// Each voxel and frame is summed element by element with the general kernel's float operations
// (combine_rows), so the signal and envelope outputs are bit-identical to every other layout's (the power
// image sums each lane's frames first, so it matches layouts with the same frames per lane). Single-transmit
// launches only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "das_device.h"

namespace mach {
namespace dev {

// N sample rows of FPL frames as loaded; FP16 pairs stay packed until they are combined (the
// conversions then fold into mixed-precision FMAs), so a window of FP16 rows holds half the registers.
template <int ST, int N, int FPL>
struct Rows {
    hvec<2 * FPL> h[N];              // FP16 storage
    float re[N][FPL], im[N][FPL];  // float32 / complex64 storage

    __device__ __forceinline__ void load(int r, const char *q) {
        if constexpr (ST == kStoreC32H)
            h[r] = *(const hvec<2 * FPL> *)q;
        else
            load_at<ST, FPL>(q, re[r], im[r]);
    }
    __device__ __forceinline__ float real(int r, int k) const {
        if constexpr (ST == kStoreC32H) return (float)h[r][2 * k];
        return re[r][k];
    }
    __device__ __forceinline__ float imag(int r, int k) const {
        if constexpr (ST == kStoreC32H) return (float)h[r][2 * k + 1];
        return im[r][k];
    }
};

// Combines rows K .. K + 2 of w (clamped to w; only the first kN are read) into one voxel.
template <int INTERP, int K, int ST, int N, int FPL>
__device__ __forceinline__ void combine_window(const Rows<ST, N, FPL> &w, const ElemCoef &cf, float (&acc_re)[FPL],
                                               float (&acc_im)[FPL]) {
    constexpr int idx[3] = {K < N ? K : N - 1, K + 1 < N ? K + 1 : N - 1, K + 2 < N ? K + 2 : N - 1};
    float r[3][FPL], i[3][FPL];
#pragma unroll
    for (int n = 0; n < 3; ++n)
#pragma unroll
        for (int k = 0; k < FPL; ++k) {
            r[n][k] = w.real(idx[n], k);
            i[n][k] = w.imag(idx[n], k);
        }
    const float *const rp[3] = {r[0], r[1], r[2]}, *const ip[3] = {i[0], i[1], i[2]};
    combine_rows<ST != kStoreF32, INTERP, FPL>(rp, ip, cf.frac, cf.qw, cf.wc, cf.ws, acc_re, acc_im);
}

template <int ST, int INTERP, int VPL, int FPL>
__device__ __forceinline__ void das_lateral_block(const DasParams &p, uint32_t bx, uint32_t by) {
    constexpr int kN = INTERP == kNearest ? 1 : (INTERP == kLinear ? 2 : 3);  // rows per voxel
    constexpr int kSpread = VPL == 2 ? 1 : 2;  // widest spread of tap starts served by one window
    constexpr int kW = kN + kSpread;           // rows per window
    constexpr int kEC = 64 / VPL;              // elements per setup round (setup lanes per voxel)
    constexpr int kE = kElemBytes<ST>;
    constexpr int kGroup = 64 * FPL;  // frames per block
    constexpr bool kPre = INTERP == kQuadratic;  // the setup lane forms the Lagrange weights
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t first = ((int64_t)bx * 4 + wave) * VPL;  // first slot of the wave's group

    // The group's voxels (wave-uniform), for the epilogue.
    int64_t vq[VPL];
    bool okq[VPL];
#pragma unroll
    for (int q = 0; q < VPL; ++q) {
        const int64_t slot = first + q;
        okq[q] = slot < p.n_vox;
        vq[q] = okq[q] ? voxel_at(p, slot) : 0;
    }

    // Setup role of this lane: voxel q_s of the group, element slot j_s of each round of kEC elements.
    const int q_s = lane / kEC, j_s = lane % kEC;
    const int64_t slot_s = first + q_s;
    const bool vok = slot_s < p.n_vox;
    const int64_t v = vok ? voxel_at(p, slot_s) : slot_s;
    double xv = 0.0, yv = 0.0, zv = 1.0, Tv = 0.0;
    if (vok) {
        xv = (double)p.vox[3 * v + 0];
        yv = (double)p.vox[3 * v + 1];
        zv = (double)p.vox[3 * v + 2];
        Tv = p.waves != nullptr ? wave_arrival(p.waves[0], xv, yv, zv) : (double)p.arrivals[v];
    }
    const double R = zv / p.two_f_number;  // aperture radius (kernel.cu:583)
    const double R2 = R * R;
    const double inv_R = 1.0 / R;

    const int64_t row_bytes = p.n_samples * p.n_frames * kE, step = p.tap_step * kE;
    const bool power = p.output == kOutPower;
    const int64_t n_blocks_f = power ? (p.n_frames + kGroup - 1) / kGroup : 1;
    float pw[VPL], vmax = 0.f;
#pragma unroll
    for (int q = 0; q < VPL; ++q) pw[q] = 0.f;
    for (int64_t fb = 0; fb < n_blocks_f; ++fb) {
        const int64_t f0 = (power ? fb : (int64_t)by) * kGroup + (int64_t)lane * FPL;
        const bool fok = f0 < p.n_frames;
        // lanes past the last frame read the first frames of the same rows and discard them
        const uint32_t fo = fok ? (uint32_t)f0 * (uint32_t)kE : 0u;
        float acc_re[VPL][FPL], acc_im[VPL][FPL];
#pragma unroll
        for (int q = 0; q < VPL; ++q)
#pragma unroll
            for (int k = 0; k < FPL; ++k) acc_re[q][k] = acc_im[q][k] = 0.f;

        for (int64_t e0 = 0; e0 < p.n_rx; e0 += kEC) {
            const int64_t e = e0 + j_s;
            PairGeom g{0.0, 0.f};
            if (vok && e < p.n_rx) g = pair_geom(p, xv, yv, zv, R2, inv_R, e);
            // a round none of whose pairs is in the aperture skips the timing
            const PairSetup s = __ballot(g.w != 0.f) != 0ull ? pair_time<INTERP>(p, g, Tv) : PairSetup{-1, 0.f, 0.f, 0.f};
            QuadWeights pw_own{0.f, 0.f, 0.f};
            if constexpr (kPre) pw_own = quad_weights(s.frac);
            // element slots some voxel of the group uses (bit q * kEC + j: voxel q, element e0 + j)
            uint64_t jm = __ballot(s.i0 >= 0);
#pragma unroll
            for (int sh = kEC; sh < 64; sh <<= 1) jm |= jm >> sh;
            jm &= (1ull << kEC) - 1ull;
            while (jm != 0ull) {
                const int j = __builtin_ctzll(jm);
                jm &= jm - 1ull;
                const char *r0 = (const char *)p.data + (e0 + j) * row_bytes;
                int i0[VPL];
                int lo = 0x7fffffff, hi = -1;
#pragma unroll
                for (int q = 0; q < VPL; ++q) {
                    i0[q] = __builtin_amdgcn_readlane(s.i0, q * kEC + j);
                    if (i0[q] >= 0) {
                        lo = i0[q] < lo ? i0[q] : lo;
                        hi = i0[q] > hi ? i0[q] : hi;
                    }
                }
                if (hi - lo <= kSpread && p.n_samples >= kW) {
                    // one window of kW rows, placed to stay inside the record
                    const int start = lo < (int)p.n_samples - kW ? lo : (int)p.n_samples - kW;
                    // rows the group reads: voxel q reads rows i0[q] - start .. + kN - 1, so hi - start + kN
                    // of them (at most kW: hi - lo <= kSpread, and a start moved down to n_samples - kW is
                    // at most kW - kN below hi, whose taps end inside the record). Groups whose tap starts
                    // coincide or differ by one (most of them) skip the last rows' gathers; the rows left
                    // unloaded are never read.
                    const int n_rows = hi - start + kN;
                    const char *rs = r0 + (int64_t)start * step;
                    Rows<ST, kW, FPL> w;
                    // each row: the window's wave-uniform 64-bit base plus a 32-bit lane offset (frame
                    // offset + row offset, added in 32 bits, which cannot overflow: a channel's block is
                    // under 4 GiB here), which the loads take directly (scalar base + vector offset). The
                    // empty asm keeps the compiler from folding the base into 64-bit per-lane addresses.
                    uint32_t fo_l = fo;
                    asm volatile("" : "+v"(fo_l));
                    const uint32_t step32 = (uint32_t)step;
#pragma unroll
                    for (int r = 0; r < kW; ++r)
                        if (r < kN || r < n_rows) w.load(r, rs + (uint32_t)(fo_l + (uint32_t)r * step32));
#pragma unroll
                    for (int q = 0; q < VPL; ++q) {
                        if (i0[q] < 0) continue;
                        const ElemCoef cf = elem_weights<64, INTERP, kPre>(s, pw_own, q * kEC + j, lane, i0[q]);
                        const int k = i0[q] - start;
                        if (k == 0) {
                            combine_window<INTERP, 0>(w, cf, acc_re[q], acc_im[q]);
                        } else if (kSpread == 1 || k == 1) {
                            combine_window<INTERP, 1>(w, cf, acc_re[q], acc_im[q]);
                        } else {
                            combine_window<INTERP, 2>(w, cf, acc_re[q], acc_im[q]);
                        }
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < VPL; ++q) {
                        if (i0[q] < 0) continue;
                        const ElemCoef cf = elem_weights<64, INTERP, kPre>(s, pw_own, q * kEC + j, lane, i0[q]);
                        const char *rq = r0 + (int64_t)i0[q] * step;
                        Rows<ST, kN, FPL> t;
#pragma unroll
                        for (int n = 0; n < kN; ++n) t.load(n, rq + n * step + fo);
                        combine_window<INTERP, 0>(t, cf, acc_re[q], acc_im[q]);
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < VPL; ++q)
            if (okq[q] && fok) emit_frames<ST, FPL>(p, vq[q], f0, acc_re[q], acc_im[q], pw[q], vmax);
    }

    if (power) {
#pragma unroll
        for (int q = 0; q < VPL; ++q) {
            const float t = lane_group_sum<64>(pw[q]);
            if (okq[q] && lane == 0) vmax = fmaxf(vmax, emit_power(p, vq[q], t));
        }
    }
    if (p.output != kOutSignal && p.peak != nullptr) {
        const float m = wave_max(vmax);
        if (lane == 0 && m > 0.f) peak_max(p.peak, m);
    }
}

}  // namespace dev
}  // namespace mach
