// Device building blocks of the delay-and-sum kernels (shared by the plain and the batched kernel).
//
// Computes, for every voxel v and frame f (reference algorithm: src/mach/kernel.cu:523-682):
//   out[v][f] (+)= sum_t sum_e  w(v,e) * interp(D[t][e][:][f], s(v,e,t)) * exp(+j 2 pi f_mod tau(v,e,t))
// with the aperture test, Tukey weight, delay and interpolation bounds of kernel.cu:106-252, 333-349
// and 444-483.
//
// Work decomposition (CDNA4, wave64):
//   * A wave owns 64/LF voxels; the LF lanes of one voxel split its frames, FPL consecutive frames per
//     lane, so channel-data rows ([channel][sample][frame], frames contiguous) are read with
//     LF*FPL-frame coalesced vector loads.
//   * Per (voxel, element, transmit) geometry is computed once, in float64, by one of the LF lanes
//     ("setup"), and broadcast to the other lanes of the voxel (v_readlane for LF = 64, ds_bpermute
//     otherwise). Setup produces the first tap index, the interpolation fraction and the complex
//     coefficient w*exp(j*phase), with the phase reduced modulo one cycle in float64 before the
//     float32 hardware sine/cosine (which take revolutions). The frame loop is pure float32 FMA work
//     on registers.
//   * Transmits are summed inside the kernel (coherent compounding) and each output element has a
//     single owner lane: one read-modify-write of `out`, no atomics.
//   * The float64 setup makes every discrete decision (aperture, tau < 0, w == 0, interpolation
//     bounds, tap index) bit-identical to a float64 evaluation of the same formulas; the library is
//     built with -ffp-contract=off so products and sums round separately.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "das_kernel.h"

namespace mach {
namespace dev {

template <int N>
using fvec = float __attribute__((ext_vector_type(N)));
template <int N>
using hvec = _Float16 __attribute__((ext_vector_type(N)));

constexpr int kNearest = 0, kLinear = 1, kQuadratic = 2;

#ifdef MACH_DEBUG
// Debug builds (MACH_DEBUG): failed bounds checks of this translation unit's kernels, counted instead
// of trapping (each translation unit holds its own counter; the host reads and clears it after every
// call through a reader defined with MACH_DEBUG_READER). A failed check also makes the access safe:
// the pair is skipped, or voxel 0 is used.
static __device__ unsigned int g_debug_violations;
__device__ __forceinline__ bool debug_ok(bool cond) {
    if (!cond) atomicAdd(&g_debug_violations, 1u);
    return cond;
}
#define MACH_DEBUG_READER(name)                                                                          \
    hipError_t name(uint32_t *count) {                                                                   \
        const uint32_t zero = 0;                                                                         \
        hipError_t e = hipMemcpyFromSymbol(count, HIP_SYMBOL(dev::g_debug_violations), sizeof(uint32_t)); \
        if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(dev::g_debug_violations), &zero, sizeof(uint32_t)); \
        return e;                                                                                        \
    }
#else
__device__ __forceinline__ constexpr bool debug_ok(bool) { return true; }
#endif

struct PairSetup {
    int i0;      // first tap sample index, -1 = no contribution
    float frac;  // linear: lerp fraction in [0, 1]; quadratic: x = s - round(s) in [-0.5, 0.5]
    float wc;    // w * cos(phase)
    float ws;    // w * sin(phase)
};

// Transmit-independent part of a (voxel, element) pair, float64: aperture test, receive delay
// d / c and Tukey weight (calculateTxRxDelayAndApodization, kernel.cu:444-483, and
// tukey_apod_weight, kernel.cu:333-349).
struct PairGeom {
    double dic;  // |r_e - r_v| * (1 / c)
    float w;     // apodisation weight; 0 = no contribution (outside the aperture or zero weight)
};

__device__ __forceinline__ PairGeom pair_geom(const DasParams &p, double xv, double yv, double zv, double R2,
                                              double inv_R, int64_t e) {
    PairGeom g{0.0, 0.f};
    const double dx = (double)p.rx[3 * e + 0] - xv;
    const double dy = (double)p.rx[3 * e + 1] - yv;
    const double dz = (double)p.rx[3 * e + 2] - zv;
    const double h2 = dx * dx + dy * dy;
    if (h2 > R2) return g;  // outside the f-number aperture (kernel.cu:462)
    g.dic = sqrt(h2 + dz * dz) * p.inv_c;
    float w = 1.f;
    // Tukey weight of r = |h| / R (kernel.cu:333-349; r = 0 when h = 0): 0 for r < 0 (a voxel behind
    // the array, R < 0); 1 on the flat top r <= 1 - alpha, decided in float64 as
    // h^2 <= (1 - alpha)^2 R^2; on the taper 0.5 - 0.5 cos(pi (1 - r) / alpha), formed as
    // sin^2(pi/2 (1 - r) / alpha) (accurate near the window edge) in float32 with the hardware sine,
    // which takes revolutions: (1 - r) / (4 alpha). The weight is continuous at both ends of the
    // taper, so the float32 r changes it only by rounding. (r > 1 cannot occur: h^2 <= R^2 here.)
    if (p.apod && h2 != 0.0) {
        if (inv_R < 0.0) {
            w = 0.f;
        } else if (!(h2 <= p.flat_r2 * R2)) {
            // v_sqrt_f32 directly (1 ulp): the correctly rounded sqrtf adds ~14 instructions of
            // range scaling and a rounding fix-up, which the continuous taper does not need
            const float rr = __builtin_amdgcn_sqrtf((float)h2) * (float)inv_R;
            const float sn = __builtin_amdgcn_sinf((1.f - rr) * p.quarter_over_alpha);
            w = sn * sn;
        }
    }
    g.w = w;
    return g;
}

// Transmit arrival of voxel (x, y, z) (float64 values of its float32 coordinates) under wavefront `w`:
// float32(distance / c + offset), the value the reference's callers compute on the host and pass as
// tx_wave_arrivals_s (src/mach/wavefront.py:8-116 divided by the sound speed; src/mach/io/uff.py:156-158
// adds the per-transmit sequence delay). Returned as float64 of that float32, like a read of the array.
__device__ __forceinline__ double wave_arrival(const WaveDesc &w, double x, double y, double z) {
    double d;
    if (w.kind == kWavePlane) {
        d = ((x - w.origin[0]) * w.vec[0] + (y - w.origin[1]) * w.vec[1]) + (z - w.origin[2]) * w.vec[2];
    } else {
        const double ax = w.origin[0] - w.vec[0], ay = w.origin[1] - w.vec[1], az = w.origin[2] - w.vec[2];
        const double bx = w.vec[0] - x, by = w.vec[1] - y, bz = w.vec[2] - z;
        const double origin_focus = sqrt((ax * ax + ay * ay) + az * az);
        const double focus_point = sqrt((bx * bx + by * by) + bz * bz);
        const double so = (w.vec[2] > w.origin[2]) ? 1.0 : ((w.vec[2] < w.origin[2]) ? -1.0 : 0.0);
        const double sp = (w.vec[2] > z) ? 1.0 : ((w.vec[2] < z) ? -1.0 : 0.0);
        d = origin_focus * so - focus_point * sp;
    }
    return (double)(float)(d / w.c + w.offset);
}

// Unit phasor exp(j 2 pi f t) of a time t (float64): the phase in cycles, f t, is reduced by its
// nearest integer in float64 (an exact subtraction), rounded to float32 and passed to v_cos_f32 /
// v_sin_f32, which take revolutions. Measured on gfx950 over [-0.5, 0.5]: max abs error 1.25e-7
// (rms 3.5e-8) against float64.
struct Phasor {
    float c, s;
};

__device__ __forceinline__ Phasor phasor(double f, double t) {
    const double cyc = f * t;
    const float red = (float)(cyc - rint(cyc));
    return Phasor{__builtin_amdgcn_cosf(red), __builtin_amdgcn_sinf(red)};
}

// Transmit-dependent part: tau = T[v] + d / c, skips (kernel.cu:634), sample index, interpolation
// taps (kernel.cu:106-252) and the complex coefficient w * exp(j 2 pi f_mod tau) (kernel.cu:648-661).
// kFactored (transmit chunks): the rotation comes factored as exp(j 2 pi f_mod T[v]) (`a`, per voxel
// and transmit) times w * exp(j 2 pi f_mod d / c) (`wb`, per voxel and element), so each (pair,
// transmit) costs one complex product instead of a phase reduction and two transcendentals. The
// discrete decisions are unchanged (tau, s and the bounds stay float64); the coefficient differs
// from the direct form by rounding only (two phasors of <= 1.25e-7 error each).
template <int INTERP, bool kFactored = false>
__device__ __forceinline__ PairSetup pair_time(const DasParams &p, const PairGeom &g, double Tv, Phasor a = {1.f, 0.f},
                                               Phasor wb = {0.f, 0.f}) {
    const double tau = Tv + g.dic;
    const double s = (tau - p.t0) * p.fs;
    const double ns = (double)p.n_samples;
    // Every skip (outside the aperture: w == 0; tau < 0; taps outside the record) is one bit mask at
    // the end, the tests combined with bitwise ands of 0/1 values and the mask made opaque to the
    // compiler: as early returns, && chains or selects they became nested exec-mask branches that
    // re-initialised the result at every level (41 VALU instructions per (pair, transmit) in the
    // compounding kernel) and held the masks in scalar registers, which spilled. Values computed for
    // a skipped pair are discarded by the mask. Callers skip a wave none of whose pairs is in the
    // aperture before calling.
    unsigned ok = (unsigned)(g.w != 0.f) & (unsigned)!(tau < 0.0);
    int i0;
    float frac = 0.f;
    if constexpr (INTERP == kNearest) {
        ok &= (unsigned)!(s < -0.5) & (unsigned)!(s > ns - 0.5);
        const double k = rint(s);  // round half to even
        i0 = (int)fmin(k, ns - 1.0);  // s in (n-1, n-0.5] would round to n: clamp to the last sample
    } else if constexpr (INTERP == kLinear) {
        ok &= (unsigned)!(s < 0.0) & (unsigned)!(s > ns - 1.0);
        // first tap floor(s), clamped to max(n-2, 0), and the fraction taken from the tap: at s == n-1
        // exactly this gives taps (n-2, n-1) with t = 1, elsewhere floor(s) and s - floor(s) (one record
        // sample: tap 0, t = 0). n_samples < 2^31: tap indices are 32-bit.
        const double k = fmin(floor(s), p.lin_tap_max);
        i0 = (int)k;
        frac = (float)(s - k);
    } else {
        ok &= (unsigned)!(s < 1.0) & (unsigned)!(s > ns - 2.0);
        const double c = rint(s);
        i0 = (int)c - 1;
        frac = (float)(s - c);
    }
    float wc = g.w, ws = 0.f;
    if constexpr (kFactored) {
        // without a rotation the callers pass a = 1 and wb = w: the product is then w and +0 exactly
        wc = fmaf(a.c, wb.c, -(a.s * wb.s));
        ws = fmaf(a.c, wb.s, a.s * wb.c);
    } else if (p.rotate) {
        const Phasor ph = phasor(p.fmod, tau);
        wc = g.w * ph.c;
        ws = g.w * ph.s;
    }
    // Debug builds: every tap of a contributing pair lies in the record (taps tap_step apart; one row
    // when n_samples == 1).
    constexpr int kTaps = INTERP == kNearest ? 1 : (INTERP == kLinear ? 2 : 3);
    if (ok && !debug_ok(i0 >= 0 && (int64_t)i0 + (p.tap_step != 0 ? kTaps : 1) <= p.n_samples)) ok = 0u;
    int m = -(int)ok;  // all ones for a contributing pair
    asm volatile("" : "+v"(m));
    PairSetup r;
    r.i0 = i0 | ~m;  // -1 when skipped
    r.frac = __int_as_float(__float_as_int(frac) & m);
    r.wc = __int_as_float(__float_as_int(wc) & m);
    r.ws = __int_as_float(__float_as_int(ws) & m);
    return r;
}

// The voxel of processing slot `slot` (< n_vox): order[slot], or the slot itself without an order.
// Debug builds check that the order's entry is a voxel index.
__device__ __forceinline__ int64_t voxel_at(const DasParams &p, int64_t slot) {
    int64_t v = p.order != nullptr ? (int64_t)p.order[slot] : slot;
    if (!debug_ok(v >= 0 && v < p.n_vox)) v = 0;
    return v;
}

// Per (voxel, element, transmit) geometry in float64.
template <int INTERP>
__device__ __forceinline__ PairSetup pair_setup(const DasParams &p, double xv, double yv, double zv,
                                                double R2, double inv_R, double Tv, int64_t e) {
    const PairGeom g = pair_geom(p, xv, yv, zv, R2, inv_R, e);
    if (__ballot(g.w != 0.f) == 0ull) return PairSetup{-1, 0.f, 0.f, 0.f};  // the whole wave outside
    return pair_time<INTERP>(p, g, Tv);
}

// Value x of lane j of this lane's aligned group of LF lanes. ds_bpermute runs on the LDS crossbar, off
// the VALU: for LF = 64 it measured 3-4 % faster than v_readlane (a VALU instruction) on the Doppler
// ensemble, although the result then takes a VGPR instead of an SGPR.
template <int LF>
__device__ __forceinline__ int bcast(int x, int j, int lane) {
    if constexpr (LF == 1) {
        return x;
    } else {
        return __builtin_amdgcn_ds_bpermute(((lane & ~(LF - 1)) + j) << 2, x);
    }
}

// Bytes per stored channel-data element.
template <int ST>
constexpr int kElemBytes = ST == kStoreC64 ? 8 : 4;

// Loads FPL consecutive frames starting at byte address q.
template <int ST, int FPL>
__device__ __forceinline__ void load_at(const char *__restrict__ q, float (&re)[FPL], float (&im)[FPL]) {
    if constexpr (ST == kStoreF32) {
        if constexpr (FPL == 1) {
            re[0] = *(const float *)q;
        } else {
            const fvec<FPL> v = *(const fvec<FPL> *)q;
#pragma unroll
            for (int k = 0; k < FPL; ++k) re[k] = v[k];
        }
#pragma unroll
        for (int k = 0; k < FPL; ++k) im[k] = 0.f;
    } else if constexpr (ST == kStoreC64) {
        if constexpr (FPL == 1) {
            const fvec<2> v = *(const fvec<2> *)q;
            re[0] = v[0];
            im[0] = v[1];
        } else if constexpr (FPL == 2) {
            const fvec<4> v = *(const fvec<4> *)q;
            re[0] = v[0]; im[0] = v[1]; re[1] = v[2]; im[1] = v[3];
        } else {
            const fvec<4> a = *(const fvec<4> *)q;
            const fvec<4> b = *(const fvec<4> *)(q + 16);
            re[0] = a[0]; im[0] = a[1]; re[1] = a[2]; im[1] = a[3];
            re[2] = b[0]; im[2] = b[1]; re[3] = b[2]; im[3] = b[3];
        }
    } else {
        const hvec<2 * FPL> v = *(const hvec<2 * FPL> *)q;
#pragma unroll
        for (int k = 0; k < FPL; ++k) {
            re[k] = (float)v[2 * k];
            im[k] = (float)v[2 * k + 1];
        }
    }
}

// Loads FPL consecutive frames of one sample row. `off` is the element offset of the first frame.
template <int ST, int FPL>
__device__ __forceinline__ void load_row(const void *__restrict__ base, int64_t off, float (&re)[FPL],
                                         float (&im)[FPL]) {
    load_at<ST, FPL>((const char *)base + off * kElemBytes<ST>, re, im);
}

// 3-point Lagrange weights about the centre tap for x = s - round(s) (kernel.cu:233-240).
struct QuadWeights {
    float l0, l1, l2;
};

__device__ __forceinline__ QuadWeights quad_weights(float x) {
    return QuadWeights{0.5f * x * (x - 1.f), (1.f - x) * (1.f + x), 0.5f * x * (x + 1.f)};
}

// One element's pair setup as the lanes of a voxel receive it from the lane that computed it (slot j of
// the voxel's LF lanes), the tap index i0 already broadcast. With kPre the setup lane has formed the
// quadratic weights `pw` itself.
struct ElemCoef {
    int i0;
    float frac;
    QuadWeights qw;
    float wc, ws;
};

template <int LF, int INTERP, bool kPre>
__device__ __forceinline__ ElemCoef elem_weights(const PairSetup &s, QuadWeights pw, int j, int lane, int i0) {
    ElemCoef c{i0, 0.f, {0.f, 0.f, 0.f}, 0.f, 0.f};
    if constexpr (kPre) {
        c.qw.l0 = __int_as_float(bcast<LF>(__float_as_int(pw.l0), j, lane));
        c.qw.l1 = __int_as_float(bcast<LF>(__float_as_int(pw.l1), j, lane));
        c.qw.l2 = __int_as_float(bcast<LF>(__float_as_int(pw.l2), j, lane));
    } else if constexpr (INTERP != kNearest) {
        c.frac = __int_as_float(bcast<LF>(__float_as_int(s.frac), j, lane));
        if constexpr (INTERP == kQuadratic) c.qw = quad_weights(c.frac);
    }
    c.wc = __int_as_float(bcast<LF>(__float_as_int(s.wc), j, lane));
    c.ws = __int_as_float(bcast<LF>(__float_as_int(s.ws), j, lane));
    return c;
}

// The same with the tap index broadcast too.
template <int LF, int INTERP, bool kPre>
__device__ __forceinline__ ElemCoef elem_coef(const PairSetup &s, QuadWeights pw, int j, int lane) {
    return elem_weights<LF, INTERP, kPre>(s, pw, j, lane, bcast<LF>(s.i0, j, lane));
}

// The taps of one (element, frame group) as loaded: FP16 pairs stay packed until they are combined, so
// taps loaded ahead of their use hold half the registers.
template <int ST, int INTERP, int FPL>
struct Taps {
    static constexpr int kN = INTERP == kNearest ? 1 : (INTERP == kLinear ? 2 : 3);
    hvec<2 * FPL> h[kN];             // FP16 storage
    float re[kN][FPL], im[kN][FPL];  // float32 / complex64 storage
};

// Loads the taps at byte addresses q0, q1, q2 (consecutive samples; q1, q2 unused by nearest
// neighbour, q2 by linear).
template <int ST, int INTERP, int FPL>
__device__ __forceinline__ void load_taps(const char *q0, const char *q1, const char *q2,
                                          Taps<ST, INTERP, FPL> &t) {
    constexpr int kN = Taps<ST, INTERP, FPL>::kN;
    const char *q[3] = {q0, q1, q2};
#pragma unroll
    for (int n = 0; n < kN; ++n) {
        if constexpr (ST == kStoreC32H)
            t.h[n] = *(const hvec<2 * FPL> *)q[n];
        else
            load_at<ST, FPL>(q[n], t.re[n], t.im[n]);
    }
}

// Interpolates kN float rows (row n = sample i0 + n; re/im per frame) and adds w * exp(j phase) * value
// to the accumulators. `frac` is the lerp fraction (linear); quadratic interpolation takes its weights
// `qw`. Rows are passed as row pointers so a caller can pass a sub-window of a larger set of rows.
template <bool kComplex, int INTERP, int FPL>
__device__ __forceinline__ void combine_rows(const float *const (&r)[3], const float *const (&i)[3], float frac,
                                             QuadWeights qw, float wc, float ws, float (&acc_re)[FPL],
                                             float (&acc_im)[FPL]) {
    float vr[FPL], vi[FPL];
    if constexpr (INTERP == kNearest) {
#pragma unroll
        for (int k = 0; k < FPL; ++k) {
            vr[k] = r[0][k];
            vi[k] = i[0][k];
        }
    } else if constexpr (INTERP == kLinear) {
        // The reference's lerp a + t (b - a) (vendor/helper_math.h:1150-1152), written as
        // (a - t a) + t b: the same value up to rounding, exact at t = 0 and t = 1, and two FMAs that
        // take FP16 taps straight from the packed halves (mixed-precision FMAs, no conversions).
#pragma unroll
        for (int k = 0; k < FPL; ++k) {
            vr[k] = fmaf(frac, r[1][k], fmaf(-frac, r[0][k], r[0][k]));
            if constexpr (kComplex) vi[k] = fmaf(frac, i[1][k], fmaf(-frac, i[0][k], i[0][k]));
        }
    } else {
        // The first product is written as fmaf(l0, a, 0), which rounds like l0 * a (a zero product
        // may lose its sign, which can change nothing but the sign of an exactly zero output): with
        // FP16 storage all three taps then feed mixed-precision FMAs (v_fma_mix_f32) straight from
        // the packed halves, with no conversions.
#pragma unroll
        for (int k = 0; k < FPL; ++k) {
            vr[k] = fmaf(qw.l2, r[2][k], fmaf(qw.l1, r[1][k], fmaf(qw.l0, r[0][k], 0.f)));
            if constexpr (kComplex) vi[k] = fmaf(qw.l2, i[2][k], fmaf(qw.l1, i[1][k], fmaf(qw.l0, i[0][k], 0.f)));
        }
    }
#pragma unroll
    for (int k = 0; k < FPL; ++k) {
        if constexpr (kComplex) {
            acc_re[k] = fmaf(wc, vr[k], fmaf(-ws, vi[k], acc_re[k]));
            acc_im[k] = fmaf(wc, vi[k], fmaf(ws, vr[k], acc_im[k]));
        } else {
            acc_re[k] = fmaf(wc, vr[k], acc_re[k]);
        }
    }
}

// Interpolates loaded taps and adds w * exp(j phase) * value to the accumulators (combine_rows).
template <int ST, int INTERP, int FPL>
__device__ __forceinline__ void combine_taps(const Taps<ST, INTERP, FPL> &t, float frac, QuadWeights qw, float wc,
                                             float ws, float (&acc_re)[FPL], float (&acc_im)[FPL]) {
    constexpr int kN = Taps<ST, INTERP, FPL>::kN;
    float r[3][FPL], i[3][FPL];
#pragma unroll
    for (int n = 0; n < kN; ++n)
#pragma unroll
        for (int k = 0; k < FPL; ++k) {
            if constexpr (ST == kStoreC32H) {
                r[n][k] = (float)t.h[n][2 * k];
                i[n][k] = (float)t.h[n][2 * k + 1];
            } else {
                r[n][k] = t.re[n][k];
                i[n][k] = t.im[n][k];
            }
        }
    const float *const rp[3] = {r[0], r[1], r[2]}, *const ip[3] = {i[0], i[1], i[2]};
    combine_rows<ST != kStoreF32, INTERP, FPL>(rp, ip, frac, qw, wc, ws, acc_re, acc_im);
}

// An empty use of loaded taps (no instructions): it pins the loads ahead of this point.
template <int ST, int INTERP, int FPL>
__device__ __forceinline__ void keep_taps(const Taps<ST, INTERP, FPL> &t) {
#pragma unroll
    for (int n = 0; n < Taps<ST, INTERP, FPL>::kN; ++n) {
        if constexpr (ST == kStoreC32H) {
            asm volatile("" ::"v"(t.h[n]));
        } else {
#pragma unroll
            for (int k = 0; k < FPL; ++k) {
                asm volatile("" ::"v"(t.re[n][k]));
                if constexpr (ST == kStoreC64) asm volatile("" ::"v"(t.im[n][k]));
            }
        }
    }
}

// load_taps then combine_taps.
template <int ST, int INTERP, int FPL>
__device__ __forceinline__ void interp_accumulate(const char *q0, const char *q1, const char *q2, float frac,
                                                  QuadWeights qw, float wc, float ws, float (&acc_re)[FPL],
                                                  float (&acc_im)[FPL]) {
    Taps<ST, INTERP, FPL> t;
    load_taps<ST, INTERP, FPL>(q0, q1, q2, t);
    combine_taps<ST, INTERP, FPL>(t, frac, qw, wc, ws, acc_re, acc_im);
}

// interp_accumulate at element offset `off` of `data`, taps `step` elements apart.
template <int ST, int INTERP, int FPL>
__device__ __forceinline__ void accumulate_pair(const void *__restrict__ data, int64_t off, int64_t step,
                                                float frac, float wc, float ws, float (&acc_re)[FPL],
                                                float (&acc_im)[FPL]) {
    constexpr int kE = kElemBytes<ST>;
    const char *q0 = (const char *)data + off * kE;
    const QuadWeights qw = INTERP == kQuadratic ? quad_weights(frac) : QuadWeights{0.f, 0.f, 0.f};
    interp_accumulate<ST, INTERP, FPL>(q0, q0 + step * kE, q0 + 2 * step * kE, frac, qw, wc, ws, acc_re, acc_im);
}

// Workgroups are dispatched round-robin over the 8 XCDs (linear id i runs on XCD i % 8), each XCD with
// its own L2. Two remaps of linear ids to logical blocks:
//   group == 0: each XCD walks one contiguous eighth of the logical blocks, so neighbouring voxels
//               (which read the same channel-data rows) share an L2. With a depth-major voxel order
//               the eighths differ in depth, and work grows with depth (the f-number aperture widens),
//               so the deepest XCD finishes last.
//   group  > 0: logical blocks in runs of `group`, run r on XCD r % 8. All XCDs sweep the depth range
//               together (balanced), each taking every eighth run; consecutive runs of one XCD are
//               8 * group blocks apart, close in depth, so their sample windows still overlap in L2.
// Both are bijections on [0, n): ids in the last incomplete set of 8 * group blocks map to themselves.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t n, uint32_t group) {
    constexpr uint32_t kXcd = 8;
    const uint32_t xcd = bid % kXcd, local = bid / kXcd;
    if (group == 0) {
        const uint32_t per = n / kXcd, rem = n % kXcd;
        return (xcd < rem) ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
    }
    const uint32_t full = n / (kXcd * group) * (kXcd * group);
    if (bid >= full) return bid;
    return ((local / group) * kXcd + xcd) * group + local % group;
}

// Maximum of `x` over the wave (every lane gets it).
__device__ __forceinline__ float wave_max(float x) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x = fmaxf(x, __shfl_xor(x, m));
    return x;
}

// Raises *peak to `x` if larger. Non-negative float32 values order like their bit patterns, so an
// unsigned atomic max gives the maximum whatever the order of the updates.
__device__ __forceinline__ void peak_max(float *peak, float x) {
    atomicMax((unsigned int *)peak, __float_as_uint(x));
}

// Epilogue of one voxel's FPL frames from frame fc on (p.output): stores the signal (out (+)= y) or
// the envelope |y|, or adds sum |y|^2 to `pw` for the power image. Raises vmax to the envelope values.
template <int ST, int FPL>
__device__ __forceinline__ void emit_frames(const DasParams &p, int64_t v, int64_t fc, const float (&acc_re)[FPL],
                                            const float (&acc_im)[FPL], float &pw, float &vmax) {
    constexpr bool kComplex = ST != kStoreF32;
    if (p.output == kOutPower) {
#pragma unroll
        for (int k = 0; k < FPL; ++k) {
            pw = fmaf(acc_re[k], acc_re[k], pw);
            if constexpr (kComplex) pw = fmaf(acc_im[k], acc_im[k], pw);
        }
    } else if (p.output == kOutEnvelope) {
        float *q = (float *)p.out + v * p.n_frames + fc;
#pragma unroll
        for (int k = 0; k < FPL; ++k) {
            const float a = kComplex ? sqrtf(fmaf(acc_re[k], acc_re[k], acc_im[k] * acc_im[k])) : fabsf(acc_re[k]);
            q[k] = a;
            vmax = fmaxf(vmax, a);
        }
    } else {
        const int64_t o = v * p.n_frames + fc;
        if constexpr (ST == kStoreF32) {
            float *q = (float *)p.out + o;
#pragma unroll
            for (int k = 0; k < FPL; ++k) q[k] = p.accumulate ? q[k] + acc_re[k] : acc_re[k];
        } else {
            float *q = (float *)p.out + 2 * o;
#pragma unroll
            for (int k = 0; k < FPL; ++k) {
                q[2 * k + 0] = p.accumulate ? q[2 * k + 0] + acc_re[k] : acc_re[k];
                q[2 * k + 1] = p.accumulate ? q[2 * k + 1] + acc_im[k] : acc_im[k];
            }
        }
    }
}

// Sum of x over each aligned group of LF lanes (xor butterfly; every lane of the group gets it).
template <int LF>
__device__ __forceinline__ float lane_group_sum(float x) {
#pragma unroll
    for (int m = 1; m < LF; m <<= 1) x += __shfl_xor(x, m);
    return x;
}

// Writes voxel v's power value (out[v] (+)= pw) and returns it.
__device__ __forceinline__ float emit_power(const DasParams &p, int64_t v, float pw) {
    float *q = (float *)p.out + v;
    const float r = p.accumulate ? *q + pw : pw;
    *q = r;
    return r;
}

// Steps of LF elements whose pair geometry a transmit split's workgroup computes together (das_block).
constexpr int kGeomSteps = 32;

// LDS of das_block, in floats: the transmit factors of a chunk for lanes with several frames (4 waves x
// TC x 64 lanes, complex), then for a transmit split the partial sums (4 waves x CH x FPL complex rows
// of 64 lanes) and the shared pair geometry of kGeomSteps element steps (64 lanes each: float64 receive
// delay and the complex receive factor with the weight).
template <int TC, int CH, int FPL>
constexpr int das_lds_tx_factor_floats() {
    return (TC > 1 && FPL * CH > 1) ? 4 * TC * 64 * 2 : 0;
}
template <int CH, int FPL, int TSX>
constexpr int das_lds_partial_floats() {
    return TSX > 1 ? 4 * CH * FPL * 2 * 64 : 0;
}
template <int TSX>
constexpr int das_lds_geom_floats() {
    return TSX > 1 ? kGeomSteps * 64 * 4 : 0;
}
// at least 1
template <int TC, int CH, int FPL, int TSX>
constexpr int das_lds_floats() {
    const int n = das_lds_tx_factor_floats<TC, CH, FPL>() + das_lds_partial_floats<CH, FPL, TSX>() +
                  das_lds_geom_floats<TSX>();
    return n > 0 ? n : 1;
}

// One 256-thread block of the delay-and-sum: voxel group `bx` (4 waves x 64/LF voxels, in processing
// order) and frame block `by`. Used by the plain kernel and the batched kernel.
//
// Frames: a lane owns CH groups of FPL consecutive frames, group c starting at frame
// (by * CH + c) * (LF * FPL) + lf * FPL; a block covers LF * FPL * CH frames. The per-pair setup, its
// broadcast and the tap offsets are computed once for all CH groups.
// Transmits are taken TC at a time: the transmit-independent pair geometry (aperture, receive delay,
// apodisation) is computed once per chunk and reused for its TC transmits. The per-voxel, per-frame
// summation order is (transmit chunk, element, transmit within chunk) whatever the layout; with a
// transmit split (TSX > 1) it is (transmit range, element chunk of kGeomSteps * LF elements, transmit
// chunk, element, transmit within chunk), the ranges added in order at the end.
//
// Epilogues (p.output, uniform over the launch):
//   kOutSignal   - out[v][f] (+)= y, the reference's output.
//   kOutEnvelope - out[v][f] = |y| (float32).
//   kOutPower    - out[v] (+)= sum_f |y|^2 (float32). The block walks every frame block itself (the
//                  launch has one block row), each lane summing its frames in frame-group order; the LF
//                  lanes of a voxel are then combined by an xor butterfly, so the result is
//                  deterministic (and the same for every CH).
// With p.peak set, the maximum value written is folded into *peak.
// WIDE selects 64-bit element offsets for channel blocks of 4 GiB or more (p.bytes32 == 0); otherwise
// each tap is a wave-uniform 64-bit row base (scalar registers) plus one 32-bit per-lane byte offset,
// which the loads take directly (scalar-base + vector-offset form, no 64-bit address arithmetic).
// Transmit chunks (TC > 1) factor the rotation into a transmit factor per (voxel, transmit) and a
// receive factor per (voxel, element) (pair_time, kFactored): one complex product per (pair,
// transmit) instead of a phase reduction and two transcendentals (75-angle compound 3.47 -> 3.24 ms).
// The chunk's transmit factors live in this lane's own slots of the workgroup's LDS (`lds`, written
// and read by the same lane, so no barrier): as a register array they cost the 8-frame volume a wave
// per SIMD (FP16 20.4 -> 22.3 ms).
// TSX (tx_split_waves of the request: kTxSplitWaves or 1): the workgroup's waves all hold one voxel
// group, wave w summing transmits [n_tx w / TSX, n_tx (w + 1) / TSX) in chunks of TC. The elements are
// walked in chunks whose pair geometry the 4 waves first compute together into LDS, each pair once
// (without the split every wave would compute every pair for each of its transmit chunks); the first wave
// adds the others' partial sums in wave order through LDS and writes the result. Summation order per
// voxel: (transmit range, chunk, element, transmit).
// `lds`: das_lds_floats<TC, CH, FPL, TSX>() floats of the workgroup's LDS.
template <int ST, int INTERP, int LF, int FPL, int TC = 1, int CH = 1, bool WIDE = false, int TSX = 1>
__device__ __forceinline__ void das_block(const DasParams &p, uint32_t bx, uint32_t by, float *lds = nullptr) {
    static_assert(TSX == 1 || (TSX == 4 && TC > 1), "transmit ranges split 4 waves of a multi-transmit launch");
    constexpr int LV = 64 / LF;
    constexpr int kGroup = LF * FPL;  // frames per group across the LF lanes of a voxel
    constexpr bool kComplex = ST != kStoreF32;
    // The empty asm makes the thread index opaque, so nothing derived from it is hoisted out of a
    // caller's loop. In the persistent batched kernel (which runs every layout variant inside one
    // loop) hoisting kept the lane constants of all variants live at once: 144 instead of 122 VGPRs.
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lv = lane / LF;
    const int lf = lane % LF;
    const int tw = wave % TSX;  // transmit range of this wave (TSX > 1)
    const int64_t vgroup = (int64_t)bx * (4 / TSX) + wave / TSX;
    const int64_t tx_lo = TSX > 1 ? p.n_tx * tw / TSX : 0, tx_hi = TSX > 1 ? p.n_tx * (tw + 1) / TSX : p.n_tx;
    const int64_t slot = vgroup * LV + lv;
    const bool vok = slot < p.n_vox;
    const int64_t v = vok ? voxel_at(p, slot) : slot;

    double xv = 0.0, yv = 0.0, zv = 1.0;
    if (vok) {
        xv = (double)p.vox[3 * v + 0];
        yv = (double)p.vox[3 * v + 1];
        zv = (double)p.vox[3 * v + 2];
    }
    const double R = zv / p.two_f_number;  // aperture radius (kernel.cu:583)
    const double R2 = R * R;
    const double inv_R = 1.0 / R;

    const bool power = p.output == kOutPower;
    const int64_t n_blocks_f = power ? (p.n_frames + kGroup * CH - 1) / (kGroup * CH) : 1;
    float pw = 0.f, vmax = 0.f;
    for (int64_t fb = 0; fb < n_blocks_f; ++fb) {
        const int64_t fblock = power ? fb : (int64_t)by;
        const int64_t f0 = fblock * (kGroup * CH) + (int64_t)lf * FPL;  // first frame of group 0
        bool fok[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) fok[c] = f0 + (int64_t)c * kGroup < p.n_frames;

        float acc_re[CH][FPL], acc_im[CH][FPL];
#pragma unroll
        for (int c = 0; c < CH; ++c)
#pragma unroll
            for (int k = 0; k < FPL; ++k) acc_re[c][k] = acc_im[c][k] = 0.f;

        // pair geometry shared by a transmit split's waves (LDS table of kGeomSteps element steps)
        constexpr bool kShareGeom = TSX > 1;
        float *const geom = lds + das_lds_tx_factor_floats<TC, CH, FPL>() + das_lds_partial_floats<CH, FPL, TSX>();
        double *const G_dic = (double *)geom;                         // [kGeomSteps][64]
        float2 *const G_wb = (float2 *)(geom + kGeomSteps * 64 * 2);  // [kGeomSteps][64]
        auto tx_chunks = [&](const int64_t e_lo, const int64_t e_hi) {
        for (int64_t t0 = tx_lo; t0 < tx_hi; t0 += TC) {
            const int nt = (int)((tx_hi - t0) < TC ? (tx_hi - t0) : TC);
            // Transmit arrivals are float32 values (array entries, or wave_arrival's rounding), so they
            // are held as float32 and widened exactly where they are used: half the registers.
            float Tv[TC];
            // transmit chunks: the transmit factor of each voxel's rotation (pair_time, kFactored). Lanes
            // with several frames keep them in their LDS slots (slot tt at A_lds[tt * 64]); single-frame
            // lanes (compounding, whose timings for the chunk are formed together) in registers, which
            // measured 3 % faster there (2.16 against 2.23 ms). The values are the same either way.
            constexpr bool kFactor = TC > 1;
            constexpr bool kALds = kFactor && FPL * CH > 1;
            float2 *const A_lds = kALds ? (float2 *)lds + wave * TC * 64 + lane : nullptr;
            Phasor A_reg[kFactor && !kALds ? TC : 1];
#pragma unroll
            for (int tt = 0; tt < TC; ++tt) {
                Tv[tt] = 0.f;
                if (vok && tt < nt)
                    Tv[tt] = p.waves != nullptr ? (float)wave_arrival(p.waves[t0 + tt], xv, yv, zv)
                                                : p.arrivals[(t0 + tt) * p.n_vox + v];
                if constexpr (kFactor) {
                    const Phasor a = p.rotate ? phasor(p.fmod, (double)Tv[tt]) : Phasor{1.f, 0.f};
                    if constexpr (kALds) {
                        if (tt < nt) A_lds[tt * 64] = make_float2(a.c, a.s);
                    } else {
                        A_reg[tt] = a;
                    }
                }
            }
            auto A_at = [&](int tt) {
                if constexpr (kALds) {
                    const float2 a = A_lds[tt * 64];
                    return Phasor{a.x, a.y};
                } else if constexpr (kFactor) {
                    return A_reg[tt];
                } else {
                    return Phasor{1.f, 0.f};
                }
            };
            for (int64_t e0 = e_lo; e0 < e_hi; e0 += LF) {
                const int64_t e = e0 + lf;
                PairGeom g{0.0, 0.f};
                Phasor wb{0.f, 0.f};
                if constexpr (kShareGeom) {
                    // the factored timing uses the weight only to skip w == 0, and the receive factor
                    // w (cos, sin) (or (w, 0) without a rotation) is 0 exactly when w is: w >= 3e-14
                    // whenever it is not 0, and max(|cos|, |sin|) > 0.7
                    const int at = (int)((e0 - e_lo) / LF) * 64 + lane;
                    const float2 b = G_wb[at];
                    wb = Phasor{b.x, b.y};
                    g = PairGeom{G_dic[at], (b.x != 0.f || b.y != 0.f) ? 1.f : 0.f};
                } else {
                    if (vok && e < p.n_rx) g = pair_geom(p, xv, yv, zv, R2, inv_R, e);
                    // the receive factor of the rotation, with the apodisation weight (transmit chunks);
                    // the weight alone without a rotation
                    wb = Phasor{g.w, 0.f};
                    if constexpr (kFactor) {
                        if (p.rotate && g.w != 0.f) {
                            const Phasor b = phasor(p.fmod, g.dic);
                            wb = Phasor{g.w * b.c, g.w * b.s};
                        }
                    }
                }
                // Multi-transmit chunks whose taps for the whole chunk fit 32 VGPRs (single frames, e.g.
                // compounding: 4 x 8 dwords) load them together. Wider taps keep one transmit at a time:
                // for the 8-frame FP16 volume (8 dwords a transmit) groups of 4 cost a wave per SIMD and
                // measured 38.5 ms against 35.8.
                constexpr int kTapDwords = Taps<ST, INTERP, FPL>::kN * CH * FPL * kElemBytes<ST> / 4;
                constexpr int kTxGroup = kTapDwords * TC <= 32 ? TC : 1;
                // One lane per voxel, one transmit at a time: the lane that computes a pair also combines
                // it (no broadcast), so each transmit's timing (pair_time) is computed right before its
                // combine instead of for the whole chunk up front. The chunk's PairSetup array (4 VGPRs a
                // transmit) is never live, and the wave walks an element when some voxel has it in its
                // aperture; out-of-record and tau < 0 pairs are skipped per transmit, as below.
                constexpr bool kLazy = LF == 1 && TC > 1 && kTxGroup == 1;
                PairSetup s[kLazy ? 1 : TC];
                if constexpr (!kLazy) {
                    // a wave none of whose pairs is in the aperture skips the timings (often: elements
                    // far from the wave's voxels)
                    const bool any = __ballot(g.w != 0.f) != 0ull;
#pragma unroll
                    for (int tt = 0; tt < TC; ++tt) {
                        s[tt] = PairSetup{-1, 0.f, 0.f, 0.f};
                        if (tt < nt && any)
                            s[tt] = pair_time<INTERP, kFactor>(p, g, (double)Tv[tt], A_at(tt), wb);
                    }
                }
                // Single-transmit quadratic launches: the setup lane forms the Lagrange weights once per
                // pair and broadcasts them, instead of every lane of the voxel forming them per element
                // (the same float operations, so the same values).
                constexpr bool kPreWeights = INTERP == kQuadratic && TC == 1;
                // Two-element pipelining (below) for single-transmit launches whose taps of one element
                // fit in 24 VGPRs, so that two elements' taps in flight leave room for 4+ waves per SIMD.
                constexpr bool kPipe = TC == 1 && !WIDE && kTapDwords <= 24;
                constexpr bool kTxPipe = TC > 1 && !WIDE && kTxGroup > 1;
                QuadWeights pw_own{0.f, 0.f, 0.f};
                if constexpr (kPreWeights) pw_own = quad_weights(s[0].frac);
                // Only elements that some voxel of this wave uses under some transmit of the chunk:
                // fold the wave's ballot (bit lv * LF + j: lane lv * LF + j computed element e0 + j of
                // voxel lv) onto the LF element slots and walk the set bits with a scalar loop. The
                // predicate is the pair's validity only: the lane that computed a pair may itself own
                // no frames.
                uint64_t jm = 0ull;
                if constexpr (kLazy) {
                    jm = __ballot(g.w != 0.f);
                } else {
#pragma unroll
                    for (int tt = 0; tt < TC; ++tt) jm |= __ballot(s[tt].i0 >= 0);
                }
                if constexpr (LF < 64) {
#pragma unroll
                    for (int sh = LF; sh < 64; sh <<= 1) jm |= jm >> sh;
                    jm &= (1ull << LF) - 1ull;
                }
                // Summation order per voxel and frame: (transmit chunk, element, transmit) -- the same
                // for every layout, so all layouts (and the batched kernel) give bit-identical results.
                if constexpr (kPipe) {
                    // Two active elements per trip: the broadcasts and tap loads of both are issued
                    // before either is combined, so a wave keeps twice the gathers in flight. They are
                    // still combined element by element, in the order above.
                    constexpr int kE = kElemBytes<ST>;
                    const uint32_t fo = (uint32_t)f0 * (uint32_t)kE;
                    const char *base = (const char *)p.data + t0 * p.n_rx * p.n_samples * p.n_frames * kE;
                    const int64_t row_bytes = p.n_samples * p.n_frames * kE, step = p.tap_step * kE;
                    // Every lane loads (straight-line code keeps the compiler's wait counts independent of
                    // which lanes contribute): a lane without a contribution reads the first frames of
                    // the element's first samples, which exist whenever some pair of the element is
                    // valid, and discards them.
                    auto load_elem = [&](int j, const ElemCoef &cf, bool ok, Taps<ST, INTERP, FPL>(&t)[CH]) {
                        const char *r0 = base + (e0 + j) * row_bytes;
                        const uint32_t o = (uint32_t)cf.i0 * (uint32_t)p.n_frames * (uint32_t)kE + fo;
#pragma unroll
                        for (int c = 0; c < CH; ++c) {
                            uint32_t oc = (ok && (c == 0 || fok[c])) ? o + (uint32_t)(c * kGroup * kE) : 0u;
                            asm volatile("" : "+v"(oc));  // keeps the select on the 32-bit lane offset
                            load_taps<ST, INTERP, FPL>(r0 + oc, r0 + step + oc, r0 + 2 * step + oc, t[c]);
                        }
                    };
                    auto combine_elem = [&](const ElemCoef &cf, const Taps<ST, INTERP, FPL>(&t)[CH]) {
#pragma unroll
                        for (int c = 0; c < CH; ++c)
                            if (c == 0 || fok[c])
                                combine_taps<ST, INTERP, FPL>(t[c], cf.frac, cf.qw, cf.wc, cf.ws, acc_re[c], acc_im[c]);
                    };
                    while (jm != 0ull) {
                        const int ja = __builtin_ctzll(jm);
                        jm &= jm - 1ull;
                        const bool two = jm != 0ull;  // wave-uniform; without a second element B repeats A
                        const int jb = two ? __builtin_ctzll(jm) : ja;
                        jm &= jm - 1ull;  // no-op when jm == 0
                        const ElemCoef ca = elem_coef<LF, INTERP, kPreWeights>(s[0], pw_own, ja, lane);
                        const ElemCoef cb = elem_coef<LF, INTERP, kPreWeights>(s[0], pw_own, jb, lane);
                        const bool oka = ca.i0 >= 0 && fok[0];
                        const bool okb = two && cb.i0 >= 0 && fok[0];
                        Taps<ST, INTERP, FPL> ta[CH], tb[CH];
                        load_elem(ja, ca, oka, ta);
                        load_elem(jb, cb, okb, tb);
                        if (oka) combine_elem(ca, ta);
                        if (okb) combine_elem(cb, tb);
                        // The taps count as used at the end of the trip, so the compiler can neither sink
                        // a load into its combine (which would serialise the two elements again) nor carry
                        // an unfinished load into the next trip.
#pragma unroll
                        for (int c = 0; c < CH; ++c) {
                            keep_taps(ta[c]);
                            keep_taps(tb[c]);
                        }
                    }
                } else if constexpr (kTxPipe) {
                    // Several transmits per element (compounding): the tap loads of a group of kTxGroup
                    // transmits are all issued before any is combined, so a wave keeps that many
                    // gathers in flight instead of waiting on each transmit in turn. They are combined
                    // in transmit order. As in the two-element loop, the loads are unconditional: a
                    // lane without a contribution (or a transmit past the chunk) reads the first
                    // frames of the element's first samples under the chunk's first transmit.
                    constexpr int kE = kElemBytes<ST>;
                    const uint32_t fo = (uint32_t)f0 * (uint32_t)kE;
                    const int64_t row_bytes = p.n_samples * p.n_frames * kE, step = p.tap_step * kE;
                    while (jm != 0ull) {
                        const int j = __builtin_ctzll(jm);
                        jm &= jm - 1ull;
#pragma unroll
                        for (int g0 = 0; g0 < TC; g0 += kTxGroup) {
                            if (g0 >= nt) break;
                            Taps<ST, INTERP, FPL> tp[kTxGroup][CH];
                            ElemCoef cf[kTxGroup];
                            bool ok[kTxGroup];
#pragma unroll
                            for (int gg = 0; gg < kTxGroup; ++gg) {
                                const int tt = g0 + gg;
                                const bool live = tt < nt;  // wave-uniform
                                cf[gg] = elem_coef<LF, INTERP, kPreWeights>(s[live ? tt : 0], pw_own, j, lane);
                                ok[gg] = live && cf[gg].i0 >= 0 && fok[0];
                                const char *r0 = (const char *)p.data +
                                                 ((t0 + (live ? tt : 0)) * p.n_rx + e0 + j) * row_bytes;
                                const uint32_t o = (uint32_t)cf[gg].i0 * (uint32_t)p.n_frames * (uint32_t)kE + fo;
#pragma unroll
                                for (int c = 0; c < CH; ++c) {
                                    uint32_t oc = (ok[gg] && (c == 0 || fok[c])) ? o + (uint32_t)(c * kGroup * kE) : 0u;
                                    asm volatile("" : "+v"(oc));  // keeps the select on the 32-bit lane offset
                                    load_taps<ST, INTERP, FPL>(r0 + oc, r0 + step + oc, r0 + 2 * step + oc, tp[gg][c]);
                                }
                            }
#pragma unroll
                            for (int gg = 0; gg < kTxGroup; ++gg) {
                                if (!ok[gg]) continue;
#pragma unroll
                                for (int c = 0; c < CH; ++c)
                                    if (c == 0 || fok[c])
                                        combine_taps<ST, INTERP, FPL>(tp[gg][c], cf[gg].frac, cf[gg].qw, cf[gg].wc,
                                                                      cf[gg].ws, acc_re[c], acc_im[c]);
                            }
#pragma unroll
                            for (int gg = 0; gg < kTxGroup; ++gg)
#pragma unroll
                                for (int c = 0; c < CH; ++c) keep_taps(tp[gg][c]);  // no load sinks into its combine
                        }
                    }
                } else {
                    while (jm != 0ull) {
                        const int j = __builtin_ctzll(jm);
                        jm &= jm - 1ull;
                        const int64_t e_abs = e0 + j;
#pragma unroll
                        for (int tt = 0; tt < TC; ++tt) {
                            if (tt >= nt) break;
                            ElemCoef cf;
                            if constexpr (kLazy) {
                                const PairSetup st = pair_time<INTERP, kFactor>(p, g, (double)Tv[tt], A_at(tt), wb);
                                cf = elem_weights<LF, INTERP, false>(st, pw_own, j, lane, st.i0);
                            } else {
                                cf = elem_coef<LF, INTERP, kPreWeights>(s[tt], pw_own, j, lane);
                            }
                            if (cf.i0 < 0 || !fok[0]) continue;
                            // first element of the channel's [sample][frame] block (wave-uniform)
                            const int64_t row0 = ((t0 + tt) * p.n_rx + e_abs) * p.n_samples * p.n_frames;
                            if constexpr (!WIDE) {
                                constexpr int kE = kElemBytes<ST>;
                                const char *r0 = (const char *)p.data + row0 * kE;
                                const char *r1 = r0 + p.tap_step * kE;
                                const char *r2 = r1 + p.tap_step * kE;
                                const uint32_t o =
                                    ((uint32_t)cf.i0 * (uint32_t)p.n_frames + (uint32_t)f0) * (uint32_t)kE;
#pragma unroll
                                for (int c = 0; c < CH; ++c) {
                                    const uint32_t oc = o + (uint32_t)(c * kGroup * kE);
                                    if (c == 0 || fok[c])
                                        interp_accumulate<ST, INTERP, FPL>(r0 + oc, r1 + oc, r2 + oc, cf.frac, cf.qw,
                                                                           cf.wc, cf.ws, acc_re[c], acc_im[c]);
                                }
                            } else {
                                const int64_t off = row0 + f0 + (int64_t)cf.i0 * p.n_frames;
                                const int64_t sb = p.tap_step * kElemBytes<ST>;
#pragma unroll
                                for (int c = 0; c < CH; ++c) {
                                    if (c == 0 || fok[c]) {
                                        const char *q0 =
                                            (const char *)p.data + (off + (int64_t)c * kGroup) * kElemBytes<ST>;
                                        interp_accumulate<ST, INTERP, FPL>(q0, q0 + sb, q0 + 2 * sb, cf.frac, cf.qw,
                                                                           cf.wc, cf.ws, acc_re[c], acc_im[c]);
                                    }
                                }
                            }
                        }
                    }
                }
            }
        }

        };
        if constexpr (kShareGeom) {
            // element chunks: the 4 waves compute the chunk's kGeomSteps steps of pair geometry together
            // (wave w steps w * kGeomSteps / 4 ..), the same pairs and values the lanes would compute
            // themselves, then every wave walks its transmit chunks over them; every wave passes both
            // barriers, also where it holds no voxel
            const int64_t e_chunk = (int64_t)kGeomSteps * LF;
            for (int64_t ec = 0; ec < p.n_rx; ec += e_chunk) {
#pragma unroll
                for (int k = 0; k < kGeomSteps / 4; ++k) {
                    const int st = wave * (kGeomSteps / 4) + k;
                    const int64_t e = ec + (int64_t)st * LF + lf;
                    PairGeom g{0.0, 0.f};
                    if (vok && e < p.n_rx) g = pair_geom(p, xv, yv, zv, R2, inv_R, e);
                    Phasor wb{g.w, 0.f};
                    if (p.rotate && g.w != 0.f) {
                        const Phasor b = phasor(p.fmod, g.dic);
                        wb = Phasor{g.w * b.c, g.w * b.s};
                    }
                    G_dic[st * 64 + lane] = g.dic;
                    G_wb[st * 64 + lane] = make_float2(wb.c, wb.s);
                }
                __syncthreads();
                tx_chunks(ec, ec + e_chunk < p.n_rx ? ec + e_chunk : p.n_rx);
                __syncthreads();  // the next chunk's geometry overwrites the table
            }
        } else {
            tx_chunks(0, p.n_rx);
        }

        if constexpr (TSX > 1) {
            // the other waves' partial sums, added in wave (transmit) order by the group's first wave;
            // every wave passes both barriers, also where it holds no voxel
            constexpr int kRows = CH * FPL * 2;  // rows of 64 floats per wave
            float *const part = lds + das_lds_tx_factor_floats<TC, CH, FPL>();  // after the transmit factors
#pragma unroll
            for (int c = 0; c < CH; ++c)
#pragma unroll
                for (int k = 0; k < FPL; ++k) {
                    part[(wave * kRows + (c * FPL + k) * 2) * 64 + lane] = acc_re[c][k];
                    part[(wave * kRows + (c * FPL + k) * 2 + 1) * 64 + lane] = acc_im[c][k];
                }
            __syncthreads();
            if (tw == 0) {
                for (int w = 1; w < TSX; ++w)
#pragma unroll
                    for (int c = 0; c < CH; ++c)
#pragma unroll
                        for (int k = 0; k < FPL; ++k) {
                            acc_re[c][k] += part[((wave + w) * kRows + (c * FPL + k) * 2) * 64 + lane];
                            acc_im[c][k] += part[((wave + w) * kRows + (c * FPL + k) * 2 + 1) * 64 + lane];
                        }
            }
            __syncthreads();  // the table is rewritten for the next frame block
        }
#pragma unroll
        for (int c = 0; c < CH; ++c)
            if (vok && fok[c] && tw == 0)
                emit_frames<ST, FPL>(p, v, f0 + (int64_t)c * kGroup, acc_re[c], acc_im[c], pw, vmax);
    }

    if (power) {
        pw = lane_group_sum<LF>(pw);  // lanes lv*LF .. lv*LF+LF-1
        if (vok && lf == 0 && tw == 0) vmax = emit_power(p, v, pw);
    }
    if (p.output != kOutSignal && p.peak != nullptr) {
        const float m = wave_max(vmax);
        if (lane == 0 && m > 0.f) peak_max(p.peak, m);
    }
}

}  // namespace dev
}  // namespace mach
