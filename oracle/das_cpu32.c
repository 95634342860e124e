/*
 * FP32 delay-and-sum on the CPU (OpenMP over voxels, frames innermost so the frame loop vectorises).
 *
 * CPU BASELINE ONLY: the timed CPU leg of bench.py (cpu_baseline) and its CPU test; nothing in the
 * mach_amd package loads it. It is a port of the reference kernel's float32 arithmetic (all paths
 * under src/mach/), restructured for a CPU:
 *   R = z / (2F), skip h^2 > R^2, in float32               kernel.cu:453-463, 581-583
 *   tau = T[v] + sqrt(h^2 + dz^2) * (1/c), float32         kernel.cu:466-471, 750 (inv_c = 1.0f / c)
 *   w = tukey(sqrt(h^2) / R, alpha) when alpha > 0         kernel.cu:333-349, 473-482, 749
 *   skip tau < 0 or w == 0                                 kernel.cu:634
 *   s = (tau - t0) * fs, float32                           kernel.cu:637
 *   nearest / linear (a + t (b - a)) / quadratic           kernel.cu:106-252, vendor/helper_math.h:1150-1152
 *   IQ rotation by phi = (2 pi f_mod) tau, float32         kernel.cu:571, 648-661
 *   out[v][f] = sum_t sum_e term                           kernel.cu:674-680, experimental.py:80-97
 * What differs from the GPU kernel is the schedule only: the pair's delay, weight, interpolation
 * index and sin/cos of phi (libm sincosf) are computed once per (voxel, element, transmit) and applied
 * to every frame, where the reference recomputes the rotation per frame. Shared edge choices with the
 * oracle: a nearest index at s in (n-1, n-0.5] is clamped to n-1, and r = 0 when h^2 == 0.
 * Out-of-range tap indices never occur: every read is behind the bounds test of its interpolation.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { DTYPE_F32 = 0, DTYPE_C64 = 1, DTYPE_F16IQ = 2 };
enum { INTERP_NEAREST = 0, INTERP_LINEAR = 1, INTERP_QUADRATIC = 2 };

typedef struct {
    const void *channel_data;   /* [T][E][S][F] (C64: float pairs; F16IQ: binary16 pairs) */
    int32_t dtype;
    int32_t interp;
    const float *rx_coords_m;   /* [E][3] */
    const float *scan_coords_m; /* [V][3] */
    const float *tx_arrivals_s; /* [T][V] */
    int64_t n_tx, n_rx, n_samples, n_frames, n_vox;
    float f_number, rx_start_s, sampling_freq_hz, sound_speed_m_s, modulation_freq_hz, tukey_alpha;
    float *out; /* [V][F] complex interleaved for IQ, real for RF; overwritten */
} das_cpu32_args;

static const float kPi = 3.14159265358979323846f;

/* IEEE binary16 -> float32 (exact; normals, subnormals, zeros, infinities and NaNs). */
static inline float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t expo = (h >> 10) & 0x1fu;
    const uint32_t mant = h & 0x3ffu;
    union { uint32_t u; float f; } r;
    if (expo == 0) {
        r.f = (float)mant * 5.9604644775390625e-8f; /* mant * 2^-24 */
        r.u |= sign;
    } else if (expo == 31) {
        r.u = sign | 0x7f800000u | (mant << 13);
    } else {
        r.u = sign | ((expo + 112u) << 23) | (mant << 13);
    }
    return r.f;
}

static float tukey(float r, float alpha) {
    if (r < 0.0f || r > 1.0f) return 0.0f;
    if (r <= 1.0f - alpha) return 1.0f;
    return 0.5f - 0.5f * cosf(kPi * (1.0f - r) / alpha);
}

/* Loads row `k` (sample index) of one channel as float re/im over all frames into re[], im[]. */
static void load_row(const das_cpu32_args *a, const void *chan, int64_t k, float *re, float *im) {
    const int64_t F = a->n_frames;
    if (a->dtype == DTYPE_F32) {
        const float *p = (const float *)chan + k * F;
        for (int64_t f = 0; f < F; ++f) { re[f] = p[f]; im[f] = 0.0f; }
    } else if (a->dtype == DTYPE_C64) {
        const float *p = (const float *)chan + 2 * k * F;
        for (int64_t f = 0; f < F; ++f) { re[f] = p[2 * f]; im[f] = p[2 * f + 1]; }
    } else {
        const uint16_t *p = (const uint16_t *)chan + 2 * k * F;
        for (int64_t f = 0; f < F; ++f) { re[f] = half_to_float(p[2 * f]); im[f] = half_to_float(p[2 * f + 1]); }
    }
}

int das_cpu32_run(const das_cpu32_args *a, int n_threads) {
    if (a == NULL) return 1;
    const int64_t T = a->n_tx, E = a->n_rx, S = a->n_samples, F = a->n_frames, V = a->n_vox;
    if (T < 0 || E < 0 || S < 0 || F < 0 || V < 0) return 1;
    const int cplx = a->dtype != DTYPE_F32;
    const int rotate = cplx && a->modulation_freq_hz != 0.0f;
    const int apod = a->tukey_alpha > 0.0f;
    const float inv_c = 1.0f / a->sound_speed_m_s;
    const float omega = 2.0f * kPi * a->modulation_freq_hz;
    const size_t elem = a->dtype == DTYPE_C64 ? 8 : 4; /* bytes per stored element */
    const int64_t chan_elems = S * F;
    if (F == 0 || V == 0) return 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    int status = 0;
#pragma omp parallel
    {
        /* per-thread rows: 3 taps x (re, im) + the accumulator (re, im) */
        float *buf = (float *)malloc(sizeof(float) * (size_t)F * 8);
        if (buf == NULL) {
#pragma omp atomic write
            status = 2;
        }
        float *r0 = buf, *i0 = buf + F, *r1 = buf + 2 * F, *i1 = buf + 3 * F, *r2 = buf + 4 * F, *i2 = buf + 5 * F;
        float *acc_re = buf + 6 * F, *acc_im = buf + 7 * F;
#pragma omp for schedule(dynamic, 16)
        for (int64_t v = 0; v < V; ++v) {
            if (buf == NULL) continue;
            for (int64_t f = 0; f < F; ++f) { acc_re[f] = 0.0f; acc_im[f] = 0.0f; }
            const float xv = a->scan_coords_m[3 * v], yv = a->scan_coords_m[3 * v + 1], zv = a->scan_coords_m[3 * v + 2];
            const float R = zv / (2.0f * a->f_number);
            const float R2 = R * R;
            for (int64_t t = 0; t < T; ++t) {
                const float Tv = a->tx_arrivals_s[t * V + v];
                for (int64_t e = 0; e < E; ++e) {
                    const float dx = a->rx_coords_m[3 * e] - xv;
                    const float dy = a->rx_coords_m[3 * e + 1] - yv;
                    const float dz = a->rx_coords_m[3 * e + 2] - zv;
                    const float h2 = dx * dx + dy * dy;
                    if (h2 > R2) continue;
                    const float tau = Tv + sqrtf(h2 + dz * dz) * inv_c;
                    float w = 1.0f;
                    if (apod) w = tukey(h2 == 0.0f ? 0.0f : sqrtf(h2) / R, a->tukey_alpha);
                    if (tau < 0.0f || w == 0.0f) continue;
                    const float s = (tau - a->rx_start_s) * a->sampling_freq_hz;
                    const char *chan = (const char *)a->channel_data + (size_t)((t * E + e) * chan_elems) * elem;
                    float wc = w, ws = 0.0f;
                    if (rotate) {
                        float sn, cs;
                        sincosf(omega * tau, &sn, &cs);
                        wc = w * cs;
                        ws = w * sn;
                    }
                    if (a->interp == INTERP_NEAREST) {
                        if (s < -0.5f || s > (float)S - 0.5f) continue;
                        int64_t k = (int64_t)rintf(s);
                        if (k > S - 1) k = S - 1;
                        load_row(a, chan, k, r0, i0);
                    } else if (a->interp == INTERP_LINEAR) {
                        if (s < 0.0f || s > (float)(S - 1)) continue;
                        const int64_t lo = (int64_t)floorf(s), hi = (int64_t)ceilf(s);
                        const float t_l = s - (float)lo;
                        load_row(a, chan, lo, r1, i1);
                        load_row(a, chan, hi, r2, i2);
                        for (int64_t f = 0; f < F; ++f) {
                            r0[f] = r1[f] + t_l * (r2[f] - r1[f]);
                            i0[f] = i1[f] + t_l * (i2[f] - i1[f]);
                        }
                    } else {
                        if (s < 1.0f || s > (float)S - 2.0f) continue;
                        const int64_t c = (int64_t)rintf(s);
                        const float x = s - (float)c;
                        const float l0 = 0.5f * x * (x - 1.0f), l1 = (1.0f - x) * (1.0f + x), l2 = 0.5f * x * (x + 1.0f);
                        load_row(a, chan, c - 1, r0, i0);
                        load_row(a, chan, c, r1, i1);
                        load_row(a, chan, c + 1, r2, i2);
                        for (int64_t f = 0; f < F; ++f) {
                            r0[f] = l0 * r0[f] + l1 * r1[f] + l2 * r2[f];
                            i0[f] = l0 * i0[f] + l1 * i1[f] + l2 * i2[f];
                        }
                    }
                    if (cplx) {
                        for (int64_t f = 0; f < F; ++f) {
                            acc_re[f] += wc * r0[f] - ws * i0[f];
                            acc_im[f] += wc * i0[f] + ws * r0[f];
                        }
                    } else {
                        for (int64_t f = 0; f < F; ++f) acc_re[f] += w * r0[f];
                    }
                }
            }
            if (cplx) {
                float *o = a->out + 2 * v * F;
                for (int64_t f = 0; f < F; ++f) { o[2 * f] = acc_re[f]; o[2 * f + 1] = acc_im[f]; }
            } else {
                float *o = a->out + v * F;
                for (int64_t f = 0; f < F; ++f) o[f] = acc_re[f];
            }
        }
        free(buf);
    }
    return status;
}

int das_cpu32_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
