/*
 * FP64 delay-and-sum oracle in C (OpenMP over voxels).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP beamformer and the timed CPU baseline. Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. It is an independent
 * restatement of the same conventions as oracle/das64.py, written loop-by-loop from the reference
 * (all paths under src/mach/):
 *   aperture R = z/(2F), skip h^2 > R^2               kernel.cu:456-463, 583
 *   tau = T[v] + |r_e - r_v| * (1/c)                  kernel.cu:466-471, 750
 *   Tukey half window when alpha > 0                  kernel.cu:333-349, 473-482
 *   skip tau < 0 or w == 0                            kernel.cu:634
 *   s = (tau - t0) * fs                               kernel.cu:637
 *   nearest / linear / quadratic interpolation        kernel.cu:106-252
 *   rotation exp(+j 2 pi f_mod tau) for IQ, f_mod!=0  kernel.cu:648-661
 *   out[v][f] += sum_t sum_e term                     kernel.cu:674-680, experimental.py:80-97
 * Shared edge decisions (see das64.py): nearest index clamped to n-1; r = 0 when h^2 == 0.
 *
 * Build with -ffp-contract=off so that h^2, R^2 and s round exactly as NumPy's separate operations.
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
    const void *channel_data; /* [T][E][S][F] (F16IQ: [T][E][S][F][2]) */
    int32_t dtype;
    int32_t interp;
    const float *rx_coords_m;  /* [E][3] */
    const float *scan_coords_m; /* [V][3] */
    const float *tx_arrivals_s; /* [T][V] */
    int64_t n_tx, n_rx, n_samples, n_frames, n_vox;
    double f_number, rx_start_s, sampling_freq_hz, sound_speed_m_s, modulation_freq_hz, tukey_alpha;
    double *out; /* [V][F] complex interleaved (re, im) for IQ, real for RF; accumulated into */
} das_oracle_args;

/* IEEE binary16 -> double, exact. */
static double half_to_double(uint16_t h) {
    int sign = (h >> 15) & 1;
    int expo = (h >> 10) & 0x1f;
    int mant = h & 0x3ff;
    double v;
    if (expo == 0) {
        v = ldexp((double)mant, -24);
    } else if (expo == 31) {
        v = mant ? NAN : INFINITY;
    } else {
        v = ldexp((double)(mant | 0x400), expo - 25);
    }
    return sign ? -v : v;
}

static inline void load_sample(const das_oracle_args *a, int64_t off, double *re, double *im) {
    if (a->dtype == DTYPE_F32) {
        *re = (double)((const float *)a->channel_data)[off];
        *im = 0.0;
    } else if (a->dtype == DTYPE_C64) {
        const float *p = (const float *)a->channel_data + 2 * off;
        *re = (double)p[0];
        *im = (double)p[1];
    } else {
        const uint16_t *p = (const uint16_t *)a->channel_data + 2 * off;
        *re = half_to_double(p[0]);
        *im = half_to_double(p[1]);
    }
}

/* round half to even, as CUDA's __float2uint_rn and numpy.rint */
static inline double rint_even(double x) { return nearbyint(x); }

int das_oracle_run(const das_oracle_args *a, int n_threads) {
    const int64_t T = a->n_tx, E = a->n_rx, S = a->n_samples, F = a->n_frames, V = a->n_vox;
    const int is_complex = a->dtype != DTYPE_F32;
    const double Fn = a->f_number, inv_c = 1.0 / a->sound_speed_m_s, alpha = a->tukey_alpha;
    const double t0 = a->rx_start_s, fs = a->sampling_freq_hz, fmod_hz = a->modulation_freq_hz;
    const int rotate = is_complex && fmod_hz != 0.0;
    const int ncomp = is_complex ? 2 : 1;
    if (T < 0 || E < 0 || S < 0 || F < 0 || V < 0) return 1;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif

#pragma omp parallel
    {
        double *acc = (double *)malloc(sizeof(double) * (size_t)(2 * (F > 0 ? F : 1)));
#pragma omp for schedule(dynamic, 16)
        for (int64_t v = 0; v < V; ++v) {
            memset(acc, 0, sizeof(double) * (size_t)(2 * F));
            const double xv = a->scan_coords_m[3 * v + 0];
            const double yv = a->scan_coords_m[3 * v + 1];
            const double zv = a->scan_coords_m[3 * v + 2];
            const double R = zv / (2.0 * Fn);
            const double R2 = R * R;
            for (int64_t t = 0; t < T; ++t) {
                const double Tv = a->tx_arrivals_s[t * V + v];
                for (int64_t e = 0; e < E; ++e) {
                    const double dx = (double)a->rx_coords_m[3 * e + 0] - xv;
                    const double dy = (double)a->rx_coords_m[3 * e + 1] - yv;
                    const double dz = (double)a->rx_coords_m[3 * e + 2] - zv;
                    const double h2 = dx * dx + dy * dy;
                    if (h2 > R2) continue;
                    const double tau = Tv + sqrt(h2 + dz * dz) * inv_c;
                    double w = 1.0;
                    if (alpha > 0.0) {
                        const double r = (h2 == 0.0) ? 0.0 : sqrt(h2) / R;
                        if (r < 0.0 || r > 1.0) w = 0.0;
                        else if (r <= 1.0 - alpha) w = 1.0;
                        else w = 0.5 - 0.5 * cos(M_PI * (1.0 - r) / alpha);
                    }
                    if (tau < 0.0 || w == 0.0) continue;
                    const double s = (tau - t0) * fs;
                    int64_t idx[3];
                    double wt[3];
                    int ntap;
                    if (a->interp == INTERP_NEAREST) {
                        if (s < -0.5 || s > (double)S - 0.5) continue;
                        int64_t k = (int64_t)rint_even(s);
                        if (k > S - 1) k = S - 1;
                        idx[0] = k; wt[0] = 1.0; ntap = 1;
                    } else if (a->interp == INTERP_LINEAR) {
                        if (s < 0.0 || s > (double)(S - 1)) continue;
                        const double lo = floor(s), hi = ceil(s);
                        idx[0] = (int64_t)lo; idx[1] = (int64_t)hi;
                        wt[0] = s - lo; /* lerp fraction; combined below as a + t (b - a) */
                        ntap = 2;
                    } else {
                        if (s < 1.0 || s > (double)S - 2.0) continue;
                        const double cc = rint_even(s);
                        const double x = s - cc;
                        idx[0] = (int64_t)cc - 1; idx[1] = (int64_t)cc; idx[2] = (int64_t)cc + 1;
                        wt[0] = 0.5 * x * (x - 1.0);
                        wt[1] = (1.0 - x) * (1.0 + x);
                        wt[2] = 0.5 * x * (x + 1.0);
                        ntap = 3;
                    }
                    double cr = w, ci = 0.0;
                    if (rotate) {
                        const double ph = 2.0 * M_PI * fmod_hz * tau;
                        cr = w * cos(ph);
                        ci = w * sin(ph);
                    }
                    const int64_t base = (t * E + e) * S;
                    for (int64_t f = 0; f < F; ++f) {
                        double vr = 0.0, vi = 0.0;
                        if (ntap == 1) {
                            load_sample(a, (base + idx[0]) * F + f, &vr, &vi);
                        } else if (ntap == 2) {
                            double ar, ai, br, bi;
                            load_sample(a, (base + idx[0]) * F + f, &ar, &ai);
                            load_sample(a, (base + idx[1]) * F + f, &br, &bi);
                            vr = ar + wt[0] * (br - ar);
                            vi = ai + wt[0] * (bi - ai);
                        } else {
                            for (int k = 0; k < 3; ++k) {
                                double xr, xi;
                                load_sample(a, (base + idx[k]) * F + f, &xr, &xi);
                                vr += wt[k] * xr;
                                vi += wt[k] * xi;
                            }
                        }
                        acc[2 * f + 0] += cr * vr - ci * vi;
                        acc[2 * f + 1] += cr * vi + ci * vr;
                    }
                }
            }
            double *o = a->out + (size_t)v * (size_t)F * (size_t)ncomp;
            for (int64_t f = 0; f < F; ++f) {
                if (is_complex) {
                    o[2 * f + 0] += acc[2 * f + 0];
                    o[2 * f + 1] += acc[2 * f + 1];
                } else {
                    o[f] += acc[2 * f + 0];
                }
            }
        }
        free(acc);
    }
    return 0;
}

int das_oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
