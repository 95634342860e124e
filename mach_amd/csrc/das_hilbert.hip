// Analytic signal of real RF channel data along the sample axis, with the transpose into the
// beamformer's [tx][rx][sample][frame] layout: the device counterpart of the reference's UFF
// preprocessing (src/mach/io/uff.py:70-103, scipy.signal.hilbert along the sample axis).
//
// scipy.signal.hilbert: X = FFT(x); X[0] and (even N) X[N/2] kept, X[1 .. ceil(N/2)-1] doubled, the
// rest zeroed; x_a = IFFT(X). Here each workgroup transforms K columns (one receive element and
// transmit, K consecutive frames) held in LDS, with a mixed-radix Stockham FFT (radices 4, 2, 3, 5;
// so N must be 2^a 3^b 5^c), float32 arithmetic and twiddles from sincospi.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "das_hilbert.h"

namespace mach {
namespace {

struct Cx {
    float re, im;
};

__device__ __forceinline__ Cx cmul(Cx a, Cx b) { return Cx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Cx cadd(Cx a, Cx b) { return Cx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cx csub(Cx a, Cx b) { return Cx{a.re - b.re, a.im - b.im}; }

// In-place R-point DFT of v[0..R) with exponent sign `sg` (-1 forward, +1 inverse).
template <int R>
__device__ __forceinline__ void dft(Cx (&v)[5], float sg) {
    if constexpr (R == 2) {
        const Cx a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const Cx a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        const Cx c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        const Cx jd = Cx{-sg * d.im, sg * d.re};  // (sg * i) * d
        v[0] = cadd(a, c);
        v[2] = csub(a, c);
        v[1] = cadd(b, jd);
        v[3] = csub(b, jd);
    } else {
        Cx y[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            y[q] = Cx{0.f, 0.f};
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float s, c;
                sincospif(sg * 2.0f * (float)((q * r) % R) / (float)R, &s, &c);
                y[q] = cadd(y[q], cmul(v[r], Cx{c, s}));
            }
        }
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = y[q];
    }
}

// One Stockham stage of radix R over K columns of length N: src -> dst. Ns is the product of the
// radices of the previous stages.
template <int R>
__device__ __forceinline__ void stage(const Cx *src, Cx *dst, int N, int K, int Ns, float sg) {
    const int nb = N / R;
    for (int i = threadIdx.x; i < K * nb; i += blockDim.x) {
        const int col = i / nb, j = i % nb;
        const Cx *s = src + col * N;
        Cx *d = dst + col * N;
        const int k = j % Ns;
        Cx v[5];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Cx x = s[j + r * nb];
            if (r > 0 && k > 0) {
                float sn, cs;
                sincospif(sg * 2.0f * (float)(r * k) / (float)(Ns * R), &sn, &cs);
                x = cmul(x, Cx{cs, sn});
            }
            v[r] = x;
        }
        dft<R>(v, sg);
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) d[base + q * Ns] = v[q];
    }
}

// Runs the planned stages; returns the buffer holding the result.
__device__ Cx *fft(Cx *a, Cx *b, int N, int K, const HilbertPlan &plan, float sg) {
    int Ns = 1;
    for (int st = 0; st < plan.n_stages; ++st) {
        const int R = plan.radix[st];
        switch (R) {
            case 4: stage<4>(a, b, N, K, Ns, sg); break;
            case 2: stage<2>(a, b, N, K, Ns, sg); break;
            case 3: stage<3>(a, b, N, K, Ns, sg); break;
            default: stage<5>(a, b, N, K, Ns, sg); break;
        }
        __syncthreads();
        Ns *= R;
        Cx *t = a;
        a = b;
        b = t;
    }
    return a;
}

template <bool kHalf>
__global__ __launch_bounds__(256) void analytic_kernel(const float *__restrict__ rf, void *__restrict__ out,
                                                       int N, int E, int T, int64_t F, int K,
                                                       const HilbertPlan plan) {
    extern __shared__ Cx lds[];
    Cx *a = lds, *b = lds + (size_t)K * N;
    const int64_t n_fblocks = (F + K - 1) / K;
    const int64_t col_block = blockIdx.x;  // over (t, e, frame block)
    const int64_t fb = col_block % n_fblocks;
    const int64_t et = col_block / n_fblocks;
    const int e = (int)(et % E), t = (int)(et / E);
    const int64_t f0 = fb * K;
    // load: rf[s][e][t][f] (the reference's (n_samples, n_elements, n_waves, n_frames) order)
    for (int i = threadIdx.x; i < K * N; i += blockDim.x) {
        const int k = i % K, s = i / K;
        const int64_t f = f0 + k;
        const float x = f < F ? rf[(((int64_t)s * E + e) * T + t) * F + f] : 0.f;
        a[k * N + s] = Cx{x, 0.f};
    }
    __syncthreads();
    Cx *X = fft(a, b, N, K, plan, -1.f);
    Cx *other = (X == a) ? b : a;
    // Hilbert weights: DC and (even N) Nyquist kept, positive frequencies doubled, negative zeroed
    const int half = N / 2;
    for (int i = threadIdx.x; i < K * N; i += blockDim.x) {
        const int bin = i % N;
        float h;
        if (bin == 0 || ((N & 1) == 0 && bin == half)) h = 1.f;
        else if (bin < (N + 1) / 2) h = 2.f;
        else h = 0.f;
        X[i].re *= h;
        X[i].im *= h;
    }
    __syncthreads();
    Cx *y = fft(X, other, N, K, plan, 1.f);
    const float scale = 1.f / (float)N;
    // store: out[t][e][s][f]
    for (int i = threadIdx.x; i < K * N; i += blockDim.x) {
        const int k = i % K, s = i / K;
        const int64_t f = f0 + k;
        if (f >= F) continue;
        const Cx v = y[k * N + s];
        const int64_t o = (((int64_t)t * E + e) * N + s) * F + f;
        if constexpr (kHalf) {
            _Float16 *q = (_Float16 *)out + 2 * o;
            q[0] = (_Float16)(v.re * scale);
            q[1] = (_Float16)(v.im * scale);
        } else {
            float *q = (float *)out + 2 * o;
            q[0] = v.re * scale;
            q[1] = v.im * scale;
        }
    }
}

}  // namespace

bool hilbert_plan(int64_t n, HilbertPlan *plan) {
    plan->n_stages = 0;
    if (n < 1 || n > kHilbertMaxSamples) return false;
    int64_t m = n;
    while (m % 4 == 0 && plan->n_stages < kHilbertMaxStages) { plan->radix[plan->n_stages++] = 4; m /= 4; }
    while (m % 2 == 0 && plan->n_stages < kHilbertMaxStages) { plan->radix[plan->n_stages++] = 2; m /= 2; }
    while (m % 3 == 0 && plan->n_stages < kHilbertMaxStages) { plan->radix[plan->n_stages++] = 3; m /= 3; }
    while (m % 5 == 0 && plan->n_stages < kHilbertMaxStages) { plan->radix[plan->n_stages++] = 5; m /= 5; }
    return m == 1;
}

int hilbert_columns_per_block(int64_t n) {
    const int64_t k = kHilbertLdsComplex / 2 / n;  // two ping-pong buffers
    return (int)(k < 1 ? 1 : (k > 8 ? 8 : k));
}

hipError_t launch_analytic_signal(const float *rf, int64_t n_samples, int64_t n_rx, int64_t n_tx, int64_t n_frames,
                                  void *out, bool half_out, hipStream_t stream) {
    HilbertPlan plan;
    if (!hilbert_plan(n_samples, &plan)) return hipErrorInvalidValue;
    const int K = hilbert_columns_per_block(n_samples);
    const int64_t blocks = n_tx * n_rx * ((n_frames + K - 1) / K);
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
    const size_t lds = 2 * (size_t)K * (size_t)n_samples * sizeof(Cx);
    if (half_out) {
        hipLaunchKernelGGL(analytic_kernel<true>, dim3((unsigned)blocks), dim3(256), lds, stream, rf, out,
                           (int)n_samples, (int)n_rx, (int)n_tx, n_frames, K, plan);
    } else {
        hipLaunchKernelGGL(analytic_kernel<false>, dim3((unsigned)blocks), dim3(256), lds, stream, rf, out,
                           (int)n_samples, (int)n_rx, (int)n_tx, n_frames, K, plan);
    }
    return hipGetLastError();
}

}  // namespace mach
