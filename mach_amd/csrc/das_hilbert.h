// Internal interface of the analytic-signal (RF -> IQ) kernel (das_hilbert.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mach {

constexpr int kHilbertMaxStages = 16;
constexpr int64_t kHilbertMaxSamples = 4096;
// complex float32 values in the two LDS buffers of one workgroup (64 KB, the default dynamic LDS limit)
constexpr int64_t kHilbertLdsComplex = 8192;

struct HilbertPlan {
    int32_t n_stages;
    int32_t radix[kHilbertMaxStages];  // 4, 2, 3 or 5, applied in order
};

// Radix plan of an n-point FFT; false unless n = 2^a 3^b 5^c <= kHilbertMaxSamples.
bool hilbert_plan(int64_t n, HilbertPlan *plan);
// Columns (frames) transformed together by one workgroup for n samples.
int hilbert_columns_per_block(int64_t n);
// out[t][e][s][f] = analytic signal of rf[s][e][t][f] along s, complex64 or FP16 pairs.
hipError_t launch_analytic_signal(const float *rf, int64_t n_samples, int64_t n_rx, int64_t n_tx, int64_t n_frames,
                                  void *out, bool half_out, hipStream_t stream);

}  // namespace mach
