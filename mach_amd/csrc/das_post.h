// Internal interface of the post-processing kernels (das_post.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "das_kernel.h"

namespace mach {

// x[i] = k log10(a(x[i])) - k log10(a(*peak)), a(v) = |v| or float64 epsilon for v == 0, k = 10 or 20.
hipError_t launch_to_db(float *x, uint64_t n, const float *peak, bool power_mode, hipStream_t stream);
// out[t][v] = transmit arrival of voxel v (scan [n_vox][3]) under waves[t] (device arrays).
hipError_t launch_tx_arrivals(const WaveDesc *waves, uint64_t n_tx, const float *scan, uint64_t n_vox, float *out,
                              hipStream_t stream);
// x[0..n) = value.
hipError_t launch_fill(float *x, uint64_t n, float value, hipStream_t stream);

}  // namespace mach
