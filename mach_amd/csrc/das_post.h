// Internal interface of the post-processing kernels (das_post.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mach {

// x[i] = k log10(a(x[i])) - k log10(a(*peak)), a(v) = |v| or float64 epsilon for v == 0, k = 10 or 20.
hipError_t launch_to_db(float *x, uint64_t n, const float *peak, bool power_mode, hipStream_t stream);
// x[0..n) = value.
hipError_t launch_fill(float *x, uint64_t n, float value, hipStream_t stream);

}  // namespace mach
