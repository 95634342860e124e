// Element-wise kernels around the delay-and-sum: transmit arrivals from wavefront descriptors,
// logarithmic compression of the epilogue outputs to dB, and the small fills the epilogues need
// (graph-safe kernels rather than hipMemsetAsync).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "das_device.h"
#include "das_post.h"

namespace mach {
namespace {

// float64 machine epsilon, the value the reference substitutes for exact zeros before taking the
// logarithm (np.finfo(float).eps in src/mach/_vis.py:25), rounded to float32.
constexpr float kZeroFloor = 2.220446049250313e-16f;

__device__ __forceinline__ float db_of(float x, float k) {
    const float a = (x == 0.f) ? kZeroFloor : fabsf(x);
    return k * log10f(a);
}

// One thread per element, grid-stride so that any n fits the launch.
__global__ __launch_bounds__(256) void to_db_kernel(float *__restrict__ x, uint64_t n, const float *__restrict__ peak,
                                                    float k) {
    const float ref = db_of(*peak, k);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = db_of(x[i], k) - ref;
}

__global__ __launch_bounds__(256) void fill_kernel(float *__restrict__ x, uint64_t n, float value) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = value;
}

// out[t][v] = arrival of voxel v under wavefront t, the value the delay-and-sum kernel computes itself.
__global__ __launch_bounds__(256) void tx_arrivals_kernel(const WaveDesc *__restrict__ waves, uint64_t n_tx,
                                                          const float *__restrict__ scan, uint64_t n_vox,
                                                          float *__restrict__ out) {
    const uint64_t n = n_tx * n_vox;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t t = i / n_vox, v = i % n_vox;
        out[i] = (float)dev::wave_arrival(waves[t], (double)scan[3 * v], (double)scan[3 * v + 1],
                                          (double)scan[3 * v + 2]);
    }
}

unsigned grid_for(uint64_t n) {
    const uint64_t blocks = (n + 255) / 256;
    return (unsigned)(blocks < 8192 ? (blocks > 0 ? blocks : 1) : 8192);
}

}  // namespace

hipError_t launch_to_db(float *x, uint64_t n, const float *peak, bool power_mode, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(to_db_kernel, dim3(grid_for(n)), dim3(256), 0, stream, x, n, peak, power_mode ? 10.f : 20.f);
    return hipGetLastError();
}

hipError_t launch_tx_arrivals(const WaveDesc *waves, uint64_t n_tx, const float *scan, uint64_t n_vox, float *out,
                              hipStream_t stream) {
    const uint64_t n = n_tx * n_vox;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(tx_arrivals_kernel, dim3(grid_for(n)), dim3(256), 0, stream, waves, n_tx, scan, n_vox, out);
    return hipGetLastError();
}

hipError_t launch_fill(float *x, uint64_t n, float value, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, stream, x, n, value);
    return hipGetLastError();
}

}  // namespace mach
