// Internal interface of the LDS-staged ("band") delay-and-sum kernel for frame ensembles.
#pragma once

#include <hip/hip_runtime.h>

#include "das_kernel.h"

namespace mach {

// Frames per lane the band kernel would use for this request, or 0 if it does not apply (needs a
// depth-major voxel order, >= 2 samples, a frame count that is a multiple of 64 * frames_per_lane,
// 16-byte aligned channel data and enough voxels to fill the GPU).
int band_frames_per_lane(const DasParams &p, int store);

hipError_t launch_das_band(const DasParams &p, int store, int interp, int fpl, hipStream_t stream);

}  // namespace mach
