// Plain delay-and-sum kernels for float32 RF channel data.
#include "das_kernel_impl.h"

namespace mach {

hipError_t launch_das_f32(const DasParams &p, int interp, DasLayout l, hipStream_t stream) {
    return launch_das_store<kStoreF32>(p, interp, l, stream);
}

#ifdef MACH_DEBUG
MACH_DEBUG_READER(debug_violations_f32)
#endif

}  // namespace mach
