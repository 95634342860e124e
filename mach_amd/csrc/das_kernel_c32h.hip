// Plain delay-and-sum kernels for FP16 IQ channel data.
#include "das_kernel_impl.h"

namespace mach {

hipError_t launch_das_c32h(const DasParams &p, int interp, DasLayout l, hipStream_t stream) {
    return launch_das_store<kStoreC32H>(p, interp, l, stream);
}

#ifdef MACH_DEBUG
MACH_DEBUG_READER(debug_violations_c32h)
#endif

}  // namespace mach
