// Plain delay-and-sum kernels for complex64 IQ channel data.
#include "das_kernel_impl.h"

namespace mach {

hipError_t launch_das_c64(const DasParams &p, int interp, DasLayout l, hipStream_t stream) {
    return launch_das_store<kStoreC64>(p, interp, l, stream);
}

#ifdef MACH_DEBUG
MACH_DEBUG_READER(debug_violations_c64)
#endif

}  // namespace mach
