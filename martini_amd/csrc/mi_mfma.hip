// BF16 split-precision MFMA form of the insertion contraction (in progress).
#include <hip/hip_runtime.h>

#include "mi_device.h"
#include "mi_internal.h"

namespace mi {

hipError_t gather_mfma(const mi_particle*, const Grid&, const mi_cube_desc&, char*, const Layout&,
                       float*, hipStream_t) {
    return hipErrorNotSupported;
}

}  // namespace mi
