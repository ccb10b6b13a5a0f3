// rms_adjoint.hip -- tangent/adjoint pass of the walk (see rms_api.cpp).
#include <hip/hip_runtime.h>

#include "rms_internal.h"

namespace rms {

hipError_t launch_adjoint(const LaunchArgs &, int, int, int, int, hipStream_t) { return hipErrorNotSupported; }
hipError_t adjoint_occupancy(int, int, int, int *blocks_per_cu)
{
    *blocks_per_cu = 1;
    return hipErrorNotSupported;
}

}  // namespace rms
