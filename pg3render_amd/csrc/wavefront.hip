// Wavefront (queue-based) organisation of the path tracer.
#include <hip/hip_runtime.h>

#include "pg3g_internal.hpp"

namespace pg3 {

hipError_t LaunchWavefront(const SceneView &, uint32_t, uint32_t, double *, unsigned long long *, void *, hipStream_t) {
    return hipErrorNotSupported;
}

}  // namespace pg3
