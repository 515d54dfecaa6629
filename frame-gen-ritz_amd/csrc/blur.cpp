// recordSynthBlur with Settings::blurSamples >= 2 (gpu/framegen.cpp:3534-3634): not built yet; refused.
#include <hip/hip_runtime.h>
#include "runtime.hpp"

bool FgrInterpolator::enqueueBlur(hipStream_t, const FgrFrame&, const FgrFrame&, const FgrFrame&, float, float) {
    return false;
}
