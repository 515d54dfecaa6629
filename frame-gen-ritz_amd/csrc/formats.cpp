// Per-format constants of the colour passes (gpu/shaders/format.glsl): filled at create time.
#include <hip/hip_runtime.h>
#include "runtime.hpp"

bool FgrInterpolator::formatInit() { return true; }

void FgrInterpolator::formatFree() {
    if (fmtLut) hipFree(fmtLut);
    fmtLut = nullptr;
}
