// Host-side launchers of the V1 kernels (defined next to their kernels in v1_estimate.hip and v1_synth.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "v1_params.hpp"

namespace fgr {

hipError_t launchLuma(const LumaParams& p, int format, hipStream_t s);
hipError_t launchLumaRgba8(const LumaParams& p, hipStream_t s);
hipError_t launchDown(const DownParams& p, hipStream_t s);
hipError_t launchMedhist(const MedParams& p, hipStream_t s);
hipError_t launchSearch(const SearchParams& p, hipStream_t s);
hipError_t launchMedian3(const Median3Params& p, hipStream_t s);
hipError_t launchSubpix(const SubpixParams& p, hipStream_t s);
hipError_t launchPixfield(const PixParams& p, hipStream_t s);

hipError_t launchLookup(const LookupParams& p, hipStream_t s);
hipError_t launchSynth(const SynthParams& p, int format, hipStream_t s);
hipError_t launchFallback(const FallbackParams& p, int format, hipStream_t s);
hipError_t launchDebugE(const SynthParams& p, uint32_t* dst, int eStride, int format, hipStream_t s);

}  // namespace fgr
