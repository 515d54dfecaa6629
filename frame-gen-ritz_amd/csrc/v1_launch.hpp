// Host-side launchers of the V1 kernels (defined next to their kernels in v1_estimate.hip and v1_synth.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "v1_params.hpp"

namespace fgr {

// passes of the batched launches (launchBatch)
enum Pass : int {
    kPassLuma = 0, kPassDown, kPassMedhist, kPassSearch2, kPassSearch6, kPassSearch12, kPassMedian3, kPassSubpix,
    kPassPixfield, kPassSynthRgba8S1, kPassSynthRgba8S2, kPassSynthRgba8S4, kPassFallback,
    kPassLumaFmt, kPassSynthFmtS1, kPassSynthFmtS2, kPassSynthFmtS4,   // Rgb10 / RgbaF16 / P010 (v1_formats.hip)
    kPassSynthRgba8S1T16, kPassSynthRgba8S2T16, kPassSynthRgba8S4T16,  // RGBA8 synth, 16-row tiles
    kPassSynthRgba8S2T64,                                              // RGBA8 synth, 64-row tiles (flowScale 2)
    kPassSynthRgba8UiS1, kPassSynthRgba8UiS2, kPassSynthRgba8UiS4,     // RGBA8 synth with CrosshairV2 tap rejection
    kPassUiPatch, kPassUiFill,                                         // CrosshairV2 paste-back, estimate fill (ui.hip)
    kPassSynthFmtUiS1, kPassSynthFmtUiS2, kPassSynthFmtUiS4,           // the other formats' synth with tap rejection
    kPassSearch2x8,                                                    // R <= 2 search, two blocks a wavefront (batches)
    kPassPixfieldWin,                                                  // pixel field from LDS luma windows (exact lumas)
    kPassSubpixInt,                                                    // sub-pixel refinement in integers (exact lumas)
    kPassPyramid, kPassPyramidFmt,                                     // level 0 and the coarser levels in one launch
    kPassSynthFmtS2T16, kPassSynthFmtS2T64,                            // the other formats' synth, 16- / 64-row tiles
    kPassSeedQ                     // learned seed network layer l, quantised forward: kPassSeedQ + 2 l (+ 1 FP8)
};
// CrosshairV2 passes (ui.hip), single launches over the work rectangle
enum UiPass : int { kUiUpdate = 0, kUiMask, kUiPatch, kUiFill };
hipError_t launchUi(int pass, const UiParams& p, hipStream_t s);
hipError_t launchBatch(int pass, const void* params, const uint32_t* work, int n, hipStream_t s);
hipError_t launchSingle(int pass, const void* params, int gx, int gy, hipStream_t s);
hipError_t launchSynthSingle(int pass, const void* params, int gx, int gy, hipStream_t s);
hipError_t launchSynthBatch(int pass, const void* params, const uint32_t* work, int n, hipStream_t s);
hipError_t launchFmtSingle(int pass, const void* params, int gx, int gy, hipStream_t s);
hipError_t launchFmtBatch(int pass, const void* params, const uint32_t* work, int n, hipStream_t s);
hipError_t launchDebugEFmt(const SynthParams& p, uint32_t* dst, int eStride, hipStream_t s);
hipError_t launchSeedSingle(int pass, const void* params, int gx, int gy, hipStream_t s);
hipError_t launchSeedBatch(int pass, const void* params, const uint32_t* work, int n, hipStream_t s);
constexpr int kSubpixBlocks = 7;   // blocks per subpix wavefront
constexpr int kSubpixIntWaves = 4; // wavefronts per workgroup of the integer subpix pass (each with its own blocks)
constexpr int kSearch2Blocks = 8;  // blocks per workgroup of the batched R <= 2 search (4 for the others)
int synthTilesX(int W);
int synthTilesY(int H, int rows = 32);
int fallbackTilesX(int W);
int fallbackTilesY(int H);


hipError_t launchLookup(const LookupParams& p, hipStream_t s);
hipError_t launchDebugCounters(const uint32_t* count, const uint32_t* sad, uint32_t* dst, hipStream_t s);
hipError_t launchDebugE(const SynthParams& p, uint32_t* dst, int eStride, int format, hipStream_t s);

}  // namespace fgr
