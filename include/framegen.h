// framegen.h -- the C++ drop-in for frame-gen-ritz's framegen::Interpolator (gpu/framegen.h:443-822), MotionVersion::V1,
// on MI355X (gfx950). Header-only: every member forwards 1:1 to the C ABI of fgr.h (libfgr.so).
//
// What changes for a host of the reference: a VkCommandBuffer becomes a hipStream_t, a Frame{view, chroma} of image
// views becomes a Frame{ptr, pitch, chroma, chromaPitch} of device memory, and InitInfo names a HIP device instead of
// the Vulkan handles. Everything else keeps the reference's meaning: the Settings fields and defaults, the refusals
// (false records nothing and leaves `out` untouched, framegen.h:537-548), kSetRing records per beginFrame, the
// pacer-facing outputs (the bool of every record, recordSynthBlur, readTimings, the debug counters and maps), and
// "not thread-safe: one stream, in recording order" (framegen.h:4-6).
#ifndef FRAMEGEN_MI355X_H
#define FRAMEGEN_MI355X_H

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>
#include "fgr.h"

namespace framegen {

enum class MotionVersion : uint32_t { V1 = 0 };                                            // framegen.h:28 (V1 only)
enum class BlurWeights : uint32_t { Even = 0, Gaussian = 1 };                              // framegen.h:125
enum class Format : uint32_t { Rgba8 = 0, Rgb10 = 1, RgbaF16 = 2, Nv12 = 3, P010 = 4 };    // framegen.h:138
enum class Transfer : uint32_t { Encoded = 0, Linear = 1 };                                // framegen.h:144
enum class YuvMatrix : uint32_t { Bt601 = 0, Bt709 = 1, Bt2020 = 2 };                      // framegen.h:150
// framegen.h:15-24 (Crosshair and WholeScreen are valid values that settingsError refuses: not built)
enum class UiProtection : uint32_t { Off = 0, Crosshair = 1, WholeScreen = 2, CrosshairV2 = 3 };

// framegen.h:154-159: one caller-owned frame (device memory; Nv12 / P010: the Y plane and the CbCr plane)
struct Frame {
    void* ptr = nullptr;
    uint64_t pitch = 0;         // bytes per row
    void* chroma = nullptr;
    uint64_t chromaPitch = 0;
    FgrFrame c() const { return FgrFrame{ptr, pitch, chroma, chromaPitch}; }
};

// framegen.h:167-409: the V1 fields with the reference's defaults (seedNet: this library's learned seed proposals)
struct Settings : FgrSettings {
    Settings() { fgr_default_settings(this); }
    Settings(const FgrSettings& s) : FgrSettings(s) {}
};

// framegen.h:411-441 with a HIP device ordinal in place of the Vulkan handles
struct InitInfo {
    int32_t device = 0;
    uint32_t width = 0, height = 0;
    Format format = Format::Rgba8;
    bool bgra = false;              // Rgba8: B8G8R8A8 byte order; Rgb10: A2R10G10B10
    Transfer transfer = Transfer::Encoded;
    float linearWhiteNits = 80.0f;
    YuvMatrix yuvMatrix = YuvMatrix::Bt709;
    bool yuvFullRange = false;
    Settings settings{};
};

struct DebugMapLayout : FgrDebugMapLayout {};   // framegen.h:736-739

class Interpolator {
public:
    static constexpr uint32_t kSetRing = FGR_SET_RING;           // framegen.h:812
    static constexpr uint32_t kMinSize = FGR_MIN_SIZE;           // framegen.h:817
    static constexpr uint32_t kBlurMaxSamples = FGR_BLUR_MAX_SAMPLES;
    static constexpr uint32_t kUiSets = FGR_UI_SETS;              // framegen.h:813
    static constexpr uint32_t kUiFrames = FGR_UI_FRAMES;          // framegen.h:815
    static constexpr uint32_t kUiFramesMax = FGR_UI_FRAMES_MAX;   // framegen.h:816

    Interpolator() = default;
    Interpolator(const Interpolator&) = delete;
    Interpolator& operator=(const Interpolator&) = delete;
    ~Interpolator() { destroy(); }

    bool init(const InitInfo& ii) {   // framegen.h:446
        destroy();
        FgrInitInfo c{};
        c.width = ii.width;
        c.height = ii.height;
        c.format = (uint32_t)ii.format;
        c.bgra = ii.bgra ? 1u : 0u;
        c.transfer = (uint32_t)ii.transfer;
        c.linearWhiteNits = ii.linearWhiteNits;
        c.yuvMatrix = (uint32_t)ii.yuvMatrix;
        c.yuvFullRange = ii.yuvFullRange ? 1u : 0u;
        c.device = ii.device;
        c.settings = ii.settings;
        h_ = fgr_create(&c);
        err_ = h_ ? nullptr : fgr_last_init_error();
        return h_ != nullptr;
    }
    const char* initError() const { return err_; }
    Format format() const { return (Format)fgr_format(h_); }
    void destroy() {
        if (h_) fgr_destroy(h_);
        h_ = nullptr;
    }
    bool resize(uint32_t w, uint32_t h) { return h_ && fgr_resize(h_, w, h); }   // framegen.h:461 (same format)

    static bool validate(const Settings& s) { return fgr_validate(&s) != 0; }
    static const char* settingsError(const Settings& s, Format f) { return fgr_settings_error_for_format(&s, (uint32_t)f); }
    const char* settingsError(const Settings& s) const { return fgr_settings_error(h_, &s); }
    const char* sizeError(uint32_t w, uint32_t h) const { return fgr_size_error(h_, w, h); }
    bool setSettings(const Settings& s) { return h_ && fgr_set_settings(h_, &s); }
    bool reconfigure(const Settings& s) { return h_ && fgr_reconfigure(h_, &s); }
    Settings settings() const {
        FgrSettings s{};
        fgr_get_settings(h_, &s);
        return Settings(s);
    }

    void beginFrame() { fgr_begin_frame(h_); }   // framegen.h:511
    bool recordEstimate(hipStream_t s, Frame prev, Frame curr) {   // framegen.h:549
        return h_ && fgr_record_estimate(h_, s, prev.c(), curr.c());
    }
    bool recordSynth(hipStream_t s, Frame prev, Frame curr, Frame out, float t) {   // framegen.h:550
        return h_ && fgr_record_synth(h_, s, prev.c(), curr.c(), out.c(), t);
    }
    bool record(hipStream_t s, Frame prev, Frame curr, Frame out, float t = 0.5f) {   // framegen.h:551
        return h_ && fgr_record(h_, s, prev.c(), curr.c(), out.c(), t);
    }
    bool recordSynthBlur(hipStream_t s, Frame prev, Frame curr, Frame out, float t0, float t1) {   // framegen.h:573
        return h_ && fgr_record_synth_blur(h_, s, prev.c(), curr.c(), out.c(), t0, t1);
    }

    // static-UI protection (framegen.h:481-494, 575-693)
    static void boxSize(const Settings& s, uint32_t w, uint32_t h, uint32_t* boxW, uint32_t* boxH) {
        fgr_box_size(&s, w, h, boxW, boxH);
    }
    static void boxRect(const Settings& s, uint32_t w, uint32_t h, int32_t rect[4]) { fgr_box_rect(&s, w, h, rect); }
    static uint32_t boxMargin(const Settings& s) { return fgr_box_margin(&s); }
    static void workRect(const Settings& s, uint32_t w, uint32_t h, int32_t rect[4]) { fgr_work_rect(&s, w, h, rect); }
    bool recordObserve(hipStream_t s, Frame frame) { return h_ && fgr_record_observe(h_, s, frame.c()); }   // :610
    bool recordTrack(hipStream_t s, Frame frame) { return h_ && fgr_record_track(h_, s, frame.c()); }       // :686
    void releaseUnusedUi() { fgr_release_unused_ui(h_); }                                                 // :693

    bool setProfiling(bool on) { return h_ && fgr_set_profiling(h_, on ? 1 : 0); }   // framegen.h:705
    bool readTimings(std::vector<const char*>& names, std::vector<uint64_t>& ticks) {   // framegen.h:717
        names.assign(64, nullptr);
        ticks.assign(64, 0);
        const int n = h_ ? fgr_read_timings(h_, names.data(), ticks.data(), 64) : -1;
        if (n < 0) return false;
        names.resize((size_t)n);
        ticks.resize((size_t)n);
        return true;
    }

    bool recordDebugCounters(hipStream_t s, void* dst) { return h_ && fgr_record_debug_counters(h_, s, dst); }   // :726
    uint32_t debugSadBlocks() const { return fgr_debug_sad_blocks(h_); }                                       // :727
    bool debugMapLayout(DebugMapLayout& l) const { return h_ && fgr_debug_map_layout(h_, &l); }                 // :740
    bool recordDebugMaps(hipStream_t s, void* dst) { return h_ && fgr_record_debug_maps(h_, s, dst); }          // :741

    // this library's additions
    bool setSeedNet(FgrSeedNet* net) { return h_ && fgr_set_seed_net(h_, net); }
    FgrInterpolator* handle() const { return h_; }

private:
    FgrInterpolator* h_ = nullptr;
    const char* err_ = nullptr;
};

}  // namespace framegen

#endif
