// Host runtime of the V1 path behind the C ABI (include/fgr.h): settings validation and refusals, the buffer layout,
// and the stream-ordered sequence of kernels for each record. Mirrors framegen::Interpolator::Impl
// (gpu/framegen.cpp): sizeLayout :658-685, validate :1679-1700, setSettings/reconfigure :1724-1755,
// recordEstimate :2866-3153, recordSynth :3467-3532, record :3636-3643, profiling :2041-2071, debug :2073-2159.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/fgr.h"
#include "v1_launch.hpp"
#include "v1_params.hpp"
#include "runtime.hpp"

using namespace fgr;

namespace {

thread_local const char* gInitError = nullptr;

const char* kErrSettings = "invalid settings: a field is out of range or not finite (see fgr_validate)";
const char* kErrVersion = "only MotionVersion::V1 is built in this library";
const char* kErrFormat = "unsupported format (Rgba8, Rgb10, RgbaF16, Nv12 and P010 are built)";
const char* kErrMinSize = "the frame size is below FGR_MIN_SIZE (16 pixels) in width or height";
const char* kErrPixels = "the frame size exceeds 2^30 pixels (32-bit indexing of the library's buffers)";
const char* kErrDim = "the frame size exceeds 32768 pixels in width or height";
const char* kErrDevice = "the HIP device could not be selected";
const char* kErrAlloc = "a device allocation failed";
const char* kErrNull = "null InitInfo";
const char* kErrWhite = "linearWhiteNits must be in (0, 10000]";
const char* kErrYuv = "invalid yuvMatrix";
const char* kErrOdd = "Nv12 and P010 frames need an even width and height (4:2:0 chroma of whole 2x2 blocks)";
const char* kErrUiMode = "only UiProtection Off and CrosshairV2 are built in this library";

bool in(float v, float lo, float hi) { return std::isfinite(v) && v >= lo && v <= hi; }

int rhe(double v) { return (int)std::nearbyint(v); }   // framegen.cpp:657 rintHalfEven (default rounding mode)

}  // namespace

// ------------------------------------------------------------------------------------------------------ settings

extern "C" void fgr_default_settings(FgrSettings* s) {
    // framegen.h:167-409 == interp.py:43-61
    s->trustLow = 12.0f;
    s->trustHigh = 40.0f;
    s->globalFallback = 0.15f;
    s->sceneCutSad = 30.0f;
    s->hudBonus = 1.0f;
    s->smoothBonus = 0.5f;
    s->searchPenalty = 0.5f;
    s->searchRadius = 2;
    s->searchRadiusTop = 6;
    s->subpixel = 1;
    s->motionVersion = FGR_MOTION_V1;
    s->blurSamples = 0;
    s->blurWeights = 1;
    s->flowScale = 2;
    s->pyramidMinSide = 48;
    s->seedNet = 0;
    // framegen.h:225-305
    s->uiProtection = FGR_UI_OFF;
    s->uiHistory = FGR_UI_FRAMES;
    s->uiBoxHeightFrac = 0.025f;
    s->uiBoxWidth = 0;
    s->uiBoxHeight = 0;
    s->uiCrosshairArea = 0.03f;
    s->uiBoxOffsetX = 0;
    s->uiBoxOffsetY = 0;
    s->uiStillFrames = 8;
    s->uiStillThreshold = 3.5f;
    s->uiInpaintReach = 12;
    s->uiMaxStillFraction = 0.75f;
    s->uiContrastThreshold = 20.0f;
    s->uiLeaveFrames = 4;
    s->uiReentryFrames = 2;
    s->uiEstimateFill = 0;
}

extern "C" int fgr_validate(const FgrSettings* s) {
    // framegen.cpp:1679-1700, the V1 fields (motionVersion <= V3 is a valid value; only V1 is built: settings_error)
    if (!s) return 0;
    return in(s->trustLow, 0.0f, 255.0f) && in(s->trustHigh, 0.0f, 255.0f) && s->trustHigh - s->trustLow >= 1.0f &&
           in(s->globalFallback, 0.0f, 1.0f) && s->globalFallback > 0.0f && in(s->sceneCutSad, 0.0f, 255.0f) &&
           s->sceneCutSad > 0.0f && in(s->hudBonus, 0.0f, 255.0f) && in(s->smoothBonus, 0.0f, 255.0f) &&
           in(s->searchPenalty, 0.0f, 255.0f) && s->searchRadius >= 1 && s->searchRadius <= 12 &&
           s->searchRadiusTop >= 1 && s->searchRadiusTop <= 12 && s->subpixel <= 1 &&
           (s->flowScale == 1 || s->flowScale == 2 || s->flowScale == 4) && s->pyramidMinSide >= 16 &&
           s->pyramidMinSide <= 256 && s->motionVersion <= 3 && s->blurSamples <= FGR_BLUR_MAX_SAMPLES &&
           s->blurWeights <= 1 && s->seedNet <= 1 && s->uiProtection <= FGR_UI_CROSSHAIR_V2 &&
           in(s->uiCrosshairArea, 0.0f, 1.0f) && s->uiCrosshairArea > 0.0f && in(s->uiBoxHeightFrac, 0.0f, 1.0f) &&
           s->uiBoxWidth <= 16384 && s->uiBoxHeight <= 16384 && s->uiBoxOffsetX >= -16384 && s->uiBoxOffsetX <= 16384 &&
           s->uiBoxOffsetY >= -16384 && s->uiBoxOffsetY <= 16384 && s->uiStillFrames >= 1 && s->uiStillFrames <= 255 &&
           in(s->uiStillThreshold, 0.0f, 255.0f) && s->uiInpaintReach >= 1 && s->uiInpaintReach <= 16 &&
           in(s->uiMaxStillFraction, 0.0f, 1.0f) && s->uiMaxStillFraction > 0.0f &&
           in(s->uiContrastThreshold, 0.0f, 255.0f) && s->uiLeaveFrames >= 1 && s->uiLeaveFrames <= 255 &&
           s->uiReentryFrames <= 255 && s->uiEstimateFill <= 1 && s->uiHistory >= FGR_UI_FRAMES &&
           s->uiHistory <= FGR_UI_FRAMES_MAX;
}

bool formatBuilt(uint32_t f) {
    return f == FGR_FORMAT_RGBA8 || f == FGR_FORMAT_RGB10 || f == FGR_FORMAT_RGBAF16 || f == FGR_FORMAT_NV12 ||
           f == FGR_FORMAT_P010;
}

extern "C" const char* fgr_settings_error_for_format(const FgrSettings* s, uint32_t format) {
    // framegen.cpp:1702-1707
    if (!fgr_validate(s)) return kErrSettings;
    if (s->motionVersion != FGR_MOTION_V1) return kErrVersion;
    if (s->uiProtection == FGR_UI_CROSSHAIR || s->uiProtection == FGR_UI_WHOLESCREEN) return kErrUiMode;
    return nullptr;
}

extern "C" int fgr_size_layout(uint32_t W, uint32_t H, uint32_t flowScale, uint32_t minSide, FgrSizeLayout* out) {
    if (!out || (flowScale != 1 && flowScale != 2 && flowScale != 4) || W < 1 || H < 1) return 0;
    std::memset(out, 0, sizeof(*out));
    out->fw = rhe(W * (1.0 / flowScale));
    out->fh = rhe(H * (1.0 / flowScale));
    int w = out->fw, h = out->fh, n = 0;
    for (;;) {
        out->w[n] = w;
        out->h[n] = h;
        out->bw[n] = (w + 7) / 8;
        out->bh[n] = (h + 7) / 8;
        n++;
        if ((uint32_t)(std::min(w, h) / 2) < minSide || n == 16) break;
        w = rhe(w * 0.5);
        h = rhe(h * 0.5);
    }
    out->levels = n;
    return 1;
}

extern "C" uint32_t fgr_version(void) { return 0x00010000u; }

const char* sizeLimitError(uint32_t W, uint32_t H, uint32_t format) {
    // framegen.cpp:760-790 (the device limits here: 32-bit indexing and the launch grid)
    if (W < FGR_MIN_SIZE || H < FGR_MIN_SIZE) return kErrMinSize;
    if ((uint64_t)W * H > (1ull << 30)) return kErrPixels;
    if (W > 32768 || H > 32768) return kErrDim;
    if ((format == FGR_FORMAT_P010 || format == FGR_FORMAT_NV12) && ((W | H) & 1)) return kErrOdd;
    return nullptr;
}

// ------------------------------------------------------------------------------------------------------ handle

void FgrInterpolator::freeBuffers() {
    if (luma) hipFree(luma);
    if (field) hipFree(field);
    if (u) hipFree(u);
    if (seedAct) hipFree(seedAct);
    if (blurScratch) hipFree(blurScratch);
    for (uint8_t* p : blurRetired) hipFree(p);
    blurRetired.clear();
    blurScratch = nullptr;
    blurCap = 0;
    luma = nullptr;
    field = nullptr;
    u = nullptr;
    seedAct = nullptr;
    uiFree();
    W = H = 0;
    haveEst = false;
}

void FgrInterpolator::layoutBuffers(uint32_t w, uint32_t h) {
    FgrSizeLayout L;
    fgr_size_layout(w, h, set.flowScale, set.pyramidMinSide, &L);
    fw = L.fw;
    fh = L.fh;
    lv.clear();
    int lumaOff = 0, fieldOff = 0;
    for (int i = 0; i < L.levels; i++) {
        Level l;
        l.w = L.w[i];
        l.h = L.h[i];
        l.bw = L.bw[i];
        l.bh = L.bh[i];
        l.offA = lumaOff;
        l.offC = lumaOff + l.w * l.h;
        lumaOff += 2 * l.w * l.h;
        l.fS = fieldOff;
        l.fM = fieldOff + l.bw * l.bh;
        fieldOff += 2 * l.bw * l.bh;
        lv.push_back(l);
    }
    fieldSub = fieldOff;
    fieldOff += lv[0].bw * lv[0].bh;
    fieldSeed = fieldOff;   // learned seed proposals (one per block of the largest level that can take them)
    fieldOff += lv[0].bw * lv[0].bh;
    fieldPF = fieldOff;
    fieldOff += fw * fh;
    fieldT = fieldOff;
    fieldOff += fw * fh;
    lumaWords = lumaOff;
    fieldEntries = fieldOff;
}

static inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

size_t FgrInterpolator::bufferBytes() const {
    return align256((size_t)lumaWords * 4) + align256((size_t)fieldEntries * 8) + (size_t)kUWords * 4;
}

size_t FgrInterpolator::bufferBytesNoDebug() const {
    return align256((size_t)lumaWords * 4) + align256((size_t)fieldT * 8) + (size_t)kUWords * 4;
}

bool FgrInterpolator::allocBuffers(uint32_t w, uint32_t h) {
    layoutBuffers(w, h);
    if (hipMalloc(&luma, (size_t)lumaWords * 4) != hipSuccess) return false;
    if (hipMalloc(&field, (size_t)fieldEntries * 8) != hipSuccess) return false;
    if (hipMalloc(&u, kUWords * 4) != hipSuccess) return false;
    if (hipMemset(u, 0, kUWords * 4) != hipSuccess) return false;
    if (hipMemset(field, 0, (size_t)fieldEntries * 8) != hipSuccess) return false;
    W = w;
    H = h;
    return true;
}

bool FgrInterpolator::rebuild(uint32_t w, uint32_t h) {
    if (hipSetDevice(device) != hipSuccess) return false;
    freeBuffers();
    if (!allocBuffers(w, h) || !allocSeedBuffers()) {
        freeBuffers();
        return false;
    }
    return true;
}

extern "C" FgrInterpolator* fgr_create(const FgrInitInfo* info) {
    gInitError = nullptr;
    if (!info) { gInitError = kErrNull; return nullptr; }
    if (!formatBuilt(info->format)) { gInitError = kErrFormat; return nullptr; }
    if (const char* e = fgr_settings_error_for_format(&info->settings, info->format)) { gInitError = e; return nullptr; }
    if (const char* e = sizeLimitError(info->width, info->height, info->format)) { gInitError = e; return nullptr; }
    if (info->format == FGR_FORMAT_RGBAF16 && info->transfer == FGR_TRANSFER_LINEAR &&
        !(info->linearWhiteNits > 0.0f && info->linearWhiteNits <= 10000.0f)) { gInitError = kErrWhite; return nullptr; }
    if ((info->format == FGR_FORMAT_P010 || info->format == FGR_FORMAT_NV12) && info->yuvMatrix > FGR_YUV_BT2020) {
        gInitError = kErrYuv;
        return nullptr;
    }
    if (hipSetDevice(info->device) != hipSuccess) { gInitError = kErrDevice; return nullptr; }
    FgrInterpolator* h = new FgrInterpolator();
    h->info = *info;
    h->set = info->settings;
    h->device = info->device;
    h->format = (int)info->format;
    if (!h->formatInit() || !h->rebuild(info->width, info->height) || !h->uiEnsure() || !h->blurEnsure()) {
        gInitError = kErrAlloc;
        fgr_destroy(h);
        return nullptr;
    }
    return h;
}

extern "C" const char* fgr_last_init_error(void) { return gInitError; }

extern "C" void fgr_destroy(FgrInterpolator* h) {
    if (!h) return;
    hipSetDevice(h->device);
    h->freeBuffers();
    h->formatFree();
    for (hipEvent_t e : h->events) hipEventDestroy(e);
    delete h;
}

extern "C" uint32_t fgr_format(const FgrInterpolator* h) { return h ? (uint32_t)h->format : FGR_FORMAT_RGBA8; }

extern "C" int fgr_resize(FgrInterpolator* h, uint32_t width, uint32_t height) {
    if (!h) return 0;
    if (sizeLimitError(width, height, (uint32_t)h->format)) {
        hipSetDevice(h->device);
        h->freeBuffers();   // framegen.h:457-458: on false the object holds no size until a later resize succeeds
        return 0;
    }
    if (!h->rebuild(width, height)) return 0;
    if (!h->uiEnsure() || !h->blurEnsure()) {   // (as a failed allocation of the size-dependent buffers: no size held)
        h->freeBuffers();
        return 0;
    }
    return 1;
}

extern "C" const char* fgr_settings_error(const FgrInterpolator* h, const FgrSettings* s) {
    if (const char* e = fgr_settings_error_for_format(s, h ? (uint32_t)h->format : FGR_FORMAT_RGBA8)) return e;
    return h && h->W ? sizeLimitError(h->W, h->H, (uint32_t)h->format) : nullptr;
}

extern "C" const char* fgr_size_error(const FgrInterpolator* h, uint32_t w, uint32_t hgt) {
    if (!h) return "not initialised";
    return sizeLimitError(w, hgt, (uint32_t)h->format);
}

extern "C" int fgr_set_settings(FgrInterpolator* h, const FgrSettings* s) {
    // framegen.cpp:1724-1740: runtime fields only; structural ones need reconfigure
    if (!h || fgr_settings_error(h, s) || s->flowScale != h->set.flowScale || s->pyramidMinSide != h->set.pyramidMinSide)
        return 0;
    // UI protection may allocate here (never frees); a change of the UI mode or its box drops the estimate
    const FgrSettings old = h->set;
    const UiRect oldBox = h->uiBox(), oldWork = h->uiWork();
    h->set = *s;
    if (!h->uiEnsure() || !h->blurEnsure()) {
        h->set = old;
        return 0;
    }
    h->uiAdopted(old);
    if (s->motionVersion != old.motionVersion || h->uiChanged(old, oldBox, oldWork)) h->haveEst = false;
    return 1;
}

extern "C" int fgr_reconfigure(FgrInterpolator* h, const FgrSettings* s) {
    // framegen.cpp:1742-1755
    if (!h || fgr_settings_error(h, s)) return 0;
    const bool structural = s->flowScale != h->set.flowScale || s->pyramidMinSide != h->set.pyramidMinSide;
    const FgrSettings old = h->set;
    const UiRect oldBox = h->uiBox(), oldWork = h->uiWork();
    h->set = *s;
    if (structural && h->W) {
        // (buffers gone: the object holds no size, as a failed resize)
        if (!h->rebuild(h->W, h->H)) return 0;
        if (!h->uiEnsure() || !h->blurEnsure()) {
            h->freeBuffers();
            return 0;
        }
        return 1;
    }
    if (!h->uiEnsure() || !h->blurEnsure()) {
        h->set = old;
        return 0;
    }
    h->uiAdopted(old);
    if (s->motionVersion != old.motionVersion || h->uiChanged(old, oldBox, oldWork)) h->haveEst = false;
    return 1;
}

extern "C" void fgr_get_settings(const FgrInterpolator* h, FgrSettings* out) {
    if (!out) return;
    if (h) *out = h->set;
    else fgr_default_settings(out);
}

extern "C" void fgr_begin_frame(FgrInterpolator* h) {
    if (h) h->ringEst = h->ringSynth = h->uiSets = 0;
}

// ------------------------------------------------------------------------------------------------------ profiling

extern "C" int fgr_set_profiling(FgrInterpolator* h, int on) {
    if (!h) return 0;
    h->profiling = false;
    if (on && h->events.empty()) {
        hipSetDevice(h->device);
        for (int i = 0; i < FgrInterpolator::kMaxQ; i++) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return 0;
            h->events.push_back(e);
        }
    }
    h->profiling = on != 0;
    return 1;
}

void FgrInterpolator::stamp(hipStream_t s, const char* name) {
    if (!profiling || events.empty() || qcount >= kMaxQ) return;
    hipEventRecord(events[qcount], s);
    qnames.push_back(name);
    qcount++;
}

void FgrInterpolator::beginProfile(hipStream_t s) {
    qcount = 0;
    qnames.clear();
    if (profiling) stamp(s, "start");
}

extern "C" int fgr_read_timings(FgrInterpolator* h, const char** names, uint64_t* ticks, int cap) {
    if (!h || h->qcount < 2) return -1;
    if (hipEventSynchronize(h->events[h->qcount - 1]) != hipSuccess) return -1;
    const int n = h->qcount - 1;
    for (int i = 0; i < n && i < cap; i++) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, h->events[i], h->events[i + 1]);
        if (names) names[i] = h->qnames[i + 1];
        if (ticks) ticks[i] = (uint64_t)std::llround((double)ms * 1e6);
    }
    return n;
}

// ------------------------------------------------------------------------------------------------------ records

static bool planeFits(const void* p, uint64_t pitch, uint64_t rowBytes, uint64_t rows) {
    return p && pitch >= rowBytes && pitch < (1ull << 24) && (rows - 1) * pitch + rowBytes < (1ull << 32);
}

bool frameFits(const FgrFrame& f, uint32_t W, uint32_t H, int format) {
    switch (format) {
        case FGR_FORMAT_RGBA8:
        case FGR_FORMAT_RGB10: return planeFits(f.ptr, f.pitch, 4ull * W, H);
        case FGR_FORMAT_RGBAF16: return planeFits(f.ptr, f.pitch, 8ull * W, H);
        case FGR_FORMAT_NV12:
            return planeFits(f.ptr, f.pitch, W, H) && planeFits(f.chroma, f.chromaPitch, 2ull * ((W + 1) / 2), (H + 1) / 2);
        case FGR_FORMAT_P010:
            return planeFits(f.ptr, f.pitch, 2ull * W, H) &&
                   planeFits(f.chroma, f.chromaPitch, 4ull * ((W + 1) / 2), (H + 1) / 2);
        default: return false;
    }
}

bool FgrInterpolator::frameOk(const FgrFrame& f) const { return W && frameFits(f, W, H, format); }

FrameRef FgrInterpolator::ref(const FgrFrame& f) const {
    FrameRef r;
    r.p = (const uint8_t*)f.ptr;
    r.c = (const uint8_t*)f.chroma;
    r.pitch = (int64_t)f.pitch;
    r.cpitch = (int64_t)f.chromaPitch;
    return r;
}

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// framegen.cpp:429-451: the longest integer half-vector (flow pixels) the search can reach at level j of an n-level
// pyramid -- the top radius Rt, then 2x the parent's + R per level -- and the median histogram of the step whose parent
// is level `parent`: -1 = bins of [-128, 127] (the default settings always, so their results never change), else the
// exact two-pass median with first-pass bins of 2^sh pixels, the smallest that hold that bound.
static int64_t reachBound(int n, int j, int R, int Rt) {
    const int64_t a = (int64_t)1 << (n - 1 - j);
    return Rt * a + (int64_t)R * (a - 1);
}

static int medhistShift(const FgrSettings& s, int n, int parent) {
    const bool nonDefault = s.searchRadius != 2 || s.searchRadiusTop != 6 || s.pyramidMinSide != 48;
    const int64_t b = reachBound(n, parent, (int)s.searchRadius, (int)s.searchRadiusTop);
    int sh = 0;
    while ((b >> sh) > 127) sh++;
    return nonDefault && b > 127 ? sh : -1;
}

void FgrInterpolator::planEstimate(std::vector<PlanOp>& plan, const FgrFrame& prev, const FgrFrame& curr,
                                   bool batched) const {
    const FgrSettings& st = set;
    float2* F = (float2*)field;
    // 1. luma + INTER_AREA 1/S (also clears the counters), and in the same launch the pyramid's levels up to
    // kPyrLevels -- unless a CrosshairV2 estimate fill must rewrite level 0 between the two
    const bool fill = estUi && st.uiEstimateFill;
    const int fused = fill ? 1 : std::min((int)lv.size(), kPyrLevels);
    if (!fill) {
        PyrParams pp{};
        pp.a = ref(prev);
        pp.c = ref(curr);
        pp.W = (int)W;
        pp.H = (int)H;
        pp.S = (int)st.flowScale;
        pp.u = u;
        pp.bgra = (int)info.bgra;
        pp.fmt = format;
        pp.lut = fmtLut;
        pp.yFull = info.yuvFullRange ? 1 : 0;
        pp.levels = fused;
        for (int i = 0; i < fused; i++) {
            pp.la[i] = luma + lv[i].offA;
            pp.lc[i] = luma + lv[i].offC;
            pp.w[i] = lv[i].w;
            pp.h[i] = lv[i].h;
        }
        pp.gx = cdiv(fw, 32);
        PlanOp op;
        op.pass = format == FGR_FORMAT_RGBA8 ? kPassPyramid : kPassPyramidFmt;
        op.stage = 0;
        op.gx = pp.gx;
        op.gy = cdiv(fh, 32);
        op.stamp = "luma";
        op.set(pp);
        plan.push_back(op);
    } else {
        LumaParams lp{};
        lp.a = ref(prev);
        lp.c = ref(curr);
        lp.W = (int)W;
        lp.H = (int)H;
        lp.S = (int)st.flowScale;
        lp.fw = fw;
        lp.fh = fh;
        lp.la = luma + lv[0].offA;
        lp.lc = luma + lv[0].offC;
        lp.u = u;
        lp.bgra = (int)info.bgra;
        lp.fmt = format;
        lp.lut = fmtLut;
        lp.yFull = info.yuvFullRange ? 1 : 0;
        lp.gx = cdiv(fw, 16);
        PlanOp op;
        op.pass = format == FGR_FORMAT_RGBA8 ? kPassLuma : kPassLumaFmt;
        op.stage = 0;
        op.gx = lp.gx;
        op.gy = cdiv(fh, 16);
        op.stamp = "luma";
        op.set(lp);
        plan.push_back(op);
    }
    // CrosshairV2 uiEstimateFill (framegen.cpp:2959-2963): the UI cells of both frames' level-0 luma, before the pyramid
    if (estUi && st.uiEstimateFill) {
        for (int k = 0; k < 2; k++) {
            const UiEntry& E = uiE[estEnt[k]];
            UiParams fp{};
            fp.f = ref(k ? curr : prev);
            fp.W = (int)W;
            fp.H = (int)H;
            fp.wx = E.wr.x;
            fp.wy = E.wr.y;
            fp.ww = E.wr.w;
            fp.wh = E.wr.h;
            fp.mask = E.mask;
            fp.mpitch = E.pitch;
            fp.luma = luma + (k ? lv[0].offC : lv[0].offA);
            fp.S = (int)st.flowScale;
            fp.fw = fw;
            fp.fh = fh;
            fp.bgra = (int)info.bgra;
            fp.reach = 8;
            fp.fmt = format;
            fp.lut = fmtLut;
            fp.yFull = info.yuvFullRange ? 1 : 0;
            PlanOp op;
            op.pass = kPassUiFill;
            op.stage = 1;
            op.stamp = k ? "ui_fill" : nullptr;
            op.set(fp);
            plan.push_back(op);
        }
    }
    // 2. the pyramid's levels past the fused ones
    for (size_t i = (size_t)fused; i < lv.size(); i++) {
        DownParams dp{};
        dp.sa = luma + lv[i - 1].offA;
        dp.sc = luma + lv[i - 1].offC;
        dp.da = luma + lv[i].offA;
        dp.dc = luma + lv[i].offC;
        dp.pw = lv[i - 1].w;
        dp.ph = lv[i - 1].h;
        dp.w = lv[i].w;
        dp.h = lv[i].h;
        dp.gx = cdiv(dp.w, 16);
        PlanOp op;
        op.pass = kPassDown;
        op.stage = 100 + (int)i;
        op.gx = dp.gx;
        op.gy = cdiv(dp.h, 16);
        op.stamp = i + 1 == lv.size() ? "pyramid" : nullptr;
        op.set(dp);
        plan.push_back(op);
    }
    // 3-5. per level, coarse -> fine: median (with the coarser level's median3), seed + search; median3 of level 0
    static const char* lvNames[] = {"level 0 (finest)", "level 1", "level 2", "level 3", "level 4", "level 5", "level 6", "level 7"};
    const int n = (int)lv.size();
    for (int i = n - 1; i >= 0; i--) {
        const int step = n - 1 - i;
        const Level& l = lv[i];
        SearchParams sp{};
        sp.la = luma + l.offA;
        sp.lc = luma + l.offC;
        sp.w = l.w;
        sp.h = l.h;
        sp.bw = l.bw;
        sp.bh = l.bh;
        sp.lam = st.searchPenalty;
        sp.u = u;
        if (i == n - 1) {
            sp.parent = nullptr;
            sp.R = (int)st.searchRadiusTop;
            if (st.seedNet && net) {   // the learned proposals: the CNN on this level's luma pair, then a seed candidate
                planSeeds(plan);
                sp.extraSeed = F + fieldSeed;
            }
        } else {
            const Level& pl = lv[i + 1];
            MedParams mp{};
            mp.parent = F + pl.fM;
            mp.raw = F + pl.fS;   // the parent level's median3 is formed here
            mp.filtered = F + pl.fM;
            mp.pw = pl.bw;
            mp.ph = pl.bh;
            mp.bw = l.bw;
            mp.bh = l.bh;
            mp.u = u;
            mp.medIdx = kUMedian + 2 * i;
            mp.shift = medhistShift(st, n, i + 1);
            mp.gx = 1;
            PlanOp op;
            op.pass = kPassMedhist;
            op.stage = 200 + 10 * step;
            op.set(mp);
            plan.push_back(op);
            sp.parent = F + pl.fM;
            sp.pw = pl.bw;
            sp.ph = pl.bh;
            sp.med = u + kUMedian + 2 * i;
            sp.R = (int)st.searchRadius;
        }
        sp.out = F + l.fS;
        sp.sadAcc = (i == 0 && !st.subpixel) ? 1 : 0;
        // level i of a dyadic level 0 (fb0 fractional bits) has fb0 + 2i; packed 16-bit SADs while 255.75 2^fb < 2^16
        const int fb0 = pixExactBits();
        sp.fb = fb0 >= 0 && fb0 + 2 * i <= 8 ? fb0 + 2 * i : -1;
        // R <= 2: two blocks a wavefront (twice the busy lanes, half the workgroups); a batch may ask for one
        const bool wide = sp.R <= 2 && !(batched && (estForm & FGR_FORM_SEARCH_NARROW));
        sp.gx = cdiv(l.bw * l.bh, wide ? kSearch2Blocks : 4);
        PlanOp op;
        op.pass = wide ? kPassSearch2x8 : (sp.R <= 2 ? kPassSearch2 : (sp.R <= 6 ? kPassSearch6 : kPassSearch12));
        op.stage = 201 + 10 * step;
        op.gx = sp.gx;
        op.set(sp);
        if (i > 0 || st.subpixel) {   // this level's median3 runs in the next level's median pass (level 0: in subpix)
            op.stamp = lvNames[std::min(i, 7)];
            plan.push_back(op);
            continue;
        }
        plan.push_back(op);
        Median3Params m3{};
        m3.in = F + l.fS;
        m3.out = F + l.fM;
        m3.bw = l.bw;
        m3.bh = l.bh;
        m3.gx = cdiv(l.bw, 16);
        PlanOp om;
        om.pass = kPassMedian3;
        om.stage = 202 + 10 * step;
        om.gx = m3.gx;
        om.gy = cdiv(l.bh, 16);
        om.stamp = lvNames[std::min(i, 7)];
        om.set(m3);
        plan.push_back(om);
    }
    // 6. sub-pixel refinement (+ the scene-cut SAD sum)
    const Level& l0 = lv[0];
    int bfield = l0.fM;
    if (st.subpixel) {
        SubpixParams q{};
        q.la = luma + l0.offA;
        q.lc = luma + l0.offC;
        q.w = l0.w;
        q.h = l0.h;
        q.bw = l0.bw;
        q.bh = l0.bh;
        q.in = F + l0.fM;
        q.raw = F + l0.fS;   // level 0's median3 is formed here
        q.filtered = F + l0.fM;
        q.out = F + fieldSub;
        q.lam = st.searchPenalty;
        q.u = u;
        q.fb = pixExactBits();
        q.gx = cdiv(l0.bw * l0.bh, kSubpixBlocks * (q.fb >= 0 ? kSubpixIntWaves : 1));
        PlanOp op;
        op.pass = q.fb >= 0 ? kPassSubpixInt : kPassSubpix;
        op.stage = 400;
        op.gx = q.gx;
        op.stamp = "subpixel";
        op.set(q);
        plan.push_back(op);
        bfield = fieldSub;
    }
    // 8. per-pixel field
    PixParams pp{};
    pp.la = luma + l0.offA;
    pp.lc = luma + l0.offC;
    pp.w = l0.w;
    pp.h = l0.h;
    pp.bf = F + bfield;
    pp.bw = l0.bw;
    pp.bh = l0.bh;
    pp.pf = F + fieldPF;
    pp.smooth = st.smoothBonus;
    pp.zero = st.hudBonus;
    pp.fb = pixExactBits();
    pp.gx = cdiv(l0.w, 32);
    PlanOp op;
    // the LDS-window kernel where latency decides (one pair's grid); in a batch the tuner's choice: by default the
    // gathering kernel, whose higher occupancy gives more throughput when the grids fill the GPU (the same floats)
    const bool win = pp.fb >= 0 && (!batched || (estForm & FGR_FORM_PIXFIELD_WINDOWS));
    op.pass = win ? kPassPixfieldWin : kPassPixfield;
    op.stage = 401;
    op.gx = pp.gx;
    op.gy = cdiv(l0.h, 32);
    op.stamp = "pixel field";
    op.set(pp);
    plan.push_back(op);
}

// The fractional bits fb of the level-0 luma when its values are k 2^-fb with 0 <= k < 2^14 (the pixel field's exact
// integer sampling), else -1. The cell mean of S x S pixel lumas adds 2 log2(S) bits to the per-pixel luma's: integer
// levels for Rgba8 (<= 255), 1/16 levels for Rgb10 (<= 4080) and full-range Nv12 / P010 (<= 4092); limited-range YCbCr
// (below 0 and above 255 levels) and RgbaF16 (up to +-4096 levels) take the float path.
// A protected estimate with uiEstimateFill rewrites the UI cells as float32 means over any number of pixels: not dyadic.
int FgrInterpolator::pixExactBits() const {
    if (estUi && set.uiEstimateFill) return -1;
    const int sb = set.flowScale == 1 ? 0 : (set.flowScale == 2 ? 2 : 4);
    if (format == FGR_FORMAT_RGBA8) return sb;
    const bool fixed16 = format == FGR_FORMAT_RGB10 ||
                         ((format == FGR_FORMAT_NV12 || format == FGR_FORMAT_P010) && info.yuvFullRange);
    return fixed16 && set.flowScale <= 2 ? 4 + sb : -1;
}

bool FgrInterpolator::runPlan(hipStream_t s, const std::vector<PlanOp>& plan) {
    for (const PlanOp& op : plan) {
        const hipError_t e = op.pass == kPassUiPatch ? launchUi(kUiPatch, *(const UiParams*)op.params, s)
                             : op.pass == kPassUiFill ? launchUi(kUiFill, *(const UiParams*)op.params, s)
                                                      : launchSingle(op.pass, op.params, op.gx, op.gy, s);
        if (e != hipSuccess) return false;
        if (op.stamp) stamp(s, op.stamp);
    }
    return true;
}

bool FgrInterpolator::enqueueEstimate(hipStream_t s, const FgrFrame& prev, const FgrFrame& curr) {
    planBuf.clear();
    planEstimate(planBuf, prev, curr);
    if (!profPending) beginProfile(s);   // else an observe opened this profile
    profPending = false;
    if (!runPlan(s, planBuf)) return false;
    blockFieldOff = set.subpixel ? fieldSub : lv[0].fM;
    return true;
}

// Trust as the synth kernel computes it (interp.py:236-237), for the host's tabulation of where it is 0, 1 and > 0.5
static float trustHost(int S, float lo, double rden) {
    const float e7 = (float)((double)S * (1.0 / 49.0));
    const float u = (float)((double)(e7 - lo) * rden);
    return std::fmin(std::fmax(u, 0.0f), 1.0f);
}

void FgrInterpolator::synthParams(SynthParams& p, const FgrFrame& prev, const FgrFrame& curr, float t, int slot,
                                  bool fusedLookup) const {
    const FgrSettings& st = set;
    p = SynthParams{};
    p.a = ref(prev);
    p.c = ref(curr);
    p.W = (int)W;
    p.H = (int)H;
    p.S = (int)st.flowScale;
    p.fw = fw;
    p.fh = fh;
    const float2* F = (const float2*)field;
    p.lookup = (t != 0.5f && fusedLookup) ? 1 : 0;
    p.field = (t != 0.5f && !fusedLookup) ? F + fieldT : F + fieldPF;
    p.s = (float)(2.0 * (double)t - 1.0);
    p.sadsum = u + kUSad;
    p.count = u + kUSynth + kSlots * slot;
    p.nblocks = lv[0].bw * lv[0].bh;
    p.cutSad = st.sceneCutSad;
    p.gfrac = st.globalFallback;
    p.trust0 = st.trustLow;
    p.den = st.trustHigh - st.trustLow;
    p.rden = 1.0 / (double)p.den;
    // where the (monotone) trust of a 7x7 sum is 0, 1, and past 0.5
    p.sZero = -1;
    p.sOne = 49 * 255 + 1;
    p.sHalf = -1;
    for (int S = 0; S <= 49 * 255; S++) {
        const float uS = trustHost(S, p.trust0, p.rden);
        if (uS == 0.0f) p.sZero = S;
        if (uS == 1.0f && p.sOne > S) p.sOne = S;
        if (uS <= 0.5f) p.sHalf = S;
    }
    // framegen.cpp:3490-3497: constants folded in double from the float t, then rounded to float
    const double td = t;
    p.wA = t != 0.5f ? (float)(1.0 - td) : 0.5f;
    p.wC = t != 0.5f ? (float)td : 0.5f;
    p.mid = t == 0.5f;
    p.nearC = p.wC > 0.5f;
    const float qa = 4.0f * p.wA, qc = 4.0f * p.wC;
    p.quarter = (qa == std::rint(qa) && qc == std::rint(qc)) ? 1 : 0;
    p.kA = (int)qa;
    p.kC = (int)qc;
    p.fmt = format;
    p.lut = fmtLut;
    for (int i = 0; i < 8; i++) p.yuv[i] = yuv[i];
    p.cw = ((int)W + 1) / 2;
    p.ch = ((int)H + 1) / 2;
    if (estUi) {   // CrosshairV2 pair: the kept masks of prev and curr (ui_reject.glsl)
        p.uiA = uiMaskRef(estEnt[0]);
        p.uiC = uiMaskRef(estEnt[1]);
    }
}

void FgrInterpolator::planSynth(std::vector<PlanOp>& plan, const FgrFrame& prev, const FgrFrame& curr,
                                const FgrFrame& out, float t, int k, bool uiPatch, bool merged) const {
    const int bank = k % kSynthBanks;
    SynthParams sp;
    synthParams(sp, prev, curr, t, bank, true);
    sp.out = (uint8_t*)out.ptr;
    sp.outc = (uint8_t*)out.chroma;
    sp.opitch = (int64_t)out.pitch;
    sp.ocpitch = (int64_t)out.chromaPitch;
    sp.gx = synthTilesX((int)W);
    PlanOp op;
    int rows = 32;
    if (format == FGR_FORMAT_RGBA8) {
        const int S = (int)set.flowScale;
        rows = (synthRows == 64 && S == 2) ? 64 : (synthRows == 16 ? 16 : 32);
        if (rows == 16) op.pass = S == 1 ? kPassSynthRgba8S1T16 : (S == 2 ? kPassSynthRgba8S2T16 : kPassSynthRgba8S4T16);
        else if (rows == 64) op.pass = kPassSynthRgba8S2T64;
        else op.pass = S == 1 ? kPassSynthRgba8S1 : (S == 2 ? kPassSynthRgba8S2 : kPassSynthRgba8S4);
        if (estUi) {   // the tap-rejecting variant (32-row tiles)
            rows = 32;
            op.pass = S == 1 ? kPassSynthRgba8UiS1 : (S == 2 ? kPassSynthRgba8UiS2 : kPassSynthRgba8UiS4);
        }
    } else if (estUi) {   // the tap-rejecting variant
        op.pass = set.flowScale == 1 ? kPassSynthFmtUiS1 : (set.flowScale == 2 ? kPassSynthFmtUiS2 : kPassSynthFmtUiS4);
    } else if (set.flowScale == 2 && (synthRows == 16 || synthRows == 64)) {
        rows = synthRows;
        op.pass = rows == 16 ? kPassSynthFmtS2T16 : kPassSynthFmtS2T64;
    } else {
        op.pass = set.flowScale == 1 ? kPassSynthFmtS1 : (set.flowScale == 2 ? kPassSynthFmtS2 : kPassSynthFmtS4);
    }
    op.stage = merged ? 500 : 500 + 10 * k;
    op.gx = sp.gx;
    op.gy = synthTilesY((int)H, rows);
    op.stamp = "synth";
    op.set(sp);
    plan.push_back(op);
    FallbackParams fp{};
    fp.a = sp.a;
    fp.c = sp.c;
    fp.out = sp.out;
    fp.outc = (format == FGR_FORMAT_P010 || format == FGR_FORMAT_NV12) ? sp.outc : nullptr;
    fp.cbpp = format == FGR_FORMAT_NV12 ? 2 : 4;
    fp.opitch = sp.opitch;
    fp.ocpitch = sp.ocpitch;
    fp.W = (int)W;
    fp.H = (int)H;
    fp.bpp = bytesPerPixel();
    fp.count = sp.count;
    fp.gfrac = set.globalFallback;
    fp.nearC = sp.nearC;
    fp.nextCount = merged ? nullptr : u + kUSynth + kSlots * ((k + 1) % kSynthBanks);
    fp.gx = fallbackTilesX((int)W);
    PlanOp of;
    of.pass = kPassFallback;
    of.stage = merged ? 590 : 501 + 10 * k;
    of.gx = fp.gx;
    of.gy = fallbackTilesY((int)H);
    of.stamp = "fallback";
    of.set(fp);
    plan.push_back(of);
    if (estUi && uiPatch) {   // framegen.cpp:3521-3527: curr's unchanged UI pixels over the finished frame
        PlanOp pp;
        pp.pass = kPassUiPatch;
        pp.stage = merged ? 595 : 502 + 10 * k;
        pp.stamp = "ui_patch";
        pp.set(uiPatchParams(curr, out));
        plan.push_back(pp);
    }
}

bool FgrInterpolator::enqueueSynth(hipStream_t s, const FgrFrame& prev, const FgrFrame& curr, const FgrFrame& out, float t) {
    planBuf.clear();
    planSynth(planBuf, prev, curr, out, t, synthK);
    if (!runPlan(s, planBuf)) return false;
    lastSlot = synthK % kSynthBanks;
    lastT = t;
    synthK++;
    return true;
}

int FgrInterpolator::bytesPerPixel() const {
    switch (format) {
    case FGR_FORMAT_RGBAF16: return 8;
    case FGR_FORMAT_P010: return 2;
    case FGR_FORMAT_NV12: return 1;
    default: return 4;
    }
}

extern "C" int fgr_record_estimate(FgrInterpolator* h, void* stream, FgrFrame prev, FgrFrame curr) {
    if (!h || !h->W || !h->frameOk(prev) || !h->frameOk(curr)) return 0;
    if (h->ringEst >= FGR_SET_RING) return 0;   // framegen.cpp:2873-2874
    // framegen.cpp:2875-2911: a protected estimate needs both frames kept; a host that never observes gets them observed
    // here (prev if unseen, curr unless it is the newest observed); with explicit observes an unseen frame means the
    // pair runs unprotected
    hipStream_t s = (hipStream_t)stream;
    bool ui = h->uiActive();
    if (ui) {
        int needObs = 0;
        if (!h->uiExplicit) {
            needObs += h->uiFind(prev.ptr) < 0;
            needObs += h->uiLast < 0 || h->uiE[h->uiLast].key != curr.ptr;
        } else if (h->uiFind(prev.ptr) < 0 || h->uiFind(curr.ptr) < 0) {
            ui = false;
        }
        if (ui && h->uiSets + (uint32_t)needObs > FGR_UI_SETS) return 0;
    }
    int ep = -1, ec = -1;
    if (ui) {
        if (!h->uiExplicit) {
            if (h->profiling && !h->profPending) h->beginProfile(s);
            h->profPending = h->profiling;
            if (h->uiFind(prev.ptr) < 0 && !h->uiObserveRecord(s, prev)) return 0;
            if ((h->uiLast < 0 || h->uiE[h->uiLast].key != curr.ptr) && !h->uiObserveRecord(s, curr)) return 0;
        }
        ep = h->uiFind(prev.ptr);
        ec = h->uiFind(curr.ptr);
        if (ep < 0 || ec < 0 || ep == ec) return 0;
    }
    h->estUi = ui;   // (read by the plan: the estimate fill)
    if (ui) {
        h->estEnt[0] = ep;
        h->estEnt[1] = ec;
        h->estSeq[0] = h->uiE[ep].seq;
        h->estSeq[1] = h->uiE[ec].seq;
    }
    if (!h->enqueueEstimate(s, prev, curr)) {
        h->estUi = false;
        h->haveEst = false;
        return 0;
    }
    h->ringEst++;
    h->haveEst = true;
    h->estPrev = prev.ptr;
    h->estCurr = curr.ptr;
    h->estPrevF = prev;
    h->estCurrF = curr;
    h->lastT = 0.0f;
    h->synthK = 0;
    return 1;
}

extern "C" int fgr_record_synth(FgrInterpolator* h, void* stream, FgrFrame prev, FgrFrame curr, FgrFrame out, float t) {
    // framegen.cpp:3467-3472
    if (!h || !h->W || !(t > 0.0f && t < 1.0f)) return 0;
    if (!h->haveEst || prev.ptr != h->estPrev || curr.ptr != h->estCurr || !h->frameOk(prev) || !h->frameOk(curr) ||
        !h->frameOk(out))
        return 0;
    if (h->ringSynth >= FGR_SET_RING) return 0;
    if (h->estUi) {   // framegen.cpp:3473-3478: the entries the estimate used must still hold those frames
        if (h->uiE[h->estEnt[0]].seq != h->estSeq[0] || h->uiE[h->estEnt[1]].seq != h->estSeq[1]) return 0;
        if (h->uiSets + 2 > FGR_UI_SETS) return 0;
    }
    if (!h->enqueueSynth((hipStream_t)stream, prev, curr, out, t)) return 0;
    h->ringSynth++;
    if (h->estUi) h->uiSets++;
    return 1;
}

extern "C" int fgr_record(FgrInterpolator* h, void* stream, FgrFrame prev, FgrFrame curr, FgrFrame out, float t) {
    // framegen.cpp:3636-3643: both rings must have room, so that a refused record records nothing
    if (!h || !h->W || !(t > 0.0f && t < 1.0f)) return 0;
    if (h->ringEst >= FGR_SET_RING || h->ringSynth >= FGR_SET_RING) return 0;
    if (!h->frameOk(out)) return 0;
    if (h->uiActive() && h->uiSets + 2 + 2 > FGR_UI_SETS) return 0;   // framegen.cpp:3641: 2 implicit observes + the patch
    return fgr_record_estimate(h, stream, prev, curr) && fgr_record_synth(h, stream, prev, curr, out, t);
}

// ------------------------------------------------------------------------------------------------------ debug

extern "C" int fgr_record_debug_counters(FgrInterpolator* h, void* stream, void* dst) {
    // framegen.cpp:2073-2084: [untrusted pixels of the latest synth, the estimate's SAD sum x 256]
    if (!h || !h->W || !dst) return 0;
    return launchDebugCounters(h->u + kUSynth + kSlots * h->lastSlot, h->u + kUSad, (uint32_t*)dst,
                               (hipStream_t)stream) == hipSuccess;
}

extern "C" uint32_t fgr_debug_sad_blocks(const FgrInterpolator* h) {
    return h && !h->lv.empty() ? (uint32_t)(h->lv[0].bw * h->lv[0].bh) : 0u;
}

extern "C" int fgr_debug_map_layout(const FgrInterpolator* h, FgrDebugMapLayout* l) {
    // framegen.cpp:2090-2104
    if (!h || !h->W || h->lv.empty() || !l) return 0;
    l->W = h->W;
    l->H = h->H;
    l->fw = (uint32_t)h->fw;
    l->fh = (uint32_t)h->fh;
    l->bw = (uint32_t)h->lv[0].bw;
    l->bh = (uint32_t)h->lv[0].bh;
    l->fscale = h->set.flowScale;
    l->eStride = (h->W + 3) / 4;
    l->blockOff = 0;
    l->pixOff = (uint64_t)l->bw * l->bh * 8;
    l->eOff = l->pixOff + (uint64_t)l->fw * l->fh * 8;
    l->size = l->eOff + (uint64_t)l->eStride * l->H * 4;
    return 1;
}

extern "C" int fgr_fgmaps_header(const FgrInterpolator* h, void* out64) {
    FgrDebugMapLayout l;
    if (!h || !out64 || !fgr_debug_map_layout(h, &l)) return 0;
    uint32_t w[16];
    w[0] = 0x504D4746u;   // 'FGMP'
    w[1] = 5;
    w[2] = l.W;
    w[3] = l.H;
    w[4] = l.fw;
    w[5] = l.fh;
    w[6] = l.bw;
    w[7] = l.bh;
    w[8] = l.fscale;
    w[9] = l.eStride;
    const float f[4] = {h->set.trustLow, h->set.trustHigh, h->set.smoothBonus, h->set.hudBonus};
    std::memcpy(w + 10, f, 16);
    w[14] = h->set.motionVersion;
    w[15] = 0;
    std::memcpy(out64, w, 64);
    return 1;
}

extern "C" int fgr_record_debug_maps(FgrInterpolator* h, void* stream, void* dst) {
    FgrDebugMapLayout l;
    if (!h || !dst || !h->haveEst || !fgr_debug_map_layout(h, &l)) return 0;
    hipStream_t s = (hipStream_t)stream;
    uint8_t* d = (uint8_t*)dst;
    const float2* F = (const float2*)h->field;
    if (hipMemcpyAsync(d + l.blockOff, F + h->blockFieldOff, (size_t)l.bw * l.bh * 8, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return 0;
    if (hipMemcpyAsync(d + l.pixOff, F + h->fieldPF, (size_t)l.fw * l.fh * 8, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return 0;
    if (h->lastT > 0.0f) {
        // the mismatch of the latest synth, recomputed from its field at t (the fused synth never stores e)
        SynthParams sp;
        h->synthParams(sp, h->estPrevF, h->estCurrF, h->lastT, h->lastSlot, false);
        if (h->lastT != 0.5f) {   // the field at t into the scratch (the fused synth computes it per tile)
            LookupParams lk{};
            lk.pf = (const float2*)h->field + h->fieldPF;
            lk.out = (float2*)h->field + h->fieldT;
            lk.fw = h->fw;
            lk.fh = h->fh;
            lk.s = sp.s;
            if (launchLookup(lk, s) != hipSuccess) return 0;
        }
        if (launchDebugE(sp, (uint32_t*)(d + l.eOff), (int)l.eStride, h->format, s) != hipSuccess) return 0;
    } else if (hipMemsetAsync(d + l.eOff, 0, (size_t)l.eStride * l.H * 4, s) != hipSuccess) {
        return 0;
    }
    return 1;
}

// ------------------------------------------------------------------------------------------------------ blur

bool FgrInterpolator::copyFrame(hipStream_t s, const FgrFrame& src, const FgrFrame& dst) {
    // t <= 0 / t >= 1 of recordSynthBlur: the real frame itself (framegen.h:557-563)
    const size_t row = (size_t)W * bytesPerPixel();
    if (hipMemcpy2DAsync(dst.ptr, dst.pitch, src.ptr, src.pitch, row, H, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return false;
    if ((format == FGR_FORMAT_P010 || format == FGR_FORMAT_NV12) &&
        hipMemcpy2DAsync(dst.chroma, dst.chromaPitch, src.chroma, src.chromaPitch,
                         (size_t)((W + 1) / 2) * (format == FGR_FORMAT_NV12 ? 2 : 4),
                         (H + 1) / 2, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return false;
    return true;
}

extern "C" int fgr_record_synth_blur(FgrInterpolator* h, void* stream, FgrFrame prev, FgrFrame curr, FgrFrame out,
                                     float t0, float t1) {
    // framegen.cpp:3534-3545: 0 <= t0 <= t1 <= 1 (false for NaN); blur off or t0 == t1 -> recordSynth(t1)
    if (!h || !h->W || !(t0 >= 0.0f && t1 <= 1.0f && t0 <= t1)) return 0;
    const uint32_t N = h->set.blurSamples;
    const bool blur = N >= 2 && t0 < t1;
    if (!blur && t1 > 0.0f && t1 < 1.0f) return fgr_record_synth(h, stream, prev, curr, out, t1);
    if (!h->haveEst || prev.ptr != h->estPrev || curr.ptr != h->estCurr || !h->frameOk(prev) || !h->frameOk(curr) ||
        !h->frameOk(out))
        return 0;
    if (h->ringSynth >= FGR_SET_RING) return 0;
    if (h->estUi) {   // framegen.cpp:3543-3548
        if (h->uiE[h->estEnt[0]].seq != h->estSeq[0] || h->uiE[h->estEnt[1]].seq != h->estSeq[1]) return 0;
        if (h->uiSets + 2 > FGR_UI_SETS) return 0;
    }
    hipStream_t s = (hipStream_t)stream;
    if (!blur) {
        if (!h->copyFrame(s, t1 <= 0.0f ? prev : curr, out)) return 0;
        if (h->estUi && launchUi(kUiPatch, h->uiPatchParams(curr, out), s) != hipSuccess) return 0;
    } else if (!h->enqueueBlur(s, prev, curr, out, t0, t1)) {
        return 0;
    }
    h->ringSynth++;
    if (h->estUi) h->uiSets++;
    return 1;
}

extern "C" int fgr_debug_internal(const FgrInterpolator* h, FgrDebugInternal* o) {
    if (!h || !h->W || !o) return 0;
    std::memset(o, 0, sizeof(*o));
    o->luma = h->luma;
    o->field = h->field;
    o->counters = h->u;
    o->levels = (int32_t)h->lv.size();
    o->fw = h->fw;
    o->fh = h->fh;
    for (size_t i = 0; i < h->lv.size() && i < 16; i++) {
        const Level& l = h->lv[i];
        o->w[i] = l.w;
        o->h[i] = l.h;
        o->bw[i] = l.bw;
        o->bh[i] = l.bh;
        o->lumaOffA[i] = l.offA;
        o->lumaOffC[i] = l.offC;
        o->fS[i] = l.fS;
        o->fM[i] = l.fM;
    }
    o->fieldSub = h->fieldSub;
    o->fieldPF = h->fieldPF;
    o->fieldT = h->fieldT;
    o->lumaWords = h->lumaWords;
    o->fieldEntries = h->fieldEntries;
    o->lastSynthSlot = h->lastSlot;
    o->fieldSeed = h->fieldSeed;
    return 1;
}
