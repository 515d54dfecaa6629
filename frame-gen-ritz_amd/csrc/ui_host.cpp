// Host side of CrosshairV2 static-UI protection (gpu/framegen.cpp:1041-1075 box geometry, :1183-1310 the kept-frame
// set, :2601-2660 observe / track, :2830-2870 the entry points, :2875-2911 the estimate's observes, :3467-3530 the
// synth's rejection and patch). The kept frames are keyed by frame pointer and recycled least recently observed first;
// every buffer is allocated when settings or sizes are applied (create, resize, set_settings, reconfigure), so no
// record allocates.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include "../../include/fgr.h"
#include "runtime.hpp"
#include "v1_launch.hpp"

using namespace fgr;

// ------------------------------------------------------------------------------------------------------ geometry

extern "C" void fgr_box_size(const FgrSettings* s, uint32_t w, uint32_t h, uint32_t* boxW, uint32_t* boxH) {
    // framegen.cpp:1041-1050: explicit sizes win over the height fraction (or the legacy area), clamped to the frame
    const int fw = (int)w, fh = (int)h;
    int side;
    if (s->uiBoxHeightFrac > 0.0f)
        side = std::min(std::max((int)std::lround((double)h * s->uiBoxHeightFrac), 8), std::min(fw, fh));
    else
        side = std::min(std::max((int)std::lround(std::sqrt((double)s->uiCrosshairArea * fw * fh)), 32), std::min(fw, fh));
    if (boxW) *boxW = s->uiBoxWidth ? (uint32_t)std::min(std::max((int)s->uiBoxWidth, 4), fw) : (uint32_t)side;
    if (boxH) *boxH = s->uiBoxHeight ? (uint32_t)std::min(std::max((int)s->uiBoxHeight, 4), fh) : (uint32_t)side;
}

extern "C" void fgr_box_rect(const FgrSettings* s, uint32_t w, uint32_t h, int32_t r[4]) {
    // framegen.cpp:1053-1060: centred on the frame centre plus the offset, clamped into the frame
    uint32_t bw, bh;
    fgr_box_size(s, w, h, &bw, &bh);
    r[2] = (int32_t)bw;
    r[3] = (int32_t)bh;
    r[0] = std::min(std::max(((int)w - r[2]) / 2 + s->uiBoxOffsetX, 0), (int)w - r[2]);
    r[1] = std::min(std::max(((int)h - r[3]) / 2 + s->uiBoxOffsetY, 0), (int)h - r[3]);
}

extern "C" uint32_t fgr_box_margin(const FgrSettings* s) {
    // framegen.cpp:1065-1067: Crosshair 2 px, CrosshairV2 1 px, otherwise 0
    return s->uiProtection == FGR_UI_CROSSHAIR ? 2u : (s->uiProtection == FGR_UI_CROSSHAIR_V2 ? 1u : 0u);
}

extern "C" void fgr_work_rect(const FgrSettings* s, uint32_t w, uint32_t h, int32_t r[4]) {
    // framegen.cpp:1069-1075: the box grown by the margin, clamped to the frame
    fgr_box_rect(s, w, h, r);
    const int m = (int)fgr_box_margin(s);
    const int x0 = std::max(r[0] - m, 0), y0 = std::max(r[1] - m, 0);
    const int x1 = std::min(r[0] + r[2] + m, (int)w), y1 = std::min(r[1] + r[3] + m, (int)h);
    r[0] = x0;
    r[1] = y0;
    r[2] = x1 - x0;
    r[3] = y1 - y0;
}

UiRect FgrInterpolator::uiBox() const {
    int32_t r[4];
    fgr_box_rect(&set, W, H, r);
    return UiRect{r[0], r[1], r[2], r[3]};
}

UiRect FgrInterpolator::uiWork() const {
    int32_t r[4];
    fgr_work_rect(&set, W, H, r);
    return UiRect{r[0], r[1], r[2], r[3]};
}

// ------------------------------------------------------------------------------------------------------ resources

static int maskPitch(const UiRect& wr) { return (wr.w + 3) & ~3; }

bool FgrInterpolator::uiEnsure() {
    // framegen.cpp:1183-1236: uiHistory active entries (a raise re-activates kept entries first, a lowering forgets the
    // least recently observed); the state and every active entry's mask sized for the current work rectangle
    if (!W || set.uiProtection == FGR_UI_OFF) return true;
    const int want = (int)std::min(std::max(set.uiHistory, (uint32_t)FGR_UI_FRAMES), (uint32_t)FGR_UI_FRAMES_MAX);
    const UiRect wr = uiWork();
    const size_t stateBytes = (size_t)wr.w * wr.h * 16, maskBytes = (size_t)maskPitch(wr) * wr.h;
    if (hipSetDevice(device) != hipSuccess) return false;
    if (stateBytes > uiStateCap) {
        uint32_t* p = nullptr;
        if (hipMalloc(&p, stateBytes) != hipSuccess) return false;
        if (uiState) hipFree(uiState);
        uiState = p;
        uiStateCap = stateBytes;
        uiBoxHist = false;   // a new state buffer describes no earlier frame
    }
    int nAct = 0;
    for (const UiEntry& e : uiE) nAct += e.active;
    for (int i = 0; nAct < want && i < FGR_UI_FRAMES_MAX; i++)
        if (!uiE[i].active) {
            uiE[i].active = true;
            uiE[i].key = nullptr;
            uiE[i].seq = 0;
            nAct++;
        }
    while (nAct > want) {
        int lo = -1;
        for (int i = 0; i < FGR_UI_FRAMES_MAX; i++)
            if (uiE[i].active && (lo < 0 || uiE[i].seq < uiE[lo].seq)) lo = i;
        uiE[lo].active = false;
        uiE[lo].key = nullptr;
        uiE[lo].seq = 0;
        if (uiLast == lo) uiLast = -1;
        nAct--;
    }
    for (int i = 0; i < FGR_UI_FRAMES_MAX; i++) {
        UiEntry& e = uiE[i];
        if (!e.active || e.cap >= maskBytes) continue;
        uint8_t* p = nullptr;
        if (hipMalloc(&p, maskBytes) != hipSuccess) return false;
        if (e.mask) hipFree(e.mask);
        e.mask = p;
        e.cap = maskBytes;
        e.key = nullptr;   // a grown mask holds nothing observed
        e.seq = 0;
        if (uiLast == i) uiLast = -1;
    }
    return true;
}

void FgrInterpolator::uiFree() {
    for (UiEntry& e : uiE) {
        if (e.mask) hipFree(e.mask);
        e = UiEntry{};
    }
    if (uiState) hipFree(uiState);
    uiState = nullptr;
    uiStateCap = 0;
    uiLast = -1;
    uiBoxHist = false;
    estUi = false;
}

extern "C" void fgr_release_unused_ui(FgrInterpolator* h) {
    // framegen.cpp:1757-1761: the masks of entries that uiHistory no longer keeps
    if (!h) return;
    hipSetDevice(h->device);
    for (UiEntry& e : h->uiE)
        if (!e.active && e.mask) {
            hipFree(e.mask);
            e = UiEntry{};
        }
}

void FgrInterpolator::uiAdopted(const FgrSettings& old) {
    // framegen.cpp:1277-1292: switching off forgets the observed frames; a gap in observation restarts the counters;
    // another mode's state and masks mean nothing to CrosshairV2
    auto forgetAll = [&]() {
        for (UiEntry& e : uiE) {
            e.key = nullptr;
            e.seq = 0;
        }
        uiLast = -1;
    };
    if (set.uiProtection == FGR_UI_OFF) forgetAll();
    if (set.uiProtection == FGR_UI_OFF || old.uiProtection == FGR_UI_OFF) uiBoxHist = false;
    if (set.uiProtection != old.uiProtection &&
        (set.uiProtection == FGR_UI_CROSSHAIR_V2 || old.uiProtection == FGR_UI_CROSSHAIR_V2)) {
        forgetAll();
        uiBoxHist = false;
    }
}

bool FgrInterpolator::uiChanged(const FgrSettings& old, const UiRect& oldBox, const UiRect& oldWork) const {
    // framegen.cpp:934-939: the mode or its box changed (set holds the new settings)
    if (set.uiProtection != old.uiProtection) return true;
    if (set.uiProtection == FGR_UI_OFF) return false;
    return !(uiBox() == oldBox) || !(uiWork() == oldWork);
}

int FgrInterpolator::uiFind(const void* key) const {
    for (int i = 0; i < FGR_UI_FRAMES_MAX; i++)
        if (key && uiE[i].active && uiE[i].key == key) return i;
    return -1;
}

int FgrInterpolator::uiPick(const void* key) const {
    // framegen.cpp:1300-1308: the entry keyed by this frame, else the least recently observed
    const int f = uiFind(key);
    if (f >= 0) return f;
    int best = -1;
    for (int i = 0; i < FGR_UI_FRAMES_MAX; i++)
        if (uiE[i].active && (best < 0 || uiE[i].seq < uiE[best].seq)) best = i;
    return best;
}

void FgrInterpolator::uiForget(const void* key) {
    // framegen.cpp:940-943: the entry keyed by this frame no longer holds its content
    const int e = uiFind(key);
    if (e >= 0) {
        uiE[e].key = nullptr;
        uiE[e].seq = 0;
        if (uiLast == e) uiLast = -1;
    }
}

// ------------------------------------------------------------------------------------------------------ passes

// the per-format constants every UI pass reads
static void uiFormatParams(const FgrInterpolator& I, UiParams& p) {
    p.fmt = I.format;
    p.lut = I.fmtLut;
    for (int i = 0; i < 5; i++) p.yuv[i] = I.yuv[i];
    p.yFull = I.info.yuvFullRange ? 1 : 0;
    p.bgra = (int)I.info.bgra;
}

static UiParams uiBaseParams(const FgrInterpolator& I, const UiRect& bx, const UiRect& wr) {
    UiParams p{};
    uiFormatParams(I, p);
    p.W = (int)I.W;
    p.H = (int)I.H;
    p.bx = bx.x;
    p.by = bx.y;
    p.bw = bx.w;
    p.bh = bx.h;
    p.wx = wr.x;
    p.wy = wr.y;
    p.ww = wr.w;
    p.wh = wr.h;
    p.thr = I.set.uiStillThreshold;
    p.cthr = I.set.uiContrastThreshold;
    p.still = (int)I.set.uiStillFrames;
    p.leave = (int)I.set.uiLeaveFrames;
    p.reentry = (int)I.set.uiReentryFrames;
    p.state = I.uiState;
    return p;
}

bool FgrInterpolator::uiObserveRecord(hipStream_t s, const FgrFrame& f) {
    // framegen.cpp:2601-2636: the state machine of the box, then the frame's mask into the entry it picks
    const UiRect bx = uiBox(), wr = uiWork();
    const int e = uiPick(f.ptr);
    if (e < 0 || (size_t)wr.w * wr.h * 16 > uiStateCap || uiE[e].cap < (size_t)maskPitch(wr) * wr.h) return false;
    if (!(bx == uiPrevBox)) uiBoxHist = false;
    uiPrevBox = bx;
    UiEntry& E = uiE[e];
    E.key = f.ptr;
    E.seq = ++uiSeq;
    E.wr = wr;
    E.pitch = maskPitch(wr);
    uiLast = e;
    UiParams p = uiBaseParams(*this, bx, wr);
    p.f = ref(f);
    p.reset = uiBoxHist ? 0 : 1;
    p.mask = E.mask;
    p.mpitch = E.pitch;
    if (launchUi(kUiUpdate, p, s) != hipSuccess || launchUi(kUiMask, p, s) != hipSuccess) return false;
    stamp(s, "ui_detect");
    uiBoxHist = true;
    uiSets++;
    return true;
}

bool FgrInterpolator::uiTrackRecord(hipStream_t s, const FgrFrame& f) {
    // framegen.cpp:2638-2656: the same state-machine pass, nothing else
    const UiRect bx = uiBox(), wr = uiWork();
    if ((size_t)wr.w * wr.h * 16 > uiStateCap) return false;
    if (!(bx == uiPrevBox)) uiBoxHist = false;
    uiPrevBox = bx;
    UiParams p = uiBaseParams(*this, bx, wr);
    p.f = ref(f);
    p.reset = uiBoxHist ? 0 : 1;
    if (launchUi(kUiUpdate, p, s) != hipSuccess) return false;
    uiBoxHist = true;
    uiSets++;
    return true;
}

UiMaskRef FgrInterpolator::uiMaskRef(int e) const {
    UiMaskRef m{};
    if (e < 0) return m;
    const UiEntry& E = uiE[e];
    m.m = E.mask;
    m.x = E.wr.x;
    m.y = E.wr.y;
    m.w = E.wr.w;
    m.h = E.wr.h;
    m.pitch = E.pitch;
    return m;
}

UiParams FgrInterpolator::uiPatchParams(const FgrFrame& curr, const FgrFrame& out) const {
    // framegen.cpp:2822-2828: curr's entry, its work rectangle, only the bytes of 255
    const UiEntry& E = uiE[estEnt[1]];
    UiParams p{};
    uiFormatParams(*this, p);
    p.f = ref(curr);
    p.out = (uint8_t*)out.ptr;
    p.opitch = (int64_t)out.pitch;
    p.outc = (uint8_t*)out.chroma;
    p.ocpitch = (int64_t)out.chromaPitch;
    p.W = (int)W;
    p.H = (int)H;
    p.wx = E.wr.x;
    p.wy = E.wr.y;
    p.ww = E.wr.w;
    p.wh = E.wr.h;
    p.mask = E.mask;
    p.mpitch = E.pitch;
    return p;
}

// ------------------------------------------------------------------------------------------------------ entry points

extern "C" int fgr_record_observe(FgrInterpolator* h, void* stream, FgrFrame frame) {
    // framegen.cpp:2830-2851
    if (!h || !h->W || !h->frameOk(frame)) return 0;
    if (h->set.uiProtection == FGR_UI_OFF) return 1;
    if (!h->uiActive()) {
        h->uiForget(frame.ptr);
        return 0;
    }
    h->uiExplicit = true;
    if (h->uiSets + 1 > FGR_UI_SETS) {
        h->uiForget(frame.ptr);
        return 0;
    }
    hipStream_t s = (hipStream_t)stream;
    if (h->profiling && !h->profPending) h->beginProfile(s);   // an observe opens a profile the next estimate continues
    h->profPending = h->profiling;
    if (!h->uiObserveRecord(s, frame)) {
        h->uiForget(frame.ptr);
        return 0;
    }
    return 1;
}

extern "C" int fgr_record_track(FgrInterpolator* h, void* stream, FgrFrame frame) {
    // framegen.cpp:2853-2864
    if (!h || !h->W || !h->frameOk(frame)) return 0;
    if (h->set.uiProtection == FGR_UI_OFF) return 1;
    h->uiForget(frame.ptr);
    if (!h->uiActive()) return 0;
    if (h->uiSets + 1 > FGR_UI_SETS) return 0;
    return h->uiTrackRecord((hipStream_t)stream, frame) ? 1 : 0;
}

extern "C" int fgr_debug_ui_mask(FgrInterpolator* h, void* stream, FgrFrame frame, void* dst, int32_t rect[4]) {
    if (!h || !h->uiActive() || !dst) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (!frame.ptr) {
        const UiRect wr = h->uiWork();
        if (rect) {
            rect[0] = wr.x;
            rect[1] = wr.y;
            rect[2] = wr.w;
            rect[3] = wr.h;
        }
        return hipMemcpyAsync(dst, h->uiState, (size_t)wr.w * wr.h * 16, hipMemcpyDeviceToDevice, s) == hipSuccess;
    }
    const int e = h->uiFind(frame.ptr);
    if (e < 0) return 0;
    const UiEntry& E = h->uiE[e];
    if (rect) {
        rect[0] = E.wr.x;
        rect[1] = E.wr.y;
        rect[2] = E.wr.w;
        rect[3] = E.wr.h;
    }
    return hipMemcpy2DAsync(dst, (size_t)E.wr.w, E.mask, (size_t)E.pitch, (size_t)E.wr.w, (size_t)E.wr.h,
                            hipMemcpyDeviceToDevice, s) == hipSuccess;
}
