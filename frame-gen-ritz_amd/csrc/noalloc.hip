// fgr_record_noalloc: fgr_record (estimate + one synth) without a handle and without any allocation. The pair's
// intermediates live in `scratch`, a file-scope device array of FGR_NOALLOC_SCRATCH_BYTES; the launch plan is built in
// a per-device host state whose containers are sized once, at the first call on that device, and reused.
//
// In host code `scratch` names a host-side stand-in, not the device array: every use below goes through the device
// address hipGetSymbolAddress returned for the current device, kept per device (each GPU holds its own copy).
#include <hip/hip_runtime.h>
#include "../../include/fgr.h"
#include "runtime.hpp"

static __device__ __align__(256) uint8_t scratch[FGR_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;

const char* kNaNull = "null InitInfo";
const char* kNaFormat = "unsupported format (Rgba8, Rgb10, Nv12 and P010 run without allocating)";
const char* kNaDevice = "the HIP device could not be selected";
const char* kNaSymbol = "the address of the static scratch array could not be read on this device";
const char* kNaF16 = "RgbaF16 needs its per-handle level table: use fgr_create + fgr_record";
const char* kNaUi = "UI protection keeps per-handle state: fgr_record_noalloc runs with uiProtection Off";
const char* kNaSeed = "learned seeds need an attached network: fgr_record_noalloc runs with seedNet 0";
const char* kNaSize = "the pair's intermediates exceed FGR_NOALLOC_SCRATCH_BYTES (48 MiB): at most 1920x1080 at "
                      "flowScale 1 or 3840x2160 at flowScale 2";
const char* kNaFrame = "a frame has a null plane or a pitch that does not fit the size";
const char* kNaT = "t must be in (0, 1)";
const char* kNaLaunch = "a kernel launch failed";

struct DeviceState {
    bool ready = false;
    uint8_t* scratch = nullptr;   // this device's address of `scratch`
    FgrInterpolator h;            // the plan's state: its luma / field / counter pointers point into `scratch`
};

DeviceState gDev[kMaxDevices];

bool refuse(const char** error, const char* why) {
    if (error) *error = why;
    return false;
}

// Once per device, with that device current: the array's device address, and the host containers sized for any plan
// (so that no later call allocates host memory either).
bool setUp(DeviceState& d) {
    if (d.ready) return true;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return false;
    d.scratch = (uint8_t*)p;
    d.h.lv.reserve(32);
    d.h.planBuf.reserve(64);
    d.ready = true;
    return true;
}

bool recordNoAlloc(void* stream, const FgrInitInfo* info, const FgrFrame& prev, const FgrFrame& curr,
                   const FgrFrame& out, float t, const char** error) {
    if (!info) return refuse(error, kNaNull);
    const uint32_t fmt = info->format;
    if (!formatBuilt(fmt)) return refuse(error, kNaFormat);
    if (const char* e = fgr_settings_error_for_format(&info->settings, fmt)) return refuse(error, e);
    if (fmt == FGR_FORMAT_RGBAF16) return refuse(error, kNaF16);
    if (info->settings.uiProtection != FGR_UI_OFF) return refuse(error, kNaUi);
    if (info->settings.seedNet) return refuse(error, kNaSeed);
    if (const char* e = sizeLimitError(info->width, info->height, fmt)) return refuse(error, e);
    if ((fmt == FGR_FORMAT_P010 || fmt == FGR_FORMAT_NV12) && info->yuvMatrix > FGR_YUV_BT2020)
        return refuse(error, "invalid yuvMatrix");
    if (!(t > 0.0f && t < 1.0f)) return refuse(error, kNaT);
    if (info->device < 0 || info->device >= kMaxDevices) return refuse(error, kNaDevice);
    DeviceState& d = gDev[info->device];
    FgrInterpolator& h = d.h;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return refuse(error, kNaDevice);
    if (!d.ready) {
        if (cur != info->device && hipSetDevice(info->device) != hipSuccess) return refuse(error, kNaDevice);
        const bool ok = setUp(d);
        if (cur != info->device) hipSetDevice(cur);
        if (!ok) return refuse(error, kNaSymbol);
    }
    // the handle state fgr_create + the layout of allocBuffers would give, pointing into the array
    h.info = *info;
    h.set = info->settings;
    h.device = info->device;
    h.format = (int)fmt;
    h.fmtLut = nullptr;
    h.formatInit();   // Nv12 / P010: host constants only (RgbaF16, the one format with a device table, is refused)
    h.layoutBuffers(info->width, info->height);
    if (h.bufferBytesNoDebug() > (size_t)FGR_NOALLOC_SCRATCH_BYTES) {
        h.W = h.H = 0;
        return refuse(error, kNaSize);
    }
    h.W = info->width;
    h.H = info->height;
    const size_t lumaBytes = ((size_t)h.lumaWords * 4 + 255) & ~(size_t)255;
    const size_t fieldBytes = ((size_t)h.fieldT * 8 + 255) & ~(size_t)255;
    h.luma = (float*)d.scratch;
    h.field = (float*)(d.scratch + lumaBytes);
    h.u = (uint32_t*)(d.scratch + lumaBytes + fieldBytes);
    if (!h.frameOk(prev) || !h.frameOk(curr) || !h.frameOk(out)) return refuse(error, kNaFrame);
    h.estUi = false;
    h.net = nullptr;
    h.synthRows = 32;
    h.profiling = false;
    // the estimate (its luma pass clears every counter, so the array's earlier contents never matter), then synth 0
    h.planBuf.clear();
    h.planEstimate(h.planBuf, prev, curr);
    h.planSynth(h.planBuf, prev, curr, out, t, 0, false);
    if (!h.runPlan((hipStream_t)stream, h.planBuf)) return refuse(error, kNaLaunch);
    return true;
}

}  // namespace

extern "C" int fgr_record_noalloc(void* stream, const FgrInitInfo* info, FgrFrame prev, FgrFrame curr, FgrFrame out,
                                  float t, const char** error) {
    if (error) *error = nullptr;
    return recordNoAlloc(stream, info, prev, curr, out, t, error) ? 1 : 0;
}
