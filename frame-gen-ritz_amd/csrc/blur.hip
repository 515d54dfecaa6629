// recordSynthBlur with blurSamples >= 2 (gpu/framegen.cpp:3534-3634, gpu/shaders/blur_resolve.comp): the output is
// the weighted mean of the frames recordSynth would produce at N sample times, rounded once.
//
// Samples: the midpoints of N equal slices of [t0, t1], t_k = t0 + (t1 - t0)(k + 1/2) / N (double, then float,
// clamped to [t0, t1]); a sample at t <= 0 / t >= 1 is the real frame prev / curr. Weights: 1, or exp(-2 x^2) with
// x_k = (2k + 1 - N) / N (Gaussian, the window ends at +-2 sigma), computed in double and rounded to float; the sum of
// the weights is a float32 sum in sample order.
//
// Each interior sample is the synthesis kernel's own output for t_k (with its scene-cut and flick fallbacks), written to
// a scratch frame of the library (N frames of the output's size, grown by the first blurred record; HBM is plentiful
// on this GPU, and the sub-frames are then bit-identical to recordSynth(t_k) by construction, flick decisions
// included, without the reference's count-then-fix-up pass). One averaging kernel then forms, per stored value,
// acc = sum_k w_k v_k (float32, sample order) and writes round(acc / wsum) in the format's own codes: 8- and 10-bit
// codes floor(x + 0.5) clamped to the code range (alpha weighted like the colour), half floats rounded to nearest even.
// Unlike the reference's V1 blur, which rounds every sub-frame to 8-bit levels in the RGB formats and not at all in the
// YUV ones (a quirk its own notes record), every format here averages the sub-frames exactly as they are stored and
// rounds once, as the reference's V2 blur does.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <vector>
#include "runtime.hpp"
#include "v1_device.hpp"
#include "v1_launch.hpp"

namespace fgr {

struct BlurParams {
    int W, H, fmt, n;
    const uint8_t* src[16];
    const uint8_t* srcC[16];
    int64_t pitch[16], cpitch[16];
    float w[16];
    float wsum;
    uint8_t* out;
    uint8_t* outc;
    int64_t opitch, ocpitch;
};

__device__ __forceinline__ float rnd(float acc, float wsum, float hi) {
    return fminf(fmaxf(floorf(acc / wsum + 0.5f), 0.0f), hi);
}

__global__ __launch_bounds__(256) void blur_average_kernel(BlurParams P) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= P.W || y >= P.H) return;
    if (P.fmt == kRgba8 || P.fmt == kRgb10) {
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        const bool ten = P.fmt == kRgb10;
        for (int k = 0; k < P.n; k++) {
            const uint32_t v = *(const uint32_t*)(P.src[k] + (int64_t)y * P.pitch[k] + 4 * (int64_t)x);
            const float w = P.w[k];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const uint32_t code = ten ? (c < 3 ? (v >> (10 * c)) & 1023u : v >> 30) : (v >> (8 * c)) & 255u;
                acc[c] = acc[c] + w * (float)code;
            }
        }
        uint32_t o;
        if (ten)
            o = (uint32_t)rnd(acc[0], P.wsum, 1023.0f) | ((uint32_t)rnd(acc[1], P.wsum, 1023.0f) << 10) |
                ((uint32_t)rnd(acc[2], P.wsum, 1023.0f) << 20) | ((uint32_t)rnd(acc[3], P.wsum, 3.0f) << 30);
        else
            o = (uint32_t)rnd(acc[0], P.wsum, 255.0f) | ((uint32_t)rnd(acc[1], P.wsum, 255.0f) << 8) |
                ((uint32_t)rnd(acc[2], P.wsum, 255.0f) << 16) | ((uint32_t)rnd(acc[3], P.wsum, 255.0f) << 24);
        *(uint32_t*)(P.out + (int64_t)y * P.opitch + 4 * (int64_t)x) = o;
        return;
    }
    if (P.fmt == kF16) {
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int k = 0; k < P.n; k++) {
            const uint2 q = *(const uint2*)(P.src[k] + (int64_t)y * P.pitch[k] + 8 * (int64_t)x);
            const uint32_t hv[4] = {q.x & 0xFFFFu, q.x >> 16, q.y & 0xFFFFu, q.y >> 16};
            const float w = P.w[k];
#pragma unroll
            for (int c = 0; c < 4; c++) acc[c] = acc[c] + w * (float)__builtin_bit_cast(_Float16, (uint16_t)hv[c]);
        }
        uint32_t h[4];
#pragma unroll
        for (int c = 0; c < 4; c++) h[c] = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)(acc[c] / P.wsum));
        uint2 o;
        o.x = h[0] | (h[1] << 16);
        o.y = h[2] | (h[3] << 16);
        *(uint2*)(P.out + (int64_t)y * P.opitch + 8 * (int64_t)x) = o;
        return;
    }
    // Nv12 / P010: the Y code of every pixel; the CbCr codes of each 2x2 block at its even-even pixel
    const bool p010 = P.fmt == kP010;
    const float top = p010 ? 1023.0f : 255.0f;
    float acc = 0.0f;
    for (int k = 0; k < P.n; k++) {
        const uint8_t* row = P.src[k] + (int64_t)y * P.pitch[k];
        acc = acc + P.w[k] * (float)(p010 ? ((const uint16_t*)row)[x] >> 6 : row[x]);
    }
    const uint32_t yc = (uint32_t)rnd(acc, P.wsum, top);
    if (p010) ((uint16_t*)(P.out + (int64_t)y * P.opitch))[x] = (uint16_t)(yc << 6);
    else P.out[(int64_t)y * P.opitch + x] = (uint8_t)yc;
    if (((x | y) & 1) == 0) {
        const int cx = x >> 1, cy = y >> 1;
        float cb = 0.0f, cr = 0.0f;
        for (int k = 0; k < P.n; k++) {
            if (p010) {
                const uint32_t v = *(const uint32_t*)(P.srcC[k] + (int64_t)cy * P.cpitch[k] + 4 * (int64_t)cx);
                cb = cb + P.w[k] * (float)((v & 0xFFFFu) >> 6);
                cr = cr + P.w[k] * (float)(v >> 22);
            } else {
                const uint32_t v = *(const uint16_t*)(P.srcC[k] + (int64_t)cy * P.cpitch[k] + 2 * (int64_t)cx);
                cb = cb + P.w[k] * (float)(v & 255u);
                cr = cr + P.w[k] * (float)(v >> 8);
            }
        }
        const uint32_t b = (uint32_t)rnd(cb, P.wsum, top), r = (uint32_t)rnd(cr, P.wsum, top);
        if (p010) *(uint32_t*)(P.outc + (int64_t)cy * P.ocpitch + 4 * (int64_t)cx) = (b << 6) | (r << 22);
        else *(uint16_t*)(P.outc + (int64_t)cy * P.ocpitch + 2 * (int64_t)cx) = (uint16_t)(b | (r << 8));
    }
}

}  // namespace fgr

using namespace fgr;

size_t FgrInterpolator::blurBytes(int n) const {
    const bool yuvFmt = format == FGR_FORMAT_P010 || format == FGR_FORMAT_NV12;
    const size_t plane = (size_t)W * H * bytesPerPixel();
    const size_t cplane = yuvFmt ? (size_t)((W + 1) / 2) * ((H + 1) / 2) * (format == FGR_FORMAT_NV12 ? 2 : 4) : 0;
    return (size_t)n * (plane + cplane);
}

bool FgrInterpolator::blurEnsure() {
    const int n = (int)set.blurSamples;
    if (n < 2 || !W) return true;
    const size_t need = blurBytes(n);
    if (need <= blurCap) return true;
    uint8_t* p = nullptr;
    if (hipMalloc(&p, need) != hipSuccess) return false;
    if (blurScratch) blurRetired.push_back(blurScratch);
    blurScratch = p;
    blurCap = need;
    return true;
}

bool FgrInterpolator::enqueueBlur(hipStream_t s, const FgrFrame& prev, const FgrFrame& curr, const FgrFrame& out,
                                  float t0, float t1) {
    const int n = (int)set.blurSamples;
    if (n < 2 || n > FGR_BLUR_MAX_SAMPLES) return false;
    // scratch: one sub-frame per sample (plane 0, then the chroma plane of Nv12 / P010), sized by blurEnsure when the
    // settings were applied
    const int bpp = bytesPerPixel();
    const bool yuvFmt = format == FGR_FORMAT_P010 || format == FGR_FORMAT_NV12;
    const int cbpp = format == FGR_FORMAT_NV12 ? 2 : 4;
    const size_t plane = (size_t)W * H * bpp;
    const size_t cplane = yuvFmt ? (size_t)((W + 1) / 2) * ((H + 1) / 2) * cbpp : 0;
    if (!blurScratch || blurBytes(n) > blurCap) return false;
    BlurParams bp{};
    bp.W = (int)W;
    bp.H = (int)H;
    bp.fmt = format;
    bp.n = n;
    bp.out = (uint8_t*)out.ptr;
    bp.outc = (uint8_t*)out.chroma;
    bp.opitch = (int64_t)out.pitch;
    bp.ocpitch = (int64_t)out.chromaPitch;
    const bool gauss = set.blurWeights == 1;
    float wsum = 0.0f;
    std::vector<PlanOp> plan;
    int k0 = synthK;
    for (int k = 0; k < n; k++) {
        const double td = (double)t0 + ((double)t1 - (double)t0) * (k + 0.5) / n;
        const float t = std::min(std::max((float)td, t0), t1);
        const double x = (2.0 * k + 1.0 - n) / n;
        bp.w[k] = gauss ? (float)std::exp(-2.0 * x * x) : 1.0f;
        wsum += bp.w[k];
        if (t <= 0.0f || t >= 1.0f) {   // the real frame
            const FgrFrame& f = t <= 0.0f ? prev : curr;
            bp.src[k] = (const uint8_t*)f.ptr;
            bp.srcC[k] = (const uint8_t*)f.chroma;
            bp.pitch[k] = (int64_t)f.pitch;
            bp.cpitch[k] = (int64_t)f.chromaPitch;
            continue;
        }
        FgrFrame sub{};
        uint8_t* base = blurScratch + (size_t)k * (plane + cplane);
        sub.ptr = base;
        sub.pitch = (uint64_t)W * bpp;
        if (cplane) {
            sub.chroma = base + plane;
            sub.chromaPitch = (uint64_t)((W + 1) / 2) * cbpp;
        }
        planSynth(plan, prev, curr, sub, t, k0++, false);   // (a protected pair: the patch follows the average)
        bp.src[k] = (const uint8_t*)sub.ptr;
        bp.srcC[k] = (const uint8_t*)sub.chroma;
        bp.pitch[k] = (int64_t)sub.pitch;
        bp.cpitch[k] = (int64_t)sub.chromaPitch;
    }
    bp.wsum = wsum;
    if (!runPlan(s, plan)) return false;
    hipLaunchKernelGGL(blur_average_kernel, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, s, bp);
    if (hipGetLastError() != hipSuccess) return false;
    stamp(s, "blur average");
    if (estUi) {   // framegen.cpp:3622-3628: curr's unchanged UI pixels over the blurred frame
        if (launchUi(kUiPatch, uiPatchParams(curr, out), s) != hipSuccess) return false;
        stamp(s, "ui_patch");
    }
    if (k0 > synthK) {
        lastSlot = (k0 - 1) % kSynthBanks;
        synthK = k0;
    }
    lastT = t1;
    return true;
}
