// CrosshairV2 static-UI protection (gpu/shaders/ui_v2_update.comp, ui_v2_mask.comp, ui_patch.comp,
// ui_patch_chroma.comp, ui_v2_lumafill.comp; the host contract is gpu/framegen.h:575-693), every format.
//
// Per observed frame, over the work rectangle (the box grown by 1 px, clamped to the frame): the detection state
// machine of every pixel (one thread a pixel, in place: a thread touches only its own state words; the neighbours of the
// contrast test are read from the frame), then the frame's mask. After a protected synth, the pixels of mask byte 255
// (UI, unchanged in the pair) take curr's colour. The box is small (27 x 27 at 1080p by default), so these are latency
// kernels of one or a few workgroups.
//
// State words per pixel (4 a pixel, the fourth unused): w0 = previous original (r | g << 8 | b << 16) | count << 24,
// w1 = last confirmed colour | flags << 24 (bit 0 in the mask, bit 1 has a stored colour, bit 2 still in this pair),
// w2 = consecutive changed frames in the mask | consecutive stored-colour matches out of it << 8.
//
// Levels (format.glsl FG_STILL / fgStillYuv: integers 0..255, pinned exactly): Rgba8 the bytes; Rgb10
// floor(255 code / 1023 + 1/2) = (510 code + 1023) / 2046 (never a tie); RgbaF16 clamp(floor(metric + 1/2), 0, 255) of the
// 1/16-level metric table, (L16 + 8) >> 4; Nv12 the Y code and the pixel's 2x2 block's Cb, Cr codes; P010 (code + 2) >> 2
// clamped to 255. A 4:2:0 contrast converts the (Y, Cb, Cr) difference to R'G'B' first (fgYuvDiff).
#include <hip/hip_runtime.h>
#include "v1_device.hpp"
#include "v1_formats.hpp"
#include "v1_launch.hpp"

namespace fgr {

__device__ __forceinline__ uint32_t pack3(int a, int b, int c) { return (uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)c << 16); }

__device__ __forceinline__ uint32_t rgbAt(const UiParams& P, int x, int y) {   // the pixel's three levels, packed
    x = clampi(x, 0, P.W - 1);
    y = clampi(y, 0, P.H - 1);
    const uint8_t* row = P.f.p + (int64_t)y * P.f.pitch;
    switch (P.fmt) {
    case kRgb10: {
        const uint32_t w = ((const uint32_t*)row)[x];
        auto l = [](uint32_t c) { return (int)((510u * c + 1023u) / 2046u); };
        return pack3(l(w & 1023u), l((w >> 10) & 1023u), l((w >> 20) & 1023u));
    }
    case kF16: {
        const uint2 q = *(const uint2*)(row + 8 * (int64_t)x);
        auto l = [&](uint32_t h) { return clampi((P.lut[h & 0xFFFFu] + 8) >> 4, 0, 255); };
        return pack3(l(q.x), l(q.x >> 16), l(q.y));
    }
    case kNv12: {
        const uint32_t c = ((const uint16_t*)(P.f.c + (int64_t)(y >> 1) * P.f.cpitch))[x >> 1];
        return pack3(row[x], (int)(c & 255u), (int)(c >> 8));
    }
    case kP010: {
        const uint32_t c = ((const uint32_t*)(P.f.c + (int64_t)(y >> 1) * P.f.cpitch))[x >> 1];
        auto l = [](uint32_t code) { return min((int)((code + 2u) >> 2), 255); };
        return pack3(l(((const uint16_t*)row)[x] >> 6), l((c & 0xFFFFu) >> 6), l(c >> 22));
    }
    default: return *(const uint32_t*)(row + 4 * (int64_t)x) & 0x00FFFFFFu;
    }
}

__device__ __forceinline__ int chan(uint32_t v, int c) { return (int)((v >> (8 * c)) & 255u); }

__device__ __forceinline__ int maxAbs(uint32_t a, uint32_t b) {
    return max(abs(chan(a, 0) - chan(b, 0)), max(abs(chan(a, 1) - chan(b, 1)), abs(chan(a, 2) - chan(b, 2))));
}

// max channel |v - median7x7| around (x, y) (frame-clamped window): the median of 49 integer levels is the smallest m
// with at least 25 of them <= m, found per channel by bisection over 0..255 with the 49 values in registers.
__device__ __forceinline__ float contrastAt(const UiParams& P, int x, int y, uint32_t v) {
    uint32_t w[49];
#pragma unroll
    for (int j = 0; j < 7; j++)
#pragma unroll
        for (int i = 0; i < 7; i++) w[j * 7 + i] = rgbAt(P, x + i - 3, y + j - 3);
    float d[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        int lo = 0, hi = 255;
        for (int it = 0; it < 8; it++) {
            const int mid = (lo + hi) >> 1;
            int cnt = 0;
#pragma unroll
            for (int j = 0; j < 49; j++) cnt += chan(w[j], c) <= mid ? 1 : 0;
            if (cnt >= 25) hi = mid;
            else lo = mid + 1;
        }
        d[c] = (float)(chan(v, c) - lo);
    }
    if (P.fmt == kNv12 || P.fmt == kP010) {   // fgYuvDiff format.glsl:98-100, the mismatch metric's operation order
        const float ys = P.yuv[0], rcr = P.yuv[1], bcb = P.yuv[2], gcb = P.yuv[3], gcr = P.yuv[4];
        const float R = ys * d[0] + rcr * d[2];
        float G = ys * d[0] + gcb * d[1];
        G = G + gcr * d[2];
        const float B = ys * d[0] + bcb * d[1];
        return fmaxf(fabsf(R), fmaxf(fabsf(G), fabsf(B)));
    }
    return fmaxf(fabsf(d[0]), fmaxf(fabsf(d[1]), fabsf(d[2])));
}

__global__ __launch_bounds__(256) void ui_update_kernel(UiParams P) {
    const int px = (int)(blockIdx.x * 16 + (threadIdx.x & 15)), py = (int)(blockIdx.y * 16 + (threadIdx.x >> 4));
    if (px >= P.ww || py >= P.wh) return;
    const int x = P.wx + px, y = P.wy + py;
    const uint32_t now = rgbAt(P, x, y);
    uint32_t* st = P.state + 4 * ((int64_t)py * P.ww + px);
    const bool inBox = x >= P.bx && y >= P.by && x < P.bx + P.bw && y < P.by + P.bh;
    if (!inBox) {   // the margin ring: the pair gate only
        const uint32_t ring = (!P.reset && (float)maxAbs(now, st[0] & 0x00FFFFFFu) <= P.thr) ? 4u : 0u;
        st[0] = now;
        st[1] = ring << 24;
        st[2] = 0u;
        return;
    }
    uint32_t cnt = 0u, flags = 0u, leave = 0u, match = 0u, stored = 0u;
    if (!P.reset) {
        const uint32_t w0 = st[0], w1 = st[1], w2 = st[2];
        flags = w1 >> 24;
        stored = w1 & 0x00FFFFFFu;
        leave = w2 & 255u;
        match = (w2 >> 8) & 255u;
        const bool still = (float)maxAbs(now, w0 & 0x00FFFFFFu) <= P.thr;
        bool inMask = (flags & 1u) != 0u, hasStored = (flags & 2u) != 0u;
        const bool matchS = !inMask && hasStored && P.reentry > 0 && (float)maxAbs(now, stored) <= P.thr;
        bool contrasty = P.cthr <= 0.0f;
        if (!inMask && !contrasty && (still || matchS)) contrasty = contrastAt(P, x, y, now) >= P.cthr;
        if (inMask) {
            leave = still ? 0u : min(leave + 1u, 255u);
            if (still) stored = now;
            if ((int)leave >= P.leave) {
                inMask = false;
                leave = 0u;
                match = 0u;
            }
            cnt = 0u;
        } else {
            cnt = (still && contrasty) ? min((w0 >> 24) + 1u, 255u) : 0u;
            match = (matchS && contrasty) ? min(match + 1u, 255u) : 0u;
            if ((int)cnt >= P.still || (P.reentry > 0 && hasStored && (int)match >= P.reentry)) {
                inMask = true;
                hasStored = true;
                stored = now;
                leave = 0u;
                match = 0u;
            }
        }
        flags = (inMask ? 1u : 0u) | (hasStored ? 2u : 0u) | (still ? 4u : 0u);
    }
    st[0] = now | (cnt << 24);
    st[1] = stored | (flags << 24);
    st[2] = leave | (match << 8);
}

// ui_v2_mask.comp: 255 = within 1 px (8-neighbourhood) of a mask pixel and still in this pair, 64 = within 1 px but it
// changed in this pair, 0 = not UI
__global__ __launch_bounds__(256) void ui_mask_kernel(UiParams P) {
    const int px = (int)(blockIdx.x * 16 + (threadIdx.x & 15)), py = (int)(blockIdx.y * 16 + (threadIdx.x >> 4));
    if (px >= P.ww || py >= P.wh) return;
    bool near = false;
    for (int j = -1; j <= 1; j++)
        for (int i = -1; i <= 1; i++) {
            const int qx = px + i, qy = py + j;
            if (qx >= 0 && qy >= 0 && qx < P.ww && qy < P.wh)
                near = near || ((P.state[4 * ((int64_t)qy * P.ww + qx) + 1] >> 24) & 1u) != 0u;
        }
    const uint32_t f = P.state[4 * ((int64_t)py * P.ww + px) + 1] >> 24;
    P.mask[(int64_t)py * P.mpitch + px] = near ? ((f & 4u) ? 255 : 64) : 0;
}

// ui_patch.comp with `binary`: a pixel of mask byte 255 takes curr's colour; the generated alpha stays (4:2:0: the Y
// code; the chroma is ui_patch_chroma_kernel's)
__global__ __launch_bounds__(256) void ui_patch_kernel(UiParams P) {
    const int px = (int)(blockIdx.x * 16 + (threadIdx.x & 15)), py = (int)(blockIdx.y * 16 + (threadIdx.x >> 4));
    if (px >= P.ww || py >= P.wh) return;
    if (P.mask[(int64_t)py * P.mpitch + px] != 255) return;
    const int x = P.wx + px, y = P.wy + py;
    uint8_t* orow = P.out + (int64_t)y * P.opitch;
    const uint8_t* crow = P.f.p + (int64_t)y * P.f.pitch;
    if (P.fmt == kF16) {
        uint2* o = (uint2*)(orow + 8 * (int64_t)x);
        const uint2 c = *(const uint2*)(crow + 8 * (int64_t)x);
        *o = make_uint2(c.x, (c.y & 0xFFFFu) | (o->y & 0xFFFF0000u));
    } else if (P.fmt == kNv12) {
        orow[x] = crow[x];
    } else if (P.fmt == kP010) {
        ((uint16_t*)orow)[x] = ((const uint16_t*)crow)[x];
    } else {
        const uint32_t keep = P.fmt == kRgb10 ? 0xC0000000u : 0xFF000000u;
        uint32_t* o = (uint32_t*)(orow + 4 * (int64_t)x);
        *o = (((const uint32_t*)crow)[x] & ~keep) | (*o & keep);
    }
}

// ui_patch_chroma.comp with `binary`: a CbCr texel whose 2x2 pixels' largest mask byte (inside the rectangle) is 255
// takes curr's chroma. One thread a texel of the rectangle's chroma span.
__global__ __launch_bounds__(256) void ui_patch_chroma_kernel(UiParams P) {
    const int tx = (P.wx >> 1) + (int)(blockIdx.x * 16 + (threadIdx.x & 15));
    const int ty = (P.wy >> 1) + (int)(blockIdx.y * 16 + (threadIdx.x >> 4));
    if (tx > (P.wx + P.ww - 1) >> 1 || ty > (P.wy + P.wh - 1) >> 1) return;
    int a = 0;
    for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) {
            const int rx = 2 * tx + i - P.wx, ry = 2 * ty + j - P.wy;
            if (rx >= 0 && ry >= 0 && rx < P.ww && ry < P.wh) a = max(a, (int)P.mask[(int64_t)ry * P.mpitch + rx]);
        }
    if (a != 255) return;
    uint8_t* orow = P.outc + (int64_t)ty * P.ocpitch;
    const uint8_t* crow = P.f.c + (int64_t)ty * P.f.cpitch;
    if (P.fmt == kP010) ((uint32_t*)orow)[tx] = ((const uint32_t*)crow)[tx];
    else ((uint16_t*)orow)[tx] = ((const uint16_t*)crow)[tx];
}

// ui_v2_lumafill.comp (uiEstimateFill): each flow cell of the work rectangle that holds a UI pixel (mask byte != 0)
// gets, in the estimate's level-0 luma only, the mean of the integer lumas (luma.comp:23-26) of the not-UI pixels of the
// cell widened by S px a side per step up to `reach` px while it has none (float32 sum of integers: exact; one
// division). Cells without UI pixels keep what the luma pass wrote. One thread a cell.
__device__ __forceinline__ bool uiPx(const UiParams& P, int x, int y) {
    const int rx = x - P.wx, ry = y - P.wy;
    return rx >= 0 && ry >= 0 && rx < P.ww && ry < P.wh && P.mask[(int64_t)ry * P.mpitch + rx] != 0;
}

__global__ __launch_bounds__(256) void ui_fill_kernel(UiParams P) {
    const int S = P.S;
    const int cx0 = P.wx / S, cy0 = P.wy / S;
    const int gx = (int)(blockIdx.x * 16 + (threadIdx.x & 15)), gy = (int)(blockIdx.y * 16 + (threadIdx.x >> 4));
    const int ox = cx0 + gx, oy = cy0 + gy;
    if (ox >= P.fw || oy >= P.fh || ox > (P.wx + P.ww - 1) / S || oy > (P.wy + P.wh - 1) / S) return;
    const int px = ox * S, py = oy * S;
    bool hasUi = false;
    for (int j = 0; j < S; j++)
        for (int i = 0; i < S; i++)
            if (px + i < P.W && py + j < P.H) hasUi = hasUi || uiPx(P, px + i, py + j);
    if (!hasUi) return;
    LumaParams lp{};
    lp.fmt = P.fmt;
    lp.lut = P.lut;
    lp.bgra = P.bgra;
    lp.yFull = P.yFull;
    for (int k = 0; k <= P.reach; k += S) {
        float sum = 0.0f;
        int n = 0;
        for (int y = py - k; y < py + S + k; y++)
            for (int x = px - k; x < px + S + k; x++) {
                if (x < 0 || y < 0 || x >= P.W || y >= P.H || uiPx(P, x, y)) continue;
                if (P.fmt == kRgba8) {
                    const uint32_t v = *(const uint32_t*)(P.f.p + (int64_t)y * P.f.pitch + 4 * (int64_t)x);
                    const uint32_t r = P.bgra ? (v >> 16) & 255u : v & 255u, g = (v >> 8) & 255u,
                                   b = P.bgra ? v & 255u : (v >> 16) & 255u;
                    sum += (float)((r * 4899u + g * 9617u + b * 1868u + 8192u) >> 14);
                } else {
                    sum += (float)luma16(lp, P.f, x, y);
                }
                n++;
            }
        if (n > 0) {   // the other formats' luma is in 1/16 level: the mean is sum / (16 n)
            P.luma[(int64_t)oy * P.fw + ox] = sum / (float)(P.fmt == kRgba8 ? n : 16 * n);
            return;
        }
    }
}

static inline unsigned cdivu(int a, int b) { return (unsigned)((a + b - 1) / b); }

hipError_t launchUi(int pass, const UiParams& p, hipStream_t s) {
    if (pass == kUiFill) {   // one thread a flow cell of the work rectangle
        const int ncx = (p.wx + p.ww - 1) / p.S - p.wx / p.S + 1, ncy = (p.wy + p.wh - 1) / p.S - p.wy / p.S + 1;
        hipLaunchKernelGGL(ui_fill_kernel, dim3(cdivu(ncx, 16), cdivu(ncy, 16)), dim3(256), 0, s, p);
        return hipGetLastError();
    }
    const dim3 g(cdivu(p.ww, 16), cdivu(p.wh, 16));
    switch (pass) {
    case kUiUpdate: hipLaunchKernelGGL(ui_update_kernel, g, dim3(256), 0, s, p); break;
    case kUiMask: hipLaunchKernelGGL(ui_mask_kernel, g, dim3(256), 0, s, p); break;
    case kUiPatch:
        hipLaunchKernelGGL(ui_patch_kernel, g, dim3(256), 0, s, p);
        if (p.fmt == kNv12 || p.fmt == kP010) {
            const int ncx = ((p.wx + p.ww - 1) >> 1) - (p.wx >> 1) + 1, ncy = ((p.wy + p.wh - 1) >> 1) - (p.wy >> 1) + 1;
            hipLaunchKernelGGL(ui_patch_chroma_kernel, dim3(cdivu(ncx, 16), cdivu(ncy, 16)), dim3(256), 0, s, p);
        }
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace fgr
