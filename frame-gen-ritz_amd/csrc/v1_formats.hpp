// Per-format pixel access shared by the format kernels (v1_formats.hip) and the static-UI passes (ui.hip): half floats,
// the fixed-point luma of each format and the two-plane YCbCr layouts.
#pragma once
#include <hip/hip_runtime.h>
#include "v1_device.hpp"
#include "v1_params.hpp"

namespace fgr {

__device__ __forceinline__ int floorDiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

__device__ __forceinline__ float halfBits(uint32_t h) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(h & 0xFFFFu));
}

__device__ __forceinline__ uint32_t toHalfBits(float v) {   // round to nearest even
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)v);
}

// Luma of pixel (x, y) in units of 1/16 level (luma.comp:9-27 y8() of each format, fixed point):
//   Rgb10   round((4899 R + 9617 G + 1868 B) * 255 / (1023 * 16384) * 16) on 10-bit codes (halves up)
//   RgbaF16 (4899 L_r + 9617 L_g + 1868 L_b + 8192) >> 14, L = the 1/16-level metric table of each half value
//   P010    round((4 code - 256) * 255 / 219) at limited range (16 (code / 4 - 16) * 255 / 219), 4 code at full range
//   Nv12    round((16 code - 256) * 255 / 219) at limited range (16 (code - 16) * 255 / 219), 16 code at full range
template <class LP>
__device__ __forceinline__ int luma16Rgb10(const LP& P, uint32_t w) {
    uint32_t r = w & 1023u, g = (w >> 10) & 1023u, b = (w >> 20) & 1023u;
    if (P.bgra) {
        const uint32_t t = r;
        r = b;
        b = t;
    }
    const uint32_t n = r * 4899u + g * 9617u + b * 1868u;   // <= 16760832: n * 255 + 523776 < 2^32
    return (int)((n * 255u + 523776u) / 1047552u);
}
template <class LP>
__device__ __forceinline__ int luma16F16(const LP& P, uint32_t lo, uint32_t hi) {
    const int lr = P.lut[lo & 0xFFFFu], lg = P.lut[lo >> 16], lb = P.lut[hi & 0xFFFFu];
    return (lr * 4899 + lg * 9617 + lb * 1868 + 8192) >> 14;   // |L| <= 65535: no overflow
}
template <class LP>
__device__ __forceinline__ int luma16Y(const LP& P, int code, bool ten) {   // a Y code of Nv12 (8-bit) or P010 (10-bit)
    if (ten) return P.yFull ? 4 * code : floorDiv((4 * code - 256) * 510 + 219, 438);
    return P.yFull ? 16 * code : floorDiv((16 * code - 256) * 510 + 219, 438);
}

template <class LP>   // LumaParams or PyrParams: the fmt, bgra, lut and yFull fields
__device__ __forceinline__ int luma16(const LP& P, const FrameRef& f, int x, int y) {
    const uint8_t* row = f.p + (int64_t)y * f.pitch;
    if (P.fmt == kRgb10) return luma16Rgb10(P, ((const uint32_t*)row)[x]);
    if (P.fmt == kF16) {
        const uint2 q = *(const uint2*)(row + 8 * (int64_t)x);
        return luma16F16(P, q.x, q.y);
    }
    if (P.fmt == kNv12) return luma16Y(P, (int)row[x], false);
    return luma16Y(P, (int)(((const uint16_t*)row)[x] >> 6), true);
}

// The lumas of the 4 pixels x .. x + 3 of row y (inside the frame): one 16-byte row load for Rgb10, two for RgbaF16
struct alignas(4) W4 {
    uint32_t x, y, z, w;
};
template <class LP>
__device__ __forceinline__ void luma16Row4(const LP& P, const FrameRef& f, int x, int y, int* l) {
    const uint8_t* row = f.p + (int64_t)y * f.pitch;
    if (P.fmt == kRgb10) {
        const W4 q = *(const W4*)(row + 4 * (int64_t)x);
        l[0] = luma16Rgb10(P, q.x);
        l[1] = luma16Rgb10(P, q.y);
        l[2] = luma16Rgb10(P, q.z);
        l[3] = luma16Rgb10(P, q.w);
    } else if (P.fmt == kF16) {
        const W4 q0 = *(const W4*)(row + 8 * (int64_t)x), q1 = *(const W4*)(row + 8 * (int64_t)x + 16);
        l[0] = luma16F16(P, q0.x, q0.y);
        l[1] = luma16F16(P, q0.z, q0.w);
        l[2] = luma16F16(P, q1.x, q1.y);
        l[3] = luma16F16(P, q1.z, q1.w);
    } else {   // the Y planes: their own element loads (their rows need not be 4-byte aligned)
#pragma unroll
        for (int i = 0; i < 4; i++) l[i] = luma16(P, f, x + i, y);
    }
}

// Two-plane YCbCr (TEN: P010's 10-bit codes in 16-bit words, else Nv12's bytes). Levels: P010 code / 4, Nv12 code.
template <bool TEN>
struct Yuv {
    static constexpr float kQ = TEN ? 4.0f : 1.0f;          // codes per level
    static constexpr int kTop = TEN ? 1023 : 255;
    static __device__ __forceinline__ int y(const uint8_t* row, int x) {
        return TEN ? ((const uint16_t*)row)[x] >> 6 : row[x];
    }
    static __device__ __forceinline__ void c(const uint8_t* row, int x, int& cb, int& cr) {
        if (TEN) {
            const uint32_t w = ((const uint32_t*)row)[x];
            cb = (int)((w & 0xFFFFu) >> 6);
            cr = (int)(w >> 22);
        } else {
            const uint32_t w = ((const uint16_t*)row)[x];
            cb = (int)(w & 255u);
            cr = (int)(w >> 8);
        }
    }
    static __device__ __forceinline__ void putY(uint8_t* row, int x, int code) {
        if (TEN) ((uint16_t*)row)[x] = (uint16_t)(code << 6);
        else row[x] = (uint8_t)code;
    }
    static __device__ __forceinline__ void putC(uint8_t* row, int x, int cb, int cr) {
        if (TEN) ((uint32_t*)row)[x] = ((uint32_t)cb << 6) | ((uint32_t)cr << 22);
        else ((uint16_t*)row)[x] = (uint16_t)(cb | (cr << 8));
    }
};

}  // namespace fgr
