// V1 luma and synth kernels of the formats other than RGBA8 (gpu/shaders/format.glsl:29-106, luma.comp:9-27,
// motion_v1.glsl:85-144, resolve.comp:48-83): Rgb10 (A2B10G10R10 / A2R10G10B10 UNORM), RgbaF16 (half floats, encoded
// or linear light), Nv12 (8-bit Y plane, half-resolution 8-bit CbCr plane) and P010 (the same with 16-bit words
// holding 10-bit codes). A level is an 8-bit-equivalent code: Nv12 codes as they are, 10-bit codes / 4.
//
// The reference reads these formats through the texture filter and leaves the luma unrounded, which no other GPU
// reproduces bit for bit. Here every step is fixed: the luma is a fixed-point value in 1/16 level; the warp taps are
// exact integer bilinear sums on the 1/32-px map of the RGBA8 path (the P010 chroma on a 1/64-texel map of its own
// plane); the mismatch, blend and stores are float32 expressions in the reference's order. The formulas, one per
// format, are stated at each function. A launch serves pairs of any of the three formats (the format is a parameter,
// uniform per workgroup), so a mixed batch needs one luma and one synth launch per pass for all of them.
#include <hip/hip_runtime.h>
#include "v1_device.hpp"
#include "v1_launch.hpp"
#include "v1_synth_common.hpp"
#include "v1_formats.hpp"
#include "v1_pyramid.hpp"

namespace fgr {

// ---------------------------------------------------------------------------------------------------------- luma

// luma.comp:58-72 on the fixed-point luma: float32(sum of the valid pixels' y16) / float32(16 n). Block (0, 0) also
// clears the estimate's counters (as the RGBA8 luma pass does).
__device__ __forceinline__ void lumaFmtBody(const LumaParams& P, int bx, int by) {
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    if (bx == 0 && by == 0)
        for (int i = threadIdx.x; i < kUWords; i += 256) P.u[i] = 0u;
    const int ox = bx * 16 + tx, oy = by * 16 + ty;
    if (ox >= P.fw || oy >= P.fh) return;
    int sa = 0, sc = 0, n = 0;
    for (int j = 0; j < P.S; j++) {
        const int py = P.S * oy + j;
        if (py >= P.H) break;
        for (int i = 0; i < P.S; i++) {
            const int px = P.S * ox + i;
            if (px >= P.W) break;
            sa += luma16(P, P.a, px, py);
            sc += luma16(P, P.c, px, py);
            n++;
        }
    }
    P.la[oy * P.fw + ox] = (float)sa / (float)(16 * n);
    P.lc[oy * P.fw + ox] = (float)sc / (float)(16 * n);
}

// The fused pyramid (v1_pyramid.hpp) of the other formats: level 0 of the thread's 2 x 2 cells as lumaFmtBody forms it
__device__ __forceinline__ void pyramidFmtBody(const PyrParams& P, int bx, int by) {
    if (bx == 0 && by == 0)
        for (int i = threadIdx.x; i < kUWords; i += 256) P.u[i] = 0u;
    const int cx0 = bx * 32 + 2 * (threadIdx.x & 15), cy0 = by * 32 + 2 * (threadIdx.x >> 4);
    float a[2][2], c[2][2];
    if (P.S == 2 && 2 * cx0 + 3 <= P.W - 1 && 2 * cy0 + 3 <= P.H - 1) {
        // flowScale 2 inside the frame: the thread's 4 x 4 pixels as four row loads a frame, the same integer sums
        int la[4][4], lc[4][4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            luma16Row4(P, P.a, 2 * cx0, 2 * cy0 + j, la[j]);
            luma16Row4(P, P.c, 2 * cx0, 2 * cy0 + j, lc[j]);
        }
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int sa = la[2 * r][2 * i] + la[2 * r][2 * i + 1] + la[2 * r + 1][2 * i] + la[2 * r + 1][2 * i + 1];
                const int sc = lc[2 * r][2 * i] + lc[2 * r][2 * i + 1] + lc[2 * r + 1][2 * i] + lc[2 * r + 1][2 * i + 1];
                a[r][i] = (float)sa / (float)(16 * 4);
                c[r][i] = (float)sc / (float)(16 * 4);
                pyrStore(P.la[0], P.w[0], P.h[0], cx0 + i, cy0 + r, a[r][i]);
                pyrStore(P.lc[0], P.w[0], P.h[0], cx0 + i, cy0 + r, c[r][i]);
            }
    } else {
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int i = 0; i < 2; i++) {
            int sa = 0, sc = 0, n = 0;
            for (int j = 0; j < P.S; j++) {
                const int py = P.S * (cy0 + r) + j;
                if (py >= P.H) break;
                for (int q = 0; q < P.S; q++) {
                    const int px = P.S * (cx0 + i) + q;
                    if (px >= P.W) break;
                    sa += luma16(P, P.a, px, py);
                    sc += luma16(P, P.c, px, py);
                    n++;
                }
            }
            a[r][i] = n ? (float)sa / (float)(16 * n) : 0.0f;
            c[r][i] = n ? (float)sc / (float)(16 * n) : 0.0f;
            pyrStore(P.la[0], P.w[0], P.h[0], cx0 + i, cy0 + r, a[r][i]);
            pyrStore(P.lc[0], P.w[0], P.h[0], cx0 + i, cy0 + r, c[r][i]);
        }
    }
    pyramidLevels(P, bx, by, a, c);   // (its barriers reached by every thread, after either form)
}

// ---------------------------------------------------------------------------------------------------------- taps

// The clamped taps (BORDER_REPLICATE) and integer weights of a bilinear sample at map (X, Y) in 1/2^SH texel units
template <int SH>
struct Taps {
    int x0, x1, y0, y1;
    int w00, w01, w10, w11;
    __device__ __forceinline__ Taps(int X, int Y, int W, int H) {
        constexpr int unit = 1 << SH;
        const int ix = X >> SH, iy = Y >> SH, fx = X & (unit - 1), fy = Y & (unit - 1);
        x0 = med3i(ix, 0, W - 1);
        x1 = med3i(ix + 1, 0, W - 1);
        y0 = med3i(iy, 0, H - 1);
        y1 = med3i(iy + 1, 0, H - 1);
        const int gx = unit - fx, gy = unit - fy;
        w00 = gx * gy;
        w01 = fx * gy;
        w10 = gx * fy;
        w11 = fx * fy;
    }
};

__device__ __forceinline__ int isum(const Taps<5>& t, int p00, int p01, int p10, int p11) {
    return t.w00 * p00 + t.w01 * p01 + t.w10 * p10 + t.w11 * p11;
}

// Rgb10: exact bilinear sums of the three 10-bit channels (code * 1024 units)
__device__ __forceinline__ void rgb10Sums(const FrameRef& f, int W, int H, int X, int Y, int s[3]) {
    const Taps<5> t(X, Y, W, H);
    const uint32_t* r0 = (const uint32_t*)(f.p + (int64_t)t.y0 * f.pitch);
    const uint32_t* r1 = (const uint32_t*)(f.p + (int64_t)t.y1 * f.pitch);
    const uint32_t p00 = r0[t.x0], p01 = r0[t.x1], p10 = r1[t.x0], p11 = r1[t.x1];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int sh = 10 * c;
        s[c] = isum(t, (int)((p00 >> sh) & 1023u), (int)((p01 >> sh) & 1023u), (int)((p10 >> sh) & 1023u),
                    (int)((p11 >> sh) & 1023u));
    }
}

// RgbaF16: the four float32 products w * v (exact) summed in the order p00, p01, p10, p11, times 1/1024
__device__ __forceinline__ void f16Warp(const FrameRef& f, int W, int H, int X, int Y, float v[3]) {
    const Taps<5> t(X, Y, W, H);
    const uint8_t* r0 = f.p + (int64_t)t.y0 * f.pitch;
    const uint8_t* r1 = f.p + (int64_t)t.y1 * f.pitch;
    const uint2 q00 = *(const uint2*)(r0 + 8 * t.x0), q01 = *(const uint2*)(r0 + 8 * t.x1);
    const uint2 q10 = *(const uint2*)(r1 + 8 * t.x0), q11 = *(const uint2*)(r1 + 8 * t.x1);
    const float w00 = (float)t.w00, w01 = (float)t.w01, w10 = (float)t.w10, w11 = (float)t.w11;
    const uint32_t a[3] = {q00.x, q00.x >> 16, q00.y}, b[3] = {q01.x, q01.x >> 16, q01.y};
    const uint32_t c[3] = {q10.x, q10.x >> 16, q10.y}, d[3] = {q11.x, q11.x >> 16, q11.y};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float s = w00 * halfBits(a[k]) + w01 * halfBits(b[k]);
        s = s + w10 * halfBits(c[k]);
        s = s + w11 * halfBits(d[k]);
        v[k] = s * (1.0f / 1024.0f);
    }
}

// luma: exact bilinear sum of the Y codes (level = sum / (1024 q))
template <bool TEN>
__device__ __forceinline__ int yuvYSum(const FrameRef& f, int W, int H, int X, int Y) {
    const Taps<5> t(X, Y, W, H);
    const uint8_t* r0 = f.p + (int64_t)t.y0 * f.pitch;
    const uint8_t* r1 = f.p + (int64_t)t.y1 * f.pitch;
    return isum(t, Yuv<TEN>::y(r0, t.x0), Yuv<TEN>::y(r0, t.x1), Yuv<TEN>::y(r1, t.x0), Yuv<TEN>::y(r1, t.x1));
}

// chroma at luma map (X, Y) (1/32 px): with centre-sited chroma on a half-resolution plane the point is (X - 16) / 64
// chroma texels; exact bilinear sums of the Cb and Cr codes (level = sum / (4096 q))
template <bool TEN>
__device__ __forceinline__ void yuvCSums(const FrameRef& f, int cw, int ch, int X, int Y, int& cb, int& cr) {
    const Taps<6> t(X - 16, Y - 16, cw, ch);
    const uint8_t* r0 = f.c + (int64_t)t.y0 * f.cpitch;
    const uint8_t* r1 = f.c + (int64_t)t.y1 * f.cpitch;
    int b00, r00, b01, r01, b10, r10, b11, r11;
    Yuv<TEN>::c(r0, t.x0, b00, r00);
    Yuv<TEN>::c(r0, t.x1, b01, r01);
    Yuv<TEN>::c(r1, t.x0, b10, r10);
    Yuv<TEN>::c(r1, t.x1, b11, r11);
    cb = t.w00 * b00 + t.w01 * b01 + t.w10 * b10 + t.w11 * b11;
    cr = t.w00 * r00 + t.w01 * r01 + t.w10 * r10 + t.w11 * r11;
}

// fgYuvDiff format.glsl:98-100: a (Y, Cb, Cr) level difference as the R'G'B' difference, the largest magnitude
__device__ __forceinline__ float yuvDiffMax(const SynthParams& P, float dY, float dCb, float dCr) {
    const float ys = P.yuv[0], rcr = P.yuv[1], bcb = P.yuv[2], gcb = P.yuv[3], gcr = P.yuv[4];
    const float R = ys * dY + rcr * dCr;
    float G = ys * dY + gcb * dCb;
    G = G + gcr * dCr;
    const float B = ys * dY + bcb * dCb;
    return fmaxf(fabsf(R), fmaxf(fabsf(G), fabsf(B)));
}

// The 1/32 maps of a and c at pixel (x, y) (as the RGBA8 synth forms them: 32 p -+ (32 S k) r)
struct Maps {
    int XA, YA, XC, YC;
};

__device__ __forceinline__ Maps mapsAt(float kA32, float kC32, float x32, float y32, float2 r) {
    Maps m;
    m.XA = (int)rintf(x32 - kA32 * r.x);
    m.YA = (int)rintf(y32 - kA32 * r.y);
    m.XC = (int)rintf(x32 + kC32 * r.x);
    m.YC = (int)rintf(y32 + kC32 * r.y);
    return m;
}

// resolve.comp:73-83 blend in float32, in two steps: the mix m = 0.5 (wa + wc) at t = 1/2, else wA wa + wC wc (taken
// with the mismatch, where the warped values are at hand), then o = m (1 - u) + near u once the trust u is known
__device__ __forceinline__ float mixF(const SynthParams& P, float wa, float wc) {
    return P.mid ? 0.5f * (wa + wc) : P.wA * wa + P.wC * wc;
}

__device__ __forceinline__ float finishF(float m, float nv, float u) { return m * (1.0f - u) + nv * u; }

// Mismatch e in 0..255 (error.comp / mismatch() with format.glsl's FG_DIFF and FG_PACKE):
//   Rgb10   uint(min(max_c |wa - wc|, 255)), levels w = float32(sum) * float32(255 / 1047552)
//   RgbaF16 min(max_c |L(half(wa)) - L(half(wc))| >> 4, 255)
//   Nv12, P010 uint(min(max |fgYuvDiff(Y, Cb, Cr difference)|, 255))
// (UI: with the CrosshairV2 tap rejection of a protected pair, on the raw sums: motion_v1.glsl:124, 141)
// MIX: also the blend's mix of the warped values into mix[] (the Y level for Nv12 / P010, the three channels else)
template <bool TEN, bool UI, bool MIX>
__device__ __forceinline__ int mismatchYuv(const SynthParams& P, const Maps& m, float* mix) {
    int ya = yuvYSum<TEN>(P.a, P.W, P.H, m.XA, m.YA), yc = yuvYSum<TEN>(P.c, P.W, P.H, m.XC, m.YC);
    int cba, cra, cbc, crc;
    yuvCSums<TEN>(P.a, P.cw, P.ch, m.XA, m.YA, cba, cra);
    yuvCSums<TEN>(P.c, P.cw, P.ch, m.XC, m.YC, cbc, crc);
    if (UI) {
        const int w = uiRejectWhich<true>(P, m.XA, m.YA, m.XC, m.YC);
        if (w == 1) {
            ya = yc;
            cba = cbc;
            cra = crc;
        } else if (w == 2) {
            yc = ya;
            cbc = cba;
            crc = cra;
        }
    }
    const float ky = 1.0f / (1024.0f * Yuv<TEN>::kQ), kc = 1.0f / (4096.0f * Yuv<TEN>::kQ);
    const float dY = (float)(ya - yc) * ky;
    const float dCb = (float)(cba - cbc) * kc, dCr = (float)(cra - crc) * kc;
    if (MIX) mix[0] = mixF(P, (float)ya * ky, (float)yc * ky);
    return (int)fminf(yuvDiffMax(P, dY, dCb, dCr), 255.0f);
}

template <int F, bool UI = false>
__device__ __forceinline__ void uiRejectArrays(const SynthParams& P, const Maps& m, int* sa, int* sc) {
    if (!UI) return;
    const int w = uiRejectWhich(P, m.XA, m.YA, m.XC, m.YC);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (w == 1) sa[c] = sc[c];
        else if (w == 2) sc[c] = sa[c];
    }
}

template <bool UI>
__device__ __forceinline__ void uiRejectArraysF(const SynthParams& P, const Maps& m, float* va, float* vc) {
    if (!UI) return;
    const int w = uiRejectWhich(P, m.XA, m.YA, m.XC, m.YC);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (w == 1) va[c] = vc[c];
        else if (w == 2) vc[c] = va[c];
    }
}

// the number of mixes a pixel keeps for its blend: the three channels, or the Y level of Nv12 / P010
template <int F>
constexpr int kMixes = (F == kRgb10 || F == kF16) ? 3 : 1;

template <int F, bool UI = false, bool MIX = false>
__device__ __forceinline__ int mismatchFmt(const SynthParams& P, const Maps& m, float* mix = nullptr) {
    if (F == kRgb10) {
        int sa[3], sc[3];
        rgb10Sums(P.a, P.W, P.H, m.XA, m.YA, sa);
        rgb10Sums(P.c, P.W, P.H, m.XC, m.YC, sc);
        uiRejectArrays<F, UI>(P, m, sa, sc);
        const float k = (float)(255.0 / 1047552.0);
        float d = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float wa = (float)sa[c] * k, wc = (float)sc[c] * k;
            d = fmaxf(d, fabsf(wa - wc));
            if (MIX) mix[c] = mixF(P, wa, wc);
        }
        return (int)fminf(d, 255.0f);
    }
    if (F == kF16) {
        float va[3], vc[3];
        f16Warp(P.a, P.W, P.H, m.XA, m.YA, va);
        f16Warp(P.c, P.W, P.H, m.XC, m.YC, vc);
        uiRejectArraysF<UI>(P, m, va, vc);
        int d = 0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            d = max(d, abs(P.lut[toHalfBits(va[c])] - P.lut[toHalfBits(vc[c])]));
            if (MIX) mix[c] = mixF(P, va[c], vc[c]);
        }
        return min(d >> 4, 255);
    }
    return mismatchYuv<F == kP010, UI, MIX>(P, m, mix);
}

// resolve.comp:73-83 blend in float32: m = 0.5 (wa + wc) at t = 1/2, else wA wa + wC wc; o = m (1 - u) + near u
__device__ __forceinline__ float blendF(const SynthParams& P, float wa, float wc, float nv, float u) {
    return finishF(mixF(P, wa, wc), nv, u);
}

// One output pixel (x, y) with trust u and the mixes its mismatch kept:
//   Rgb10   code = floor(clamp(o, 0, 255) * float32(1023 / 255) + 0.5), alpha of the nearer frame
//   RgbaF16 half(o), alpha of the nearer frame
//   Nv12 / P010 Y code = clamp(floor(q o + 0.5), 0, top) (q codes per level); at an even-even pixel also the CbCr
//           of its 2x2 block, warped at the block centre p + 1/2 (warpChroma motion_v1.glsl:134-144) and blended with
//           this pixel's trust
template <bool TEN, bool UI>
__device__ __forceinline__ void resolveYuv(const SynthParams& P, float kA32, float kC32, int x, int y, float2 r, float u,
                                           float my) {
    using T = Yuv<TEN>;
    const FrameRef& nf = P.nearC ? P.c : P.a;
    const float kc = 1.0f / (4096.0f * T::kQ), kn = 1.0f / T::kQ;
    const float ny = (float)T::y(nf.p + (int64_t)y * nf.pitch, x) * kn;
    const float oy = finishF(my, ny, u);
    T::putY(P.out + (int64_t)y * P.opitch, x, med3i((int)floorf(oy * T::kQ + 0.5f), 0, T::kTop));
    if (((x | y) & 1) == 0) {
        const Maps b = mapsAt(kA32, kC32, (float)(32 * x + 16), (float)(32 * y + 16), r);
        int cba, cra, cbc, crc, nb, nr;
        yuvCSums<TEN>(P.a, P.cw, P.ch, b.XA, b.YA, cba, cra);
        yuvCSums<TEN>(P.c, P.cw, P.ch, b.XC, b.YC, cbc, crc);
        if (UI) {   // warpChroma's rejection at the block centre (motion_v1.glsl:141)
            const int w = uiRejectWhich<true>(P, b.XA, b.YA, b.XC, b.YC);
            if (w == 1) {
                cba = cbc;
                cra = crc;
            } else if (w == 2) {
                cbc = cba;
                crc = cra;
            }
        }
        const int cx = x >> 1, cy = y >> 1;
        T::c(nf.c + (int64_t)cy * nf.cpitch, cx, nb, nr);
        const float ob = blendF(P, (float)cba * kc, (float)cbc * kc, (float)nb * kn, u);
        const float orr = blendF(P, (float)cra * kc, (float)crc * kc, (float)nr * kn, u);
        T::putC(P.outc + (int64_t)cy * P.ocpitch, cx, med3i((int)floorf(ob * T::kQ + 0.5f), 0, T::kTop),
                med3i((int)floorf(orr * T::kQ + 0.5f), 0, T::kTop));
    }
}

template <int F, bool UI = false>
__device__ __forceinline__ void resolveFmt(const SynthParams& P, float kA32, float kC32, int x, int y, float2 r, float u,
                                           const float* mix) {
    const FrameRef& nf = P.nearC ? P.c : P.a;
    if (F == kRgb10) {
        const uint32_t nw = *(const uint32_t*)(nf.p + (int64_t)y * nf.pitch + 4 * (int64_t)x);
        const float kn = (float)(255.0 / 1023.0), ko = (float)(1023.0 / 255.0);
        uint32_t o = nw & 0xC0000000u;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float nv = (float)((nw >> (10 * c)) & 1023u) * kn;
            const float v = finishF(mix[c], nv, u);
            o |= (uint32_t)(int)floorf(fminf(fmaxf(v, 0.0f), 255.0f) * ko + 0.5f) << (10 * c);
        }
        *(uint32_t*)(P.out + (int64_t)y * P.opitch + 4 * (int64_t)x) = o;
        return;
    }
    if (F == kF16) {
        const uint2 nq = *(const uint2*)(nf.p + (int64_t)y * nf.pitch + 8 * (int64_t)x);
        const float n0 = halfBits(nq.x), n1 = halfBits(nq.x >> 16), n2 = halfBits(nq.y);
        uint2 o;
        o.x = toHalfBits(finishF(mix[0], n0, u)) | (toHalfBits(finishF(mix[1], n1, u)) << 16);
        o.y = toHalfBits(finishF(mix[2], n2, u)) | (nq.y & 0xFFFF0000u);
        *(uint2*)(P.out + (int64_t)y * P.opitch + 8 * (int64_t)x) = o;
        return;
    }
    resolveYuv<F == kP010, UI>(P, kA32, kC32, x, y, r, u, mix[0]);
}

// the nearer frame's pixel (x, y) into the output (scene cut): plane 0, and the chroma of an even-even pixel (P010)
template <int F>
__device__ __forceinline__ void copyNear(const SynthParams& P, int x, int y) {
    const FrameRef& nf = P.nearC ? P.c : P.a;
    if (F == kRgb10) {
        *(uint32_t*)(P.out + (int64_t)y * P.opitch + 4 * (int64_t)x) = ld32(nf, x, y);
    } else if (F == kF16) {
        *(uint2*)(P.out + (int64_t)y * P.opitch + 8 * (int64_t)x) = *(const uint2*)(nf.p + (int64_t)y * nf.pitch + 8 * (int64_t)x);
    } else if (F == kP010) {
        ((uint16_t*)(P.out + (int64_t)y * P.opitch))[x] = ((const uint16_t*)(nf.p + (int64_t)y * nf.pitch))[x];
        if (((x | y) & 1) == 0)
            *(uint32_t*)(P.outc + (int64_t)(y >> 1) * P.ocpitch + 4 * (int64_t)(x >> 1)) =
                *(const uint32_t*)(nf.c + (int64_t)(y >> 1) * nf.cpitch + 4 * (int64_t)(x >> 1));
    } else {
        P.out[(int64_t)y * P.opitch + x] = nf.p[(int64_t)y * nf.pitch + x];
        if (((x | y) & 1) == 0)
            *(uint16_t*)(P.outc + (int64_t)(y >> 1) * P.ocpitch + 2 * (int64_t)(x >> 1)) =
                *(const uint16_t*)(nf.c + (int64_t)(y >> 1) * nf.cpitch + 2 * (int64_t)(x >> 1));
    }
}

// One 64 x RT output tile (RT = 16, 32 or 64 rows; the RGBA8 kernel's structure): field tile, mismatch of the tile and
// its 3-pixel apron into LDS, 7x7 box, trust, untrusted count, blend. A thread warps the RT / 4 pixels it resolves
// first and keeps their mixes in registers (three floats, or the Y level's one), so the blend reads only the nearer
// frame; the apron's 2 x 3 x 70 + 6 RT pixels follow, spread over the workgroup.
template <int S, int F, bool UI = false, int RT = TH>
__device__ __forceinline__ void synthFmtTile(const SynthParams& P, int tileX, int tileY, float2* ft, float2* hrow,
                                             uint8_t* es, uint16_t* hs, uint32_t* cnt) {
    const int tid = threadIdx.x;
    const int x0 = tileX * TW, y0 = tileY * RT;
    const int cx = tid & (TW - 1), ry = tid >> 6;
    if (sceneCut(P)) {
        for (int k = 0; k < RT / 4; k++) {
            const int x = x0 + cx, y = y0 + ry + 4 * k;
            if (x < P.W && y < P.H) copyNear<F>(P, x, y);
        }
        return;
    }
    if (tid == 0) *cnt = 0u;
    const int ty0 = loadFieldTile<S, RT>(P, x0, y0, ft, hrow);
    const float kA32 = (float)(32 * S) * (2.0f * P.wC), kC32 = (float)(32 * S) * (2.0f * P.wA);
    // the tile's own pixels (region column cx + AP, row ly + AP; clamped into the image as the apron is)
    float mix[RT / 4][kMixes<F>];
#pragma unroll
    for (int k = 0; k < RT / 4; k++) {
        const int ly = ry + 4 * k;
        const int x = min(x0 + cx, P.W - 1), y = min(y0 + ly, P.H - 1);
        const float2 r = fieldAtRegion<S>(P, hrow, ty0, cx + AP, y);
        es[(ly + AP) * EWB + cx + AP] =
            (uint8_t)mismatchFmt<F, UI, true>(P, mapsAt(kA32, kC32, (float)(32 * x), (float)(32 * y), r), mix[k]);
    }
    // the apron: AP rows above and below the tile (EW wide), then AP columns either side of its RT rows
    constexpr int kRowsAp = 2 * AP * EW, kApron = kRowsAp + 2 * AP * RT;
    for (int i = tid; i < kApron; i += 256) {
        int ex, ey;
        if (i < kRowsAp) {
            const int j = i / EW;
            ex = i - j * EW;
            ey = j < AP ? j : j + RT;
        } else {
            const int j = (i - kRowsAp) / (2 * AP), c = i - kRowsAp - j * (2 * AP);
            ey = j + AP;
            ex = c < AP ? c : c + TW;
        }
        const int x = med3i(x0 - AP + ex, 0, P.W - 1), y = med3i(y0 - AP + ey, 0, P.H - 1);
        const float2 r = fieldAtRegion<S>(P, hrow, ty0, ex, y);
        es[ey * EWB + ex] = (uint8_t)mismatchFmt<F, UI>(P, mapsAt(kA32, kC32, (float)(32 * x), (float)(32 * y), r));
    }
    __syncthreads();
    boxRows7B<RT + 2 * AP>(es, hs, tid);
    __syncthreads();
    uint32_t mine = 0u;
#pragma unroll
    for (int k = 0; k < RT / 4; k++) {
        const int ly = ry + 4 * k, x = x0 + cx, y = y0 + ly;
        if (x >= P.W || y >= P.H) continue;
        const uint16_t* h = &hs[ly * TW + cx];
        const int S49 = (int)h[0] + h[TW] + h[2 * TW] + h[3 * TW] + h[4 * TW] + h[5 * TW] + h[6 * TW];
        mine += S49 > P.sHalf ? 1u : 0u;
        // the field vector only for the chroma of an even-even Nv12 / P010 pixel
        constexpr bool kYuv = F == kNv12 || F == kP010;
        const float2 r = kYuv && ((x | y) & 1) == 0 ? fieldAtRegion<S>(P, hrow, ty0, cx + AP, y) : make_float2(0.0f, 0.0f);
        // a wavefront whose pixels are all trusted (S <= sZero: u = 0) skips the trust ramp's per-pixel branches
        const float u = __all(S49 <= P.sZero) ? 0.0f : trustOf(P, S49);
        resolveFmt<F, UI>(P, kA32, kC32, x, y, r, u, mix[k]);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mine += __shfl_xor(mine, o, 64);
    if ((tid & 63) == 0 && mine) atomicAdd(cnt, mine);
    __syncthreads();
    if (tid == 0 && *cnt) atomicAdd(P.count + ((tileY * P.gx + tileX) & (kSlots - 1)), *cnt);
}

template <int S, bool UI = false, int RT = TH>
__device__ __forceinline__ void synthFmtBody(const SynthParams& P, int bx, int by) {
    constexpr int FTW = SynthTile<S, RT>::FTW, FTH = SynthTile<S, RT>::FTH, EHR = RT + 2 * AP;
    __shared__ float2 ft[FTW * FTH];
    __shared__ float2 hrow[FTH * EW];
    __shared__ __align__(16) uint8_t es[EHR * EWB];   // the mismatch (0 .. 255), one byte a pixel
    __shared__ __align__(16) uint16_t hs[EHR * TW];
    __shared__ uint32_t cnt;
    if (P.fmt == kRgb10) synthFmtTile<S, kRgb10, UI, RT>(P, bx, by, ft, hrow, es, hs, &cnt);
    else if (P.fmt == kF16) synthFmtTile<S, kF16, UI, RT>(P, bx, by, ft, hrow, es, hs, &cnt);
    else if (P.fmt == kNv12) synthFmtTile<S, kNv12, UI, RT>(P, bx, by, ft, hrow, es, hs, &cnt);
    else synthFmtTile<S, kP010, UI, RT>(P, bx, by, ft, hrow, es, hs, &cnt);
}

FGR_PASS(luma_fmt, LumaParams, lumaFmtBody, 256)
FGR_PASS(pyramid_fmt, PyrParams, pyramidFmtBody, 256)
FGR_PASS(synth_fmt_s1, SynthParams, synthFmtBody<1>, 256)
FGR_PASS_OCC(synth_fmt_s2, SynthParams, synthFmtBody<2>, 256, 5)   // 5 workgroups a CU (24.6 KB of LDS each)
FGR_PASS_OCC(synth_fmt_s4, SynthParams, synthFmtBody<4>, 256, 5)
#define FGR_SYNTH_FMT_ROWS_BODY(S, RT) synthFmtBody<S, false, RT>
FGR_PASS(synth_fmt_s2_t16, SynthParams, FGR_SYNTH_FMT_ROWS_BODY(2, 16), 256)
FGR_PASS(synth_fmt_s2_t64, SynthParams, FGR_SYNTH_FMT_ROWS_BODY(2, 64), 256)
#define FGR_SYNTH_FMT_UI_BODY(S) synthFmtBody<S, true>
FGR_PASS_SINGLE(synth_fmt_ui_s1, SynthParams, FGR_SYNTH_FMT_UI_BODY(1), 256)
FGR_PASS_SINGLE(synth_fmt_ui_s2, SynthParams, FGR_SYNTH_FMT_UI_BODY(2), 256)
FGR_PASS_SINGLE(synth_fmt_ui_s4, SynthParams, FGR_SYNTH_FMT_UI_BODY(4), 256)

// the packed mismatch e of the latest synth (debug maps), 4 pixels per uint32, low byte first
__global__ __launch_bounds__(256) void debug_e_fmt_kernel(SynthParams P, uint32_t* dst, int eStride) {
    const int gx = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (gx >= eStride || y >= P.H) return;
    const float kA32 = 32.0f * (2.0f * P.wC), kC32 = 32.0f * (2.0f * P.wA);
    uint32_t packed = 0u;
    for (int i = 0; i < 4; i++) {
        const int x = gx * 4 + i;
        if (x >= P.W) continue;
        // hv itself (not r): the per-pixel halfVec, whose factor S is folded into the map constants as 1
        const float2 hv = halfVec(P.field, x, y, P.S, P.fw, P.fh);
        const Maps m = mapsAt(kA32, kC32, (float)(32 * x), (float)(32 * y), hv);
        int e;
        if (P.uiA.w > 0 || P.uiC.w > 0) {   // a protected pair: with the tap rejection
            if (P.fmt == kRgb10) e = mismatchFmt<kRgb10, true>(P, m);
            else if (P.fmt == kF16) e = mismatchFmt<kF16, true>(P, m);
            else if (P.fmt == kNv12) e = mismatchFmt<kNv12, true>(P, m);
            else e = mismatchFmt<kP010, true>(P, m);
        } else {
            if (P.fmt == kRgb10) e = mismatchFmt<kRgb10>(P, m);
            else if (P.fmt == kF16) e = mismatchFmt<kF16>(P, m);
            else if (P.fmt == kNv12) e = mismatchFmt<kNv12>(P, m);
            else e = mismatchFmt<kP010>(P, m);
        }
        packed |= (uint32_t)e << (8 * i);
    }
    dst[(int64_t)y * eStride + gx] = packed;
}

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

hipError_t launchFmtSingle(int pass, const void* params, int gx, int gy, hipStream_t s) {
    const dim3 g((unsigned)gx, (unsigned)gy);
    switch (pass) {
    case kPassLumaFmt: hipLaunchKernelGGL(luma_fmt_kernel, g, dim3(256), 0, s, *(const LumaParams*)params); break;
    case kPassPyramidFmt: hipLaunchKernelGGL(pyramid_fmt_kernel, g, dim3(256), 0, s, *(const PyrParams*)params); break;
    case kPassSynthFmtS1: hipLaunchKernelGGL(synth_fmt_s1_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthFmtS2: hipLaunchKernelGGL(synth_fmt_s2_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthFmtS4: hipLaunchKernelGGL(synth_fmt_s4_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthFmtS2T16: hipLaunchKernelGGL(synth_fmt_s2_t16_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthFmtS2T64: hipLaunchKernelGGL(synth_fmt_s2_t64_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthFmtUiS1: hipLaunchKernelGGL(synth_fmt_ui_s1_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthFmtUiS2: hipLaunchKernelGGL(synth_fmt_ui_s2_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthFmtUiS4: hipLaunchKernelGGL(synth_fmt_ui_s4_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    default: return launchSeedSingle(pass, params, gx, gy, s);
    }
    return hipGetLastError();
}

hipError_t launchFmtBatch(int pass, const void* params, const uint32_t* work, int n, hipStream_t s) {
    const dim3 g((unsigned)n);
    switch (pass) {
    case kPassLumaFmt: hipLaunchKernelGGL(luma_fmt_batch_kernel, g, dim3(256), 0, s, (const LumaParams*)params, work); break;
    case kPassPyramidFmt: hipLaunchKernelGGL(pyramid_fmt_batch_kernel, g, dim3(256), 0, s, (const PyrParams*)params, work); break;
    case kPassSynthFmtS1: hipLaunchKernelGGL(synth_fmt_s1_batch_kernel, g, dim3(256), 0, s, (const SynthParams*)params, work); break;
    case kPassSynthFmtS2: hipLaunchKernelGGL(synth_fmt_s2_batch_kernel, g, dim3(256), 0, s, (const SynthParams*)params, work); break;
    case kPassSynthFmtS4: hipLaunchKernelGGL(synth_fmt_s4_batch_kernel, g, dim3(256), 0, s, (const SynthParams*)params, work); break;
    case kPassSynthFmtS2T16: hipLaunchKernelGGL(synth_fmt_s2_t16_batch_kernel, g, dim3(256), 0, s, (const SynthParams*)params, work); break;
    case kPassSynthFmtS2T64: hipLaunchKernelGGL(synth_fmt_s2_t64_batch_kernel, g, dim3(256), 0, s, (const SynthParams*)params, work); break;
    default: return launchSeedBatch(pass, params, work, n, s);
    }
    return hipGetLastError();
}

hipError_t launchDebugEFmt(const SynthParams& p, uint32_t* dst, int eStride, hipStream_t s) {
    hipLaunchKernelGGL(debug_e_fmt_kernel, dim3((unsigned)cdiv(eStride, 16), (unsigned)cdiv(p.H, 16)), dim3(256), 0, s,
                       p, dst, eStride);
    return hipGetLastError();
}

}  // namespace fgr
