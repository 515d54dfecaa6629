// V1 synthesis kernels for gfx950 (passes 9-11 of proto/interp.py; gpu/shaders/{lookup,error,resolve,fallback}.comp).
//
// The reference runs the mismatch, the 7x7 box + resolve and the whole-frame fallback as three dispatches with the
// packed mismatch e stored in between (error.comp, resolve.comp, fallback.comp). Here one kernel does the warp, the
// mismatch, the 7x7 box (an LDS tile with a 3-pixel apron whose e is recomputed rather than stored), the trust ramp,
// the untrusted count (a ballot per wavefront, one global atomic per workgroup) and the blend, so the only HBM traffic
// is the two frames, the field and the output. The scene cut is known from the estimate and turns the kernel into a
// copy of the nearer frame; the flick decision needs every workgroup's count and is applied by a second, normally
// empty, kernel.
#include <hip/hip_runtime.h>
#include "v1_device.hpp"
#include "v1_launch.hpp"

namespace fgr {

constexpr int STW = 32, STH = 16, SAP = 3, SEW = STW + 2 * SAP, SEH = STH + 2 * SAP;

__device__ __forceinline__ uint32_t ld32(const FrameRef& f, int x, int y) {
    return *(const uint32_t*)(f.p + (int64_t)y * f.pitch + 4 * (int64_t)x);
}

__device__ __forceinline__ int byteOf(uint32_t v, int c) { return (int)((v >> (8 * c)) & 255u); }

// cv2.remap INTER_LINEAR / BORDER_REPLICATE of an 8-bit RGB frame (interp.py:214): map rounded to 1/32 px (half even),
// integer weights (32-fx)(32-fy), fx(32-fy), (32-fx)fy, fx fy, result (sum + 512) >> 10 -- the exact value rounded half
// up, as cv2's fixed-point remap. Returns the three warped channels packed as bytes (byte order of the frame).
__device__ __forceinline__ uint32_t warpRgba8(const FrameRef& f, int W, int H, float px, float py) {
    const int X = (int)rintf(px * 32.0f), Y = (int)rintf(py * 32.0f);
    const int ix = X >> 5, iy = Y >> 5, fx = X & 31, fy = Y & 31;
    const int x0 = clampi(ix, 0, W - 1), x1 = clampi(ix + 1, 0, W - 1);
    const int y0 = clampi(iy, 0, H - 1), y1 = clampi(iy + 1, 0, H - 1);
    const uint32_t p00 = ld32(f, x0, y0), p01 = ld32(f, x1, y0), p10 = ld32(f, x0, y1), p11 = ld32(f, x1, y1);
    const int w00 = (32 - fx) * (32 - fy), w01 = fx * (32 - fy), w10 = (32 - fx) * fy, w11 = fx * fy;
    uint32_t r = 0u;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int v = (w00 * byteOf(p00, c) + w01 * byteOf(p01, c) + w10 * byteOf(p10, c) + w11 * byteOf(p11, c) + 512) >> 10;
        r |= (uint32_t)v << (8 * c);
    }
    return r;
}

// error.comp:9-21 / mismatch() motion_v1.glsl:117-129: e = max-channel |wa - wc| (integers 0..255)
__device__ __forceinline__ int emaxRgba8(uint32_t wa, uint32_t wc) {
    const int d0 = abs(byteOf(wa, 0) - byteOf(wc, 0)), d1 = abs(byteOf(wa, 1) - byteOf(wc, 1)),
              d2 = abs(byteOf(wa, 2) - byteOf(wc, 2));
    return max(d0, max(d1, d2));
}

// the warp pair at full-res pixel (x, y): a at p - 2t h, c at p + 2(1-t) h (float32 product, then the difference)
__device__ __forceinline__ void warpPairRgba8(const SynthParams& P, int x, int y, uint32_t& wa, uint32_t& wc) {
    const float2 hv = halfVec(P.field, x, y, P.S, P.fw, P.fh);
    const float ka = 2.0f * P.wC, kc = 2.0f * P.wA;
    wa = warpRgba8(P.a, P.W, P.H, (float)x - ka * hv.x, (float)y - ka * hv.y);
    wc = warpRgba8(P.c, P.W, P.H, (float)x + kc * hv.x, (float)y + kc * hv.y);
}

__device__ __forceinline__ bool sceneCut(const SynthParams& P) {
    // fallback.comp:10-11: float(sum) * (1/256) / float(blocks) > sceneCutSad
    const float meanSad = (float)(*P.sadsum) * (1.0f / 256.0f) / (float)P.nblocks;
    return meanSad > P.cutSad;
}

// interp.py:236-237 trust from the exact 7x7 sum S: e7 = float32(S * (1/49)), u = clip((e7 - lo) / (hi - lo), 0, 1).
// The float32 quotient is taken as float(double(e7 - lo) * (1 / double(hi - lo))): with a 24-bit numerator and
// denominator the quotient is never a rounding midpoint unless the denominator is a power of two (then both are
// exact), and the double product is close enough to round the same way.
__device__ __forceinline__ float trustOf(const SynthParams& P, int S) {
    const float e7 = (float)((double)S * (1.0 / 49.0));
    const float u = (float)((double)(e7 - P.trust0) * P.rden);
    return fminf(fmaxf(u, 0.0f), 1.0f);
}

// resolve.comp:73-82 (interp.py:241-246): m = 0.5 (wa + wc) at t = 0.5, else (1-t) wa + t wc; o = m (1 - u) + near u;
// out = clamp(floor(o + 0.5), 0, 255); alpha from the nearer frame
__device__ __forceinline__ uint32_t blendRgba8(const SynthParams& P, uint32_t wa, uint32_t wc, uint32_t nr, float u) {
    const float omu = 1.0f - u;
    uint32_t r = nr & 0xFF000000u;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int a8 = byteOf(wa, c), c8 = byteOf(wc, c), n8 = byteOf(nr, c);
        float m;
        if (P.mid) m = 0.5f * (float)(a8 + c8);
        else m = P.wA * (float)a8 + P.wC * (float)c8;
        const float o = m * omu + (float)n8 * u;
        const int v = clampi((int)floorf(o + 0.5f), 0, 255);
        r |= (uint32_t)v << (8 * c);
    }
    return r;
}

__global__ __launch_bounds__(256) void synth_rgba8_kernel(SynthParams P) {
    __shared__ int es[SEH * SEW];
    __shared__ int hs[SEH * STW];
    __shared__ uint32_t cnt;
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * STW, y0 = blockIdx.y * STH;
    const FrameRef& nf = P.nearC ? P.c : P.a;
    if (sceneCut(P)) {   // uniform over the grid: the whole output is the nearer real frame (interp.py:225-226)
        for (int i = tid; i < STW * STH; i += 256) {
            const int x = x0 + i % STW, y = y0 + i / STW;
            if (x < P.W && y < P.H) *(uint32_t*)(P.out + (int64_t)y * P.opitch + 4 * (int64_t)x) = ld32(nf, x, y);
        }
        return;
    }
    if (tid == 0) cnt = 0u;
    const int ix = tid % STW, iy = tid / STW;
    uint32_t wa[2], wc[2];
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const int x = min(x0 + ix, P.W - 1), y = min(y0 + iy + 8 * r, P.H - 1);
        warpPairRgba8(P, x, y, wa[r], wc[r]);
        es[(iy + 8 * r + SAP) * SEW + ix + SAP] = emaxRgba8(wa[r], wc[r]);
    }
    for (int i = tid; i < SEH * SEW; i += 256) {
        const int ex = i % SEW, ey = i / SEW;
        if (ex >= SAP && ex < SAP + STW && ey >= SAP && ey < SAP + STH) continue;
        const int x = clampi(x0 - SAP + ex, 0, P.W - 1), y = clampi(y0 - SAP + ey, 0, P.H - 1);
        uint32_t a, c;
        warpPairRgba8(P, x, y, a, c);
        es[i] = emaxRgba8(a, c);
    }
    __syncthreads();
    for (int i = tid; i < SEH * STW; i += 256) {
        const int r = i / STW, c = i % STW;
        const int* e = &es[r * SEW + c];
        hs[i] = e[0] + e[1] + e[2] + e[3] + e[4] + e[5] + e[6];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const int x = x0 + ix, y = y0 + iy + 8 * r;
        const bool inside = x < P.W && y < P.H;
        const int* h = &hs[(iy + 8 * r) * STW + ix];
        const int S = h[0] + h[STW] + h[2 * STW] + h[3 * STW] + h[4 * STW] + h[5 * STW] + h[6 * STW];
        const float u = trustOf(P, S);
        const uint64_t b = __ballot(inside && u > 0.5f);
        if ((tid & 63) == 0 && b) atomicAdd(&cnt, (uint32_t)__popcll(b));
        if (inside) {
            const uint32_t nr = ld32(nf, x, y);
            *(uint32_t*)(P.out + (int64_t)y * P.opitch + 4 * (int64_t)x) = blendRgba8(P, wa[r], wc[r], nr, u);
        }
    }
    __syncthreads();
    if (tid == 0 && cnt) atomicAdd(P.count, cnt);
}

// lookup.comp:10-19 (interp.py:194-205): the per-pixel field at time t, g <- pf(q - s g) twice (LOOKUP_ITERS = 2)
__global__ __launch_bounds__(256) void lookup_kernel(LookupParams P) {
    const int qx = blockIdx.x * 16 + threadIdx.x, qy = blockIdx.y * 16 + threadIdx.y;
    if (qx >= P.fw || qy >= P.fh) return;
    float2 g = P.pf[qy * P.fw + qx];
    for (int k = 0; k < 2; k++) {
        const float px = (float)qx - P.s * g.x, py = (float)qy - P.s * g.y;
        g = sampleF(P.pf, px, py, P.fw, P.fh);
    }
    P.out[qy * P.fw + qx] = g;
}

// fallback.comp:9-24: more than globalFallback x W x H untrusted pixels -> the nearer real frame. Every workgroup
// exits at once in the normal case; 64x64 pixels per workgroup.
__global__ __launch_bounds__(256) void fallback_kernel(FallbackParams P) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *P.nextCount = 0u;
    const bool flick = (float)(*P.count) > P.gfrac * (float)P.W * (float)P.H;
    if (!flick) return;
    const FrameRef& nf = P.nearC ? P.c : P.a;
    const int bx = blockIdx.x * 64, by = blockIdx.y * 64;
    for (int i = threadIdx.x; i < 64 * 64; i += 256) {
        const int x = bx + (i & 63), y = by + (i >> 6);
        if (x >= P.W || y >= P.H) continue;
        const uint8_t* src = nf.p + (int64_t)y * nf.pitch + (int64_t)x * P.bpp;
        uint8_t* dst = P.out + (int64_t)y * P.opitch + (int64_t)x * P.bpp;
        if (P.bpp == 4) *(uint32_t*)dst = *(const uint32_t*)src;
        else if (P.bpp == 8) *(uint2*)dst = *(const uint2*)src;
        else if (P.bpp == 2) *(uint16_t*)dst = *(const uint16_t*)src;
        if (P.outc && ((x | y) & 1) == 0) {   // 4:2:0 chroma of the even-even pixel (fallback.comp:19-22)
            const int cx = x >> 1, cy = y >> 1;
            *(uint32_t*)(P.outc + (int64_t)cy * P.ocpitch + 4 * (int64_t)cx) =
                *(const uint32_t*)(nf.c + (int64_t)cy * nf.cpitch + 4 * (int64_t)cx);
        }
    }
}

// error.comp:9-21 as a debug pass: the packed mismatch e of the latest synth (4 pixels per uint32, low byte first)
__global__ __launch_bounds__(256) void debug_e_rgba8_kernel(SynthParams P, uint32_t* dst, int eStride) {
    const int gx = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
    if (gx >= eStride || y >= P.H) return;
    uint32_t packed = 0u;
    for (int i = 0; i < 4; i++) {
        const int x = gx * 4 + i;
        if (x < P.W) {
            uint32_t a, c;
            warpPairRgba8(P, x, y, a, c);
            packed |= (uint32_t)emaxRgba8(a, c) << (8 * i);
        }
    }
    dst[(int64_t)y * eStride + gx] = packed;
}

hipError_t launchLookup(const LookupParams& p, hipStream_t s) {
    hipLaunchKernelGGL(lookup_kernel, dim3((unsigned)((p.fw + 15) / 16), (unsigned)((p.fh + 15) / 16)), dim3(16, 16), 0, s, p);
    return hipGetLastError();
}

hipError_t launchSynth(const SynthParams& p, int format, hipStream_t s) {
    const dim3 g((unsigned)((p.W + STW - 1) / STW), (unsigned)((p.H + STH - 1) / STH));
    switch (format) {
    case kRgba8: hipLaunchKernelGGL(synth_rgba8_kernel, g, dim3(256), 0, s, p); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launchFallback(const FallbackParams& p, int format, hipStream_t s) {
    (void)format;
    hipLaunchKernelGGL(fallback_kernel, dim3((unsigned)((p.W + 63) / 64), (unsigned)((p.H + 63) / 64)), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launchDebugE(const SynthParams& p, uint32_t* dst, int eStride, int format, hipStream_t s) {
    const dim3 g((unsigned)((eStride + 15) / 16), (unsigned)((p.H + 15) / 16));
    switch (format) {
    case kRgba8: hipLaunchKernelGGL(debug_e_rgba8_kernel, g, dim3(16, 16), 0, s, p, dst, eStride); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launchLuma(const LumaParams& p, int format, hipStream_t s) {
    switch (format) {
    case kRgba8: return launchLumaRgba8(p, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace fgr
