// Device helpers shared by the V1 kernels: clamped fetches and the cv2-exact samplers of the reference
// (gpu/shaders/motion_v1.glsl:42-84). Every float expression is written in the order the reference evaluates it; the
// library is compiled with -ffp-contract=off so that no product is fused into an add.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "v1_params.hpp"

namespace fgr {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// L() motion_v1.glsl:42-44: clamp to edge == cv2 BORDER_REPLICATE == interp.py pad_to_block
__device__ __forceinline__ float lumaAt(const float* img, int x, int y, int w, int h) {
    return img[clampi(y, 0, h - 1) * w + clampi(x, 0, w - 1)];
}

__device__ __forceinline__ float2 fieldAt(const float2* f, int x, int y, int w, int h) {
    return f[clampi(y, 0, h - 1) * w + clampi(x, 0, w - 1)];
}

// mix(x, y, a) as pinned by the oracle: x * (1 - a) + y * a
__device__ __forceinline__ float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ float2 mix2(float2 x, float2 y, float a) {
    return make_float2(mixf(x.x, y.x, a), mixf(x.y, y.y, a));
}

// sampleL() motion_v1.glsl:48-56: cv2.remap INTER_LINEAR / BORDER_REPLICATE on a float image; the map is rounded to
// 1/32 px (round half even), weights (1-fy)(1-fx), (1-fy)fx, fy(1-fx), fy fx, summed in that order.
__device__ __forceinline__ float sampleL(const float* img, float px, float py, int w, int h) {
    const int X = (int)rintf(px * 32.0f), Y = (int)rintf(py * 32.0f);
    const int ix = X >> 5, iy = Y >> 5;
    const float fx = (float)(X & 31) * (1.0f / 32.0f), fy = (float)(Y & 31) * (1.0f / 32.0f);
    const int x0 = clampi(ix, 0, w - 1), x1 = clampi(ix + 1, 0, w - 1);
    const int y0 = clampi(iy, 0, h - 1) * w, y1 = clampi(iy + 1, 0, h - 1) * w;
    const float p00 = img[y0 + x0], p01 = img[y0 + x1], p10 = img[y1 + x0], p11 = img[y1 + x1];
    return p00 * ((1.0f - fy) * (1.0f - fx)) + p01 * ((1.0f - fy) * fx) + p10 * (fy * (1.0f - fx)) + p11 * (fy * fx);
}

// sampleF() motion_v1.glsl:62-70: the same on the two-channel field
__device__ __forceinline__ float2 sampleF(const float2* f, float px, float py, int w, int h) {
    const int X = (int)rintf(px * 32.0f), Y = (int)rintf(py * 32.0f);
    const int ix = X >> 5, iy = Y >> 5;
    const float fx = (float)(X & 31) * (1.0f / 32.0f), fy = (float)(Y & 31) * (1.0f / 32.0f);
    const int x0 = clampi(ix, 0, w - 1), x1 = clampi(ix + 1, 0, w - 1);
    const int y0 = clampi(iy, 0, h - 1) * w, y1 = clampi(iy + 1, 0, h - 1) * w;
    const float2 p00 = f[y0 + x0], p01 = f[y0 + x1], p10 = f[y1 + x0], p11 = f[y1 + x1];
    const float w00 = (1.0f - fy) * (1.0f - fx), w01 = (1.0f - fy) * fx, w10 = fy * (1.0f - fx), w11 = fy * fx;
    return make_float2(p00.x * w00 + p01.x * w01 + p10.x * w10 + p11.x * w11,
                       p00.y * w00 + p01.y * w01 + p10.y * w10 + p11.y * w11);
}

// halfVec() motion_v1.glsl:73-84: cv2.resize(field, x S, INTER_LINEAR) + edge pad, times S, at full-res pixel (x, y)
__device__ __forceinline__ float2 halfVec(const float2* f, int x, int y, int S, int fw, int fh) {
    x = min(x, S * fw - 1);
    y = min(y, S * fh - 1);
    const float inv = 1.0f / (float)S;
    const float sx = ((float)x + 0.5f) * inv - 0.5f, sy = ((float)y + 0.5f) * inv - 0.5f;
    const float flx = floorf(sx), fly = floorf(sy);
    const float fx = sx - flx, fy = sy - fly;
    const int i0x = clampi((int)flx, 0, fw - 1), i1x = clampi((int)flx + 1, 0, fw - 1);
    const int i0y = clampi((int)fly, 0, fh - 1) * fw, i1y = clampi((int)fly + 1, 0, fh - 1) * fw;
    const float2 top = mix2(f[i0y + i0x], f[i0y + i1x], fx);
    const float2 bot = mix2(f[i1y + i0x], f[i1y + i1x], fx);
    const float2 r = mix2(top, bot, fy);
    return make_float2(r.x * (float)S, r.y * (float)S);
}

// Wave-wide sum in the reference's shared-memory tree order (search.comp:45-49): lane i += lane i + s, s = 32 .. 1.
// The total ends in lane 0 and is broadcast.
__device__ __forceinline__ float treeSum64(float d) {
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) d = d + __shfl_down(d, s, 64);
    return __shfl(d, 0, 64);
}

// Sum of the kSlots partial counters (the whole wavefront must be active): lane i reads slot i, integer wave reduce.
__device__ __forceinline__ uint32_t slotSum(const uint32_t* slots) {
    uint32_t v = slots[threadIdx.x & 63];
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s, 64);
    return v;
}

__device__ __forceinline__ int linearBlock() { return (int)(blockIdx.y * gridDim.x + blockIdx.x); }

// XCD-aware tile order. Workgroup i of a launch runs on XCD i mod 8, and each XCD has its own L2. The tiles are
// dealt to the XCDs in chunks of kXcdChunk consecutive tiles (chunk c to XCD c mod 8), so neighbouring tiles -- which
// read overlapping aprons and search windows -- mostly share an L2, while every XCD still gets an even share of each
// pair (whole-pair chunks would leave an XCD idle on a cheap scene-cut pair). The tail beyond the last whole round of
// 8 chunks keeps the plain order. The batched launches get the same order from their host-built work lists.
constexpr int kXcds = 8, kXcdChunk = 64;
__host__ __device__ __forceinline__ int xcdChunkTile(int i, int n, int chunk = kXcdChunk) {
    const int span = kXcds * chunk;
    if (i >= n - n % span) return i;
    const int m = i / kXcds, x = i % kXcds;
    return ((m / chunk) * kXcds + x) * chunk + m % chunk;
}

// One device body, two kernels: name_kernel runs one pass of one pair on a (gx, gy) grid; name_batch_kernel runs the
// same pass for many pairs, workgroup i taking work[i] = (parameter block << kWorkShift) | tile.
// The parameter blocks and work list of a batch are read-only for a whole launch. Read through the constant address
// space, the pointers inside a block are known to address global memory, so the body's accesses compile to global (not
// flat) loads and stores; flat accesses would also count against the LDS wait counter. (The host compilation pass only
// type-checks kernel bodies and has no such address space.)
#if defined(__HIP_DEVICE_COMPILE__)
#define FGR_CONST_SPACE __attribute__((address_space(4)))
#else
#define FGR_CONST_SPACE
#endif

#define FGR_PASS_SINGLE(name, Params, body, threads)                                                              \
    __global__ __launch_bounds__(threads) void name##_kernel(Params P) {                                          \
        const int l = xcdChunkTile(linearBlock(), (int)(gridDim.x * gridDim.y));                                 \
        body(P, l % (int)gridDim.x, l / (int)gridDim.x);                                                          \
    }

#define FGR_PASS_BATCH(name, Params, body, threads)                                                               \
    __global__ __launch_bounds__(threads) void name##_batch_kernel(const Params* __restrict__ P,                  \
                                                                   const uint32_t* __restrict__ work) {          \
        const uint32_t w = ((const FGR_CONST_SPACE uint32_t*)work)[blockIdx.x];                                   \
        const Params p = ((const FGR_CONST_SPACE Params*)P)[w >> kWorkShift];                                     \
        const int l = (int)(w & kWorkMask);                                                                       \
        body(p, l % p.gx, l / p.gx);                                                                              \
    }

// both kernels of a pass (spelled out rather than through the two macros above, so that a body naming a template with
// several arguments passes through one macro expansion only)
#define FGR_PASS(name, Params, body, threads)                                                                     \
    __global__ __launch_bounds__(threads) void name##_kernel(Params P) {                                          \
        const int l = xcdChunkTile(linearBlock(), (int)(gridDim.x * gridDim.y));                                 \
        body(P, l % (int)gridDim.x, l / (int)gridDim.x);                                                          \
    }                                                                                                             \
    __global__ __launch_bounds__(threads) void name##_batch_kernel(const Params* __restrict__ P,                  \
                                                                   const uint32_t* __restrict__ work) {          \
        const uint32_t w = ((const FGR_CONST_SPACE uint32_t*)work)[blockIdx.x];                                   \
        const Params p = ((const FGR_CONST_SPACE Params*)P)[w >> kWorkShift];                                     \
        const int l = (int)(w & kWorkMask);                                                                       \
        body(p, l % p.gx, l / p.gx);                                                                              \
    }

}  // namespace fgr
