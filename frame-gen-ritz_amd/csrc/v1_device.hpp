// Device helpers shared by the V1 kernels: clamped fetches and the cv2-exact samplers of the reference
// (gpu/shaders/motion_v1.glsl:42-84). Every float expression is written in the order the reference evaluates it; the
// library is compiled with -ffp-contract=off so that no product is fused into an add.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "v1_params.hpp"

namespace fgr {

// v clamped to [lo, hi] (lo <= hi at every use): one v_med3_i32
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// |a - b| + c of unsigned values as one v_sad_u32. Written out: HIP's __sad forms the signed difference, its negation,
// their maximum and the sum (four instructions; the same value for operands below 2^31), and umax - umin + c inside a
// running sum is reassociated by the compiler into separate maxima, minima and adds (about three an element).
__device__ __forceinline__ uint32_t sadU32(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_sad_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Row-major element index y w + x of in-range coordinates as a 32-bit unsigned value (y, w < 2^24: a full-rate 24-bit
// multiply), and the element at it: the byte offset stays 32-bit, so a plane whose base is uniform is read with the
// base in scalar registers plus one vector offset.
__device__ __forceinline__ uint32_t idx2(int x, int y, int w) { return __umul24((uint32_t)y, (uint32_t)w) + (uint32_t)x; }
template <class T>
__device__ __forceinline__ T ldIdx(const T* base, uint32_t i) {
    return *(const T*)((const char*)base + i * (uint32_t)sizeof(T));
}

// L() motion_v1.glsl:42-44: clamp to edge == cv2 BORDER_REPLICATE == interp.py pad_to_block
__device__ __forceinline__ float lumaAt(const float* img, int x, int y, int w, int h) {
    return ldIdx(img, idx2(clampi(x, 0, w - 1), clampi(y, 0, h - 1), w));
}

__device__ __forceinline__ float2 fieldAt(const float2* f, int x, int y, int w, int h) {
    return ldIdx(f, idx2(clampi(x, 0, w - 1), clampi(y, 0, h - 1), w));
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
    const uint32_t x0 = (uint32_t)clampi(ix, 0, w - 1), x1 = (uint32_t)clampi(ix + 1, 0, w - 1);
    const uint32_t y0 = idx2(0, clampi(iy, 0, h - 1), w), y1 = idx2(0, clampi(iy + 1, 0, h - 1), w);
    const float p00 = ldIdx(img, y0 + x0), p01 = ldIdx(img, y0 + x1), p10 = ldIdx(img, y1 + x0), p11 = ldIdx(img, y1 + x1);
    return p00 * ((1.0f - fy) * (1.0f - fx)) + p01 * ((1.0f - fy) * fx) + p10 * (fy * (1.0f - fx)) + p11 * (fy * fx);
}

// sampleF() motion_v1.glsl:62-70: the same on the two-channel field, at the map already scaled by 32 (X32 = 32 px)
__device__ __forceinline__ float2 sampleF32(const float2* f, float X32, float Y32, int w, int h) {
    const int X = (int)rintf(X32), Y = (int)rintf(Y32);
    const int ix = X >> 5, iy = Y >> 5;
    const float fx = (float)(X & 31) * (1.0f / 32.0f), fy = (float)(Y & 31) * (1.0f / 32.0f);
    const uint32_t x0 = (uint32_t)clampi(ix, 0, w - 1), x1 = (uint32_t)clampi(ix + 1, 0, w - 1);
    const uint32_t y0 = idx2(0, clampi(iy, 0, h - 1), w), y1 = idx2(0, clampi(iy + 1, 0, h - 1), w);
    const float2 p00 = ldIdx(f, y0 + x0), p01 = ldIdx(f, y0 + x1), p10 = ldIdx(f, y1 + x0), p11 = ldIdx(f, y1 + x1);
    const float w00 = (1.0f - fy) * (1.0f - fx), w01 = (1.0f - fy) * fx, w10 = fy * (1.0f - fx), w11 = fy * fx;
    return make_float2(p00.x * w00 + p01.x * w01 + p10.x * w10 + p11.x * w11,
                       p00.y * w00 + p01.y * w01 + p10.y * w10 + p11.y * w11);
}
__device__ __forceinline__ float2 sampleF(const float2* f, float px, float py, int w, int h) {
    return sampleF32(f, px * 32.0f, py * 32.0f, w, h);
}

// halfVec() motion_v1.glsl:73-84: cv2.resize(field, x S, INTER_LINEAR) + edge pad, times S, at full-res pixel (x, y)
__device__ __forceinline__ float2 halfVec(const float2* f, int x, int y, int S, int fw, int fh) {
    x = min(x, S * fw - 1);
    y = min(y, S * fh - 1);
    const float inv = 1.0f / (float)S;
    const float sx = ((float)x + 0.5f) * inv - 0.5f, sy = ((float)y + 0.5f) * inv - 0.5f;
    const float flx = floorf(sx), fly = floorf(sy);
    const float fx = sx - flx, fy = sy - fly;
    const uint32_t i0x = (uint32_t)clampi((int)flx, 0, fw - 1), i1x = (uint32_t)clampi((int)flx + 1, 0, fw - 1);
    const uint32_t i0y = idx2(0, clampi((int)fly, 0, fh - 1), fw), i1y = idx2(0, clampi((int)fly + 1, 0, fh - 1), fw);
    const float2 top = mix2(ldIdx(f, i0y + i0x), ldIdx(f, i0y + i1x), fx);
    const float2 bot = mix2(ldIdx(f, i1y + i0x), ldIdx(f, i1y + i1x), fx);
    const float2 r = mix2(top, bot, fy);
    return make_float2(r.x * (float)S, r.y * (float)S);
}

// v of lane i + s at lane i (what a step of treeSum64 reads; right only for the lanes i < s that the step feeds):
// DPP row_shl:s inside 16-lane rows for s <= 8, a swizzle (lane ^ 16 within groups of 32 lanes) for s = 16.
template <int s>
__device__ __forceinline__ float laneAbove(float v) {
    static_assert(s == 1 || s == 2 || s == 4 || s == 8 || s == 16, "row shifts and the 16-lane swizzle");
    if constexpr (s == 16) return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));
    else return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x100 | s, 0xF, 0xF, true));
}

// Wave-wide sum in the reference's shared-memory tree order (search.comp:45-49): lane i += lane i + s, s = 32 .. 1.
// Only lanes i < s of a step feed the total, so a step needs lane i + s only there. The whole wavefront must be
// active. The total ends in lane 0 and is read from there by every lane.
__device__ __forceinline__ float treeSum64(float d) {
    d = d + __shfl_down(d, 32, 64);
    d = d + laneAbove<16>(d);
    d = d + laneAbove<8>(d);
    d = d + laneAbove<4>(d);
    d = d + laneAbove<2>(d);
    d = d + laneAbove<1>(d);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(d), 0));
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

// FGR_PASS_BATCH asking the register allocator for `waves` wavefronts per SIMD
#define FGR_PASS_BATCH_OCC(name, Params, body, threads, waves)                                                    \
    __global__ __launch_bounds__(threads) __attribute__((amdgpu_waves_per_eu(waves))) void name##_batch_kernel(   \
        const Params* __restrict__ P, const uint32_t* __restrict__ work) {                                        \
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

// FGR_PASS asking the register allocator for `waves` wavefronts per SIMD in both kernels
#define FGR_PASS_OCC(name, Params, body, threads, waves)                                                           \
    __global__ __launch_bounds__(threads) __attribute__((amdgpu_waves_per_eu(waves))) void name##_kernel(Params P) { \
        const int l = xcdChunkTile(linearBlock(), (int)(gridDim.x * gridDim.y));                                 \
        body(P, l % (int)gridDim.x, l / (int)gridDim.x);                                                          \
    }                                                                                                             \
    __global__ __launch_bounds__(threads) __attribute__((amdgpu_waves_per_eu(waves))) void name##_batch_kernel(   \
        const Params* __restrict__ P, const uint32_t* __restrict__ work) {                                        \
        const uint32_t w = ((const FGR_CONST_SPACE uint32_t*)work)[blockIdx.x];                                   \
        const Params p = ((const FGR_CONST_SPACE Params*)P)[w >> kWorkShift];                                     \
        const int l = (int)(w & kWorkMask);                                                                       \
        body(p, l % p.gx, l / p.gx);                                                                              \
    }

}  // namespace fgr
