// Pieces shared by the synth kernels of every format (v1_synth.hip for RGBA8, v1_formats.hip for the others): the
// tile geometry, the scene-cut test, the trust ramp, the field lookup and the field tile of a workgroup.
#pragma once
#include <hip/hip_runtime.h>
#include "v1_device.hpp"

namespace fgr {

// Tile of the fused synth kernel: 64 x 32 output pixels per 256-thread workgroup (8 per thread), a 3-pixel apron for
// the 7x7 box (70 x 38 mismatch values, 1.30 per output pixel).
constexpr int TW = 64, TH = 32, AP = 3, EW = TW + 2 * AP, EH = TH + 2 * AP;
// the same geometry for a tile of TH_ rows (the RGBA8 kernel comes in 16, 32 and 64 rows; the tile height is tuned)
template <int TH_>
struct TileRows {
    static constexpr int H = TH_, EH = TH_ + 2 * AP;
};

struct alignas(4) U32x2 {
    uint32_t x, y;
};

__device__ __forceinline__ int med3i(int v, int lo, int hi) { return min(max(v, lo), hi); }
__device__ __forceinline__ int byteOf(uint32_t v, int c) { return (int)((v >> (8 * c)) & 255u); }

__device__ __forceinline__ uint32_t ld32(const FrameRef& f, int x, int y) {
    return *(const uint32_t*)(f.p + (int64_t)y * f.pitch + 4 * (int64_t)x);
}

struct alignas(4) U32x4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ bool sceneCut(const SynthParams& P) {
    // fallback.comp:10-11: float(sum) * (1/256) / float(blocks) > sceneCutSad (the wavefront adds the partial sums)
    const float meanSad = (float)slotSum(P.sadsum) * (1.0f / 256.0f) / (float)P.nblocks;
    return meanSad > P.cutSad;
}

// interp.py:236-237 trust from the exact 7x7 sum S: e7 = float32(S * (1/49)), u = clip((e7 - lo) / (hi - lo), 0, 1).
// The float32 quotient is taken as float(double(e7 - lo) * (1 / double(hi - lo))): with a 24-bit numerator and
// denominator the quotient is never a rounding midpoint unless the denominator is a power of two (then both are
// exact), and the double product is close enough to round the same way. u is monotone in S, so the host tabulates
// where it is 0 and 1 and the kernel evaluates the ramp only in between.
__device__ __forceinline__ float trustOf(const SynthParams& P, int S) {
    if (S <= P.sZero) return 0.0f;
    if (S >= P.sOne) return 1.0f;
    const float e7 = (float)((double)S * (1.0 / 49.0));
    const float u = (float)((double)(e7 - P.trust0) * P.rden);
    return fminf(fmaxf(u, 0.0f), 1.0f);
}

// lookup.comp:10-19 (interp.py:194-205): the field at time t at flow pixel (qx, qy): g <- pf(q - s g) twice
__device__ __forceinline__ float2 lookupAt(const float2* pf, int fw, int fh, float s, int qx, int qy) {
    float2 g = ldIdx(pf, idx2(qx, qy, fw));
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const float px = (float)qx - s * g.x, py = (float)qy - s * g.y;
        g = sampleF(pf, px, py, fw, fh);
    }
    return g;
}

// Flow column of full-res pixel x: floor((x + 0.5) / S - 0.5) = floor((2x + 1 - S) / 2S), and its weight
template <int S>
struct FlowMap {
    static constexpr int kShift = S == 1 ? 1 : (S == 2 ? 2 : 3);
    static constexpr int indexC(int x) { return (2 * x + 1 - S) >> kShift; }
    static __device__ __forceinline__ int index(int x) { return (2 * x + 1 - S) >> kShift; }
    static __device__ __forceinline__ float frac(int x) {
        return (float)((2 * x + 1 - S) & (2 * S - 1)) * (1.0f / (float)(2 * S));
    }
};

template <int S, int TH_ = TH>
struct SynthTile {
    // field tile (flow res) covering the tile and its apron, and its rows interpolated to full-res columns: tiles start
    // at multiples of TW columns and TH_ rows, so 2 x0 and 2 y0 are multiples of 2S and the flow cells that the region
    // x0 - AP .. x0 + TW + AP - 1 interpolates are index(x0 - AP) .. index(x0 + TW + AP - 1) + 1 (fewer at a clamped
    // frame edge); the same for rows
    static constexpr int FTW = FlowMap<S>::indexC(TW + AP - 1) - FlowMap<S>::indexC(-AP) + 2;
    static constexpr int FTH = FlowMap<S>::indexC(TH_ + AP - 1) - FlowMap<S>::indexC(-AP) + 2;
};

// The field tile of an output tile at (x0, y0) and its apron (clamped to the frame, then to S * fw - 1), looked up at t
// when P.lookup is set, and its rows interpolated at every full-res column x0 - AP .. x0 + TW + AP - 1 of the region
// (hrow[j * EW + ex]; the first half of halfVec, motion_v1.glsl:73-84). Returns the flow row of ft's first row.
// Ends with a barrier.
template <int S, int TH_ = TH>
__device__ __forceinline__ int loadFieldTile(const SynthParams& P, int x0, int y0, float2* ft, float2* hrow) {
    constexpr int FTW = SynthTile<S, TH_>::FTW, FTH = SynthTile<S, TH_>::FTH;
    const int tid = threadIdx.x;
    const int xlim = S * P.fw - 1, ylim = S * P.fh - 1;
    const int xmin = min(max(x0 - AP, 0), xlim), ymin = min(max(y0 - AP, 0), ylim);
    const int tx0 = FlowMap<S>::index(xmin), ty0 = FlowMap<S>::index(ymin);
    // the thread's (at most) IT entries side by side, so that their dependent loads overlap: the index is clamped
    // rather than branched on (an entry past the end repeats the last one and is not stored)
    constexpr int N = FTW * FTH, IT = (N + 255) / 256;
    // the last step's entries past N are not stored; a wavefront with none left there skips that step's lookup
    const bool lastIt = 256 * (IT - 1) + (tid & ~63) < N;
    float2 v[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int i = min(tid + 256 * it, N - 1);
        const int qx = med3i(tx0 + i % FTW, 0, P.fw - 1), qy = med3i(ty0 + i / FTW, 0, P.fh - 1);
        v[it] = ldIdx(P.field, idx2(qx, qy, P.fw));
    }
    if (P.lookup) {   // lookup.comp:10-19, g <- pf(q - s g) twice, the IT entries interleaved step by step
        // the map 32 (q - s g) as 32 q - (32 s) g: scaling by 32 commutes with the rounding of s g and of the difference
        // (field values are never subnormal), so it is the same float with one multiply fewer a coordinate
        const float s32 = 32.0f * P.s;
#pragma unroll
        for (int k = 0; k < 2; k++) {
#pragma unroll
            for (int it = 0; it < IT; it++) {
                if (it == IT - 1 && !lastIt) continue;   // (wave-uniform)
                const int i = min(tid + 256 * it, N - 1);
                const int qx = med3i(tx0 + i % FTW, 0, P.fw - 1), qy = med3i(ty0 + i / FTW, 0, P.fh - 1);
                v[it] = sampleF32(P.field, (float)(32 * qx) - s32 * v[it].x, (float)(32 * qy) - s32 * v[it].y, P.fw,
                                  P.fh);
            }
        }
    }
#pragma unroll
    for (int it = 0; it < IT; it++)
        if (tid + 256 * it < N) ft[tid + 256 * it] = v[it];
    __syncthreads();
    // column by column: thread t takes region column t % EW of the rows t / EW, t / EW + G, ..., so the column's
    // clamped position, weight and flow cell are formed once
    constexpr int G = 256 / EW;
    if (tid < G * EW) {
        const int ex = tid % EW;
        const int x = min(med3i(x0 - AP + ex, 0, P.W - 1), xlim);
        const float fx = FlowMap<S>::frac(x);
        const float2* f = ft + (FlowMap<S>::index(x) - tx0);
        for (int j = tid / EW; j < FTH; j += G) hrow[j * EW + ex] = mix2(f[j * FTW], f[j * FTW + 1], fx);
    }
    __syncthreads();
    return ty0;
}

// The interpolated field r (hv = S r) at region column ex of frame row y (the second half of halfVec)
template <int S>
__device__ __forceinline__ float2 fieldAtRegion(const SynthParams& P, const float2* hrow, int ty0, int ex, int y) {
    const int yc = min(y, S * P.fh - 1);
    const float2* h = hrow + (FlowMap<S>::index(yc) - ty0) * EW + ex;
    return mix2(h[0], h[EW], FlowMap<S>::frac(yc));
}


// The 7-tap horizontal sums of a tile's mismatch rows: hs[r][c] = e[r][c] + .. + e[r][c + 6] (exact integers) for the
// EH_ rows of es (EW ints a row) and the TW output columns. A task is 8 output columns of one row: its 14 values read
// once (as 7 aligned pairs: EW and the task's first column are even), then a running sum, stored two per word.
template <int EH_>
__device__ __forceinline__ void boxRows7(const int* es, uint16_t* hs, int tid) {
    static_assert(EW % 2 == 0 && TW % 8 == 0, "aligned pairs");
    for (int t = tid; t < EH_ * (TW / 8); t += 256) {
        const int r = t / (TW / 8), c0 = (t % (TW / 8)) * 8;
        const int2* e2 = (const int2*)&es[r * EW + c0];
        int v[14];
#pragma unroll
        for (int i = 0; i < 7; i++) {
            const int2 q = e2[i];
            v[2 * i] = q.x;
            v[2 * i + 1] = q.y;
        }
        int h[8];
        h[0] = v[0] + v[1] + v[2] + v[3] + v[4] + v[5] + v[6];
#pragma unroll
        for (int i = 1; i < 8; i++) h[i] = h[i - 1] + v[i + 6] - v[i - 1];
        uint32_t* o = (uint32_t*)&hs[r * TW + c0];
#pragma unroll
        for (int j = 0; j < 4; j++) o[j] = (uint32_t)h[2 * j] | ((uint32_t)h[2 * j + 1] << 16);
    }
}

// boxRows7 for a mismatch tile of bytes (every e is 0 .. 255) in rows EWB bytes apart: a task's 14 values are one
// 16-byte read (two aligned 8-byte halves)
constexpr int EWB = 72;
template <int EH_>
__device__ __forceinline__ void boxRows7B(const uint8_t* es, uint16_t* hs, int tid) {
    static_assert(EWB % 8 == 0 && EWB >= EW + 2 && TW % 8 == 0, "aligned 16-byte reads inside the row");
    for (int t = tid; t < EH_ * (TW / 8); t += 256) {
        const int r = t / (TW / 8), c0 = (t % (TW / 8)) * 8;
        const uint2* e8 = (const uint2*)&es[r * EWB + c0];
        const uint2 q0 = e8[0], q1 = e8[1];
        const uint32_t w[4] = {q0.x, q0.y, q1.x, q1.y};
        int v[14];
#pragma unroll
        for (int i = 0; i < 14; i++) v[i] = (int)((w[i >> 2] >> (8 * (i & 3))) & 255u);
        int h[8];
        h[0] = v[0] + v[1] + v[2] + v[3] + v[4] + v[5] + v[6];
#pragma unroll
        for (int i = 1; i < 8; i++) h[i] = h[i - 1] + v[i + 6] - v[i - 1];
        uint32_t* o = (uint32_t*)&hs[r * TW + c0];
#pragma unroll
        for (int j = 0; j < 4; j++) o[j] = (uint32_t)h[2 * j] | ((uint32_t)h[2 * j + 1] << 16);
    }
}

// CrosshairV2 tap rejection (ui_reject.glsl). The footprint of a tap at the 1/32 map (X, Y): RGB the texels X >> 5 .. + 1
// and Y >> 5 .. + 1, those of weight 0 dropped; 4:2:0 the pixels of the 2x2 chroma texels around it (the chroma point
// is (X - 16) / 64 texels), up to 4x4 pixels, which hold the luma taps too. A tap hits when its footprint touches a UI
// pixel of its frame's kept mask.
__device__ __forceinline__ bool uiHitRect(const UiMaskRef& m, int lx, int ly, int hx, int hy) {
    const int x0 = max(lx - m.x, 0), y0 = max(ly - m.y, 0);
    const int x1 = min(hx - m.x, m.w - 1), y1 = min(hy - m.y, m.h - 1);
    bool hit = false;
    for (int y = y0; y <= y1; y++)
        for (int x = x0; x <= x1; x++) hit = hit || m.m[y * m.pitch + x] != 0;
    return hit;
}

template <bool YUV = false>
__device__ __forceinline__ bool uiHit(const UiMaskRef& m, int X, int Y) {
    if (YUV) {
        const int cx = X - 16, cy = Y - 16;
        const int ix = cx >> 6, iy = cy >> 6;
        return uiHitRect(m, 2 * ix, 2 * iy, 2 * (ix + ((cx & 63) != 0)) + 1, 2 * (iy + ((cy & 63) != 0)) + 1);
    }
    const int lx = X >> 5, ly = Y >> 5;
    return uiHitRect(m, lx, ly, lx + ((X & 31) != 0), ly + ((Y & 31) != 0));
}

// which of a tap pair to replace: 0 none (neither or both on their frame's UI), 1 a's (a takes c's), 2 c's
template <bool YUV = false>
__device__ __forceinline__ int uiRejectWhich(const SynthParams& P, int Xa, int Ya, int Xc, int Yc) {
    const bool ha = uiHit<YUV>(P.uiA, Xa, Ya), hc = uiHit<YUV>(P.uiC, Xc, Yc);
    return ha == hc ? 0 : (ha ? 1 : 2);
}

// a tap on its own frame's UI (the other not) takes the other frame's tap; both or neither: unchanged
template <class T>
__device__ __forceinline__ void uiReject(const SynthParams& P, int Xa, int Ya, int Xc, int Yc, T& wa, T& wc) {
    const int w = uiRejectWhich(P, Xa, Ya, Xc, Yc);
    if (w == 1) wa = wc;
    else if (w == 2) wc = wa;
}

}  // namespace fgr
