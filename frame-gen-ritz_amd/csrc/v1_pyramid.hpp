// The luma pyramid in one launch (luma.comp + down.comp of every level, framegen.cpp:2932-2958): a workgroup of 256
// threads owns 32 x 32 level-0 cells, a thread 2 x 2 of them. Level k + 1 cell (x, y) is the mean of the valid level-k
// pixels (2x + i, 2y + j), so a tile's coarser levels come from the tile alone: level 1 (one cell a thread) from the
// thread's own level-0 cells, levels 2-5 through small LDS grids. Every value is the float the separate passes give:
// the same sums in the same order (j outer, i inner, from 0), times 0.25 or float(1 / n), validity from each level's
// size. Levels past kPyrLevels (frames above 16K pixels) are left to the down pass.
#pragma once
#include <hip/hip_runtime.h>
#include "v1_device.hpp"

namespace fgr {

// mean of the valid ones of four level-k values (v[j][i] at level-k pixel (x0 + i, y0 + j), valid inside w x h)
__device__ __forceinline__ float down4(float v00, float v01, float v10, float v11, int x0, int y0, int w, int h) {
    float s = 0.0f;
    int n = 0;
    const bool c0 = x0 < w, c1 = x0 + 1 < w, r0 = y0 < h, r1 = y0 + 1 < h;
    if (r0 && c0) { s = s + v00; n++; }
    if (r0 && c1) { s = s + v01; n++; }
    if (r1 && c0) { s = s + v10; n++; }
    if (r1 && c1) { s = s + v11; n++; }
    return s * (n == 4 ? 0.25f : 1.0f / (float)n);
}

__device__ __forceinline__ void pyrStore(float* plane, int w, int h, int x, int y, float v) {
    if (x < w && y < h) plane[y * w + x] = v;
}

// Levels 0 (stored by the caller) .. P.levels - 1 of the tile (bx, by) from the thread's 2 x 2 level-0 cells of both
// frames (a[j][i], c[j][i] at cells (cx0 + i, cy0 + j)); every thread of the workgroup calls it (it holds barriers)
__device__ __forceinline__ void pyramidLevels(const PyrParams& P, int bx, int by, const float (&a)[2][2],
                                              const float (&c)[2][2]) {
    __shared__ float ga[16 * 16 + 8 * 8 + 4 * 4 + 2 * 2 + 1], gc[16 * 16 + 8 * 8 + 4 * 4 + 2 * 2 + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    // level 1: the thread's cell (bx * 16 + tx, by * 16 + ty)
    const int x1 = bx * 16 + tx, y1 = by * 16 + ty;
    const float a1 = down4(a[0][0], a[0][1], a[1][0], a[1][1], 2 * x1, 2 * y1, P.w[0], P.h[0]);
    const float c1 = down4(c[0][0], c[0][1], c[1][0], c[1][1], 2 * x1, 2 * y1, P.w[0], P.h[0]);
    if (P.levels > 1) {
        pyrStore(P.la[1], P.w[1], P.h[1], x1, y1, a1);
        pyrStore(P.lc[1], P.w[1], P.h[1], x1, y1, c1);
    }
    if (P.levels <= 2) return;
    ga[ty * 16 + tx] = a1;
    gc[ty * 16 + tx] = c1;
    // levels 2 .. : grid g = 8, 4, 2, 1 cells a tile, read from the previous level's 2g x 2g grid in LDS
    int src = 0, off = 256;
#pragma unroll
    for (int k = 2, g = 8; k < kPyrLevels; k++, g >>= 1) {
        __syncthreads();
        if (k >= P.levels) return;   // (uniform)
        const int t = threadIdx.x;
        if (t < g * g) {
            const int lx = t % g, ly = t / g, sg = 2 * g;
            const int xk = bx * g + lx, yk = by * g + ly;
            const int s0 = src + (2 * ly) * sg + 2 * lx;
            const float ak = down4(ga[s0], ga[s0 + 1], ga[s0 + sg], ga[s0 + sg + 1], 2 * xk, 2 * yk, P.w[k - 1], P.h[k - 1]);
            const float ck = down4(gc[s0], gc[s0 + 1], gc[s0 + sg], gc[s0 + sg + 1], 2 * xk, 2 * yk, P.w[k - 1], P.h[k - 1]);
            pyrStore(P.la[k], P.w[k], P.h[k], xk, yk, ak);
            pyrStore(P.lc[k], P.w[k], P.h[k], xk, yk, ck);
            ga[off + t] = ak;
            gc[off + t] = ck;
        }
        src = off;
        off += g * g;
    }
}

}  // namespace fgr
