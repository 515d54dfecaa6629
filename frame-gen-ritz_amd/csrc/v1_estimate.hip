// V1 motion estimate kernels for gfx950 (passes 1-8 of proto/interp.py; gpu/shaders/{luma,down,medhist,search,
// median3,subpix,pixfield}.comp). One launch per pass and level, in the reference's order (framegen.cpp:2932-3153).
//
// Mapping to CDNA4: a 64-lane wavefront owns one 8x8 block in the block passes (search, subpix): the block's 64 pixels
// are the 64 lanes, reductions run over the wavefront with cross-lane shuffles in the reference's tree order, and the
// candidate windows are staged in LDS with their search apron. Per-pixel passes use 16x16 tiles with an LDS apron.
#include <hip/hip_runtime.h>
#include <float.h>
#include <limits.h>
#include "v1_device.hpp"
#include "v1_launch.hpp"
#include "v1_pyramid.hpp"

namespace fgr {

// ---------------------------------------------------------------------------------------------------------- luma

// luma.comp:23-26: cv2 BGR2GRAY in 14-bit fixed point on 8-bit R, G, B
__device__ __forceinline__ int luma8(uint32_t px, int bgra) {
    int r = px & 255, g = (px >> 8) & 255, b = (px >> 16) & 255;
    if (bgra) { const int t = r; r = b; b = t; }
    return (r * 4899 + g * 9617 + b * 1868 + 8192) >> 14;
}

// luma.comp:58-72 + interp.py:150: level-0 luma = mean of the valid integer lumas of each S x S cell (integer sums,
// one float32 division). Block (0, 0) also clears the estimate's counters (framegen.cpp:2929 clearCounters(64)).
__device__ __forceinline__ void lumaRgba8Body(const LumaParams& P, int bx, int by) {
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    if (bx == 0 && by == 0)
        for (int i = threadIdx.x; i < kUWords; i += 256) P.u[i] = 0u;
    const int ox = bx * 16 + tx, oy = by * 16 + ty;
    if (ox >= P.fw || oy >= P.fh) return;
    int sa = 0, sc = 0, n = 0;
    for (int j = 0; j < P.S; j++) {
        const int py = P.S * oy + j;
        if (py >= P.H) break;
        const uint32_t* ra = (const uint32_t*)(P.a.p + (int64_t)py * P.a.pitch);
        const uint32_t* rc = (const uint32_t*)(P.c.p + (int64_t)py * P.c.pitch);
        for (int i = 0; i < P.S; i++) {
            const int px = P.S * ox + i;
            if (px >= P.W) break;
            sa += luma8(ra[px], P.bgra);
            sc += luma8(rc[px], P.bgra);
            n++;
        }
    }
    P.la[oy * P.fw + ox] = (float)sa / (float)n;
    P.lc[oy * P.fw + ox] = (float)sc / (float)n;
}

// The fused pyramid's level 0 (RGBA8): the thread's 2 x 2 cells (cx0 .. cx0 + 1, cy0 .. cy0 + 1) as lumaRgba8Body
// forms them; a pixel row of the cells' 2S-pixel segment is read as S/2 16-byte words (S = 1: one 8-byte word) when the
// segment lies inside the frame, else pixel by pixel
template <int S>
__device__ __forceinline__ void pyrCellsRgba8(const PyrParams& P, int cx0, int cy0, float (&a)[2][2], float (&c)[2][2]) {
    constexpr int SEG = 2 * S;
    const int px0 = S * cx0;
    const bool inside = px0 + SEG <= P.W;
#pragma unroll
    for (int r = 0; r < 2; r++) {
        int sa[2] = {0, 0}, sc[2] = {0, 0};
#pragma unroll
        for (int j = 0; j < S; j++) {
            const int py = S * (cy0 + r) + j;
            if (py >= P.H) break;
            const uint8_t* ra = P.a.p + (int64_t)py * P.a.pitch + 4 * (int64_t)px0;
            const uint8_t* rc = P.c.p + (int64_t)py * P.c.pitch + 4 * (int64_t)px0;
            uint32_t qa[SEG], qc[SEG];
            if (inside) {
                if constexpr (S == 1) {
                    const uint2 wa = *(const uint2*)ra, wc = *(const uint2*)rc;
                    qa[0] = wa.x; qa[1] = wa.y;
                    qc[0] = wc.x; qc[1] = wc.y;
                } else {
#pragma unroll
                    for (int q = 0; q < SEG / 4; q++) {
                        const uint4 wa = ((const uint4*)ra)[q], wc = ((const uint4*)rc)[q];
                        qa[4 * q] = wa.x; qa[4 * q + 1] = wa.y; qa[4 * q + 2] = wa.z; qa[4 * q + 3] = wa.w;
                        qc[4 * q] = wc.x; qc[4 * q + 1] = wc.y; qc[4 * q + 2] = wc.z; qc[4 * q + 3] = wc.w;
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < SEG; k++) {
                    const bool ok = px0 + k < P.W;
                    qa[k] = ok ? ((const uint32_t*)ra)[k] : 0u;
                    qc[k] = ok ? ((const uint32_t*)rc)[k] : 0u;
                }
            }
#pragma unroll
            for (int k = 0; k < SEG; k++) {
                if (px0 + k < P.W) {
                    sa[k / S] += luma8(qa[k], P.bgra);
                    sc[k / S] += luma8(qc[k], P.bgra);
                }
            }
        }
        const int ny = min(max(P.H - S * (cy0 + r), 0), S);
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int n = min(max(P.W - S * (cx0 + i), 0), S) * ny;
            a[r][i] = n ? (float)sa[i] / (float)n : 0.0f;
            c[r][i] = n ? (float)sc[i] / (float)n : 0.0f;
        }
    }
}

// the thread's level-0 cells into the planes (inside the level), then the coarser levels
__device__ __forceinline__ void pyramidStore0(const PyrParams& P, int bx, int by, int cx0, int cy0, const float (&a)[2][2],
                                              const float (&c)[2][2]) {
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int i = 0; i < 2; i++) {
            pyrStore(P.la[0], P.w[0], P.h[0], cx0 + i, cy0 + r, a[r][i]);
            pyrStore(P.lc[0], P.w[0], P.h[0], cx0 + i, cy0 + r, c[r][i]);
        }
    pyramidLevels(P, bx, by, a, c);
}

__device__ __forceinline__ void pyramidRgba8Body(const PyrParams& P, int bx, int by) {
    if (bx == 0 && by == 0)
        for (int i = threadIdx.x; i < kUWords; i += 256) P.u[i] = 0u;
    const int cx0 = bx * 32 + 2 * (threadIdx.x & 15), cy0 = by * 32 + 2 * (threadIdx.x >> 4);
    float a[2][2], c[2][2];
    if (P.S == 1) pyrCellsRgba8<1>(P, cx0, cy0, a, c);
    else if (P.S == 2) pyrCellsRgba8<2>(P, cx0, cy0, a, c);
    else pyrCellsRgba8<4>(P, cx0, cy0, a, c);
    pyramidStore0(P, bx, by, cx0, cy0, a, c);
}

// down.comp:9-26: 2x2 mean of the valid pixels, summed j-outer / i-inner, times 0.25 or float(1/n)
__device__ __forceinline__ void downBody(const DownParams& P, int bx, int by) {
    const int ox = bx * 16 + (threadIdx.x & 15), oy = by * 16 + (threadIdx.x >> 4);
    if (ox >= P.w || oy >= P.h) return;
    float sa = 0.0f, sc = 0.0f;
    int n = 0;
    for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) {
            const int px = 2 * ox + i, py = 2 * oy + j;
            if (px < P.pw && py < P.ph) {
                sa = sa + P.sa[py * P.pw + px];
                sc = sc + P.sc[py * P.pw + px];
                n++;
            }
        }
    const float s = n == 4 ? 0.25f : 1.0f / (float)n;
    P.da[oy * P.w + ox] = sa * s;
    P.dc[oy * P.w + ox] = sc * s;
}

// ---------------------------------------------------------------------------------------------------------- median

// medhist.comp:73-106 (interp.py:142): exact per-component median of the fine level's own-parent vectors. Each
// parent is counted with the number of fine blocks that map to it (min(ix/2, pw-1), min(iy/2, ph-1)), so the loop runs
// over the parent grid. Lanes holding the same (bin, weight) add once per wavefront (ballot aggregation), which keeps
// the LDS atomics off the one hot bin of a global pan. Writes 2 * 0.5 * (k1-th + k2-th) as float bits.
// median3.comp:8-28: per-component 3x3 median, replicate border. The shader's partial selection sort returns the 5th
// smallest of the 9 values; here a 19-exchange min/max network selects it (38 instructions instead of 90; the network
// is checked on all 2^9 0/1 inputs by tests/test_kernel_arith.py, which by the 0-1 principle covers every input). Field
// values are finite and never -0, so min and max order them as the shader's < does.
__device__ __forceinline__ float med9(float* v) {
    constexpr int kNet[19][2] = {{1, 2}, {4, 5}, {7, 8}, {0, 1}, {3, 4}, {6, 7}, {1, 2}, {4, 5}, {7, 8}, {0, 3},
                                 {5, 8}, {4, 7}, {3, 6}, {1, 4}, {2, 5}, {4, 7}, {4, 2}, {6, 4}, {4, 2}};
#pragma unroll
    for (int k = 0; k < 19; k++) {
        const float a = v[kNet[k][0]], b = v[kNet[k][1]];
        v[kNet[k][0]] = fminf(a, b);
        v[kNet[k][1]] = fmaxf(a, b);
    }
    return v[4];
}

__device__ __forceinline__ float2 median3At(const float2* in, int bx, int by, int bw, int bh) {
    float xs[9], ys[9];
    int k = 0;
    for (int j = -1; j <= 1; j++)
        for (int i = -1; i <= 1; i++) {
            const float2 v = fieldAt(in, bx + i, by + j, bw, bh);
            xs[k] = v.x;
            ys[k] = v.y;
            k++;
        }
    return make_float2(med9(xs), med9(ys));
}

__device__ __forceinline__ int parentWeight(int p, int np, int nb) {
    // number of fine indices i in [0, nb) with min(i / 2, np - 1) == p
    const int lo = 2 * p;
    const int hi = (p == np - 1) ? nb : min(2 * p + 2, nb);
    return max(hi - lo, 0);
}

__device__ __forceinline__ void histAdd(uint32_t* hist, int bin, int weight) {
    // group the active lanes by key; one LDS atomic per distinct key
    const int key = bin * 1024 + weight;
    uint64_t todo = __ballot(1);
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const int lk = __shfl(key, leader, 64);
        const uint64_t same = __ballot(key == lk) & todo;
        if ((int)(threadIdx.x & 63) == leader) atomicAdd(&hist[lk >> 10], (uint32_t)((lk & 1023) * __popcll(same)));
        todo &= ~same;
    }
}

// The median of the wide reach settings (medhist.comp FG_WIDE): first-pass bins of 2^sh pixels, bin = (v >> sh) + 128,
// hold every vector the settings can reach; each of the four order statistics lies in one bin (tb, with tbase elements
// before it); a second histogram of the low sh bits of the vectors in those bins gives the exact value.
__device__ __forceinline__ void medhistFine(const MedParams& P, const float* kvBins, const uint32_t* tbase, float* kv) {
    __shared__ uint32_t fine[4][256];
    const int t = threadIdx.x, sh = P.shift, mask = (1 << sh) - 1;
    fine[t >> 8][t & 255] = 0u;
    __syncthreads();
    const int np = P.pw * P.ph;
    for (int i0 = 0; i0 < np; i0 += 1024) {
        const int i = i0 + t;
        if (i < np) {
            const int px = i % P.pw, py = i / P.pw;
            const int wgt = parentWeight(px, P.pw, P.bw) * parentWeight(py, P.ph, P.bh);
            if (wgt > 0) {
                const float2 v = P.parent[i];
                const int vx = (int)rintf(v.x), vy = (int)rintf(v.y);
                for (int j = 0; j < 4; j++) {
                    const int c = j < 2 ? vx : vy;
                    const bool in = clampi((c >> sh) + 128, 0, 255) == (int)kvBins[j] + 128;
                    if (__any(in)) histAdd(fine[j], in ? (c & mask) : 0, in ? wgt : 0);
                }
            }
        }
    }
    __syncthreads();
    if (t < 4) {
        const uint32_t n = (uint32_t)(P.bw * P.bh), k = (t & 1) == 0 ? (n - 1u) / 2u : n / 2u;
        uint32_t acc = tbase[t];
        int f = 0;
        for (; f < mask; f++) {
            acc += fine[t][f];
            if (acc > k) break;
        }
        kv[t] = (float)((int)kvBins[t] * (1 << sh) + f);
    }
}

__device__ __forceinline__ void medhistBody(const MedParams& P, int, int) {
    __shared__ uint32_t hx[256], hy[256];
    __shared__ uint32_t wsum[2][4];
    __shared__ float kv[4];
    __shared__ uint32_t tbase[4];
    const int t = threadIdx.x;
    const int sh = P.shift < 0 ? 0 : P.shift;
    if (t < 256) {
        hx[t] = 0u;
        hy[t] = 0u;
    }
    if (t < 4) {
        kv[t] = 127.0f;
        tbase[t] = 0u;
    }
    __syncthreads();
    const int np = P.pw * P.ph;
    for (int i0 = 0; i0 < np; i0 += 1024) {
        const int i = i0 + t;
        if (i < np) {
            const int px = i % P.pw, py = i / P.pw;
            float2 v;
            if (P.raw) {   // (uniform) the parent level's median3, written for the search and binned here
                v = median3At(P.raw, px, py, P.pw, P.ph);
                P.filtered[i] = v;
            }
            const int wgt = parentWeight(px, P.pw, P.bw) * parentWeight(py, P.ph, P.bh);
            if (wgt > 0) {
                if (!P.raw) v = P.parent[i];
                histAdd(hx, clampi(((int)rintf(v.x) >> sh) + 128, 0, 255), wgt);
                histAdd(hy, clampi(((int)rintf(v.y) >> sh) + 128, 0, 255), wgt);
            }
        }
    }
    __syncthreads();   // (also makes the filtered field the second pass reads visible to the workgroup)
    // inclusive prefix over the 256 bins: a scan per wavefront, then the wavefront totals
    uint32_t cx = 0u, cy = 0u;
    if (t < 256) {
        cx = hx[t];
        cy = hy[t];
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t vx = __shfl_up(cx, o, 64), vy = __shfl_up(cy, o, 64);
            if ((t & 63) >= o) {
                cx += vx;
                cy += vy;
            }
        }
        if ((t & 63) == 63) {
            wsum[0][t >> 6] = cx;
            wsum[1][t >> 6] = cy;
        }
    }
    __syncthreads();
    if (t < 256) {
        for (int j = 0; j < (t >> 6); j++) {
            cx += wsum[0][j];
            cy += wsum[1][j];
        }
        // the k-th element (0-based) lies in the first bin whose cumulative count exceeds k (medhist.comp:77-84)
        const uint32_t n = (uint32_t)(P.bw * P.bh), k1 = (n - 1u) / 2u, k2 = n / 2u;
        const uint32_t ex = cx - hx[t], ey = cy - hy[t];
        const float b = (float)(t - 128);
        if (ex <= k1 && k1 < cx) { kv[0] = b; tbase[0] = ex; }
        if (ex <= k2 && k2 < cx) { kv[1] = b; tbase[1] = ex; }
        if (ey <= k1 && k1 < cy) { kv[2] = b; tbase[2] = ey; }
        if (ey <= k2 && k2 < cy) { kv[3] = b; tbase[3] = ey; }
    }
    __syncthreads();
    if (P.shift >= 0) {
        __shared__ float kb[4];
        if (t < 4) kb[t] = kv[t];
        __syncthreads();
        medhistFine(P, kb, tbase, kv);
        __syncthreads();
    }
    if (t < 2) {
        const float m = 0.5f * (kv[2 * t] + kv[2 * t + 1]);
        P.u[P.medIdx + t] = __float_as_uint(2.0f * m);
    }
}

// ---------------------------------------------------------------------------------------------------------- search

// search.comp:23-87 (interp.py:116-144): one wavefront per 8x8 block, four blocks per workgroup.
//  1. seed (not on the coarsest level): 6 candidates (own parent x2, vertical / horizontal / diagonal neighbour
//     parents x2, zero, global median) [+ one learned proposal], lane = pixel, SAD by the 64-lane tree;
//     cost SAD/64 (+ lambda for all but own), strict <. On the coarsest level the seed is zero, or with learned
//     proposals the better of zero (no lambda) and the proposal (+ lambda);
//  2. the (8 + 2R)^2 windows of a (at q - seed) and c (at q + seed) into LDS;
//  3. lane = candidate (dy-major), sequential y-outer / x-inner SAD over LDS, cost SAD/64 + lambda |d|_1;
//  4. wavefront argmin on (cost, candidate index): the first of the minimal costs, as the sequential strict <.
// The seed of block (bx, by) (step 1 above), wave-uniform: lane = pixel, SAD by the 64-lane tree. In three steps, so
// that a wave can form the seeds of two blocks side by side: the candidates' vectors, then all their texel pairs
// loaded together (one memory latency, not one per candidate), then the tree sums and the first-minimum choice in
// candidate order. Slots past nc repeat the zero vector and are not considered. A candidate whose vector repeats an
// earlier one's is skipped (no loads, no sum): its cost is that earlier candidate's SAD plus lambda >= 0, or the same
// float, so under the strict < it can never be chosen -- the choice is unchanged.
constexpr int kMaxSeeds = 7;

struct SeedSet {
    int nc;
    int cx[kMaxSeeds], cy[kMaxSeeds];
    bool dup[kMaxSeeds];
    float d[kMaxSeeds];
};

__device__ __forceinline__ void seedVectors(const SearchParams& P, int blk, bool active, int bx, int by, SeedSet& S) {
    const int px = P.parent ? min(bx / 2, P.pw - 1) : 0, py = P.parent ? min(by / 2, P.ph - 1) : 0;
    const int ssy = (by & 1) == 0 ? -1 : 1, ssx = (bx & 1) == 0 ? -1 : 1;
    S.nc = P.parent ? (P.extraSeed ? 7 : 6) : 2;
#pragma unroll
    for (int kk = 0; kk < kMaxSeeds; kk++) {
        const int k = P.parent ? kk : (kk == 0 ? 4 : 6);   // coarsest level: zero, then the proposal
        int cx = 0, cy = 0;
        if (kk >= S.nc) {
        } else if (k < 4) {
            const int ox = (k == 2 || k == 3) ? ssx : 0, oy = (k == 1 || k == 3) ? ssy : 0;
            const float2 v = P.parent[clampi(py + oy, 0, P.ph - 1) * P.pw + clampi(px + ox, 0, P.pw - 1)];
            cx = 2 * (int)rintf(v.x);
            cy = 2 * (int)rintf(v.y);
        } else if (k == 5) {
            cx = (int)rintf(__uint_as_float(P.med[0]));
            cy = (int)rintf(__uint_as_float(P.med[1]));
        } else if (k == 6) {
            const float2 v = P.extraSeed[active ? blk : 0];
            cx = (int)rintf(v.x);
            cy = (int)rintf(v.y);
        }
        S.cx[kk] = cx;
        S.cy[kk] = cy;
    }
    // repeats, each vector as one 32-bit key (|cx|, |cy| < 2^15), compared without short-circuit
    uint32_t key[kMaxSeeds];
#pragma unroll
    for (int kk = 0; kk < kMaxSeeds; kk++) key[kk] = ((uint32_t)S.cx[kk] & 0xFFFFu) | ((uint32_t)S.cy[kk] << 16);
#pragma unroll
    for (int kk = 0; kk < kMaxSeeds; kk++) {
        bool dup = kk >= S.nc;
#pragma unroll
        for (int j = 0; j < kk; j++) dup = dup | (key[j] == key[kk]);
        S.dup[kk] = dup;
    }
}

__device__ __forceinline__ void seedTexels(const SearchParams& P, int bx, int by, int lane, SeedSet& S) {
    const int qx = bx * kBlock + (lane & 7), qy = by * kBlock + (lane >> 3);
#pragma unroll
    for (int kk = 0; kk < kMaxSeeds; kk++)
        if (!S.dup[kk])   // (wave-uniform)
            S.d[kk] = fabsf(lumaAt(P.la, qx - S.cx[kk], qy - S.cy[kk], P.w, P.h) -
                            lumaAt(P.lc, qx + S.cx[kk], qy + S.cy[kk], P.w, P.h));
}

__device__ __forceinline__ void seedPick(const SearchParams& P, const SeedSet& S, int& sx, int& sy) {
    float best = 0.0f;
#pragma unroll
    for (int kk = 0; kk < kMaxSeeds; kk++) {
        if (kk >= S.nc) break;
        if (S.dup[kk]) continue;   // (wave-uniform; never kk == 0)
        float c = treeSum64(S.d[kk]) * (1.0f / 64.0f);
        if (kk == 0) {
            best = c;
            sx = S.cx[kk];
            sy = S.cy[kk];
        } else {
            c = c + P.lam;
            if (c < best) {
                best = c;
                sx = S.cx[kk];
                sy = S.cy[kk];
            }
        }
    }
}

__device__ __forceinline__ void searchSeed(const SearchParams& P, int blk, bool active, int bx, int by, int lane, int& sx,
                                           int& sy) {
    sx = 0;
    sy = 0;
    if (!P.parent && !P.extraSeed) return;
    SeedSet S;
    seedVectors(P, blk, active, bx, by, S);
    seedTexels(P, bx, by, lane, S);
    seedPick(P, S, sx, sy);
}

// step 2 for a dyadic level (P.fb >= 0): the window values as 16-bit integers k (value k 2^-fb), each stored twice,
// E16[i] = k_i and O16[i - 1] = k_i, so that the texel pair (k_i, k_i+1) is word i / 2 of E16 (i even) or (i - 1) / 2 of
// O16 (i odd); the windows' rows (T = 8 + 2R) are even
__device__ __forceinline__ void searchStoreInt(uint16_t* E, int i, int N, float v, float up) {
    if (i < N) {
        const uint16_t k = (uint16_t)(uint32_t)(v * up);
        E[i] = k;
        if (i > 0) E[N + i - 1] = k;   // O16 follows E16
    }
}

// step 2: the (8 + 2R)^2 windows of a (at q - seed) and c (at q + seed), 64 lanes
__device__ __forceinline__ void searchWindows(const SearchParams& P, float* A, float* C, int b0x, int b0y, int sx, int sy,
                                             int R, int lane, bool ints) {
    const int T = 8 + 2 * R;
    const int q = 64 / T, r = 64 - q * T;   // element i + 64 is (x + r, y + q), carried
    const float up = __int_as_float((127 + max(P.fb, 0)) << 23);
    int x = lane % T, y = lane / T;
    for (int i = lane; i < T * T; i += 64) {
        const float va = lumaAt(P.la, b0x - sx - R + x, b0y - sy - R + y, P.w, P.h);
        const float vc = lumaAt(P.lc, b0x + sx - R + x, b0y + sy - R + y, P.w, P.h);
        if (ints) {
            searchStoreInt((uint16_t*)A, i, T * T, va, up);
            searchStoreInt((uint16_t*)C, i, T * T, vc, up);
        } else {
            A[i] = va;
            C[i] = vc;
        }
        x += r;
        y += q;
        if (x >= T) {
            x -= T;
            y++;
        }
    }
}

// step 2 with every lane's loads issued before any store (T = 8 + 2R <= MAXT; one memory latency for the window)
template <int MAXT>
__device__ __forceinline__ void searchWindowsIssued(const SearchParams& P, float* A, float* C, int b0x, int b0y, int sx,
                                                    int sy, int R, int lane, bool ints) {
    constexpr int IT = (MAXT * MAXT + 63) / 64;
    const int T = 8 + 2 * R, N = T * T;
    const int q = 64 / T, r = 64 - q * T;   // element i + 64 is (x + r, y + q), carried
    float va[IT], vc[IT];
    int x = lane % T, y = lane / T;
#pragma unroll
    for (int it = 0; it < IT; it++) {   // past N: a clamped read that is not stored
        va[it] = lumaAt(P.la, b0x - sx - R + x, b0y - sy - R + y, P.w, P.h);
        vc[it] = lumaAt(P.lc, b0x + sx - R + x, b0y + sy - R + y, P.w, P.h);
        x += r;
        y += q;
        if (x >= T) {
            x -= T;
            y++;
        }
    }
    if (ints) {   // (uniform) the 16-bit pairs of a dyadic level
        const float up = __int_as_float((127 + P.fb) << 23);
#pragma unroll
        for (int it = 0; it < IT; it++) {
            searchStoreInt((uint16_t*)A, lane + 64 * it, N, va[it], up);
            searchStoreInt((uint16_t*)C, lane + 64 * it, N, vc[it], up);
        }
        return;
    }
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int i = lane + 64 * it;
        if (i < N) {
            A[i] = va[it];
            C[i] = vc[it];
        }
    }
}

// step 2 at radius R (T = 8 + 2R) in two halves, so that a wave can issue its loads early: a lane's IT values of the
// window of a (at q - seed) and c (at q + seed) of one block (element i of a window as in searchWindows; past N a
// clamped read that is not stored), then their stores
template <int R>
struct SearchWin {
    static constexpr int T = 8 + 2 * R, N = T * T, IT = (N + 63) / 64;
    float va[IT], vc[IT];
};

template <int R>
__device__ __forceinline__ void searchWinLoad(const SearchParams& P, int bx, int by, int sx, int sy, int lane,
                                              SearchWin<R>& w) {
    constexpr int T = SearchWin<R>::T, IT = SearchWin<R>::IT;
    constexpr int q = 64 / T, r = 64 - q * T;   // element i + 64 is (x + r, y + q), carried
    int x = lane % T, y = lane / T;
    const int b0x = bx * kBlock, b0y = by * kBlock;
#pragma unroll
    for (int it = 0; it < IT; it++) {
        w.va[it] = lumaAt(P.la, b0x - sx - R + x, b0y - sy - R + y, P.w, P.h);
        w.vc[it] = lumaAt(P.lc, b0x + sx - R + x, b0y + sy - R + y, P.w, P.h);
        x += r;
        y += q;
        if (x >= T) {
            x -= T;
            y++;
        }
    }
}

template <int R>
__device__ __forceinline__ void searchWinStore(const SearchParams& P, const SearchWin<R>& w, float* A, float* C, int lane,
                                               bool ints) {
    constexpr int N = SearchWin<R>::N, IT = SearchWin<R>::IT;
    if (ints) {   // (uniform) the 16-bit pairs of a dyadic level
        const float up = __int_as_float((127 + P.fb) << 23);
#pragma unroll
        for (int it = 0; it < IT; it++) {
            searchStoreInt((uint16_t*)A, lane + 64 * it, N, w.va[it], up);
            searchStoreInt((uint16_t*)C, lane + 64 * it, N, w.vc[it], up);
        }
        return;
    }
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int i = lane + 64 * it;
        if (i < N) {
            A[i] = w.va[it];
            C[i] = w.vc[it];
        }
    }
}

// step 3 for candidate ci: the SAD over the 64 pixels (y-outer, x-inner) and the cost
__device__ __forceinline__ void searchCost(const SearchParams& P, const float* A, const float* C, int R, int ci, float& cost,
                                           float& sadc) {
    const int T = 8 + 2 * R, D = 2 * R + 1;
    const int dx = ci % D - R, dy = ci / D - R;
    const float* a = A + (R - dy) * T + (R - dx);
    const float* c = C + (R + dy) * T + (R + dx);
    float sum = 0.0f;
    for (int y = 0; y < 8; y++)
        for (int x = 0; x < 8; x++) sum = sum + fabsf(a[y * T + x] - c[y * T + x]);
    sadc = sum * (1.0f / 64.0f);
    cost = sadc + P.lam * (float)(abs(dx) + abs(dy));
}

// step 3 for a dyadic level: the 64 |a - c| of candidate ci in 16-bit pairs (v_sad_u16, two pixels an instruction). The
// candidate's a and c row offsets R - dx and R + dx have one parity, so both rows are four aligned words of E (even) or
// O (odd). Every level-k SAD below 2^24 units is exact in float32 (fb <= 8), so sum k 2^-fb (1/64) is searchCost's float.
__device__ __forceinline__ void searchCostInt(const SearchParams& P, const uint16_t* A, const uint16_t* C, int R, int ci,
                                              float& cost, float& sadc) {
    const int T = 8 + 2 * R, D = 2 * R + 1, N = T * T;
    const int dx = ci % D - R, dy = ci / D - R;
    const int ca = R - dx, cc = R + dx, par = ca & 1;
    const uint32_t* wa = (const uint32_t*)(A + (par ? N : 0)) + (((R - dy) * T + ca - par) >> 1);
    const uint32_t* wc = (const uint32_t*)(C + (par ? N : 0)) + (((R + dy) * T + cc - par) >> 1);
    uint32_t acc = 0u;
#pragma unroll
    for (int y = 0; y < 8; y++)
#pragma unroll
        for (int m = 0; m < 4; m++) acc = __builtin_amdgcn_sad_u16(wa[y * (T / 2) + m], wc[y * (T / 2) + m], acc);
    sadc = (float)acc * (__int_as_float((127 - P.fb) << 23) * (1.0f / 64.0f));
    cost = sadc + P.lam * (float)(abs(dx) + abs(dy));
}

// searchCostInt at a compile-time radius: every LDS offset an immediate
template <int R>
__device__ __forceinline__ void searchCostIntR(const SearchParams& P, const uint16_t* A, const uint16_t* C, int ci,
                                               float& cost, float& sadc) {
    constexpr int T = 8 + 2 * R, D = 2 * R + 1, N = T * T;
    const int dx = ci % D - R, dy = ci / D - R;
    const int ca = R - dx, cc = R + dx, par = ca & 1;
    const uint32_t* wa = (const uint32_t*)(A + (par ? N : 0)) + (((R - dy) * T + ca - par) >> 1);
    const uint32_t* wc = (const uint32_t*)(C + (par ? N : 0)) + (((R + dy) * T + cc - par) >> 1);
    uint32_t acc = 0u;
#pragma unroll
    for (int y = 0; y < 8; y++)
#pragma unroll
        for (int m = 0; m < 4; m++) acc = __builtin_amdgcn_sad_u16(wa[y * (T / 2) + m], wc[y * (T / 2) + m], acc);
    sadc = (float)acc * (__int_as_float((127 - P.fb) << 23) * (1.0f / 64.0f));
    cost = sadc + P.lam * (float)(abs(dx) + abs(dy));
}

// step 4 within groups of `width` lanes: the first of the minimal costs
__device__ __forceinline__ void searchArgmin(float& bestc, int& besti, float& bests, int width) {
    for (int off = width / 2; off >= 1; off >>= 1) {
        const float oc = __shfl_xor(bestc, off, 64);
        const int oi = __shfl_xor(besti, off, 64);
        const float os = __shfl_xor(bests, off, 64);
        if (oc < bestc || (oc == bestc && oi < besti)) {
            bestc = oc;
            besti = oi;
            bests = os;
        }
    }
}

template <int MAXR>
__device__ __forceinline__ void searchBody(const SearchParams& P, int bx0, int) {
    constexpr int MAXT = 8 + 2 * MAXR;
    __shared__ float ta[4][MAXT * MAXT], tc[4][MAXT * MAXT];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;   // wave: uniform
    const int nb = P.bw * P.bh;
    const int blk = bx0 * 4 + wave;
    const bool active = blk < nb;
    const int bx = active ? blk % P.bw : 0, by = active ? blk / P.bw : 0;
    int sx, sy;
    searchSeed(P, blk, active, bx, by, lane, sx, sy);
    const int R = P.R, D = 2 * R + 1;
    float* A = ta[wave];
    float* C = tc[wave];
    const bool ints = P.fb >= 0;   // (uniform) packed 16-bit SADs
    if constexpr (MAXR <= 6) searchWindowsIssued<MAXT>(P, A, C, bx * kBlock, by * kBlock, sx, sy, R, lane, ints);
    else searchWindows(P, A, C, bx * kBlock, by * kBlock, sx, sy, R, lane, ints);
    __syncthreads();
    float bestc = FLT_MAX, bests = 0.0f;
    int besti = INT_MAX;
    for (int ci = lane; ci < D * D; ci += 64) {
        float cost, sadc;
        if (ints) searchCostInt(P, (const uint16_t*)A, (const uint16_t*)C, R, ci, cost, sadc);
        else searchCost(P, A, C, R, ci, cost, sadc);
        if (cost < bestc) {
            bestc = cost;
            besti = ci;
            bests = sadc;
        }
    }
    searchArgmin(bestc, besti, bests, 64);
    if (active && lane == 0) P.out[blk] = make_float2((float)(sx + besti % D - R), (float)(sy + besti / D - R));
    if (P.sadAcc) {   // search.comp:85, one atomic per workgroup into its slot
        __shared__ uint32_t wsad[4];
        if (lane == 0) wsad[wave] = active ? (uint32_t)rintf(bests * 256.0f) : 0u;
        __syncthreads();
        if (threadIdx.x == 0) atomicAdd(P.u + kUSad + (bx0 & (kSlots - 1)), wsad[0] + wsad[1] + wsad[2] + wsad[3]);
    }
}

// The seeds of a wave's two blocks on a dyadic level (values k 2^-fb: every SAD below 2^24 units is exact in float32,
// so it does not depend on the order of its sum): lane 32 h + 4 kk + rp takes candidate slot kk (0..7) of block h and
// its rows 2 rp, 2 rp + 1, two 8-texel rows of each frame (four 16-byte loads when the rows lie inside the plane),
// then the sums of the slot's four lanes, the cost SAD/64 (+ lambda for all but slot 0; slots past nc never chosen)
// and the first minimum over the slots in order, as seedPick's strict < chooses. A slot whose vector repeats an earlier
// one's costs that one's SAD plus lambda >= 0 and is never chosen, so repeats need no test. In two steps, so that a
// caller can issue other loads (the windows at slot 0's vector) between them: the slot's vector (seedVectors'
// candidate kk), then the sums and the choice, the chosen vector of block lane >> 5 at every lane.
struct alignas(4) F4 {
    float x, y, z, w;
};

__device__ __forceinline__ void seedSlotVector(const SearchParams& P, const int* bxs, const int* bys, const bool* acts,
                                               int blk0, int lane, int& cx, int& cy) {
    const int h = lane >> 5, kk = (lane >> 2) & 7;
    const int bx = h ? bxs[1] : bxs[0], by = h ? bys[1] : bys[0];
    const int nc = P.parent ? (P.extraSeed ? 7 : 6) : 2;
    const int k = P.parent ? kk : (kk == 0 ? 4 : 6);
    cx = 0;
    cy = 0;
    if (kk >= nc) return;
    if (k < 4) {
        const int px = min(bx / 2, P.pw - 1), py = min(by / 2, P.ph - 1);
        const int ssy = (by & 1) == 0 ? -1 : 1, ssx = (bx & 1) == 0 ? -1 : 1;
        const int ox = (k == 2 || k == 3) ? ssx : 0, oy = (k == 1 || k == 3) ? ssy : 0;
        const float2 v = P.parent[clampi(py + oy, 0, P.ph - 1) * P.pw + clampi(px + ox, 0, P.pw - 1)];
        cx = 2 * (int)rintf(v.x);
        cy = 2 * (int)rintf(v.y);
    } else if (k == 5) {
        cx = (int)rintf(__uint_as_float(P.med[0]));
        cy = (int)rintf(__uint_as_float(P.med[1]));
    } else if (k == 6) {
        const bool active = h ? acts[1] : acts[0];
        const float2 v = P.extraSeed[active ? blk0 + h : 0];
        cx = (int)rintf(v.x);
        cy = (int)rintf(v.y);
    }
}

__device__ __forceinline__ void seedPickRows(const SearchParams& P, const int* bxs, const int* bys, int lane, int cx,
                                             int cy, int& sx, int& sy) {
    const int h = lane >> 5, kk = (lane >> 2) & 7, rp = lane & 3;
    const int bx = h ? bxs[1] : bxs[0], by = h ? bys[1] : bys[0];
    const int nc = P.parent ? (P.extraSeed ? 7 : 6) : 2;
    const int xa = bx * kBlock - cx, xc = bx * kBlock + cx;
    float sum = 0.0f;
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int y = by * kBlock + 2 * rp + j, ya = y - cy, yc = y + cy;
        if (xa >= 0 && xa + 7 <= P.w - 1 && xc >= 0 && xc + 7 <= P.w - 1 && ya >= 0 && ya <= P.h - 1 && yc >= 0 &&
            yc <= P.h - 1) {
            const F4* ra = (const F4*)(P.la + idx2(xa, ya, P.w));
            const F4* rc = (const F4*)(P.lc + idx2(xc, yc, P.w));
            const F4 a0 = ra[0], a1 = ra[1], c0 = rc[0], c1 = rc[1];
            sum = sum + fabsf(a0.x - c0.x) + fabsf(a0.y - c0.y) + fabsf(a0.z - c0.z) + fabsf(a0.w - c0.w) +
                  fabsf(a1.x - c1.x) + fabsf(a1.y - c1.y) + fabsf(a1.z - c1.z) + fabsf(a1.w - c1.w);
        } else {   // a row that leaves the plane: the clamped texels
#pragma unroll
            for (int x = 0; x < 8; x++)
                sum = sum + fabsf(lumaAt(P.la, xa + x, ya, P.w, P.h) - lumaAt(P.lc, xc + x, yc, P.w, P.h));
        }
    }
    sum = sum + __shfl_xor(sum, 1, 64);
    sum = sum + __shfl_xor(sum, 2, 64);
    float c = sum * (1.0f / 64.0f);
    if (kk > 0) c = c + P.lam;
    if (kk >= nc) c = FLT_MAX;
    // the first minimum over the block's 8 slots (lanes 4 kk of its half)
    int bk = kk;
    sx = cx;
    sy = cy;
#pragma unroll
    for (int o = 4; o <= 16; o <<= 1) {
        const float oc = __shfl_xor(c, o, 64);
        const int ok = __shfl_xor(bk, o, 64), ox = __shfl_xor(sx, o, 64), oy = __shfl_xor(sy, o, 64);
        if (oc < c || (oc == c && ok < bk)) {
            c = oc;
            bk = ok;
            sx = ox;
            sy = oy;
        }
    }
}

// R <= 2 ((2R + 1)^2 <= 25 candidates): two blocks a wavefront, eight a workgroup. The seeds and windows of a wave's two
// blocks are formed one after the other with all 64 lanes, then lanes 0-31 take the first block's candidates and lanes
// 32-63 the second's (each lane one candidate, its SAD in the order above), and the argmin runs within each half:
// the same vectors with twice the lanes busy. R is a template argument (1 or 2), so window and LDS offsets are
// constants.
constexpr int kSearch2MaxT = 12;
template <int R>
__device__ __forceinline__ void search2BodyR(const SearchParams& P, int bx0, float (*ta)[kSearch2MaxT * kSearch2MaxT],
                                             float (*tc)[kSearch2MaxT * kSearch2MaxT]) {
    constexpr int D = 2 * R + 1;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;   // wave: uniform
    const int nb = P.bw * P.bh;
    int sxs[2] = {0, 0}, sys[2] = {0, 0}, bxs[2], bys[2];
    bool acts[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int blk = bx0 * kSearch2Blocks + 2 * wave + h;
        acts[h] = blk < nb;
        bxs[h] = acts[h] ? blk % P.bw : 0;
        bys[h] = acts[h] ? blk / P.bw : 0;
    }
    SearchWin<R> win[2];
    const bool ints = P.fb >= 0;   // (uniform) packed 16-bit SADs
    if ((P.parent || P.extraSeed) && ints) {   // the seeds by slot and row pairs (exact sums)
        int cx, cy, sx, sy, c0x[2], c0y[2];
        seedSlotVector(P, bxs, bys, acts, bx0 * kSearch2Blocks + 2 * wave, lane, cx, cy);
        // the windows at slot 0's vector (the own parent's, or zero on the coarsest level), usually the seed, loaded
        // with the slots' rows
#pragma unroll
        for (int h = 0; h < 2; h++) {
            c0x[h] = __builtin_amdgcn_readlane(cx, 32 * h);
            c0y[h] = __builtin_amdgcn_readlane(cy, 32 * h);
            searchWinLoad<R>(P, bxs[h], bys[h], c0x[h], c0y[h], lane, win[h]);
        }
        seedPickRows(P, bxs, bys, lane, cx, cy, sx, sy);
#pragma unroll
        for (int h = 0; h < 2; h++) {
            sxs[h] = __builtin_amdgcn_readlane(sx, 32 * h);
            sys[h] = __builtin_amdgcn_readlane(sy, 32 * h);
            if (sxs[h] != c0x[h] || sys[h] != c0y[h])   // (wave-uniform) another seed: its windows
                searchWinLoad<R>(P, bxs[h], bys[h], sxs[h], sys[h], lane, win[h]);
        }
    } else if (P.parent || P.extraSeed) {   // both blocks' seeds side by side, each step for the two together
        SeedSet S[2];
#pragma unroll
        for (int h = 0; h < 2; h++) seedVectors(P, bx0 * kSearch2Blocks + 2 * wave + h, acts[h], bxs[h], bys[h], S[h]);
        // the windows at the first candidate's vector (the own parent's, or zero on the coarsest level), which is
        // usually the seed, loaded with the candidates' texels: one memory latency instead of two
#pragma unroll
        for (int h = 0; h < 2; h++) searchWinLoad<R>(P, bxs[h], bys[h], S[h].cx[0], S[h].cy[0], lane, win[h]);
#pragma unroll
        for (int h = 0; h < 2; h++) seedTexels(P, bxs[h], bys[h], lane, S[h]);
#pragma unroll
        for (int h = 0; h < 2; h++) seedPick(P, S[h], sxs[h], sys[h]);
#pragma unroll
        for (int h = 0; h < 2; h++)
            if (sxs[h] != S[h].cx[0] || sys[h] != S[h].cy[0])   // (wave-uniform) another seed: its windows
                searchWinLoad<R>(P, bxs[h], bys[h], sxs[h], sys[h], lane, win[h]);
    } else {
#pragma unroll
        for (int h = 0; h < 2; h++) searchWinLoad<R>(P, bxs[h], bys[h], 0, 0, lane, win[h]);
    }
#pragma unroll
    for (int h = 0; h < 2; h++) searchWinStore<R>(P, win[h], ta[2 * wave + h], tc[2 * wave + h], lane, ints);
    __syncthreads();
    const int half = lane >> 5, ci = lane & 31;
    const int blk = bx0 * kSearch2Blocks + 2 * wave + half;
    const bool active = blk < nb;
    float bestc = FLT_MAX, bests = 0.0f;
    int besti = INT_MAX;
    if (ci < D * D) {
        float cost, sadc;
        if (ints) searchCostIntR<R>(P, (const uint16_t*)ta[2 * wave + half], (const uint16_t*)tc[2 * wave + half], ci, cost, sadc);
        else searchCost(P, ta[2 * wave + half], tc[2 * wave + half], R, ci, cost, sadc);
        bestc = cost;
        besti = ci;
        bests = sadc;
    }
    searchArgmin(bestc, besti, bests, 32);
    const int sx = half ? sxs[1] : sxs[0], sy = half ? sys[1] : sys[0];
    if (active && ci == 0) P.out[blk] = make_float2((float)(sx + besti % D - R), (float)(sy + besti / D - R));
    if (P.sadAcc) {   // search.comp:85, one atomic per workgroup into its slot
        __shared__ uint32_t wsad[kSearch2Blocks];
        if (ci == 0) wsad[2 * wave + half] = active ? (uint32_t)rintf(bests * 256.0f) : 0u;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t sum = 0u;
            for (int i = 0; i < kSearch2Blocks; i++) sum += wsad[i];
            atomicAdd(P.u + kUSad + (bx0 & (kSlots - 1)), sum);
        }
    }
}

__device__ __forceinline__ void search2Body(const SearchParams& P, int bx0, int) {
    // one LDS allocation for both radii (each instantiation's own arrays would both be allocated)
    __shared__ float ta[kSearch2Blocks][kSearch2MaxT * kSearch2MaxT], tc[kSearch2Blocks][kSearch2MaxT * kSearch2MaxT];
    if (P.R == 2) search2BodyR<2>(P, bx0, ta, tc);   // (uniform) the radius of the level
    else search2BodyR<1>(P, bx0, ta, tc);
}

__device__ __forceinline__ void median3Body(const Median3Params& P, int gbx, int gby) {
    const int bx = gbx * 16 + (threadIdx.x & 15), by = gby * 16 + (threadIdx.x >> 4);
    if (bx >= P.bw || by >= P.bh) return;
    P.out[by * P.bw + bx] = median3At(P.in, bx, by, P.bw, P.bh);
}

// Two u16 lanes dotted with two u16 weights plus c (v_dot2_u32_u16): the exact integer form of a bilinear tap row
typedef unsigned short U16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pairDot(uint32_t pair, uint32_t w, uint32_t c = 0u) {
    return __builtin_amdgcn_udot2(__builtin_bit_cast(U16x2, pair), __builtin_bit_cast(U16x2, w), c, false);
}

// ---------------------------------------------------------------------------------------------------------- subpix

// subpix.comp:15-24 tap(): bilinear in tile coordinates (p >= 0), weights exact for quarter steps
__device__ __forceinline__ float tileTap(const float* t, int T, float px, float py) {
    const float flx = floorf(px), fly = floorf(py);
    const float fx = px - flx, fy = py - fly;
    const int i = (int)fly * T + (int)flx;
    const float p00 = t[i], p01 = t[i + 1], p10 = t[i + T], p11 = t[i + T + 1];
    return p00 * ((1.0f - fy) * (1.0f - fx)) + p01 * ((1.0f - fy) * fx) + p10 * (fy * (1.0f - fx)) + p11 * (fy * fx);
}

// subpix.comp:26-73 (interp.py:162-163): 3x3 at 0.5 px then 3x3 at 0.25 px around the median-filtered integer vector,
// cost = SAD/64 + (lambda * step) * |d|_1, strict <; adds roundEven(256 * winning SAD) to the scene-cut sum.
// One 64-lane workgroup takes 7 blocks; lane 9j + k computes candidate k of block j over all 64 pixels, summing each
// row over x and then the rows (subpix.comp:42-55), so every SAD is the shader's float. The taps of one candidate
// share their fractional offset, so the four bilinear weights are formed once per candidate (the products tap()
// forms, subpix.comp:15-24). Both frames' 13x13 windows of the 7 blocks sit in LDS.


__device__ __forceinline__ void subpixBody(const SubpixParams& P, int g, int) {
    constexpr int T = 13, O = 2, NB = kSubpixBlocks;
    __shared__ float ta[NB][T * T], tc[NB][T * T];
    __shared__ float cst[NB][9], sads[NB][9];
    __shared__ int sv[NB][2], sb[NB][2];
    __shared__ uint32_t wsad[NB];
    const int lane = threadIdx.x;
    const int nb = P.bw * P.bh;
    if (lane < NB) {
        const int blk = g * NB + lane;
        const int b0 = blk < nb ? blk : 0;
        float2 v0;
        if (P.raw) {   // (uniform) level 0's median3, formed here for this workgroup's blocks
            v0 = median3At(P.raw, b0 % P.bw, b0 / P.bw, P.bw, P.bh);
            if (blk < nb) P.filtered[blk] = v0;
        } else {
            v0 = P.in[b0];
        }
        sv[lane][0] = (int)rintf(v0.x);
        sv[lane][1] = (int)rintf(v0.y);
        const int b = min(blk, nb - 1);
        sb[lane][0] = (b % P.bw) * kBlock;
        sb[lane][1] = (b / P.bw) * kBlock;
    }
    __syncthreads();
    // the 7 blocks' windows: every lane's loads issued before any store (one memory latency for the fill); element i
    // of the flattened [NB][T * T] windows, the lanes past the end reading the last element again without storing it
    constexpr int N = NB * T * T, IT = (N + 63) / 64;
    float va[IT], vc[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int i = min(lane + 64 * it, N - 1);
        const int j = i / (T * T), r = i - j * (T * T);
        const int x = r % T - O, y = r / T - O;
        va[it] = lumaAt(P.la, sb[j][0] - sv[j][0] + x, sb[j][1] - sv[j][1] + y, P.w, P.h);
        vc[it] = lumaAt(P.lc, sb[j][0] + sv[j][0] + x, sb[j][1] + sv[j][1] + y, P.w, P.h);
    }
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int i = lane + 64 * it;
        if (i < N) {
            (&ta[0][0])[i] = va[it];
            (&tc[0][0])[i] = vc[it];
        }
    }
    __syncthreads();
    const int j = min(lane / 9, NB - 1), k = lane - 9 * (lane / 9);   // lane 63: a duplicate of lane 62's work
    const float* A = ta[j];
    const float* C = tc[j];
    const float dx = (float)(k % 3 - 1), dy = (float)(k / 3 - 1);
    float basex = 0.0f, basey = 0.0f, bestsad = 0.0f;
    for (int pass = 0; pass < 2; pass++) {
        const float step = pass == 0 ? 0.5f : 0.25f;
        const float ox = basex + dx * step, oy = basey + dy * step;
        // tap coordinates (x + O) -/+ o: integer part x + floor(O -/+ o), one fractional part per candidate
        const float pax = (float)O - ox, pay = (float)O - oy, pcx = (float)O + ox, pcy = (float)O + oy;
        const float flax = floorf(pax), flay = floorf(pay), flcx = floorf(pcx), flcy = floorf(pcy);
        const float fax = pax - flax, fay = pay - flay, fcx = pcx - flcx, fcy = pcy - flcy;
        const float a00 = (1.0f - fay) * (1.0f - fax), a01 = (1.0f - fay) * fax, a10 = fay * (1.0f - fax), a11 = fay * fax;
        const float c00 = (1.0f - fcy) * (1.0f - fcx), c01 = (1.0f - fcy) * fcx, c10 = fcy * (1.0f - fcx), c11 = fcy * fcx;
        const int ia = (int)flay * T + (int)flax, ic = (int)flcy * T + (int)flcx;
        float total = 0.0f;
        // the 9 taps of window row y and y + 1 per frame; row y + 1 is carried into the next row (one LDS row a step)
        float ta0[9], tc0[9];
#pragma unroll
        for (int x = 0; x < 9; x++) {
            ta0[x] = A[ia + x];
            tc0[x] = C[ic + x];
        }
#pragma unroll 2
        for (int y = 0; y < 8; y++) {
            float ta1[9], tc1[9];
#pragma unroll
            for (int x = 0; x < 9; x++) {
                ta1[x] = A[ia + (y + 1) * T + x];
                tc1[x] = C[ic + (y + 1) * T + x];
            }
            float row = 0.0f;
#pragma unroll
            for (int x = 0; x < 8; x++) {
                const float va = ta0[x] * a00 + ta0[x + 1] * a01 + ta1[x] * a10 + ta1[x + 1] * a11;
                const float vc = tc0[x] * c00 + tc0[x + 1] * c01 + tc1[x] * c10 + tc1[x + 1] * c11;
                row = row + fabsf(va - vc);
            }
            total = total + row;
#pragma unroll
            for (int x = 0; x < 9; x++) {
                ta0[x] = ta1[x];
                tc0[x] = tc1[x];
            }
        }
        const float sad = total * (1.0f / 64.0f);
        if (lane < 9 * NB) {
            sads[j][k] = sad;
            cst[j][k] = sad + P.lam * step * (fabsf(dx) + fabsf(dy));
        }
        __syncthreads();
        int bi = 0;
        float best = cst[j][0];
        for (int kk = 1; kk < 9; kk++)
            if (cst[j][kk] < best) {
                best = cst[j][kk];
                bi = kk;
            }
        basex = basex + (float)(bi % 3 - 1) * step;
        basey = basey + (float)(bi / 3 - 1) * step;
        bestsad = sads[j][bi];
        __syncthreads();
    }
    const int blk = g * NB + j;
    const bool active = lane < 9 * NB && blk < nb;
    if (k == 0 && lane < 9 * NB) {
        if (active) P.out[blk] = make_float2((float)sv[j][0] + basex, (float)sv[j][1] + basey);
        wsad[j] = active ? (uint32_t)rintf(bestsad * 256.0f) : 0u;
    }
    __syncthreads();
    // subpix.comp:65-69: the blocks' SADs into the scene-cut sum, one atomic per workgroup into its slot
    if (lane == 0) {
        uint32_t sum = 0u;
        for (int i = 0; i < NB; i++) sum += wsad[i];
        atomicAdd(P.u + kUSad + (g & (kSlots - 1)), sum);
    }
}

// The same refinement in exact integers, for a level-0 luma of values k 2^-fb with 0 <= k < 2^14 and fb <= 6
// (SubpixParams.fb): at quarter-pixel offsets every bilinear weight is a multiple of 1/16, so a tap is exact in float32
// and, in units of 2^-(fb + 4), the integer sum of t W over its four texels (t = k, W = 16 w in 0..16); the row sums and
// the total of the 64 errors are exact too (at most 14 + fb + 4 <= 24 bits), so the float SAD is the integer total times
// 2^-(fb + 4) and every cost, choice and scene-cut term is the float path's. A tap is two v_dot2_u32_u16 over the texel
// pairs of its rows with the weight pairs (W00, W01) and (W10, W11), an error one v_sad_u32 into the running total.
//
// Windows: the offsets stay within 0.75 px of the integer vector, so a pixel x + O of the float path's 13 x 13 tile taps
// columns floor(x + O -/+ o) .. + 1 within 1 .. 10 (rows alike): each frame's 10 x 10 texels of a block (the tile's
// rows and columns 1 .. 10) as u16 rows of 10 stored twice, E16[i] = t[i] and O16[i] = t[i + 1], so that the texel pair
// (t[i], t[i + 1]) is the aligned word i / 2 of E16 (i even) or (i - 1) / 2 of O16 (i odd).
//
// A workgroup is kSubpixIntWaves wavefronts, each with its own 7 blocks (group gw = g kSubpixIntWaves + wave) and its
// own 6.2 KB of LDS: single-wavefront workgroups would be held to the CU's workgroup slots rather than its registers and
// LDS. Every wavefront runs the same barriers.
__device__ __forceinline__ void subpixIntBody(const SubpixParams& P, int gwg, int) {
    constexpr int T = 10, TS = 10, O = 2, NB = kSubpixBlocks, WB = T * TS, NW = kSubpixIntWaves;   // WB u16 a block window
    static_assert(TS % 2 == 0 && TS >= T, "rows of whole texel-pair words");
    __shared__ __align__(16) uint16_t eaW[NW][NB * WB], oaW[NW][NB * WB], ecW[NW][NB * WB], ocW[NW][NB * WB];
    __shared__ float cstW[NW][NB][9], sadsW[NW][NB][9];
    __shared__ int svW[NW][NB][2], sbW[NW][NB][2];
    __shared__ uint32_t wsadW[NW][NB];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = gwg * NW + wave;
    uint16_t* ea = eaW[wave];
    uint16_t* oa = oaW[wave];
    uint16_t* ec = ecW[wave];
    uint16_t* oc = ocW[wave];
    float(*cst)[9] = cstW[wave];
    float(*sads)[9] = sadsW[wave];
    int(*sv)[2] = svW[wave];
    int(*sb)[2] = sbW[wave];
    uint32_t* wsad = wsadW[wave];
    const int nb = P.bw * P.bh;
    if (lane < NB) {
        const int blk = g * NB + lane;
        const int b0 = blk < nb ? blk : 0;
        float2 v0;
        if (P.raw) {   // (uniform) level 0's median3, formed here for this workgroup's blocks
            v0 = median3At(P.raw, b0 % P.bw, b0 / P.bw, P.bw, P.bh);
            if (blk < nb) P.filtered[blk] = v0;
        } else {
            v0 = P.in[b0];
        }
        sv[lane][0] = (int)rintf(v0.x);
        sv[lane][1] = (int)rintf(v0.y);
        const int b = min(blk, nb - 1);
        sb[lane][0] = (b % P.bw) * kBlock;
        sb[lane][1] = (b / P.bw) * kBlock;
    }
    __syncthreads();
    // the 7 blocks' windows, block by block (the block's corner and seed wave-uniform), two 64-texel chunks of its
    // 10 x 10 texels a block (a lane's row and column fixed per chunk); every load issued before any store
    constexpr int KC = (T * T + 63) / 64;
    const float up = __int_as_float((127 + P.fb) << 23);
    int cr[KC], cc[KC];
#pragma unroll
    for (int kc = 0; kc < KC; kc++) {
        const int r = min(lane + 64 * kc, T * T - 1);   // past the window: the last texel again, not stored
        cr[kc] = r / T;
        cc[kc] = r - cr[kc] * T;
    }
    float va[NB][KC], vc[NB][KC];
#pragma unroll
    for (int jb = 0; jb < NB; jb++) {
        const int bx = __builtin_amdgcn_readfirstlane(sb[jb][0]), by = __builtin_amdgcn_readfirstlane(sb[jb][1]);
        const int vx = __builtin_amdgcn_readfirstlane(sv[jb][0]), vy = __builtin_amdgcn_readfirstlane(sv[jb][1]);
#pragma unroll
        for (int kc = 0; kc < KC; kc++) {
            va[jb][kc] = lumaAt(P.la, bx - vx - O + 1 + cc[kc], by - vy - O + 1 + cr[kc], P.w, P.h);
            vc[jb][kc] = lumaAt(P.lc, bx + vx - O + 1 + cc[kc], by + vy - O + 1 + cr[kc], P.w, P.h);
        }
    }
#pragma unroll
    for (int kc = 0; kc < KC; kc++) {
        if (lane + 64 * kc < T * T) {
            const int e0 = cr[kc] * TS + cc[kc];   // the texel's u16 index in its block's window
#pragma unroll
            for (int jb = 0; jb < NB; jb++) {
                const int e = jb * WB + e0;
                const uint16_t ta = (uint16_t)(uint32_t)(va[jb][kc] * up), tc = (uint16_t)(uint32_t)(vc[jb][kc] * up);
                ea[e] = ta;
                ec[e] = tc;
                if (e > 0) {
                    oa[e - 1] = ta;
                    oc[e - 1] = tc;
                }
            }
        }
    }
    __syncthreads();
    const int j = min(lane / 9, NB - 1), k = lane - 9 * (lane / 9);   // lane 63: a duplicate of lane 62's work
    const uint32_t* EA = (const uint32_t*)(ea + j * WB);
    const uint32_t* OA = (const uint32_t*)(oa + j * WB);
    const uint32_t* EC = (const uint32_t*)(ec + j * WB);
    const uint32_t* OC = (const uint32_t*)(oc + j * WB);
    const float dx = (float)(k % 3 - 1), dy = (float)(k / 3 - 1);
    const float sadScale = __int_as_float((127 - 4 - P.fb) << 23) * (1.0f / 64.0f);
    float basex = 0.0f, basey = 0.0f, bestsad = 0.0f;
    for (int pass = 0; pass < 2; pass++) {
        const float step = pass == 0 ? 0.5f : 0.25f;
        const float ox = basex + dx * step, oy = basey + dy * step;
        const float pax = (float)O - ox, pay = (float)O - oy, pcx = (float)O + ox, pcy = (float)O + oy;
        const float flax = floorf(pax), flay = floorf(pay), flcx = floorf(pcx), flcy = floorf(pcy);
        // quarter units of the fractional parts, the 16 w weight pairs of each tap row
        const uint32_t qax = (uint32_t)((pax - flax) * 4.0f), qay = (uint32_t)((pay - flay) * 4.0f);
        const uint32_t qcx = (uint32_t)((pcx - flcx) * 4.0f), qcy = (uint32_t)((pcy - flcy) * 4.0f);
        const uint32_t wa0 = ((4u - qay) * (4u - qax)) | (((4u - qay) * qax) << 16), wa1 = (qay * (4u - qax)) | ((qay * qax) << 16);
        const uint32_t wc0 = ((4u - qcy) * (4u - qcx)) | (((4u - qcy) * qcx) << 16), wc1 = (qcy * (4u - qcx)) | ((qcy * qcx) << 16);
        // the texel pair at window column c of a row: word (row TS + c) >> 1 of E (c even) or O (c odd); pixel x reads
        // window column col + x (the tile's column col + 1 + x), so even x keeps col's parity and odd x flips it
        const int colA = (int)flax - 1, colC = (int)flcx - 1;   // 0 or 1
        const uint32_t* A0 = (colA & 1) ? OA : EA;
        const uint32_t* A1 = (colA & 1) ? EA : OA;
        const uint32_t* C0 = (colC & 1) ? OC : EC;
        const uint32_t* C1 = (colC & 1) ? EC : OC;
        const int ra = ((int)flay - 1) * (TS / 2), rc = ((int)flcy - 1) * (TS / 2);
        const int a0 = ra + (colA >> 1), a1 = ra + ((colA + 1) >> 1), c0 = rc + (colC >> 1), c1 = rc + ((colC + 1) >> 1);
        uint32_t pa0[8], pc0[8];
#pragma unroll
        for (int x = 0; x < 8; x++) {
            pa0[x] = (x & 1) ? A1[a1 + (x >> 1)] : A0[a0 + (x >> 1)];
            pc0[x] = (x & 1) ? C1[c1 + (x >> 1)] : C0[c0 + (x >> 1)];
        }
        uint32_t total = 0u;
#pragma unroll
        for (int y = 0; y < 8; y++) {
            const int o = (y + 1) * (TS / 2);
            uint32_t pa1[8], pc1[8];
#pragma unroll
            for (int x = 0; x < 8; x++) {
                pa1[x] = (x & 1) ? A1[a1 + o + (x >> 1)] : A0[a0 + o + (x >> 1)];
                pc1[x] = (x & 1) ? C1[c1 + o + (x >> 1)] : C0[c0 + o + (x >> 1)];
            }
#pragma unroll
            for (int x = 0; x < 8; x++) {
                const uint32_t tA = pairDot(pa1[x], wa1, pairDot(pa0[x], wa0, 0u));
                const uint32_t tC = pairDot(pc1[x], wc1, pairDot(pc0[x], wc0, 0u));
                total = sadU32(tA, tC, total);
            }
#pragma unroll
            for (int x = 0; x < 8; x++) {
                pa0[x] = pa1[x];
                pc0[x] = pc1[x];
            }
        }
        const float sad = (float)total * sadScale;
        if (lane < 9 * NB) {
            sads[j][k] = sad;
            cst[j][k] = sad + P.lam * step * (fabsf(dx) + fabsf(dy));
        }
        __syncthreads();
        int bi = 0;
        float best = cst[j][0];
        for (int kk = 1; kk < 9; kk++)
            if (cst[j][kk] < best) {
                best = cst[j][kk];
                bi = kk;
            }
        basex = basex + (float)(bi % 3 - 1) * step;
        basey = basey + (float)(bi / 3 - 1) * step;
        bestsad = sads[j][bi];
        __syncthreads();
    }
    const int blk = g * NB + j;
    const bool active = lane < 9 * NB && blk < nb;
    if (k == 0 && lane < 9 * NB) {
        if (active) P.out[blk] = make_float2((float)sv[j][0] + basex, (float)sv[j][1] + basey);
        wsad[j] = active ? (uint32_t)rintf(bestsad * 256.0f) : 0u;
    }
    __syncthreads();
    if (lane == 0) {
        uint32_t sum = 0u;
        for (int i = 0; i < NB; i++) sum += wsad[i];
        atomicAdd(P.u + kUSad + (g & (kSlots - 1)), sum);
    }
}

// ---------------------------------------------------------------------------------------------------------- pixfield

// The 8 x 8 block vectors around a 32 x 32 tile at (x0, y0), blocks x0 / 8 - 2 .. x0 / 8 + 5 and y0 / 8 - 2 .. y0 / 8 + 5
// (clamped into the block field as fieldAt clamps): every block vector and bilinear corner that the tile's candidates read
// (its region's pixels, clamped into the frame, lie in blocks x0 / 8 - 1 .. x0 / 8 + 4), staged in LDS once
struct BlockWin {
    const float2* v;
    int bx0, by0;
    __device__ __forceinline__ float2 at(int bx, int by) const { return v[(by - by0) * 8 + (bx - bx0)]; }
};

// pixfield.comp:19-35 cand(): per-pixel candidate k at flow pixel (qx, qy)
__device__ __forceinline__ float2 pixCand(int k, int qx, int qy, const BlockWin& bw) {
    const int bx = qx / kBlock, by = qy / kBlock;
    switch (k) {
    case 0: {
        const float sx = ((float)qx + 0.5f) / 8.0f - 0.5f, sy = ((float)qy + 0.5f) / 8.0f - 0.5f;
        const float flx = floorf(sx), fly = floorf(sy);
        const float fx = sx - flx, fy = sy - fly;
        const int ix = (int)flx, iy = (int)fly;
        const float2 top = mix2(bw.at(ix, iy), bw.at(ix + 1, iy), fx);
        const float2 bot = mix2(bw.at(ix, iy + 1), bw.at(ix + 1, iy + 1), fx);
        return mix2(top, bot, fy);
    }
    case 1: return bw.at(bx, by);
    case 2: return bw.at(bx, by - 1);
    case 3: return bw.at(bx, by + 1);
    case 4: return bw.at(bx - 1, by);
    case 5: return bw.at(bx + 1, by);
    default: return make_float2(0.0f, 0.0f);
    }
}

// Four horizontally adjacent pixels sampled at p - v (sign -1) or p + v (sign +1) with one vector v, each exactly as
// sampleL() would: when the four 1/32 maps are 32 apart with one fractional part and no tap needs clamping, the four
// samples share their weights and read 5 consecutive texels per row. Returns false when that does not hold.
__device__ __forceinline__ bool sampleRun4(const float* img, int w, int h, int qx, int qy, float2 v, float sign, float* out) {
    const float px0 = (float)qx + sign * v.x, px3 = (float)(qx + 3) + sign * v.x, py = (float)qy + sign * v.y;
    const int X0 = (int)rintf(px0 * 32.0f), X3 = (int)rintf(px3 * 32.0f), Y = (int)rintf(py * 32.0f);
    const int ix = X0 >> 5, iy = Y >> 5;
    if (X3 != X0 + 96 || ix < 0 || ix + 4 > w - 1 || iy < 0 || iy + 1 > h - 1) return false;
    const float fx = (float)(X0 & 31) * (1.0f / 32.0f), fy = (float)(Y & 31) * (1.0f / 32.0f);
    const float w00 = (1.0f - fy) * (1.0f - fx), w01 = (1.0f - fy) * fx, w10 = fy * (1.0f - fx), w11 = fy * fx;
    const float* r0 = img + idx2(ix, iy, w);   // a row pointer, so that the 5 taps of a row load as dwordx4 + dword
    const float* r1 = r0 + w;
    float a[5], b[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
        a[i] = r0[i];
        b[i] = r1[i];
    }
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = a[i] * w00 + a[i + 1] * w01 + b[i] * w10 + b[i + 1] * w11;
    return true;
}

// pixfield.comp:37-63 (interp.py:167-183): per flow pixel, the candidate whose 5x5 window of |a(q - v) - c(q + v)|
// (each window pixel with its own candidate, replicate border) costs least; cost = float32(exact window sum * (1/25))
// with the sum in float64 (cv2.blur's CV_32F path), - smoothBonus for the bilinear candidate, - hudBonus for zero,
// strict < (first wins).
//
// 32 x 32 flow pixels per workgroup. The candidates are taken one at a time: the errors of the tile and its apron go to
// LDS in runs of 4 pixels of one row (inside one 8x8 block, since runs start at multiples of 4), then the horizontal
// 5-sums (float64, 4 per task), then each pixel's vertical sum (a thread owns a column segment) and the running minimum
// in registers. Work a run shares:
//   * candidates 1-5 (a block vector): one vector for the run, one set of bilinear weights (sampleRun4); a run whose
//     vector equals its own-block vector (candidate 1) reuses those errors -- a smooth field skips most of the work;
//   * candidate 0 (the bilinear field): the run's pixels share the four block vectors they interpolate (their sample
//     positions (q + 1/2) / 8 - 1/2 have one integer part within a run of 4 starting at a multiple of 4);
//   * candidate 6 (zero): a texel difference (weights 1, 0, 0, 0).
constexpr int kPfTile = 32, kPfRW = 40, kPfRH = kPfTile + 4, kPfRuns = (kPfRW / 4) * kPfRH;
constexpr int kPfHsStride = 36;   // words a row of the horizontal sums (16-byte rows; rows 4 apart in other banks)
constexpr int kPfThreads = 512, kPfBatchThreads = 256;

// the block offset of candidate k = 1..5 (own, above, below, left, right; pixfield.comp:29-33)
__device__ __forceinline__ int2 candOffset(int k) {
    const int dy = k == 2 ? -1 : (k == 3 ? 1 : 0);
    const int dx = k == 4 ? -1 : (k == 5 ? 1 : 0);
    return make_int2(dx, dy);
}

// the errors of the run of 4 pixels (gx0 .. gx0 + 3, row qy) under candidate k. own / ownV: the run's errors and
// vector under candidate 1, kept from k = 1 (computed first) for runs inside the frame (ownOk). Returns true when the
// run's errors are own's because every pixel's vector equals ownV (then the errors are the same floats).
__device__ __forceinline__ bool pixErrRun(const PixParams& P, const BlockWin& bw, int k, int gx0, int qy, float* e,
                                          float* own, float2& ownV, bool& ownOk) {
    if (k == 6) {   // zero vector: sampleL at an integer position is the texel (weights 1, 0, 0, 0)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint32_t q = idx2(clampi(gx0 + i, 0, P.w - 1), qy, P.w);
            e[i] = fabsf(ldIdx(P.la, q) - ldIdx(P.lc, q));
        }
        return false;
    }
    const bool inside = gx0 >= 0 && gx0 + 3 <= P.w - 1;
    if (k == 0 && inside) {   // pixCand(0, ..) of each pixel with the four block vectors loaded once
        const float sx0 = ((float)gx0 + 0.5f) / 8.0f - 0.5f, sy = ((float)qy + 0.5f) / 8.0f - 0.5f;
        const float flx = floorf(sx0), fly = floorf(sy);
        const float fy = sy - fly;
        const int ix = (int)flx, iy = (int)fly;
        const float2 f00 = bw.at(ix, iy), f10 = bw.at(ix + 1, iy), f01 = bw.at(ix, iy + 1), f11 = bw.at(ix + 1, iy + 1);
        auto cand0 = [&](int qx) {
            const float fx = (((float)qx + 0.5f) / 8.0f - 0.5f) - flx;
            return mix2(mix2(f00, f10, fx), mix2(f01, f11, fx), fy);
        };
        bool same = ownOk;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float2 v = cand0(gx0 + i);
            same = same && v.x == ownV.x && v.y == ownV.y;
        }
        if (same) {
#pragma unroll
            for (int i = 0; i < 4; i++) e[i] = own[i];
            return true;
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {   // the vectors formed again rather than held in registers
            const int qx = gx0 + i;
            const float2 v = cand0(qx);
            e[i] = fabsf(sampleL(P.la, (float)qx - v.x, (float)qy - v.y, P.w, P.h) -
                         sampleL(P.lc, (float)qx + v.x, (float)qy + v.y, P.w, P.h));
        }
        return false;
    }
    if (k > 0 && inside) {
        const int2 o = candOffset(k);
        const float2 v = bw.at(gx0 / kBlock + o.x, qy / kBlock + o.y);   // one block, one vector
        if (k > 1 && ownOk && v.x == ownV.x && v.y == ownV.y) {
#pragma unroll
            for (int i = 0; i < 4; i++) e[i] = own[i];
            return true;
        }
        float sa[4], sc[4];
        if (sampleRun4(P.la, P.w, P.h, gx0, qy, v, -1.0f, sa) && sampleRun4(P.lc, P.w, P.h, gx0, qy, v, 1.0f, sc)) {
#pragma unroll
            for (int i = 0; i < 4; i++) e[i] = fabsf(sa[i] - sc[i]);
        } else {
#pragma unroll 1
            for (int i = 0; i < 4; i++) {
                const int qx = gx0 + i;
                e[i] = fabsf(sampleL(P.la, (float)qx - v.x, (float)qy - v.y, P.w, P.h) -
                             sampleL(P.lc, (float)qx + v.x, (float)qy + v.y, P.w, P.h));
            }
        }
        if (k == 1) {
#pragma unroll
            for (int i = 0; i < 4; i++) own[i] = e[i];
            ownV = v;
            ownOk = true;
        }
        return false;
    }
#pragma unroll 1
    for (int i = 0; i < 4; i++) {
        const int qx = clampi(gx0 + i, 0, P.w - 1);
        const float2 v = pixCand(k, qx, qy, bw);
        e[i] = fabsf(sampleL(P.la, (float)qx - v.x, (float)qy - v.y, P.w, P.h) -
                     sampleL(P.lc, (float)qx + v.x, (float)qy + v.y, P.w, P.h));
    }
    return false;
}

// The horizontal sums' LDS slot of row y, column c: columns with (c & 2) == 0 in one plane, the others in a second,
// each plane's rows 16 doubles wide, so that a task's two 16-byte stores land in consecutive slots across the lanes
__device__ __forceinline__ int hsumAt(int y, int c) {
    return ((c & 2) ? kPfRH * kPfTile / 2 : 0) + y * (kPfTile / 2) + ((c >> 2) << 1) + (c & 1);
}

// NT = 512 (the per-pair kernel: few workgroups, so latency decides): one run and 2 pixels a thread. NT = 256 (the
// batched kernel: many workgroups, so occupancy decides): up to 2 runs and 4 pixels a thread, fewer registers.
//
// Candidates are taken in the order 1, 0, 2, 3, 4, 5, 6 (candidate 1's errors are what the others reuse) with the
// reference's first-minimum rule kept: 0 replaces 1 on a tie (c0 <= c1), every later one needs strict <. When every
// run of the tile's region reused candidate 1's errors for candidate k (a workgroup vote), every window sum is
// candidate 1's, so candidate 0 costs c1 - smoothBonus and candidates 2-5 cannot win: the sums are skipped.
// ---- the pixel field in exact integers, from LDS windows of the two planes ----
// For a level-0 luma of values k 2^-fb with 0 <= k < 2^14 (RGBA8 at every flowScale; PixParams.fb) every sampleL is
// exact in float32: a tap k m / 1024 (m <= 1024) has at most 14 + 11 significant bits, and the sum, a convex combination,
// at most 14 + 10 bits; so is the error |a - c|, and every window sum of 25 errors fits 32 bits. In integers (units of
// 2^-(fb + 10)) a sample is the sum of t m over its taps (t = k): one v_dot2_u32_u16 per tap row over the texel pair with
// the weights (32 - fx, fx), then the vertical blend; the error is one v_sad_u32. The window sum S gives the cost
// float32(S 2^-(fb + 10) (1/25)), the float32 rounding of cv2.blur's float64 mean, as the float path computes it.
//
// The texels come from LDS windows of both planes covering the tile's region plus the reach R of its largest candidate
// vector: word x of a window row is the texel pair t[x] | t[x + 1] << 16 of the edge-clamped plane, so a window tap is
// the clamped global tap. A tile whose candidates reach further than kPfWinR flow px samples the planes in global memory.
// Bounds: a region pixel q (columns x0 - 4 .. x0 + 35, rows y0 - 2 .. y0 + 33) is sampled at q -+ v with |v| < R, so
// its taps' words lie in columns x0 - 4 - R .. x0 + 35 + R and rows y0 - 2 - R .. y0 + 34 + R of the plane, inside the
// windows' columns x0 - 5 - R .. x0 + 36 + R and rows y0 - 3 - R .. y0 + 35 + R.
constexpr int kPfWinR = 7;
constexpr int kPfWinWords = 42 + 2 * kPfWinR, kPfWinRows = 39 + 2 * kPfWinR;   // at R: 42 + 2R words, 39 + 2R rows
constexpr int kPfWinPool = 2 * kPfWinWords * kPfWinRows * 4;

struct PixWin {
    const uint32_t* a;
    const uint32_t* c;
    int x0, y0, stride;   // plane coordinates of the windows' word 0, words a row
};

// sampleL at the 1/32 map (X, Y), in units of 2^-(fb + 10)
__device__ __forceinline__ uint32_t winSample(const uint32_t* w, const PixWin& W, int X, int Y) {
    const uint32_t fx = (uint32_t)(X & 31), fy = (uint32_t)(Y & 31);
    const uint32_t* p = w + ((Y >> 5) - W.y0) * W.stride + ((X >> 5) - W.x0);
    const uint32_t wx = (32u - fx) | (fx << 16);
    return pairDot(p[0], wx) * (32u - fy) + pairDot(p[W.stride], wx) * fy;
}

// the four samples at X0 + 32 i, Y (one fractional part, 5 texels a row)
__device__ __forceinline__ void winSample4(const uint32_t* w, const PixWin& W, int X0, int Y, uint32_t* out) {
    const uint32_t fx = (uint32_t)(X0 & 31), fy = (uint32_t)(Y & 31);
    const uint32_t* p = w + ((Y >> 5) - W.y0) * W.stride + ((X0 >> 5) - W.x0);
    const uint32_t wx = (32u - fx) | (fx << 16);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = pairDot(p[i], wx) * (32u - fy) + pairDot(p[W.stride + i], wx) * fy;
}

// |sampleL(a, q - v) - sampleL(c, q + v)| with the maps formed as sampleL forms them
__device__ __forceinline__ uint32_t winErr(const PixWin& W, int qx, int qy, float2 v) {
    const int Xa = (int)rintf(((float)qx - v.x) * 32.0f), Ya = (int)rintf(((float)qy - v.y) * 32.0f);
    const int Xc = (int)rintf(((float)qx + v.x) * 32.0f), Yc = (int)rintf(((float)qy + v.y) * 32.0f);
    return sadU32(winSample(W.a, W, Xa, Ya), winSample(W.c, W, Xc, Yc), 0u);
}

// sampleRun4 in two steps, so that a caller issues the loads of both frames (and more) before any arithmetic: the plan
// (the maps, their check, the taps' row and weights), then the sums over the 5 + 5 texels it loads
struct RunPlan {
    uint32_t row;   // element index of the first texel of the top tap row
    float w00, w01, w10, w11;
    bool ok;
};

__device__ __forceinline__ RunPlan runPlan(int w, int h, int qx, int qy, float2 v, float sign) {
    const float px0 = (float)qx + sign * v.x, px3 = (float)(qx + 3) + sign * v.x, py = (float)qy + sign * v.y;
    const int X0 = (int)rintf(px0 * 32.0f), X3 = (int)rintf(px3 * 32.0f), Y = (int)rintf(py * 32.0f);
    const int ix = X0 >> 5, iy = Y >> 5;
    RunPlan p;
    p.ok = X3 == X0 + 96 && ix >= 0 && ix + 4 <= w - 1 && iy >= 0 && iy + 1 <= h - 1;
    const float fx = (float)(X0 & 31) * (1.0f / 32.0f), fy = (float)(Y & 31) * (1.0f / 32.0f);
    p.w00 = (1.0f - fy) * (1.0f - fx);
    p.w01 = (1.0f - fy) * fx;
    p.w10 = fy * (1.0f - fx);
    p.w11 = fy * fx;
    p.row = p.ok ? idx2(ix, iy, w) : 0u;
    return p;
}

__device__ __forceinline__ void runSums(const float* a, const float* b, const RunPlan& p, float* out) {
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = a[i] * p.w00 + a[i + 1] * p.w01 + b[i] * p.w10 + b[i + 1] * p.w11;
}

// |sampleL(a, q - v_i) - sampleL(c, q + v_i)| at the pixels qx_i of row qy, each with its own vector (the loads of all
// four pixels are issued together)
__device__ __forceinline__ void pixErr4(const PixParams& P, const int* qx, int qy, const float2* v, float* ev) {
#pragma unroll
    for (int i = 0; i < 4; i++)
        ev[i] = fabsf(sampleL(P.la, (float)qx[i] - v[i].x, (float)qy - v[i].y, P.w, P.h) -
                      sampleL(P.lc, (float)qx[i] + v[i].x, (float)qy + v[i].y, P.w, P.h));
}

// |sampleL(a, q - v) - sampleL(c, q + v)| at the 4 pixels gx0 .. gx0 + 3 of row qy (inside the plane), one vector
__device__ __forceinline__ void pixRunErrs(const PixParams& P, int gx0, int qy, float2 v, float* ev) {
    const RunPlan pa = runPlan(P.w, P.h, gx0, qy, v, -1.0f), pc = runPlan(P.w, P.h, gx0, qy, v, 1.0f);
    if (pa.ok && pc.ok) {
        const float* ra = P.la + pa.row;
        const float* rc = P.lc + pc.row;
        float a0[5], a1[5], c0[5], c1[5];
#pragma unroll
        for (int i = 0; i < 5; i++) {
            a0[i] = ra[i];
            a1[i] = ra[P.w + i];
            c0[i] = rc[i];
            c1[i] = rc[P.w + i];
        }
        float sa[4], sc[4];
        runSums(a0, a1, pa, sa);
        runSums(c0, c1, pc, sc);
#pragma unroll
        for (int i = 0; i < 4; i++) ev[i] = fabsf(sa[i] - sc[i]);
    } else {
        const int qx[4] = {gx0, gx0 + 1, gx0 + 2, gx0 + 3};
        const float2 vv[4] = {v, v, v, v};
        pixErr4(P, qx, qy, vv, ev);
    }
}

__device__ __forceinline__ float pixErr1(const PixParams& P, int qx, int qy, float2 v) {
    return fabsf(sampleL(P.la, (float)qx - v.x, (float)qy - v.y, P.w, P.h) -
                 sampleL(P.lc, (float)qx + v.x, (float)qy + v.y, P.w, P.h));
}

__device__ __forceinline__ bool sameVec(float2 a, float2 b) { return a.x == b.x && a.y == b.y; }

// The exact path's two sources of errors (integers in units of 2^-(fb + 10)): the float planes in global memory, whose
// samples are exact, or the LDS windows. zero: the texel differences at gx0 .. gx0 + 3 (clamped), run: the errors of
// one vector on a run inside the plane, pix4: four pixels each with its own vector.
struct GlobalErrs {
    const PixParams* P;
    float up;
    __device__ __forceinline__ void zero(int gx0, int qy, uint32_t* e) const {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint32_t q = idx2(clampi(gx0 + i, 0, P->w - 1), qy, P->w);
            e[i] = (uint32_t)(fabsf(ldIdx(P->la, q) - ldIdx(P->lc, q)) * up);
        }
    }
    __device__ __forceinline__ void run(int gx0, int qy, float2 v, uint32_t* e) const {
        float ev[4];
        pixRunErrs(*P, gx0, qy, v, ev);
#pragma unroll
        for (int i = 0; i < 4; i++) e[i] = (uint32_t)(ev[i] * up);
    }
    __device__ __forceinline__ void pix4(const int* qx, int qy, const float2* v, uint32_t* e) const {
        float ev[4];
        pixErr4(*P, qx, qy, v, ev);
#pragma unroll
        for (int i = 0; i < 4; i++) e[i] = (uint32_t)(ev[i] * up);
    }
};

struct WinErrs {
    PixWin W;
    int w;
    __device__ __forceinline__ void zero(int gx0, int qy, uint32_t* e) const {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int o = (qy - W.y0) * W.stride + (clampi(gx0 + i, 0, w - 1) - W.x0);
            e[i] = sadU32(W.a[o] & 0xFFFFu, W.c[o] & 0xFFFFu, 0u) << 10;
        }
    }
    __device__ __forceinline__ void run(int gx0, int qy, float2 v, uint32_t* e) const {
        const int Xa = (int)rintf(((float)gx0 - v.x) * 32.0f), Xa3 = (int)rintf(((float)(gx0 + 3) - v.x) * 32.0f);
        const int Xc = (int)rintf(((float)gx0 + v.x) * 32.0f), Xc3 = (int)rintf(((float)(gx0 + 3) + v.x) * 32.0f);
        if (Xa3 == Xa + 96 && Xc3 == Xc + 96) {   // one map fraction a frame: the run form
            const int Ya = (int)rintf(((float)qy - v.y) * 32.0f), Yc = (int)rintf(((float)qy + v.y) * 32.0f);
            uint32_t sa[4], sc[4];
            winSample4(W.a, W, Xa, Ya, sa);
            winSample4(W.c, W, Xc, Yc, sc);
#pragma unroll
            for (int i = 0; i < 4; i++) e[i] = sadU32(sa[i], sc[i], 0u);
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) e[i] = winErr(W, gx0 + i, qy, v);
        }
    }
    __device__ __forceinline__ void pix4(const int* qx, int qy, const float2* v, uint32_t* e) const {
#pragma unroll
        for (int i = 0; i < 4; i++) e[i] = winErr(W, qx[i], qy, v[i]);
    }
};

// The errors of candidate k (0 or 2-5) on the run of 4 pixels gx0 .. gx0 + 3 of row qy
template <class Errs>
__device__ __forceinline__ void pixCandRunErrs(const PixParams& P, const BlockWin& bw, const Errs& E, int k, int gx0,
                                               int qy, uint32_t* e) {
    if (gx0 >= 0 && gx0 + 3 <= P.w - 1) {
        if (k > 0) {
            const int2 o = candOffset(k);
            E.run(gx0, qy, bw.at((gx0 >> 3) + o.x, (qy >> 3) + o.y), e);   // one block, one vector
            return;
        }
        const float sx0 = ((float)gx0 + 0.5f) / 8.0f - 0.5f, sy = ((float)qy + 0.5f) / 8.0f - 0.5f;
        const float flx = floorf(sx0), fly = floorf(sy);
        const float fy = sy - fly;
        const int ix = (int)flx, iy = (int)fly;
        const float2 f00 = bw.at(ix, iy), f10 = bw.at(ix + 1, iy), f01 = bw.at(ix, iy + 1), f11 = bw.at(ix + 1, iy + 1);
        float2 v[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float fx = (((float)(gx0 + i) + 0.5f) / 8.0f - 0.5f) - flx;
            v[i] = mix2(mix2(f00, f10, fx), mix2(f01, f11, fx), fy);
        }
        if (sameVec(v[1], v[0]) && sameVec(v[2], v[0]) && sameVec(v[3], v[0])) {
            E.run(gx0, qy, v[0], e);
        } else {
            const int qx[4] = {gx0, gx0 + 1, gx0 + 2, gx0 + 3};
            E.pix4(qx, qy, v, e);
        }
        return;
    }
    int qx[4];
    float2 v[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        qx[i] = clampi(gx0 + i, 0, P.w - 1);
        v[i] = pixCand(k, qx[i], qy, bw);
    }
    E.pix4(qx, qy, v, e);
}

// ---- uniform tiles ----
// Every candidate vector of the tile's region (its pixels lie in blocks x0 / 8 - 1 .. x0 / 8 + 4, the same rows, and
// reach one block further in the plus shape of candidates 2-5) is an entry of the 8 x 8 block window other than its
// four corners. When those 60 entries are one vector V bit for bit, and the bilinear candidate of V with itself is V at
// every fraction it uses (odd sixteenths, from (q + 1/2) / 8 - 1/2), candidates 0-5 are V at every pixel and their
// window sums are one sum: candidate 0 costs c1 - smoothBonus, candidates 2-5 cost c1 and (strict <) cannot win. Only
// V's and the zero vector's errors are formed, and the winner is V or zero. The values are those of the seven-candidate
// pass. Integer sums: for exact lumas (fb >= 0) every sampleL is a multiple of 2^-(fb + 10) below 2^24 of them.
constexpr int kPfUW = 36, kPfUH = kPfTile + 4;   // the windows' region: columns x0 - 2 .. x0 + 33, rows y0 - 2 .. y0 + 33
constexpr int kPfUPlane = kPfUH * kPfUW;
constexpr int kPfUniPool = 4 * kPfUPlane * 4;    // errors of V and of zero, then their horizontal 5-sums

// the uniform test, by the wavefront that staged the window (lane i holds entry i): returns V's bits in (vx, vy)
__device__ __forceinline__ bool pixWindowUniform(float2 v, int lane, uint32_t& vx, uint32_t& vy) {
    vx = (uint32_t)__shfl((int)__float_as_uint(v.x), 9, 64);
    vy = (uint32_t)__shfl((int)__float_as_uint(v.y), 9, 64);
    const bool corner = lane == 0 || lane == 7 || lane == 56 || lane == 63;
    bool ok = corner || (__float_as_uint(v.x) == vx && __float_as_uint(v.y) == vy);
    if (lane < 8) {
        const float t = (float)(2 * lane + 1) * (1.0f / 16.0f), Vx = __uint_as_float(vx), Vy = __uint_as_float(vy);
        ok = ok && mixf(Vx, Vx, t) == Vx && mixf(Vy, Vy, t) == Vy;
    }
    return __all(ok);
}

template <int NT, class Errs>
__device__ __forceinline__ void pixfieldUniform(const PixParams& P, int x0, int y0, float2 V, uint32_t* pool, const Errs& E) {
    constexpr int kRows = kPfTile * kPfTile / NT;
    uint32_t* eu = pool;                   // [2][36][36]: |errors| of V, of zero, in units of 2^-(fb + 10)
    uint32_t* hu = pool + 2 * kPfUPlane;   // [2][36][36]: their horizontal 5-sums (columns 0 .. 31 of a row)
    const int tid = threadIdx.x;
    for (int r = tid; r < kPfUH * (kPfUW / 4); r += NT) {
        const int ry = r / (kPfUW / 4), rc = r - ry * (kPfUW / 4);
        const int qy = clampi(y0 - 2 + ry, 0, P.h - 1), gx0 = x0 - 2 + 4 * rc;
        uint32_t iv[4], iz[4];
        E.zero(gx0, qy, iz);
        if (gx0 >= 0 && gx0 + 3 <= P.w - 1) {
            E.run(gx0, qy, V, iv);
        } else {
            const int qx[4] = {clampi(gx0, 0, P.w - 1), clampi(gx0 + 1, 0, P.w - 1), clampi(gx0 + 2, 0, P.w - 1),
                               clampi(gx0 + 3, 0, P.w - 1)};
            const float2 vv[4] = {V, V, V, V};
            E.pix4(qx, qy, vv, iv);
        }
        *(uint4*)&eu[ry * kPfUW + 4 * rc] = make_uint4(iv[0], iv[1], iv[2], iv[3]);
        *(uint4*)&eu[kPfUPlane + ry * kPfUW + 4 * rc] = make_uint4(iz[0], iz[1], iz[2], iz[3]);
    }
    __syncthreads();
    // horizontal 5-sums: output column c of a row sums region columns c .. c + 4; four outputs a task (two quads)
    for (int t = tid; t < 2 * kPfUH * (kPfTile / 4); t += NT) {
        const int pl = t >= kPfUH * (kPfTile / 4) ? 1 : 0, tt = t - pl * kPfUH * (kPfTile / 4);
        const int hy = tt >> 3, c0 = (tt & 7) * 4;
        const uint4* e4 = (const uint4*)&eu[pl * kPfUPlane + hy * kPfUW + c0];
        const uint4 q0 = e4[0], q1 = e4[1];
        uint4 h;
        h.x = q0.x + q0.y + q0.z + q0.w + q1.x;
        h.y = h.x + q1.y - q0.x;
        h.z = h.y + q1.z - q0.y;
        h.w = h.z + q1.w - q0.z;
        *(uint4*)&hu[pl * kPfUPlane + hy * kPfUW + c0] = h;
    }
    __syncthreads();
    const double cscale = (1.0 / 25.0) * (double)__int_as_float((127 - 10 - P.fb) << 23);
    const int cx = tid & 31, r0 = (tid >> 5) * kRows;
    uint32_t hv[kRows + 4], hz[kRows + 4];
#pragma unroll
    for (int y = 0; y < kRows + 4; y++) {
        hv[y] = hu[(r0 + y) * kPfUW + cx];
        hz[y] = hu[kPfUPlane + (r0 + y) * kPfUW + cx];
    }
    uint32_t sv = hv[0] + hv[1] + hv[2] + hv[3] + hv[4], sz = hz[0] + hz[1] + hz[2] + hz[3] + hz[4];
    const int qx = x0 + cx;
    if (qx >= P.w) return;
#pragma unroll
    for (int j = 0; j < kRows; j++) {
        if (j > 0) {
            sv = sv + hv[j + 4] - hv[j - 1];
            sz = sz + hz[j + 4] - hz[j - 1];
        }
        const int py = y0 + r0 + j;
        if (py >= P.h) break;
        // the seven-candidate order: 1, then 0 on c0 <= best, then 2-5 (never below c1), then 6 on strict <
        const float c1 = (float)((double)sv * cscale);
        float best = c1;
        const float c0 = c1 - P.smooth;
        if (c0 <= best) best = c0;
        const float c6 = (float)((double)sz * cscale) - P.zero;
        P.pf[py * P.w + qx] = c6 < best ? make_float2(0.0f, 0.0f) : V;
    }
}

// The seven-candidate pass for exact lumas (fb >= 0): errors as integers in units of 2^-(fb + 10), window sums in 32
// bits, each cost formed as the float path forms it. Candidates 1 and 6 are always formed, together (their errors, then
// their sums). A run of 4 pixels whose candidate k (0 or 2-5) is candidate 1's vector at each of its pixels has
// candidate 1's errors for k (the same floats) and copies them; a candidate that every run of the region copies has
// candidate 1's window sums, so it is decided without a pass of its own (0 costs c1 - smoothBonus, 2-5 cannot win).
// The others get a pass each, in the reference's order 0, 2, 3, 4, 5; candidate 6 is compared last (strict <). The runs
// a pass forms (rather than copies) are listed per candidate in LDS when candidate 1 is formed, and a pass deals its list
// to the threads densely: a tile with one odd block forms a few runs, and no wavefront runs the copy and the forming
// paths side by side.
constexpr int kPfCandLists = 5;   // candidates 0, 2, 3, 4, 5
constexpr int kPfExactPool = 2 * kPfRH * kPfRW * 4 + 2 * kPfRH * kPfHsStride * 4 + kPfCandLists * kPfRuns * 2;

// list slot of candidate k (0, 2-5)
__device__ __forceinline__ int candList(int k) { return k == 0 ? 0 : k - 1; }

template <int NT, class Errs>
__device__ __forceinline__ void pixfieldExact(const PixParams& P, int x0, int y0, const BlockWin& bw, uint8_t* pool,
                                              uint32_t* cnt, const Errs& E) {
    constexpr int kRows = kPfTile * kPfTile / NT;
    constexpr int kRunsRow = kPfRW / 4;
    uint32_t* e1 = (uint32_t*)pool;              // [36][40] candidate 1's errors (kept)
    uint32_t* ek = e1 + kPfRH * kPfRW;           // [36][40] candidate 6's errors, then candidate k's
    uint32_t* h1 = ek + kPfRH * kPfRW;           // horizontal 5-sums of e1, rows kPfHsStride apart
    uint32_t* hk = h1 + kPfRH * kPfHsStride;     // ... of ek
    uint16_t* lists = (uint16_t*)(hk + kPfRH * kPfHsStride);   // [5][360] runs a candidate's pass forms
    const int tid = threadIdx.x;
    const double cscale = (1.0 / 25.0) * (double)__int_as_float((127 - 10 - P.fb) << 23);
    static_assert((kPfRuns + NT - 1) / NT <= 2, "at most two runs a thread");
    uint32_t reuse0 = 0u, reuse1 = 0u;   // bit k: the run's candidate k is candidate 1 at every pixel
    for (int r = tid; r < kPfRuns; r += NT) {
        const int ry = r / kRunsRow, rc = r - ry * kRunsRow;
        const int qy = clampi(y0 - 2 + ry, 0, P.h - 1), gx0 = x0 - 4 + 4 * rc;
        uint32_t i1[4], i6[4];
        E.zero(gx0, qy, i6);
        uint32_t bits = 0u;
        if (gx0 >= 0 && gx0 + 3 <= P.w - 1) {
            const int bx = gx0 >> 3, by = qy >> 3;
            const float2 v1 = bw.at(bx, by);
            E.run(gx0, qy, v1, i1);
#pragma unroll
            for (int k = 2; k <= 5; k++) {
                const int2 o = candOffset(k);
                bits |= sameVec(bw.at(bx + o.x, by + o.y), v1) ? 1u << k : 0u;
            }
            const float sx0 = ((float)gx0 + 0.5f) / 8.0f - 0.5f, sy = ((float)qy + 0.5f) / 8.0f - 0.5f;
            const float flx = floorf(sx0), fly = floorf(sy);
            const float fy = sy - fly;
            const int ix = (int)flx, iy = (int)fly;
            const float2 f00 = bw.at(ix, iy), f10 = bw.at(ix + 1, iy), f01 = bw.at(ix, iy + 1), f11 = bw.at(ix + 1, iy + 1);
            bool same0 = true;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float fx = (((float)(gx0 + i) + 0.5f) / 8.0f - 0.5f) - flx;
                same0 = same0 && sameVec(mix2(mix2(f00, f10, fx), mix2(f01, f11, fx), fy), v1);
            }
            bits |= same0 ? 1u : 0u;
        } else {   // a run across the frame edge: pixel by pixel at the clamped column
            uint32_t all = 0x3Du;
            int qx[4];
            float2 v1[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                qx[i] = clampi(gx0 + i, 0, P.w - 1);
                v1[i] = pixCand(1, qx[i], qy, bw);
                uint32_t b = 0u;
#pragma unroll
                for (int k = 0; k <= 5; k++)
                    if (k != 1) b |= sameVec(pixCand(k, qx[i], qy, bw), v1[i]) ? 1u << k : 0u;
                all &= b;
            }
            E.pix4(qx, qy, v1, i1);
            bits = all;
        }
        *(uint4*)&e1[ry * kPfRW + 4 * rc] = make_uint4(i1[0], i1[1], i1[2], i1[3]);
        *(uint4*)&ek[ry * kPfRW + 4 * rc] = make_uint4(i6[0], i6[1], i6[2], i6[3]);
        if (r >= NT) reuse1 = bits;
        else reuse0 = bits;
    }
    // the lists: per candidate, a wavefront's runs that do not copy, placed by a ballot and one LDS atomic
    const int lane = tid & 63;
    const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
    for (int slot = 0; slot < (kPfRuns + NT - 1) / NT; slot++) {
        const int r = tid + slot * NT;
        const uint32_t bits = slot ? reuse1 : reuse0;
#pragma unroll
        for (int l = 0; l < kPfCandLists; l++) {
            const int k = l == 0 ? 0 : l + 1;
            const bool forms = r < kPfRuns && !((bits >> k) & 1u);
            const uint64_t m = __ballot(forms);
            if (m == 0ull) continue;   // (uniform)
            uint32_t base = 0u;
            if (lane == 0) base = atomicAdd(&cnt[l], (uint32_t)__popcll(m));
            base = (uint32_t)__shfl((int)base, 0, 64);
            if (forms) lists[l * kPfRuns + base + (uint32_t)__popcll(m & below)] = (uint16_t)r;
        }
    }
    __syncthreads();
    // horizontal 5-sums of both planes: region columns c0 + 2 .. c0 + 9 of a row (three aligned quads), four a task
    auto hsums = [&](const uint32_t* e, uint32_t* h, int t) {
        const int hy = t >> 3, c0 = (t & 7) * 4;
        const uint4* e4 = (const uint4*)&e[hy * kPfRW + c0];
        const uint4 q0 = e4[0], q1 = e4[1], q2 = e4[2];
        uint4 s;
        s.x = q0.z + q0.w + q1.x + q1.y + q1.z;
        s.y = s.x + q1.w - q0.z;
        s.z = s.y + q2.x - q0.w;
        s.w = s.z + q2.y - q1.x;
        *(uint4*)&h[hy * kPfHsStride + c0] = s;
    };
    constexpr int kTasks = kPfRH * (kPfTile / 4);
    for (int t = tid; t < 2 * kTasks; t += NT) {
        if (t < kTasks) hsums(e1, h1, t);
        else hsums(ek, hk, t - kTasks);
    }
    __syncthreads();
    const int cx = tid & 31, r0 = (tid >> 5) * kRows;
    float best[kRows], c6[kRows];
    int bk[kRows];
    {
        uint32_t a[kRows + 4], b[kRows + 4];
#pragma unroll
        for (int y = 0; y < kRows + 4; y++) {
            a[y] = h1[(r0 + y) * kPfHsStride + cx];
            b[y] = hk[(r0 + y) * kPfHsStride + cx];
        }
        uint32_t sa = a[0] + a[1] + a[2] + a[3] + a[4], sb = b[0] + b[1] + b[2] + b[3] + b[4];
#pragma unroll
        for (int j = 0; j < kRows; j++) {
            if (j > 0) {
                sa = sa + a[j + 4] - a[j - 1];
                sb = sb + b[j + 4] - b[j - 1];
            }
            best[j] = (float)((double)sa * cscale);
            bk[j] = 1;
            c6[j] = (float)((double)sb * cscale) - P.zero;
        }
    }
#pragma unroll 1
    for (int step = 0; step < 5; step++) {
        const int k = step == 0 ? 0 : step + 1;
        const int nk = (int)cnt[step];
        if (nk == 0) {   // candidate 1's sums: 0 costs c1 - smoothBonus (best is c1 here), 2-5 cannot win
            if (k == 0) {
#pragma unroll
                for (int j = 0; j < kRows; j++) {
                    const float c = best[j] - P.smooth;
                    if (c <= best[j]) {
                        best[j] = c;
                        bk[j] = 0;
                    }
                }
            }
            continue;
        }
        // the copies, then the listed runs (disjoint runs: no barrier between)
        for (int r = tid; r < kPfRuns; r += NT) {
            const int ry = r / kRunsRow, rc = r - ry * kRunsRow;
            if ((((r >= NT) ? reuse1 : reuse0) >> k) & 1u)
                *(uint4*)&ek[ry * kPfRW + 4 * rc] = *(const uint4*)&e1[ry * kPfRW + 4 * rc];
        }
        for (int li = tid; li < nk; li += NT) {
            const int r = lists[step * kPfRuns + li];
            const int ry = r / kRunsRow, rc = r - ry * kRunsRow;
            uint32_t e[4];
            pixCandRunErrs(P, bw, E, k, x0 - 4 + 4 * rc, clampi(y0 - 2 + ry, 0, P.h - 1), e);
            *(uint4*)&ek[ry * kPfRW + 4 * rc] = make_uint4(e[0], e[1], e[2], e[3]);
        }
        __syncthreads();
        for (int t = tid; t < kTasks; t += NT) hsums(ek, hk, t);
        __syncthreads();
        uint32_t a[kRows + 4];
#pragma unroll
        for (int y = 0; y < kRows + 4; y++) a[y] = hk[(r0 + y) * kPfHsStride + cx];
        uint32_t s = a[0] + a[1] + a[2] + a[3] + a[4];
#pragma unroll
        for (int j = 0; j < kRows; j++) {
            if (j > 0) s = s + a[j + 4] - a[j - 1];
            float c = (float)((double)s * cscale);
            if (k == 0) c = c - P.smooth;
            if (k == 0 ? c <= best[j] : c < best[j]) {
                best[j] = c;
                bk[j] = k;
            }
        }
    }
    const int qx = x0 + cx;
    if (qx >= P.w) return;
#pragma unroll
    for (int j = 0; j < kRows; j++) {
        const int py = y0 + r0 + j;
        if (py >= P.h) break;
        int k = bk[j];
        if (c6[j] < best[j]) k = 6;
        float2 v;
        if (k == 0) {
            v = pixCand(0, qx, py, bw);
        } else if (k == 6) {
            v = make_float2(0.0f, 0.0f);
        } else {
            const int2 o = candOffset(k);
            v = bw.at(qx / kBlock + o.x, py / kBlock + o.y);
        }
        P.pf[py * P.w + qx] = v;
    }
}

constexpr int kPfFloatWork = kPfRH * kPfRW * 4 + kPfRH * kPfTile * 8;   // errs, hsum
constexpr int kPfWork0 = kPfFloatWork > kPfUniPool ? kPfFloatWork : kPfUniPool;
constexpr int kPfWork = kPfWork0 > kPfExactPool ? kPfWork0 : kPfExactPool;
constexpr int kPfFloatPool = kPfWork + 64 * 8 + 32;   // + the block window, the uniform flag and the list lengths

// The LDS windows of a tile at reach kPfWinR, NT threads: a wavefront a row (one texel load a lane, the pair's second
// texel from the next lane), every load of a thread issued before its stores
template <int NT>
__device__ __forceinline__ PixWin pixFillWindows(const PixParams& P, int x0, int y0, uint32_t* win) {
    constexpr int R = kPfWinR, kWaves = NT / 64, kIt = (kPfWinRows + kWaves - 1) / kWaves;
    const PixWin W{win, win + kPfWinWords * kPfWinRows, x0 - 5 - R, y0 - 3 - R, kPfWinWords};
    const float up = __int_as_float((127 + P.fb) << 23);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t gx = (uint32_t)clampi(W.x0 + lane, 0, P.w - 1);
    float a[kIt], c[kIt];
#pragma unroll
    for (int j = 0; j < kIt; j++) {
        const uint32_t row = idx2(0, clampi(W.y0 + min(wave + kWaves * j, kPfWinRows - 1), 0, P.h - 1), P.w);
        a[j] = ldIdx(P.la, row + gx);
        c[j] = ldIdx(P.lc, row + gx);
    }
#pragma unroll
    for (int j = 0; j < kIt; j++) {
        const int wy = wave + kWaves * j;
        const uint32_t ta = (uint32_t)(a[j] * up), tc = (uint32_t)(c[j] * up);
        const uint32_t ta1 = (uint32_t)__shfl_down((int)ta, 1, 64), tc1 = (uint32_t)__shfl_down((int)tc, 1, 64);
        if (wy < kPfWinRows && lane < kPfWinWords) {
            win[wy * kPfWinWords + lane] = ta | (ta1 << 16);
            win[kPfWinWords * kPfWinRows + wy * kPfWinWords + lane] = tc | (tc1 << 16);
        }
    }
    return W;
}

// A 32 x 32 tile of the pixel field. Win: the exact path samples LDS windows of the planes when the tile's candidates
// reach at most kPfWinR flow px (win: kPfWinPool bytes of LDS), else the planes in global memory.
template <int NT, bool Win>
__device__ __forceinline__ void pixfieldTile(const PixParams& P, int gbx, int gby, uint8_t* pool, uint32_t* win) {
    static_assert(NT == 256 || NT == 512, "two layouts");
    static_assert(kPfWinWords + 1 <= 64, "a window row is one wavefront's texels");
    constexpr int kRows = kPfTile * kPfTile / NT;      // pixels a thread: column cx, rows r0 .. r0 + kRows - 1
    double* hsum = (double*)pool;
    float* errs = (float*)(pool + kPfRH * kPfTile * 8);
    float2* bwin = (float2*)(pool + kPfWork);
    int* uni = (int*)(pool + kPfWork + 64 * 8);
    uint32_t* cnt = (uint32_t*)(pool + kPfWork + 64 * 8 + 4);   // [5] list lengths of pixfieldExact
    int* reach = (int*)(pool + kPfWork + 64 * 8 + 24);
    const int tid = threadIdx.x;
    const int x0 = gbx * kPfTile, y0 = gby * kPfTile;
    const BlockWin bw{bwin, x0 / kBlock - 2, y0 / kBlock - 2};
    // the windows (exact lumas) at the fixed reach, their loads issued with the block vectors' (a tile whose candidates
    // reach further samples global memory instead)
    PixWin W{};
    if (Win && P.fb >= 0) W = pixFillWindows<NT>(P, x0, y0, win);
    if (tid < 64) {   // (one wavefront) the block vectors, the uniform test, R = floor(max |component|) + 1
        const float2 v = fieldAt(P.bf, bw.bx0 + (tid & 7), bw.by0 + (tid >> 3), P.bw, P.bh);
        bwin[tid] = v;
        uint32_t vx, vy;
        const bool u = pixWindowUniform(v, tid, vx, vy);
        float m = fmaxf(fabsf(v.x), fabsf(v.y));
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        if (tid == 0) {
            *uni = u && P.fb >= 0;
            *reach = (int)fminf(floorf(m) + 1.0f, (float)(kPfWinR + 1));
        }
        if (tid < kPfCandLists) cnt[tid] = 0u;
    }
    __syncthreads();
    if (P.fb >= 0) {
        const bool inWin = Win && *reach <= kPfWinR;
        const GlobalErrs G{&P, __int_as_float((127 + 10 + P.fb) << 23)};
        if (*uni) {   // (uniform over the workgroup)
            if (inWin) pixfieldUniform<NT>(P, x0, y0, bwin[9], (uint32_t*)pool, WinErrs{W, P.w});
            else pixfieldUniform<NT>(P, x0, y0, bwin[9], (uint32_t*)pool, G);
        } else {
            if (inWin) pixfieldExact<NT>(P, x0, y0, bw, pool, cnt, WinErrs{W, P.w});
            else pixfieldExact<NT>(P, x0, y0, bw, pool, cnt, G);
        }
        return;
    }
    const int cx = tid & 31, r0 = (tid >> 5) * kRows;
    float best[kRows], c1[kRows];
    int bk[kRows];
#pragma unroll
    for (int j = 0; j < kRows; j++) {
        best[j] = 0.0f;
        c1[j] = 0.0f;
        bk[j] = 0;
    }
    // candidate 1's errors of this thread's runs (r = tid and tid + NT), kept for candidates 0 and 2-5
    float own0[4] = {0.0f, 0.0f, 0.0f, 0.0f}, own1[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float2 ownV0 = make_float2(0.0f, 0.0f), ownV1 = make_float2(0.0f, 0.0f);
    bool ownOk0 = false, ownOk1 = false;
#pragma unroll 1
    for (int step = 0; step < 7; step++) {
        const int k = step == 0 ? 1 : (step == 1 ? 0 : step);
        bool reusedAll = true;
#pragma unroll 1
        for (int r = tid; r < kPfRuns; r += NT) {
            const bool s1 = r >= NT;   // one code path for both runs; the kept values are selected, not indexed
            float own[4];
#pragma unroll
            for (int i = 0; i < 4; i++) own[i] = s1 ? own1[i] : own0[i];
            float2 ownV = s1 ? ownV1 : ownV0;
            bool ownOk = s1 ? ownOk1 : ownOk0;
            const int ry = r / (kPfRW / 4), rc = r - ry * (kPfRW / 4);
            float e[4];
            reusedAll = pixErrRun(P, bw, k, x0 - 4 + 4 * rc, clampi(y0 - 2 + ry, 0, P.h - 1), e, own, ownV, ownOk) &&
                        reusedAll;
            *(float4*)&errs[ry * kPfRW + 4 * rc] = make_float4(e[0], e[1], e[2], e[3]);
            if (k == 1) {
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    own0[i] = s1 ? own0[i] : own[i];
                    own1[i] = s1 ? own[i] : own1[i];
                }
                ownV0 = s1 ? ownV0 : ownV;
                ownV1 = s1 ? ownV : ownV1;
                ownOk0 = s1 ? ownOk0 : ownOk;
                ownOk1 = s1 ? ownOk : ownOk1;
            }
        }
        // a barrier (errs complete) and the vote: only candidates 0 and 2-5 ever reuse, so k = 1 and 6 never skip
        if (__syncthreads_and(reusedAll)) {
            if (k == 0) {
#pragma unroll
                for (int j = 0; j < kRows; j++) {
                    const float c = c1[j] - P.smooth;
                    if (c <= best[j]) {
                        best[j] = c;
                        bk[j] = 0;
                    }
                }
            }
            continue;
        }
        // horizontal 5-sums, 4 output columns a task: region columns c0 + 2 .. c0 + 10 of one row (read as three
        // aligned 16-byte quads; a row's 8 tasks cover every LDS bank once)
        for (int t = tid; t < kPfRH * (kPfTile / 4); t += NT) {
            const int hy = t >> 3, c0 = (t & 7) * 4;
            const float4* e4 = (const float4*)&errs[hy * kPfRW + c0];
            const float4 q0 = e4[0], q1 = e4[1], q2 = e4[2];
            const float e[8] = {q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y};
            double h[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                double hs = (double)e[i];
#pragma unroll
                for (int x = 1; x < 5; x++) hs = hs + (double)e[i + x];
                h[i] = hs;
            }
            *(double2*)&hsum[hsumAt(hy, c0)] = make_double2(h[0], h[1]);
            *(double2*)&hsum[hsumAt(hy, c0 + 2)] = make_double2(h[2], h[3]);
        }
        __syncthreads();
        double hv[kRows + 4];
#pragma unroll
        for (int y = 0; y < kRows + 4; y++) hv[y] = hsum[hsumAt(r0 + y, cx)];
#pragma unroll
        for (int j = 0; j < kRows; j++) {
            double vs = hv[j];
#pragma unroll
            for (int y = 1; y < 5; y++) vs = vs + hv[j + y];
            float c = (float)(vs * (1.0 / 25.0));
            if (k == 0) c = c - P.smooth;
            if (k == 6) c = c - P.zero;
            if (k == 1) {
                c1[j] = c;
                best[j] = c;
                bk[j] = 1;
            } else if (k == 0 ? c <= best[j] : c < best[j]) {
                best[j] = c;
                bk[j] = k;
            }
        }
    }
    const int qx = x0 + cx;
    if (qx >= P.w) return;
#pragma unroll
    for (int j = 0; j < kRows; j++) {
        const int py = y0 + r0 + j;
        if (py >= P.h) break;
        const int k = bk[j];
        float2 v;
        if (k == 0) {
            v = pixCand(0, qx, py, bw);
        } else {
            const int2 o = candOffset(k);
            v = bw.at(qx / kBlock + o.x, py / kBlock + o.y);
            if (k == 6) v = make_float2(0.0f, 0.0f);
        }
        P.pf[py * P.w + qx] = v;
    }
}

template <int NT, bool Win>
__device__ __forceinline__ void pixfieldBody(const PixParams& P, int gbx, int gby) {
    __shared__ __align__(16) uint8_t pool[kPfFloatPool];
    __shared__ __align__(16) uint32_t win[Win ? kPfWinPool / 4 : 1];
    pixfieldTile<NT, Win>(P, gbx, gby, pool, win);
}

// ---------------------------------------------------------------------------------------------------------- kernels

// Each pass has a per-pair kernel (grid = that pair's blocks) and a batched kernel over a work list spanning many pairs
// (one launch per pass for a whole batch); both run the same body.

FGR_PASS(luma_rgba8, LumaParams, lumaRgba8Body, 256)
FGR_PASS(pyramid_rgba8, PyrParams, pyramidRgba8Body, 256)
FGR_PASS(down, DownParams, downBody, 256)
FGR_PASS(medhist, MedParams, medhistBody, 1024)
FGR_PASS(search2, SearchParams, searchBody<2>, 256)
FGR_PASS(search2x8, SearchParams, search2Body, 256)
FGR_PASS(search6, SearchParams, searchBody<6>, 256)
FGR_PASS(search12, SearchParams, searchBody<12>, 256)
FGR_PASS(median3, Median3Params, median3Body, 256)
FGR_PASS(subpix, SubpixParams, subpixBody, 64)
FGR_PASS(subpix_int, SubpixParams, subpixIntBody, 64 * kSubpixIntWaves)
#define FGR_PIX_BODY(NT, WIN) pixfieldBody<NT, WIN>
FGR_PASS_SINGLE(pixfield, PixParams, FGR_PIX_BODY(kPfThreads, false), kPfThreads)
FGR_PASS_BATCH_OCC(pixfield, PixParams, FGR_PIX_BODY(kPfBatchThreads, false), kPfBatchThreads, 6)
FGR_PASS_SINGLE(pixfield_win, PixParams, FGR_PIX_BODY(kPfThreads, true), kPfThreads)
FGR_PASS_BATCH_OCC(pixfield_win, PixParams, FGR_PIX_BODY(kPfThreads, true), kPfThreads, 6)

// ---------------------------------------------------------------------------------------------------------- launches

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// one plan op of one pair: grid gx x gy of the per-pair kernel
hipError_t launchSingle(int pass, const void* params, int gx, int gy, hipStream_t s) {
    const dim3 g((unsigned)gx, (unsigned)gy);
    switch (pass) {
    case kPassLuma: hipLaunchKernelGGL(luma_rgba8_kernel, g, dim3(256), 0, s, *(const LumaParams*)params); break;
    case kPassPyramid: hipLaunchKernelGGL(pyramid_rgba8_kernel, g, dim3(256), 0, s, *(const PyrParams*)params); break;
    case kPassDown: hipLaunchKernelGGL(down_kernel, g, dim3(256), 0, s, *(const DownParams*)params); break;
    case kPassMedhist: hipLaunchKernelGGL(medhist_kernel, g, dim3(1024), 0, s, *(const MedParams*)params); break;
    case kPassSearch2: hipLaunchKernelGGL(search2_kernel, g, dim3(256), 0, s, *(const SearchParams*)params); break;
    case kPassSearch2x8: hipLaunchKernelGGL(search2x8_kernel, g, dim3(256), 0, s, *(const SearchParams*)params); break;
    case kPassSearch6: hipLaunchKernelGGL(search6_kernel, g, dim3(256), 0, s, *(const SearchParams*)params); break;
    case kPassSearch12: hipLaunchKernelGGL(search12_kernel, g, dim3(256), 0, s, *(const SearchParams*)params); break;
    case kPassMedian3: hipLaunchKernelGGL(median3_kernel, g, dim3(256), 0, s, *(const Median3Params*)params); break;
    case kPassSubpix: hipLaunchKernelGGL(subpix_kernel, g, dim3(64), 0, s, *(const SubpixParams*)params); break;
    case kPassSubpixInt: hipLaunchKernelGGL(subpix_int_kernel, g, dim3(64 * kSubpixIntWaves), 0, s, *(const SubpixParams*)params); break;
    case kPassPixfield: hipLaunchKernelGGL(pixfield_kernel, g, dim3(kPfThreads), 0, s, *(const PixParams*)params); break;
    case kPassPixfieldWin: hipLaunchKernelGGL(pixfield_win_kernel, g, dim3(kPfThreads), 0, s, *(const PixParams*)params); break;
    default: return launchSynthSingle(pass, params, gx, gy, s);
    }
    return hipGetLastError();
}

// batched launches: n work items, one workgroup each
hipError_t launchBatch(int pass, const void* params, const uint32_t* work, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const dim3 g((unsigned)n);
    switch (pass) {
    case kPassLuma: hipLaunchKernelGGL(luma_rgba8_batch_kernel, g, dim3(256), 0, s, (const LumaParams*)params, work); break;
    case kPassPyramid: hipLaunchKernelGGL(pyramid_rgba8_batch_kernel, g, dim3(256), 0, s, (const PyrParams*)params, work); break;
    case kPassDown: hipLaunchKernelGGL(down_batch_kernel, g, dim3(256), 0, s, (const DownParams*)params, work); break;
    case kPassMedhist: hipLaunchKernelGGL(medhist_batch_kernel, g, dim3(1024), 0, s, (const MedParams*)params, work); break;
    case kPassSearch2: hipLaunchKernelGGL(search2_batch_kernel, g, dim3(256), 0, s, (const SearchParams*)params, work); break;
    case kPassSearch2x8: hipLaunchKernelGGL(search2x8_batch_kernel, g, dim3(256), 0, s, (const SearchParams*)params, work); break;
    case kPassSearch6: hipLaunchKernelGGL(search6_batch_kernel, g, dim3(256), 0, s, (const SearchParams*)params, work); break;
    case kPassSearch12: hipLaunchKernelGGL(search12_batch_kernel, g, dim3(256), 0, s, (const SearchParams*)params, work); break;
    case kPassMedian3: hipLaunchKernelGGL(median3_batch_kernel, g, dim3(256), 0, s, (const Median3Params*)params, work); break;
    case kPassSubpix: hipLaunchKernelGGL(subpix_batch_kernel, g, dim3(64), 0, s, (const SubpixParams*)params, work); break;
    case kPassSubpixInt: hipLaunchKernelGGL(subpix_int_batch_kernel, g, dim3(64 * kSubpixIntWaves), 0, s, (const SubpixParams*)params, work); break;
    case kPassPixfield: hipLaunchKernelGGL(pixfield_batch_kernel, g, dim3(kPfBatchThreads), 0, s, (const PixParams*)params, work); break;
    case kPassPixfieldWin: hipLaunchKernelGGL(pixfield_win_batch_kernel, g, dim3(kPfThreads), 0, s, (const PixParams*)params, work); break;
    default: return launchSynthBatch(pass, params, work, n, s);
    }
    return hipGetLastError();
}

}  // namespace fgr
