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
#include "v1_synth_common.hpp"

namespace fgr {

// cv2.remap INTER_LINEAR / BORDER_REPLICATE of an 8-bit RGB frame (interp.py:214), at the map position already scaled
// by 32 (X = 32 px, Y = 32 py, float): map rounded to 1/32 px (half even), integer weights (32-fx)(32-fy), fx(32-fy),
// (32-fx)fy, fx fy, result (sum + 512) >> 10 -- the exact value rounded half up, as cv2's fixed-point remap. Each tap
// row is one 8-byte load at the clamped column (W >= 2): the two taps are (x, x+1), or the edge texel twice.
struct Rgb {
    uint32_t r, g, b;
};

__device__ __forceinline__ Rgb warpRgba8At(const uint8_t* base, uint32_t pitch, int W, int H, int X, int Y) {
    const int ix = X >> 5, iy = Y >> 5;
    const uint32_t fx = (uint32_t)(X & 31), fy = (uint32_t)(Y & 31);
    const uint32_t y0 = (uint32_t)med3i(iy, 0, H - 1), y1 = (uint32_t)med3i(iy + 1, 0, H - 1);
    const uint32_t xo = 4u * (uint32_t)med3i(ix, 0, W - 2);
    const U32x2 q0 = *(const U32x2*)(base + (__umul24(y0, pitch) + xo));
    const U32x2 q1 = *(const U32x2*)(base + (__umul24(y1, pitch) + xo));
    const bool right = ix >= W - 1, left = ix < 0;
    const uint32_t p00 = right ? q0.y : q0.x, p01 = left ? q0.x : q0.y;
    const uint32_t p10 = right ? q1.y : q1.x, p11 = left ? q1.x : q1.y;
    const uint32_t gx = 32u - fx, gy = 32u - fy;
    const uint32_t w00 = gx * gy, w01 = fx * gy, w10 = gx * fy, w11 = fx * fy;
    Rgb o;
    o.r = (w00 * (p00 & 255u) + w01 * (p01 & 255u) + w10 * (p10 & 255u) + w11 * (p11 & 255u) + 512u) >> 10;
    o.g = (w00 * ((p00 >> 8) & 255u) + w01 * ((p01 >> 8) & 255u) + w10 * ((p10 >> 8) & 255u) +
           w11 * ((p11 >> 8) & 255u) + 512u) >> 10;
    o.b = (w00 * ((p00 >> 16) & 255u) + w01 * ((p01 >> 16) & 255u) + w10 * ((p10 >> 16) & 255u) +
           w11 * ((p11 >> 16) & 255u) + 512u) >> 10;
    return o;
}

__device__ __forceinline__ Rgb warpRgba8(const uint8_t* base, uint32_t pitch, int W, int H, float X32, float Y32) {
    return warpRgba8At(base, pitch, W, H, (int)rintf(X32), (int)rintf(Y32));
}

__device__ __forceinline__ uint32_t packRgb(const Rgb& v) { return v.r | (v.g << 8) | (v.b << 16); }

// Four horizontally adjacent pixels (x .. x+3, one row) warped with one interpolated field value r. When their maps
// X_i = rint(32 (x + i) - k r.x) are 32 apart (same fractional part) and no tap column needs clamping, the four share
// Y, the weights and 5 consecutive texels per tap row, and each result equals warpRgba8 of that pixel.
// The maps are 32 apart whenever d_0 = fl(32 x - k r.x) and d_3 = fl(32 (x + 3) - k r.x) have one sign and one exponent:
// then the exact differences D_i = D_0 + 32 i lie in one binade, whose float grid (spacing <= 32 below 2^29) holds
// 32 i, so fl(D_i) = fl(D_0) + 32 i, and rint (ties to even) commutes with adding the even integer 32 i.
struct RunMap {
    int X0, Y;
    bool ok;
};

__device__ __forceinline__ RunMap runMap(int W, int x, int y, float k, float2 r) {
    const float kx = k * r.x;
    RunMap m;
    const float d0 = (float)(32 * x) - kx, d3 = (float)(32 * (x + 3)) - kx;
    m.X0 = (int)rintf(d0);
    const int ix = m.X0 >> 5;
    m.ok = (__float_as_uint(d0) ^ __float_as_uint(d3)) < 0x00800000u && ix >= 0 && ix + 4 <= W - 1;
    m.Y = (int)rintf((float)(32 * y) - k * r.y);
    return m;
}

// Two 16-bit lanes of one register added lane by lane (no carry between lanes; every value here < 2^16), and
// a.lo b.lo + a.hi b.hi + c in 32 bits
typedef unsigned short U16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pkAdd16(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(U16x2, a) + __builtin_bit_cast(U16x2, b));
}
__device__ __forceinline__ uint32_t dot2u16(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_udot2(__builtin_bit_cast(U16x2, a), __builtin_bit_cast(U16x2, b), c, false);
}

// The four warps of a run whose pixels share the map row Y and take the tap columns X_i >> 5 = ix + i (in the frame:
// ix >= 0, ix + 4 <= W - 1), each with its own column weight fx_i = X_i & 31 -- a run whose map passed runMap (ok) is
// the case X_i = X_0 + 32 i. The run shares fy and 5 consecutive texels a row, so each texel column is blended
// vertically once per channel: the column's top and bottom bytes interleaved by one byte permute ([R, R', G, G'] and
// [B, B', 0, 0]) and dotted with the bytes [32 - fy, fy] (v_dot4_u32_u8), V = (32 - fy) top + fy bottom <= 8160.
// Each pixel then takes one dot product per channel over its two columns' V with its horizontal weights x 64 and
// 64 x 512 added: 64 ((32-fx)(32-fy) p00 + fx(32-fy) p01 + (32-fx)fy p10 + fx fy p11 + 512) < 2^24, whose byte 2 is
// warpRgba8's (sum + 512) >> 10 -- the same integers, regrouped.
__device__ __forceinline__ void warpRow4(const uint8_t* base, uint32_t pitch, int H, int ix, int Y, const uint32_t* fx,
                                         uint32_t* out) {
    const int iy = Y >> 5;
    const uint32_t fy = (uint32_t)(Y & 31);
    const uint32_t y0 = (uint32_t)med3i(iy, 0, H - 1), y1 = (uint32_t)med3i(iy + 1, 0, H - 1);
    const uint8_t* r0 = base + (__umul24(y0, pitch) + 4u * (uint32_t)ix);
    const uint8_t* r1 = base + (__umul24(y1, pitch) + 4u * (uint32_t)ix);
    const U32x4 a4 = *(const U32x4*)r0;
    const uint32_t a5 = *(const uint32_t*)(r0 + 16);
    const U32x4 b4 = *(const U32x4*)r1;
    const uint32_t b5 = *(const uint32_t*)(r1 + 16);
    const uint32_t ta[5] = {a4.x, a4.y, a4.z, a4.w, a5}, tb[5] = {b4.x, b4.y, b4.z, b4.w, b5};
    const uint32_t wv = (32u - fy) | (fy << 8);   // bytes [32 - fy, fy, 0, 0]
    uint32_t vR[5], vG[5], vB[5];
#pragma unroll
    for (int j = 0; j < 5; j++) {
        const uint32_t rg = __builtin_amdgcn_perm(tb[j], ta[j], 0x05010400u);   // [top R, bottom R, top G, bottom G]
        const uint32_t bb = __builtin_amdgcn_perm(tb[j], ta[j], 0x0C0C0602u);   // [top B, bottom B, 0, 0]
        vR[j] = __builtin_amdgcn_udot4(rg, wv, 0u, false);
        vG[j] = __builtin_amdgcn_udot4(rg, wv << 16, 0u, false);
        vB[j] = __builtin_amdgcn_udot4(bb, wv, 0u, false);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t hw = ((32u - fx[i]) << 6) | (fx[i] << 22);   // [64 (32 - fx), 64 fx]
        // [V(i), V(i + 1)] of each channel as two 16-bit lanes
        const uint32_t sr = dot2u16(__builtin_amdgcn_perm(vR[i + 1], vR[i], 0x05040100u), hw, 32768u);
        const uint32_t sg = dot2u16(__builtin_amdgcn_perm(vG[i + 1], vG[i], 0x05040100u), hw, 32768u);
        const uint32_t sb = dot2u16(__builtin_amdgcn_perm(vB[i + 1], vB[i], 0x05040100u), hw, 32768u);
        out[i] = __builtin_amdgcn_perm(sg, sr, 0x0C0C0602u) | (sb & 0x00FF0000u);   // bytes R, G, B, alpha 0
    }
}

// warpRgba8 of one pixel whose tap columns ix, ix + 1 lie in the frame (0 <= ix <= W - 2), as warpRow4 forms it: each
// tap column blended vertically per channel by byte dot products, then one 16-bit pair dot product per channel
__device__ __forceinline__ uint32_t warpPix(const uint8_t* base, uint32_t pitch, int H, int X, int Y) {
    const int ix = X >> 5, iy = Y >> 5;
    const uint32_t fx = (uint32_t)(X & 31), fy = (uint32_t)(Y & 31);
    const uint32_t y0 = (uint32_t)med3i(iy, 0, H - 1), y1 = (uint32_t)med3i(iy + 1, 0, H - 1);
    const U32x2 q0 = *(const U32x2*)(base + (__umul24(y0, pitch) + 4u * (uint32_t)ix));
    const U32x2 q1 = *(const U32x2*)(base + (__umul24(y1, pitch) + 4u * (uint32_t)ix));
    const uint32_t wv = (32u - fy) | (fy << 8);
    const uint32_t hw = ((32u - fx) << 6) | (fx << 22);
    const uint32_t rg0 = __builtin_amdgcn_perm(q1.x, q0.x, 0x05010400u), bb0 = __builtin_amdgcn_perm(q1.x, q0.x, 0x0C0C0602u);
    const uint32_t rg1 = __builtin_amdgcn_perm(q1.y, q0.y, 0x05010400u), bb1 = __builtin_amdgcn_perm(q1.y, q0.y, 0x0C0C0602u);
    const uint32_t r0 = __builtin_amdgcn_udot4(rg0, wv, 0u, false), r1 = __builtin_amdgcn_udot4(rg1, wv, 0u, false);
    const uint32_t g0 = __builtin_amdgcn_udot4(rg0, wv << 16, 0u, false), g1 = __builtin_amdgcn_udot4(rg1, wv << 16, 0u, false);
    const uint32_t b0 = __builtin_amdgcn_udot4(bb0, wv, 0u, false), b1 = __builtin_amdgcn_udot4(bb1, wv, 0u, false);
    const uint32_t sr = dot2u16(__builtin_amdgcn_perm(r1, r0, 0x05040100u), hw, 32768u);
    const uint32_t sg = dot2u16(__builtin_amdgcn_perm(g1, g0, 0x05040100u), hw, 32768u);
    const uint32_t sb = dot2u16(__builtin_amdgcn_perm(b1, b0, 0x05040100u), hw, 32768u);
    return __builtin_amdgcn_perm(sg, sr, 0x0C0C0602u) | (sb & 0x00FF0000u);
}

// a run whose map passed runMap: one column weight for the four
__device__ __forceinline__ void warpRun4(const uint8_t* base, uint32_t pitch, int H, const RunMap& m, uint32_t* out) {
    const uint32_t f = (uint32_t)(m.X0 & 31);
    const uint32_t fx[4] = {f, f, f, f};
    warpRow4(base, pitch, H, m.X0 >> 5, m.Y, fx, out);
}

// the same from the packed bytes: max |d| over the three channel differences d = max(max d, -min d)
__device__ __forceinline__ int emaxPacked(uint32_t a, uint32_t c) {
    const int dr = (int)(a & 255u) - (int)(c & 255u), dg = (int)((a >> 8) & 255u) - (int)((c >> 8) & 255u),
              db = (int)((a >> 16) & 255u) - (int)((c >> 16) & 255u);
    return max(max(max(dr, dg), db), -min(min(dr, dg), db));
}

// error.comp:9-21 / mismatch() motion_v1.glsl:117-129: e = max-channel |wa - wc| (integers 0..255)
__device__ __forceinline__ int emaxRgb(const Rgb& a, const Rgb& c) {
    return (int)max(max(sadU32(a.r, c.r, 0u), sadU32(a.g, c.g, 0u)), sadU32(a.b, c.b, 0u));
}

__device__ __forceinline__ int emaxRgba8(uint32_t wa, uint32_t wc) {
    const int d0 = abs(byteOf(wa, 0) - byteOf(wc, 0)), d1 = abs(byteOf(wa, 1) - byteOf(wc, 1)),
              d2 = abs(byteOf(wa, 2) - byteOf(wc, 2));
    return max(d0, max(d1, d2));
}

// the same pair from the plain per-pixel halfVec (debug pass)
__device__ __forceinline__ void warpPairRgba8(const SynthParams& P, int x, int y, float2 hv, uint32_t& wa, uint32_t& wc) {
    const float ka = 2.0f * P.wC, kc = 2.0f * P.wA;
    const float xa = 32.0f * ((float)x - ka * hv.x), ya = 32.0f * ((float)y - ka * hv.y);
    const float xc = 32.0f * ((float)x + kc * hv.x), yc = 32.0f * ((float)y + kc * hv.y);
    wa = packRgb(warpRgba8(P.a.p, (uint32_t)P.a.pitch, P.W, P.H, xa, ya));
    wc = packRgb(warpRgba8(P.c.p, (uint32_t)P.c.pitch, P.W, P.H, xc, yc));
    if (P.uiA.w > 0 || P.uiC.w > 0) uiReject(P, (int)rintf(xa), (int)rintf(ya), (int)rintf(xc), (int)rintf(yc), wa, wc);
}

// The trusted quarter-t blend (kA a + kC c + 2) >> 2 of every byte, (kA, kC) = (3, 1) or (1, 3): with the per-byte
// floor mean u = (a + c) >> 1 (v_lerp_u8, rounding bits 0), 3a + c + 2 = 2 (a + u + 1) + ((a + c) & 1), whose quotient
// by 4 rounded down is (a + u + 1) >> 1: a second v_lerp_u8 with rounding bits 1 (a + 3c the same with c). Alpha
// bytes of wa and wc are 0, so the result's alpha byte is 0; other (kA, kC): per channel dot products.
__device__ __forceinline__ uint32_t quarterRgba8(const SynthParams& P, uint32_t wa, uint32_t wc) {
    if (P.kA == 3 || P.kA == 1) {
        const uint32_t u = __builtin_amdgcn_lerp(wa, wc, 0u);
        return __builtin_amdgcn_lerp(P.kA == 3 ? wa : wc, u, 0x01010101u);
    }
    const uint32_t k = (uint32_t)P.kA | ((uint32_t)P.kC << 8);
    uint32_t r = 0u;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const uint32_t ac = __builtin_amdgcn_perm(wc, wa, 0x0C0C0400u + 0x0101u * (uint32_t)c);   // [a_c, c_c, 0, 0]
        r |= (__builtin_amdgcn_udot4(ac, k, 2u, false) >> 2) << (8 * c);
    }
    return r;
}

// resolve.comp:73-82 (interp.py:241-246): m = 0.5 (wa + wc) at t = 0.5, else (1-t) wa + t wc; o = m (1 - u) + near u;
// out = clamp(floor(o + 0.5), 0, 255); alpha from the nearer frame. Two exact shortcuts of the same float32 formula:
// u = 1 gives o = near; u = 0 gives o = m, and when 4(1-t) and 4t are integers m + 0.5 is exact, so
// floor(m + 0.5) = (kA a + kC c + 2) >> 2 with kA = 4(1-t), kC = 4t.
__device__ __forceinline__ uint32_t blendRgba8(const SynthParams& P, uint32_t wa, uint32_t wc, uint32_t nr, float u) {
    if (u == 1.0f) return nr;
    if (u == 0.0f && P.mid)   // (a + c + 1) >> 1 per byte (v_lerp_u8), alpha from the nearer frame
        return (__builtin_amdgcn_lerp(wa, wc, 0x01010101u) & 0x00FFFFFFu) | (nr & 0xFF000000u);
    if (u == 0.0f && P.quarter) return quarterRgba8(P, wa, wc) | (nr & 0xFF000000u);
    uint32_t r = nr & 0xFF000000u;
    const float omu = 1.0f - u;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int a8 = byteOf(wa, c), c8 = byteOf(wc, c), n8 = byteOf(nr, c);
        float m;
        if (P.mid) m = 0.5f * (float)(a8 + c8);
        else m = P.wA * (float)a8 + P.wC * (float)c8;
        const float o = m * omu + (float)n8 * u;
        const int v = med3i((int)floorf(o + 0.5f), 0, 255);
        r |= (uint32_t)v << (8 * c);
    }
    return r;
}

// The warp pairs of the 4 region pixels at columns ex0 .. ex0 + 3 (frame columns x0 - AP + ex0 + i, clamped) of frame
// row y (already clamped), from the interpolated field rows hrow. Each form is taken when every lane of the wavefront
// can take it (no divergent second path): one map for the run (warpRun4), else one map row and consecutive tap columns
// (warpRow4), else pixel by pixel (warpPix with every tap column inside the frame, else warpRgba8At with its edge
// clamps); every form gives the same values.
template <int S, bool UI>
__device__ __forceinline__ void warpRegionRun(const SynthParams& P, const float2* hrow, int ty0, float kA32, float kC32,
                                              int x0, int ex0, int y, uint32_t* oa, uint32_t* oc) {
    const int yc = min(y, S * P.fh - 1);
    const float2* h = hrow + (FlowMap<S>::index(yc) - ty0) * EW + ex0;
    const float fy = FlowMap<S>::frac(yc);
    int xs[4];
    float2 ht[4], hb[4];
    bool hsame = true;   // the run's two interpolated rows hold one value each, bit for bit: one field value
#pragma unroll
    for (int i = 0; i < 4; i++) {
        xs[i] = med3i(x0 - AP + ex0 + i, 0, P.W - 1);
        // a clamped column repeats the edge pixel: its hrow entry is that pixel's (clamped the same way)
        ht[i] = h[i];
        hb[i] = h[i + EW];
        if (i > 0)   // (no short-circuit: every load issued before the first compare)
            hsame = hsame & (__builtin_bit_cast(uint64_t, ht[i]) == __builtin_bit_cast(uint64_t, ht[0])) &
                    (__builtin_bit_cast(uint64_t, hb[i]) == __builtin_bit_cast(uint64_t, hb[0]));
    }
    // the run's field value; the others are formed (and compared) only when the rows are not one value each
    const float2 r0 = mix2(ht[0], hb[0], fy);
    bool same = xs[3] == xs[0] + 3;
    if (!hsame) {
#pragma unroll
        for (int i = 1; i < 4; i++) {
            const float2 ri = mix2(ht[i], hb[i], fy);
            same = same && ri.x == r0.x && ri.y == r0.y;
        }
    }
    const RunMap ma = runMap(P.W, xs[0], y, kA32, r0), mc = runMap(P.W, xs[0], y, -kC32, r0);
    const bool fast = same && ma.ok && mc.ok;
    if (__all(fast)) {
        warpRun4(P.a.p, (uint32_t)P.a.pitch, P.H, ma, oa);
        warpRun4(P.c.p, (uint32_t)P.c.pitch, P.H, mc, oc);
        if (UI) {   // the run's maps are X0 + 32 i at one Y
#pragma unroll
            for (int i = 0; i < 4; i++) uiReject(P, ma.X0 + 32 * i, ma.Y, mc.X0 + 32 * i, mc.Y, oa[i], oc[i]);
        }
        return;
    }
    // Each pixel's maps: from r, the interpolated field before the x S of halfVec (hv = S r), a at p - 2t hv and c at
    // p + 2(1-t) hv, each the float32 product then the difference (motion_v1.glsl:120-122), then x 32 for the map;
    // scaling by a power of two commutes with rounding, so 32 (p - ka (S r)) = 32 p - (32 S ka) r exactly (kA32 and
    // kC32 are those folded constants). UI: with the CrosshairV2 tap rejection (motion_v1.glsl:124). A run whose pixels
    // share the map row and take consecutive tap columns (a field that varies by less than a pixel along the run)
    // shares its texels and vertical blends (warpRow4)
    int Xa[4], Xc[4], Ya[4], Yc[4];
    bool row = xs[3] == xs[0] + 3;
    const float y32 = (float)(32 * y);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float2 ri = mix2(ht[i], hb[i], fy);
        const float x32 = (float)(32 * xs[i]);
        Xa[i] = (int)rintf(x32 - kA32 * ri.x);
        Xc[i] = (int)rintf(x32 + kC32 * ri.x);
        Ya[i] = (int)rintf(y32 - kA32 * ri.y);
        Yc[i] = (int)rintf(y32 + kC32 * ri.y);
        row = row & (Ya[i] == Ya[0]) & (Yc[i] == Yc[0]) & ((Xa[i] >> 5) == (Xa[0] >> 5) + i) &
              ((Xc[i] >> 5) == (Xc[0] >> 5) + i);
    }
    row = row & (Xa[0] >= 0) & ((Xa[0] >> 5) + 4 <= P.W - 1) & (Xc[0] >= 0) & ((Xc[0] >> 5) + 4 <= P.W - 1);
    if (__all(row)) {
        uint32_t fa[4], fc[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            fa[i] = (uint32_t)(Xa[i] & 31);
            fc[i] = (uint32_t)(Xc[i] & 31);
        }
        warpRow4(P.a.p, (uint32_t)P.a.pitch, P.H, Xa[0] >> 5, Ya[0], fa, oa);
        warpRow4(P.c.p, (uint32_t)P.c.pitch, P.H, Xc[0] >> 5, Yc[0], fc, oc);
    } else {
        // pixel by pixel: the byte-dot-product form when every tap column of the wavefront lies in the frame
        bool inside = true;
#pragma unroll
        for (int i = 0; i < 4; i++)
            inside = inside & ((uint32_t)(Xa[i] >> 5) <= (uint32_t)(P.W - 2)) & ((uint32_t)(Xc[i] >> 5) <= (uint32_t)(P.W - 2));
        if (__all(inside)) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                oa[i] = warpPix(P.a.p, (uint32_t)P.a.pitch, P.H, Xa[i], Ya[i]);
                oc[i] = warpPix(P.c.p, (uint32_t)P.c.pitch, P.H, Xc[i], Yc[i]);
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                oa[i] = packRgb(warpRgba8At(P.a.p, (uint32_t)P.a.pitch, P.W, P.H, Xa[i], Ya[i]));
                oc[i] = packRgb(warpRgba8At(P.c.p, (uint32_t)P.c.pitch, P.W, P.H, Xc[i], Yc[i]));
            }
        }
    }
    if (UI) {
#pragma unroll
        for (int i = 0; i < 4; i++) uiReject(P, Xa[i], Ya[i], Xc[i], Yc[i], oa[i], oc[i]);
    }
}

// One 64 x RT tile (RT = 16, 32 or 64 rows; 4, 8 or 16 pixels a thread) of the synthesised frame (passes 9-11 fused):
// field tile (+ lookup at t), mismatch of the tile and its apron, 7x7 box, trust, untrusted count, blend. Scene cut: a
// copy of the nearer frame.
//
// halfVec() motion_v1.glsl:73-84 is split in its two stages: the horizontal mixes of each field row at every full-res
// column of the tile (row[j][x] = mix(F[j][i0(x)], F[j][i1(x)], fx(x))), then per pixel the vertical mix of two of
// those rows; the operations and their order are the shader's, so the result is the same float.
template <int S, int RT, bool UI = false>
__device__ __forceinline__ void synthRgba8Tile(const SynthParams& P, int tileX, int tileY, int nTilesX) {
    constexpr int FTW = SynthTile<S, RT>::FTW, FTH = SynthTile<S, RT>::FTH;
    constexpr int TH = RT, EH = TileRows<RT>::EH, RK = RT / 16;
    // LDS in two regions, each holding two arrays whose lifetimes do not overlap: the field tile ft (until the rows
    // are interpolated) then the mismatch es (from the warps to the box), and the interpolated rows hrow (until the
    // warps are done) then the box's row sums hs; a barrier separates each pair
    constexpr int kRegA = ((FTW * FTH * 8 > EH * EWB ? FTW * FTH * 8 : EH * EWB) + 15) / 16 * 16;
    constexpr int kRegB = FTH * EW * 8 > EH * TW * 2 ? FTH * EW * 8 : EH * TW * 2;
    __shared__ __align__(16) uint8_t pool[kRegA + kRegB];
    float2* ft = (float2*)pool;
    uint8_t* es = pool;   // the mismatch, one byte a pixel, rows EWB bytes apart
    float2* hrow = (float2*)(pool + kRegA);
    uint16_t* hs = (uint16_t*)(pool + kRegA);
    __shared__ uint32_t cnt;
    const int tid = threadIdx.x;
    const int x0 = tileX * TW, y0 = tileY * TH;
    const FrameRef& nf = P.nearC ? P.c : P.a;
    const int cx = tid & (TW - 1), ry = tid >> 6;   // this thread's output column and first row (rows ry + 4k)
    if (sceneCut(P)) {   // uniform over the grid: the whole output is the nearer real frame (interp.py:225-226)
        for (int k = 0; k < TH / 4; k++) {
            const int x = x0 + cx, y = y0 + ry + 4 * k;
            if (x < P.W && y < P.H) *(uint32_t*)(P.out + (int64_t)y * P.opitch + 4 * (int64_t)x) = ld32(nf, x, y);
        }
        return;
    }
    if (tid == 0) cnt = 0u;
    const int ty0 = loadFieldTile<S, RT>(P, x0, y0, ft, hrow);
    const float kA32 = (float)(32 * S) * (2.0f * P.wC), kC32 = (float)(32 * S) * (2.0f * P.wA);
    // the apron first (no interior warps held in registers yet), in runs of 4 as well: rows 0..2 and EH-3..EH-1 as 18
    // runs a row (the last at column EW - 4, overlapping its neighbour), then columns 0..3 and EW-4..EW-1 of each row
    // between (overlapping the interior's first and last column). A pixel taken twice gets the same value twice.
    constexpr int kRowRuns = (EW + 3) / 4, kApRuns = 2 * AP * kRowRuns + 2 * TH;
    static_assert(kRowRuns == 18 && 2 * AP * kRowRuns < 256, "one run a thread for the apron rows");
    for (int j = tid; j < kApRuns; j += 256) {
        int ex0, ey;
        if (j < 2 * AP * kRowRuns) {
            const int r = j / kRowRuns, c = j - r * kRowRuns;
            ey = r < AP ? r : r + TH;
            ex0 = min(4 * c, EW - 4);
        } else {
            const int jj = j - 2 * AP * kRowRuns;
            ey = AP + (jj >> 1);
            ex0 = (jj & 1) ? EW - 4 : 0;
        }
        const int y = med3i(y0 - AP + ey, 0, P.H - 1);
        uint32_t oa[4], oc[4];
        warpRegionRun<S, UI>(P, hrow, ty0, kA32, kC32, x0, ex0, y, oa, oc);
#pragma unroll
        for (int i = 0; i < 4; i++) es[ey * EWB + ex0 + i] = (uint8_t)emaxPacked(oa[i], oc[i]);
    }
    // interior pixels in runs of 4 (columns 4 rc .. 4 rc + 3, rows rr + 16 k): warps kept in registers (packed)
    const int rc = tid & 15, rr = tid >> 4;
    uint32_t wa[4 * RK], wc[4 * RK];
#pragma unroll
    for (int k = 0; k < RK; k++) {
        const int y = min(y0 + rr + 16 * k, P.H - 1);
        uint32_t* oa = wa + 4 * k;
        uint32_t* oc = wc + 4 * k;
        warpRegionRun<S, UI>(P, hrow, ty0, kA32, kC32, x0, 4 * rc + AP, y, oa, oc);
#pragma unroll
        for (int i = 0; i < 4; i++) es[(rr + 16 * k + AP) * EWB + 4 * rc + i + AP] = (uint8_t)emaxPacked(oa[i], oc[i]);
    }
    __syncthreads();
    boxRows7B<EH>(es, hs, tid);
    __syncthreads();
    uint32_t mine = 0u;
#pragma unroll
    for (int k = 0; k < RK; k++) {
        const int x = x0 + 4 * rc, y = y0 + rr + 16 * k;
        if (y >= P.H || x >= P.W) continue;
        const uint16_t* h = &hs[(rr + 16 * k) * TW + 4 * rc];
        uint32_t res[4];
        uint32_t nr[4];
        uint8_t* dst = P.out + (int64_t)y * P.opitch + 4 * (int64_t)x;
        const uint8_t* src = nf.p + (int64_t)y * nf.pitch + 4 * (int64_t)x;
        const bool full = x + 3 < P.W;
        if (full) {
            const U32x4 n4 = *(const U32x4*)src;
            nr[0] = n4.x; nr[1] = n4.y; nr[2] = n4.z; nr[3] = n4.w;
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) nr[i] = x + i < P.W ? ((const uint32_t*)src)[i] : 0u;
        }
        // the 7-row sums of the run's 4 columns: 8-byte rows added in 16-bit lanes (every sum <= 49 x 255 < 2^16)
        uint32_t s01 = 0u, s23 = 0u;
#pragma unroll
        for (int m = 0; m < 7; m++) {
            const uint2 q = *(const uint2*)&h[m * TW];
            s01 = pkAdd16(s01, q.x);
            s23 = pkAdd16(s23, q.y);
        }
        const int sums[4] = {(int)(s01 & 0xFFFFu), (int)(s01 >> 16), (int)(s23 & 0xFFFFu), (int)(s23 >> 16)};
        // every pixel of the wavefront's runs trusted (u = 0 at S <= sZero <= sHalf: none counted): the blend of
        // u = 0 for all four without a branch a pixel; the same values as blendRgba8(.., 0)
        const bool trusted = max(max(sums[0], sums[1]), max(sums[2], sums[3])) <= P.sZero;
        if (__all(trusted)) {
            if (P.mid) {
#pragma unroll
                for (int i = 0; i < 4; i++)
                    res[i] = (__builtin_amdgcn_lerp(wa[4 * k + i], wc[4 * k + i], 0x01010101u) & 0x00FFFFFFu) |
                             (nr[i] & 0xFF000000u);
            } else if (P.quarter) {
#pragma unroll
                for (int i = 0; i < 4; i++) res[i] = quarterRgba8(P, wa[4 * k + i], wc[4 * k + i]) | (nr[i] & 0xFF000000u);
            } else {
#pragma unroll
                for (int i = 0; i < 4; i++) res[i] = blendRgba8(P, wa[4 * k + i], wc[4 * k + i], nr[i], 0.0f);
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int S49 = sums[i];
                const bool inside = x + i < P.W;
                mine += (inside && S49 > P.sHalf) ? 1u : 0u;
                res[i] = blendRgba8(P, wa[4 * k + i], wc[4 * k + i], nr[i], trustOf(P, S49));
            }
        }
        if (full) {
            U32x4 o;
            o.x = res[0]; o.y = res[1]; o.z = res[2]; o.w = res[3];
            *(U32x4*)dst = o;
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++)
                if (x + i < P.W) ((uint32_t*)dst)[i] = res[i];
        }
    }
    // untrusted count: a wavefront sum, one LDS atomic per wavefront, one global atomic per workgroup into its slot
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mine += __shfl_xor(mine, o, 64);
    if ((tid & 63) == 0 && mine) atomicAdd(&cnt, mine);
    __syncthreads();
    if (tid == 0 && cnt) atomicAdd(P.count + ((tileY * nTilesX + tileX) & (kSlots - 1)), cnt);
}

template <int S, int RT, bool UI = false>
__device__ __forceinline__ void synthRgba8Body(const SynthParams& P, int bx, int by) {
    synthRgba8Tile<S, RT, UI>(P, bx, by, P.gx);
}

// lookup.comp:10-19 (interp.py:194-205): the per-pixel field at time t, g <- pf(q - s g) twice (LOOKUP_ITERS = 2)
__global__ __launch_bounds__(256) void lookup_kernel(LookupParams P) {
    const int qx = blockIdx.x * 16 + (threadIdx.x & 15), qy = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (qx >= P.fw || qy >= P.fh) return;
    P.out[qy * P.fw + qx] = lookupAt(P.pf, P.fw, P.fh, P.s, qx, qy);
}

// fallback.comp:9-24: more than globalFallback x W x H untrusted pixels -> the nearer real frame. Every workgroup
// exits at once in the normal case; kFbTile x kFbTile pixels a workgroup keep those workgroups few (40 a 1080p frame).
constexpr int kFbTile = 256;
__device__ __forceinline__ void fallbackBody(const FallbackParams& P, int gbx, int gby) {
    if (P.nextCount && gbx == 0 && gby == 0 && threadIdx.x < kSlots) P.nextCount[threadIdx.x] = 0u;
    const bool flick = (float)slotSum(P.count) > P.gfrac * (float)P.W * (float)P.H;
    if (!flick) return;
    const FrameRef& nf = P.nearC ? P.c : P.a;
    const int bx = gbx * kFbTile, by = gby * kFbTile;
    if (P.bpp >= 4 && !P.outc) {   // one plane of 4- or 8-byte pixels: the tile's rows as 16-byte chunks
        const int w = min(kFbTile, P.W - bx), h = min(kFbTile, P.H - by);
        const int words = w * P.bpp / 4, chunks = (words + 3) >> 2;
#pragma unroll 4
        for (int i = threadIdx.x; i < chunks * h; i += 256) {
            const int r = i / chunks, c = i - r * chunks;
            const uint32_t* src = (const uint32_t*)(nf.p + (int64_t)(by + r) * nf.pitch + (int64_t)bx * P.bpp) + 4 * c;
            uint32_t* dst = (uint32_t*)(P.out + (int64_t)(by + r) * P.opitch + (int64_t)bx * P.bpp) + 4 * c;
            if (4 * c + 4 <= words) {
                *(U32x4*)dst = *(const U32x4*)src;
            } else {
                for (int k = 0; k < words - 4 * c; k++) dst[k] = src[k];
            }
        }
        return;
    }
#pragma unroll 4
    for (int i = threadIdx.x; i < kFbTile * kFbTile; i += 256) {
        const int x = bx + (i % kFbTile), y = by + (i / kFbTile);
        if (x >= P.W || y >= P.H) continue;
        const uint8_t* src = nf.p + (int64_t)y * nf.pitch + (int64_t)x * P.bpp;
        uint8_t* dst = P.out + (int64_t)y * P.opitch + (int64_t)x * P.bpp;
        if (P.bpp == 4) *(uint32_t*)dst = *(const uint32_t*)src;
        else if (P.bpp == 8) *(uint2*)dst = *(const uint2*)src;
        else if (P.bpp == 2) *(uint16_t*)dst = *(const uint16_t*)src;
        else *dst = *src;
        if (P.outc && ((x | y) & 1) == 0) {   // 4:2:0 chroma of the even-even pixel (fallback.comp:19-22)
            const int cx = x >> 1, cy = y >> 1;
            if (P.cbpp == 4)
                *(uint32_t*)(P.outc + (int64_t)cy * P.ocpitch + 4 * (int64_t)cx) =
                    *(const uint32_t*)(nf.c + (int64_t)cy * nf.cpitch + 4 * (int64_t)cx);
            else
                *(uint16_t*)(P.outc + (int64_t)cy * P.ocpitch + 2 * (int64_t)cx) =
                    *(const uint16_t*)(nf.c + (int64_t)cy * nf.cpitch + 2 * (int64_t)cx);
        }
    }
}

// error.comp:9-21 as a debug pass: the packed mismatch e of the latest synth (4 pixels per uint32, low byte first)
__global__ __launch_bounds__(256) void debug_e_rgba8_kernel(SynthParams P, uint32_t* dst, int eStride) {
    const int gx = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (gx >= eStride || y >= P.H) return;
    uint32_t packed = 0u;
    for (int i = 0; i < 4; i++) {
        const int x = gx * 4 + i;
        if (x < P.W) {
            uint32_t a, c;
            warpPairRgba8(P, x, y, halfVec(P.field, x, y, P.S, P.fw, P.fh), a, c);
            packed |= (uint32_t)emaxRgba8(a, c) << (8 * i);
        }
    }
    dst[(int64_t)y * eStride + gx] = packed;
}

// [untrusted count of a synth, scene-cut SAD sum] from their partial sums (framegen.cpp:2073-2084)
__global__ __launch_bounds__(64) void debug_counters_kernel(const uint32_t* count, const uint32_t* sad, uint32_t* dst) {
    const uint32_t c = slotSum(count), s = slotSum(sad);
    if (threadIdx.x == 0) {
        dst[0] = c;
        dst[1] = s;
    }
}

hipError_t launchDebugCounters(const uint32_t* count, const uint32_t* sad, uint32_t* dst, hipStream_t s) {
    hipLaunchKernelGGL(debug_counters_kernel, dim3(1), dim3(64), 0, s, count, sad, dst);
    return hipGetLastError();
}


#define FGR_SYNTH_BODY(S, RT) synthRgba8Body<S, RT>
FGR_PASS(synth_rgba8_s1, SynthParams, FGR_SYNTH_BODY(1, 32), 256)
FGR_PASS(synth_rgba8_s2, SynthParams, FGR_SYNTH_BODY(2, 32), 256)
FGR_PASS(synth_rgba8_s4, SynthParams, FGR_SYNTH_BODY(4, 32), 256)
FGR_PASS(synth_rgba8_s1_t16, SynthParams, FGR_SYNTH_BODY(1, 16), 256)
FGR_PASS(synth_rgba8_s2_t16, SynthParams, FGR_SYNTH_BODY(2, 16), 256)
FGR_PASS(synth_rgba8_s4_t16, SynthParams, FGR_SYNTH_BODY(4, 16), 256)
FGR_PASS_OCC(synth_rgba8_s2_t64, SynthParams, FGR_SYNTH_BODY(2, 64), 256, 5)   // 5 workgroups a CU (30.5 KB of LDS each)
#define FGR_SYNTH_UI_BODY(S) synthRgba8Body<S, 32, true>
FGR_PASS_SINGLE(synth_rgba8_ui_s1, SynthParams, FGR_SYNTH_UI_BODY(1), 256)
FGR_PASS_SINGLE(synth_rgba8_ui_s2, SynthParams, FGR_SYNTH_UI_BODY(2), 256)
FGR_PASS_SINGLE(synth_rgba8_ui_s4, SynthParams, FGR_SYNTH_UI_BODY(4), 256)
FGR_PASS(fallback, FallbackParams, fallbackBody, 256)

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

hipError_t launchLookup(const LookupParams& p, hipStream_t s) {
    hipLaunchKernelGGL(lookup_kernel, dim3((unsigned)cdiv(p.fw, 16), (unsigned)cdiv(p.fh, 16)), dim3(256), 0, s, p);
    return hipGetLastError();
}

int synthTilesX(int W) { return cdiv(W, TW); }
int synthTilesY(int H, int rows) { return cdiv(H, rows); }
int fallbackTilesX(int W) { return cdiv(W, kFbTile); }
int fallbackTilesY(int H) { return cdiv(H, kFbTile); }

hipError_t launchSynthSingle(int pass, const void* params, int gx, int gy, hipStream_t s) {
    const dim3 g((unsigned)gx, (unsigned)gy);
    switch (pass) {
    case kPassSynthRgba8S1: hipLaunchKernelGGL(synth_rgba8_s1_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthRgba8S2: hipLaunchKernelGGL(synth_rgba8_s2_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthRgba8S4: hipLaunchKernelGGL(synth_rgba8_s4_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthRgba8S1T16: hipLaunchKernelGGL(synth_rgba8_s1_t16_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthRgba8S2T16: hipLaunchKernelGGL(synth_rgba8_s2_t16_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthRgba8S4T16: hipLaunchKernelGGL(synth_rgba8_s4_t16_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthRgba8S2T64: hipLaunchKernelGGL(synth_rgba8_s2_t64_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthRgba8UiS1: hipLaunchKernelGGL(synth_rgba8_ui_s1_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthRgba8UiS2: hipLaunchKernelGGL(synth_rgba8_ui_s2_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassSynthRgba8UiS4: hipLaunchKernelGGL(synth_rgba8_ui_s4_kernel, g, dim3(256), 0, s, *(const SynthParams*)params); break;
    case kPassFallback: hipLaunchKernelGGL(fallback_kernel, g, dim3(256), 0, s, *(const FallbackParams*)params); break;
    default: return launchFmtSingle(pass, params, gx, gy, s);
    }
    return hipGetLastError();
}

hipError_t launchSynthBatch(int pass, const void* params, const uint32_t* work, int n, hipStream_t s) {
    const dim3 g((unsigned)n);
    switch (pass) {
    case kPassSynthRgba8S1: hipLaunchKernelGGL(synth_rgba8_s1_batch_kernel, g, dim3(256), 0, s, (const SynthParams*)params, work); break;
    case kPassSynthRgba8S2: hipLaunchKernelGGL(synth_rgba8_s2_batch_kernel, g, dim3(256), 0, s, (const SynthParams*)params, work); break;
    case kPassSynthRgba8S4: hipLaunchKernelGGL(synth_rgba8_s4_batch_kernel, g, dim3(256), 0, s, (const SynthParams*)params, work); break;
    case kPassSynthRgba8S1T16: hipLaunchKernelGGL(synth_rgba8_s1_t16_batch_kernel, g, dim3(256), 0, s, (const SynthParams*)params, work); break;
    case kPassSynthRgba8S2T16: hipLaunchKernelGGL(synth_rgba8_s2_t16_batch_kernel, g, dim3(256), 0, s, (const SynthParams*)params, work); break;
    case kPassSynthRgba8S4T16: hipLaunchKernelGGL(synth_rgba8_s4_t16_batch_kernel, g, dim3(256), 0, s, (const SynthParams*)params, work); break;
    case kPassSynthRgba8S2T64: hipLaunchKernelGGL(synth_rgba8_s2_t64_batch_kernel, g, dim3(256), 0, s, (const SynthParams*)params, work); break;
    case kPassFallback: hipLaunchKernelGGL(fallback_batch_kernel, g, dim3(256), 0, s, (const FallbackParams*)params, work); break;
    default: return launchFmtBatch(pass, params, work, n, s);
    }
    return hipGetLastError();
}

hipError_t launchDebugE(const SynthParams& p, uint32_t* dst, int eStride, int format, hipStream_t s) {
    const dim3 g((unsigned)cdiv(eStride, 16), (unsigned)cdiv(p.H, 16));
    switch (format) {
    case kRgba8: hipLaunchKernelGGL(debug_e_rgba8_kernel, g, dim3(256), 0, s, p, dst, eStride); break;
    default: return launchDebugEFmt(p, dst, eStride, s);
    }
    return hipGetLastError();
}

}  // namespace fgr
