// Parameter blocks of the V1 kernels (host and device). Each kernel takes one block by value, the way the reference's
// passes take one push-constant block (gpu/shaders/motion_v1.glsl:5-25, gpu/framegen.cpp:417-422); the fields are the
// subset each kernel reads, with device pointers in place of buffer offsets.
#pragma once
#include <stdint.h>

namespace fgr {

constexpr int kBlock = 8;            // interp.py:43 BLOCK
constexpr int kMaxLevels = 16;
// u[] counter words. The reference keeps one word per global sum (motion_v1.glsl:39 U_UNTRUSTED, U_SADSUM) and
// every workgroup adds to it. Device-scope atomics on one address serialise across the chip, so here each sum is
// spread over kSlots words (a workgroup adds to slot blockId & 63) and its readers add the 64 partial sums (integers:
// the order does not matter).
constexpr int kSlots = 64;
constexpr int kUMedian = 4;          // level i's global median at u[4 + 2i], u[5 + 2i] (float bits)
constexpr int kUSad = 64;            // scene-cut SAD partial sums, u[64 .. 127]
constexpr int kUSynth = 128;         // untrusted partial counts of synth k at u[128 + 64 (k % 16) ..]
constexpr int kSynthBanks = 16;
constexpr int kUWords = kUSynth + kSynthBanks * kSlots;

// Batched launches: each workgroup reads its work item, (parameter block index << kWorkShift) | block index within that
// pass of that pair, and runs the same body as the per-pair kernel.
constexpr int kWorkShift = 22;
constexpr uint32_t kWorkMask = (1u << kWorkShift) - 1u;

// Frame formats (fgr.h FGR_FORMAT_*)
enum Fmt : int { kRgba8 = 0, kRgb10 = 1, kF16 = 2, kNv12 = 3, kP010 = 4 };

struct FrameRef {
    const uint8_t* p;   // plane 0
    const uint8_t* c;   // CbCr plane (Nv12, P010)
    int64_t pitch, cpitch;
};

struct LumaParams {
    int gx;             // blocks per grid row (the batched kernels map a linear block index to (bx, by))
    FrameRef a, c;
    int W, H, S;        // frame size, flow scale
    int fw, fh;         // level-0 size
    float* la;          // level-0 luma of a (fw x fh)
    float* lc;
    uint32_t* u;        // counters, cleared by block (0, 0)
    int bgra;
    int fmt;            // Fmt (the Rgba8 kernel ignores it)
    const int* lut;     // RgbaF16: metric level of every half value, in 1/16 level (65536 entries)
    int yFull;          // Nv12 / P010: full-range codes (else limited range: black 16 / 64, white 235 / 940)
};

// The fused luma pyramid (v1_pyramid.hpp): level 0 from the frames and the coarser levels, kPyrLevels at most
constexpr int kPyrLevels = 6;
struct PyrParams {
    int gx;             // blocks per grid row (the batched kernels map a linear block index to (bx, by))
    FrameRef a, c;
    int W, H, S;        // frame size, flow scale
    uint32_t* u;        // counters, cleared by block (0, 0)
    int bgra;
    int fmt;            // Fmt (the Rgba8 kernel ignores it)
    const int* lut;     // RgbaF16 level table
    int yFull;          // Nv12 / P010 full-range codes
    int levels;         // levels written here (1 .. kPyrLevels)
    float* la[kPyrLevels];
    float* lc[kPyrLevels];
    int w[kPyrLevels], h[kPyrLevels];
};

struct DownParams {
    int gx;             // blocks per grid row (the batched kernels map a linear block index to (bx, by))
    const float* sa; const float* sc;   // source level
    float* da; float* dc;               // destination level
    int pw, ph, w, h;
};

struct MedParams {
    int gx;             // blocks per grid row (the batched kernels map a linear block index to (bx, by))
    const float2* parent; int pw, ph;   // parent (coarser) block field, median3 output
    const float2* raw; float2* filtered; // when raw is set: the parent's search field, whose 3x3 median the kernel first
                                         // writes to filtered (== parent) -- the parent level's median3 pass
    int bw, bh;                         // this level's block grid
    uint32_t* u; int medIdx;
    int shift;                          // -1: bins of [-128, 127] parent pixels (clamped); else the exact two-pass median
                                        // with first-pass bins of 2^shift pixels (medhist.comp FG_WIDE)
};

struct SearchParams {
    int gx;             // blocks per grid row (the batched kernels map a linear block index to (bx, by))
    const float* la; const float* lc; int w, h;   // this level's luma
    int bw, bh;
    const float2* parent; int pw, ph;             // null on the coarsest level (seed 0)
    const uint32_t* med;                          // u + medIdx (2 words, float bits)
    const float2* extraSeed;                      // optional 7th seed candidate per block (learned proposals), or null
    int R;
    float lam;
    float2* out;                                  // raw search field (bw x bh)
    uint32_t* u; int sadAcc;
    int fb;             // this level's luma is k 2^-fb with 0 <= k < 2^16 (packed 16-bit SADs), or -1 (float SADs)
};

struct Median3Params {
    int gx;             // blocks per grid row (the batched kernels map a linear block index to (bx, by))
    const float2* in; float2* out; int bw, bh;
};

struct SubpixParams {
    int gx;             // blocks per grid row (the batched kernels map a linear block index to (bx, by))
    const float* la; const float* lc; int w, h, bw, bh;
    const float2* in; float2* out;
    const float2* raw; float2* filtered;   // when raw is set: level 0's search field, whose 3x3 median (the median3
                                           // pass) the kernel forms for its blocks and writes to filtered (== in)
    float lam;
    uint32_t* u;
    int fb;             // as PixParams.fb, for the integer kernel (fb <= 6)
};

struct PixParams {
    int gx;             // blocks per grid row (the batched kernels map a linear block index to (bx, by))
    const float* la; const float* lc; int w, h;   // level 0
    const float2* bf; int bw, bh;                 // block field
    float2* pf;                                   // per-pixel field, w x h
    float smooth, zero;
    int fb;             // the level-0 luma's fractional bits when it is k 2^-fb with 0 <= k 2^-fb <= 255 and
                        // k < 2^14 (sampling it is then exact integer arithmetic); -1 otherwise
};

struct LookupParams {
    int gx;             // blocks per grid row (the batched kernels map a linear block index to (bx, by))
    const float2* pf; float2* out; int fw, fh;
    float s;
};

// A kept CrosshairV2 mask of one frame: bytes over the rectangle (x, y, w, h) of the frame, `pitch` bytes a row;
// non-zero = a UI pixel. w = 0: no mask (nothing is ever hit).
struct UiMaskRef {
    const uint8_t* m;
    int x, y, w, h, pitch;
};

struct SynthParams {
    int gx;             // blocks per grid row (the batched kernels map a linear block index to (bx, by))
    FrameRef a, c;
    uint8_t* out; uint8_t* outc; int64_t opitch, ocpitch;
    int W, H, S, fw, fh;
    const float2* field;      // per-pixel field at time t (flow res): the midpoint field pf, or its lookup at t
    int lookup;               // 1: field is pf and the kernel looks it up at t itself (lookup.comp), s = 2t - 1
    float s;
    int sZero, sOne, sHalf;   // 7x7 sums: trust is 0 for S <= sZero, 1 for S >= sOne, > 0.5 for S > sHalf
    int quarter;              // 4(1 - t) and 4t are integers: the trusted blend is (kA a + kC c + 2) >> 2 exactly
    int kA, kC;
    const uint32_t* sadsum;   // the estimate's scene-cut SAD partial sums (kSlots words)
    uint32_t* count;          // this synth's untrusted partial counts (kSlots words)
    int nblocks;
    float cutSad, gfrac, trust0, den;
    double rden;              // 1 / (double)den
    float wA, wC;             // 1 - t, t (float, folded in double)
    int mid;                  // t == 0.5
    int nearC;                // t > 0.5
    int fmt;                  // Fmt
    const int* lut;           // RgbaF16: metric level of every half value, in 1/16 level (65536 entries)
    float yuv[8];             // Nv12 / P010: ys, rcr, bcb, gcb, gcr (format.glsl YU_YS, YU_RCR, YU_BCB, YU_GCB, YU_GCR)
    int cw, ch;               // Nv12 / P010: chroma plane size
    UiMaskRef uiA, uiC;       // CrosshairV2 pair: the masks of prev and curr for the tap rejection (the UI kernels)
};

// CrosshairV2 detection of one observed / tracked frame (ui.hip), the patch of a generated frame
struct UiParams {
    FrameRef f;               // the observed frame (update, mask), or curr (patch)
    uint8_t* out; uint8_t* outc; int64_t opitch, ocpitch;   // patch: the generated frame (and its CbCr plane)
    int W, H;
    int bx, by, bw, bh;       // the box: detection
    int wx, wy, ww, wh;       // the work rectangle: state layout and mask
    int reset;                // the state describes no previous frame
    float thr, cthr;          // uiStillThreshold, uiContrastThreshold (8-bit levels)
    int still, leave, reentry;   // uiStillFrames, uiLeaveFrames, uiReentryFrames
    uint32_t* state;          // 4 words per work-rectangle pixel
    uint8_t* mask; int mpitch;   // the entry's mask (ww x wh bytes)
    float* luma; int S, fw, fh, bgra, reach;   // estimate fill: this frame's level-0 luma (fw x fh), flowScale
    int fmt;                  // Fmt
    const int* lut;           // RgbaF16: the 1/16-level metric of every half value
    float yuv[5];             // Nv12 / P010: ys, rcr, bcb, gcb, gcr (the contrast's R'G'B' difference)
    int yFull;                // Nv12 / P010 full range (the fill's luma)
};

struct FallbackParams {
    int gx;             // blocks per grid row (the batched kernels map a linear block index to (bx, by))
    FrameRef a, c;
    uint8_t* out; uint8_t* outc; int64_t opitch, ocpitch;
    int W, H, bpp;            // bytes per pixel of plane 0
    int cbpp;                 // bytes per chroma texel (Nv12 2, P010 4)
    const uint32_t* count;    // kSlots partial counts
    uint32_t* nextCount;      // the next synth's bank (kSlots words), cleared here for it; null: none (batch)
    float gfrac;
    int nearC;
};

}  // namespace fgr
