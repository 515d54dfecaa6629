// Parameter blocks of the V1 kernels (host and device). Each kernel takes one block by value, the way the reference's
// passes take one push-constant block (gpu/shaders/motion_v1.glsl:5-25, gpu/framegen.cpp:417-422); the fields are the
// subset each kernel reads, with device pointers in place of buffer offsets.
#pragma once
#include <stdint.h>

namespace fgr {

constexpr int kBlock = 8;            // interp.py:43 BLOCK
constexpr int kMaxLevels = 16;
// u[] counter words (motion_v1.glsl:39 U_UNTRUSTED, U_SADSUM, U_MEDIAN), plus the two per-synth slots of this build.
constexpr int kUSadSum = 1;
constexpr int kUMedian = 4;          // level i's global median at u[4 + 2i], u[5 + 2i] (float bits)
constexpr int kUSynth = 40;          // untrusted count of synth k at u[40 + (k & 1)]
constexpr int kUWords = 64;

// Frame formats (fgr.h FGR_FORMAT_*)
enum Fmt : int { kRgba8 = 0, kRgb10 = 1, kF16 = 2, kP010 = 4 };

struct FrameRef {
    const uint8_t* p;   // plane 0
    const uint8_t* c;   // CbCr plane (P010)
    int64_t pitch, cpitch;
};

struct LumaParams {
    FrameRef a, c;
    int W, H, S;        // frame size, flow scale
    int fw, fh;         // level-0 size
    float* la;          // level-0 luma of a (fw x fh)
    float* lc;
    uint32_t* u;        // counters, cleared by block (0, 0)
    int bgra;
    // format conversion (non-Rgba8): per-format constants, see luma kernels
    const float* lut;   // RgbaF16 linear: 65536-entry half -> level table (null otherwise)
    float yBlack, yGain;
};

struct DownParams {
    const float* sa; const float* sc;   // source level
    float* da; float* dc;               // destination level
    int pw, ph, w, h;
};

struct MedParams {
    const float2* parent; int pw, ph;   // parent (coarser) block field, median3 output
    int bw, bh;                         // this level's block grid
    uint32_t* u; int medIdx;
};

struct SearchParams {
    const float* la; const float* lc; int w, h;   // this level's luma
    int bw, bh;
    const float2* parent; int pw, ph;             // null on the coarsest level (seed 0)
    const uint32_t* med;                          // u + medIdx (2 words, float bits)
    const float2* extraSeed;                      // optional 7th seed candidate per block (learned proposals), or null
    int R;
    float lam;
    float2* out;                                  // raw search field (bw x bh)
    uint32_t* u; int sadAcc;
};

struct Median3Params {
    const float2* in; float2* out; int bw, bh;
};

struct SubpixParams {
    const float* la; const float* lc; int w, h, bw, bh;
    const float2* in; float2* out;
    float lam;
    uint32_t* u;
};

struct PixParams {
    const float* la; const float* lc; int w, h;   // level 0
    const float2* bf; int bw, bh;                 // block field
    float2* pf;                                   // per-pixel field, w x h
    float smooth, zero;
};

struct LookupParams {
    const float2* pf; float2* out; int fw, fh;
    float s;
};

struct SynthParams {
    FrameRef a, c;
    uint8_t* out; uint8_t* outc; int64_t opitch, ocpitch;
    int W, H, S, fw, fh;
    const float2* field;      // per-pixel field at time t (flow res)
    const uint32_t* sadsum;   // u + 1 (the estimate's scene-cut sum)
    uint32_t* count;          // this synth's untrusted counter
    int nblocks;
    float cutSad, gfrac, trust0, den;
    double rden;              // 1 / (double)den
    float wA, wC;             // 1 - t, t (float, folded in double)
    int mid;                  // t == 0.5
    int nearC;                // t > 0.5
    const float* lut;         // RgbaF16 linear: half -> level table
    float yuv[8];             // P010: ys, rcr, bcb, gcb, gcr, black, scale
};

struct FallbackParams {
    FrameRef a, c;
    uint8_t* out; uint8_t* outc; int64_t opitch, ocpitch;
    int W, H, bpp;            // bytes per pixel of plane 0
    const uint32_t* count;
    uint32_t* nextCount;      // the other synth slot, cleared here for the next synth
    float gfrac;
    int nearC;
};

}  // namespace fgr
