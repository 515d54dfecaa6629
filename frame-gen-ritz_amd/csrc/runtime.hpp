// The state behind an FgrInterpolator handle (framegen::Interpolator::Impl, gpu/framegen.cpp:791-994 for V1).
#pragma once
#include <hip/hip_runtime.h>
#include <cstring>
#include <vector>
#include "../../include/fgr.h"
#include "v1_params.hpp"

// One pyramid level (framegen.cpp:658-685 Level): size, block grid, luma offsets of a and c, field offsets of the
// raw search output (fS) and the median3 output (fM), in float / float2 units of the shared buffers.
struct Level {
    int w, h, bw, bh, offA, offC, fS, fM;
};

// One kernel launch of a record, before it is launched: the pass, its stage in the record (ops of the same stage of
// different pairs are independent and run as one batched launch), the per-pair grid and the parameter block.
struct PlanOp {
    static constexpr size_t kMaxParams = 512;
    int pass = 0, stage = 0, gx = 1, gy = 1;
    const char* stamp = nullptr;   // profile group that ends with this op (the single-pair path)
    size_t size = 0;
    alignas(16) uint8_t params[kMaxParams];   // the parameter block, inline: planning a record allocates nothing
    template <class T>
    void set(const T& p) {
        static_assert(sizeof(T) <= kMaxParams, "parameter block larger than PlanOp::kMaxParams");
        size = sizeof(T);
        std::memcpy(params, &p, sizeof(T));
    }
};

// CrosshairV2 (framegen.cpp:893-950): a rectangle of the frame, and one kept observed frame: its key (the frame
// pointer), its observe sequence number (LRU; 0 = holds nothing), its mask over the work rectangle it was observed with.
struct UiRect {
    int x = 0, y = 0, w = 0, h = 0;
    bool operator==(const UiRect& o) const { return x == o.x && y == o.y && w == o.w && h == o.h; }
};

struct UiEntry {
    const void* key = nullptr;
    uint64_t seq = 0;
    bool active = false;   // one of the uiHistory kept frames (an inactive entry keeps its memory)
    uint8_t* mask = nullptr;
    size_t cap = 0;
    UiRect wr;
    int pitch = 0;
};

// A frame the kernels can address for a W x H pair of `format`: non-null planes (the CbCr plane of Nv12 / P010 too),
// a pitch of at least one row's bytes, below 2^24 (rows are addressed as 24-bit products), and every row offset plus
// a row's bytes below 2^32 (32-bit plane offsets).
bool frameFits(const FgrFrame& f, uint32_t W, uint32_t H, int format);
// fgr_create's checks: a format this library builds, and a size its indexing and launch grids hold (nullptr = fine)
bool formatBuilt(uint32_t format);
const char* sizeLimitError(uint32_t W, uint32_t H, uint32_t format);

struct FgrInterpolator {
    static constexpr int kMaxQ = 64;

    FgrInitInfo info{};
    FgrSettings set{};
    int device = 0;
    int format = 0;
    uint32_t W = 0, H = 0;
    int fw = 0, fh = 0;
    std::vector<Level> lv;
    int lumaWords = 0, fieldEntries = 0;
    int fieldSub = 0, fieldSeed = 0, fieldPF = 0, fieldT = 0, blockFieldOff = 0;
    float* luma = nullptr;        // luma pyramids of a and c
    float* field = nullptr;       // float2 fields
    uint32_t* u = nullptr;        // counters and medians

    // per-format constants (formats.cpp)
    int* fmtLut = nullptr;        // RgbaF16: metric level of every half value in 1/16 level (65536 entries)
    float yuv[8] = {};            // Nv12 / P010: YU_YS, YU_RCR, YU_BCB, YU_GCB, YU_GCR
    bool formatInit();
    void formatFree();

    // record state
    bool haveEst = false;
    const void* estPrev = nullptr;
    const void* estCurr = nullptr;
    FgrFrame estPrevF{}, estCurrF{};
    uint32_t ringEst = 0, ringSynth = 0;
    int synthK = 0, lastSlot = 0;
    float lastT = 0.0f;

    // learned seed proposals (seednet.hip): the attached network and its activations on the coarsest level
    FgrSeedNet* net = nullptr;
    int synthRows = 32;               // synth tile height (16, 32, 64; the other formats: 16 and 64 at flowScale 2 only),
                                      // chosen by the batch tuner
    int estForm = 0;                  // batched estimate kernels, chosen by the batch tuner (FGR_FORM_*)
    // recordSynthBlur's sub-frames (blur.hip): sized when settings or sizes are applied with blurSamples >= 2 and only
    // grown; a replaced buffer is kept (records queued before the settings change may still read it) until the
    // handle's buffers are freed, so no record allocates
    uint8_t* blurScratch = nullptr;
    size_t blurCap = 0;
    std::vector<uint8_t*> blurRetired;
    size_t blurBytes(int n) const;
    bool blurEnsure();
    float* seedAct = nullptr;
    bool allocSeedBuffers();
    void planSeeds(std::vector<PlanOp>& plan) const;

    // static-UI protection, CrosshairV2 (ui_host.cpp)
    UiEntry uiE[FGR_UI_FRAMES_MAX];
    uint64_t uiSeq = 0;
    int uiLast = -1;                  // the entry observed last
    bool uiExplicit = false;          // fgr_record_observe has been called: estimates no longer observe implicitly
    bool uiBoxHist = false;           // the state describes the previous real frame
    UiRect uiPrevBox;
    uint32_t* uiState = nullptr;      // 4 words per work-rectangle pixel
    size_t uiStateCap = 0;
    uint32_t uiSets = 0;              // UI passes since fgr_begin_frame (FGR_UI_SETS)
    bool estUi = false;               // the latest estimate ran protected: its entries and their sequence numbers
    int estEnt[2] = {-1, -1};
    uint64_t estSeq[2] = {0, 0};
    bool uiActive() const { return W && set.uiProtection == FGR_UI_CROSSHAIR_V2 && uiState; }
    UiRect uiBox() const;
    UiRect uiWork() const;
    bool uiEnsure();
    void uiFree();
    void uiAdopted(const FgrSettings& old);
    bool uiChanged(const FgrSettings& old, const UiRect& oldBox, const UiRect& oldWork) const;
    int uiFind(const void* key) const;
    int uiPick(const void* key) const;
    void uiForget(const void* key);
    bool uiObserveRecord(hipStream_t s, const FgrFrame& f);
    bool uiTrackRecord(hipStream_t s, const FgrFrame& f);
    fgr::UiMaskRef uiMaskRef(int e) const;
    fgr::UiParams uiPatchParams(const FgrFrame& curr, const FgrFrame& out) const;

    // profiling
    bool profiling = false;
    bool profPending = false;         // an observe opened a profile that the next estimate continues
    std::vector<hipEvent_t> events;
    std::vector<const char*> qnames;
    int qcount = 0;

    std::vector<PlanOp> planBuf;      // the launch plan of the latest record, its capacity reused by the next
    size_t bufferBytes() const;       // luma + fields + counters of the current layout, each 256-byte aligned
    size_t bufferBytesNoDebug() const;   // the same without the debug maps' field at t (last in the field buffer)

    void freeBuffers();
    void layoutBuffers(uint32_t w, uint32_t h);
    bool allocBuffers(uint32_t w, uint32_t h);
    bool rebuild(uint32_t w, uint32_t h);
    bool frameOk(const FgrFrame& f) const;
    fgr::FrameRef ref(const FgrFrame& f) const;
    int bytesPerPixel() const;
    int pixExactBits() const;
    void stamp(hipStream_t s, const char* name);
    void beginProfile(hipStream_t s);
    // batched: the plan of a batch (fgr_batch_*): kernels laid out for many workgroups rather than for latency
    void planEstimate(std::vector<PlanOp>& plan, const FgrFrame& prev, const FgrFrame& curr, bool batched = false) const;
    // merged: the batch's form -- every synth of the pair at one stage (one launch for all its synths), every fallback
    // at one later stage (its count bank was cleared by the estimate, so it clears none)
    void planSynth(std::vector<PlanOp>& plan, const FgrFrame& prev, const FgrFrame& curr, const FgrFrame& out, float t,
                   int k, bool uiPatch = true, bool merged = false) const;
    bool runPlan(hipStream_t s, const std::vector<PlanOp>& plan);
    bool enqueueEstimate(hipStream_t s, const FgrFrame& prev, const FgrFrame& curr);
    bool enqueueSynth(hipStream_t s, const FgrFrame& prev, const FgrFrame& curr, const FgrFrame& out, float t);
    void synthParams(fgr::SynthParams& p, const FgrFrame& prev, const FgrFrame& curr, float t, int slot,
                     bool fusedLookup) const;
    bool copyFrame(hipStream_t s, const FgrFrame& src, const FgrFrame& dst);
    bool enqueueBlur(hipStream_t s, const FgrFrame& prev, const FgrFrame& curr, const FgrFrame& out, float t0, float t1);
};
