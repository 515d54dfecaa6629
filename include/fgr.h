/* fgr.h -- C ABI of the MI355X (gfx950) implementation of frame-gen-ritz's MotionVersion::V1 path.
 *
 * The reference's interface is the C++ class framegen::Interpolator (gpu/framegen.h:443-822), which records Vulkan
 * compute work into a caller's VkCommandBuffer. Here the same operations enqueue HIP kernels on a caller's
 * hipStream_t (passed as void*), in stream order, and never synchronise the device inside a record call.
 * include/framegen.h wraps these entry points 1:1 as the C++ drop-in.
 *
 * Conventions (as the reference's):
 *   - Not thread-safe. All records of one handle go to ONE stream, in recording order: the handle's scratch buffers
 *     are shared between records and only stream order serialises them (framegen.h:4-6).
 *   - A record function returns 1 on success, 0 if it was refused; a refused record enqueues nothing and leaves `out`
 *     untouched (framegen.h:537-548). Reasons are string literals (fgr_init_error, fgr_settings_error, fgr_size_error).
 *   - Frames are caller-owned device memory: FgrFrame{ptr, pitch} (pitch in bytes), plus the CbCr plane for Nv12 /
 *     P010.
 *     A frame is refused when a plane is null, its pitch is below one row's bytes or 2^24 or more, or its last row
 *     ends at 2^32 bytes or beyond.
 *   - The library allocates its own buffers with hipMalloc only in fgr_create / fgr_resize / fgr_reconfigure /
 *     fgr_set_settings (the last two when a runtime setting needs a larger buffer: the UI box, blurSamples); it never
 *     allocates or frees inside a record.
 */
#ifndef FGR_H
#define FGR_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* framegen.h:138 Format. Rgba8 = R8G8B8A8 / B8G8R8A8 UNORM (FgrInitInfo.bgra picks the byte order), Rgb10 =
 * A2B10G10R10 UNORM (R in bits 0-9; bgra: A2R10G10B10), RgbaF16 = R16G16B16A16 SFLOAT, Nv12 = 8-bit 4:2:0 (Y plane +
 * interleaved CbCr plane at half resolution), P010 = the same with 10-bit codes in the high bits of 16-bit words. */
enum { FGR_FORMAT_RGBA8 = 0, FGR_FORMAT_RGB10 = 1, FGR_FORMAT_RGBAF16 = 2, FGR_FORMAT_NV12 = 3, FGR_FORMAT_P010 = 4 };
/* framegen.h:144 Transfer (RgbaF16 only): 0 = display-encoded values (level = value * 255), 1 = linear light. */
enum { FGR_TRANSFER_ENCODED = 0, FGR_TRANSFER_LINEAR = 1 };
/* framegen.h:150 YuvMatrix (Nv12 and P010). */
enum { FGR_YUV_BT601 = 0, FGR_YUV_BT709 = 1, FGR_YUV_BT2020 = 2 };
/* framegen.h:28 MotionVersion: only V1 (0) is built. */
enum { FGR_MOTION_V1 = 0 };

/* framegen.h:167-409 Settings: the fields V1 reads, with the reference's defaults (fgr_default_settings) and ranges
 * (fgr_validate). seedNet is this library's own addition (no reference counterpart): 1 adds the learned seed
 * proposal of an attached seed network (fgr_set_seed_net) as a seventh seed candidate on the level it was trained
 * for; 0 (default) is the plain V1 search, bit-identical to a library without it. */
typedef struct FgrSettings {
    float trustLow;           /* [0, 255]; trustHigh >= trustLow + 1                       (framegen.h:173) */
    float trustHigh;          /* [0, 255]                                                  (framegen.h:174) */
    float globalFallback;     /* (0, 1]                                                    (framegen.h:177) */
    float sceneCutSad;        /* (0, 255]                                                  (framegen.h:180) */
    float hudBonus;           /* [0, 255]                                                  (framegen.h:186) */
    float smoothBonus;        /* [0, 255]                                                  (framegen.h:189) */
    float searchPenalty;      /* [0, 255]                                                  (framegen.h:193) */
    uint32_t searchRadius;    /* [1, 12]                                                   (framegen.h:202) */
    uint32_t searchRadiusTop; /* [1, 12]                                                   (framegen.h:203) */
    uint32_t subpixel;        /* 0 / 1                                                     (framegen.h:206) */
    uint32_t motionVersion;   /* V1 only                                                   (framegen.h:325) */
    uint32_t blurSamples;     /* 0..16; < 2 = off                                          (framegen.h:315) */
    uint32_t blurWeights;     /* 0 Even, 1 Gaussian                                        (framegen.h:316) */
    uint32_t flowScale;       /* structural: 1, 2, 4                                       (framegen.h:403) */
    uint32_t pyramidMinSide;  /* structural: [16, 256]                                     (framegen.h:408) */
    uint32_t seedNet;         /* 0 / 1 (no reference counterpart, see above)                             */
    /* Static-UI protection (framegen.h:207-305). Built: Off and CrosshairV2 (every format); Crosshair and WholeScreen
     * are valid values that fgr_settings_error refuses. */
    uint32_t uiProtection;        /* FGR_UI_*                                                  (framegen.h:225) */
    uint32_t uiHistory;           /* [FGR_UI_FRAMES, FGR_UI_FRAMES_MAX] kept observed frames   (framegen.h:234) */
    float uiBoxHeightFrac;        /* [0, 1]: box side = round(H * frac), 0 = uiCrosshairArea   (framegen.h:250) */
    uint32_t uiBoxWidth;          /* [0, 16384], 0 = from the fraction                         (framegen.h:253) */
    uint32_t uiBoxHeight;         /* [0, 16384]                                                (framegen.h:253) */
    float uiCrosshairArea;        /* (0, 1], legacy box area (used when uiBoxHeightFrac = 0)   (framegen.h:257) */
    int32_t uiBoxOffsetX;         /* [-16384, 16384] from the frame centre                     (framegen.h:260) */
    int32_t uiBoxOffsetY;         /* [-16384, 16384]                                           (framegen.h:260) */
    uint32_t uiStillFrames;       /* [1, 255]                                                  (framegen.h:265) */
    float uiStillThreshold;       /* [0, 255] 8-bit levels                                     (framegen.h:266) */
    uint32_t uiInpaintReach;      /* [1, 16] (Crosshair / WholeScreen only, validated)         (framegen.h:270) */
    float uiMaxStillFraction;     /* (0, 1] (WholeScreen only, validated)                      (framegen.h:274) */
    float uiContrastThreshold;    /* [0, 255] (CrosshairV2)                                    (framegen.h:282) */
    uint32_t uiLeaveFrames;       /* [1, 255] (CrosshairV2)                                    (framegen.h:286) */
    uint32_t uiReentryFrames;     /* [0, 255] (CrosshairV2)                                    (framegen.h:291) */
    uint32_t uiEstimateFill;      /* 0 / 1 (CrosshairV2)                                       (framegen.h:305) */
} FgrSettings;

/* framegen.h:15-24 UiProtection */
enum { FGR_UI_OFF = 0, FGR_UI_CROSSHAIR = 1, FGR_UI_WHOLESCREEN = 2, FGR_UI_CROSSHAIR_V2 = 3 };
#define FGR_UI_FRAMES 3        /* framegen.h:815 kUiFrames: observed frames kept at least */
#define FGR_UI_FRAMES_MAX 65   /* framegen.h:816 kUiFramesMax */
#define FGR_UI_SETS 64         /* framegen.h:813 kUiSets: UI-protection passes per fgr_begin_frame */

/* framegen.h:411-441 InitInfo, with a HIP device index in place of the Vulkan handles. */
typedef struct FgrInitInfo {
    uint32_t width, height;   /* >= FGR_MIN_SIZE each */
    uint32_t format;          /* FGR_FORMAT_* */
    uint32_t bgra;            /* Rgba8 only: 1 = B8G8R8A8 byte order (the luma weights follow the channels) */
    uint32_t transfer;        /* RgbaF16: FGR_TRANSFER_* */
    float linearWhiteNits;    /* RgbaF16 linear: luminance of 1.0, (0, 10000]; 80 = scRGB */
    uint32_t yuvMatrix;       /* Nv12 / P010: FGR_YUV_* */
    uint32_t yuvFullRange;    /* Nv12 / P010: 0 = limited (video) range */
    int32_t device;           /* HIP device ordinal */
    FgrSettings settings;
} FgrInitInfo;

/* framegen.h:154-159 Frame: one caller-owned image. Nv12 / P010: ptr = Y plane (W x H bytes / uint16 words), chroma =
 * CbCr plane (W/2 x H/2 pairs of bytes / uint16 words; W and H even). Frames are identified by ptr (the reference
 * compares image views). */
typedef struct FgrFrame {
    void* ptr;
    uint64_t pitch;           /* bytes per row of ptr */
    void* chroma;
    uint64_t chromaPitch;     /* bytes per row of chroma */
} FgrFrame;

/* framegen.h:736-739 DebugMapLayout (byte offsets relative to the destination offset). */
typedef struct FgrDebugMapLayout {
    uint32_t W, H, fw, fh, bw, bh, fscale, eStride;   /* eStride in uint32 words */
    uint64_t blockOff, pixOff, eOff, size;
} FgrDebugMapLayout;

/* Sizes of one frame's motion estimate (framegen.cpp:658-685 sizeLayout). */
typedef struct FgrSizeLayout {
    int32_t fw, fh, levels;
    int32_t w[16], h[16], bw[16], bh[16];
} FgrSizeLayout;

#define FGR_MIN_SIZE 16        /* framegen.h:817 kMinSize */
#define FGR_SET_RING 32        /* framegen.h:812 kSetRing: records of each kind per fgr_begin_frame */
#define FGR_BLUR_MAX_SAMPLES 16 /* framegen.h:814 kBlurMaxSamples */

typedef struct FgrInterpolator FgrInterpolator;

/* ---- settings (host only, no device work) ---- */
void fgr_default_settings(FgrSettings* out);                                   /* framegen.h:167-409 defaults */
int fgr_validate(const FgrSettings* s);                                        /* framegen.h:464 validate() */
const char* fgr_settings_error_for_format(const FgrSettings* s, uint32_t format); /* framegen.h:468 settingsError(s, f) */
int fgr_size_layout(uint32_t W, uint32_t H, uint32_t flowScale, uint32_t minSide, FgrSizeLayout* out);
uint32_t fgr_version(void);

/* ---- lifetime (framegen.h:451-461) ---- */
/* init(): a new handle, or null (fgr_last_init_error() says why: a string literal). */
FgrInterpolator* fgr_create(const FgrInitInfo* info);
const char* fgr_last_init_error(void);                                         /* framegen.h:453 initError() */
void fgr_destroy(FgrInterpolator* h);                                          /* framegen.h:456 destroy() */
uint32_t fgr_format(const FgrInterpolator* h);                                 /* framegen.h:455 format() */
/* resize(): re-creates the size-dependent buffers; the caller guarantees earlier records have retired. */
int fgr_resize(FgrInterpolator* h, uint32_t width, uint32_t height);           /* framegen.h:461 */

/* ---- settings of a handle (framegen.h:472-507) ---- */
const char* fgr_settings_error(const FgrInterpolator* h, const FgrSettings* s); /* framegen.h:472 */
const char* fgr_size_error(const FgrInterpolator* h, uint32_t w, uint32_t hgt); /* framegen.h:480 sizeError() */
int fgr_set_settings(FgrInterpolator* h, const FgrSettings* s);                /* framegen.h:500 setSettings() */
int fgr_reconfigure(FgrInterpolator* h, const FgrSettings* s);                 /* framegen.h:505 reconfigure() */
void fgr_get_settings(const FgrInterpolator* h, FgrSettings* out);             /* framegen.h:507 settings() */

/* ---- records (framegen.h:511-573) ---- */
void fgr_begin_frame(FgrInterpolator* h);                                      /* framegen.h:511 beginFrame() */
int fgr_record_estimate(FgrInterpolator* h, void* stream, FgrFrame prev, FgrFrame curr);            /* :549 */
int fgr_record_synth(FgrInterpolator* h, void* stream, FgrFrame prev, FgrFrame curr, FgrFrame out, float t); /* :550 */
int fgr_record(FgrInterpolator* h, void* stream, FgrFrame prev, FgrFrame curr, FgrFrame out, float t);       /* :551 */
int fgr_record_synth_blur(FgrInterpolator* h, void* stream, FgrFrame prev, FgrFrame curr, FgrFrame out,
                          float t0, float t1);                                                       /* :573 */

/* ---- a record that allocates nothing (this library's addition; no counterpart in the reference) ----
 * For callers that cannot allocate device memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc. The same work and results as fgr_record (one estimate, then one synth at t) without a handle: the
 * pair's size, format and settings come in `info` (its device and format fields as fgr_create reads them), and every
 * intermediate (luma pyramid, block and per-pixel fields, counters) lives in one file-scope device array of
 * FGR_NOALLOC_SCRATCH_BYTES per GPU, held by each GPU the library runs on until the process ends. The call allocates
 * and frees nothing, on the device or the host. Refused (returns 0, records nothing, *error = a string literal reason
 * when error is not null): what fgr_create or fgr_record would refuse; a pair whose intermediates exceed the array
 * (1920x1080 at flowScale 1 or 3840x2160 at flowScale 2 fit); RgbaF16 (its level table is a per-handle device
 * buffer); uiProtection other than Off and seedNet 1 (per-handle state). The array keeps the last call's contents and
 * is shared by every call on that GPU: calls in one stream are ordered, calls whose work overlaps on one GPU (two
 * streams, two threads) race. */
#define FGR_NOALLOC_SCRATCH_BYTES (48u << 20)
int fgr_record_noalloc(void* stream, const FgrInitInfo* info, FgrFrame prev, FgrFrame curr, FgrFrame out, float t,
                       const char** error);

/* ---- static-UI protection, CrosshairV2 (framegen.h:575-693) ----
 * recordObserve: once per new real frame (in presentation order), after its pixels are final: advances the box's
 * per-pixel state machine and keeps the frame's mask (the frame pointer is the key; uiHistory frames are kept, least
 * recently observed evicted). recordTrack: a real frame passed through without generating: the state machine only. With
 * protection on, an estimate of a pair whose frames are both kept runs protected: its synths reject warp taps that land
 * on a frame's own UI pixels and paste curr's UI pixels that did not change in the pair. A host that never observes gets
 * the implicit observes of recordEstimate (framegen.h:602-609). Off: both return 1 and enqueue nothing. */
int fgr_record_observe(FgrInterpolator* h, void* stream, FgrFrame frame);      /* framegen.h:610 recordObserve() */
int fgr_record_track(FgrInterpolator* h, void* stream, FgrFrame frame);        /* framegen.h:686 recordTrack() */
void fgr_release_unused_ui(FgrInterpolator* h);                                /* framegen.h:693 releaseUnusedUi() */
/* box geometry (host only, no validity check): framegen.h:481-494 */
void fgr_box_size(const FgrSettings* s, uint32_t w, uint32_t h, uint32_t* boxW, uint32_t* boxH);   /* boxSize() */
void fgr_box_rect(const FgrSettings* s, uint32_t w, uint32_t h, int32_t rect[4]);                  /* boxRect() {x, y, w, h} */
uint32_t fgr_box_margin(const FgrSettings* s);                                                     /* boxMargin() */
void fgr_work_rect(const FgrSettings* s, uint32_t w, uint32_t h, int32_t rect[4]);                 /* workRect() */
/* Debug (this library's addition): the kept mask of `frame` (rect = its work rectangle {x, y, w, h}; rect[2] bytes a
 * row: 255 = UI unchanged in its pair, 64 = UI that changed, 0 = not UI) copied to dst (device, rect[2] * rect[3] bytes),
 * or the box state (4 uint32 words per work-rectangle pixel) when frame.ptr is null. Returns 0 if there is none. */
int fgr_debug_ui_mask(FgrInterpolator* h, void* stream, FgrFrame frame, void* dst, int32_t rect[4]);

/* ---- profiling (framegen.h:705-717): HIP events after every pass group; ticks are nanoseconds ---- */
int fgr_set_profiling(FgrInterpolator* h, int on);
/* Waits for the last profile. Fills up to cap entries; returns the number of pass groups (-1 if none). */
int fgr_read_timings(FgrInterpolator* h, const char** names, uint64_t* ticks, int cap);

/* ---- debug (framegen.h:719-741) ---- */
/* Enqueues a copy of [untrusted pixels, sum of roundEven(256 * blockSAD)] of the latest record into dst (device). */
int fgr_record_debug_counters(FgrInterpolator* h, void* stream, void* dst);
uint32_t fgr_debug_sad_blocks(const FgrInterpolator* h);
int fgr_debug_map_layout(const FgrInterpolator* h, FgrDebugMapLayout* out);
/* Block field (bw x bh float2, after sub-pixel), per-pixel field (fw x fh float2) and the packed mismatch e of the
 * latest synth (W x H bytes, eStride words per row, low byte first) into dst (device memory of layout.size bytes). */
int fgr_record_debug_maps(FgrInterpolator* h, void* stream, void* dst);
/* The 64-byte header of an fgtest --dump file (fgtest.cpp:24-28, the .fgmaps format read by gpu/fgmaps.py) for the
 * maps fgr_record_debug_maps writes: uint32 magic 'FGMP', version 5, W, H, fw, fh, bw, bh, fscale, eStride, float32
 * trustLow, trustHigh, smoothBonus, hudBonus, uint32 motionVersion, v2Mode (0, 0 for V1). The file is this header
 * followed by the layout.size bytes of the maps. */
int fgr_fgmaps_header(const FgrInterpolator* h, void* out64);

/* ---- batched V1 (this library's addition; no counterpart in the reference) ----
 * Many pairs of any sizes, each with its own multiplier N (outputs at t = k/N, k = 1..N-1), processed as one launch
 * per pass for the whole batch: the estimate of every pair, then every pair's k-th synth. The frames are bound at
 * create time (fgr_batch_set_frames rebinds pointers of the same sizes); a record enqueues the whole batch on a stream,
 * captured once into a hipGraph and replayed when useGraph is set. Results are bit-identical to fgr_record_estimate +
 * fgr_record_synth per pair. Not thread-safe; one stream, like a handle. */
typedef struct FgrBatchPair {
    FgrFrame prev, curr;
    uint32_t width, height, format;
    uint32_t multiplier;       /* N in [2, 16] */
    FgrFrame out[15];          /* the N - 1 outputs */
} FgrBatchPair;

typedef struct FgrBatchInfo {
    int32_t device;
    uint32_t bgra, transfer;   /* as FgrInitInfo, for every pair */
    float linearWhiteNits;
    uint32_t yuvMatrix, yuvFullRange;
    uint32_t useGraph;         /* 1: capture the batch into a hipGraph on the first record, replay it after */
    FgrSettings settings;
    struct FgrSeedNet* seedNet; /* attached to every pair (fgr_set_seed_net) when settings.seedNet is 1; may be null */
} FgrBatchInfo;

typedef struct FgrBatch FgrBatch;
FgrBatch* fgr_batch_create(const FgrBatchInfo* info, const FgrBatchPair* pairs, uint32_t n);
const char* fgr_batch_last_error(void);
void fgr_batch_destroy(FgrBatch* b);
int fgr_batch_set_frames(FgrBatch* b, void* stream, const FgrBatchPair* pairs, uint32_t n);
int fgr_batch_record(FgrBatch* b, void* stream);
uint32_t fgr_batch_launches(const FgrBatch* b);
/* While profiling is on, a record enqueues the launches eagerly (no graph), each between two HIP events on the stream;
 * read_timings waits for the last one and fills, per launch, its stage (0 luma, 100 + i down to level i + 1,
 * 200 + 10 s medhist / 201 + 10 s search / 202 + 10 s median3 of the s-th level from the coarsest, 400 subpix,
 * 401 pixfield, 500 + 10 k synth k, 501 + 10 k its fallback), its pass and its duration in ns. Returns the number of
 * launches (-1 if nothing was profiled). */
int fgr_batch_set_profiling(FgrBatch* b, int on);
/* Tile autotuner: the tile configurations are the synth's tile height (16, 32 or 64 rows: RGBA8 at any flowScale but
 * 64 rows at 2 only, the other formats at flowScale 2, 32 rows otherwise) and the number of
 * consecutive tiles each XCD takes from a launch's work list (16, 64 or 256). autotune times all nine on this batch's
 * own records (`reps` each; the outputs are written), keeps the fastest, then the estimate forms (below) one knob at a
 * time, then 1 to 4 lanes (below), and remembers the choice for batches of the same pairs (sizes, formats,
 * multipliers, flowScale, pyramidMinSide), which then start with it. Results are bit-identical under every
 * configuration. ns: 17 times per record (ns; index rows * 3 + chunk, then 1 .. 4 lanes, then the forms), or null. */
int fgr_batch_autotune(FgrBatch* b, void* stream, int32_t reps, uint64_t* ns);
int fgr_batch_set_tuning(FgrBatch* b, int32_t rows, int32_t chunk);
int fgr_batch_get_tuning(const FgrBatch* b, int32_t* rows, int32_t* chunk);
/* Lanes (1 to 4; tuned by fgr_batch_autotune after the tiles: ns[9 .. 12] = 1 .. 4 lanes): L lanes split the pairs in
 * L groups of about equal cost, each with its own launches, recorded on L streams that fork from and join into the
 * record's stream, so the short passes of one group run beside the long ones of another. Results are unchanged. */
int fgr_batch_set_lanes(FgrBatch* b, int32_t lanes);
int32_t fgr_batch_get_lanes(const FgrBatch* b);
/* Estimate forms of a batch (bits; bit-identical results; fgr_batch_autotune times each, ns[13 .. 16] = pixel field
 * gathering / LDS windows, then search two blocks a wavefront / one block a wavefront, with the best tiles):
 *   FGR_FORM_PIXFIELD_WINDOWS: the pixel field from LDS windows of the luma (lumas with exact integer sampling) instead
 *     of gathering every tap from the planes; FGR_FORM_SEARCH_NARROW: the radius <= 2 searches one block a wavefront.
 * (The per-pair records always take the LDS windows and two blocks a wavefront.) */
enum { FGR_FORM_PIXFIELD_WINDOWS = 1, FGR_FORM_SEARCH_NARROW = 2 };
int fgr_batch_set_form(FgrBatch* b, int32_t form);
int32_t fgr_batch_get_form(const FgrBatch* b);
int fgr_tuning_save(const char* path);
int fgr_tuning_load(const char* path);   /* returns the number of entries read */
int fgr_batch_read_timings(FgrBatch* b, int32_t* stage, int32_t* pass, uint64_t* ticks, int cap);
/* [untrusted pixels of synth `synth` of pair `pair`, the pair's scene-cut SAD sum x 256] into dst (device) */
int fgr_batch_debug_counters(FgrBatch* b, void* stream, uint32_t pair, uint32_t synth, void* dst);

/* ---- learned seed proposals (this library's addition; no counterpart in the reference) ----
 * A 5-layer 3x3 CNN (2 -> 16 -> 32 -> 64 -> 64 -> 2 channels, stride 2 in layers 1-3, ReLU after 1-4) on the coarsest
 * level's luma pair proposes one vector per 8x8 block of that level; with settings.seedNet = 1 and a network attached
 * the coarsest search takes the better (by SAD, the proposal paying the seed penalty) of zero and the proposal as its
 * seed. Forward and both gradients run on the BF16 matrix cores with FP32 accumulation; training (Adam, mean squared
 * error against the plain search's field) runs on the GPU. Luma arguments are device pointers to n planes of w x h
 * float32 levels (plane i at + i w h); fields are n x ceil(h/8) x ceil(w/8) float2 (x, y) in that level's pixels. */
typedef struct FgrSeedNet FgrSeedNet;
FgrSeedNet* fgr_seednet_create(int32_t device, uint64_t seed);   /* He-normal initial weights from the seed */
void fgr_seednet_destroy(FgrSeedNet* net);
uint32_t fgr_seednet_param_count(void);
int fgr_seednet_get_weights(const FgrSeedNet* net, float* host, uint32_t n);
int fgr_seednet_set_weights(FgrSeedNet* net, const float* host, uint32_t n);    /* also resets the Adam state */
int fgr_seednet_forward(FgrSeedNet* net, void* stream, const float* la, const float* lc, int32_t w, int32_t h, int32_t n,
                        float* out);
/* gradients of the mean squared error against target into grad (device, param_count floats), the loss into loss
 * (device float); either may be null */
int fgr_seednet_gradients(FgrSeedNet* net, void* stream, const float* la, const float* lc, int32_t w, int32_t h,
                          int32_t n, const float* target, float* grad, float* loss);
int fgr_seednet_train_step(FgrSeedNet* net, void* stream, const float* la, const float* lc, int32_t w, int32_t h,
                           int32_t n, const float* target, float lr, float* loss);
/* the coarsest-level field of n luma pairs: the search with radius R around seed zero -- or, given proposals (n x bh x
 * bw float2, else null), around the better of zero and the proposal -- then median3; raw and median-filtered fields.
 * With null proposals this is the plain field, the training target. */
int fgr_seednet_search_field(void* stream, const float* la, const float* lc, int32_t w, int32_t h, int32_t n, int32_t R,
                             float lam, const float* seeds, float* raw, float* out);
/* FP8 inference (OCP E4M3 on v_mfma_f32_16x16x32_fp8_fp8, FP32 accumulation): calibrate on a batch of luma pairs
 * (per-channel input scales sa = the smallest power of two at or above max|x| / 448, folded into the weights,
 * per-output-channel weight scales sw), then switch the network's forward (and the proposals of attached handles,
 * from their next record) between FP8 and BF16 (v_mfma_f32_16x16x32_bf16). Both store each layer's activations once,
 * quantised, for the next layer. fp8_scales copies all layers' sa (2 + 16 + 32 + 64 + 64 floats) and sw
 * (16 + 32 + 64 + 64 + 2 floats) to the host. Training always runs in BF16; calibrate again after training. */
int fgr_seednet_calibrate_fp8(FgrSeedNet* net, void* stream, const float* la, const float* lc, int32_t w, int32_t h,
                              int32_t n);
int fgr_seednet_set_fp8(FgrSeedNet* net, int on);
int fgr_seednet_fp8_scales(const FgrSeedNet* net, float* sa, float* sw);
/* attach (or detach with null) a network to a handle; it must outlive the handle's records that use it */
int fgr_set_seed_net(FgrInterpolator* h, FgrSeedNet* net);

/* ---- test / debug only; no counterpart in the reference ----
 * Device pointers of the handle's internal buffers and their layout (float / float2 offsets), so a test can read every
 * intermediate of the latest estimate (luma pyramid, per-level raw and median-filtered block fields, medians, the
 * sub-pixel field, the per-pixel field, the counters). Valid until the next resize / reconfigure / destroy. */
typedef struct FgrDebugInternal {
    const void* luma;       /* float */
    const void* field;      /* float2 */
    const void* counters;   /* uint32[1152]: [4 + 2i] level i median (float bits), [64..127] partial sums of the scene-cut
                               SAD, [128 + 64 (k % 16) ..] partial untrusted counts of the k-th synth of the estimate */
    int32_t levels, fw, fh;
    int32_t w[16], h[16], bw[16], bh[16], lumaOffA[16], lumaOffC[16], fS[16], fM[16];
    int32_t fieldSub, fieldPF, fieldT, lumaWords, fieldEntries, lastSynthSlot;
    int32_t fieldSeed;      /* float2 offset of the learned proposals of the coarsest level's blocks */
} FgrDebugInternal;
int fgr_debug_internal(const FgrInterpolator* h, FgrDebugInternal* out);
/* the FP8 E4M3 conversion of the seed network's FP8 forward (round to nearest even, saturating at +-448) on n device
 * floats, into n device bytes */
int fgr_debug_fp8_e4m3(void* stream, const float* in, uint8_t* out, uint32_t n);
int fgr_batch_debug_internal(const FgrBatch* b, uint32_t pair, FgrDebugInternal* out);

#ifdef __cplusplus
}
#endif
#endif
