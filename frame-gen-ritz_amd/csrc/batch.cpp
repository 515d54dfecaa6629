// Batched V1: many frame pairs of any sizes (and any multipliers) in one launch per pass, replayed as a hipGraph.
//
// Each pair keeps the state of a plain handle (its own pyramid, fields and counters) and produces the same plan of
// kernel launches as fgr_record_estimate + fgr_record_synth would (FgrInterpolator::planEstimate / planSynth). The
// batch merges the ops of equal (stage, pass) across pairs into one launch of the pass's batched kernel, whose work
// list maps each workgroup to (pair's parameter block, block within that pair's grid). Stages keep every pair's own
// order (coarse-to-fine levels align at each pair's coarsest level), so the results are bit-identical to the
// per-pair path. The parameter blocks and work lists are built once, at create time, in one device allocation; a
// record is a fixed sequence of launches, captured once into a hipGraph and replayed.
//
// Lanes: with L lanes (1, 2, 3 or 4) the pairs are split in L groups of about equal cost, each with its own launches,
// recorded on L streams (a fork and a join around them). Pairs are independent, so the groups' kernels may overlap:
// the short latency-bound passes of one group (level medians, coarse levels, fallbacks) run beside the long passes of
// another. The autotuner picks the lane count by timing each.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <cstdio>
#include <map>
#include <tuple>
#include <string>
#include <vector>
#include "../../include/fgr.h"
#include "runtime.hpp"
#include "v1_device.hpp"
#include "v1_launch.hpp"
#include "v1_params.hpp"

using namespace fgr;

namespace {
thread_local const char* gBatchError = nullptr;
const char* kBErrUi = "static-UI protection needs the per-frame observes of a handle: batches run with uiProtection Off";

struct Launch {
    int lane = 0;
    int pass = 0;
    int stage = 0;
    int n = 0;             // workgroups
    size_t paramOff = 0;   // byte offsets in the device blob
    size_t workOff = 0;
    size_t paramBytes = 0;
};

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }
}  // namespace

struct FgrBatch {
    FgrBatchInfo info{};
    std::vector<FgrInterpolator*> pairs;
    std::vector<FgrBatchPair> desc;
    std::vector<Launch> launches;
    std::vector<uint8_t> blob;   // host copy of the device blob
    uint8_t* dev = nullptr;
    size_t devCap = 0;
    hipStream_t cap = nullptr;
    static constexpr int kMaxLanes = 4;
    hipStream_t side[kMaxLanes - 1] = {};   // lanes 1.. (lane 0 records on the record's stream)
    hipEvent_t fork = nullptr, join[kMaxLanes - 1] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool profiling = false;
    std::vector<hipEvent_t> events;   // a (start, end) pair per launch while profiling
    bool profiled = false;
    int chunk = kXcdChunk;            // tiles per XCD chunk of the work lists (tuned)
    int rows = 32;                    // synth tile rows (tuned)
    int lanes = 1;                    // 1 .. kMaxLanes (tuned)
    int form = 0;                     // estimate forms, FGR_FORM_* (tuned)
    std::vector<int> laneOf;          // each pair's lane

    ~FgrBatch() {
        for (hipEvent_t e : events) hipEventDestroy(e);
        if (fork) hipEventDestroy(fork);
        for (int i = 0; i < kMaxLanes - 1; i++) {
            if (join[i]) hipEventDestroy(join[i]);
            if (side[i]) hipStreamDestroy(side[i]);
        }
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        if (cap) hipStreamDestroy(cap);
        if (dev) hipFree(dev);
        for (FgrInterpolator* p : pairs) fgr_destroy(p);
    }

    // the plan of every pair, merged by (stage, pass); fills `blob` (params + work lists) and `launches`
    void build() {
        // lanes: the pairs by cost (pixels x multiplier), each to the lighter lane
        laneOf.assign(pairs.size(), 0);
        if (lanes > 1) {
            std::vector<int> idx(pairs.size());
            for (size_t p = 0; p < idx.size(); p++) idx[p] = (int)p;
            auto cost = [&](int p) { return (double)desc[p].width * desc[p].height * desc[p].multiplier; };
            std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cost(a) > cost(b); });
            double load[kMaxLanes] = {};
            for (int p : idx) {
                int l = 0;
                for (int i = 1; i < lanes; i++)
                    if (load[i] < load[l]) l = i;
                laneOf[p] = l;
                load[l] += cost(p);
            }
        }
        std::map<std::tuple<int, int, int>, std::vector<std::pair<int, const PlanOp*>>> groups;
        std::vector<std::vector<PlanOp>> plans(pairs.size());
        for (size_t p = 0; p < pairs.size(); p++) {
            FgrInterpolator& I = *pairs[p];
            const FgrBatchPair& d = desc[p];
            I.planEstimate(plans[p], d.prev, d.curr, true);
            for (uint32_t k = 1; k < d.multiplier; k++)
                I.planSynth(plans[p], d.prev, d.curr, d.out[k - 1], (float)((double)k / (double)d.multiplier), (int)k - 1,
                            true, true);
            for (const PlanOp& op : plans[p]) groups[{op.stage, op.pass, laneOf[p]}].push_back({(int)p, &op});
        }
        launches.clear();
        blob.clear();
        for (auto& g : groups) {
            Launch L;
            L.stage = std::get<0>(g.first);
            L.pass = std::get<1>(g.first);
            L.lane = std::get<2>(g.first);
            const size_t psz = g.second[0].second->size;
            L.paramOff = blob.size();
            L.paramBytes = psz * g.second.size();
            blob.resize(align256(blob.size() + L.paramBytes));
            for (size_t i = 0; i < g.second.size(); i++)
                std::memcpy(blob.data() + L.paramOff + i * psz, g.second[i].second->params, psz);
            // a pair's ops in one group (its synths, its fallbacks) have one grid: their tiles interleave, so that the
            // workgroups that read the same region of the pair's frames run side by side (and share an L2)
            std::vector<uint32_t> work;
            for (size_t i = 0; i < g.second.size();) {
                size_t j = i + 1;
                const PlanOp& op = *g.second[i].second;
                const int tiles = op.gx * op.gy;
                while (j < g.second.size() && g.second[j].first == g.second[i].first &&
                       g.second[j].second->gx * g.second[j].second->gy == tiles)
                    j++;
                for (int l = 0; l < tiles; l++)
                    for (size_t m = i; m < j; m++) work.push_back(((uint32_t)m << kWorkShift) | (uint32_t)l);
                i = j;
            }
            // XCD-aware order (xcdChunkTile, as the per-pair kernels use): tiles that read overlapping regions share
            // an L2
            const size_t n = work.size();
            std::vector<uint32_t> order(n);
            for (size_t i = 0; i < n; i++) order[i] = work[xcdChunkTile((int)i, (int)n, chunk)];
            L.n = (int)n;
            L.workOff = blob.size();
            blob.resize(align256(blob.size() + n * 4));
            std::memcpy(blob.data() + L.workOff, order.data(), n * 4);
            launches.push_back(L);
        }
    }

    // one stream in (stage, pass, lane) order (and timed: each launch between two events), or the two lanes forked
    bool enqueue(hipStream_t s, bool timed = false) {
        const int nl = timed ? 1 : lanes;
        if (nl > 1 && hipEventRecord(fork, s) != hipSuccess) return false;
        for (int i = 0; i + 1 < nl; i++)
            if (hipStreamWaitEvent(side[i], fork, 0) != hipSuccess) return false;
        for (size_t i = 0; i < launches.size(); i++) {
            const Launch& L = launches[i];
            hipStream_t ls = nl > 1 && L.lane > 0 ? side[L.lane - 1] : s;
            if (timed) hipEventRecord(events[2 * i], ls);
            if (launchBatch(L.pass, dev + L.paramOff, (const uint32_t*)(dev + L.workOff), L.n, ls) != hipSuccess)
                return false;
            if (timed) hipEventRecord(events[2 * i + 1], ls);
        }
        for (int i = 0; i + 1 < nl; i++)
            if (hipEventRecord(join[i], side[i]) != hipSuccess || hipStreamWaitEvent(s, join[i], 0) != hipSuccess)
                return false;
        return true;
    }

    bool makeEvents() {
        while (events.size() < 2 * launches.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return false;
            events.push_back(e);
        }
        return true;
    }

    void dropGraph() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }

    // a tile configuration: rebuild the launches and the device blob (its size changes with the grids)
    bool tune(int r, int c, int l = -1, int f = -1) {
        rows = r;
        chunk = c;
        if (l > 0) lanes = l;
        if (f >= 0) form = f;
        for (FgrInterpolator* p : pairs) {
            p->synthRows = r;
            p->estForm = form;
        }
        build();
        dropGraph();
        if (blob.size() > devCap) {
            if (dev) hipFree(dev);
            dev = nullptr;
            devCap = 0;
            if (hipMalloc(&dev, blob.size()) != hipSuccess) return false;
            devCap = blob.size();
        }
        return hipMemcpy(dev, blob.data(), blob.size(), hipMemcpyHostToDevice) == hipSuccess;
    }

    // the tuning cache key: every pair's size, format, multiplier and the settings that shape the grids
    std::string key() const {
        std::string k = std::to_string(info.settings.flowScale) + "/" + std::to_string(info.settings.pyramidMinSide);
        for (const FgrBatchPair& d : desc)
            k += ";" + std::to_string(d.width) + "x" + std::to_string(d.height) + "f" + std::to_string(d.format) + "n" +
                 std::to_string(d.multiplier);
        return k;
    }
};

// Tile configurations chosen by fgr_batch_autotune, per batch key (process-wide; fgr_tuning_save / load persist it)
struct Tuning {
    int rows, chunk, lanes, form;
};
static std::map<std::string, Tuning>& tuningCache() {
    static std::map<std::string, Tuning> m;
    return m;
}

static const char* kBErrNull = "null batch description";
static const char* kBErrCount = "a batch holds 1 to 1024 pairs";
static const char* kBErrMult = "a pair's multiplier must be in [2, 16]";
static const char* kBErrFrames =
    "a pair's frame has a null plane or a pitch below a row's bytes, of 2^24 or more, or with a last row past 2^32";

// every frame of a pair (prev, curr and the multiplier - 1 outputs) addressable at the pair's size and format
static bool pairFramesFit(const FgrBatchPair& d) {
    if (!frameFits(d.prev, d.width, d.height, (int)d.format) || !frameFits(d.curr, d.width, d.height, (int)d.format))
        return false;
    for (uint32_t k = 0; k + 1 < d.multiplier; k++)
        if (!frameFits(d.out[k], d.width, d.height, (int)d.format)) return false;
    return true;
}
static const char* kBErrAlloc = "a device allocation of the batch failed";
static const char* kBErrStream = "the batch's capture stream could not be created";

extern "C" FgrBatch* fgr_batch_create(const FgrBatchInfo* info, const FgrBatchPair* pairs, uint32_t n) {
    gBatchError = nullptr;
    if (!info || !pairs) { gBatchError = kBErrNull; return nullptr; }
    if (n < 1 || n > 1024) { gBatchError = kBErrCount; return nullptr; }
    if (info->settings.uiProtection != FGR_UI_OFF) { gBatchError = kBErrUi; return nullptr; }
    for (uint32_t p = 0; p < n; p++) {
        if (pairs[p].multiplier < 2 || pairs[p].multiplier > 16) { gBatchError = kBErrMult; return nullptr; }
        if (!pairFramesFit(pairs[p])) { gBatchError = kBErrFrames; return nullptr; }
    }
    FgrBatch* b = new FgrBatch();
    b->info = *info;
    b->desc.assign(pairs, pairs + n);
    for (uint32_t p = 0; p < n; p++) {
        FgrInitInfo ii{};
        ii.width = pairs[p].width;
        ii.height = pairs[p].height;
        ii.format = pairs[p].format;
        ii.bgra = info->bgra;
        ii.transfer = info->transfer;
        ii.linearWhiteNits = info->linearWhiteNits;
        ii.yuvMatrix = info->yuvMatrix;
        ii.yuvFullRange = info->yuvFullRange;
        ii.device = info->device;
        ii.settings = info->settings;
        FgrInterpolator* h = fgr_create(&ii);
        if (!h) {
            gBatchError = fgr_last_init_error();
            delete b;
            return nullptr;
        }
        b->pairs.push_back(h);
        if (info->seedNet && !fgr_set_seed_net(h, info->seedNet)) {
            gBatchError = kBErrAlloc;
            delete b;
            return nullptr;
        }
    }
    const auto hit = tuningCache().find(b->key());
    if (hit != tuningCache().end()) {
        b->rows = hit->second.rows;
        b->chunk = hit->second.chunk;
        b->lanes = hit->second.lanes;
        b->form = hit->second.form;
        for (FgrInterpolator* p : b->pairs) {
            p->synthRows = b->rows;
            p->estForm = b->form;
        }
    }
    b->build();
    if (hipMalloc(&b->dev, b->blob.size()) != hipSuccess) { gBatchError = kBErrAlloc; delete b; return nullptr; }
    b->devCap = b->blob.size();
    if (hipMemcpy(b->dev, b->blob.data(), b->blob.size(), hipMemcpyHostToDevice) != hipSuccess) {
        gBatchError = kBErrAlloc;
        delete b;
        return nullptr;
    }
    bool streams = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) == hipSuccess &&
                   hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; streams && i < FgrBatch::kMaxLanes - 1; i++)
        streams = hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) == hipSuccess;
    if (!streams) {
        gBatchError = kBErrStream;
        delete b;
        return nullptr;
    }
    return b;
}

extern "C" const char* fgr_batch_last_error(void) { return gBatchError; }

extern "C" void fgr_batch_destroy(FgrBatch* b) { delete b; }

extern "C" uint32_t fgr_batch_launches(const FgrBatch* b) { return b ? (uint32_t)b->launches.size() : 0u; }

extern "C" int fgr_batch_set_frames(FgrBatch* b, void* stream, const FgrBatchPair* pairs, uint32_t n) {
    // same sizes, formats and multipliers; new frame pointers: rebuild the parameter blocks, stream-ordered upload
    if (!b || !pairs || n != b->desc.size()) return 0;
    for (uint32_t p = 0; p < n; p++)
        if (pairs[p].width != b->desc[p].width || pairs[p].height != b->desc[p].height ||
            pairs[p].multiplier != b->desc[p].multiplier || pairs[p].format != b->desc[p].format ||
            !pairFramesFit(pairs[p]))
            return 0;
    b->desc.assign(pairs, pairs + n);
    b->build();
    if (b->blob.size() > b->devCap) return 0;   // (same sizes: the same layout; a captured graph stays valid)
    return hipMemcpyAsync(b->dev, b->blob.data(), b->blob.size(), hipMemcpyHostToDevice, (hipStream_t)stream) ==
           hipSuccess;
}

extern "C" int fgr_batch_record(FgrBatch* b, void* stream) {
    if (!b) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (b->profiling) {   // eager launches bracketed by events (a graph replay has no per-launch timestamps)
        b->profiled = b->enqueue(s, true);
        return b->profiled ? 1 : 0;
    }
    if (!b->info.useGraph) return b->enqueue(s) ? 1 : 0;
    if (!b->exec) {
        if (hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal) != hipSuccess) return 0;
        const bool ok = b->enqueue(b->cap);
        if (hipStreamEndCapture(b->cap, &b->graph) != hipSuccess || !ok) {
            b->dropGraph();
            return 0;
        }
        if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
            b->dropGraph();
            return 0;
        }
    }
    return hipGraphLaunch(b->exec, s) == hipSuccess ? 1 : 0;
}

extern "C" int fgr_batch_set_tuning(FgrBatch* b, int32_t rows, int32_t chunk) {
    if (!b || !(rows == 16 || rows == 32 || rows == 64) || chunk < 1 || chunk > 4096) return 0;
    hipSetDevice(b->info.device);
    return b->tune(rows, chunk) ? 1 : 0;
}

extern "C" int fgr_batch_get_tuning(const FgrBatch* b, int32_t* rows, int32_t* chunk) {
    if (!b) return 0;
    if (rows) *rows = b->rows;
    if (chunk) *chunk = b->chunk;
    return 1;
}

extern "C" int fgr_batch_set_lanes(FgrBatch* b, int32_t lanes) {
    if (!b || lanes < 1 || lanes > FgrBatch::kMaxLanes) return 0;
    hipSetDevice(b->info.device);
    return b->tune(b->rows, b->chunk, lanes) ? 1 : 0;
}

extern "C" int32_t fgr_batch_get_lanes(const FgrBatch* b) { return b ? b->lanes : 0; }

// Times every tile configuration (synth tiles of 16, 32, 64 rows x work-list XCD chunks of 16, 64, 256 tiles,
// at one lane and the default forms) on this batch's own records -- `reps` eager records each, after one untimed --
// keeps the fastest; then each estimate form knob on and off with those tiles; then 1, 2, 3 and 4 lanes with the best
// of both; and caches the result for batches of the same key. The records write the batch's outputs. ns[i] receives
// configuration i's time per record (i = rowsIndex * 3 + chunkIndex; then ns[9 .. 12] 1 .. 4 lanes; ns[13 .. 16] the
// forms; 17 entries) when given.
extern "C" int fgr_batch_autotune(FgrBatch* b, void* stream, int32_t reps, uint64_t* ns) {
    if (!b || reps < 1) return 0;
    hipSetDevice(b->info.device);
    hipStream_t s = (hipStream_t)stream;
    static const int kRows[3] = {16, 32, 64}, kChunks[3] = {16, 64, 256};
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 0;
    double best = 1e30;
    int bestR = b->rows, bestC = b->chunk;
    bool ok = true;
    for (int ri = 0; ri < 3 && ok; ri++)
        for (int ci = 0; ci < 3 && ok; ci++) {
            if (!b->tune(kRows[ri], kChunks[ci], 1, 0) || !b->enqueue(s)) { ok = false; break; }
            hipEventRecord(e0, s);
            for (int r = 0; r < reps && ok; r++) ok = b->enqueue(s);
            hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) { ok = false; break; }
            float ms = 0.0f;
            hipEventElapsedTime(&ms, e0, e1);
            const double t = (double)ms / reps;
            if (ns) ns[ri * 3 + ci] = (uint64_t)(t * 1e6 + 0.5);
            if (t < best) {
                best = t;
                bestR = kRows[ri];
                bestC = kChunks[ci];
            }
        }
    // the estimate forms with the best tiles at one lane, a knob at a time (ns[13 .. 16] when given): the pixel field
    // gathering / from LDS windows, then the radius <= 2 search two blocks / one block a wavefront
    int bestF = 0;
    static const int kKnobs[2] = {FGR_FORM_PIXFIELD_WINDOWS, FGR_FORM_SEARCH_NARROW};
    for (int kb = 0; kb < 2 && ok; kb++) {
        double tk[2] = {0.0, 0.0};
        for (int on = 0; on < 2 && ok; on++) {
            const int f = on ? (bestF | kKnobs[kb]) : (bestF & ~kKnobs[kb]);
            if (!b->tune(bestR, bestC, 1, f) || !b->enqueue(s)) { ok = false; break; }
            hipEventRecord(e0, s);
            for (int r = 0; r < reps && ok; r++) ok = b->enqueue(s);
            hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) { ok = false; break; }
            float ms = 0.0f;
            hipEventElapsedTime(&ms, e0, e1);
            tk[on] = (double)ms / reps;
            if (ns) ns[13 + 2 * kb + on] = (uint64_t)(tk[on] * 1e6 + 0.5);
        }
        if (ok && tk[1] < tk[0]) bestF |= kKnobs[kb];
        else bestF &= ~kKnobs[kb];
    }
    // then 1 .. 4 lanes with the best tiles and forms (ns[9 .. 12] when given)
    double laneT[FgrBatch::kMaxLanes + 1] = {};
    for (int l = 1; l <= FgrBatch::kMaxLanes && ok; l++) {
        if (!b->tune(bestR, bestC, l, bestF) || !b->enqueue(s)) { ok = false; break; }
        hipEventRecord(e0, s);
        for (int r = 0; r < reps && ok; r++) ok = b->enqueue(s);
        hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) { ok = false; break; }
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e0, e1);
        const double t = (double)ms / reps;
        if (ns) ns[8 + l] = (uint64_t)(t * 1e6 + 0.5);
        laneT[l] = t;
    }
    int bestL = 1;
    for (int l = 2; l <= FgrBatch::kMaxLanes; l++)
        if (laneT[l] < laneT[bestL]) bestL = l;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (!ok || !b->tune(bestR, bestC, bestL, bestF)) return 0;
    tuningCache()[b->key()] = {bestR, bestC, bestL, bestF};
    return 1;
}

extern "C" int fgr_batch_set_form(FgrBatch* b, int32_t form) {
    if (!b || form < 0 || form > (FGR_FORM_PIXFIELD_WINDOWS | FGR_FORM_SEARCH_NARROW)) return 0;
    hipSetDevice(b->info.device);
    return b->tune(b->rows, b->chunk, b->lanes, form) ? 1 : 0;
}

extern "C" int32_t fgr_batch_get_form(const FgrBatch* b) { return b ? b->form : 0; }

// the tuning cache as text, one "key rows chunk lanes form" line per entry (a line without the form reads as form 0)
extern "C" int fgr_tuning_save(const char* path) {
    FILE* f = path ? std::fopen(path, "w") : nullptr;
    if (!f) return 0;
    for (const auto& e : tuningCache())
        std::fprintf(f, "%s %d %d %d %d\n", e.first.c_str(), e.second.rows, e.second.chunk, e.second.lanes,
                     e.second.form);
    std::fclose(f);
    return 1;
}

extern "C" int fgr_tuning_load(const char* path) {
    FILE* f = path ? std::fopen(path, "r") : nullptr;
    if (!f) return 0;
    static thread_local char line[65600], key[65600];
    int n = 0;
    while (std::fgets(line, sizeof line, f)) {
        int r, c, l, fm = 0;
        if (std::sscanf(line, "%65535s %d %d %d %d", key, &r, &c, &l, &fm) < 4) continue;
        tuningCache()[key] = {r, c, std::min(std::max(l, 1), FgrBatch::kMaxLanes),
                              fm & (FGR_FORM_PIXFIELD_WINDOWS | FGR_FORM_SEARCH_NARROW)};
        n++;
    }
    std::fclose(f);
    return n;
}

extern "C" int fgr_batch_set_profiling(FgrBatch* b, int on) {
    if (!b) return 0;
    b->profiled = false;
    b->profiling = false;
    if (on) {
        hipSetDevice(b->info.device);
        if (!b->makeEvents()) return 0;
    }
    b->profiling = on != 0;
    return 1;
}

extern "C" int fgr_batch_read_timings(FgrBatch* b, int32_t* stage, int32_t* pass, uint64_t* ticks, int cap) {
    if (!b || !b->profiled || b->launches.empty()) return -1;
    if (hipEventSynchronize(b->events[2 * b->launches.size() - 1]) != hipSuccess) return -1;
    const int n = (int)b->launches.size();
    for (int i = 0; i < n && i < cap; i++) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, b->events[2 * i], b->events[2 * i + 1]);
        if (stage) stage[i] = b->launches[i].stage;
        if (pass) pass[i] = b->launches[i].pass;
        if (ticks) ticks[i] = (uint64_t)((double)ms * 1e6 + 0.5);
    }
    return n;
}

extern "C" int fgr_batch_debug_counters(FgrBatch* b, void* stream, uint32_t pair, uint32_t synth, void* dst) {
    // [untrusted pixels of the pair's synth `synth` (0-based, t = (synth + 1) / multiplier), SAD sum x 256]
    if (!b || pair >= b->pairs.size() || synth + 1 >= b->desc[pair].multiplier || !dst) return 0;
    FgrInterpolator* h = b->pairs[pair];
    return launchDebugCounters(h->u + kUSynth + kSlots * ((int)synth % kSynthBanks), h->u + kUSad, (uint32_t*)dst,
                               (hipStream_t)stream) == hipSuccess;
}

extern "C" int fgr_batch_debug_internal(const FgrBatch* b, uint32_t pair, FgrDebugInternal* out) {
    if (!b || pair >= b->pairs.size()) return 0;
    return fgr_debug_internal(b->pairs[pair], out);
}
