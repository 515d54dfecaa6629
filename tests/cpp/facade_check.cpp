// Drives the C++ drop-in (include/framegen.h) the way a host of the reference would: init, beginFrame, recordEstimate,
// recordSynth at each t, one stream; reads prev / curr RGBA8 files and writes each output frame.
//   facade_check W H prev.rgba curr.rgba out_prefix t1 [t2 ...]
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "framegen.h"

static bool readAll(const char* path, std::vector<uint8_t>& buf) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    const size_t n = std::fread(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    return n == buf.size();
}

int main(int argc, char** argv) {
    if (argc < 7) return 2;
    const uint32_t W = (uint32_t)std::atoi(argv[1]), H = (uint32_t)std::atoi(argv[2]);
    const size_t bytes = (size_t)W * H * 4;
    std::vector<uint8_t> a(bytes), c(bytes), o(bytes);
    if (!readAll(argv[3], a) || !readAll(argv[4], c)) return 3;
    framegen::InitInfo ii;
    ii.width = W;
    ii.height = H;
    framegen::Interpolator fg;
    if (!fg.init(ii)) {
        std::fprintf(stderr, "init: %s\n", fg.initError());
        return 4;
    }
    void *da, *dc, *dout;
    if (hipMalloc(&da, bytes) != hipSuccess || hipMalloc(&dc, bytes) != hipSuccess || hipMalloc(&dout, bytes) != hipSuccess)
        return 5;
    hipMemcpy(da, a.data(), bytes, hipMemcpyHostToDevice);
    hipMemcpy(dc, c.data(), bytes, hipMemcpyHostToDevice);
    hipStream_t s;
    hipStreamCreate(&s);
    const framegen::Frame prev{da, W * 4ull}, curr{dc, W * 4ull}, out{dout, W * 4ull};
    fg.beginFrame();
    if (!fg.recordEstimate(s, prev, curr)) return 6;
    for (int i = 6; i < argc; i++) {
        if (!fg.recordSynth(s, prev, curr, out, (float)std::atof(argv[i]))) return 7;
        hipMemcpyAsync(o.data(), dout, bytes, hipMemcpyDeviceToHost, s);
        hipStreamSynchronize(s);
        char path[1024];
        std::snprintf(path, sizeof(path), "%s_%d.rgba", argv[5], i - 6);
        FILE* f = std::fopen(path, "wb");
        if (!f || std::fwrite(o.data(), 1, bytes, f) != bytes) return 8;
        std::fclose(f);
    }
    std::printf("sad blocks %u\n", fg.debugSadBlocks());
    fg.destroy();
    hipFree(da);
    hipFree(dc);
    hipFree(dout);
    hipStreamDestroy(s);
    return 0;
}
