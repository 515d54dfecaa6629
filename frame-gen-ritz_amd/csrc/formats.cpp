// Per-format constants of the colour passes, fixed at create time (gpu/framegen.cpp:1508-1540 specialization
// constants of gpu/shaders/format.glsl).
//   RgbaF16: the metric level of every half value (fgMetric format.glsl:70-73) in units of 1/16 level, built in float64:
//            encoded M = v * 255; linear M = 255 PQ(clamp(v * linearWhiteNits / 10000, 0, 1)) (fgPq :65-69);
//            L16 = floor(16 M + 0.5), clamped to +-65535 (+-4096 levels), NaN -> 0. One 256 KB device table.
//   Nv12 / P010: the YCbCr -> R'G'B' difference constants YU_YS, YU_RCR, YU_BCB, YU_GCB, YU_GCR (framegen.cpp:1519-1533).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <vector>
#include "runtime.hpp"

namespace {

double halfToDouble(uint16_t h) {
    const int s = h >> 15, e = (h >> 10) & 31, m = h & 1023;
    double v;
    if (e == 0) v = std::ldexp((double)m, -24);
    else if (e == 31) v = m ? NAN : INFINITY;
    else v = std::ldexp((double)(m | 1024), e - 25);
    return s ? -v : v;
}

double pqCode(double y) {
    const double p = std::pow(y, 0.1593017578125);
    return std::pow((0.8359375 + 18.8515625 * p) / (1.0 + 18.6875 * p), 78.84375);
}

}  // namespace

bool FgrInterpolator::formatInit() {
    if (format == FGR_FORMAT_RGBAF16) {
        const bool linear = info.transfer == FGR_TRANSFER_LINEAR;
        const double white = info.linearWhiteNits;
        std::vector<int> lut(65536);
        for (uint32_t i = 0; i < 65536; i++) {
            const double x = halfToDouble((uint16_t)i);
            int q = 0;
            if (!std::isnan(x)) {
                double m;
                if (linear) m = 255.0 * pqCode(std::min(std::max(x * (white / 10000.0), 0.0), 1.0));
                else m = x * 255.0;
                const double f = std::isfinite(m) ? std::floor(16.0 * m + 0.5) : (m > 0 ? 65535.0 : -65535.0);
                q = (int)std::min(std::max(f, -65535.0), 65535.0);
            }
            lut[i] = q;
        }
        if (hipMalloc(&fmtLut, 65536 * sizeof(int)) != hipSuccess) {
            fmtLut = nullptr;
            return false;
        }
        if (hipMemcpy(fmtLut, lut.data(), 65536 * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return false;
    }
    if (format == FGR_FORMAT_P010 || format == FGR_FORMAT_NV12) {
        double kr = 0.2126, kb = 0.0722;   // BT.709
        if (info.yuvMatrix == FGR_YUV_BT601) { kr = 0.299; kb = 0.114; }
        if (info.yuvMatrix == FGR_YUV_BT2020) { kr = 0.2627; kb = 0.0593; }
        const double kg = 1.0 - kr - kb;
        const double sY = info.yuvFullRange ? 1.0 : 255.0 / 219.0, sC = info.yuvFullRange ? 1.0 : 255.0 / 224.0;
        yuv[0] = (float)sY;
        yuv[1] = (float)(sC * 2.0 * (1.0 - kr));
        yuv[2] = (float)(sC * 2.0 * (1.0 - kb));
        yuv[3] = (float)(-sC * 2.0 * (1.0 - kb) * kb / kg);
        yuv[4] = (float)(-sC * 2.0 * (1.0 - kr) * kr / kg);
    }
    return true;
}

void FgrInterpolator::formatFree() {
    if (fmtLut) hipFree(fmtLut);
    fmtLut = nullptr;
}
