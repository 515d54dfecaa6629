// Parameter block of the seed network's convolution kernels (seednet.hip).
#pragma once
#include <stdint.h>

namespace fgr {

constexpr int kNetLayers = 5;
constexpr int kDwRows = 1024;   // output pixels per split of the weight-gradient sum

struct ConvParams {
    int gx;                       // workgroups per image (4 wave tiles each)
    int n;                        // images
    int mode;                     // 0: input = NHWC activations `in`; 1: input = luma planes la, lc (layer 1)
    const float* in;
    const float* la; const float* lc; int64_t lumaStride;   // mode 1: plane of image i at la + i * lumaStride
    int Cin, inW, inH;
    float* out; int Cout, outW, outH;                       // NHWC; also the ReLU mask of the backward
    const float* w; const float* b;                         // W [Cout][9 Cin] (k = (ky * 3 + kx) * Cin + ci), b [Cout]
    int stride, relu;
    // backward
    const float* dy;              // gradient of the layer's output (before the ReLU mask)
    float* dx;                    // gradient of its input
    float* dw; float* db;         // gradients of W and b
    float* dwPart;                // per-split partial sums of dW
    // FP8 inference (E4M3, per-channel scales): input channel ci is quantised as e4m3(x / sa[ci]); the weights, with
    // sa folded in, as wq[co][k] = e4m3(W[co][k] sa[ci] / sw[co]) (kp bytes a row); out = sw[co] sum + b[co]
    int fp8;
    const uint8_t* wq; int kp;
    const float* sa; const float* sw;
};

int convFwdTiles(const ConvParams& p);

}  // namespace fgr

void netLayerSizes(int w, int h, int* lw, int* lh);
