// Parameter block of the seed network's convolution kernels (seednet.hip).
#pragma once
#include <stdint.h>

namespace fgr {

constexpr int kNetLayers = 5;
constexpr int kNetCin[kNetLayers] = {2, 16, 32, 64, 64};
constexpr int kNetCout[kNetLayers] = {16, 32, 64, 64, 2};
constexpr int kNetStride[kNetLayers] = {2, 2, 2, 1, 1};
// quantised forward: 16-pixel tiles per wavefront of layers 2-5 (layer 1 has a kernel of its own)
constexpr int kSeedTpw[kNetLayers] = {0, 2, 1, 1, 1};
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
    // quantised inference forward (passes kPassSeedQ + 2 layer + FP8): activations NHWC, stored once per layer as
    // BF16, or as OCP FP8 E4M3 with input channel ci quantised as e4m3(x / sa[ci]), sa a power of two (so x / sa is
    // the exact product x qs, qs = 1 / sa); FP8 weights with sa folded in, wq[co][k] = e4m3(W[co][k] sa[ci] / sw[co]),
    // and out = sw[co] sum + b[co]
    const void* qin;              // the layer's input (layers 2-5)
    void* qout;                   // its output (layers 1-4; layer 5 writes float32 `out`)
    const void* wq; int kp;       // weights [Cout][kp] (BF16 W, or FP8 as above), zero from 9 Cin to kp
    const float* sw;              // FP8: per output channel weight scale
    const float* qsIn;            // FP8 layer 1: qs of the two luma channels
    const float* qsOut;           // FP8 layers 1-4: qs of the next layer's input channels
};

int convQTiles(int layer, const ConvParams& p);

}  // namespace fgr

void netLayerSizes(int w, int h, int* lw, int* lh);
