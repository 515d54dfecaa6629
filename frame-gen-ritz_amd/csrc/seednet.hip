// Learned seed proposals for the coarsest search level (this library's addition; no counterpart in the reference).
//
// A 5-layer 3x3 CNN on the coarsest-level luma pair (a, c): channels 2 -> 16 -> 32 -> 64 -> 64 -> 2, stride 2 in
// layers 1-3, padding 1, ReLU after layers 1-4. Its output grid is ceil(w / 8) x ceil(h / 8), one vector per 8x8 block
// of that level, in that level's pixels. The vector is rounded half-even and offered to the level's search as a second
// seed candidate beside zero (the plain seed): the exact SAD search picks the better of the two (the proposal pays the
// seed penalty lambda, as every non-own candidate does) and then searches +-R around it, so a wrong proposal costs
// nothing but the candidate's SAD, and the exact SAD comparison against the plain seed decides.
//
// Every convolution -- forward, and both gradients -- is an implicit GEMM on the BF16 matrix cores
// (v_mfma_f32_16x16x16_bf16, FP32 accumulation): a wavefront owns a 16 x 16 output tile and gathers its A and B operands
// straight from the NHWC activations and the weights, rounding them to bf16 (round to nearest even). Activations,
// weights, gradients and the Adam state are FP32. Training runs on the GPU with these kernels: forward, mean squared
// error against the plain search's coarsest-level field, backward, Adam.
//
// Layouts: activations NHWC float32 [n][h][w][C]; layer l's weights W_l [Cout][9 Cin] with k = (ky * 3 + kx) * Cin + ci,
// then its bias b_l [Cout]; all layers' parameters in one float array in layer order.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>
#include "../../include/fgr.h"
#include "runtime.hpp"
#include "v1_device.hpp"
#include "v1_launch.hpp"
#include "seednet.hpp"

namespace fgr {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float32 -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ short bf16Bits(float f) {
    const uint32_t u = __float_as_uint(f);
    return (short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// one operand value of the input of a convolution at input pixel (ix, iy), channel ci (zero outside the image)
__device__ __forceinline__ float inputAt(const ConvParams& P, int img, int ix, int iy, int ci) {
    if (ix < 0 || iy < 0 || ix >= P.inW || iy >= P.inH) return 0.0f;
    if (P.mode == 1) {   // luma planes, normalised (v - 128) / 64
        const float* pl = (ci == 0 ? P.la : P.lc) + (int64_t)img * P.lumaStride;
        return (pl[iy * P.inW + ix] - 128.0f) * (1.0f / 64.0f);
    }
    return P.in[(((int64_t)img * P.inH + iy) * P.inW + ix) * P.Cin + ci];
}

// dY of the layer's output at (img, pixel p, channel co), times the ReLU mask (output > 0) when the layer has one
__device__ __forceinline__ float gradAt(const ConvParams& P, int img, int p, int co) {
    const int64_t i = ((int64_t)img * P.outW * P.outH + p) * P.Cout + co;
    const float g = P.dy[i];
    return (P.relu && !(P.out[i] > 0.0f)) ? 0.0f : g;
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// float32 -> OCP FP8 E4M3 (bias 7, largest finite 448), round to nearest even, saturating; written out so that the CPU
// oracle can follow it bit for bit
__device__ __forceinline__ uint32_t fp8e4m3(float v) {
    const uint32_t s = v < 0.0f ? 0x80u : 0u;
    const float a = fabsf(v);
    if (!(a < 448.0f)) return s | 0x7Eu;
    if (a < 0.015625f) return s | (uint32_t)(int)rintf(a * 512.0f);   // subnormals: multiples of 2^-9 (8 -> 2^-6)
    int e;
    frexpf(a, &e);
    const int eb = e - 1;                                              // a in [2^eb, 2^(eb + 1))
    int M = (int)rintf(ldexpf(a, 3 - eb)) - 8, E = eb + 7;
    if (M == 8) {
        M = 0;
        E++;
    }
    if (E > 15 || (E == 15 && M == 7)) return s | 0x7Eu;
    return s | (uint32_t)(E << 3) | (uint32_t)M;
}

// FP8 forward (v_mfma_f32_16x16x32_fp8_fp8): the tile of convFwdBody with 32-deep K steps, 8 E4M3 operands a lane
__device__ __forceinline__ void convFwdFp8(const ConvParams& P, int bx, int by) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int Pn = P.outW * P.outH, K = 9 * P.Cin;
    const int nPT = (Pn + 15) >> 4, nCT = (P.Cout + 15) >> 4;
    const int t = bx * 4 + wave;
    if (t >= nPT * nCT) return;
    const int ct = t % nCT, pt = t / nCT, img = by;
    const int row = lane & 15, quad = lane >> 4;
    const int p = pt * 16 + row;
    const int ox = p % P.outW, oy = p / P.outW;
    const int co = ct * 16 + row;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < K; k0 += 32) {
        uint64_t a = 0, b = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = k0 + 8 * quad + j;
            uint32_t aq = 0, bq = 0;
            if (k < K) {
                const int tap = k / P.Cin, ci = k - tap * P.Cin;
                const int ky = tap / 3, kx = tap - 3 * ky;
                if (p < Pn) aq = fp8e4m3(inputAt(P, img, ox * P.stride + kx - 1, oy * P.stride + ky - 1, ci) / P.sa[ci]);
                if (co < P.Cout) bq = P.wq[co * P.kp + k];
            }
            a |= (uint64_t)aq << (8 * j);
            b |= (uint64_t)bq << (8 * j);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8((long)a, (long)b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int pp = pt * 16 + 4 * quad + j, cc = ct * 16 + row;
        if (pp < Pn && cc < P.Cout) {
            float v = acc[j] * P.sw[cc] + P.b[cc];
            if (P.relu) v = fmaxf(v, 0.0f);
            P.out[((int64_t)img * Pn + pp) * P.Cout + cc] = v;
        }
    }
}

// per-channel largest |input| of a layer over all its pixels and images (one workgroup per channel), for the FP8 scales
__global__ __launch_bounds__(256) void conv_absmax_kernel(ConvParams P, float* out) {
    __shared__ float red[256];
    const int ci = blockIdx.x, Q = P.inW * P.inH;
    float m = 0.0f;
    for (int r = threadIdx.x; r < P.n * Q; r += 256) {
        const int img = r / Q, q = r - img * Q;
        m = fmaxf(m, fabsf(inputAt(P, img, q % P.inW, q / P.inW, ci)));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[ci] = red[0];
}

// Forward: out[img][p][co] = relu?(b[co] + sum_k A[p][k] W[co][k]), A = the 3x3 patch of input pixel
// (ox s + kx - 1, oy s + ky - 1). Wave tile: 16 output pixels x 16 output channels; workgroup: 4 tiles.
__device__ __forceinline__ void convFwdBody(const ConvParams& P, int bx, int by) {
    if (P.fp8) {
        convFwdFp8(P, bx, by);
        return;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int Pn = P.outW * P.outH, K = 9 * P.Cin;
    const int nPT = (Pn + 15) >> 4, nCT = (P.Cout + 15) >> 4;
    const int t = bx * 4 + wave;
    if (t >= nPT * nCT) return;
    const int ct = t % nCT, pt = t / nCT, img = by;
    const int row = lane & 15, quad = lane >> 4;
    const int p = pt * 16 + row;
    const int ox = p % P.outW, oy = p / P.outW;
    const int co = ct * 16 + row;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = k0 + 4 * quad + j;
            float av = 0.0f, bv = 0.0f;
            if (k < K) {
                const int tap = k / P.Cin, ci = k - tap * P.Cin;
                const int ky = tap / 3, kx = tap - 3 * ky;
                if (p < Pn) av = inputAt(P, img, ox * P.stride + kx - 1, oy * P.stride + ky - 1, ci);
                if (co < P.Cout) bv = P.w[co * K + k];
            }
            a[j] = bf16Bits(av);
            b[j] = bf16Bits(bv);
        }
        acc = mfma16(a, b, acc);
    }
    // D[row 4 quad + j][col row]: pixel pt * 16 + 4 quad + j, channel ct * 16 + row
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int pp = pt * 16 + 4 * quad + j, cc = ct * 16 + row;
        if (pp < Pn && cc < P.Cout) {
            float v = acc[j] + P.b[cc];
            if (P.relu) v = fmaxf(v, 0.0f);
            P.out[((int64_t)img * Pn + pp) * P.Cout + cc] = v;
        }
    }
}

// Gradient of the input: dx[img][q][ci] = sum over (tap, co) of W[co][tap Cin + ci] dYm[(q - tap + 1) / s][co] where
// the output pixel is whole and inside (a transposed convolution), as an implicit GEMM with k = tap Cout + co.
__device__ __forceinline__ void convDxBody(const ConvParams& P, int bx, int by) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int Qn = P.inW * P.inH, K = 9 * P.Cout, KW = 9 * P.Cin;
    const int nPT = (Qn + 15) >> 4, nCT = (P.Cin + 15) >> 4;
    const int t = bx * 4 + wave;
    if (t >= nPT * nCT) return;
    const int ct = t % nCT, pt = t / nCT, img = by;
    const int row = lane & 15, quad = lane >> 4;
    const int q = pt * 16 + row;
    const int ix = q % P.inW, iy = q / P.inW;
    const int ci = ct * 16 + row;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = k0 + 4 * quad + j;
            float av = 0.0f, bv = 0.0f;
            if (k < K) {
                const int tap = k / P.Cout, co = k - tap * P.Cout;
                const int ky = tap / 3, kx = tap - 3 * ky;
                const int sy = iy + 1 - ky, sx = ix + 1 - kx;
                if (q < Qn && sx >= 0 && sy >= 0 && sx % P.stride == 0 && sy % P.stride == 0) {
                    const int ox = sx / P.stride, oy = sy / P.stride;
                    if (ox < P.outW && oy < P.outH) av = gradAt(P, img, oy * P.outW + ox, co);
                }
                if (ci < P.Cin) bv = P.w[co * KW + tap * P.Cin + ci];
            }
            a[j] = bf16Bits(av);
            b[j] = bf16Bits(bv);
        }
        acc = mfma16(a, b, acc);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int qq = pt * 16 + 4 * quad + j, cc = ct * 16 + row;
        if (qq < Qn && cc < P.Cin) P.dx[((int64_t)img * Qn + qq) * P.Cin + cc] = acc[j];
    }
}

// Gradient of the weights, one split of the pixel sum: part[split][co][k] = sum over the split's output pixels r of
// dYm[r][co] A[r][k] (A the forward patch matrix). Wave tile: 16 channels x 16 k; the pixels of all images are one
// reduction axis, cut into splits of kDwRows; the splits are added in order by convDwReduce.
__device__ __forceinline__ void convDwBody(const ConvParams& P, int bx, int by) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int Pn = P.outW * P.outH, K = 9 * P.Cin, R = P.n * Pn;
    const int nCT = (P.Cout + 15) >> 4, nKT = (K + 15) >> 4;
    const int t = bx * 4 + wave;
    if (t >= nCT * nKT) return;
    const int ct = t % nCT, kt = t / nCT, split = by;
    const int row = lane & 15, quad = lane >> 4;
    const int co = ct * 16 + row, k = kt * 16 + row;
    int tap = 0, ci = 0, ky = 0, kx = 0;
    if (k < K) {
        tap = k / P.Cin;
        ci = k - tap * P.Cin;
        ky = tap / 3;
        kx = tap - 3 * ky;
    }
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int r0 = split * kDwRows, r1 = min(R, r0 + kDwRows);
    for (int rb = r0; rb < r1; rb += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = rb + 4 * quad + j;
            float av = 0.0f, bv = 0.0f;
            if (r < r1) {
                const int img = r / Pn, p = r - img * Pn;
                if (co < P.Cout) av = gradAt(P, img, p, co);
                if (k < K) {
                    const int ox = p % P.outW, oy = p / P.outW;
                    bv = inputAt(P, img, ox * P.stride + kx - 1, oy * P.stride + ky - 1, ci);
                }
            }
            a[j] = bf16Bits(av);
            b[j] = bf16Bits(bv);
        }
        acc = mfma16(a, b, acc);
    }
    // D[row 4 quad + j][col row]: channel ct * 16 + 4 quad + j, k kt * 16 + row
    float* part = P.dwPart + (int64_t)split * P.Cout * K;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int cc = ct * 16 + 4 * quad + j, kk = kt * 16 + row;
        if (cc < P.Cout && kk < K) part[cc * K + kk] = acc[j];
    }
}

// dW = sum of the splits in order; db[co] = sum of dYm over all pixels (one workgroup per output channel, fixed order)
__global__ __launch_bounds__(256) void conv_dw_reduce_kernel(ConvParams P, int nSplit) {
    const int K = 9 * P.Cin;
    const int co = blockIdx.x;
    for (int k = threadIdx.x; k < K; k += 256) {
        float s = 0.0f;
        for (int sp = 0; sp < nSplit; sp++) s += P.dwPart[((int64_t)sp * P.Cout + co) * K + k];
        P.dw[co * K + k] = s;
    }
    __shared__ float red[256];
    const int Pn = P.outW * P.outH, R = P.n * Pn;
    float s = 0.0f;
    for (int r = threadIdx.x; r < R; r += 256) s += gradAt(P, r / Pn, r % Pn, co);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) P.db[co] = red[0];
}

// mean squared error over the blocks of all images: loss = sum |pred - target|^2 / nBlocks, dpred = 2 (pred - target)
// / nBlocks. The loss sum is one workgroup, fixed order.
__global__ __launch_bounds__(256) void mse_kernel(const float* pred, const float* target, float* dpred, float* loss,
                                                  int count) {
    __shared__ float red[256];
    const float inv = 1.0f / (float)(count / 2);
    float s = 0.0f;
    for (int i = threadIdx.x; i < count; i += 256) {
        const float d = pred[i] - target[i];
        dpred[i] = 2.0f * d * inv;
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss) *loss = red[0] * inv;
}

// Adam (Kingma & Ba 2015), bias-corrected: m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
// w -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps)
__global__ __launch_bounds__(256) void adam_kernel(float* w, const float* g, float* m, float* v, int n, float lr,
                                                   float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

FGR_PASS(seed_conv, ConvParams, convFwdBody, 256)

__global__ __launch_bounds__(256) void conv_dx_kernel(ConvParams P) { convDxBody(P, (int)blockIdx.x, (int)blockIdx.y); }
__global__ __launch_bounds__(256) void conv_dw_kernel(ConvParams P) { convDwBody(P, (int)blockIdx.x, (int)blockIdx.y); }

hipError_t launchSeedSingle(int pass, const void* params, int gx, int gy, hipStream_t s) {
    if (pass != kPassSeedConv) return hipErrorInvalidValue;
    hipLaunchKernelGGL(seed_conv_kernel, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s, *(const ConvParams*)params);
    return hipGetLastError();
}

hipError_t launchSeedBatch(int pass, const void* params, const uint32_t* work, int n, hipStream_t s) {
    if (pass != kPassSeedConv) return hipErrorInvalidValue;
    hipLaunchKernelGGL(seed_conv_batch_kernel, dim3((unsigned)n), dim3(256), 0, s, (const ConvParams*)params, work);
    return hipGetLastError();
}

int convFwdTiles(const ConvParams& p) {
    const int nPT = (p.outW * p.outH + 15) / 16, nCT = (p.Cout + 15) / 16;
    return (nPT * nCT + 3) / 4;
}

}  // namespace fgr

using namespace fgr;

// ------------------------------------------------------------------------------------------------------ host side

static const int kCin[kNetLayers] = {2, 16, 32, 64, 64};
static const int kCout[kNetLayers] = {16, 32, 64, 64, 2};
static const int kStride[kNetLayers] = {2, 2, 2, 1, 1};

static int layerParams(int l) { return kCout[l] * 9 * kCin[l] + kCout[l]; }
static int paramCount() {
    int n = 0;
    for (int l = 0; l < kNetLayers; l++) n += layerParams(l);
    return n;
}
static int layerOffset(int l) {
    int n = 0;
    for (int i = 0; i < l; i++) n += layerParams(i);
    return n;
}

void netLayerSizes(int w, int h, int* lw, int* lh) {
    lw[0] = w;
    lh[0] = h;
    for (int l = 0; l < kNetLayers; l++) {
        lw[l + 1] = (lw[l] + kStride[l] - 1) / kStride[l];
        lh[l + 1] = (lh[l] + kStride[l] - 1) / kStride[l];
    }
}

struct FgrSeedNet {
    int device = 0;
    float* w = nullptr;      // parameters (paramCount floats)
    float* g = nullptr;      // gradients
    float* m = nullptr;      // Adam moments
    float* v = nullptr;
    int steps = 0;
    // training buffers, grown on demand
    float* act = nullptr;    // activations of layers 1..5 (NHWC)
    float* grad = nullptr;   // their gradients
    float* part = nullptr;   // weight-gradient partial sums
    float* loss = nullptr;   // one float
    size_t actCap = 0, partCap = 0;
    // FP8 inference: per layer E4M3 weights (Cout x kp bytes), input scales sa (Cin) and output scales sw (Cout)
    uint8_t* wq = nullptr;
    float* sa = nullptr;     // all layers' sa, then all layers' sw
    bool fp8Ready = false, fp8On = false;

    ~FgrSeedNet() {
        hipSetDevice(device);
        for (float* p : {w, g, m, v, act, grad, part, loss, sa})
            if (p) hipFree(p);
        if (wq) hipFree(wq);
    }

    bool grow(int w0, int h0, int n) {
        int lw[kNetLayers + 1], lh[kNetLayers + 1];
        netLayerSizes(w0, h0, lw, lh);
        size_t need = 0, partNeed = 0;
        for (int l = 0; l < kNetLayers; l++) {
            need += (size_t)n * lw[l + 1] * lh[l + 1] * kCout[l];
            const size_t R = (size_t)n * lw[l + 1] * lh[l + 1];
            partNeed = std::max(partNeed, ((R + kDwRows - 1) / kDwRows) * kCout[l] * 9 * kCin[l]);
        }
        if (need > actCap) {
            if (act) hipFree(act);
            if (grad) hipFree(grad);
            act = grad = nullptr;
            actCap = 0;
            if (hipMalloc(&act, need * 4) != hipSuccess || hipMalloc(&grad, need * 4) != hipSuccess) return false;
            actCap = need;
        }
        if (partNeed > partCap) {
            if (part) hipFree(part);
            part = nullptr;
            partCap = 0;
            if (hipMalloc(&part, partNeed * 4) != hipSuccess) return false;
            partCap = partNeed;
        }
        return true;
    }

    // the ConvParams of layer l over n images of a w0 x h0 luma pair, activations at act (layer outputs in order)
    ConvParams layer(int l, const float* la, const float* lc, int w0, int h0, int n, float* acts, float* grads,
                     const float* weights, bool fp8 = false) const {
        int lw[kNetLayers + 1], lh[kNetLayers + 1];
        netLayerSizes(w0, h0, lw, lh);
        size_t off[kNetLayers + 1];
        off[0] = 0;
        for (int i = 0; i < kNetLayers; i++) off[i + 1] = off[i] + (size_t)n * lw[i + 1] * lh[i + 1] * kCout[i];
        ConvParams p{};
        p.n = n;
        p.mode = l == 0 ? 1 : 0;
        p.la = la;
        p.lc = lc;
        p.lumaStride = (int64_t)w0 * h0;
        p.in = l == 0 ? nullptr : acts + off[l - 1];
        p.Cin = kCin[l];
        p.inW = lw[l];
        p.inH = lh[l];
        p.out = acts + off[l];
        p.Cout = kCout[l];
        p.outW = lw[l + 1];
        p.outH = lh[l + 1];
        p.w = weights + layerOffset(l);
        p.b = p.w + kCout[l] * 9 * kCin[l];
        p.stride = kStride[l];
        p.relu = l < kNetLayers - 1 ? 1 : 0;
        if (grads) {
            p.dy = grads + off[l];
            p.dx = l == 0 ? nullptr : grads + off[l - 1];
        }
        if (fp8 && fp8On && fp8Ready) {
            p.fp8 = 1;
            p.wq = wq + wqOffset(l);
            p.kp = kpOf(l);
            p.sa = sa + chanOffset(l, true);
            p.sw = sa + chanOffset(l, false);
        }
        p.gx = convFwdTiles(p);
        return p;
    }

    static int kpOf(int l) { return (9 * kCinL(l) + 31) / 32 * 32; }
    static int kCinL(int l);
    static int kCoutL(int l);
    static size_t wqOffset(int l) {
        size_t o = 0;
        for (int i = 0; i < l; i++) o += (size_t)kCoutL(i) * kpOf(i);
        return o;
    }
    static int chanOffset(int l, bool in) {   // sa: the Cin of every layer, then sw: the Cout of every layer
        int o = 0, total = 0;
        for (int i = 0; i < kNetLayers; i++) total += kCinL(i);
        if (in) {
            for (int i = 0; i < l; i++) o += kCinL(i);
            return o;
        }
        for (int i = 0; i < l; i++) o += kCoutL(i);
        return total + o;
    }
};

int FgrSeedNet::kCinL(int l) { return kCin[l]; }
int FgrSeedNet::kCoutL(int l) { return kCout[l]; }

// OCP FP8 E4M3 on the host: the same steps as fp8e4m3 on the device
static uint8_t fp8e4m3Host(float v) {
    const uint32_t s = v < 0.0f ? 0x80u : 0u;
    const float a = std::fabs(v);
    if (!(a < 448.0f)) return (uint8_t)(s | 0x7Eu);
    if (a < 0.015625f) return (uint8_t)(s | (uint32_t)(int)std::nearbyint(a * 512.0f));
    int e;
    std::frexp(a, &e);
    const int eb = e - 1;
    int M = (int)std::nearbyint(std::ldexp(a, 3 - eb)) - 8, E = eb + 7;
    if (M == 8) {
        M = 0;
        E++;
    }
    if (E > 15 || (E == 15 && M == 7)) return (uint8_t)(s | 0x7Eu);
    return (uint8_t)(s | (uint32_t)(E << 3) | (uint32_t)M);
}

static bool launchFwd(const ConvParams& p, hipStream_t s) {
    hipLaunchKernelGGL(seed_conv_kernel, dim3((unsigned)p.gx, (unsigned)p.n), dim3(256), 0, s, p);
    return hipGetLastError() == hipSuccess;
}

extern "C" uint32_t fgr_seednet_param_count(void) { return (uint32_t)paramCount(); }

extern "C" FgrSeedNet* fgr_seednet_create(int32_t device, uint64_t seed) {
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    FgrSeedNet* net = new FgrSeedNet();
    net->device = device;
    const int n = paramCount();
    for (float** p : {&net->w, &net->g, &net->m, &net->v})
        if (hipMalloc(p, (size_t)n * 4) != hipSuccess) {
            *p = nullptr;
            delete net;
            return nullptr;
        }
    if (hipMalloc(&net->loss, 4) != hipSuccess) {
        net->loss = nullptr;
        delete net;
        return nullptr;
    }
    // He-normal weights (std sqrt(2 / fan_in)), zero biases; the last layer scaled down (small initial proposals)
    std::mt19937_64 rng(seed);
    std::vector<float> h(n, 0.0f);
    for (int l = 0; l < kNetLayers; l++) {
        const int fan = 9 * kCin[l];
        std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / fan) * (l == kNetLayers - 1 ? 0.1 : 1.0));
        float* wl = h.data() + layerOffset(l);
        for (int i = 0; i < kCout[l] * fan; i++) wl[i] = (float)nd(rng);
    }
    hipMemcpy(net->w, h.data(), (size_t)n * 4, hipMemcpyHostToDevice);
    hipMemset(net->m, 0, (size_t)n * 4);
    hipMemset(net->v, 0, (size_t)n * 4);
    hipMemset(net->g, 0, (size_t)n * 4);
    return net;
}

extern "C" void fgr_seednet_destroy(FgrSeedNet* net) { delete net; }

extern "C" int fgr_seednet_get_weights(const FgrSeedNet* net, float* host, uint32_t n) {
    if (!net || !host || n != (uint32_t)paramCount()) return 0;
    hipSetDevice(net->device);
    return hipMemcpy(host, net->w, (size_t)n * 4, hipMemcpyDeviceToHost) == hipSuccess;
}

extern "C" int fgr_seednet_set_weights(FgrSeedNet* net, const float* host, uint32_t n) {
    if (!net || !host || n != (uint32_t)paramCount()) return 0;
    hipSetDevice(net->device);
    net->steps = 0;
    hipMemset(net->m, 0, (size_t)n * 4);
    hipMemset(net->v, 0, (size_t)n * 4);
    return hipMemcpy(net->w, host, (size_t)n * 4, hipMemcpyHostToDevice) == hipSuccess;
}

extern "C" int fgr_seednet_forward(FgrSeedNet* net, void* stream, const float* la, const float* lc, int32_t w,
                                   int32_t h, int32_t n, float* out) {
    if (!net || !la || !lc || !out || w < 8 || h < 8 || n < 1) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (!net->grow(w, h, n)) return 0;
    for (int l = 0; l < kNetLayers; l++) {
        ConvParams p = net->layer(l, la, lc, w, h, n, net->act, nullptr, net->w, true);
        if (l == kNetLayers - 1) p.out = out;
        if (!launchFwd(p, s)) return 0;
    }
    return 1;
}

// FP8 calibration: a BF16 forward over the batch, then per layer the largest |input| of each channel c, the scale
// sa[c] = max / 448 (the largest value maps onto E4M3's largest); the weights with sa folded in, W'[co][k] = W[co][k]
// sa[ci], get per output channel sw[co] = max_k |W'| / 448 and are stored as e4m3(W' / sw).
extern "C" int fgr_seednet_calibrate_fp8(FgrSeedNet* net, void* stream, const float* la, const float* lc, int32_t w,
                                         int32_t h, int32_t n) {
    if (!net || !la || !lc || w < 8 || h < 8 || n < 1) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (!net->grow(w, h, n)) return 0;
    int nIn = 0, nOut = 0;
    for (int l = 0; l < kNetLayers; l++) {
        nIn += kCin[l];
        nOut += kCout[l];
    }
    if (!net->sa && hipMalloc(&net->sa, (size_t)(nIn + nOut) * 4) != hipSuccess) {
        net->sa = nullptr;
        return 0;
    }
    const size_t wqBytes = FgrSeedNet::wqOffset(kNetLayers);
    if (!net->wq && hipMalloc(&net->wq, wqBytes) != hipSuccess) {
        net->wq = nullptr;
        return 0;
    }
    for (int l = 0; l < kNetLayers; l++) {
        const ConvParams p = net->layer(l, la, lc, w, h, n, net->act, nullptr, net->w);
        if (!launchFwd(p, s)) return 0;
        hipLaunchKernelGGL(conv_absmax_kernel, dim3((unsigned)p.Cin), dim3(256), 0, s, p,
                           net->sa + FgrSeedNet::chanOffset(l, true));
    }
    std::vector<float> sa(nIn + nOut), wts(paramCount());
    if (hipMemcpyAsync(sa.data(), net->sa, (size_t)nIn * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(wts.data(), net->w, wts.size() * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return 0;
    for (int i = 0; i < nIn; i++) sa[i] = sa[i] > 0.0f ? sa[i] / 448.0f : 1.0f;
    std::vector<uint8_t> wq(wqBytes, 0);
    for (int l = 0; l < kNetLayers; l++) {
        const int K = 9 * kCin[l], kp = FgrSeedNet::kpOf(l);
        const float* W = wts.data() + layerOffset(l);
        const float* sal = sa.data() + FgrSeedNet::chanOffset(l, true);
        float* swl = sa.data() + FgrSeedNet::chanOffset(l, false);
        uint8_t* q = wq.data() + FgrSeedNet::wqOffset(l);
        std::vector<float> wp(K);
        for (int co = 0; co < kCout[l]; co++) {
            float m = 0.0f;
            for (int k = 0; k < K; k++) {
                wp[k] = W[co * K + k] * sal[k % kCin[l]];
                m = std::max(m, std::fabs(wp[k]));
            }
            swl[co] = m > 0.0f ? m / 448.0f : 1.0f;
            for (int k = 0; k < K; k++) q[co * kp + k] = fp8e4m3Host(wp[k] / swl[co]);
        }
    }
    if (hipMemcpyAsync(net->sa, sa.data(), sa.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(net->wq, wq.data(), wqBytes, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return 0;
    net->fp8Ready = true;
    return 1;
}

extern "C" int fgr_seednet_set_fp8(FgrSeedNet* net, int on) {
    if (!net || (on && !net->fp8Ready)) return 0;
    net->fp8On = on != 0;
    return 1;
}

extern "C" int fgr_seednet_fp8_scales(const FgrSeedNet* net, float* sa, float* sw) {
    if (!net || !net->fp8Ready) return 0;
    int nIn = 0, nOut = 0;
    for (int l = 0; l < kNetLayers; l++) {
        nIn += kCin[l];
        nOut += kCout[l];
    }
    hipSetDevice(net->device);
    if (sa && hipMemcpy(sa, net->sa, (size_t)nIn * 4, hipMemcpyDeviceToHost) != hipSuccess) return 0;
    if (sw && hipMemcpy(sw, net->sa + nIn, (size_t)nOut * 4, hipMemcpyDeviceToHost) != hipSuccess) return 0;
    return 1;
}

// forward + loss + backward into net->g; the loss into net->loss (device)
static bool gradients(FgrSeedNet* net, hipStream_t s, const float* la, const float* lc, int w, int h, int n,
                      const float* target) {
    if (!net->grow(w, h, n)) return false;
    for (int l = 0; l < kNetLayers; l++)
        if (!launchFwd(net->layer(l, la, lc, w, h, n, net->act, nullptr, net->w), s)) return false;
    const ConvParams last = net->layer(kNetLayers - 1, la, lc, w, h, n, net->act, net->grad, net->w);
    const int count = n * last.outW * last.outH * 2;
    float* dpred = net->grad + (last.dy - net->grad);   // the last layer's output gradient
    hipLaunchKernelGGL(mse_kernel, dim3(1), dim3(256), 0, s, last.out, target, dpred, net->loss, count);
    for (int l = kNetLayers - 1; l >= 0; l--) {
        ConvParams p = net->layer(l, la, lc, w, h, n, net->act, net->grad, net->w);
        p.dw = net->g + layerOffset(l);
        p.db = p.dw + kCout[l] * 9 * kCin[l];
        p.dwPart = net->part;
        const int R = n * p.outW * p.outH, nSplit = (R + kDwRows - 1) / kDwRows;
        const int tiles = ((p.Cout + 15) / 16) * ((9 * p.Cin + 15) / 16);
        hipLaunchKernelGGL(conv_dw_kernel, dim3((unsigned)((tiles + 3) / 4), (unsigned)nSplit), dim3(256), 0, s, p);
        hipLaunchKernelGGL(conv_dw_reduce_kernel, dim3((unsigned)p.Cout), dim3(256), 0, s, p, nSplit);
        if (l > 0) {
            const int dt = ((p.inW * p.inH + 15) / 16) * ((p.Cin + 15) / 16);
            hipLaunchKernelGGL(conv_dx_kernel, dim3((unsigned)((dt + 3) / 4), (unsigned)n), dim3(256), 0, s, p);
        }
        if (hipGetLastError() != hipSuccess) return false;
    }
    return true;
}

extern "C" int fgr_seednet_gradients(FgrSeedNet* net, void* stream, const float* la, const float* lc, int32_t w,
                                     int32_t h, int32_t n, const float* target, float* grad, float* loss) {
    if (!net || !la || !lc || !target || w < 8 || h < 8 || n < 1) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (!gradients(net, s, la, lc, w, h, n, target)) return 0;
    if (grad && hipMemcpyAsync(grad, net->g, (size_t)paramCount() * 4, hipMemcpyDeviceToDevice, s) != hipSuccess) return 0;
    if (loss && hipMemcpyAsync(loss, net->loss, 4, hipMemcpyDeviceToDevice, s) != hipSuccess) return 0;
    return 1;
}

extern "C" int fgr_seednet_train_step(FgrSeedNet* net, void* stream, const float* la, const float* lc, int32_t w,
                                      int32_t h, int32_t n, const float* target, float lr, float* loss) {
    if (!net || !la || !lc || !target || w < 8 || h < 8 || n < 1 || !(lr > 0.0f)) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (!gradients(net, s, la, lc, w, h, n, target)) return 0;
    net->steps++;
    const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, net->steps)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, net->steps)));
    const int np = paramCount();
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, net->w, net->g, net->m, net->v,
                       np, lr, c1, c2);
    if (loss && hipMemcpyAsync(loss, net->loss, 4, hipMemcpyDeviceToDevice, s) != hipSuccess) return 0;
    return hipGetLastError() == hipSuccess;
}

// The coarsest-level field of each image: the level's search with seed zero (or, given proposals, the better of zero
// and the proposal) and radius R, then median3 -- the kernels and parameters of the estimate's coarsest step. With
// seeds null this is the plain field, the training target.
extern "C" int fgr_seednet_search_field(void* stream, const float* la, const float* lc, int32_t w, int32_t h, int32_t n,
                                        int32_t R, float lam, const float* seeds, float* raw, float* out) {
    if (!la || !lc || !raw || !out || w < 8 || h < 8 || n < 1 || R < 1 || R > 12) return 0;
    hipStream_t s = (hipStream_t)stream;
    const int bw = (w + 7) / 8, bh = (h + 7) / 8;
    for (int i = 0; i < n; i++) {
        SearchParams sp{};
        sp.la = la + (int64_t)i * w * h;
        sp.lc = lc + (int64_t)i * w * h;
        sp.w = w;
        sp.h = h;
        sp.bw = bw;
        sp.bh = bh;
        sp.R = R;
        sp.lam = lam;
        sp.out = (float2*)raw + (int64_t)i * bw * bh;
        sp.extraSeed = seeds ? (const float2*)seeds + (int64_t)i * bw * bh : nullptr;
        sp.gx = (bw * bh + 3) / 4;
        const int pass = R <= 2 ? kPassSearch2 : (R <= 6 ? kPassSearch6 : kPassSearch12);
        if (launchSingle(pass, &sp, sp.gx, 1, s) != hipSuccess) return 0;
        Median3Params m3{};
        m3.in = sp.out;
        m3.out = (float2*)out + (int64_t)i * bw * bh;
        m3.bw = bw;
        m3.bh = bh;
        m3.gx = (bw + 15) / 16;
        if (launchSingle(kPassMedian3, &m3, m3.gx, (bh + 15) / 16, s) != hipSuccess) return 0;
    }
    return 1;
}

// ------------------------------------------------------------------------------------------------------ inference

extern "C" int fgr_set_seed_net(FgrInterpolator* h, FgrSeedNet* net) {
    if (!h) return 0;
    if (net && net->device != h->device) return 0;
    h->net = net;
    return h->allocSeedBuffers() ? 1 : 0;
}

bool FgrInterpolator::allocSeedBuffers() {
    if (seedAct) hipFree(seedAct);
    seedAct = nullptr;
    if (!net || lv.empty()) return true;
    const Level& c = lv.back();
    int lw[kNetLayers + 1], lh[kNetLayers + 1];
    netLayerSizes(c.w, c.h, lw, lh);
    size_t need = 0;
    for (int l = 0; l < kNetLayers - 1; l++) need += (size_t)lw[l + 1] * lh[l + 1] * kCout[l];
    hipSetDevice(device);
    return hipMalloc(&seedAct, need * 4) == hipSuccess;
}

// the CNN's launches on the coarsest level, before its search (stages 190-194); the proposals into fieldSeed
void FgrInterpolator::planSeeds(std::vector<PlanOp>& plan) const {
    const Level& c = lv.back();
    float2* F = (float2*)field;
    for (int l = 0; l < kNetLayers; l++) {
        ConvParams p = net->layer(l, luma + c.offA, luma + c.offC, c.w, c.h, 1, seedAct, nullptr, net->w, true);
        if (l == kNetLayers - 1) p.out = (float*)(F + fieldSeed);
        PlanOp op;
        op.pass = kPassSeedConv;
        op.stage = 190 + l;
        op.gx = p.gx;
        op.gy = 1;
        op.stamp = l == kNetLayers - 1 ? "seeds" : nullptr;
        op.set(p);
        plan.push_back(op);
    }
}
