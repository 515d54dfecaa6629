// Learned seed proposals for the coarsest search level (this library's addition; no counterpart in the reference).
//
// A 5-layer 3x3 CNN on the coarsest-level luma pair (a, c): channels 2 -> 16 -> 32 -> 64 -> 64 -> 2, stride 2 in
// layers 1-3, padding 1, ReLU after layers 1-4. Its output grid is ceil(w / 8) x ceil(h / 8), one vector per 8x8 block
// of that level, in that level's pixels. The vector is rounded half-even and offered to the level's search as a second
// seed candidate beside zero (the plain seed): the exact SAD search picks the better of the two (the proposal pays the
// seed penalty lambda, as every non-own candidate does) and then searches +-R around it, so a wrong proposal costs
// nothing but the candidate's SAD, and the exact SAD comparison against the plain seed decides.
//
// Every convolution -- forward, and both gradients -- is an implicit GEMM on the matrix cores with FP32 accumulation.
// Training (forward, mean squared error against the plain search's coarsest-level field, backward, Adam) keeps FP32
// activations, weights, gradients and Adam state: a wavefront owns a 16 x 16 output tile and gathers its operands
// straight from the NHWC activations and the weights, rounding them to bf16 (round to nearest even), 32-deep K steps
// on v_mfma_f32_16x16x32_bf16. Inference (the proposals) stores each layer's activations once, quantised for the next
// layer -- BF16, or OCP FP8 E4M3 with per-channel power-of-two scales -- so that a lane's operand of one K step is a
// single 8- or 16-byte load, with the weights staged in LDS (v_mfma_f32_16x16x32_bf16 / _fp8_fp8).
//
// Layouts: activations NHWC float32 [n][h][w][C]; layer l's weights W_l [Cout][9 Cin] with k = (ky * 3 + kx) * Cin + ci,
// then its bias b_l [Cout]; all layers' parameters in one float array in layer order.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <random>
#include <type_traits>
#include <vector>
#include "../../include/fgr.h"
#include "runtime.hpp"
#include "v1_device.hpp"
#include "v1_launch.hpp"
#include "seednet.hpp"

namespace fgr {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// two float32 -> two bf16 (round to nearest even; v_cvt_pk_bf16_f32), `lo` in the low half
__device__ __forceinline__ uint32_t bf16Pair(float lo, float hi) {
    const f32x2 v = {lo, hi};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// eight float32 -> the BF16 operand of one 16x16x32 MFMA step (k ascending)
__device__ __forceinline__ bf16x8 bf16Op(const float (&v)[8]) {
    const u32x4 u = {bf16Pair(v[0], v[1]), bf16Pair(v[2], v[3]), bf16Pair(v[4], v[5]), bf16Pair(v[6], v[7])};
    return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ f32x4 mfmaBf16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x4 mfmaFp8(uint64_t a, uint64_t b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8((long)a, (long)b, c, 0, 0, 0);
}

// float32 -> OCP FP8 E4M3 (bias 7, largest finite 448) on the conversion instruction (v_cvt_pk_fp8_f32, round to
// nearest even), four values packed a..d from the low byte. The values are clamped to +-448 first, so that larger
// magnitudes saturate to the largest finite value.
__device__ __forceinline__ float sat448(float v) { return fminf(fmaxf(v, -448.0f), 448.0f); }
__device__ __forceinline__ uint32_t fp8Quad(float a, float b, float c, float d) {
    const int lo = __builtin_amdgcn_cvt_pk_fp8_f32(sat448(a), sat448(b), 0, false);
    return (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(sat448(c), sat448(d), lo, true);
}

// ------------------------------------------------------------------------------------ float32 activations (training)

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

// Forward on float32 activations (training and FP8 calibration): out[img][p][co] = relu?(b[co] + sum_k A[p][k]
// W[co][k]), A = the 3x3 patch of input pixel (ox s + kx - 1, oy s + ky - 1), operands rounded to bf16. Wave tile:
// 16 output pixels x 16 output channels, 32-deep K steps (lane quad q holds k = k0 + 8 q .. + 7); workgroup: 4 tiles.
__global__ __launch_bounds__(256) void conv_fwd_kernel(ConvParams P) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int Pn = P.outW * P.outH, K = 9 * P.Cin;
    const int nPT = (Pn + 15) >> 4, nCT = (P.Cout + 15) >> 4;
    const int t = (int)blockIdx.x * 4 + wave;
    if (t >= nPT * nCT) return;
    const int ct = t % nCT, pt = t / nCT, img = (int)blockIdx.y;
    const int row = lane & 15, quad = lane >> 4;
    const int p = pt * 16 + row;
    const int ox = p % P.outW, oy = p / P.outW;
    const int co = ct * 16 + row;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < K; k0 += 32) {
        float av[8], bv[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = k0 + 8 * quad + j;
            av[j] = bv[j] = 0.0f;
            if (k < K) {
                const int tap = k / P.Cin, ci = k - tap * P.Cin;
                const int ky = tap / 3, kx = tap - 3 * ky;
                if (p < Pn) av[j] = inputAt(P, img, ox * P.stride + kx - 1, oy * P.stride + ky - 1, ci);
                if (co < P.Cout) bv[j] = P.w[co * K + k];
            }
        }
        acc = mfmaBf16(bf16Op(av), bf16Op(bv), acc);
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
__global__ __launch_bounds__(256) void conv_dx_kernel(ConvParams P) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int Qn = P.inW * P.inH, K = 9 * P.Cout, KW = 9 * P.Cin;
    const int nPT = (Qn + 15) >> 4, nCT = (P.Cin + 15) >> 4;
    const int t = (int)blockIdx.x * 4 + wave;
    if (t >= nPT * nCT) return;
    const int ct = t % nCT, pt = t / nCT, img = (int)blockIdx.y;
    const int row = lane & 15, quad = lane >> 4;
    const int q = pt * 16 + row;
    const int ix = q % P.inW, iy = q / P.inW;
    const int ci = ct * 16 + row;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < K; k0 += 32) {
        float av[8], bv[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = k0 + 8 * quad + j;
            av[j] = bv[j] = 0.0f;
            if (k < K) {
                const int tap = k / P.Cout, co = k - tap * P.Cout;
                const int ky = tap / 3, kx = tap - 3 * ky;
                const int sy = iy + 1 - ky, sx = ix + 1 - kx;
                if (q < Qn && sx >= 0 && sy >= 0 && sx % P.stride == 0 && sy % P.stride == 0) {
                    const int ox = sx / P.stride, oy = sy / P.stride;
                    if (ox < P.outW && oy < P.outH) av[j] = gradAt(P, img, oy * P.outW + ox, co);
                }
                if (ci < P.Cin) bv[j] = P.w[co * KW + tap * P.Cin + ci];
            }
        }
        acc = mfmaBf16(bf16Op(av), bf16Op(bv), acc);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int qq = pt * 16 + 4 * quad + j, cc = ct * 16 + row;
        if (qq < Qn && cc < P.Cin) P.dx[((int64_t)img * Qn + qq) * P.Cin + cc] = acc[j];
    }
}

// Gradient of the weights, one split of the pixel sum: part[split][co][k] = sum over the split's output pixels r of
// dYm[r][co] A[r][k] (A the forward patch matrix). Wave tile: 16 channels x 16 k; the pixels of all images are one
// reduction axis, cut into splits of kDwRows; the splits are added in order by conv_dw_reduce_kernel.
__global__ __launch_bounds__(256) void conv_dw_kernel(ConvParams P) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int Pn = P.outW * P.outH, K = 9 * P.Cin, R = P.n * Pn;
    const int nCT = (P.Cout + 15) >> 4, nKT = (K + 15) >> 4;
    const int t = (int)blockIdx.x * 4 + wave;
    if (t >= nCT * nKT) return;
    const int ct = t % nCT, kt = t / nCT, split = (int)blockIdx.y;
    const int row = lane & 15, quad = lane >> 4;
    const int co = ct * 16 + row, k = kt * 16 + row;
    int ci = 0, ky = 0, kx = 0;
    if (k < K) {
        const int tap = k / P.Cin;
        ci = k - tap * P.Cin;
        ky = tap / 3;
        kx = tap - 3 * ky;
    }
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int r0 = split * kDwRows, r1 = min(R, r0 + kDwRows);
    for (int rb = r0; rb < r1; rb += 32) {
        float av[8], bv[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int r = rb + 8 * quad + j;
            av[j] = bv[j] = 0.0f;
            if (r < r1) {
                const int img = r / Pn, p = r - img * Pn;
                if (co < P.Cout) av[j] = gradAt(P, img, p, co);
                if (k < K) {
                    const int ox = p % P.outW, oy = p / P.outW;
                    bv[j] = inputAt(P, img, ox * P.stride + kx - 1, oy * P.stride + ky - 1, ci);
                }
            }
        }
        acc = mfmaBf16(bf16Op(av), bf16Op(bv), acc);
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

// BF16 copy of one layer's weights for the quantised forward: wb[co][k] = bf16(W[co][k]) for k < K, zero up to kp
__global__ __launch_bounds__(256) void weights_bf16_kernel(const float* w, uint16_t* wb, int cout, int K, int kp) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= cout * kp) return;
    const int co = i / kp, k = i - co * kp;
    wb[i] = k < K ? (uint16_t)(bf16Pair(w[co * K + k], 0.0f) & 0xFFFFu) : (uint16_t)0;
}

// the FP8 conversion of the quantised forward on n floats (test / debug)
__global__ __launch_bounds__(256) void fp8_convert_kernel(const float* in, uint8_t* out, int n) {
    const int i = (int)(blockIdx.x * 256 + threadIdx.x) * 4;
    if (i >= n) return;
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = i + j < n ? in[i + j] : 0.0f;
    const uint32_t q = fp8Quad(v[0], v[1], v[2], v[3]);
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (i + j < n) out[i + j] = (uint8_t)(q >> (8 * j));
}

// ------------------------------------------------------------------------------ quantised activations (inference)
//
// Each layer reads its input activations as stored by the layer before (NHWC, BF16 or FP8 E4M3 quantised with this
// layer's per-channel scale), and writes its output quantised once for the next layer; the last layer writes float32.
// Layer 1 reads the luma pair and quantises the normalised values as it gathers them. With k = (ky 3 + kx) Cin + ci
// and Cin a multiple of 8, the 8 k values a lane feeds one 32-deep MFMA step are 8 consecutive channels of one input
// pixel: one 8-byte (FP8) or 16-byte (BF16) load. The weights ([Cout][kp], zero past 9 Cin) are staged in LDS once per
// workgroup. A wavefront computes 16 output pixels x every output channel (Cout / 16 accumulators), the weights as the
// MFMA's A operand so that a lane ends up holding 4 consecutive channels of one pixel: one packed store.

template <bool F8> struct QOperand;
template <> struct QOperand<true> { typedef uint64_t T; };   // 8 E4M3
template <> struct QOperand<false> { typedef bf16x8 T; };    // 8 BF16

// the input operand of the K step at k .. k + 7 (k a multiple of 8) for output pixel (ox, oy) of image img
template <bool F8, int CIN, int STRIDE>
__device__ __forceinline__ typename QOperand<F8>::T qInput(const ConvParams& P, int img, int ox, int oy, bool pv,
                                                           int k) {
    constexpr int K = 9 * CIN, EB = F8 ? 1 : 2;
    typedef typename QOperand<F8>::T Op;
    if constexpr (CIN >= 8) {
        Op v = {};
        const int tap = k / CIN, ci = k % CIN;
        const int ky = tap / 3, kx = tap - 3 * ky;
        const int ix = ox * STRIDE + kx - 1, iy = oy * STRIDE + ky - 1;
        if (pv && k < K && ix >= 0 && iy >= 0 && ix < P.inW && iy < P.inH) {
            const uint8_t* src = (const uint8_t*)P.qin + ((((int64_t)img * P.inH + iy) * P.inW + ix) * CIN + ci) * EB;
            v = *(const Op*)src;
        }
        return v;
    } else {
        // the luma pair (CIN 2: k even, so value j is channel j & 1), normalised (v - 128) / 64
        float x[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int kk = k + j, tap = kk >> 1;
            const int ky = tap / 3, kx = tap - 3 * ky;
            const int ix = ox * STRIDE + kx - 1, iy = oy * STRIDE + ky - 1;
            x[j] = 0.0f;
            if (pv && kk < K && ix >= 0 && iy >= 0 && ix < P.inW && iy < P.inH) {
                const float* pl = ((j & 1) ? P.lc : P.la) + (int64_t)img * P.lumaStride;
                x[j] = (pl[iy * P.inW + ix] - 128.0f) * (1.0f / 64.0f);
            }
        }
        if constexpr (F8) {
            const float q0 = P.qsIn[0], q1 = P.qsIn[1];
            const uint32_t lo = fp8Quad(x[0] * q0, x[1] * q1, x[2] * q0, x[3] * q1);
            const uint32_t hi = fp8Quad(x[4] * q0, x[5] * q1, x[6] * q0, x[7] * q1);
            return (uint64_t)lo | ((uint64_t)hi << 32);
        } else {
            return bf16Op(x);
        }
    }
}

// the layer's weights ([COUT][kp] elements, kp * element bytes = RB bytes a row) into LDS rows RS bytes apart, rows
// from COUT to ROWS zero
template <int COUT, int ROWS, int RB, int RS>
__device__ __forceinline__ void stageWeights(const ConvParams& P, uint8_t* wl) {
    constexpr int CH = RB / 16, N = ROWS * CH, IT = (N + 255) / 256;   // 16-byte chunks a row, in all, a thread
    const uint8_t* wg = (const uint8_t*)P.wq;
    u32x4 v[IT];   // every load in flight before the first LDS store
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int i = it * 256 + (int)threadIdx.x, r = i / CH, c = i - r * CH;
        v[it] = u32x4{0u, 0u, 0u, 0u};
        if (i < N && r < COUT) v[it] = *(const u32x4*)(wg + (size_t)r * RB + 16 * c);
    }
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int i = it * 256 + (int)threadIdx.x, r = i / CH, c = i - r * CH;
        if (i < N) *(u32x4*)(wl + r * RS + 16 * c) = v[it];
    }
}

// The first layer (luma pair -> 16 channels, stride 2, 18-deep K in one 32-deep step). A workgroup owns a block of
// kQ0W x kQ0H output pixels: it stages the block's input window of both planes in LDS with coalesced loads,
// normalised and quantised once per value, then every wavefront gathers its operands from LDS for 4 tiles of 16
// pixels (one output row segment each).
constexpr int kQ0W = 32, kQ0H = 8;
constexpr int kQ0InW = 2 * kQ0W + 1, kQ0InH = 2 * kQ0H + 1, kQ0Pitch = 68;

int convQ0Blocks(int outW, int outH) { return ((outW + kQ0W - 1) / kQ0W) * ((outH + kQ0H - 1) / kQ0H); }

template <bool F8>
__device__ __forceinline__ void convQ0(const ConvParams& P, int bx, int img) {
    typedef typename std::conditional<F8, uint8_t, uint16_t>::type E;
    typedef typename QOperand<F8>::T Op;
    constexpr int EB = F8 ? 1 : 2, RS = 32 * EB + 16;
    __shared__ __align__(16) uint8_t wl[16 * RS];
    __shared__ E win[2][kQ0InH][kQ0Pitch];
    const int nbx = (P.outW + kQ0W - 1) / kQ0W;
    const int ox0 = (bx % nbx) * kQ0W, oy0 = (bx / nbx) * kQ0H;
    const int ix0 = 2 * ox0 - 1, iy0 = 2 * oy0 - 1;
    stageWeights<16, 16, 32 * EB, RS>(P, wl);
    const float qs0 = F8 ? P.qsIn[0] : 1.0f, qs1 = F8 ? P.qsIn[1] : 1.0f;
    constexpr int NV = 2 * kQ0InH * kQ0InW, IT = (NV + 255) / 256;
    float lum[IT];   // every load in flight before the first conversion; 128 (normalised 0) outside the image
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int i = it * 256 + (int)threadIdx.x;
        const int c = i / (kQ0InH * kQ0InW), r = (i / kQ0InW) % kQ0InH, x = i % kQ0InW;
        const int ix = ix0 + x, iy = iy0 + r;
        lum[it] = 128.0f;
        if (i < NV && ix >= 0 && iy >= 0 && ix < P.inW && iy < P.inH)
            lum[it] = (c ? P.lc : P.la)[(int64_t)img * P.lumaStride + (int64_t)iy * P.inW + ix];
    }
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int i = it * 256 + (int)threadIdx.x;
        if (i >= NV) break;
        const int c = i / (kQ0InH * kQ0InW), r = (i / kQ0InW) % kQ0InH, x = i % kQ0InW;
        const float v = (lum[it] - 128.0f) * (1.0f / 64.0f);
        if constexpr (F8) win[c][r][x] = (uint8_t)fp8Quad(v * (c ? qs1 : qs0), 0.0f, 0.0f, 0.0f);
        else win[c][r][x] = (uint16_t)bf16Pair(v, 0.0f);
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, row = lane & 15, quad = lane >> 4;
    // the lane's 8 operand values k = 8 quad + j (tap k >> 1 of channel k & 1; zero from k = 18): window offsets
    // relative to the top left input pixel of its output pixel
    int off[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int k = 8 * quad + j, tap = k >> 1, ky = tap / 3, kx = tap - 3 * ky;
        off[j] = k < 18 ? ((k & 1) * kQ0InH + ky) * kQ0Pitch + kx : -1;
    }
    const int Pn = P.outW * P.outH, c0 = 4 * quad;
    // the epilogue's per-channel constants, read once (the output stores could alias them)
    float bias[4], sw[4], qs[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        bias[r] = P.b[c0 + r];
        sw[r] = F8 ? P.sw[c0 + r] : 1.0f;
        qs[r] = F8 ? P.qsOut[c0 + r] : 1.0f;
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int yl = 2 * wave + (t >> 1), xl = 16 * (t & 1) + row;
        const E* base = &win[0][0][0] + (2 * yl) * kQ0Pitch + 2 * xl;
        uint32_t e[8];
#pragma unroll
        for (int j = 0; j < 8; j++) e[j] = off[j] >= 0 ? (uint32_t)base[off[j]] : 0u;
        Op in;
        if constexpr (F8) {
            in = (uint64_t)(e[0] | e[1] << 8 | e[2] << 16 | e[3] << 24) |
                 (uint64_t)(e[4] | e[5] << 8 | e[6] << 16 | e[7] << 24) << 32;
        } else {
            const u32x4 u = {e[0] | e[1] << 16, e[2] | e[3] << 16, e[4] | e[5] << 16, e[6] | e[7] << 16};
            in = __builtin_bit_cast(bf16x8, u);
        }
        const Op w = *(const Op*)(wl + row * RS + 8 * quad * EB);
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        if constexpr (F8) acc = mfmaFp8(w, in, acc);
        else acc = mfmaBf16(w, in, acc);
        const int ox = ox0 + xl, oy = oy0 + yl;
        if (ox >= P.outW || oy >= P.outH) continue;
        // D[row 4 quad + r][col row]: channel 4 quad + r of this pixel
        const int64_t pix = (int64_t)img * Pn + (int64_t)oy * P.outW + ox;
        float v[4];
#pragma unroll
        for (int r = 0; r < 4; r++) v[r] = fmaxf(acc[r] * sw[r] + bias[r], 0.0f);
        if constexpr (F8) {
            *(uint32_t*)((uint8_t*)P.qout + pix * 16 + c0) = fp8Quad(v[0] * qs[0], v[1] * qs[1], v[2] * qs[2], v[3] * qs[3]);
        } else {
            const u32x2 u = {bf16Pair(v[0], v[1]), bf16Pair(v[2], v[3])};
            *(u32x2*)((uint8_t*)P.qout + (pix * 16 + c0) * 2) = u;
        }
    }
}

template <bool F8, int CIN, int COUT, int STRIDE, int TPW>
__device__ __forceinline__ void convQ(const ConvParams& P, int bx, int img) {
    constexpr int EB = F8 ? 1 : 2;
    constexpr int K = 9 * CIN, KS = (K + 31) / 32, KP = 32 * KS, NCT = (COUT + 15) / 16;
    constexpr int RS = KP * EB + 16;   // LDS bytes a weight row (16 more than the row, against bank conflicts)
    constexpr bool LAST = COUT < 16;   // the proposals: float32 out, no ReLU
    typedef typename QOperand<F8>::T Op;
    __shared__ __align__(16) uint8_t wl[NCT * 16 * RS];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, row = lane & 15, quad = lane >> 4;
    const int Pn = P.outW * P.outH, nPT = (Pn + 15) >> 4;
    Op in[KS];
    const auto loadTile = [&](int pt) {
        const int p = pt * 16 + row;
        const bool pv = p < Pn;
        const int ox = pv ? p % P.outW : 0, oy = pv ? p / P.outW : 0;
#pragma unroll
        for (int s = 0; s < KS; s++) in[s] = qInput<F8, CIN, STRIDE>(P, img, ox, oy, pv, 32 * s + 8 * quad);
    };
    loadTile(bx * TPW * 4 + wave);   // the first tile's loads in flight with the weight staging
    stageWeights<COUT, NCT * 16, KP * EB, RS>(P, wl);
    __syncthreads();
    for (int i = 0; i < TPW; i++) {
        const int pt = (bx * TPW + i) * 4 + wave;
        if (pt >= nPT) break;
        const int p = pt * 16 + row;
        const bool pv = p < Pn;
        if (i > 0) loadTile(pt);
        f32x4 acc[NCT];
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) acc[ct] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < KS; s++) {
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) {
                const Op w = *(const Op*)(wl + (ct * 16 + row) * RS + (32 * s + 8 * quad) * EB);
                if constexpr (F8) acc[ct] = mfmaFp8(w, in[s], acc[ct]);
                else acc[ct] = mfmaBf16(w, in[s], acc[ct]);
            }
        }
        if (!pv) continue;
        // D[row 4 quad + r][col row]: channel ct * 16 + 4 quad + r of pixel p
        const int64_t pix = (int64_t)img * Pn + p;
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
            const int c0 = ct * 16 + 4 * quad;
            if constexpr (LAST) {
                if (quad == 0) {
#pragma unroll
                    for (int r = 0; r < COUT; r++) P.out[pix * COUT + r] = acc[ct][r] * (F8 ? P.sw[r] : 1.0f) + P.b[r];
                }
            } else {
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] = fmaxf(acc[ct][r] * (F8 ? P.sw[c0 + r] : 1.0f) + P.b[c0 + r], 0.0f);
                if constexpr (F8) {
                    *(uint32_t*)((uint8_t*)P.qout + pix * COUT + c0) =
                        fp8Quad(v[0] * P.qsOut[c0], v[1] * P.qsOut[c0 + 1], v[2] * P.qsOut[c0 + 2], v[3] * P.qsOut[c0 + 3]);
                } else {
                    const u32x2 u = {bf16Pair(v[0], v[1]), bf16Pair(v[2], v[3])};
                    *(u32x2*)((uint8_t*)P.qout + (pix * COUT + c0) * 2) = u;
                }
            }
        }
    }
}

// the ten quantised layer passes: kPassSeedQ + 2 layer + (FP8 ? 1 : 0)
#define FGR_SEED_Q(name, F8, L)                                                                                     \
    __device__ __forceinline__ void name##Body(const ConvParams& P, int bx, int by) {                               \
        convQ<F8, kNetCin[L], kNetCout[L], kNetStride[L], kSeedTpw[L]>(P, bx, by);                                  \
    }                                                                                                                \
    FGR_PASS(name, ConvParams, name##Body, 256)

__device__ __forceinline__ void seedQ0Bf16Body(const ConvParams& P, int bx, int by) { convQ0<false>(P, bx, by); }
__device__ __forceinline__ void seedQ0Fp8Body(const ConvParams& P, int bx, int by) { convQ0<true>(P, bx, by); }
FGR_PASS(seed_q0_bf16, ConvParams, seedQ0Bf16Body, 256)
FGR_PASS(seed_q0_fp8, ConvParams, seedQ0Fp8Body, 256)
FGR_SEED_Q(seed_q1_bf16, false, 1)
FGR_SEED_Q(seed_q1_fp8, true, 1)
FGR_SEED_Q(seed_q2_bf16, false, 2)
FGR_SEED_Q(seed_q2_fp8, true, 2)
FGR_SEED_Q(seed_q3_bf16, false, 3)
FGR_SEED_Q(seed_q3_fp8, true, 3)
FGR_SEED_Q(seed_q4_bf16, false, 4)
FGR_SEED_Q(seed_q4_fp8, true, 4)

#define FGR_SEED_CASES(launch)                                                                                      \
    launch(0, seed_q0_bf16) launch(1, seed_q0_fp8) launch(2, seed_q1_bf16) launch(3, seed_q1_fp8)                 \
    launch(4, seed_q2_bf16) launch(5, seed_q2_fp8) launch(6, seed_q3_bf16) launch(7, seed_q3_fp8)                 \
    launch(8, seed_q4_bf16) launch(9, seed_q4_fp8)

hipError_t launchSeedSingle(int pass, const void* params, int gx, int gy, hipStream_t s) {
    const dim3 g((unsigned)gx, (unsigned)gy);
    const ConvParams& P = *(const ConvParams*)params;
#define FGR_SEED_SINGLE(i, name)                                                                                    \
    case i: hipLaunchKernelGGL(name##_kernel, g, dim3(256), 0, s, P); break;
    switch (pass - kPassSeedQ) {
        FGR_SEED_CASES(FGR_SEED_SINGLE)
    default: return hipErrorInvalidValue;
    }
#undef FGR_SEED_SINGLE
    return hipGetLastError();
}

hipError_t launchSeedBatch(int pass, const void* params, const uint32_t* work, int n, hipStream_t s) {
    const dim3 g((unsigned)n);
    const ConvParams* P = (const ConvParams*)params;
#define FGR_SEED_BATCH(i, name)                                                                                     \
    case i: hipLaunchKernelGGL(name##_batch_kernel, g, dim3(256), 0, s, P, work); break;
    switch (pass - kPassSeedQ) {
        FGR_SEED_CASES(FGR_SEED_BATCH)
    default: return hipErrorInvalidValue;
    }
#undef FGR_SEED_BATCH
    return hipGetLastError();
}

int convQTiles(int layer, const ConvParams& p) {
    if (layer == 0) return convQ0Blocks(p.outW, p.outH);
    const int nPT = (p.outW * p.outH + 15) / 16, per = 4 * kSeedTpw[layer];
    return (nPT + per - 1) / per;
}

}  // namespace fgr

using namespace fgr;

// ------------------------------------------------------------------------------------------------------ host side

static const int* const kCin = kNetCin;
static const int* const kCout = kNetCout;
static const int* const kStride = kNetStride;

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

// byte offsets of the quantised activations of layers 1-4 for n images of a w0 x h0 pair (room for BF16, each layer
// 256-byte aligned); returns the total
static size_t qActLayout(int w0, int h0, int n, size_t* off) {
    int lw[kNetLayers + 1], lh[kNetLayers + 1];
    netLayerSizes(w0, h0, lw, lh);
    size_t o = 0;
    for (int l = 0; l < kNetLayers - 1; l++) {
        if (off) off[l] = o;
        o += ((size_t)n * lw[l + 1] * lh[l + 1] * kCout[l] * 2 + 255) & ~(size_t)255;
    }
    return o;
}

static int nInChannels() {
    int c = 0;
    for (int l = 0; l < kNetLayers; l++) c += kCin[l];
    return c;
}
static int nOutChannels() {
    int c = 0;
    for (int l = 0; l < kNetLayers; l++) c += kCout[l];
    return c;
}

struct FgrSeedNet {
    int device = 0;
    float* w = nullptr;      // parameters (paramCount floats)
    float* g = nullptr;      // gradients
    float* m = nullptr;      // Adam moments
    float* v = nullptr;
    int steps = 0;
    // training buffers, grown on demand
    float* act = nullptr;    // activations of layers 1..5 (NHWC); the quantised forward's activations use it too
    float* grad = nullptr;   // their gradients
    float* part = nullptr;   // weight-gradient partial sums
    float* loss = nullptr;   // one float
    size_t actCap = 0, partCap = 0;
    // BF16 inference: per layer the weights as bf16 (Cout x kp), refreshed whenever the weights change
    uint16_t* wb = nullptr;
    // FP8 inference: per layer E4M3 weights (Cout x kp bytes); scales: the input scales sa (every layer's Cin), the
    // weight scales sw (every layer's Cout), then qs = 1 / sa (every layer's Cin)
    uint8_t* wq = nullptr;
    float* sa = nullptr;
    bool fp8Ready = false, fp8On = false;

    ~FgrSeedNet() {
        hipSetDevice(device);
        for (float* p : {w, g, m, v, act, grad, part, loss, sa})
            if (p) hipFree(p);
        if (wq) hipFree(wq);
        if (wb) hipFree(wb);
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
        need = std::max(need, (qActLayout(w0, h0, n, nullptr) + 3) / 4);
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

    // the ConvParams of layer l over n images of a w0 x h0 luma pair on float32 activations (training, calibration),
    // activations at acts (layer outputs in order)
    ConvParams layer(int l, const float* la, const float* lc, int w0, int h0, int n, float* acts, float* grads,
                     const float* weights) const {
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
        p.gx = (((p.outW * p.outH + 15) / 16) * ((p.Cout + 15) / 16) + 3) / 4;
        return p;
    }

    bool fp8Active() const { return fp8On && fp8Ready; }

    // the pass of the quantised forward's layer l (BF16, or FP8 when it is on)
    int qPass(int l) const { return kPassSeedQ + 2 * l + (fp8Active() ? 1 : 0); }

    // the ConvParams of layer l of the quantised forward over n images: quantised activations at qact (qActLayout),
    // the last layer's float32 output at out
    ConvParams qlayer(int l, const float* la, const float* lc, int w0, int h0, int n, uint8_t* qact, float* out) const {
        int lw[kNetLayers + 1], lh[kNetLayers + 1];
        netLayerSizes(w0, h0, lw, lh);
        size_t off[kNetLayers];
        qActLayout(w0, h0, n, off);
        const bool f8 = fp8Active();
        ConvParams p{};
        p.n = n;
        p.la = la;
        p.lc = lc;
        p.lumaStride = (int64_t)w0 * h0;
        p.Cin = kCin[l];
        p.inW = lw[l];
        p.inH = lh[l];
        p.Cout = kCout[l];
        p.outW = lw[l + 1];
        p.outH = lh[l + 1];
        p.b = w + layerOffset(l) + kCout[l] * 9 * kCin[l];
        p.stride = kStride[l];
        p.relu = l < kNetLayers - 1 ? 1 : 0;
        p.qin = l == 0 ? nullptr : qact + off[l - 1];
        p.qout = l == kNetLayers - 1 ? nullptr : qact + off[l];
        p.out = l == kNetLayers - 1 ? out : nullptr;
        p.kp = kpOf(l);
        if (f8) {
            p.wq = wq + wqOffset(l);
            p.sw = sa + chanOffset(l, false);
            p.qsIn = sa + qsOffset(0);
            p.qsOut = l == kNetLayers - 1 ? nullptr : sa + qsOffset(l + 1);
        } else {
            p.wq = wb + wqOffset(l);
        }
        p.gx = convQTiles(l, p);
        return p;
    }

    static int kpOf(int l) { return (9 * kCin[l] + 31) / 32 * 32; }
    static size_t wqOffset(int l) {   // elements
        size_t o = 0;
        for (int i = 0; i < l; i++) o += (size_t)kCout[i] * kpOf(i);
        return o;
    }
    static int chanOffset(int l, bool in) {   // sa: the Cin of every layer, then sw: the Cout of every layer
        int o = 0;
        if (in) {
            for (int i = 0; i < l; i++) o += kCin[i];
            return o;
        }
        for (int i = 0; i < l; i++) o += kCout[i];
        return nInChannels() + o;
    }
    static int qsOffset(int l) { return nInChannels() + nOutChannels() + chanOffset(l, true); }

    // the BF16 weights from the float32 ones (after every change of the weights), on stream s
    bool refreshBf16(hipStream_t s) const {
        for (int l = 0; l < kNetLayers; l++) {
            const int kp = kpOf(l), n = kCout[l] * kp;
            hipLaunchKernelGGL(weights_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                               w + layerOffset(l), wb + wqOffset(l), kCout[l], 9 * kCin[l], kp);
        }
        return hipGetLastError() == hipSuccess;
    }
};

// the smallest power of two at or above v (v > 0, finite)
static float pow2AtLeast(float v) {
    int e;
    const float m = std::frexp(v, &e);   // v = m 2^e, m in [0.5, 1)
    return m == 0.5f ? std::ldexp(1.0f, e - 1) : std::ldexp(1.0f, e);
}

// OCP FP8 E4M3 (bias 7, largest finite 448), round to nearest even, saturating: the weights' quantisation on the host
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
    hipLaunchKernelGGL(conv_fwd_kernel, dim3((unsigned)p.gx, (unsigned)p.n), dim3(256), 0, s, p);
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
    if (hipMalloc(&net->loss, 4) != hipSuccess || hipMalloc(&net->wb, FgrSeedNet::wqOffset(kNetLayers) * 2) != hipSuccess) {
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
    if (!net->refreshBf16(nullptr) || hipStreamSynchronize(nullptr) != hipSuccess) {
        delete net;
        return nullptr;
    }
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
    if (hipMemcpy(net->w, host, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess) return 0;
    return net->refreshBf16(nullptr) && hipStreamSynchronize(nullptr) == hipSuccess;
}

extern "C" int fgr_seednet_forward(FgrSeedNet* net, void* stream, const float* la, const float* lc, int32_t w,
                                   int32_t h, int32_t n, float* out) {
    if (!net || !la || !lc || !out || w < 8 || h < 8 || n < 1) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (!net->grow(w, h, n)) return 0;
    for (int l = 0; l < kNetLayers; l++) {
        const ConvParams p = net->qlayer(l, la, lc, w, h, n, (uint8_t*)net->act, out);
        if (launchSeedSingle(net->qPass(l), &p, p.gx, n, s) != hipSuccess) return 0;
    }
    return 1;
}

// FP8 calibration: a float32 forward over the batch, then per layer the largest |input| of each channel c, the scale
// sa[c] = the smallest power of two at or above max / 448 (the largest value maps at most onto E4M3's largest, and
// x / sa is exact); the weights, with sa folded in, W'[co][k] = W[co][k] sa[ci], get per output channel
// sw[co] = max_k |W'| / 448 and are stored as e4m3(W' / sw).
extern "C" int fgr_seednet_calibrate_fp8(FgrSeedNet* net, void* stream, const float* la, const float* lc, int32_t w,
                                         int32_t h, int32_t n) {
    if (!net || !la || !lc || w < 8 || h < 8 || n < 1) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (!net->grow(w, h, n)) return 0;
    const int nIn = nInChannels(), nOut = nOutChannels();
    if (!net->sa && hipMalloc(&net->sa, (size_t)(2 * nIn + nOut) * 4) != hipSuccess) {
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
    std::vector<float> sa(2 * nIn + nOut), wts(paramCount());
    if (hipMemcpyAsync(sa.data(), net->sa, (size_t)nIn * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(wts.data(), net->w, wts.size() * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return 0;
    for (int i = 0; i < nIn; i++) {
        sa[i] = sa[i] > 0.0f ? pow2AtLeast(sa[i] / 448.0f) : 1.0f;
        sa[nIn + nOut + i] = 1.0f / sa[i];
    }
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
    const int nIn = nInChannels(), nOut = nOutChannels();
    hipSetDevice(net->device);
    if (sa && hipMemcpy(sa, net->sa, (size_t)nIn * 4, hipMemcpyDeviceToHost) != hipSuccess) return 0;
    if (sw && hipMemcpy(sw, net->sa + nIn, (size_t)nOut * 4, hipMemcpyDeviceToHost) != hipSuccess) return 0;
    return 1;
}

extern "C" int fgr_debug_fp8_e4m3(void* stream, const float* in, uint8_t* out, uint32_t n) {
    if (!in || !out || n == 0 || n > (1u << 30)) return 0;
    const unsigned groups = (n + 3) / 4;
    hipLaunchKernelGGL(fp8_convert_kernel, dim3((groups + 255) / 256), dim3(256), 0, (hipStream_t)stream, in, out, (int)n);
    return hipGetLastError() == hipSuccess;
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
    if (!net->refreshBf16(s)) return 0;
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
    hipSetDevice(device);
    return hipMalloc(&seedAct, qActLayout(c.w, c.h, 1, nullptr)) == hipSuccess;
}

// the CNN's launches on the coarsest level, before its search (stages 190-194); the proposals into fieldSeed
void FgrInterpolator::planSeeds(std::vector<PlanOp>& plan) const {
    const Level& c = lv.back();
    float2* F = (float2*)field;
    for (int l = 0; l < kNetLayers; l++) {
        const ConvParams p = net->qlayer(l, luma + c.offA, luma + c.offC, c.w, c.h, 1, (uint8_t*)seedAct,
                                         (float*)(F + fieldSeed));
        PlanOp op;
        op.pass = net->qPass(l);
        op.stage = 190 + l;
        op.gx = p.gx;
        op.gy = 1;
        op.stamp = l == kNetLayers - 1 ? "seeds" : nullptr;
        op.set(p);
        plan.push_back(op);
    }
}
