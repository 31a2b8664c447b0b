// Path-guiding network shared by the render kernels and the training kernels (guide.hip).
//
// A small MLP maps a shading point to a mixture of 8 von Mises-Fisher lobes over incident
// directions:
//   input  : 9 scalars (position scaled to [0,1]^3 by the scene box, outward normal, wo),
//            each encoded as sin/cos(2^k * pi * x) for k = 0..11  -> 216 features (padded to
//            224, in the order of guidePair)
//   hidden : 4 layers of 64 ReLU units
//   output : 40 = 8 lobes x (mean 3, log kappa 1, logit 1), and output 40 the logit of the
//            probability of sampling the guide rather than the BSDF (kGSel)
// Weights live as W[out][in] in BF16 for the matrix cores (with an FP32 master copy and
// FP32 Adam moments); all products accumulate in FP32.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rbr {

constexpr int kGIn = 216, kGInPad = 224, kGHid = 64, kGOut = 40, kGOutPad = 48, kGLobes = 8, kGFreq = 12;
// Parameter layout (FP32 master, same order as the BF16 copies): W0[64][224] W1[64][64]
// W2[64][64] W3[64][64] W4[48][64] then b0[64] b1[64] b2[64] b3[64] b4[48].
constexpr int kGW0 = kGHid * kGInPad, kGW1 = kGHid * kGHid, kGW4 = kGOutPad * kGHid;
constexpr int kGWeights = kGW0 + 3 * kGW1 + kGW4;           // 29696 (incl. padding)
constexpr int kGBias = 4 * kGHid + kGOutPad;                // 304
constexpr int kGParams = kGWeights + kGBias;
constexpr int kGOffW[5] = {0, kGW0, kGW0 + kGW1, kGW0 + 2 * kGW1, kGW0 + 3 * kGW1};
constexpr int kGOffB[5] = {kGWeights, kGWeights + 64, kGWeights + 128, kGWeights + 192, kGWeights + 256};
constexpr float kGLogKappaMin = -5.0f, kGLogKappaMax = 8.0f;
// Output kGSel sets the selection probability s = kGSelMin + (kGSelMax - kGSelMin) sigmoid(o)
// with which the renderer draws a guided vertex's direction from the guide rather than from
// the BSDF (one of layer 4's padding rows, zero at initialisation: s starts at 1/2).  The
// range keeps either strategy at 10 % or more, which bounds the weight of a direction the
// other one covers and the learned one misses.  Only the row's own weights and bias learn it:
// the backward pass leaves it out of the gradient of the hidden layers, which the lobes train.
constexpr int kGSel = kGOut;
constexpr float kGSelMin = 0.1f, kGSelMax = 0.9f;

// Input feature order.  Pair p (0..111) is features 2p = sin(2^k pi s_j), 2p+1 = cos(...);
// guidePair gives (j, k) for it, or returns false for the 4 zero padding pairs.  The order
// follows the matrix-core operand layout: with k-block kb = p / 16 and slot group
// g = (p % 8) / 2, the lane group g owns local pairs q = 4 kb + 2 ((p % 16) / 8) + p % 2 and
// produces scalar j = q / 3 at frequencies 3g, 3g+1, 3g+2 from one sine/cosine by doubling.
__host__ __device__ inline bool guidePair(int p, int& j, int& k) {
    const int kb = p >> 4, rem = p & 15, half = rem >> 3, g = (rem & 7) >> 1, t = rem & 1;
    const int q = 4 * kb + 2 * half + t;
    if (q >= 27) return false;
    j = q / 3;
    k = 3 * g + q % 3;
    return true;
}

// Matrix-core fragment order of the BF16 weights used by inference: layer l, output row,
// input column -> element index.  Fragment (l, mb, kb) = 64 lanes x 8 BF16; lane r + 16 g,
// slot i holds W_l[16 mb + r][32 kb + 4 g + (i & 3) + 16 (i >> 2)].
constexpr int kGFragKB[5] = {7, 2, 2, 2, 2};
constexpr int kGFragOff[5] = {0, 28, 36, 44, 52};   // 58 fragments in all
__host__ __device__ inline int guideFragPos(int layer, int row, int col) {
    const int mb = row >> 4, r = row & 15, kb = col >> 5, c = col & 31;
    const int g = (c & 15) >> 2, i = 4 * (c >> 4) + (c & 3);
    return ((kGFragOff[layer] + mb * kGFragKB[layer] + kb) * 64 + r + 16 * g) * 8 + i;
}

// Transposed fragments of layers 1-4 for the backward pass (dA^T = W^T dZ^T): fragment
// (l, mb, kb), mb over the layer's 4 input blocks of 16, kb over its output blocks of 32 (layer
// 4's 48 outputs padded to 64 with zeros): lane r + 16 g, slot i holds
// W_l[32 kb + 4 g + (i & 3) + 16 (i >> 2)][16 mb + r].
constexpr int kGTFrags = 32;
constexpr int kGTWeights = kGTFrags * 64 * 8;   // 16384 BF16
__host__ __device__ inline int guideTFragPos(int layer, int row, int col) {
    const int kb = row >> 5, c = row & 31, g = (c & 15) >> 2, i = 4 * (c >> 4) + (c & 3);
    const int mb = col >> 4, r = col & 15;
    return ((8 * (layer - 1) + mb * 2 + kb) * 64 + r + 16 * g) * 8 + i;
}

// One vMF lobe as the samplers use it.
struct GLobe {
    float mu[3];
    float kappa;
    float w;
};

struct GuideDev {
    const uint16_t* w;     // BF16 weights [kGWeights], row-major W[out][in]
    const uint16_t* wf;    // the same in matrix-core fragment order (guideFragPos)
    const uint16_t* wt;    // layers 1-4 in transposed fragment order (guideTFragPos)
    const float* b;        // FP32 biases [kGBias]
    float lo[3], inv[3];   // position scaling to [0,1]
};

}  // namespace rbr
