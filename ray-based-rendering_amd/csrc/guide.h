// Path-guiding network shared by the render kernels and the training kernels (guide.hip).
//
// A small MLP maps a shading point to a mixture of 8 von Mises-Fisher lobes over incident
// directions:
//   input  : 9 scalars (position scaled to [0,1]^3 by the scene box, outward normal, wo),
//            each encoded as sin/cos(2^k * pi * x) for k = 0..11  -> 216 features (padded to 224)
//   hidden : 4 layers of 64 ReLU units
//   output : 40 = 8 lobes x (mean 3, log kappa 1, logit 1)
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

// One vMF lobe as the samplers use it.
struct GLobe {
    float mu[3];
    float kappa;
    float w;
};

struct GuideDev {
    const uint16_t* w;     // BF16 weights [kGWeights]
    const float* b;        // FP32 biases [kGBias]
    float lo[3], inv[3];   // position scaling to [0,1]
};

}  // namespace rbr
