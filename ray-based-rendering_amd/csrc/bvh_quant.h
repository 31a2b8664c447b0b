// Quantization of one axis of a wide node's child boxes, shared by the host quantizer
// (bvh_build.cpp) and the device builder (bvh_gpu.hip) so both produce the same node bits.
// RBR_QUANT_HD compiles the functions for the device as well when the file is built by hipcc.
#pragma once

#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#define RBR_QUANT_HD __host__ __device__
#else
#define RBR_QUANT_HD
#endif

namespace rbr {

// ceil(log2(v)) for finite v > 0, exact: v = m 2^x with m in [0.5, 1).
RBR_QUANT_HD inline int ceilLog2(double v) {
    int x = 0;
    const double m = frexp(v, &x);
    return m == 0.5 ? x - 1 : x;
}

// Offset of the stored plane codes: the traversal reads a plane byte q as the half-precision
// number 1024 + q (the byte placed in the low mantissa bits under the exponent byte 0x64, one
// v_perm_b32 for two children), so the stored origin carries the offset: plane = p' +
// (1024 + q) s with p' = p - 1024 s.
constexpr int kQuantBias = 1024;

// The largest float p' with fma(1024, s, p') <= lo, so that code 0 decodes to at most lo.
RBR_QUANT_HD inline float biasedOrigin(float lo, float s) {
    float p = float(double(lo) - double(kQuantBias) * double(s));
    while (fmaf(float(kQuantBias), s, p) > lo) p = nextafterf(p, -INFINITY);
    return p;
}

// Per axis: scale s = 2^e, the smallest power of two (from ceil(log2(extent / 255)) up) with
// fma(1024 + 255, s, p') >= the union's upper corner, p' = biasedOrigin(the union's lower
// corner of the live children, s); every child plane lands in codes 0..255 after outward
// rounding (each plane checked with the single rounding the device uses: lo' = fma(1024 +
// q_lo, s, p') <= lo, hi' = fma(1024 + q_hi, s, p') >= hi).  Dead children get q_lo = 255,
// q_hi = 0.
RBR_QUANT_HD inline void quantAxis(const float* lo, const float* hi, const bool* live, int W, float& p, int& ebias,
                             uint8_t* qlo, uint8_t* qhi) {
    float plo = INFINITY, phi = -INFINITY;
    for (int k = 0; k < W; ++k)
        if (live[k]) {
            plo = lo[k] < plo ? lo[k] : plo;
            phi = hi[k] > phi ? hi[k] : phi;
        }
    if (!(plo <= phi)) {   // no live child
        p = -float(kQuantBias);
        ebias = 127;
        for (int k = 0; k < W; ++k) {
            qlo[k] = 255;
            qhi[k] = 0;
        }
        return;
    }
    const double ext = double(phi) - double(plo);
    int e = -100;
    if (ext > 0.0) {
        const int c = ceilLog2(ext / 255.0);
        e = c > -100 ? c : -100;
    }
    float s = ldexpf(1.0f, e);
    p = biasedOrigin(plo, s);
    while (fmaf(float(kQuantBias + 255), s, p) < phi) {
        ++e;
        s = ldexpf(1.0f, e);
        p = biasedOrigin(plo, s);
    }
    ebias = e + 127;
    const double base = double(p) + double(kQuantBias) * double(s);
    for (int k = 0; k < W; ++k) {
        if (!live[k]) {
            qlo[k] = 255;
            qhi[k] = 0;
            continue;
        }
        int a = int(floor((double(lo[k]) - base) / double(s)));
        a = a < 0 ? 0 : (a > 255 ? 255 : a);
        while (a > 0 && fmaf(float(kQuantBias + a), s, p) > lo[k]) --a;
        int b = int(ceil((double(hi[k]) - base) / double(s)));
        b = b < 0 ? 0 : (b > 255 ? 255 : b);
        while (b < 255 && fmaf(float(kQuantBias + b), s, p) < hi[k]) ++b;
        qlo[k] = uint8_t(a);
        qhi[k] = uint8_t(b);
    }
}

// Dword of the plane codes of child k on axis ax: children 2g and 2g + 1 share dword
// 4 + ax * W/2 + g, bytes (q_lo[2g], q_lo[2g+1], q_hi[2g], q_hi[2g+1]).
RBR_QUANT_HD inline int qPlaneDword(int W, int ax, int k) { return 4 + ax * (W / 2) + (k >> 1); }
RBR_QUANT_HD inline int qLoShift(int k) { return 8 * (k & 1); }
RBR_QUANT_HD inline int qHiShift(int k) { return 8 * (2 + (k & 1)); }

// One quantized node (qNodeDwords(W) dwords, zero-filled by the caller) from a float node
// (bvh.h layouts).
RBR_QUANT_HD inline void quantizeNode(const float* f, int W, uint32_t* q) {
    const int q4 = W / 4;
    uint32_t ref[8];
    bool live[8];
    for (int k = 0; k < W; ++k) {
        float x = f[6 * W + k];
        uint32_t u;
        __builtin_memcpy(&u, &x, 4);
        ref[k] = u;
        live[k] = u != 0xFFFFFFFFu;
    }
    uint32_t exps = 0, scales[3];
    for (int ax = 0; ax < 3; ++ax) {
        float p;
        int eb;
        uint8_t qlo[8], qhi[8];
        quantAxis(&f[(2 * ax) * W], &f[(2 * ax + 1) * W], live, W, p, eb, qlo, qhi);
        uint32_t pu;
        __builtin_memcpy(&pu, &p, 4);
        q[ax] = pu;
        exps |= uint32_t(eb) << (8 * ax);
        scales[ax] = uint32_t(eb) << 23;   // 2^(eb - 127) as a float (eb in 1..254)
        for (int k = 0; k < W; ++k) {
            q[qPlaneDword(W, ax, k)] |= uint32_t(qlo[k]) << qLoShift(k) | uint32_t(qhi[k]) << qHiShift(k);
        }
    }
    if (W == 4) {
        q[3] = scales[0];
        q[14] = scales[1];
        q[15] = scales[2];
    } else {
        q[3] = exps;
    }
    for (int k = 0; k < W; ++k) q[4 + 6 * q4 + k] = ref[k];
}

}  // namespace rbr
