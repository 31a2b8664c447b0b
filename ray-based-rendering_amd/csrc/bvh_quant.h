// Quantization of one axis of a wide node's child boxes, shared by the host quantizer
// (bvh_build.cpp) and the device builder (bvh_gpu.hip) so both produce the same node bits.
// RBR_HD compiles the functions for the device as well when the file is built by hipcc.
#pragma once

#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#define RBR_HD __host__ __device__
#else
#define RBR_HD
#endif

namespace rbr {

// ceil(log2(v)) for finite v > 0, exact: v = m 2^x with m in [0.5, 1).
RBR_HD inline int ceilLog2(double v) {
    int x = 0;
    const double m = frexp(v, &x);
    return m == 0.5 ? x - 1 : x;
}

// Per axis: origin p = the union's lower corner of the live children; scale s = 2^e, the
// smallest power of two (from ceil(log2(extent / 255)) up) with fma(255, s, p) >= the
// union's upper corner, so every child plane lands in 0..255 after outward rounding (each
// plane checked with the same single rounding the device uses: lo' = fma(q_lo, s, p) <= lo,
// hi' = fma(q_hi, s, p) >= hi).  Dead children get q_lo = 255, q_hi = 0.
RBR_HD inline void quantAxis(const float* lo, const float* hi, const bool* live, int W, float& p, int& ebias,
                             uint8_t* qlo, uint8_t* qhi) {
    float plo = INFINITY, phi = -INFINITY;
    for (int k = 0; k < W; ++k)
        if (live[k]) {
            plo = lo[k] < plo ? lo[k] : plo;
            phi = hi[k] > phi ? hi[k] : phi;
        }
    if (!(plo <= phi)) {   // no live child
        p = 0.0f;
        ebias = 127;
        for (int k = 0; k < W; ++k) {
            qlo[k] = 255;
            qhi[k] = 0;
        }
        return;
    }
    p = plo;
    const double ext = double(phi) - double(plo);
    int e = -100;
    if (ext > 0.0) {
        const int c = ceilLog2(ext / 255.0);
        e = c > -100 ? c : -100;
    }
    while (fmaf(255.0f, ldexpf(1.0f, e), p) < phi) ++e;
    const float s = ldexpf(1.0f, e);
    ebias = e + 127;
    for (int k = 0; k < W; ++k) {
        if (!live[k]) {
            qlo[k] = 255;
            qhi[k] = 0;
            continue;
        }
        int a = int(floor((double(lo[k]) - double(p)) / double(s)));
        a = a < 0 ? 0 : (a > 255 ? 255 : a);
        while (a > 0 && fmaf(float(a), s, p) > lo[k]) --a;
        int b = int(ceil((double(hi[k]) - double(p)) / double(s)));
        b = b < 0 ? 0 : (b > 255 ? 255 : b);
        while (b < 255 && fmaf(float(b), s, p) < hi[k]) ++b;
        qlo[k] = uint8_t(a);
        qhi[k] = uint8_t(b);
    }
}

// One quantized node (qNodeDwords(W) dwords, zero-filled by the caller) from a float node
// (bvh.h layouts).
RBR_HD inline void quantizeNode(const float* f, int W, uint32_t* q) {
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
            q[4 + (2 * ax) * q4 + (k >> 2)] |= uint32_t(qlo[k]) << (8 * (k & 3));
            q[4 + (2 * ax + 1) * q4 + (k >> 2)] |= uint32_t(qhi[k]) << (8 * (k & 3));
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
