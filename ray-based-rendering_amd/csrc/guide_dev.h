// Device helpers and host state of the guiding network (guide.hip, render.hip).
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "context.h"
#include "guide.h"
#include "shading.h"

namespace rbr {

typedef short v4s __attribute__((ext_vector_type(4)));

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f)); }
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Two FP32 values -> BF16 pair (lo in bits 0-15), one v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t packBf2(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{lo, hi}, bf16x2));
}

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

// Two FP32 values -> FP16 pair (lo in bits 0-15), round to nearest even, one v_cvt_pk_f16_f32.
__device__ __forceinline__ uint32_t packH2(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{lo, hi}, f16x2));
}

typedef short s16x2 __attribute__((ext_vector_type(2)));

// ReLU of a BF16 pair on its bits: a BF16 number read as a 16-bit integer has the sign of the
// number, so max(bits, 0) keeps the non-negative halves and turns the negative ones (and -0)
// into +0 (one v_pk_max_i16).  Equal to packing fmaxf(x, 0) but for NaN inputs, and it needs
// none of the canonicalising v_max_f32 that fmaxf of a matrix-core result costs.
__device__ __forceinline__ uint32_t reluBf2(uint32_t w) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, w), s16x2{0, 0}));
}

// Features of one 32-wide k-block for lane group g (0..3) of a query with scalars s9: word w
// is the BF16 (sin, cos) pair of local pair q = 4 kb + w (guidePair's order).  Scalar j = q / 3
// runs over frequencies 3g, 3g+1, 3g+2: the first is evaluated directly as
// sin/cos(2 pi frac(2^(k-1) x)) (scaling by a power of two and the fraction are exact; the
// hardware sine takes revolutions), the next two by angle doubling,
//   sin 2a = (2 sin a) cos a,  cos 2a = 1 - (2 sin a) sin a   (three instructions),
// carrying (sin, cos) in st between calls.  The absolute error of cos 2a stays within a few
// FP32 ulps of 1, far below the BF16 rounding of the features.  kb must be a compile-time
// constant after unrolling.
__device__ __forceinline__ u32x4 guideEncodeBlock(const float* s9, int g, int kb, float2& st) {
    u32x4 out;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const int q = 4 * kb + w;
        if (q >= 27) {
            out[w] = 0u;
            continue;
        }
        if (q % 3 == 0) {
            const float r = s9[q / 3] * __builtin_ldexpf(1.0f, 3 * g - 1);
            // r - floor(r) in one v_fract_f32 (which returns the largest float below 1 where
            // that difference rounds to 1, for r just below an integer)
            const float f = __builtin_amdgcn_fractf(r);
            st.x = __builtin_amdgcn_sinf(f);
            st.y = __builtin_amdgcn_cosf(f);
        } else {
            const float t2 = st.x + st.x;
            const float sn = t2 * st.y;
            const float cs = __builtin_fmaf(-t2, st.x, 1.0f);
            st.x = sn;
            st.y = cs;
        }
        out[w] = packBf2(st.x, st.y);
    }
    return out;
}

// Features (sin, cos)(2 pi 2^(k-1) x) of input scalar x (frequency k = 0..11), by the operations
// guideEncodeBlock applies for them: the direct evaluation at frequency 3 (k / 3), then k % 3
// angle doublings.  Equal to that function's FP32 values bit for bit (same operations, same
// order), so the BF16 features match too.  Both doublings are evaluated and one result
// selected, so lanes with different k do not diverge.
__device__ __forceinline__ float2 guideFeaturePair(float x, int k) {
    const int g = k / 3, d = k - 3 * g;
    const float r = x * __builtin_ldexpf(1.0f, 3 * g - 1);
    const float f = __builtin_amdgcn_fractf(r);
    const float s0 = __builtin_amdgcn_sinf(f), c0 = __builtin_amdgcn_cosf(f);
    const float t0 = s0 + s0;
    const float s1 = t0 * c0, c1 = __builtin_fmaf(-t0, s0, 1.0f);
    const float t1 = s1 + s1;
    const float s2 = t1 * c1, c2 = __builtin_fmaf(-t1, s1, 1.0f);
    return make_float2(d == 0 ? s0 : d == 1 ? s1 : s2, d == 0 ? c0 : d == 1 ? c1 : c2);
}

// Per-query guide record written by inference for the shading kernel: 5 x 8 FP16,
// [mu.x of lobes 0..7][mu.y][mu.z][kappa][weight].
constexpr uint32_t kGuideRecBytes = 80;

struct GuideInferArgs {
    GuideDev g;
    // explicit queries: 9 floats each (world position, outward normal, wo)
    const float* queries;
    uint32_t n;
    const uint32_t* nDev;
    // render queue: rays, hits, path slots, jobs
    const float4* rays;
    const uint2* hits;
    const uint32_t* paths;
    const DJob* jobs;
    int njobs;
    uint32_t depth;           // bounce of the queue (jobs guide bounces below their guideDepth)
    uint4* out;               // kGuideRecBytes per query (5 x uint4)
};

struct GuideTrainArgs {
    GuideDev g;
    const float4* samples;    // per sample: (pos, c) (normal, -) (wo, -) (dir, -) [+ stride padding]
    uint32_t stride;          // float4 per sample
    uint32_t n;               // samples in this batch
    // Batch sample s reads record (permAdd + s * permMul) mod total; permMul is a prime above
    // total reduced mod total, coprime to it, so a batch never repeats a record (s < 2^26 and
    // total <= 2^27 keep s * permMul below 2^53).  total = 0: record s.
    uint32_t total, permMul, permAdd;   // permMul already reduced mod total
    double invTotal;                    // 1 / total
    float cmax;               // training targets are clamped to cmax before scaling
    size_t cap;               // samples the activation arrays hold (multiple of 32)
    float scale;              // multiplies c (1 / (batch * mean c))
    uint16_t* act[5];         // A_l (l >= 1), BF16, blocked [cap / 32][K_l][32 samples]
    float* scal;              // the 9 input scalars, FP32, blocked [cap / 32][9][32 samples]:
                              // the weight gradient derives A_0 from them (guideFeaturePair)
    uint16_t* dz[5];          // dZ_l, BF16, blocked [cap / 32][N_l][32 samples]
    float* grad;              // FP32 gradient [kGParams]
    double* loss;
};

__device__ __forceinline__ int findJobD(const DJob* jobs, int njobs, uint32_t p) {
    int j = 0;
    while (j + 1 < njobs && p >= jobs[j + 1].pathBase) ++j;
    return j;
}

__device__ __forceinline__ void scaleQuery(const GuideDev& g, const float* pos, const float* n, const float* wo, float* s9) {
    for (int k = 0; k < 3; ++k) {
        s9[k] = (pos[k] - g.lo[k]) * g.inv[k];
        s9[3 + k] = n[k];
        s9[6 + k] = wo[k];
    }
}

__device__ __forceinline__ bool guideQuery(const GuideInferArgs& a, uint32_t i, float* s9) {
    if (a.queries) {
        const float* q = a.queries + size_t(i) * 9;
        scaleQuery(a.g, q, q + 3, q + 6, s9);
        return true;
    }
    const uint32_t p = a.paths[i];
    const DJob& jb = a.jobs[findJobD(a.jobs, a.njobs, p)];
    if (!(jb.guided & 1u) || a.depth >= jb.guideDepth) return false;
    const uint2 h = a.hits[i];
    if (h.x == 0xFFFFFFFFu) return false;
    const float4 s1 = jb.primShade[2 * size_t(h.x) + 1];
    if (__float_as_int(s1.w) < 0) return false;
    const float4 r0 = a.rays[2 * size_t(i)], r1 = a.rays[2 * size_t(i) + 1];
    const float t = __uint_as_float(h.y);
    const float P[3] = {r0.x + r0.w * t, r0.y + r1.x * t, r0.z + r1.y * t};
    const V ow = hitOutward(jb, h.x, V{r0.x, r0.y, r0.z}, V{r0.w, r1.x, r1.y}, V{P[0], P[1], P[2]});
    const float n[3] = {ow.x, ow.y, ow.z};
    const float wo[3] = {-r0.w, -r1.x, -r1.y};
    scaleQuery(a.g, P, n, wo, s9);
    return true;
}

__device__ __forceinline__ void guideSample(const GuideTrainArgs& a, uint32_t i, float* s9, float* dir, float& c,
                                            float& pb, float& pdraw) {
    uint32_t r = i;
    if (a.total) {
        // i * permMul < 2^53 is exact in FP64; the quotient from the reciprocal is off by at
        // most one, which the two corrections take back
        const double x = double(i) * double(a.permMul), tot = double(a.total);
        double m = x - floor(x * a.invTotal) * tot;
        if (m < 0.0) m += tot;
        if (m >= tot) m -= tot;
        r = uint32_t(m) + a.permAdd;
        if (r >= a.total) r -= a.total;
    }
    const float4* s = a.samples + size_t(r) * a.stride;
    const float4 p = s[0], n = s[1], wo = s[2], d = s[3];
    const float pp[3] = {p.x, p.y, p.z}, nn[3] = {n.x, n.y, n.z}, ww[3] = {wo.x, wo.y, wo.z};
    scaleQuery(a.g, pp, nn, ww, s9);
    dir[0] = d.x;
    dir[1] = d.y;
    dir[2] = d.z;
    c = isfinite(p.w) ? fminf(fmaxf(p.w, 0.0f), a.cmax) : 0.0f;
    pb = isfinite(d.w) ? d.w : 0.0f;          // BSDF density of the direction; <= 0: none
    pdraw = isfinite(wo.w) ? wo.w : 0.0f;     // density the direction was drawn with
}

// log(1 - exp(-x)) for x > 0
__device__ __forceinline__ float log1mexp(float x) { return x < 0.693f ? logf(-expm1f(-x)) : log1pf(-expf(-x)); }

constexpr float kLog2Pi = 1.8378770664093453f;

// Lobe k of the network output o (per lobe: mean 3, log kappa, logit): unit mean, kappa
// clamped to [e^-5, e^8], softmax weight.
__device__ __forceinline__ void decodeLobe(const float* o, int k, GLobe& lobe) {
    float mx = -INFINITY;
    for (int j = 0; j < kGLobes; ++j) mx = fmaxf(mx, o[5 * j + 4]);
    float se = 0.0f;
    for (int j = 0; j < kGLobes; ++j) se += expf(o[5 * j + 4] - mx);
    const float mx0 = o[5 * k], my = o[5 * k + 1], mz = o[5 * k + 2];
    const float ln = fmaxf(sqrtf(mx0 * mx0 + my * my + mz * mz), 1e-6f);
    lobe.mu[0] = mx0 / ln;
    lobe.mu[1] = my / ln;
    lobe.mu[2] = mz / ln;
    lobe.kappa = expf(fminf(fmaxf(o[5 * k + 3], kGLogKappaMin), kGLogKappaMax));
    lobe.w = expf(o[5 * k + 4] - mx) / se;
}

// Max / sum over the four lanes 16 apart that share a sample (lane = 16 part + sample).
__device__ __forceinline__ float partMax(float v) {
    v = fmaxf(v, __shfl_xor(v, 16));
    return fmaxf(v, __shfl_xor(v, 32));
}
__device__ __forceinline__ float partSum(float v) {
    v += __shfl_xor(v, 16);
    return v + __shfl_xor(v, 32);
}

// Loss -c*log q(dir) of one sample, q = s p + (1 - s) pb with s the selection probability of
// o[kGSel] (guide.h) when the sample carries the BSDF density pb > 0 of its direction (the
// renderer draws from the guide with probability s, from the BSDF otherwise), else q = p.
// Gradients: the 10 outputs of lobes 2*part and 2*part+1 (layout per lobe: mu 3, log kappa,
// logit) get those of -c log p, the guide's own cross-entropy, so that a small s does not
// starve the lobes; gsel, for the selection logit, is the derivative of the sample's share of
// the one-sample MIS estimator's second moment, c c0 g / q with p held constant (c0 = c
// before the batch scale, g the density the direction was drawn with; 0 without pb).
// Called by the whole wave: the four lanes of a sample (lane = 16 part + sample) each evaluate
// their two lobes and combine the log-sum-exps of the mixture weights and of the lobe
// densities over the eight lobes with cross-lane max / sum.  Hardware exp / log / rsq where
// only relative accuracy matters (the gradient is rounded to BF16); log(1 - e^-x) keeps
// expm1 / log1p.
__device__ __forceinline__ float outputGrad(const float* o, const float* dir, float c, float c0, float pb, float g,
                                            int part, float* grad, float& gsel) {
    float lg[2], pre[2], il[2], dd[2], kap[2];
    bool clamped[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int k = 2 * part + j;
        lg[j] = o[5 * k + 4];
        il[j] = __builtin_amdgcn_rsqf(
            fmaxf(o[5 * k] * o[5 * k] + o[5 * k + 1] * o[5 * k + 1] + o[5 * k + 2] * o[5 * k + 2], 1e-12f));
        dd[j] = (o[5 * k] * dir[0] + o[5 * k + 1] * dir[1] + o[5 * k + 2] * dir[2]) * il[j];
        const float lraw = o[5 * k + 3];
        clamped[j] = lraw < kGLogKappaMin || lraw > kGLogKappaMax;
        const float l = fminf(fmaxf(lraw, kGLogKappaMin), kGLogKappaMax);
        kap[j] = __expf(l);
        pre[j] = l - kLog2Pi - log1mexp(2.0f * kap[j]) + kap[j] * (dd[j] - 1.0f);
    }
    const float mx = partMax(fmaxf(lg[0], lg[1]));
    const float lse = mx + __logf(partSum(__expf(lg[0] - mx) + __expf(lg[1] - mx)));
    float a[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) a[j] = (lg[j] - lse) + pre[j];
    const float amax = partMax(fmaxf(a[0], a[1]));
    const float logp = amax + __logf(partSum(__expf(a[0] - amax) + __expf(a[1] - amax)));
    // Selection: d(c c0 g / q)/d o_sel = -c c0 g (p - pb) / q^2 * ds/do_sel, ds/do_sel =
    // (kGSelMax - kGSelMin) sigma (1 - sigma); q in logs so that a density far below the other
    // neither overflows nor loses the sum.
    float logq = logp;
    gsel = 0.0f;
    if (pb > 0.0f) {
        const float sg = __builtin_amdgcn_rcpf(1.0f + __expf(-o[kGSel]));
        const float s = kGSelMin + (kGSelMax - kGSelMin) * sg;
        const float x1 = __logf(s) + logp, x2 = __logf(1.0f - s) + __logf(pb);
        const float xm = fmaxf(x1, x2);
        logq = xm + __logf(__expf(x1 - xm) + __expf(x2 - xm));
        const float rq = __expf(-logq);
        gsel = -(c * c0) * g * (__expf(logp) - pb) * rq * rq * (kGSelMax - kGSelMin) * sg * (1.0f - sg);
        if (!isfinite(gsel)) gsel = 0.0f;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int k = 2 * part + j;
        const float r = __expf(a[j] - logp);
        const float w = __expf(lg[j] - lse);
        const float mu[3] = {o[5 * k] * il[j], o[5 * k + 1] * il[j], o[5 * k + 2] * il[j]};
        const float x = 2.0f * kap[j];
        const float ratio = x < 1e-4f ? 1.0f - 0.5f * x : x / expm1f(x);   // 2k e^{-2k} / (1 - e^{-2k})
        const float dl = clamped[j] ? 0.0f : -c * r * (1.0f - ratio + kap[j] * (dd[j] - 1.0f));
        const float gm = -c * r * kap[j] * il[j];
        for (int q = 0; q < 3; ++q) grad[5 * j + q] = gm * (dir[q] - dd[j] * mu[q]);
        grad[5 * j + 3] = dl;
        grad[5 * j + 4] = c * (w - r);
    }
    return isfinite(logq) ? -c * logq : 0.0f;
}

// ---------------------------------------------------------------- sampling the mixture

constexpr uint32_t kRecStride = 6;     // float4 per recorded training sample
constexpr double kTargetClamp = 50.0;  // recorded targets above this multiple of their mean are clamped

// Lobe k of array a (0..2 mean, 3 kappa, 4 weight) of a guide record; k static after unrolling.
__device__ __forceinline__ float guideVal(const uint4* rec, int a, int k) {
    const uint4 v = rec[a];
    const uint32_t w = k < 2 ? v.x : k < 4 ? v.y : k < 6 ? v.z : v.w;
    return __half2float(__ushort_as_half(uint16_t((k & 1) ? (w >> 16) : (w & 0xFFFFu))));
}

// Lobe parameters as both the sampler and the density use them: the FP16 mean renormalised
// in FP32, kappa and the weight as stored.
__device__ __forceinline__ void guideLobe(const uint4* rec, int k, float* mu, float& kap, float& w) {
    const float x = guideVal(rec, 0, k), y = guideVal(rec, 1, k), z = guideVal(rec, 2, k);
    const float inv = 1.0f / sqrtf(fmaxf(x * x + y * y + z * z, 1e-30f));
    mu[0] = x * inv;
    mu[1] = y * inv;
    mu[2] = z * inv;
    kap = guideVal(rec, 3, k);
    w = guideVal(rec, 4, k);
}

// vMF density kappa e^{kappa (mu.w - 1)} / (2 pi (1 - e^{-2 kappa})).
__device__ __forceinline__ float vmfDensity(const float* mu, float kap, const float* w) {
    const float d = mu[0] * w[0] + mu[1] * w[1] + mu[2] * w[2];
    return kap * expf(kap * (d - 1.0f)) / (6.28318530717958647f * -expm1f(-2.0f * kap));
}

// Mixture density at unit direction w, weights renormalised by their float sum exactly as
// guideSampleDir selects lobes.
__device__ __forceinline__ float guideMixPdf(const uint4* rec, const float* w) {
    float p = 0.0f, tot = 0.0f;
#pragma unroll
    for (int k = 0; k < kGLobes; ++k) {
        float mu[3], kap, wk;
        guideLobe(rec, k, mu, kap, wk);
        tot += wk;
        if (wk > 0.0f) p += wk * vmfDensity(mu, kap, w);
    }
    return tot > 0.0f ? p / tot : 0.0f;
}

// Probability of drawing a guided vertex's direction from the guide (one-sample MIS with the
// BSDF): the sum of the stored weights, which inference scales by the network's selection
// probability.  The same value picks the strategy and weights the combined density, so the
// estimator stays unbiased whatever FP16 made of it.
__device__ __forceinline__ float guideSelectProb(const uint4* rec) {
    float tot = 0.0f;
#pragma unroll
    for (int k = 0; k < kGLobes; ++k) tot += guideVal(rec, 4, k);
    return fminf(fmaxf(tot, 0.0f), 1.0f);
}

// Draws a unit direction from the mixture: u0 picks the lobe, (u1, u2) the direction within
// it.  Polar angle by the numerically stable inversion of Jakob 2012 ("Numerically stable
// sampling of the von Mises Fisher distribution on S^2"):
//   1 - cos = -log(u1 + (1 - u1) e^{-2 kappa}) / kappa,  sin from (1 - cos)(1 + cos);
// tangent frame of Duff et al. 2017 ("Building an orthonormal basis, revisited").
__device__ __forceinline__ void guideSampleDir(const uint4* rec, float u0, float u1, float u2, float* w) {
    float tot = 0.0f;
#pragma unroll
    for (int k = 0; k < kGLobes; ++k) tot += guideVal(rec, 4, k);
    float u = u0 * tot;
    int pick = -1, last = 0;
#pragma unroll
    for (int k = 0; k < kGLobes; ++k) {
        const float wk = guideVal(rec, 4, k);
        if (wk > 0.0f) {
            last = k;
            if (pick < 0 && u < wk) pick = k;
        }
        u -= wk;
    }
    if (pick < 0) pick = last;   // rounding at the top of the cumulative sum
    float mu[3] = {0.0f, 0.0f, 1.0f}, kap = 1.0f, wsel = 0.0f;
#pragma unroll
    for (int k = 0; k < kGLobes; ++k)
        if (k == pick) guideLobe(rec, k, mu, kap, wsel);
    float d = -logf(fmaxf(u1 + (1.0f - u1) * expf(-2.0f * kap), 1e-38f)) / kap;   // 1 - cos
    d = fminf(fmaxf(d, 0.0f), 2.0f);
    const float c = 1.0f - d;
    const float s = sqrtf(fmaxf(0.0f, d * (2.0f - d)));
    float sp, cp;
    sincosf(6.28318530717958647f * u2, &sp, &cp);
    const float sg = mu[2] >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sg + mu[2]);
    const float b = mu[0] * mu[1] * a;
    const float t[3] = {1.0f + sg * mu[0] * mu[0] * a, sg * b, -sg * mu[0]};
    const float bb[3] = {b, sg + mu[1] * mu[1] * a, -mu[1]};
    for (int q = 0; q < 3; ++q) w[q] = c * mu[q] + s * (cp * t[q] + sp * bb[q]);
}

struct GuideState {
    bool ready = false, adamInit = false, masterDirty = false, attrSet = false;
    int step = 0;
    float lo[3] = {0, 0, 0}, inv[3] = {1, 1, 1};
    std::vector<float> master;
    DevBuf dW, dWf, dWt, dMaster, dM, dV, dGrad, dLoss, dQ, dOut;
    DevBuf act[5], dz[5], dPartial;
    size_t cap = 0, partialChunks = 0;
    // Samples recorded by the renderer, kRecStride float4 each:
    //   (position, c) (outward normal, path slot) (wo, density the direction was drawn with)
    //   (direction, BSDF density of it) (throughput, -) (L, -)
    // dRecCount holds {appended, finalised}; dCsum the sum of c over finalised records.
    DevBuf dRec, dRecCount, dCsum;
    uint32_t recCap = 0, recReserve = 1u << 22;
    uint64_t trainRng = 0x1234567ull;
};

GuideDev guideDev(const GuideState& g);
int launchGuideInfer(rbr_ctx* ctx, const GuideInferArgs& a, uint32_t maxN);
// One Adam step on n samples.  total = 0: samples[0..n); total > 0: a batch of n records
// drawn without repetition from samples[0..total) by a fresh random permutation.
int guideTrainDevice(rbr_ctx* ctx, const float4* samples, uint32_t stride, uint32_t n, uint32_t total, float scale,
                     float cmax, float lr, double* lossOut);
int guideSetAttributes(rbr_ctx* ctx);
// Allocates the record buffers (recReserve records) if absent.
int guideEnsureRecords(rbr_ctx* ctx, GuideState& g);

}  // namespace rbr
