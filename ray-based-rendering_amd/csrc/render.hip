// Wavefront path tracer for gfx950: Integrator::render (integrator.cpp:7-50) with
// PathTracingIntegrator::Li (path_tracing_integrator.cpp:5-138) and
// VolumetricPathIntegrator::Li (volumetric_path_integrator.cpp:6-143), restructured as
// one kernel per stage over a device pool of paths:
//
//   camera  -> [ extend (closest hit) -> shade -> occlusion (any hit) -> gather ] x depth
//           -> accumulate film
//
// * extend / occlusion are the persistent traversal kernels of trace.hip, launched with the
//   queue length read from device memory, so no bounce needs a host round trip;
// * shade writes continuing paths to the next queue and NEE shadow rays to the shadow queue,
//   compacted per wave with a ballot + one atomicAdd per wave;
// * gather adds each path's unoccluded NEE terms in light order (deterministic);
// * several jobs (different scenes, resolutions, spp, depths, media) share the pool: their
//   BVHs are packed into one node arena and every queued ray carries its root node.
// Random numbers are a counter-based hash of (seed, path, bounce, dimension), so results
// do not depend on queue order or pool size.  The estimator is the reference's, including its
// quirks: BSDF-sampled emission counted only at depth 0 / after delta bounces, NEE weighted
// by powerHeuristic(p_light, p_bsdf) for area lights; the volumetric integrator's emission
// on every hit, point-light-only volume NEE and its medium weights.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "context.h"
#include "guide_dev.h"
#include "shading.h"
#include "trace.h"
#include "trace_dev.h"

// Geometry that decides occlusion (hit points, light sample points, shadow-ray origins and
// lengths) must round exactly as the reference does: an emitter quad coplanar with its
// AreaLight puts NEE shadow rays on a knife edge (t_quad ~ dist - EPSILON), so contraction
// stays off and every operation keeps the reference's operand order.
#pragma clang fp contract(off)

namespace rbr {

namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 1.0f / kPi;
constexpr float kTwoPi = 2.0f * kPi;
constexpr float kEpsD = 1e-4f;
constexpr uint32_t kMaxLights = 32;

// ---------------------------------------------------------------- vectors

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 mk(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 operator-(F3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ F3 operator*(F3 a, F3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float len(F3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ F3 normalize(F3 a) {
    float l = len(a);
    return l > 0.0f ? a * (1.0f / l) : mk(0, 0, 0);
}
__device__ __forceinline__ bool black(F3 a) { return a.x <= 0.0f && a.y <= 0.0f && a.z <= 0.0f; }
__device__ __forceinline__ float maxc(F3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
__device__ __forceinline__ F3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

// coordinateSystem (vector.h:282-289) and sampling::Frame (sampler.h:103-118)
struct Frame {
    F3 t, b, n;
    __device__ explicit Frame(F3 nn) : n(nn) {
        if (fabsf(nn.x) > fabsf(nn.y)) t = normalize(mk(-nn.z, 0.0f, nn.x));
        else t = normalize(mk(0.0f, nn.z, -nn.y));
        b = cross(nn, t);
    }
    __device__ F3 toWorld(F3 l) const { return t * l.x + b * l.y + n * l.z; }
    __device__ F3 toLocal(F3 w) const { return mk(dot(w, t), dot(w, b), dot(w, n)); }
};

// ---------------------------------------------------------------- random numbers

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t key;
    __device__ float u(uint32_t dim) const {
        uint64_t h = mix64(key ^ (uint64_t(dim) * 0xD6E8FEB86659FD93ull));
        return float(uint32_t(h >> 40)) * (1.0f / 16777216.0f);
    }
};

__device__ __forceinline__ Rng pathRng(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t pixels) {
    return Rng{mix64(seed * 0xA0761D6478BD642Full + uint64_t(sample) * pixels + pixel)};
}

// Dimensions: camera 0..1; bounce d uses base(d) + {0,1 BSDF, 2 RR, 3 distance, 4,5 phase
// (or the guide's direction), 6 guide-or-BSDF choice, 7 guide lobe, 8+2l, 9+2l light l}.
__device__ __forceinline__ uint32_t dimBase(uint32_t depth) { return 8u + depth * 80u; }

// ---------------------------------------------------------------- Fresnel and lobes

// Weight arithmetic (BSDF values and densities, Fresnel terms, MIS weights) uses the hardware
// reciprocal and square root (1 ulp).  Geometry that decides visibility -- hit points, light
// sample points, shadow-ray origins and lengths -- keeps IEEE division and square root.
__device__ __forceinline__ float frcp(float b) { return __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ F3 fnormalize(F3 a) {
    const float l2 = dot(a, a);
    return l2 > 0.0f ? a * __builtin_amdgcn_rsqf(l2) : mk(0, 0, 0);
}

// fresnelDielectric, fresnel.h:16-32
__device__ float fresnelDielectric(float c, float etaI, float etaT) {
    c = fminf(fmaxf(c, -1.0f), 1.0f);
    if (c < 0.0f) {
        float tmp = etaI;
        etaI = etaT;
        etaT = tmp;
        c = -c;
    }
    const float sinI = fsqrt(fmaxf(0.0f, 1.0f - c * c));
    const float sinT = fdiv(etaI, etaT) * sinI;
    if (sinT >= 1.0f) return 1.0f;
    const float cosT = fsqrt(fmaxf(0.0f, 1.0f - sinT * sinT));
    const float rParl = fdiv(etaT * c - etaI * cosT, etaT * c + etaI * cosT);
    const float rPerp = fdiv(etaI * c - etaT * cosT, etaI * c + etaT * cosT);
    return 0.5f * (rParl * rParl + rPerp * rPerp);
}

__device__ F3 fresnelEval(uint32_t kind, float etaI, float etaT, const float* f0, float c) {
    if (kind == DF_DIELECTRIC) {
        float f = fresnelDielectric(c, etaI, etaT);
        return mk(f, f, f);
    }
    if (kind == DF_SCHLICK) {   // fresnelSchlick, fresnel.h:41-45
        const float m = 1.0f - fminf(fmaxf(fabsf(c), 0.0f), 1.0f);
        const float m5 = m * m * m * m * m;
        return mk(f0[0] + (1.0f - f0[0]) * m5, f0[1] + (1.0f - f0[1]) * m5, f0[2] + (1.0f - f0[2]) * m5);
    }
    return mk(1.0f, 1.0f, 1.0f);
}

__device__ __forceinline__ bool sameHemi(F3 a, F3 b) { return a.z * b.z > 0.0f; }

// GGX (microfacet_bsdf.cpp:9-58)
__device__ float ggxD(float a, F3 wh) {
    const float c2 = wh.z * wh.z;
    if (c2 <= 0.0f) return 0.0f;
    const float a2 = a * a;
    const float d = c2 * (a2 - 1.0f) + 1.0f;
    return fdiv(a2, kPi * d * d);
}
__device__ float ggxLambda(float a, F3 w) {
    const float c2 = w.z * w.z;
    if (c2 <= 0.0f) return 0.0f;
    const float t2 = fdiv(fmaxf(0.0f, 1.0f - c2), c2);
    return 0.5f * (fsqrt(1.0f + a * a * t2) - 1.0f);
}
__device__ float ggxPdf(float a, F3 wo, F3 wh) {
    const float cosO = fabsf(wo.z);
    if (cosO == 0.0f) return 0.0f;
    return fdiv(ggxD(a, wh) * frcp(1.0f + ggxLambda(a, wo)) * fabsf(dot(wo, wh)), cosO);
}
__device__ F3 ggxSample(float a, F3 wo, float u1, float u2) {   // Heitz 2018 VNDF
    const bool flip = wo.z < 0.0f;
    const F3 v = flip ? -wo : wo;
    const F3 vh = fnormalize(mk(a * v.x, a * v.y, v.z));
    const F3 t1 = vh.z < 0.9999f ? fnormalize(cross(mk(0, 0, 1), vh)) : mk(1, 0, 0);
    const F3 t2 = cross(vh, t1);
    const float r = fsqrt(u1);
    const float phi = kTwoPi * u2;
    float sp, cp;
    sincosf(phi, &sp, &cp);
    const float p1 = r * cp;
    float p2 = r * sp;
    const float s = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - s) * fsqrt(fmaxf(0.0f, 1.0f - p1 * p1)) + s * p2;
    const F3 nh = t1 * p1 + t2 * p2 + vh * fsqrt(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2));
    const F3 wh = fnormalize(mk(a * nh.x, a * nh.y, fmaxf(1e-6f, nh.z)));
    return flip ? -wh : wh;
}

__device__ bool halfVecT(const DLobe& l, F3 wo, F3 wi, F3& wh, float& eta) {
    if (sameHemi(wo, wi)) return false;
    if (wo.z == 0.0f || wi.z == 0.0f) return false;
    eta = wo.z > 0.0f ? fdiv(l.etaIn, l.etaOut) : fdiv(l.etaOut, l.etaIn);
    wh = wo + wi * eta;
    if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return false;
    wh = fnormalize(wh);
    if (dot(wo, wh) < 0.0f) wh = -wh;
    return wh.z * wo.z > 0.0f && dot(wi, wh) <= 0.0f;
}

__device__ F3 lobeEval(const DLobe& l, F3 wo, F3 wi) {
    const F3 R = ld3(l.R);
    if (l.type == DL_LAMBERT) return sameHemi(wo, wi) ? R * kInvPi : mk(0, 0, 0);
    if (l.type == DL_MF_REFL) {
        if (!sameHemi(wo, wi)) return mk(0, 0, 0);
        const float cosO = fabsf(wo.z), cosI = fabsf(wi.z);
        if (cosO == 0.0f || cosI == 0.0f) return mk(0, 0, 0);
        F3 wh = wo + wi;
        if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return mk(0, 0, 0);
        wh = fnormalize(wh);
        if (wh.z < 0.0f) wh = -wh;
        const F3 F = fresnelEval(l.fkind, l.fEtaI, l.fEtaT, l.f0, dot(wo, wh));
        const float G = frcp(1.0f + ggxLambda(l.alpha, wo) + ggxLambda(l.alpha, wi));
        return R * F * fdiv(ggxD(l.alpha, wh) * G, 4.0f * cosO * cosI);
    }
    if (l.type == DL_MF_TRANS) {
        F3 wh;
        float eta;
        if (!halfVecT(l, wo, wi, wh, eta)) return mk(0, 0, 0);
        const float dotOH = dot(wo, wh), dotIH = dot(wi, wh);
        const float F = fresnelDielectric(wo.z > 0.0f ? dotOH : -dotOH, l.etaOut, l.etaIn);
        const float denom = dotOH + eta * dotIH;
        if (denom == 0.0f) return mk(0, 0, 0);
        const float G = frcp(1.0f + ggxLambda(l.alpha, wo) + ggxLambda(l.alpha, wi));
        const float v = fdiv(ggxD(l.alpha, wh) * G * fabsf(dotIH * dotOH), fabsf(wi.z * wo.z * denom * denom));
        return R * ((1.0f - F) * v);
    }
    return mk(0, 0, 0);
}

__device__ float lobePdf(const DLobe& l, F3 wo, F3 wi) {
    if (l.type == DL_LAMBERT) return sameHemi(wo, wi) ? fabsf(wi.z) * kInvPi : 0.0f;
    if (l.type == DL_MF_REFL) {
        if (!sameHemi(wo, wi)) return 0.0f;
        F3 wh = wo + wi;
        if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return 0.0f;
        wh = fnormalize(wh);
        if (wh.z < 0.0f) wh = -wh;
        const float dotOH = dot(wo, wh);
        if (dotOH == 0.0f) return 0.0f;
        return fdiv(ggxPdf(l.alpha, wo, wh), 4.0f * fabsf(dotOH));
    }
    if (l.type == DL_MF_TRANS) {
        F3 wh;
        float eta;
        if (!halfVecT(l, wo, wi, wh, eta)) return 0.0f;
        const float dotOH = dot(wo, wh), dotIH = dot(wi, wh);
        const float denom = dotOH + eta * dotIH;
        if (denom == 0.0f) return 0.0f;
        return ggxPdf(l.alpha, wo, wh) * fdiv(fabsf(eta * eta * dotIH), denom * denom);
    }
    return 0.0f;
}

__device__ F3 lobeRho(const DLobe& l, F3 wo) {
    const F3 R = ld3(l.R);
    if (l.type == DL_LAMBERT) return R;
    if (l.type == DL_SPEC_REFL || l.type == DL_MF_REFL) return R * fresnelEval(l.fkind, l.fEtaI, l.fEtaT, l.f0, wo.z);
    return R * (1.0f - fresnelDielectric(wo.z, l.etaOut, l.etaIn));
}

struct BSample {
    F3 wi, f;
    float pdf;
    bool delta;
};

__device__ BSample lobeSample(const DLobe& l, F3 wo, float u1, float u2) {
    BSample r{mk(0, 0, 0), mk(0, 0, 0), 0.0f, false};
    const F3 R = ld3(l.R);
    if (l.type == DL_LAMBERT) {   // cosineHemisphere, sampler.h:73-81
        const float rr = fsqrt(u1);
        float s, c;
        sincosf(kTwoPi * u2, &s, &c);
        r.wi = mk(rr * c, rr * s, fsqrt(fmaxf(0.0f, 1.0f - u1)));
        if (wo.z < 0.0f) r.wi.z = -r.wi.z;
        r.f = R * kInvPi;
        r.pdf = fabsf(r.wi.z) * kInvPi;
        return r;
    }
    if (l.type == DL_SPEC_REFL) {
        const float cosO = fabsf(wo.z);
        if (cosO == 0.0f) return r;
        r.wi = mk(-wo.x, -wo.y, wo.z);
        r.f = R * fresnelEval(l.fkind, l.fEtaI, l.fEtaT, l.f0, wo.z) * frcp(cosO);
        r.pdf = 1.0f;
        r.delta = true;
        return r;
    }
    if (l.type == DL_SPEC_TRANS) {
        if (wo.z == 0.0f) return r;
        const bool entering = wo.z > 0.0f;
        const float eta = entering ? fdiv(l.etaOut, l.etaIn) : fdiv(l.etaIn, l.etaOut);
        const float cosI = fabsf(wo.z);
        const float sin2T = eta * eta * fmaxf(0.0f, 1.0f - cosI * cosI);
        if (sin2T >= 1.0f) return r;
        const float cosT = fsqrt(1.0f - sin2T);
        r.wi = mk(-eta * wo.x, -eta * wo.y, entering ? -cosT : cosT);
        const float fr = fresnelDielectric(wo.z, l.etaOut, l.etaIn);
        r.f = R * fdiv((1.0f - fr) * eta * eta, cosT);
        r.pdf = 1.0f;
        r.delta = true;
        return r;
    }
    if (wo.z == 0.0f) return r;
    const F3 wh = ggxSample(l.alpha, wo, u1, u2);
    F3 wi;
    if (l.type == DL_MF_REFL) {
        wi = wh * (2.0f * dot(wo, wh)) - wo;
        if (!sameHemi(wo, wi)) return r;
    } else {
        const float dotOH = dot(wo, wh);
        if (dotOH <= 0.0f) return r;
        const bool entering = wo.z > 0.0f;
        const float er = entering ? fdiv(l.etaOut, l.etaIn) : fdiv(l.etaIn, l.etaOut);
        const float sin2T = er * er * fmaxf(0.0f, 1.0f - dotOH * dotOH);
        if (sin2T >= 1.0f) return r;
        const float cosT = fsqrt(1.0f - sin2T);
        wi = wh * (er * dotOH - cosT) - wo * er;
        if (sameHemi(wo, wi) || wi.z == 0.0f) return r;
    }
    const float p = lobePdf(l, wo, wi);
    if (p <= 0.0f) return r;
    r.wi = wi;
    r.pdf = p;
    r.f = lobeEval(l, wo, wi);
    return r;
}

// MixtureBSDF (mixture_bsdf.cpp:5-104)
__device__ float effWeight(const DMaterial& m, int i, F3 wo) {
    const float w = m.lobe[i].weight;
    if (!((m.substrateMask >> i) & 1u)) return w;
    const float f0[3] = {1.0f, 1.0f, 1.0f};
    return w * fmaxf(0.0f, 1.0f - fresnelEval(m.sfKind, m.sfEtaI, m.sfEtaT, f0, wo.z).x);
}

// The mixture's per-vertex state for one wo: effective lobe weights w (substrate lobes
// attenuated by the coat's Fresnel) and selection scores s = w * max(rho), total = sum s.  The
// shading kernel computes it once per vertex and keeps it in LDS (element i at [i * kShadeBlock])
// for all of the vertex's light samples and its BSDF sample.
constexpr int kShadeBlock = 256;

struct MixW {
    float* w;
    float* s;
    float total;
};

__device__ MixW mixWeights(const DMaterial& m, F3 wo, float* w, float* s) {
    float total = 0.0f;
    for (uint32_t i = 0; i < m.n; ++i) {
        const float wi = effWeight(m, int(i), wo);
        const float si = wi * maxc(lobeRho(m.lobe[i], wo));
        w[i * kShadeBlock] = wi;
        s[i * kShadeBlock] = si;
        total += si;
    }
    return MixW{w, s, total};
}

__device__ __forceinline__ bool lobeDelta(const DLobe& l) { return l.type == DL_SPEC_REFL || l.type == DL_SPEC_TRANS; }

__device__ F3 mixEval(const DMaterial& m, const MixW& x, F3 wo, F3 wi) {
    F3 f = mk(0, 0, 0);
    for (uint32_t i = 0; i < m.n; ++i) {
        if (lobeDelta(m.lobe[i])) continue;
        f = f + lobeEval(m.lobe[i], wo, wi) * x.w[i * kShadeBlock];
    }
    return f;
}

__device__ float mixPdf(const DMaterial& m, const MixW& x, F3 wo, F3 wi) {
    if (x.total <= 0.0f) return 0.0f;
    float p = 0.0f;
    for (uint32_t i = 0; i < m.n; ++i) {
        if (lobeDelta(m.lobe[i])) continue;
        p += x.s[i * kShadeBlock] * lobePdf(m.lobe[i], wo, wi);
    }
    return fdiv(p, x.total);
}

__device__ BSample mixSample(const DMaterial& m, const MixW& x, F3 wo, float u1, float u2) {
    BSample none{mk(0, 0, 0), mk(0, 0, 0), 0.0f, false};
    if (x.total <= 0.0f) return none;
    float u = u1 * x.total;
    int chosen = int(m.n) - 1;
    for (uint32_t i = 0; i < m.n; ++i) {
        const float si = x.s[i * kShadeBlock];
        if (si <= 0.0f) continue;
        if (u < si || i + 1 == m.n) {
            chosen = int(i);
            break;
        }
        u -= si;
    }
    const float sc = x.s[chosen * kShadeBlock];
    if (sc <= 0.0f) return none;
    const float v1 = fminf(fdiv(u, sc), 0.99999994f);
    const DLobe& l = m.lobe[chosen];
    BSample r = lobeSample(l, wo, v1, u2);
    if (r.pdf <= 0.0f || black(r.f)) return none;
    if (lobeDelta(l)) {
        r.f = r.f * x.w[chosen * kShadeBlock];
        r.pdf *= fdiv(sc, x.total);
        return r;
    }
    r.f = mixEval(m, x, wo, r.wi);
    r.pdf = mixPdf(m, x, wo, r.wi);
    if (r.pdf <= 0.0f) return none;
    return r;
}

// sampling::powerHeuristic (sampler.h:135-149)
__device__ float powerHeuristic(float f, float g) {
    if (isinf(f)) return 1.0f;
    if (isinf(g)) return 0.0f;
    const float m = fmaxf(f, g);
    if (m <= 0.0f) return 0.0f;
    const float im = frcp(m);
    f *= im;
    g *= im;
    return fdiv(f * f, f * f + g * g);
}

// ---------------------------------------------------------------- lights and media

struct LSample {
    F3 wi, Li;
    float pdf, dist;
};

__device__ LSample sampleLight(const DLight& l, F3 p, float u1, float u2) {
    LSample s{mk(0, 0, 0), mk(0, 0, 0), 0.0f, 0.0f};
    if (l.type == RBR_LIGHT_AREA) {   // area_light.cpp:13-39
        const F3 lp = ld3(l.position) + ld3(l.u) * (u1 - 0.5f) + ld3(l.v) * (u2 - 0.5f);
        const F3 to = lp - p;
        const float d2 = dot(to, to);
        s.dist = sqrtf(d2);
        s.wi = to * (1.0f / s.dist);
        const float c = -dot(ld3(l.normal), s.wi);
        if (c <= 0.0f) return s;
        s.pdf = d2 / (c * l.area);
        s.Li = ld3(l.color) * l.intensity;
    } else if (l.type == RBR_LIGHT_POINT) {   // point_light.cpp:5-19
        const F3 to = ld3(l.position) - p;
        const float d2 = dot(to, to);
        s.dist = sqrtf(d2);
        s.wi = to * (1.0f / s.dist);
        s.Li = ld3(l.color) * l.intensity * (1.0f / d2);
        s.pdf = 1.0f;
    } else {   // directional_light.cpp:5-15
        s.wi = -ld3(l.direction);
        s.dist = INFINITY;
        s.Li = ld3(l.color) * l.intensity;
        s.pdf = 1.0f;
    }
    return s;
}

__device__ __forceinline__ float densityScale(const DMedium& m, F3 p) {
    return expf(-m.falloff * fmaxf(0.0f, p.y - m.base));
}

// HomogeneousMedium / ExponentialMedium::transmittance (medium.cpp:35-43, 87-103)
__device__ F3 transmittance(const DMedium& m, F3 o, F3 d, float dist) {
    const F3 st = ld3(m.sa) + ld3(m.ss);
    if (m.type == RBR_MEDIUM_EXPONENTIAL) {
        const float step = dist / 64.0f;
        float od = 0.0f;
        for (int i = 0; i < 64; ++i) od += densityScale(m, o + d * ((float(i) + 0.5f) * step)) * step;
        return mk(expf(-st.x * od), expf(-st.y * od), expf(-st.z * od));
    }
    return mk(expf(-st.x * dist), expf(-st.y * dist), expf(-st.z * dist));
}

// sampleDistance (medium.cpp:9-33, 61-85).
// The exponential medium's march re-uses the same u at every step (medium.cpp:68-84), so
// its steps are a constant delta = -ln(1-u)/avg and a step at height y is accepted iff
// u < exp(-falloff*max(0, y - base)), i.e. iff y < base + ln(1/u)/falloff.  Heights along
// the ray are linear in the step index, so the first accepted step is found in closed form
// instead of by iterating (the reference's loop runs ~maxDist/delta times and never ends
// for a ray with no surface behind it; such a ray ends here with zero weight).
__device__ bool sampleDistance(const DMedium& m, F3 o, F3 d, float maxDist, float u, float& t, F3& w) {
    const F3 st = ld3(m.sa) + ld3(m.ss);
    const float avg = (st.x + st.y + st.z) / 3.0f;
    if (m.type == RBR_MEDIUM_EXPONENTIAL) {
        const float step = -logf(1.0f - u) / avg;
        const F3 ratio = mk(m.ss[0] / st.x, m.ss[1] / st.y, m.ss[2] / st.z);   // sigma_s(p) / sigma_t(p)
        bool accept = false;
        if (step == 0.0f) {   // u == 0: the march stays at t = 0, where u < density always holds
            t = 0.0f;
            accept = true;
        } else if (m.falloff <= 0.0f || u <= 0.0f) {
            t = step;
            accept = true;
        } else {
            const float ystar = m.base + (-logf(u)) / m.falloff;
            const float y1 = o.y + d.y * step;
            if (y1 < ystar) {
                t = step;
                accept = true;
            } else if (d.y < 0.0f) {
                const float k = ceilf((ystar - o.y) / (d.y * step));
                t = fmaxf(k, 1.0f) * step;
                if (o.y + d.y * t >= ystar) t += step;
                accept = true;
            }
        }
        if (accept && t < maxDist) {
            w = ratio;
            return true;
        }
        w = isinf(maxDist) ? mk(0, 0, 0) : transmittance(m, o, d, maxDist);
        return false;
    }
    if (avg < 1e-6f) {
        w = mk(1, 1, 1);
        return false;
    }
    t = -logf(1.0f - u) / avg;
    if (t < maxDist) {
        const F3 tr = transmittance(m, o, d, t);
        w = ld3(m.ss) * tr * (1.0f / (avg * avg));
        return true;
    }
    w = transmittance(m, o, d, maxDist);
    return false;
}

// HenyeyGreensteinPhase (phase_function.cpp:17-47), wo = propagation direction.
__device__ float hgEval(float g, F3 wo, F3 wi) {
    const float c = dot(wo, wi);
    const float den = 1.0f + g * g + 2.0f * g * c;
    return (1.0f - g * g) / (4.0f * kPi * den * sqrtf(den));
}
__device__ F3 hgSample(float g, F3 wo, float u1, float u2) {
    float c;
    if (fabsf(g) < 1e-3f) c = 1.0f - 2.0f * u1;
    else {
        const float sq = (1.0f - g * g) / (1.0f + g - 2.0f * g * u1);
        c = -(1.0f + g * g - sq * sq) / (2.0f * g);
    }
    const float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    float sp, cp;
    sincosf(kTwoPi * u2, &sp, &cp);
    Frame f(wo);
    return f.t * (s * cp) + f.b * (s * sp) + wo * c;
}

// ---------------------------------------------------------------- queues

struct Queues {
    // extension queue (ping-pong): rays, hits, path slot, BVH root
    float4* ray[2];
    uint32_t* path[2];
    uint32_t* root[2];
    uint2* hit;
    // shadow queue
    float4* sray;
    uint32_t* sroot;
    float4* scontrib;
    uint8_t* socc;
    int32_t* sindex;      // [path * maxLights + light] -> shadow entry or -1
    // path state
    float4* T;            // throughput, w = specular-bounce flag
    float4* L;            // radiance
    uint32_t* counts;     // [0],[1] extension queues, [2] shadow queue
    // path guiding (null when no job is guided / records)
    const uint4* guide;   // per extension-queue entry: guide record (5 x uint4, FP16 lobes)
    float4* rec;          // training records, kRecStride float4 each
    uint32_t* recOf;      // [path] record written at this bounce or ~0
    uint32_t* recCount;   // {appended, finalised}
    uint32_t recCap;
    // material-ordered shading (null when off): shade entry i is queue entry perm[i]
    uint32_t* perm;
    uint32_t* sortKey;    // [entry] material bin
    uint32_t* bins;       // [kSortBins] counts, then running offsets
};

constexpr uint32_t kSortBins = 1024;   // material bins (last = miss / no material)


__device__ __forceinline__ uint32_t waveAppend(uint32_t* counter, bool want) {
    const unsigned long long m = __ballot(want ? 1 : 0);
    if (m == 0ull) return 0u;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, uint32_t(__popcll(m)));
    base = __shfl(base, leader);
    return base + uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
}

// Block-wide appends of the shading kernel.  Atomics on one address are serialized (about
// 11 ns apart on MI355X), and with one atomic per wave and queue the shading of tens of
// millions of paths queued behind its own counters; a block of kShadeBlock threads now
// issues one atomic per call (blockAppend), or one per kShadowChunk slots (blockAppendChunk).
// Every thread of the block must make the same sequence of calls.  `sh` is block-shared
// scratch of kApWords words; a call's two barriers order the reuse of `sh` by the next one.
constexpr uint32_t kApWaves = kShadeBlock / 64;
constexpr uint32_t kApWords = kApWaves + 3;
constexpr uint32_t kShadowChunk = 1024;   // >= kShadeBlock: one chunk always holds a call's slots

__device__ __forceinline__ uint32_t blockAppend(uint32_t* counter, bool want, uint32_t* sh) {
    const unsigned long long m = __ballot(want ? 1 : 0);
    const uint32_t wave = threadIdx.x >> 6, lane = __lane_id();
    if (lane == 0) sh[wave] = uint32_t(__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0u;
        for (uint32_t w = 0; w < kApWaves; ++w) {
            const uint32_t c = sh[w];
            sh[w] = tot;
            tot += c;
        }
        sh[kApWaves] = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    return sh[kApWaves] + sh[wave] + uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
}

// Same, with the block's slots taken from a chunk of kShadowChunk slots it reserved earlier
// (chunk[0..1] = next free slot, end; block-shared), reserving the next chunk when the current
// one runs out (a call's slots may straddle the two).  The slots left in the block's last
// chunk are never written by shading; the caller marks them.
__device__ __forceinline__ uint32_t blockAppendChunk(uint32_t* counter, bool want, uint32_t* sh, uint32_t* chunk) {
    const unsigned long long m = __ballot(want ? 1 : 0);
    const uint32_t wave = threadIdx.x >> 6, lane = __lane_id();
    if (lane == 0) sh[wave] = uint32_t(__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0u;
        for (uint32_t w = 0; w < kApWaves; ++w) {
            const uint32_t c = sh[w];
            sh[w] = tot;
            tot += c;
        }
        const uint32_t cur = chunk[0], avail = chunk[1] - cur;
        uint32_t nb = 0u;
        if (tot > avail) {
            nb = atomicAdd(counter, kShadowChunk);
            chunk[0] = nb + (tot - avail);
            chunk[1] = nb + kShadowChunk;
        } else {
            chunk[0] = cur + tot;
        }
        sh[kApWaves] = cur;
        sh[kApWaves + 1] = avail;
        sh[kApWaves + 2] = nb;
    }
    __syncthreads();
    const uint32_t r = sh[wave] + uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
    return r < sh[kApWaves + 1] ? sh[kApWaves] + r : sh[kApWaves + 2] + (r - sh[kApWaves + 1]);
}

__device__ __forceinline__ int findJob(const DJob* jobs, int njobs, uint32_t p) {
    int j = 0;
    while (j + 1 < njobs && p >= jobs[j + 1].pathBase) ++j;
    return j;
}

__device__ __forceinline__ void putRay(float4* q, uint32_t i, F3 o, F3 d, float tmin, float tmax) {
    q[2 * size_t(i)] = make_float4(o.x, o.y, o.z, d.x);
    q[2 * size_t(i) + 1] = make_float4(d.y, d.z, tmin, tmax);
}

// spawnRay (interaction.h:32-39): offset along the ray-facing normal toward the new
// direction; the direction is re-normalized as Ray's constructor does.
__device__ __forceinline__ void spawn(F3 p, F3 n, F3 dir, F3& o, F3& d) {
    F3 off = n * kEpsD;
    if (dot(dir, n) < 0.0f) off = -off;
    o = p + off;
    d = normalize(dir);
}

// ---------------------------------------------------------------- kernels

// Integrator::render's per-sample camera ray (integrator.cpp:22-31, camera.cpp:5-9).
__global__ __launch_bounds__(256) void cameraKernel(const DJob* __restrict__ jobs, int njobs, uint32_t total, Queues q) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i == 0) q.counts[0] = total;
    if (i >= total) return;
    const int j = findJob(jobs, njobs, i);
    const DJob& jb = jobs[j];
    const uint32_t local = i - jb.pathBase;
    const uint32_t pixel = local % jb.pixels;
    const uint32_t sample = jb.sampleBase + local / jb.pixels;
    const Rng rng = pathRng(jb.seed, pixel, sample, jb.pixels);
    const DCamera& c = jb.cam;
    const float px = float(pixel % c.width) + rng.u(0);
    const float py = float(pixel / c.width) + rng.u(1);
    const F3 local3 = mk((px - c.cx) / c.fx, (c.cy - py) / c.fy, 1.0f);
    // normalized in generateIntrinsicRay and again by Ray's constructor (camera.cpp:5-9)
    const F3 d = normalize(normalize(ld3(c.right) * local3.x + ld3(c.up) * local3.y + ld3(c.fwd) * local3.z));
    putRay(q.ray[0], i, ld3(c.pos), d, kEpsD, INFINITY);
    q.path[0][i] = i;
    q.root[0][i] = jb.root;
    q.T[i] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    q.L[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// One bounce of shading for queue entry i (lanes with i >= count only take part in the wave
// collectives).  mixLds: this block's LDS for the per-vertex mixture weights.
__device__ __forceinline__ void shadeEntry(const DJob* __restrict__ jobs, int njobs, uint32_t depth, uint32_t maxLights,
                                           int cur, const Queues& q, uint32_t i, uint32_t count, float* mixLds,
                                           uint32_t* apSh, uint32_t* shadowChunk) {
    const bool active = i < count;
    const int nxt = cur ^ 1;

    uint32_t p = 0;
    int j = 0;
    F3 T = mk(0, 0, 0), L = mk(0, 0, 0);
    bool spec = false;
    // mode: 0 done, 1 surface, 2 volume scatter
    int mode = 0;
    F3 P = mk(0, 0, 0), n = mk(0, 0, 0), outward = mk(0, 0, 0), rd = mk(0, 0, 0);
    int mat = -1;
    Rng rng{0};
    uint32_t base = dimBase(depth);
    bool nee = false;

    const uint32_t src = (active && q.perm) ? q.perm[i] : i;   // queue entry shaded by this lane
    if (active) {
        p = q.path[cur][src];
        j = findJob(jobs, njobs, p);
        const DJob& jb = jobs[j];
        const float4 r0 = q.ray[cur][2 * size_t(src)], r1 = q.ray[cur][2 * size_t(src) + 1];
        const F3 ro = mk(r0.x, r0.y, r0.z);
        rd = mk(r0.w, r1.x, r1.y);
        const uint2 h = q.hit[src];
        const float4 t4 = q.T[p], l4 = q.L[p];
        T = mk(t4.x, t4.y, t4.z);
        spec = t4.w != 0.0f;
        L = mk(l4.x, l4.y, l4.z);
        const uint32_t local = p - jb.pathBase;
        rng = pathRng(jb.seed, local % jb.pixels, jb.sampleBase + local / jb.pixels, jb.pixels);
        for (uint32_t l = 0; l < maxLights; ++l) q.sindex[size_t(p) * maxLights + l] = -1;

        const bool hitSurf = h.x != RBR_MISS_ID;
        const float th = __uint_as_float(h.y);
        bool scattered = false;
        if (jb.integrator == RBR_INTEGRATOR_VOLPATH && jb.medium.type != RBR_MEDIUM_NONE) {
            float ts;
            F3 w;
            scattered = sampleDistance(jb.medium, ro, rd, hitSurf ? th : INFINITY, rng.u(base + 3), ts, w);
            T = T * w;
            if (scattered) {
                P = ro + rd * ts;
                mode = 2;
                nee = jb.useNEE != 0;
            }
        }
        if (!scattered) {
            if (!hitSurf) {
                L = L + T * ld3(jb.bg);
            } else {
                const float4 s1 = jb.primShade[2 * size_t(h.x) + 1];
                mat = __float_as_int(s1.w);
                if (mat >= 0) {
                    P = ro + rd * th;
                    const V ow = hitOutward(jb, h.x, V{ro.x, ro.y, ro.z}, V{rd.x, rd.y, rd.z}, V{P.x, P.y, P.z});
                    outward = mk(ow.x, ow.y, ow.z);
                    const bool front = dot(rd, outward) < 0.0f;
                    n = front ? outward : -outward;
                    const DMaterial& m = jb.mats[mat];
                    const bool vol = jb.integrator == RBR_INTEGRATOR_VOLPATH;
                    if (m.emissive && (vol || depth == 0 || spec)) L = L + T * ld3(m.emission);
                    mode = 1;
                    nee = jb.useNEE && !m.allDelta;
                }
            }
        }
    }

    // Next-event estimation: one shadow ray per light, appended block by block.
    const DJob& jb = jobs[j];
    const bool volInt = active && jb.integrator == RBR_INTEGRATOR_VOLPATH;
    Frame frame(mode == 1 ? outward : mk(0, 0, 1));
    const F3 woL = frame.toLocal(-rd);
    MixW mx{mixLds + threadIdx.x, mixLds + 6 * kShadeBlock + threadIdx.x, 0.0f};
    if (mode == 1) mx = mixWeights(jb.mats[mat], woL, mx.w, mx.s);
    for (uint32_t l = 0; l < maxLights; ++l) {
        bool want = false;
        F3 so = mk(0, 0, 0), sd = mk(0, 0, 0), contrib = mk(0, 0, 0);
        float smax = 0.0f;
        if (active && nee && l < jb.nLights) {
            const DLight& lt = jb.lights[l];
            if (mode == 2) {   // directLightingVolume: point lights only, no offset
                if (lt.type == RBR_LIGHT_POINT) {
                    const F3 to = ld3(lt.position) - P;
                    const float dist = len(to);
                    const F3 wi = to * (1.0f / dist);
                    so = P;
                    sd = normalize(wi);
                    smax = dist;
                    const F3 tr = transmittance(jb.medium, so, sd, dist);
                    const F3 Li = ld3(lt.color) * lt.intensity * (1.0f / (dist * dist));
                    contrib = T * Li * tr * hgEval(jb.medium.g, rd, wi);
                    want = true;
                }
            } else {
                const LSample ls = sampleLight(lt, P, rng.u(base + 8 + 2 * l), rng.u(base + 9 + 2 * l));
                if (ls.pdf > 0.0f && !black(ls.Li)) {
                    const DMaterial& m = jb.mats[mat];
                    const F3 wiL = frame.toLocal(ls.wi);
                    const F3 f = mixEval(m, mx, woL, wiL);
                    if (!black(f)) {
                        const float c = fabsf(dot(ls.wi, n));
                        float w = 1.0f;
                        if (!volInt && jb.useMIS && lt.type == RBR_LIGHT_AREA) w = powerHeuristic(ls.pdf, mixPdf(m, mx, woL, wiL));
                        contrib = T * f * ls.Li * fdiv(c * w, ls.pdf);
                        spawn(P, n, ls.wi, so, sd);
                        smax = ls.dist - kEpsD;
                        want = true;
                    }
                }
            }
        }
        const uint32_t slot = blockAppendChunk(&q.counts[2], want, apSh, shadowChunk);
        if (want) {
            putRay(q.sray, slot, so, sd, kEpsD, smax);
            q.sroot[slot] = jb.root;
            q.scontrib[slot] = make_float4(contrib.x, contrib.y, contrib.z, 0.0f);
            q.sindex[size_t(p) * maxLights + l] = int32_t(slot);
        }
    }

    // Continuation: phase or BSDF sample, Russian roulette, spawn.
    bool cont = false, record = false;
    F3 no = mk(0, 0, 0), nd = mk(0, 0, 0), wiRec = mk(0, 0, 0), Tin = T;
    float pdfRec = 0.0f, pdfDraw = 0.0f;   // BSDF density of the recorded direction, density it was drawn with
    if (mode == 2) {
        nd = normalize(hgSample(jb.medium.g, rd, rng.u(base + 4), rng.u(base + 5)));
        no = P;
        cont = depth + 1 < jb.maxDepth;
    } else if (mode == 1) {
        const DMaterial& m = jb.mats[mat];
        // Guided jobs: one-sample MIS between the guide's vMF mixture (probability s, which
        // the network learns: the record's weights sum to it) and the BSDF; the direction's
        // density is the mixture of both, a delta BSDF sample keeps its discrete probability.
        // All-delta materials are never guided.  Recording jobs write a training sample per
        // guidable vertex, whichever density drew the direction (guided = 2 records plain
        // BSDF-sampled paths), with the BSDF density of that direction.
        const bool guidable = !m.allDelta && depth < jb.guideDepth;
        const bool guided = (jb.guided & 1u) && guidable;
        uint4 gl[5];
        float sel = 0.0f;
        if (guided) {
            for (int k = 0; k < 5; ++k) gl[k] = q.guide[size_t(src) * 5 + k];
            sel = guideSelectProb(gl);
        }
        const bool fromGuide = guided && rng.u(base + 6) < sel;
        BSample bs{mk(0, 0, 0), mk(0, 0, 0), 0.0f, false};
        F3 wi = mk(0, 0, 0);
        float pdf = 0.0f, pdfBsdf = 0.0f;
        if (!fromGuide) {
            bs = mixSample(m, mx, woL, rng.u(base + 0), rng.u(base + 1));
            if (bs.pdf > 0.0f && !black(bs.f)) {
                wi = frame.toWorld(bs.wi);
                pdf = bs.pdf;
                pdfBsdf = bs.pdf;
                if (guided) {
                    const float wa[3] = {wi.x, wi.y, wi.z};
                    pdf = bs.delta ? (1.0f - sel) * bs.pdf : sel * guideMixPdf(gl, wa) + (1.0f - sel) * bs.pdf;
                }
            }
        } else {
            float w3[3];
            guideSampleDir(gl, rng.u(base + 7), rng.u(base + 4), rng.u(base + 5), w3);
            wi = normalize(mk(w3[0], w3[1], w3[2]));
            const F3 wiL = frame.toLocal(wi);
            bs.f = mixEval(m, mx, woL, wiL);
            const float wa[3] = {wi.x, wi.y, wi.z};
            pdfBsdf = mixPdf(m, mx, woL, wiL);
            pdf = sel * guideMixPdf(gl, wa) + (1.0f - sel) * pdfBsdf;
            if (black(bs.f)) pdf = 0.0f;
        }
        if (pdf > 0.0f) {
            T = T * bs.f * fdiv(fabsf(dot(wi, n)), pdf);
            bool alive = true;
            if (depth > 3) {
                const float qq = fmaxf(0.05f, 1.0f - maxc(T));
                if (rng.u(base + 2) < qq) alive = false;
                else T = T * frcp(1.0f - qq);
            }
            if (alive) {
                spec = bs.delta;
                spawn(P, n, wi, no, nd);
                cont = depth + 1 < jb.maxDepth;
                record = guidable && (jb.guided & 2u) && !bs.delta && cont;
                wiRec = wi;
                pdfRec = pdfBsdf;
                pdfDraw = pdf;
            }
        }
    }
    if (active) {
        q.L[p] = make_float4(L.x, L.y, L.z, 0.0f);
        if (cont) q.T[p] = make_float4(T.x, T.y, T.z, spec ? 1.0f : 0.0f);
    }
    if (q.recCount) {
        // Training record of this vertex: the radiance that arrives later along wi is
        // L_final - L(after this vertex's NEE), the gather kernel stores the latter.
        const uint32_t r = blockAppend(q.recCount, record, apSh);
        const bool keep = record && r < q.recCap;
        if (active) q.recOf[p] = keep ? r : 0xFFFFFFFFu;
        if (keep) {
            float4* o = q.rec + size_t(r) * kRecStride;
            o[0] = make_float4(P.x, P.y, P.z, 0.0f);
            o[1] = make_float4(outward.x, outward.y, outward.z, __uint_as_float(p));
            o[2] = make_float4(-rd.x, -rd.y, -rd.z, pdfDraw);
            o[3] = make_float4(wiRec.x, wiRec.y, wiRec.z, pdfRec);
            o[4] = make_float4(Tin.x, Tin.y, Tin.z, 0.0f);
            o[5] = make_float4(L.x, L.y, L.z, 0.0f);
        }
    }
    const uint32_t slot = blockAppend(&q.counts[nxt], cont, apSh);
    if (cont) {
        putRay(q.ray[nxt], slot, no, nd, kEpsD, INFINITY);
        q.path[nxt][slot] = p;
        q.root[nxt][slot] = jb.root;
    }
}

// Persistent shading: the grid (sized to the device, not to the queue, whose length is only
// known on the device) walks the queue in block-sized steps.
__global__ __launch_bounds__(256, 4) void shadeKernel(const DJob* __restrict__ jobs, int njobs, uint32_t depth,
                                                   uint32_t maxLights, int cur, Queues q) {
    __shared__ float mixLds[2 * 6 * kShadeBlock];
    __shared__ uint32_t apSh[kApWords];
    __shared__ uint32_t shadowChunk[2];
    if (threadIdx.x == 0) shadowChunk[0] = shadowChunk[1] = 0u;
    __syncthreads();
    const uint32_t count = q.counts[cur];
    for (uint32_t b = blockIdx.x * 256u; b < count; b += gridDim.x * 256u)
        shadeEntry(jobs, njobs, depth, maxLights, cur, q, b + threadIdx.x, count, mixLds, apSh, shadowChunk);
    // Shadow slots of the block's last chunk that no shading point took: rays that test
    // nothing (no BVH root, empty [tmin, tmax]); no path's shadow index refers to them.
    __syncthreads();
    for (uint32_t t = shadowChunk[0] + threadIdx.x; t < shadowChunk[1]; t += 256u) {
        putRay(q.sray, t, mk(0, 0, 0), mk(0, 0, 1), 1.0f, 0.0f);
        q.sroot[t] = kEmptyD;
    }
}

// Material-ordered shading: bin of each extension-queue entry (its hit's material; misses and
// primitives without material share the last bin) and the count per bin, from per-block
// histograms (one global atomic per block and bin).
__global__ __launch_bounds__(256) void sortKeyKernel(const DJob* __restrict__ jobs, int cur, Queues q) {
    __shared__ uint32_t cnt[kSortBins];
    for (uint32_t t = threadIdx.x; t < kSortBins; t += 256u) cnt[t] = 0u;
    __syncthreads();
    const uint32_t count = q.counts[cur];
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < count; i += gridDim.x * 256u) {
        uint32_t key = kSortBins - 1;
        const uint2 h = q.hit[i];
        if (h.x != RBR_MISS_ID) {
            const int32_t m = __float_as_int(jobs[0].primShade[2 * size_t(h.x) + 1].w);   // arena-global ids
            if (m >= 0) key = uint32_t(m) % (kSortBins - 1);
        }
        q.sortKey[i] = key;
        atomicAdd(&cnt[key], 1u);
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < kSortBins; t += 256u)
        if (cnt[t]) atomicAdd(&q.bins[t], cnt[t]);
}

// Exclusive prefix sum of the bin counts (one block).
__global__ __launch_bounds__(1024) void sortScanKernel(uint32_t* bins) {
    __shared__ uint32_t sh[kSortBins];
    const uint32_t t = threadIdx.x;
    sh[t] = bins[t];
    __syncthreads();
    for (uint32_t off = 1; off < kSortBins; off <<= 1) {
        const uint32_t v = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    bins[t] = t ? sh[t - 1] : 0u;
}

// Scatter in chunks of 4096 entries per block: ranks within the chunk from an LDS histogram,
// one global atomic per (chunk, bin) reserves the chunk's range in the bin.  Atomics on one
// bin are serialized, so with few materials the chunk size sets the scatter's floor (1024-entry
// chunks: ~32 K atomics per hot bin and bounce of a 32 Mi-path round).
constexpr uint32_t kSortChunk = 4096;

__global__ __launch_bounds__(256) void sortScatterKernel(int cur, Queues q) {
    __shared__ uint32_t cnt[kSortBins];
    const uint32_t count = q.counts[cur];
    for (uint32_t c0 = blockIdx.x * kSortChunk; c0 < count; c0 += gridDim.x * kSortChunk) {
        for (uint32_t t = threadIdx.x; t < kSortBins; t += 256u) cnt[t] = 0u;
        __syncthreads();
        uint32_t key[kSortChunk / 256], rank[kSortChunk / 256];
#pragma unroll
        for (uint32_t k = 0; k < kSortChunk / 256; ++k) {
            const uint32_t i = c0 + k * 256u + threadIdx.x;
            key[k] = 0u;
            rank[k] = 0u;
            if (i < count) {
                key[k] = q.sortKey[i];
                rank[k] = atomicAdd(&cnt[key[k]], 1u);
            }
        }
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < kSortBins; t += 256u)
            if (cnt[t]) cnt[t] = atomicAdd(&q.bins[t], cnt[t]);   // now the chunk's base in bin t
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kSortChunk / 256; ++k) {
            const uint32_t i = c0 + k * 256u + threadIdx.x;
            if (i < count) q.perm[cnt[key[k]] + rank[k]] = i;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void zeroBinsKernel(uint32_t* bins) {
    for (uint32_t t = threadIdx.x; t < kSortBins; t += 256u) bins[t] = 0u;
}

// Adds each shaded path's unoccluded NEE terms in light order.
__device__ __forceinline__ void gatherEntry(uint32_t maxLights, const Queues& q, uint32_t p) {
    float4 l4 = q.L[p];
    for (uint32_t l = 0; l < maxLights; ++l) {
        const int32_t s = q.sindex[size_t(p) * maxLights + l];
        if (s >= 0 && !q.socc[s]) {
            const float4 c = q.scontrib[s];
            l4.x += c.x;
            l4.y += c.y;
            l4.z += c.z;
        }
    }
    q.L[p] = l4;
    if (q.recOf) {
        const uint32_t r = q.recOf[p];
        if (r != 0xFFFFFFFFu) q.rec[size_t(r) * kRecStride + 5] = l4;
    }
}

__global__ __launch_bounds__(256) void gatherKernel(uint32_t maxLights, int cur, Queues q) {
    const uint32_t count = q.counts[cur];
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < count; i += gridDim.x * 256u)
        gatherEntry(maxLights, q, q.path[cur][i]);
}

// Turns this round's records into training targets once their paths have finished:
// c = lum((L_final - L_vertex) / T_vertex) = lum(f cos L_i) / pdf of the sampled direction,
// so that minimising E[-c log p(wi)] over wi ~ pdf drives p towards f cos L_i.
__global__ __launch_bounds__(256) void guideFinalizeKernel(float4* __restrict__ rec, const float4* __restrict__ L,
                                                           const uint32_t* __restrict__ counts, uint32_t cap,
                                                           double* csum) {
    const uint32_t lo = counts[1], hi = min(counts[0], cap);
    float acc = 0.0f;
    for (uint32_t r = lo + blockIdx.x * 256u + threadIdx.x; r < hi; r += gridDim.x * 256u) {
        float4* o = rec + size_t(r) * kRecStride;
        const uint32_t p = __float_as_uint(o[1].w);
        const float4 lf = L[p], ls = o[5], t = o[4];
        const float dx = t.x > 0.0f ? (lf.x - ls.x) / t.x : 0.0f;
        const float dy = t.y > 0.0f ? (lf.y - ls.y) / t.y : 0.0f;
        const float dz = t.z > 0.0f ? (lf.z - ls.z) / t.z : 0.0f;
        float c = 0.2126f * dx + 0.7152f * dy + 0.0722f * dz;
        if (!(c > 0.0f) || !isfinite(c)) c = 0.0f;
        o[0].w = c;
        acc += c;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
    if ((threadIdx.x & 63) == 0 && acc > 0.0f) atomicAdd(csum, double(acc));
}

__global__ void guideFinalizeEndKernel(uint32_t* counts, uint32_t cap) { counts[1] = min(counts[0], cap); }

// ---------------------------------------------------------------- debug integrators

// Hit point and ray-facing normal of extension-queue entry i (the surface part of shadeEntry:
// SurfaceInteraction's normal after setFaceNormal).
__device__ __forceinline__ bool debugHit(const DJob& jb, const Queues& q, int cur, uint32_t i, F3& P, F3& n, float& t) {
    const uint2 h = q.hit[i];
    if (h.x == RBR_MISS_ID) return false;
    const float4 r0 = q.ray[cur][2 * size_t(i)], r1 = q.ray[cur][2 * size_t(i) + 1];
    const F3 ro = mk(r0.x, r0.y, r0.z), rd = mk(r0.w, r1.x, r1.y);
    t = __uint_as_float(h.y);
    P = ro + rd * t;
    const V ow = hitOutward(jb, h.x, V{ro.x, ro.y, ro.z}, V{rd.x, rd.y, rd.z}, V{P.x, P.y, P.z});
    const F3 outward = mk(ow.x, ow.y, ow.z);
    n = dot(rd, outward) < 0.0f ? outward : -outward;
    return true;
}

// NormalIntegrator / DepthIntegrator / AOIntegrator::Li (debug_integrators.cpp:5-60) for the
// camera rays of debug jobs.  Normal and depth are final here; AO sends its rays (cosine-
// weighted about the normal, spawned as spawnRay does, length aoDistance) through the
// shadow queue, traced for the closest hit, and debugGatherKernel weighs them.
__global__ __launch_bounds__(256) void debugShadeKernel(const DJob* __restrict__ jobs, int njobs, uint32_t maxSlots,
                                                        int cur, Queues q) {
    const uint32_t count = q.counts[cur];
    for (uint32_t b = blockIdx.x * 256u; b < count; b += gridDim.x * 256u) {
        const uint32_t i = b + threadIdx.x;
        const bool active = i < count;
        uint32_t p = 0;
        int j = 0;
        bool hit = false;
        F3 P = mk(0, 0, 0), n = mk(0, 0, 1), L = mk(0, 0, 0);
        float t = 0.0f;
        Rng rng{0};
        if (active) {
            p = q.path[cur][i];
            j = findJob(jobs, njobs, p);
            const DJob& jb = jobs[j];
            const uint32_t local = p - jb.pathBase;
            rng = pathRng(jb.seed, local % jb.pixels, jb.sampleBase + local / jb.pixels, jb.pixels);
            hit = debugHit(jb, q, cur, i, P, n, t);
            if (jb.integrator == RBR_INTEGRATOR_NORMAL) {
                if (hit) L = (n + mk(1.0f, 1.0f, 1.0f)) * 0.5f;
            } else if (jb.integrator == RBR_INTEGRATOR_DEPTH) {
                float v = 1.0f;
                if (hit) {
                    float d = (t - jb.depthNear) / (jb.depthFar - jb.depthNear);
                    d = fmaxf(0.0f, fminf(1.0f, d));
                    v = 1.0f - d;
                }
                L = mk(v, v, v);
            } else if (!hit) {
                L = mk(1.0f, 1.0f, 1.0f);   // AO miss
            }
            q.L[p] = make_float4(L.x, L.y, L.z, 0.0f);
            for (uint32_t k = 0; k < maxSlots; ++k) q.sindex[size_t(p) * maxSlots + k] = -1;
        }
        const DJob& jb = jobs[j];
        const bool ao = active && hit && jb.integrator == RBR_INTEGRATOR_AO;
        const Frame frame(n);
        for (uint32_t k = 0; k < maxSlots; ++k) {
            const bool want = ao && k < jb.aoSamples;
            F3 so = mk(0, 0, 0), sd = mk(0, 0, 1);
            if (want) {   // sampling::cosineHemisphere (sampler.h:73-81) about the normal
                const float u1 = rng.u(16 + 2 * k), u2 = rng.u(17 + 2 * k);
                const float rr = sqrtf(u1);
                float sn, cs;
                sincosf(kTwoPi * u2, &sn, &cs);
                spawn(P, n, frame.toWorld(mk(rr * cs, rr * sn, sqrtf(fmaxf(0.0f, 1.0f - u1)))), so, sd);
            }
            const uint32_t slot = waveAppend(&q.counts[2], want);
            if (want) {
                putRay(q.sray, slot, so, sd, kEpsD, jb.aoDistance);
                q.sroot[slot] = jb.root;
                q.sindex[size_t(p) * maxSlots + k] = int32_t(slot);
            }
        }
    }
}

// AO: 1 - (sum over the path's rays that hit something of 1 - t / distance) / samples.
__global__ __launch_bounds__(256) void debugGatherKernel(const DJob* __restrict__ jobs, int njobs, uint32_t maxSlots,
                                                         int cur, Queues q, const uint2* __restrict__ shit) {
    const uint32_t count = q.counts[cur];
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < count; i += gridDim.x * 256u) {
        const uint32_t p = q.path[cur][i];
        const DJob& jb = jobs[findJob(jobs, njobs, p)];
        if (jb.integrator != RBR_INTEGRATOR_AO || q.sindex[size_t(p) * maxSlots] < 0) continue;
        float occlusion = 0.0f;
        for (uint32_t k = 0; k < jb.aoSamples; ++k) {
            const uint2 h = shit[q.sindex[size_t(p) * maxSlots + k]];
            if (h.x != RBR_MISS_ID) occlusion += 1.0f - (__uint_as_float(h.y) / jb.aoDistance);
        }
        const float ao = 1.0f - (occlusion / float(jb.aoSamples));
        q.L[p] = make_float4(ao, ao, ao, 0.0f);
    }
}

__global__ __launch_bounds__(256) void resetCountsKernel(uint32_t* counts, int nxt) {
    counts[nxt] = 0u;
    counts[2] = 0u;
}

// Traversal census: an evenly strided sample of up to k entries of a queue of *count rays
// (entries j * max(1, count / k), j < min(k, count)), with their BVH roots, and the count.
__global__ __launch_bounds__(256) void censusKernel(const float4* __restrict__ rays, const uint32_t* __restrict__ roots,
                                                    const uint32_t* __restrict__ count, uint32_t k, float4* outRays,
                                                    uint32_t* outRoots, uint32_t* outCount) {
    const uint32_t n = *count;
    const uint32_t stride = n > k ? n / k : 1u;
    const uint32_t m = n < k ? n : k;
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j == 0) *outCount = n;
    if (j >= m) return;
    const size_t src = size_t(j) * stride;
    outRays[2 * size_t(j)] = rays[2 * src];
    outRays[2 * size_t(j) + 1] = rays[2 * src + 1];
    outRoots[j] = roots[src];
}

// Adds this round's samples of one job to its film (and optional moments), in sample order.
// ---------------------------------------------------------------- Whitted

constexpr uint32_t kWhittedMaxDepth = 16;
constexpr uint32_t kWhittedStack = kWhittedMaxDepth + 2;   // pending rays of a depth-first walk

// refract of whitted_integrator.cpp:102-110 (eta = etaI / etaT, normal facing the incident ray)
__device__ __forceinline__ bool whittedRefract(F3 incident, F3 n, float eta, F3& out) {
    const float cosI = -dot(incident, n);
    const float sin2T = eta * eta * (1.0f - cosI * cosI);
    if (sin2T > 1.0f) return false;
    const float cosT = sqrtf(1.0f - sin2T);
    out = normalize(incident * eta + n * (eta * cosI - cosT));
    return true;
}

// WhittedIntegrator::Li (whitted_integrator.cpp:7-110) for one camera sample per thread.  The
// reference's recursion is a tree whose value is linear in its leaves, so a thread walks it
// depth first with an explicit stack of (ray, weight, depth): each surface adds its classic
// local shading (classic_shading.h: ambient, then per light a shadow ray and Lambert + Blinn-
// Phong terms) times the weight of the path to it, a smooth conductor pushes its mirror ray
// (weight x baseColor x metallic), a transmissive dielectric pushes the reflected ray
// (x (1-m) t F) and, below total internal reflection, the refracted ray
// (x (1-m) t (1-F) baseColor); misses, surfaces without material and rays deeper than the
// maximum depth return the background, emitters their emission.  Rays are traced inline with
// the traversal of trace_dev.h (closest hit; any hit for shadow rays).
template <int W, bool Q, bool MIXED>
__global__ __launch_bounds__(256) void whittedKernel(const DJob* __restrict__ jobs, int njobs, uint32_t total,
                                                     TraceArgs a, Queues q) {
    __shared__ uint2 ldsStack[kLdsDepthDefault * 256];
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= total) return;
    const int j = findJob(jobs, njobs, i);
    const DJob& jb = jobs[j];
    const uint32_t local = i - jb.pathBase;
    const uint32_t pixel = local % jb.pixels;
    const uint32_t sample = jb.sampleBase + local / jb.pixels;
    const Rng rng = pathRng(jb.seed, pixel, sample, jb.pixels);
    const DCamera& c = jb.cam;
    const float px = float(pixel % c.width) + rng.u(0);
    const float py = float(pixel / c.width) + rng.u(1);
    const F3 local3 = mk((px - c.cx) / c.fx, (c.cy - py) / c.fy, 1.0f);
    const F3 d0 = normalize(normalize(ld3(c.right) * local3.x + ld3(c.up) * local3.y + ld3(c.fwd) * local3.z));

    tdev::PrivateSpillStack<kLdsDepthDefault, 256> st;
    st.lds = ldsStack;
    F3 so[kWhittedStack], sd[kWhittedStack], sw[kWhittedStack];
    uint32_t sdep[kWhittedStack];
    int n = 0;
    so[0] = ld3(c.pos);
    sd[0] = d0;
    sw[0] = mk(1.0f, 1.0f, 1.0f);
    sdep[0] = 0;
    n = 1;
    F3 L = mk(0, 0, 0);
    const F3 bg = ld3(jb.bg);
    uint32_t vtx = 0;   // shaded vertices so far: random dimensions of their light samples
    while (n > 0) {
        --n;
        const F3 o = so[n], d = sd[n], w = sw[n];
        const uint32_t depth = sdep[n];
        if (depth > jb.maxDepth) {
            L = L + w * bg;
            continue;
        }
        tdev::DRay r;
        r.o = V{o.x, o.y, o.z};
        r.d = V{d.x, d.y, d.z};
        r.tmin = kEpsD;
        r.tmax = INFINITY;
        tdev::setupRay(r);
        float t;
        uint32_t id;
        bool occ;
        st.sp = 0;
        tdev::traceRay<W, Q, false, MIXED>(a, r, jb.root, st, t, id, occ);
        if (id == RBR_MISS_ID) {
            L = L + w * bg;
            continue;
        }
        const int32_t mi = __float_as_int(jb.primShade[2 * size_t(id) + 1].w);
        if (mi < 0) {
            L = L + w * bg;
            continue;
        }
        const DMaterial& m = jb.mats[mi];
        if (m.emissive) {
            L = L + w * ld3(m.emission);
            continue;
        }
        const F3 P = o + d * t;
        const V ow = hitOutward(jb, id, r.o, r.d, V{P.x, P.y, P.z});
        const F3 outward = mk(ow.x, ow.y, ow.z);
        const bool front = dot(d, outward) < 0.0f;
        const F3 nn = front ? outward : -outward;
        const F3 wo = -d;
        const F3 base = ld3(m.base);
        const float mm = fminf(fmaxf(m.metallic, 0.0f), 1.0f);
        const float tt = fminf(fmaxf(m.transmission, 0.0f), 1.0f);
        const bool smooth = m.roughness < 0.03f;
        const float dw = (1.0f - mm) * (1.0f - tt);
        const float cw = smooth ? 0.0f : mm;
        if (dw > 0.0f || cw > 0.0f) {   // classic::shade (classic_shading.h:27-58)
            F3 res = base * (0.1f * dw) * ld3(jb.ambient);
            const float al = fmaxf(m.roughness * m.roughness, 1e-2f);
            const float shin = fmaxf(2.0f / (al * al) - 2.0f, 1.0f);
            const float f0 = fabsf(m.ior - 1.0f) > 1e-3f ? ((m.ior - 1.0f) / (m.ior + 1.0f)) * ((m.ior - 1.0f) / (m.ior + 1.0f))
                                                         : 0.0f;
            const F3 spec = mk(f0 * dw, f0 * dw, f0 * dw) + base * cw;
            const float specNorm = (shin + 8.0f) / (8.0f * 3.14159265358979323846f);
            for (uint32_t l = 0; l < jb.nLights; ++l) {
                const uint32_t dim = 2u + 2u * (vtx * kMaxLights + l);
                const LSample ls = sampleLight(jb.lights[l], P, rng.u(dim), rng.u(dim + 1));
                if (ls.pdf <= 0.0f) continue;
                F3 sro, srd;
                spawn(P, nn, ls.wi, sro, srd);
                tdev::DRay sr;
                sr.o = V{sro.x, sro.y, sro.z};
                sr.d = V{srd.x, srd.y, srd.z};
                sr.tmin = kEpsD;
                sr.tmax = ls.dist - kEpsD;
                tdev::setupRay(sr);
                float st2;
                uint32_t sid;
                bool blocked;
                st.sp = 0;
                tdev::traceRay<W, Q, true, MIXED>(a, sr, jb.root, st, st2, sid, blocked);
                if (blocked) continue;
                const float ndl = fmaxf(0.0f, dot(nn, ls.wi));
                if (ndl <= 0.0f) continue;
                F3 shaded = base * (0.7f * dw * ndl);
                const F3 hv = normalize(ls.wi + wo);
                const float ndh = fmaxf(0.0f, dot(nn, hv));
                shaded = shaded + spec * (specNorm * powf(ndh, shin) * ndl);
                res = res + shaded * ls.Li;
            }
            L = L + w * res;
        }
        ++vtx;
        const F3 rdir = d - nn * (2.0f * dot(d, nn));   // reflect (vector.h:277-279)
        if (smooth && mm > 0.0f && n < int(kWhittedStack)) {   // traceReflection
            spawn(P, nn, rdir, so[n], sd[n]);
            sw[n] = w * base * mm;
            sdep[n] = depth + 1;
            ++n;
        }
        if (tt > 0.0f && mm < 1.0f) {   // traceRefraction
            const float cosI = fabsf(dot(d, nn));
            float etaI = 1.0f, etaT = m.ior;
            if (!front) {
                etaI = m.ior;
                etaT = 1.0f;
            }
            const float kr = fresnelDielectric(cosI, etaI, etaT);
            const float share = (1.0f - mm) * tt;
            if (n < int(kWhittedStack)) {
                spawn(P, nn, rdir, so[n], sd[n]);
                sw[n] = w * (share * kr);
                sdep[n] = depth + 1;
                ++n;
            }
            F3 tdir;
            if (kr < 1.0f && whittedRefract(d, nn, etaI / etaT, tdir) && n < int(kWhittedStack)) {
                so[n] = P - nn * kEpsD;
                sd[n] = normalize(tdir);
                sw[n] = w * base * (share * (1.0f - kr));
                sdep[n] = depth + 1;
                ++n;
            }
        }
    }
    q.L[i] = make_float4(L.x, L.y, L.z, 0.0f);
}

template <int W, bool Q>
hipError_t launchWhittedW(const DJob* jobs, int njobs, uint32_t total, const TraceArgs& a, const Queues& q, hipStream_t s) {
    const dim3 g((total + 255) / 256);
    if (a.mixed) hipLaunchKernelGGL((whittedKernel<W, Q, true>), g, dim3(256), 0, s, jobs, njobs, total, a, q);
    else hipLaunchKernelGGL((whittedKernel<W, Q, false>), g, dim3(256), 0, s, jobs, njobs, total, a, q);
    return hipGetLastError();
}

hipError_t launchWhitted(int width, const DJob* jobs, int njobs, uint32_t total, const TraceArgs& a, const Queues& q,
                         hipStream_t s) {
    switch (width) {
        case 4: return launchWhittedW<4, false>(jobs, njobs, total, a, q, s);
        case 8: return launchWhittedW<8, false>(jobs, njobs, total, a, q, s);
        case 4 | kQuantFlag: return launchWhittedW<4, true>(jobs, njobs, total, a, q, s);
        case 8 | kQuantFlag: return launchWhittedW<8, true>(jobs, njobs, total, a, q, s);
        default: return hipErrorInvalidValue;
    }
}

__global__ __launch_bounds__(256) void accumulateKernel(const DJob* __restrict__ jobs, int j, uint32_t samples,
                                                        const float4* __restrict__ L) {
    const DJob& jb = jobs[j];
    const uint32_t px = blockIdx.x * 256u + threadIdx.x;
    if (px >= jb.pixels) return;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    double mx = 0, my = 0, mz = 0, qx = 0, qy = 0, qz = 0;
    for (uint32_t s = 0; s < samples; ++s) {
        const float4 v = L[jb.pathBase + size_t(s) * jb.pixels + px];
        sx += v.x;
        sy += v.y;
        sz += v.z;
        if (jb.momSum) {
            mx += v.x; my += v.y; mz += v.z;
            qx += double(v.x) * v.x; qy += double(v.y) * v.y; qz += double(v.z) * v.z;
        }
    }
    jb.filmSum[3 * size_t(px) + 0] += sx;
    jb.filmSum[3 * size_t(px) + 1] += sy;
    jb.filmSum[3 * size_t(px) + 2] += sz;
    if (jb.momSum) {
        jb.momSum[3 * size_t(px) + 0] += mx;
        jb.momSum[3 * size_t(px) + 1] += my;
        jb.momSum[3 * size_t(px) + 2] += mz;
        jb.momSq[3 * size_t(px) + 0] += qx;
        jb.momSq[3 * size_t(px) + 1] += qy;
        jb.momSq[3 * size_t(px) + 2] += qz;
    }
}

__global__ void zeroKernel(float* p, size_t n) {
    for (size_t i = blockIdx.x * size_t(256) + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) p[i] = 0.0f;
}
__global__ void zeroKernelD(double* p, size_t n) {
    for (size_t i = blockIdx.x * size_t(256) + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) p[i] = 0.0;
}

// ---------------------------------------------------------------- host side

DMaterial makeMaterial(const rbr_material& in) {
    DMaterial m{};
    for (int k = 0; k < 3; ++k) m.base[k] = in.base_color[k];
    m.metallic = in.metallic;
    m.roughness = in.roughness;
    m.ior = in.ior;
    m.transmission = in.transmission;
    const float metallic = std::min(std::max(in.metallic, 0.0f), 1.0f);
    const float t = std::min(std::max(in.transmission, 0.0f), 1.0f);
    const float dielectric = 1.0f - metallic;
    const bool smooth = in.roughness < 0.03f;
    const bool hasInterface = std::fabs(in.ior - 1.0f) > 1e-3f;
    const bool tinted = !(in.base_color[0] <= 0.0f && in.base_color[1] <= 0.0f && in.base_color[2] <= 0.0f);
    const float alpha = std::max(in.roughness * in.roughness, 1e-3f);
    auto add = [&](DLobe l, float w, bool substrate) {
        if (w <= 0.0f || m.n >= 6) return;
        l.weight = w;
        if (substrate) m.substrateMask |= 1u << m.n;
        m.lobe[m.n++] = l;
    };
    auto reflLobe = [&](uint32_t fk, float etaT, const float* f0) {
        DLobe l{};
        l.type = smooth ? DL_SPEC_REFL : DL_MF_REFL;
        l.R[0] = l.R[1] = l.R[2] = 1.0f;
        l.fkind = fk;
        l.fEtaI = 1.0f;
        l.fEtaT = etaT;
        for (int k = 0; k < 3; ++k) l.f0[k] = f0 ? f0[k] : 1.0f;
        l.alpha = alpha;
        return l;
    };
    if (metallic > 0.0f && tinted) add(reflLobe(DF_SCHLICK, 1.0f, in.base_color), metallic, false);
    if (dielectric > 0.0f) {
        if (hasInterface) {
            add(reflLobe(DF_DIELECTRIC, in.ior, nullptr), dielectric, false);
            m.sfKind = DF_DIELECTRIC;
            m.sfEtaI = 1.0f;
            m.sfEtaT = in.ior;
            if (t > 0.0f && tinted) {
                DLobe l{};
                l.type = smooth ? DL_SPEC_TRANS : DL_MF_TRANS;
                for (int k = 0; k < 3; ++k) l.R[k] = in.base_color[k];
                l.etaOut = 1.0f;
                l.etaIn = in.ior;
                l.alpha = alpha;
                add(l, dielectric * t, false);
            }
            if (t < 1.0f && tinted) {
                DLobe l{};
                l.type = DL_LAMBERT;
                for (int k = 0; k < 3; ++k) l.R[k] = in.base_color[k];
                add(l, dielectric * (1.0f - t), true);
            }
        } else {
            if (t > 0.0f && tinted) {
                DLobe l{};
                l.type = DL_SPEC_TRANS;
                for (int k = 0; k < 3; ++k) l.R[k] = in.base_color[k];
                l.etaOut = l.etaIn = 1.0f;
                add(l, dielectric * t, false);
            }
            if (t < 1.0f && tinted) {
                DLobe l{};
                l.type = DL_LAMBERT;
                for (int k = 0; k < 3; ++k) l.R[k] = in.base_color[k];
                add(l, dielectric * (1.0f - t), false);
            }
        }
    }
    m.allDelta = m.n > 0 ? 1u : 0u;
    for (uint32_t i = 0; i < m.n; ++i)
        if (m.lobe[i].type != DL_SPEC_REFL && m.lobe[i].type != DL_SPEC_TRANS) m.allDelta = 0u;
    m.emissive = (in.emission[0] > 0.0f || in.emission[1] > 0.0f || in.emission[2] > 0.0f) ? 1u : 0u;
    for (int k = 0; k < 3; ++k) m.emission[k] = in.emission[k];
    return m;
}

}  // namespace

// Per-context render state: the path pool and queues, grown on demand.
struct RenderState {
    DevBuf ray[2], path[2], root[2], hit, sray, sroot, scontrib, socc, sindex, T, L, counts, jobs, guideOut, recOf, shit;
    DevBuf arenaNodes, arenaPrims, arenaQuadN, arenaShade, arenaSmooth, arenaMats, arenaLights;
    std::vector<DevBuf> films, momS, momQ;
    double lastTotalMs = 0.0;
    double stageMs[4] = {0, 0, 0, 0};
    std::vector<hipEvent_t> events;
    // Arena of the last render: (scene, version) per slot, slot offsets and traversal args.
    std::vector<std::pair<const rbr_scene*, uint64_t>> arenaKey;
    std::vector<uint32_t> arenaLightOff, arenaNodeOff;
    TraceArgs arenaTA{};
    // material sort buffers
    DevBuf perm, sortKey, bins;
    // recent rounds' launch sequences as executable graphs, with the values each was built
    // for (most recently used first)
    struct Graph {
        std::vector<uint64_t> key;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    // traversal census (rbr_set_render_census): sampled rays and roots per (bounce, queue)
    // slot, the queue length of each slot, and the result of the last census render
    DevBuf censusRays, censusRoots, censusCounts;
    rbr_render_stats stats{};
    bool statsValid = false;
};

constexpr uint32_t kCensusSlots = 512;   // (round, bounce, queue) samples per render

constexpr size_t kGraphCache = 8;

}  // namespace rbr

using namespace rbr;

void rbr::freeShading(rbr_scene* s) {
    freeBuffer(s->dMaterials);
    freeBuffer(s->dPrimMaterial);
    freeBuffer(s->dLights);
}

void rbr::destroyRenderState(rbr_ctx* ctx) {
    RenderState* r = ctx->render;
    if (!r) return;
    for (DevBuf* b : {&r->ray[0], &r->ray[1], &r->path[0], &r->path[1], &r->root[0], &r->root[1], &r->hit, &r->sray,
                      &r->sroot, &r->scontrib, &r->socc, &r->sindex, &r->T, &r->L, &r->counts, &r->jobs,
                      &r->guideOut, &r->recOf,
                      &r->arenaNodes, &r->arenaPrims, &r->arenaQuadN, &r->arenaShade, &r->arenaSmooth, &r->arenaMats,
                      &r->arenaLights})
        freeBuffer(*b);
    for (auto* v : {&r->films, &r->momS, &r->momQ})
        for (DevBuf& b : *v) freeBuffer(b);
    for (hipEvent_t e : r->events) (void)hipEventDestroy(e);
    for (DevBuf* b : {&r->perm, &r->sortKey, &r->bins, &r->censusRays, &r->censusRoots, &r->censusCounts}) freeBuffer(*b);
    for (auto& g : r->graphs) (void)hipGraphExecDestroy(g.exec);
    delete r;
    ctx->render = nullptr;
}

extern "C" {

int rbr_scene_set_shading(rbr_scene* s, const rbr_material* materials, uint32_t nmat, const int32_t* primMaterial,
                          uint32_t nprimIds, const rbr_light* lights, uint32_t nlights, const rbr_medium* medium,
                          const float background[3]) {
    if (!s) return RBR_ERR_INVALID;
    rbr_ctx* ctx = s->ctx;
    const HostGeometry& g = s->geom;
    const uint32_t needIds = g.anyId ? g.maxId + 1 : 0;
    if (nprimIds < needIds || (nprimIds && !primMaterial)) return fail(ctx, RBR_ERR_INVALID, "prim_material must cover every primitive id");
    if ((nmat && !materials) || (nlights && !lights)) return fail(ctx, RBR_ERR_INVALID, "null material or light array");
    if (nlights > kMaxLights) return fail(ctx, RBR_ERR_UNSUPPORTED, "more than 32 lights");
    for (uint32_t i = 0; i < nprimIds; ++i)
        if (primMaterial[i] < -1 || primMaterial[i] >= int32_t(nmat)) return fail(ctx, RBR_ERR_INVALID, "material index out of range");
    if (medium && (medium->type < RBR_MEDIUM_NONE || medium->type > RBR_MEDIUM_EXPONENTIAL))
        return fail(ctx, RBR_ERR_INVALID, "bad medium type");
    for (uint32_t i = 0; i < nlights; ++i)
        if (lights[i].type < RBR_LIGHT_AREA || lights[i].type > RBR_LIGHT_DIRECTIONAL) return fail(ctx, RBR_ERR_INVALID, "bad light type");
    s->materials.assign(materials, materials + nmat);
    s->primMaterial.assign(primMaterial, primMaterial + nprimIds);
    s->lights.assign(lights, lights + nlights);
    s->medium = medium ? *medium : rbr_medium{};
    for (int k = 0; k < 3; ++k) s->background[k] = background ? background[k] : 0.0f;
    s->shaded = true;
    bumpVersion(s);
    return RBR_OK;
}

}  // extern "C"

namespace {

// Shading records by primitive id: (outward normal or sphere centre, kind) (radius, smooth
// index, -, material); triangles with vertex normals (kind 4) also get a record of 6 F4 in
// `smooth` (v0, e1, e2 as the traversal's primitive records hold them, unit n0, n1, n2).
void buildShade(const rbr_scene* s, uint32_t idBase, int32_t matBase, std::vector<F4>& out, std::vector<F4>& smooth) {
    const HostGeometry& g = s->geom;
    const uint32_t n = g.anyId ? g.maxId + 1 : 0;
    const size_t o = out.size();
    out.resize(o + 2 * size_t(n));
    auto asF = [](uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; };
    for (uint32_t i = 0; i < n; ++i) {
        out[o + 2 * i] = {0, 0, 1, asF(0xFFu)};
        out[o + 2 * i + 1] = {0, 0, 0, asF(0xFFFFFFFFu)};
    }
    auto mat = [&](uint32_t id) -> float {
        int32_t m = id < s->primMaterial.size() ? s->primMaterial[id] : -1;
        return asF(uint32_t(m < 0 ? -1 : m + matBase));
    };
    for (size_t i = 0; i < g.triIds.size(); ++i) {
        const float* t = &g.tris[9 * i];
        // Triangle geometric normal: cross(v1-v0, v2-v0).normalized() (triangle.cpp:33).
        float e1[3] = {t[3] - t[0], t[4] - t[1], t[5] - t[2]}, e2[3] = {t[6] - t[0], t[7] - t[1], t[8] - t[2]};
        float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        float l = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        float inv = l > 0.0f ? 1.0f / l : 0.0f;
        const uint32_t id = g.triIds[i];
        const bool flat = s->triNormals.empty() ||
                          std::all_of(&s->triNormals[9 * i], &s->triNormals[9 * i] + 9, [](float v) { return v == 0.0f; });
        if (flat) {
            out[o + 2 * id] = {c[0] * inv, c[1] * inv, c[2] * inv, asF(0u)};
            out[o + 2 * id + 1] = {0, 0, 0, mat(id)};
        } else {
            const float* n = &s->triNormals[9 * i];
            out[o + 2 * id] = {c[0] * inv, c[1] * inv, c[2] * inv, asF(4u)};
            out[o + 2 * id + 1] = {0, asF(uint32_t(smooth.size() / 6)), 0, mat(id)};
            smooth.push_back({t[0], t[1], t[2], 0});
            smooth.push_back({e1[0], e1[1], e1[2], 0});
            smooth.push_back({e2[0], e2[1], e2[2], 0});
            smooth.push_back({n[0], n[1], n[2], 0});
            smooth.push_back({n[3], n[4], n[5], 0});
            smooth.push_back({n[6], n[7], n[8], 0});
        }
    }
    for (size_t i = 0; i < g.sphIds.size(); ++i) {
        const float* sp = &g.sphs[4 * i];
        const uint32_t id = g.sphIds[i];
        out[o + 2 * id] = {sp[0], sp[1], sp[2], asF(1u)};
        out[o + 2 * id + 1] = {sp[3], 0, 0, mat(id)};
    }
    for (size_t i = 0; i < g.quadIds.size(); ++i) {
        const float* qd = &g.quads[12 * i];
        const uint32_t id = g.quadIds[i];
        out[o + 2 * id] = {qd[9], qd[10], qd[11], asF(2u)};
        out[o + 2 * id + 1] = {0, 0, 0, mat(id)};
    }
    for (size_t i = 0; i < g.planeIds.size(); ++i) {
        const float* pl = &g.planes[4 * i];
        const uint32_t id = g.planeIds[i];
        out[o + 2 * id] = {pl[0], pl[1], pl[2], asF(3u)};
        out[o + 2 * id + 1] = {0, 0, 0, mat(id)};
    }
    (void)idBase;
}

DLight makeLight(const rbr_light& l) {
    DLight d{};
    d.type = uint32_t(l.type);
    for (int k = 0; k < 3; ++k) {
        d.color[k] = l.color[k];
        d.position[k] = l.position[k];
        d.u[k] = l.u[k];
        d.v[k] = l.v[k];
        d.normal[k] = l.normal[k];
        d.direction[k] = l.direction[k];
    }
    d.intensity = l.intensity;
    const float c[3] = {l.u[1] * l.v[2] - l.u[2] * l.v[1], l.u[2] * l.v[0] - l.u[0] * l.v[2], l.u[0] * l.v[1] - l.u[1] * l.v[0]};
    d.area = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    return d;
}

struct SceneSlot {
    const rbr_scene* s;
    uint32_t nodeOff, primOff, quadOff, idOff, matOff, lightOff;
};

}  // namespace

namespace {

// Host half of the traversal census: downloads the sampled rays and counts node visits and
// primitive tests with the host traversal of each ray's scene (found by its BVH root), each
// slot's sample weighted by queue length / sample size.
int censusStats(rbr_ctx* ctx, RenderState& R, const std::vector<SceneSlot>& scenes, const std::vector<uint8_t>& kind,
                uint32_t k) {
    const size_t slots = kind.size();
    std::vector<uint32_t> counts(slots), roots(slots * k);
    std::vector<float> rays(slots * k * 8);
    hipError_t e = hipSuccess;
    if (slots) {
        e = hipMemcpyAsync(counts.data(), R.censusCounts.ptr, slots * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(roots.data(), R.censusRoots.ptr, slots * k * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(rays.data(), R.censusRays.ptr, slots * k * 32, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return failHip(ctx, e, "census download");
    }
    double n[2] = {0, 0}, nodes[2] = {0, 0}, prims[2] = {0, 0}, empty = 0;
    uint64_t sampled[2] = {0, 0};
    std::vector<float> group;
    for (size_t sl = 0; sl < slots; ++sl) {
        const uint32_t cnt = counts[sl], m = std::min(cnt, k);
        if (!m) continue;
        const double w = double(cnt) / double(m);
        const int kd = kind[sl];
        n[kd] += cnt;
        sampled[kd] += m;
        for (const SceneSlot& sc : scenes) {
            group.clear();
            for (uint32_t j = 0; j < m; ++j)
                if (roots[sl * k + j] == sc.nodeOff)
                    group.insert(group.end(), &rays[(sl * k + j) * 8], &rays[(sl * k + j) * 8] + 8);
            if (group.empty() || sc.s->bvh.empty()) continue;
            if (int rc = ensureHostBvh(sc.s)) return rc;
            const TraceCounters c = hostCount(sc.s->bvh, sc.s->geom, group.data(), group.size() / 8, kd == 1);
            nodes[kd] += w * double(c.nodes);
            prims[kd] += w * double(c.prims);
        }
        if (kd == 1)
            for (uint32_t j = 0; j < m; ++j) empty += roots[sl * k + j] == kEmptyD ? w : 0.0;
    }
    rbr_render_stats& st = R.stats;
    st = rbr_render_stats{};
    st.shadow_empty = uint64_t(empty + 0.5);
    st.extension_rays = uint64_t(n[0] + 0.5);
    st.shadow_rays = uint64_t(std::max(0.0, n[1] - empty) + 0.5);
    st.sampled_extension = sampled[0];
    st.sampled_shadow = sampled[1];
    st.ext_nodes_per_ray = n[0] > 0 ? nodes[0] / n[0] : 0.0;
    st.ext_prims_per_ray = n[0] > 0 ? prims[0] / n[0] : 0.0;
    const double ns = n[1] - empty;
    st.shadow_nodes_per_ray = ns > 0 ? nodes[1] / ns : 0.0;
    st.shadow_prims_per_ray = ns > 0 ? prims[1] / ns : 0.0;
    R.statsValid = true;
    return RBR_OK;
}

// Progress reports of one render: the fraction of its path samples whose round has finished,
// handed to the caller's function by a host callback the stream runs after each round (on a
// runtime thread).  The items outlive every callback: renderImpl waits for the stream before
// they go (the guard also covers early error returns).
struct ProgressItem {
    rbr_progress_fn fn;
    void* user;
    float fraction;
};

void runProgress(void* data) {
    const ProgressItem* it = static_cast<const ProgressItem*>(data);
    it->fn(it->fraction, it->user);
}

struct ProgressQueue {
    hipStream_t stream = nullptr;
    std::deque<ProgressItem> items;
    ~ProgressQueue() {
        if (!items.empty()) (void)hipStreamSynchronize(stream);
    }
};

int renderImpl(rbr_ctx* ctx, const rbr_job* jobs, int njobs, double* const* momSum, double* const* momSq,
               rbr_progress_fn progress = nullptr, void* progressUser = nullptr) {
    if (!ctx || !jobs || njobs < 1) return RBR_ERR_INVALID;
    if (njobs > 64) return fail(ctx, RBR_ERR_UNSUPPORTED, "at most 64 jobs per render");
    if (int rc = useDevice(ctx)) return rc;
    auto t0 = std::chrono::steady_clock::now();
    // ---- validate jobs and collect distinct scenes
    std::vector<SceneSlot> scenes;
    std::map<const rbr_scene*, int> sceneIdx;
    int width = 0;
    uint32_t maxStack = 1, maxDepthAll = 0, maxLights = 0;
    bool anyMixed = false, anyPlanes = false, anyGuided = false, anyRecord = false, anyDebug = false, anyAO = false,
         anyWhitted = false;
    uint32_t aoSlots = 0;
    for (int j = 0; j < njobs; ++j) {
        const rbr_job& jb = jobs[j];
        if (!jb.scene || jb.scene->ctx != ctx) return fail(ctx, RBR_ERR_INVALID, "job scene missing or from another context");
        if (!jb.scene->built || !jb.scene->shaded) return fail(ctx, RBR_ERR_STATE, "job scene needs rbr_scene_build_bvh and rbr_scene_set_shading");
        if (!jb.film_rgb) return fail(ctx, RBR_ERR_INVALID, "job film is null");
        if (jb.camera.width < 1 || jb.camera.height < 1 || jb.config.spp < 1 || jb.config.max_depth < 1)
            return fail(ctx, RBR_ERR_INVALID, "job needs width, height, spp and max_depth >= 1");
        if (jb.config.integrator < RBR_INTEGRATOR_PATH || jb.config.integrator > RBR_INTEGRATOR_WHITTED)
            return fail(ctx, RBR_ERR_INVALID, "unknown integrator");
        const bool dbg = jb.config.integrator >= RBR_INTEGRATOR_NORMAL && jb.config.integrator <= RBR_INTEGRATOR_AO;
        const bool wht = jb.config.integrator == RBR_INTEGRATOR_WHITTED;
        if (j > 0 && (dbg != anyDebug || wht != anyWhitted))
            return fail(ctx, RBR_ERR_UNSUPPORTED,
                        "path, debug and Whitted integrators cannot share a launch (one kind per rbr_render)");
        anyDebug = dbg;
        anyWhitted = wht;
        if (wht && jb.config.max_depth > int32_t(kWhittedMaxDepth))
            return fail(ctx, RBR_ERR_UNSUPPORTED, "Whitted max_depth above 16");
        anyAO |= jb.config.integrator == RBR_INTEGRATOR_AO;
        if (jb.config.integrator == RBR_INTEGRATOR_AO) {
            if (jb.config.ao_samples < 0 || jb.config.ao_samples > 256 || jb.config.ao_distance < 0.0f)
                return fail(ctx, RBR_ERR_INVALID, "ao_samples must be in 0..256 and ao_distance >= 0");
            aoSlots = std::max(aoSlots, uint32_t(jb.config.ao_samples ? jb.config.ao_samples : 16));
        }
        if (jb.config.guided < 0 || jb.config.guided > 3) return fail(ctx, RBR_ERR_INVALID, "guided must be 0..3");
        if (jb.config.guided) {
            if (jb.config.integrator != RBR_INTEGRATOR_PATH)
                return fail(ctx, RBR_ERR_UNSUPPORTED, "guiding is implemented for the path integrator");
            if (!ctx->guide || !ctx->guide->ready) return fail(ctx, RBR_ERR_STATE, "guided job needs rbr_guide_init");
            anyGuided |= (jb.config.guided & 1) != 0;
            anyRecord |= (jb.config.guided & 2) != 0;
        }
        if (uint64_t(jb.camera.width) * jb.camera.height > (1ull << 26)) return fail(ctx, RBR_ERR_INVALID, "film too large");
        if (sceneIdx.find(jb.scene) == sceneIdx.end()) {
            sceneIdx[jb.scene] = int(scenes.size());
            scenes.push_back({jb.scene, 0, 0, 0, 0, 0, 0});
        }
        const rbr_scene* s = jb.scene;
        if (!s->bvh.empty()) {
            if (width && width != s->bvh.code()) return fail(ctx, RBR_ERR_UNSUPPORTED, "jobs mix BVH layouts");
            width = s->bvh.code();
            maxStack = std::max(maxStack, uint32_t(s->bvh.width - 1) * s->bvh.maxDepth + 1u);
        }
        anyMixed |= (s->geom.sphIds.size() + s->geom.quadIds.size()) > 0;
        anyPlanes |= !s->geom.planeIds.empty();
        maxDepthAll = std::max(maxDepthAll, uint32_t(jb.config.max_depth));
        maxLights = std::max(maxLights, uint32_t(s->lights.size()));
    }
    if (anyDebug) maxLights = aoSlots;   // debug launches use the shadow queue for AO rays only
    if (width == 0) width = 4;
    if (anyPlanes && scenes.size() > 1) return fail(ctx, RBR_ERR_UNSUPPORTED, "infinite planes in a multi-scene render");
    if (anyWhitted && maxStack > uint32_t(kMaxStack)) return fail(ctx, RBR_ERR_UNSUPPORTED, "BVH too deep for the Whitted stack");
    if (!ctx->render) ctx->render = new RenderState();
    RenderState& R = *ctx->render;
    // node layout of the launch (all scenes share it): dwords per node, child-ref offset
    const int nodeDw = (width & kQuantFlag) ? qNodeDwords(width & 0xFF) : 4 * ((width & 0xFF) == 4 ? 8 : 16);
    const int refOff = (width & kQuantFlag) ? qRefOffset(width & 0xFF) : 6 * (width & 0xFF);

    // ---- node/primitive arena (packed only when several scenes share the launch); reused
    // while the same scenes, unchanged, are rendered again
    std::vector<F4> shade, smooth;
    std::vector<DMaterial> mats;
    std::vector<DLight> lights;
    TraceArgs ta{};
    int rc;
    std::vector<std::pair<const rbr_scene*, uint64_t>> key;
    for (const SceneSlot& sl : scenes) key.emplace_back(sl.s, sl.s->version);
    const bool arenaValid = key == R.arenaKey;
    if (arenaValid) {
        for (size_t k = 0; k < scenes.size(); ++k) {
            scenes[k].lightOff = R.arenaLightOff[k];
            scenes[k].nodeOff = R.arenaNodeOff[k];
        }
        ta = R.arenaTA;
    }
    uint32_t nodeOff = 0, primOff = 0, quadOff = 0, idOff = 0;
    for (SceneSlot& sl : scenes) {
        if (arenaValid) break;
        sl.nodeOff = nodeOff;
        sl.primOff = primOff;
        sl.quadOff = quadOff;
        sl.idOff = idOff;
        sl.matOff = uint32_t(mats.size());
        sl.lightOff = uint32_t(lights.size());
        buildShade(sl.s, idOff, int32_t(sl.matOff), shade, smooth);
        for (const rbr_material& m : sl.s->materials) mats.push_back(makeMaterial(m));
        for (const rbr_light& l : sl.s->lights) lights.push_back(makeLight(l));
        nodeOff += sl.s->bvh.nodeCount;
        primOff += sl.s->bvh.primCount;
        quadOff += uint32_t(sl.s->bvh.quadNormals.size());
        idOff += sl.s->geom.anyId ? sl.s->geom.maxId + 1 : 0;
    }
    if (primOff >= (1u << 27)) return fail(ctx, RBR_ERR_UNSUPPORTED, "arena exceeds 2^27 primitives");
    if (arenaValid) {
        // device arena, shading records, materials and lights are current
    } else if (scenes.size() == 1) {
        ta = sceneTraceArgs(scenes[0].s);
    } else {
        for (const SceneSlot& sl : scenes)
            if ((rc = ensureHostBvh(sl.s))) return rc;
        std::vector<uint32_t> nodes(size_t(nodeOff) * nodeDw);
        std::vector<F4> prims(size_t(primOff) * 3), quadN(quadOff);
        for (const SceneSlot& sl : scenes) {
            const WideBVH& b = sl.s->bvh;
            std::memcpy(&nodes[size_t(sl.nodeOff) * nodeDw], b.deviceNodes(), b.deviceNodeBytes());
            for (uint32_t nd = 0; nd < b.nodeCount; ++nd) {
                uint32_t* ch = &nodes[size_t(sl.nodeOff + nd) * nodeDw] + refOff;
                for (int k = 0; k < (width & 0xFF); ++k) {
                    if (ch[k] == kEmpty) continue;
                    if (isLeaf(ch[k])) ch[k] = makeLeaf(leafStart(ch[k]) + sl.primOff, leafCount(ch[k]));
                    else ch[k] += sl.nodeOff;
                }
            }
            for (uint32_t pi = 0; pi < b.primCount; ++pi) {
                F4* r = &prims[size_t(sl.primOff + pi) * 3];
                std::memcpy(r, &b.prims[size_t(pi) * 3], 3 * sizeof(F4));
                uint32_t id, kind;
                std::memcpy(&id, &r[0].w, 4);
                id += sl.idOff;
                std::memcpy(&r[0].w, &id, 4);
                std::memcpy(&kind, &r[1].w, 4);
                if (kind == KIND_QUAD) {
                    uint32_t qi;
                    std::memcpy(&qi, &r[2].w, 4);
                    qi += sl.quadOff;
                    std::memcpy(&r[2].w, &qi, 4);
                }
            }
            std::copy(b.quadNormals.begin(), b.quadNormals.end(), quadN.begin() + sl.quadOff);
        }
        if ((rc = uploadBuffer(ctx, R.arenaNodes, nodes.data(), nodes.size() * sizeof(uint32_t)))) return rc;
        if ((rc = uploadBuffer(ctx, R.arenaPrims, prims.data(), prims.size() * sizeof(F4)))) return rc;
        if ((rc = uploadBuffer(ctx, R.arenaQuadN, quadN.data(), quadN.size() * sizeof(F4)))) return rc;
        ta.nodes = nodeOff ? static_cast<const float4*>(R.arenaNodes.ptr) : nullptr;
        ta.prims = static_cast<const float4*>(R.arenaPrims.ptr);
        ta.quadN = static_cast<const float4*>(R.arenaQuadN.ptr);
        ta.nplanes = 0;
        ta.mixed = anyMixed ? 1u : 0u;
        // watchdog bound of the arena: a ray stays in its job's BVH, which has fewer nodes than
        // the arena
        ta.stepCap = nodeOff + 1u;
    }
    if (!arenaValid) {
        R.arenaKey.clear();   // stays invalid if an upload below fails
        if ((rc = uploadBuffer(ctx, R.arenaShade, shade.data(), std::max<size_t>(1, shade.size()) * sizeof(F4)))) return rc;
        if ((rc = uploadBuffer(ctx, R.arenaSmooth, smooth.data(), std::max<size_t>(1, smooth.size()) * sizeof(F4))))
            return rc;
        if ((rc = uploadBuffer(ctx, R.arenaMats, mats.data(), std::max<size_t>(1, mats.size()) * sizeof(DMaterial)))) return rc;
        if ((rc = uploadBuffer(ctx, R.arenaLights, lights.data(), std::max<size_t>(1, lights.size()) * sizeof(DLight)))) return rc;
        R.arenaKey = key;
        R.arenaTA = ta;
        R.arenaLightOff.clear();
        R.arenaNodeOff.clear();
        for (const SceneSlot& sl : scenes) {
            R.arenaLightOff.push_back(sl.lightOff);
            R.arenaNodeOff.push_back(sl.nodeOff);
        }
    }

    // ---- per-job device descriptors and films
    std::vector<DJob> dj(njobs);
    R.films.resize(std::max<size_t>(R.films.size(), size_t(njobs)));
    R.momS.resize(std::max<size_t>(R.momS.size(), size_t(njobs)));
    R.momQ.resize(std::max<size_t>(R.momQ.size(), size_t(njobs)));
    uint64_t totalPaths = 0;
    for (int j = 0; j < njobs; ++j) {
        const rbr_job& jb = jobs[j];
        const SceneSlot& sl = scenes[sceneIdx[jb.scene]];
        DJob& d = dj[j];
        d = DJob{};
        for (int k = 0; k < 3; ++k) {
            d.cam.pos[k] = jb.camera.position[k];
            d.cam.fwd[k] = jb.camera.forward[k];
            d.cam.right[k] = jb.camera.right[k];
            d.cam.up[k] = jb.camera.up[k];
            d.bg[k] = jb.scene->background[k];
            d.medium.sa[k] = jb.scene->medium.sigma_a[k];
            d.medium.ss[k] = jb.scene->medium.sigma_s[k];
        }
        d.cam.fx = jb.camera.fx;
        d.cam.fy = jb.camera.fy;
        d.cam.cx = jb.camera.cx;
        d.cam.cy = jb.camera.cy;
        d.cam.width = uint32_t(jb.camera.width);
        d.cam.height = uint32_t(jb.camera.height);
        d.medium.type = uint32_t(jb.scene->medium.type);
        d.medium.falloff = jb.scene->medium.falloff;
        d.medium.base = jb.scene->medium.base_height;
        d.medium.g = jb.scene->medium.g;
        d.primShade = static_cast<const float4*>(R.arenaShade.ptr);   // ids are arena-global
        d.smooth = static_cast<const float4*>(R.arenaSmooth.ptr);
        d.mats = static_cast<const DMaterial*>(R.arenaMats.ptr);
        d.lights = static_cast<const DLight*>(R.arenaLights.ptr) + sl.lightOff;
        d.nLights = uint32_t(jb.scene->lights.size());
        d.integrator = uint32_t(jb.config.integrator);
        d.maxDepth = uint32_t(jb.config.max_depth);
        d.useNEE = jb.config.use_nee ? 1u : 0u;
        d.useMIS = jb.config.use_mis ? 1u : 0u;
        d.guided = uint32_t(jb.config.guided);
        d.guideDepth = jb.config.guide_depth ? jb.config.guide_depth : 0xFFFFFFFFu;
        d.aoSamples = uint32_t(jb.config.ao_samples ? jb.config.ao_samples : 16);
        d.aoDistance = jb.config.ao_distance > 0.0f ? jb.config.ao_distance : 2.0f;
        d.depthNear = jb.config.depth_near != 0.0f || jb.config.depth_far != 0.0f ? jb.config.depth_near : 0.1f;
        d.depthFar = jb.config.depth_near != 0.0f || jb.config.depth_far != 0.0f ? jb.config.depth_far : 100.0f;
        for (int k = 0; k < 3; ++k) d.ambient[k] = jb.config.ambient[k];
        d.root = jb.scene->bvh.empty() ? kEmpty : sl.nodeOff;
        d.pixels = uint32_t(jb.camera.width) * uint32_t(jb.camera.height);
        d.seed = jb.config.seed;
        const size_t fb = size_t(d.pixels) * 3;
        if ((rc = ensureBuffer(ctx, R.films[j], fb * sizeof(float)))) return rc;
        d.filmSum = static_cast<float*>(R.films[j].ptr);
        hipLaunchKernelGGL(zeroKernel, dim3(1024), dim3(256), 0, ctx->stream, d.filmSum, fb);
        if (momSum && momSum[j]) {
            if ((rc = ensureBuffer(ctx, R.momS[j], fb * sizeof(double)))) return rc;
            if ((rc = ensureBuffer(ctx, R.momQ[j], fb * sizeof(double)))) return rc;
            d.momSum = static_cast<double*>(R.momS[j].ptr);
            d.momSq = static_cast<double*>(R.momQ[j].ptr);
            hipLaunchKernelGGL(zeroKernelD, dim3(1024), dim3(256), 0, ctx->stream, d.momSum, fb);
            hipLaunchKernelGGL(zeroKernelD, dim3(1024), dim3(256), 0, ctx->stream, d.momSq, fb);
        }
        totalPaths += uint64_t(d.pixels) * uint32_t(jb.config.spp);
    }

    // ---- pool
    uint32_t maxPixels = 0;
    for (const DJob& d : dj) maxPixels = std::max(maxPixels, d.pixels);
    const uint32_t P = uint32_t(std::min<uint64_t>(std::max<uint64_t>(ctx->renderPool, maxPixels), totalPaths));
    const uint32_t ML = std::max(1u, maxLights);
    // shadow queue: one slot per path and light, plus the unused tail of each shading block's
    // last slot chunk (at most one chunk per persistent block)
    const size_t SQ = size_t(P) * ML + size_t(std::max(1, ctx->cuCount)) * 8u * kShadowChunk;
    struct { DevBuf* b; size_t bytes; } allocs[] = {
        {&R.ray[0], size_t(P) * 32}, {&R.ray[1], size_t(P) * 32}, {&R.path[0], size_t(P) * 4}, {&R.path[1], size_t(P) * 4},
        {&R.root[0], size_t(P) * 4}, {&R.root[1], size_t(P) * 4}, {&R.hit, size_t(P) * 8},
        {&R.sray, SQ * 32}, {&R.sroot, SQ * 4}, {&R.scontrib, SQ * 16},
        {&R.socc, SQ}, {&R.sindex, size_t(P) * ML * 4}, {&R.T, size_t(P) * 16}, {&R.L, size_t(P) * 16},
        {&R.counts, 16}, {&R.jobs, sizeof(DJob) * size_t(njobs)}};
    for (auto& a : allocs)
        if ((rc = ensureBuffer(ctx, *a.b, a.bytes))) return rc;
    if (anyGuided && (rc = ensureBuffer(ctx, R.guideOut, size_t(P) * kGuideRecBytes))) return rc;
    if (anyAO && (rc = ensureBuffer(ctx, R.shit, size_t(P) * ML * sizeof(uint2)))) return rc;
    if (anyRecord) {
        if ((rc = ensureBuffer(ctx, R.recOf, size_t(P) * 4))) return rc;
        if ((rc = guideEnsureRecords(ctx, *ctx->guide))) return rc;
    }
    Queues q;
    for (int k = 0; k < 2; ++k) {
        q.ray[k] = static_cast<float4*>(R.ray[k].ptr);
        q.path[k] = static_cast<uint32_t*>(R.path[k].ptr);
        q.root[k] = static_cast<uint32_t*>(R.root[k].ptr);
    }
    q.hit = static_cast<uint2*>(R.hit.ptr);
    q.sray = static_cast<float4*>(R.sray.ptr);
    q.sroot = static_cast<uint32_t*>(R.sroot.ptr);
    q.scontrib = static_cast<float4*>(R.scontrib.ptr);
    q.socc = static_cast<uint8_t*>(R.socc.ptr);
    q.sindex = static_cast<int32_t*>(R.sindex.ptr);
    q.T = static_cast<float4*>(R.T.ptr);
    q.L = static_cast<float4*>(R.L.ptr);
    q.counts = static_cast<uint32_t*>(R.counts.ptr);
    q.guide = anyGuided ? static_cast<const uint4*>(R.guideOut.ptr) : nullptr;
    q.rec = anyRecord ? static_cast<float4*>(ctx->guide->dRec.ptr) : nullptr;
    q.recOf = anyRecord ? static_cast<uint32_t*>(R.recOf.ptr) : nullptr;
    q.recCount = anyRecord ? static_cast<uint32_t*>(ctx->guide->dRecCount.ptr) : nullptr;
    q.recCap = anyRecord ? ctx->guide->recCap : 0u;
    q.perm = q.sortKey = q.bins = nullptr;
    if (ctx->renderSort && !anyDebug) {
        if ((rc = ensureBuffer(ctx, R.perm, size_t(P) * 4)) || (rc = ensureBuffer(ctx, R.sortKey, size_t(P) * 4)) ||
            (rc = ensureBuffer(ctx, R.bins, kSortBins * 4)))
            return rc;
        q.perm = static_cast<uint32_t*>(R.perm.ptr);
        q.sortKey = static_cast<uint32_t*>(R.sortKey.ptr);
        q.bins = static_cast<uint32_t*>(R.bins.ptr);
    }
    if ((rc = prepareTraversal(ctx, maxStack))) return rc;
    if (anyGuided && (rc = guideSetAttributes(ctx))) return rc;

    // ---- traversal census (profiling): sample buffers; graphs are not used meanwhile
    const bool census = ctx->renderCensus > 0 && !anyDebug && !anyWhitted;
    const uint32_t censusK = ctx->renderCensus;
    std::vector<uint8_t> censusKind;   // per slot: 0 extension queue, 1 shadow queue
    R.statsValid = false;
    if (census) {
        if ((rc = ensureBuffer(ctx, R.censusRays, size_t(kCensusSlots) * censusK * 32)) ||
            (rc = ensureBuffer(ctx, R.censusRoots, size_t(kCensusSlots) * censusK * 4)) ||
            (rc = ensureBuffer(ctx, R.censusCounts, size_t(kCensusSlots) * 4)))
            return rc;
    }
    auto censusSample = [&](const float4* rays, const uint32_t* roots, const uint32_t* count, uint8_t kind) {
        if (!census || censusKind.size() >= kCensusSlots) return;
        const size_t slot = censusKind.size();
        censusKind.push_back(kind);
        hipLaunchKernelGGL(censusKernel, dim3((censusK + 255) / 256), dim3(256), 0, ctx->stream, rays, roots, count,
                           censusK, static_cast<float4*>(R.censusRays.ptr) + slot * censusK * 2,
                           static_cast<uint32_t*>(R.censusRoots.ptr) + slot * censusK,
                           static_cast<uint32_t*>(R.censusCounts.ptr) + slot);
    };

    // ---- rounds: give every unfinished job as many whole samples as the pool holds
    std::vector<uint32_t> done(njobs, 0);
    hipError_t e = hipSuccess;
    hipEvent_t ev[2];
    while (R.events.size() < 2) {
        hipEvent_t x;
        if ((e = hipEventCreate(&x)) != hipSuccess) return failHip(ctx, e, "hipEventCreate");
        R.events.push_back(x);
    }
    ev[0] = R.events[0];
    ev[1] = R.events[1];
    ProgressQueue progressQ;
    progressQ.stream = ctx->stream;
    if ((e = hipEventRecord(ev[0], ctx->stream)) != hipSuccess) return failHip(ctx, e, "hipEventRecord");
    while (true) {
        uint32_t used = 0;
        std::vector<uint32_t> take(njobs, 0);
        for (int j = 0; j < njobs; ++j) {
            const uint32_t left = uint32_t(jobs[j].config.spp) - done[j];
            if (!left) continue;
            const uint32_t fit = (P - used) / dj[j].pixels;
            take[j] = std::min(left, fit);
            used += take[j] * dj[j].pixels;
        }
        if (used == 0) break;
        uint32_t depthRound = 0;
        uint32_t base = 0;
        for (int j = 0; j < njobs; ++j) {
            dj[j].pathBase = base;
            dj[j].pathCount = take[j] * dj[j].pixels;
            dj[j].sampleBase = jobs[j].config.sample_offset + done[j];
            base += dj[j].pathCount;
            if (take[j]) depthRound = std::max(depthRound, anyDebug ? 1u : uint32_t(jobs[j].config.max_depth));
        }
        e = hipMemcpyAsync(R.jobs.ptr, dj.data(), sizeof(DJob) * njobs, hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) return failHip(ctx, e, "job upload");
        const DJob* djd = static_cast<const DJob*>(R.jobs.ptr);
        const unsigned grid = (used + 255) / 256;
        // persistent kernels: enough blocks to fill the device, never more than the queue needs
        const unsigned cus = unsigned(std::max(1, ctx->cuCount));
        const unsigned gridPersist = std::min(grid, cus * 8u);
        const unsigned gridSort = std::min((used + kSortChunk - 1) / kSortChunk, cus * 4u);
        armWatchdog(ctx, ta);   // the Whitted kernel traces inline with ta
        // The round's launch sequence; it allocates nothing, so it can be captured.
        auto launchRound = [&]() -> int {
            if (anyWhitted) {
                if (hipError_t we = launchWhitted(width, djd, njobs, used, ta, q, ctx->stream); we != hipSuccess)
                    return failHip(ctx, we, "Whitted launch");
                for (int j = 0; j < njobs; ++j) {
                    if (!take[j]) continue;
                    hipLaunchKernelGGL(accumulateKernel, dim3((dj[j].pixels + 255) / 256), dim3(256), 0, ctx->stream,
                                       djd, j, take[j], q.L);
                }
                return RBR_OK;
            }
            hipLaunchKernelGGL(cameraKernel, dim3(grid), dim3(256), 0, ctx->stream, djd, njobs, used, q);
            int cur = 0;
            for (uint32_t depth = 0; depth < depthRound; ++depth) {
                hipLaunchKernelGGL(resetCountsKernel, dim3(1), dim3(1), 0, ctx->stream, q.counts, cur ^ 1);
                TraceArgs a = ta;
                a.rays = q.ray[cur];
                a.n = used;
                a.nDev = q.counts + cur;
                a.rootOf = q.root[cur];
                a.hits = q.hit;
                a.occ = nullptr;
                if (int r2 = launchTraversal(ctx, a, width, maxStack, false)) return r2;
                censusSample(q.ray[cur], q.root[cur], q.counts + cur, 0);
                if (anyDebug) {
                    hipLaunchKernelGGL(debugShadeKernel, dim3(gridPersist), dim3(256), 0, ctx->stream, djd, njobs, ML, cur, q);
                    if (anyAO) {
                        TraceArgs b = ta;
                        b.rays = q.sray;
                        b.n = used * ML;
                        b.nDev = q.counts + 2;
                        b.rootOf = q.sroot;
                        b.hits = static_cast<uint2*>(R.shit.ptr);
                        b.occ = nullptr;
                        if (int r2 = launchTraversal(ctx, b, width, maxStack, false)) return r2;
                        hipLaunchKernelGGL(debugGatherKernel, dim3(gridPersist), dim3(256), 0, ctx->stream, djd, njobs, ML,
                                           cur, q, static_cast<const uint2*>(R.shit.ptr));
                    }
                    cur ^= 1;
                    continue;
                }
                if (anyGuided) {
                    GuideInferArgs gi{};
                    gi.g = guideDev(*ctx->guide);
                    gi.n = used;
                    gi.nDev = q.counts + cur;
                    gi.rays = q.ray[cur];
                    gi.hits = q.hit;
                    gi.paths = q.path[cur];
                    gi.jobs = djd;
                    gi.njobs = njobs;
                    gi.depth = depth;
                    gi.out = static_cast<uint4*>(R.guideOut.ptr);
                    if (int r2 = launchGuideInfer(ctx, gi, used)) return r2;
                }
                if (q.perm) {
                    hipLaunchKernelGGL(zeroBinsKernel, dim3(1), dim3(256), 0, ctx->stream, q.bins);
                    hipLaunchKernelGGL(sortKeyKernel, dim3(gridSort), dim3(256), 0, ctx->stream, djd, cur, q);
                    hipLaunchKernelGGL(sortScanKernel, dim3(1), dim3(kSortBins), 0, ctx->stream, q.bins);
                    hipLaunchKernelGGL(sortScatterKernel, dim3(gridSort), dim3(256), 0, ctx->stream, cur, q);
                }
                hipLaunchKernelGGL(shadeKernel, dim3(gridPersist), dim3(256), 0, ctx->stream, djd, njobs, depth, maxLights,
                                   cur, q);
                if (maxLights) {
                    TraceArgs b = ta;
                    b.rays = q.sray;
                    b.n = used * ML;
                    b.nDev = q.counts + 2;
                    b.rootOf = q.sroot;
                    b.hits = nullptr;
                    b.occ = q.socc;
                    if (int r2 = launchTraversal(ctx, b, width, maxStack, true)) return r2;
                    censusSample(q.sray, q.sroot, q.counts + 2, 1);
                    hipLaunchKernelGGL(gatherKernel, dim3(gridPersist), dim3(256), 0, ctx->stream, maxLights, cur, q);
                }
                cur ^= 1;
            }
            if (anyRecord) {
                GuideState& gs = *ctx->guide;
                hipLaunchKernelGGL(guideFinalizeKernel, dim3(1024), dim3(256), 0, ctx->stream, q.rec, q.L,
                                   static_cast<const uint32_t*>(gs.dRecCount.ptr), gs.recCap,
                                   static_cast<double*>(gs.dCsum.ptr));
                hipLaunchKernelGGL(guideFinalizeEndKernel, dim3(1), dim3(1), 0, ctx->stream,
                                   static_cast<uint32_t*>(gs.dRecCount.ptr), gs.recCap);
            }
            for (int j = 0; j < njobs; ++j) {
                if (!take[j]) continue;
                hipLaunchKernelGGL(accumulateKernel, dim3((dj[j].pixels + 255) / 256), dim3(256), 0, ctx->stream, djd, j,
                                   take[j], q.L);
            }
            hipError_t le = hipGetLastError();
            return le == hipSuccess ? RBR_OK : failHip(ctx, le, "render launch");
        };
        if (!ctx->renderGraphs || census) {
            if ((rc = launchRound())) return rc;
        } else {
            // Every value a launch of the round depends on; the captured graph is replayed while
            // it is unchanged (device-side job table and counters are rewritten every round).
            std::vector<uint64_t> key;
            auto put = [&](uint64_t v) { key.push_back(v); };
            auto putp = [&](const void* ptr) { key.push_back(uint64_t(reinterpret_cast<uintptr_t>(ptr))); };
            auto putf = [&](float f) { uint32_t u; std::memcpy(&u, &f, 4); key.push_back(u); };
            put(used); put(depthRound); put(maxLights); put(uint64_t(width)); put(maxStack); put(uint64_t(njobs));
            put(anyGuided); put(anyRecord); put(anyDebug); put(anyAO); putp(R.shit.ptr);
            for (int j = 0; j < njobs; ++j) { put(take[j]); put(dj[j].pixels); }
            putp(ta.nodes); putp(ta.prims); putp(ta.quadN); putp(ta.planes); putp(ta.planeIds); put(ta.nplanes); put(ta.mixed);
            for (int k = 0; k < 2; ++k) { putp(q.ray[k]); putp(q.path[k]); putp(q.root[k]); }
            putp(q.hit); putp(q.sray); putp(q.sroot); putp(q.scontrib); putp(q.socc); putp(q.sindex); putp(q.T); putp(q.L);
            putp(q.counts); putp(q.guide); putp(q.rec); putp(q.recOf); putp(q.recCount); put(q.recCap);
            putp(q.perm); putp(q.sortKey); putp(q.bins); putp(djd);
            if (anyGuided) {
                const GuideDev gd = guideDev(*ctx->guide);
                putp(gd.w); putp(gd.wf); putp(gd.b);
                for (int k = 0; k < 3; ++k) { putf(gd.lo[k]); putf(gd.inv[k]); }
            }
            if (anyRecord) { putp(ctx->guide->dRecCount.ptr); putp(ctx->guide->dCsum.ptr); }
            put(uint64_t(ctx->traceMode)); put(uint64_t(uint32_t(ctx->traceModeAny))); put(uint64_t(ctx->ldsDepth));
            put(uint64_t(ctx->wavesPerCu));
            putp(ctx->dCounter.ptr); putp(ctx->dSpill.ptr); put(ctx->dSpill.bytes);
            put(ta.stepCap); putp(ta.fault);
            size_t hit = R.graphs.size();
            for (size_t k = 0; k < R.graphs.size(); ++k)
                if (R.graphs[k].key == key) hit = k;
            if (hit < R.graphs.size()) {
                std::rotate(R.graphs.begin(), R.graphs.begin() + hit, R.graphs.begin() + hit + 1);
            } else {
                if (R.graphs.size() >= kGraphCache) {
                    (void)hipGraphExecDestroy(R.graphs.back().exec);
                    R.graphs.pop_back();
                }
                if ((e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
                    return failHip(ctx, e, "hipStreamBeginCapture");
                const int lrc = launchRound();
                hipGraph_t graph = nullptr;
                e = hipStreamEndCapture(ctx->stream, &graph);
                if (lrc) {
                    if (graph) (void)hipGraphDestroy(graph);
                    return lrc;
                }
                if (e != hipSuccess) return failHip(ctx, e, "hipStreamEndCapture");
                hipGraphExec_t exec = nullptr;
                e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (e != hipSuccess) return failHip(ctx, e, "hipGraphInstantiate");
                R.graphs.insert(R.graphs.begin(), RenderState::Graph{key, exec});
            }
            if ((e = hipGraphLaunch(R.graphs.front().exec, ctx->stream)) != hipSuccess) return failHip(ctx, e, "hipGraphLaunch");
        }
        for (int j = 0; j < njobs; ++j) done[j] += take[j];
        if (progress) {
            uint64_t finished = 0;
            for (int j = 0; j < njobs; ++j) finished += uint64_t(done[j]) * dj[j].pixels;
            progressQ.items.push_back({progress, progressUser, float(double(finished) / double(totalPaths))});
            if ((e = hipLaunchHostFunc(ctx->stream, runProgress, &progressQ.items.back())) != hipSuccess)
                return failHip(ctx, e, "hipLaunchHostFunc");
        }
    }
    if ((e = hipEventRecord(ev[1], ctx->stream)) != hipSuccess) return failHip(ctx, e, "hipEventRecord");
    // ---- films
    std::vector<float> host;
    for (int j = 0; j < njobs; ++j) {
        const size_t fb = size_t(dj[j].pixels) * 3;
        host.resize(fb);
        e = hipMemcpyAsync(host.data(), R.films[j].ptr, fb * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return failHip(ctx, e, "film download");
        const float inv = 1.0f / float(jobs[j].config.spp);
        for (size_t k = 0; k < fb; ++k) jobs[j].film_rgb[k] = host[k] * inv;
        if (momSum && momSum[j]) {
            e = hipMemcpyAsync(momSum[j], R.momS[j].ptr, fb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(momSq[j], R.momQ[j].ptr, fb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) return failHip(ctx, e, "moment download");
        }
    }
    if (census && (rc = censusStats(ctx, R, scenes, censusKind, censusK))) return rc;
    float ms = 0.0f;
    if ((e = hipEventElapsedTime(&ms, ev[0], ev[1])) != hipSuccess) return failHip(ctx, e, "hipEventElapsedTime");
    R.lastTotalMs = ms;
    R.stageMs[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if ((rc = queueFaultRead(ctx))) return rc;
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return failHip(ctx, e, "hipStreamSynchronize");
    return checkFault(ctx);
}

}  // namespace

extern "C" {

int rbr_render(rbr_ctx* ctx, const rbr_job* jobs, int njobs) { return renderImpl(ctx, jobs, njobs, nullptr, nullptr); }

int rbr_render_progress(rbr_ctx* ctx, const rbr_job* jobs, int njobs, rbr_progress_fn progress, void* user) {
    return renderImpl(ctx, jobs, njobs, nullptr, nullptr, progress, user);
}

int rbr_render_moments(rbr_ctx* ctx, const rbr_job* job, double* sum_rgb, double* sumsq_rgb) {
    if (!ctx || !job || !sum_rgb || !sumsq_rgb) return RBR_ERR_INVALID;
    double* s[1] = {sum_rgb};
    double* q[1] = {sumsq_rgb};
    return renderImpl(ctx, job, 1, s, q);
}

int rbr_set_render_tuning(rbr_ctx* ctx, uint32_t pool_paths, int use_graphs, int sort_materials) {
    if (!ctx) return RBR_ERR_INVALID;
    if (pool_paths < (1u << 16) || pool_paths > (1u << 26)) return fail(ctx, RBR_ERR_INVALID, "pool_paths must be in 2^16..2^26");
    if ((use_graphs != 0 && use_graphs != 1) || (sort_materials != 0 && sort_materials != 1))
        return fail(ctx, RBR_ERR_INVALID, "use_graphs and sort_materials are 0 or 1");
    ctx->renderPool = pool_paths;
    ctx->renderGraphs = use_graphs;
    ctx->renderSort = sort_materials;
    return RBR_OK;
}

int rbr_set_render_census(rbr_ctx* ctx, uint32_t rays_per_queue) {
    if (!ctx) return RBR_ERR_INVALID;
    if (rays_per_queue > (1u << 20)) return fail(ctx, RBR_ERR_INVALID, "rays_per_queue must be at most 2^20");
    ctx->renderCensus = rays_per_queue;
    return RBR_OK;
}

int rbr_last_render_stats(rbr_ctx* ctx, rbr_render_stats* out) {
    if (!ctx || !out) return RBR_ERR_INVALID;
    if (!ctx->render || !ctx->render->statsValid) return fail(ctx, RBR_ERR_STATE, "no census render yet");
    *out = ctx->render->stats;
    return RBR_OK;
}

int rbr_last_render_timing(rbr_ctx* ctx, double* total_ms, double* stage_ms) {
    if (!ctx) return RBR_ERR_INVALID;
    if (!ctx->render) return fail(ctx, RBR_ERR_STATE, "no render yet");
    if (total_ms) *total_ms = ctx->render->lastTotalMs;
    if (stage_ms)
        for (int k = 0; k < 4; ++k) stage_ms[k] = ctx->render->stageMs[k];
    return RBR_OK;
}

}  // extern "C"
