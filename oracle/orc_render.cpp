// CPU restatement of the reference's shading and integrators.  TEST INFRASTRUCTURE ONLY
// (file:line map in orc_render.h).  Operand order follows the reference so that, with the
// same sampler stream, radiance values agree with the reference build.
#include "orc_render.h"

#include <cmath>

namespace orc {

namespace {

inline bool sameHemisphere(const V3& a, const V3& b) { return a.z * b.z > 0.0f; }
inline float absCos(const V3& w) { return std::abs(w.z); }

// sampling::cosineHemisphere, sampler.h:73-81
inline V3 cosineHemisphere(float u1, float u2) {
    float r = std::sqrt(u1);
    float theta = TWO_PI * u2;
    float x = r * std::cos(theta);
    float y = r * std::sin(theta);
    float z = std::sqrt(std::max(0.0f, 1.0f - u1));
    return V3(x, y, z);
}

// sampling::powerHeuristic, sampler.h:135-149
inline float powerHeuristic(int nf, float fPdf, int ng, float gPdf) {
    float f = nf * fPdf, g = ng * gPdf;
    if (std::isinf(f)) return 1.0f;
    if (std::isinf(g)) return 0.0f;
    const float m = std::max(f, g);
    if (m <= 0.0f) return 0.0f;
    f /= m;
    g /= m;
    return (f * f) / (f * f + g * g);
}

// sampling::Frame, sampler.h:103-118
struct Frame {
    V3 t, b, n;
    explicit Frame(const V3& nn) : n(nn) { coordinateSystem(nn, t, b); }
    V3 toWorld(const V3& l) const { return t * l.x + b * l.y + n * l.z; }
    V3 toLocal(const V3& w) const { return V3(dot(w, t), dot(w, b), dot(w, n)); }
};

// GGXDistribution, microfacet_bsdf.cpp:9-58
struct GGX {
    float alpha;
    float D(const V3& wh) const {
        const float cos2 = wh.z * wh.z;
        if (cos2 <= 0.0f) return 0.0f;
        const float a2 = alpha * alpha;
        const float d = cos2 * (a2 - 1.0f) + 1.0f;
        return a2 / (PI * d * d);
    }
    float lambda(const V3& w) const {
        const float cos2 = w.z * w.z;
        if (cos2 <= 0.0f) return 0.0f;
        const float tan2 = std::max(0.0f, 1.0f - cos2) / cos2;
        return 0.5f * (std::sqrt(1.0f + alpha * alpha * tan2) - 1.0f);
    }
    float G1(const V3& w) const { return 1.0f / (1.0f + lambda(w)); }
    float G(const V3& wo, const V3& wi) const { return 1.0f / (1.0f + lambda(wo) + lambda(wi)); }
    float pdf(const V3& wo, const V3& wh) const {
        const float cosO = std::abs(wo.z);
        if (cosO == 0.0f) return 0.0f;
        return D(wh) * G1(wo) * std::abs(dot(wo, wh)) / cosO;
    }
    V3 sampleWh(const V3& wo, float u1, float u2) const {
        const bool flip = wo.z < 0.0f;
        const V3 v = flip ? -wo : wo;
        const V3 vh = V3(alpha * v.x, alpha * v.y, v.z).normalized();
        const V3 t1 = vh.z < 0.9999f ? cross(V3(0.0f, 0.0f, 1.0f), vh).normalized() : V3(1.0f, 0.0f, 0.0f);
        const V3 t2 = cross(vh, t1);
        const float r = std::sqrt(u1);
        const float phi = TWO_PI * u2;
        const float p1 = r * std::cos(phi);
        float p2 = r * std::sin(phi);
        const float s = 0.5f * (1.0f + vh.z);
        p2 = (1.0f - s) * std::sqrt(std::max(0.0f, 1.0f - p1 * p1)) + s * p2;
        const V3 nh = t1 * p1 + t2 * p2 + vh * std::sqrt(std::max(0.0f, 1.0f - p1 * p1 - p2 * p2));
        const V3 wh = V3(alpha * nh.x, alpha * nh.y, std::max(1e-6f, nh.z)).normalized();
        return flip ? -wh : wh;
    }
};

// MicrofacetTransmissionBSDF::halfVector, microfacet_bsdf.cpp:110-124
bool halfVectorT(const Lobe& l, const V3& wo, const V3& wi, V3& wh, float& eta) {
    if (sameHemisphere(wo, wi)) return false;
    if (wo.z == 0.0f || wi.z == 0.0f) return false;
    eta = wo.z > 0.0f ? l.etaIn / l.etaOut : l.etaOut / l.etaIn;
    wh = wo + wi * eta;
    if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return false;
    wh = wh.normalized();
    if (dot(wo, wh) < 0.0f) wh = -wh;
    return wh.z * wo.z > 0.0f && dot(wi, wh) <= 0.0f;
}

}  // namespace

// fresnel.h:16-32
float fresnelDielectric(float cosThetaI, float etaI, float etaT) {
    cosThetaI = clampf(cosThetaI, -1.0f, 1.0f);
    if (cosThetaI < 0.0f) {
        std::swap(etaI, etaT);
        cosThetaI = -cosThetaI;
    }
    const float sinThetaI = std::sqrt(std::max(0.0f, 1.0f - cosThetaI * cosThetaI));
    const float sinThetaT = etaI / etaT * sinThetaI;
    if (sinThetaT >= 1.0f) return 1.0f;
    const float cosThetaT = std::sqrt(std::max(0.0f, 1.0f - sinThetaT * sinThetaT));
    const float rParl = (etaT * cosThetaI - etaI * cosThetaT) / (etaT * cosThetaI + etaI * cosThetaT);
    const float rPerp = (etaI * cosThetaI - etaT * cosThetaT) / (etaI * cosThetaI + etaT * cosThetaT);
    return 0.5f * (rParl * rParl + rPerp * rPerp);
}

// fresnel.h:41-80
Color Fresnel::operator()(float cosThetaI) const {
    switch (kind) {
        case Dielectric: return Color(fresnelDielectric(cosThetaI, etaI, etaT));
        case Schlick: {
            const float m = 1.0f - clampf(std::abs(cosThetaI), 0.0f, 1.0f);
            const float m2 = m * m;
            return f0 + (Color(1.0f) - f0) * (m2 * m2 * m);
        }
        default: return Color(1.0f);
    }
}

// ---------------------------------------------------------------- lobes

Color Lobe::evaluate(const V3& wo, const V3& wi) const {
    switch (type) {
        case L_LAMBERT:
            if (!sameHemisphere(wo, wi)) return Color(0.0f);
            return R * INV_PI;
        case L_MF_REFL: {
            if (!sameHemisphere(wo, wi)) return Color(0.0f);
            const float cosO = absCos(wo), cosI = absCos(wi);
            if (cosO == 0.0f || cosI == 0.0f) return Color(0.0f);
            V3 wh = wo + wi;
            if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return Color(0.0f);
            wh = wh.normalized();
            if (wh.z < 0.0f) wh = -wh;
            const Color F = fresnel(dot(wo, wh));
            GGX d{alpha};
            return R * F * (d.D(wh) * d.G(wo, wi) / (4.0f * cosO * cosI));
        }
        case L_MF_TRANS: {
            V3 wh;
            float eta;
            if (!halfVectorT(*this, wo, wi, wh, eta)) return Color(0.0f);
            const float cosO = wo.z, cosI = wi.z;
            const float dotOH = dot(wo, wh), dotIH = dot(wi, wh);
            const float F = fresnelDielectric(wo.z > 0.0f ? dotOH : -dotOH, etaOut, etaIn);
            const float denom = dotOH + eta * dotIH;
            if (denom == 0.0f) return Color(0.0f);
            GGX d{alpha};
            const float value = d.D(wh) * d.G(wo, wi) * std::abs(dotIH * dotOH) / std::abs(cosI * cosO * denom * denom);
            return R * ((1.0f - F) * value);
        }
        default:
            return Color(0.0f);  // delta lobes
    }
}

float Lobe::pdf(const V3& wo, const V3& wi) const {
    switch (type) {
        case L_LAMBERT:
            if (!sameHemisphere(wo, wi)) return 0.0f;
            return absCos(wi) * INV_PI;
        case L_MF_REFL: {
            if (!sameHemisphere(wo, wi)) return 0.0f;
            V3 wh = wo + wi;
            if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return 0.0f;
            wh = wh.normalized();
            if (wh.z < 0.0f) wh = -wh;
            const float dotOH = dot(wo, wh);
            if (dotOH == 0.0f) return 0.0f;
            return GGX{alpha}.pdf(wo, wh) / (4.0f * std::abs(dotOH));
        }
        case L_MF_TRANS: {
            V3 wh;
            float eta;
            if (!halfVectorT(*this, wo, wi, wh, eta)) return 0.0f;
            const float dotOH = dot(wo, wh), dotIH = dot(wi, wh);
            const float denom = dotOH + eta * dotIH;
            if (denom == 0.0f) return 0.0f;
            const float dwhDwi = std::abs(eta * eta * dotIH) / (denom * denom);
            return GGX{alpha}.pdf(wo, wh) * dwhDwi;
        }
        default:
            return 0.0f;
    }
}

BSDFSample Lobe::sample(const V3& wo, float u1, float u2) const {
    BSDFSample r;
    switch (type) {
        case L_LAMBERT:
            r.wi = cosineHemisphere(u1, u2);
            if (wo.z < 0.0f) r.wi.z = -r.wi.z;
            r.f = R * INV_PI;
            r.pdf = absCos(r.wi) * INV_PI;
            return r;
        case L_SPEC_REFL: {
            const float cosO = absCos(wo);
            if (cosO == 0.0f) return r;
            r.wi = V3(-wo.x, -wo.y, wo.z);
            r.f = R * fresnel(wo.z) / cosO;
            r.pdf = 1.0f;
            r.isDelta = true;
            return r;
        }
        case L_SPEC_TRANS: {
            if (wo.z == 0.0f) return r;
            const bool entering = wo.z > 0.0f;
            const float etaI = entering ? etaOut : etaIn;
            const float etaT = entering ? etaIn : etaOut;
            const float eta = etaI / etaT;
            const float cosI = absCos(wo);
            const float sin2T = eta * eta * std::max(0.0f, 1.0f - cosI * cosI);
            if (sin2T >= 1.0f) return r;
            const float cosT = std::sqrt(1.0f - sin2T);
            r.wi = V3(-eta * wo.x, -eta * wo.y, entering ? -cosT : cosT);
            const float fr = fresnelDielectric(wo.z, etaOut, etaIn);
            r.f = R * ((1.0f - fr) * eta * eta / cosT);
            r.pdf = 1.0f;
            r.isDelta = true;
            return r;
        }
        case L_MF_REFL: {
            if (wo.z == 0.0f) return r;
            const V3 wh = GGX{alpha}.sampleWh(wo, u1, u2);
            const V3 wi = wh * (2.0f * dot(wo, wh)) - wo;
            if (!sameHemisphere(wo, wi)) return r;
            const float p = pdf(wo, wi);
            if (p <= 0.0f) return r;
            r.wi = wi;
            r.pdf = p;
            r.f = evaluate(wo, wi);
            return r;
        }
        case L_MF_TRANS: {
            if (wo.z == 0.0f) return r;
            const V3 wh = GGX{alpha}.sampleWh(wo, u1, u2);
            const float dotOH = dot(wo, wh);
            if (dotOH <= 0.0f) return r;
            const bool entering = wo.z > 0.0f;
            const float etaRatio = entering ? etaOut / etaIn : etaIn / etaOut;
            const float sin2T = etaRatio * etaRatio * std::max(0.0f, 1.0f - dotOH * dotOH);
            if (sin2T >= 1.0f) return r;
            const float cosT = std::sqrt(1.0f - sin2T);
            const V3 wi = wh * (etaRatio * dotOH - cosT) - wo * etaRatio;
            if (sameHemisphere(wo, wi) || wi.z == 0.0f) return r;
            const float p = pdf(wo, wi);
            if (p <= 0.0f) return r;
            r.wi = wi;
            r.pdf = p;
            r.f = evaluate(wo, wi);
            return r;
        }
    }
    return r;
}

Color Lobe::rho(const V3& wo) const {
    switch (type) {
        case L_LAMBERT: return R;
        case L_SPEC_REFL:
        case L_MF_REFL: return R * fresnel(wo.z);
        default: return R * (1.0f - fresnelDielectric(wo.z, etaOut, etaIn));
    }
}

// ---------------------------------------------------------------- mixture

void Mixture::add(const Lobe& l, float weight, bool substrate) {
    if (weight <= 0.0f || n >= 6) return;
    lobes[n] = l;
    lobes[n].weight = weight;
    lobes[n].substrate = substrate;
    ++n;
}

float Mixture::effectiveWeight(const Lobe& l, const V3& wo) const {
    if (!l.substrate) return l.weight;
    return l.weight * std::max(0.0f, 1.0f - substrateFresnel(wo.z).x);
}

float Mixture::selection(const V3& wo, float* w, float* s) const {
    float total = 0.0f;
    for (int i = 0; i < n; ++i) {
        w[i] = effectiveWeight(lobes[i], wo);
        s[i] = w[i] * lobes[i].rho(wo).maxComponent();
        total += s[i];
    }
    return total;
}

Color Mixture::evaluate(const V3& wo, const V3& wi) const {
    Color f(0.0f);
    for (int i = 0; i < n; ++i) {
        if (lobes[i].isDelta()) continue;
        f += lobes[i].evaluate(wo, wi) * effectiveWeight(lobes[i], wo);
    }
    return f;
}

float Mixture::pdf(const V3& wo, const V3& wi) const {
    float w[6], s[6];
    const float total = selection(wo, w, s);
    if (total <= 0.0f) return 0.0f;
    float p = 0.0f;
    for (int i = 0; i < n; ++i) {
        if (lobes[i].isDelta()) continue;
        p += s[i] * lobes[i].pdf(wo, wi);
    }
    return p / total;
}

BSDFSample Mixture::sample(const V3& wo, float u1, float u2) const {
    BSDFSample none;
    float w[6], s[6];
    const float total = selection(wo, w, s);
    if (total <= 0.0f) return none;
    float u = u1 * total;
    int chosen = n - 1;
    for (int i = 0; i < n; ++i) {
        if (s[i] <= 0.0f) continue;
        if (u < s[i] || i + 1 == n) { chosen = i; break; }
        u -= s[i];
    }
    if (s[chosen] <= 0.0f) return none;
    const float v1 = std::min(u / s[chosen], 0.99999994f);
    const Lobe& lobe = lobes[chosen];
    BSDFSample r = lobe.sample(wo, v1, u2);
    if (r.pdf <= 0.0f || r.f.isBlack()) return none;
    const float pChosen = s[chosen] / total;
    if (lobe.isDelta()) {
        r.f = r.f * w[chosen];
        r.pdf *= pChosen;
        return r;
    }
    r.f = evaluate(wo, r.wi);
    r.pdf = pdf(wo, r.wi);
    if (r.pdf <= 0.0f) return none;
    return r;
}

bool Mixture::isDelta() const {
    if (n == 0) return false;
    for (int i = 0; i < n; ++i)
        if (!lobes[i].isDelta()) return false;
    return true;
}

// material.cpp:69-108
Mixture Material::createBSDF() const {
    const float m = clampf(metallic, 0.0f, 1.0f);
    const float t = clampf(transmission, 0.0f, 1.0f);
    const float dielectric = 1.0f - m;
    const bool smooth = roughness < 0.03f;
    const bool hasInterface = std::abs(ior - 1.0f) > 1e-3f;
    const bool tinted = !baseColor.isBlack();
    const float a = std::max(roughness * roughness, 1e-3f);
    Mixture mix;
    auto refl = [&](const Fresnel& f) {
        Lobe l{};
        l.type = smooth ? L_SPEC_REFL : L_MF_REFL;
        l.R = Color(1.0f);
        l.fresnel = f;
        l.alpha = a;
        return l;
    };
    if (m > 0.0f && tinted) {
        Fresnel f;
        f.kind = Fresnel::Schlick;
        f.f0 = baseColor;
        mix.add(refl(f), m);
    }
    if (dielectric > 0.0f) {
        if (hasInterface) {
            Fresnel f;
            f.kind = Fresnel::Dielectric;
            f.etaI = 1.0f;
            f.etaT = ior;
            mix.add(refl(f), dielectric);
            mix.substrateFresnel = f;
            if (t > 0.0f && tinted) {
                Lobe l{};
                l.type = smooth ? L_SPEC_TRANS : L_MF_TRANS;
                l.R = baseColor;
                l.etaOut = 1.0f;
                l.etaIn = ior;
                l.alpha = a;
                mix.add(l, dielectric * t);
            }
            if (t < 1.0f && tinted) {
                Lobe l{};
                l.type = L_LAMBERT;
                l.R = baseColor;
                mix.add(l, dielectric * (1.0f - t), true);
            }
        } else {
            if (t > 0.0f && tinted) {
                Lobe l{};
                l.type = L_SPEC_TRANS;
                l.R = baseColor;
                l.etaOut = 1.0f;
                l.etaIn = 1.0f;
                mix.add(l, dielectric * t);
            }
            if (t < 1.0f && tinted) {
                Lobe l{};
                l.type = L_LAMBERT;
                l.R = baseColor;
                mix.add(l, dielectric * (1.0f - t));
            }
        }
    }
    return mix;
}

// ---------------------------------------------------------------- lights, media, camera

LightSample Light::sample(const V3& p, float u1, float u2) const {
    LightSample s;
    if (type == RBR_LIGHT_AREA) {  // area_light.cpp:13-39
        V3 lp = position + u * (u1 - 0.5f) + v * (u2 - 0.5f);
        V3 toLight = lp - p;
        float dist2 = dot(toLight, toLight);
        s.distance = std::sqrt(dist2);
        s.wi = toLight / s.distance;
        float cosTheta = -dot(normal, s.wi);
        if (cosTheta <= 0.0f) {
            s.Li = Color(0.0f);
            s.pdf = 0.0f;
            return s;
        }
        float area = cross(u, v).length();
        s.pdf = dist2 / (cosTheta * area);
        s.Li = color * intensity;
        s.isDelta = false;
    } else if (type == RBR_LIGHT_POINT) {  // point_light.cpp:5-19
        V3 toLight = position - p;
        float dist2 = dot(toLight, toLight);
        s.distance = std::sqrt(dist2);
        s.wi = toLight / s.distance;
        s.Li = color * intensity / dist2;
        s.pdf = 1.0f;
        s.isDelta = true;
    } else {  // directional_light.cpp:5-15
        s.wi = -direction;
        s.distance = INF;
        s.Li = color * intensity;
        s.pdf = 1.0f;
        s.isDelta = true;
    }
    return s;
}

float Medium::densityScale(const V3& p) const { return std::exp(-falloff * std::max(0.0f, p.y - baseHeight)); }

Color Medium::sigmaT(const V3& p) const {
    if (type == RBR_MEDIUM_EXPONENTIAL) {
        float d = densityScale(p);
        return sigma_a * d + sigma_s * d;   // Medium::getSigmaT = getSigmaA + getSigmaS
    }
    return sigma_a + sigma_s;
}

Color Medium::sigmaS(const V3& p) const {
    return type == RBR_MEDIUM_EXPONENTIAL ? sigma_s * densityScale(p) : sigma_s;
}

// medium.cpp:35-43 and 87-103
Color Medium::transmittance(const Ray& r, float distance) const {
    if (type == RBR_MEDIUM_EXPONENTIAL) {
        const int steps = 64;
        float stepSize = distance / steps;
        Color od(0.0f);
        for (int i = 0; i < steps; ++i) {
            V3 p = r.at((i + 0.5f) * stepSize);
            od += sigmaT(p) * stepSize;
        }
        return Color(std::exp(-od.x), std::exp(-od.y), std::exp(-od.z));
    }
    Color st = sigma_a + sigma_s;
    return Color(std::exp(-st.x * distance), std::exp(-st.y * distance), std::exp(-st.z * distance));
}

// medium.cpp:9-33 and 61-85
bool Medium::sampleDistance(const Ray& r, float maxDist, float u, float& t, Color& weight) const {
    Color st = sigma_a + sigma_s;
    float avg = (st.x + st.y + st.z) / 3.0f;
    if (type == RBR_MEDIUM_EXPONENTIAL) {
        // The reference's march re-uses u, so its step is the constant -ln(1-u)/avg.  It never
        // ends for a ray with nothing behind it (maxDist = inf) and takes maxDist/step steps
        // for tiny u; both are resolved in closed form (first step whose height satisfies
        // u < exp(-falloff*max(0, y - base))), an escaping ray ending with zero weight.
        const float step = -std::log(1.0f - u) / avg;
        if (step > 0.0f && (std::isinf(maxDist) || maxDist / step > 1e7f)) {
            float tt = INF;
            if (falloff <= 0.0f) {
                tt = step;
            } else {
                const float ystar = baseHeight + (-std::log(u)) / falloff;
                if (r.o.y + r.d.y * step < ystar) tt = step;
                else if (r.d.y < 0.0f) {
                    float k = std::ceil((ystar - r.o.y) / (r.d.y * step));
                    tt = std::max(k, 1.0f) * step;
                    if (r.o.y + r.d.y * tt >= ystar) tt += step;
                }
            }
            if (tt < maxDist) {
                t = tt;
                V3 p = r.at(t);
                weight = sigmaS(p) / sigmaT(p);
                return true;
            }
            weight = std::isinf(maxDist) ? Color(0.0f) : transmittance(r, maxDist);
            return false;
        }
        t = 0.0f;
        while (true) {
            t -= std::log(1.0f - u) / avg;
            if (t >= maxDist) {
                weight = transmittance(r, maxDist);
                return false;
            }
            V3 p = r.at(t);
            Color s = sigmaT(p);
            float density = (s.x + s.y + s.z) / 3.0f;
            if (u < density / avg) {
                weight = sigmaS(p) / s;
                return true;
            }
        }
    }
    if (avg < 1e-6f) {
        weight = Color(1.0f);
        return false;
    }
    t = -std::log(1.0f - u) / avg;
    if (t < maxDist) {
        Color tr = transmittance(r, t);
        weight = sigma_s * tr / (avg * avg);
        return true;
    }
    weight = transmittance(r, maxDist);
    return false;
}

// phase_function.cpp:17-47
float Medium::hgEvaluate(const V3& wo, const V3& wi) const {
    float cosTheta = dot(wo, wi);
    float denom = 1.0f + g * g + 2.0f * g * cosTheta;
    return (1.0f - g * g) / (4.0f * PI * denom * std::sqrt(denom));
}

V3 Medium::hgSample(const V3& wo, float u1, float u2) const {
    float cosTheta;
    if (std::abs(g) < 1e-3f) {
        cosTheta = 1.0f - 2.0f * u1;
    } else {
        float sqrTerm = (1.0f - g * g) / (1.0f + g - 2.0f * g * u1);
        cosTheta = -(1.0f + g * g - sqrTerm * sqrTerm) / (2.0f * g);
    }
    float sinTheta = std::sqrt(std::max(0.0f, 1.0f - cosTheta * cosTheta));
    float phi = TWO_PI * u2;
    V3 t, b;
    coordinateSystem(wo, t, b);
    return t * (sinTheta * std::cos(phi)) + b * (sinTheta * std::sin(phi)) + wo * cosTheta;
}

// camera.cpp:5-9 with PerspectiveIntrinsics::generateRay (perspective_intrinsics.h:43-45);
// the direction is normalized there and again by Ray's constructor.
Ray Camera::generateRay(float px, float py) const {
    V3 local((px - cx) / fx, (cy - py) / fy, 1.0f);
    V3 d = right * local.x + up * local.y + forward * local.z;
    Ray r;
    r.o = position + (right * 0.0f + up * 0.0f + forward * 0.0f);
    r.d = d.normalized().normalized();
    r.tmin = EPSILON;
    r.tMax = INF;
    return r;
}

// interaction.h:32-39
Ray spawnRay(const Isect& is, const V3& dir) {
    V3 offset = is.n * EPSILON;
    if (dot(dir, is.n) < 0.0f) offset = -offset;
    Ray r;
    r.o = is.p + offset;
    r.d = dir.normalized();
    r.tmin = EPSILON;
    r.tMax = INF;
    return r;
}

bool RenderScene::intersect(const Ray& r, Isect& is, int& mat) const {
    float t;
    uint32_t id = geo.closestBVH(r, t);
    if (id == 0xFFFFFFFFu) return false;
    geo.fill(id, r, t, is);
    mat = id < primMaterial.size() ? primMaterial[id] : -1;
    return true;
}

// ---------------------------------------------------------------- integrators

namespace {

// PathTracingIntegrator::sampleLightDirect, path_tracing_integrator.cpp:83-138
Color sampleLightDirect(const RenderScene& s, const Config& c, const Isect& is, const Mixture& bsdf, const Light& light,
                        Sampler& smp) {
    Color Ld(0.0f);
    auto ls = light.sample(is.p, smp.get1D(), smp.get1D());
    if (ls.pdf > 0.0f && !ls.Li.isBlack()) {
        const Frame frame(is.outward());
        V3 woL = frame.toLocal(is.wo);
        V3 wiL = frame.toLocal(ls.wi);
        Color f = bsdf.evaluate(woL, wiL);
        float bsdfPdf = bsdf.pdf(woL, wiL);
        if (!f.isBlack()) {
            Ray sh = spawnRay(is, ls.wi);
            sh.tMax = ls.distance - EPSILON;
            if (!s.occluded(sh)) {
                float cosTheta = std::abs(dot(ls.wi, is.n));
                if (c.useMIS && !light.isDelta()) {
                    float w = powerHeuristic(1, ls.pdf, 1, bsdfPdf);
                    Ld += f * ls.Li * cosTheta * w / ls.pdf;
                } else {
                    Ld += f * ls.Li * cosTheta / ls.pdf;
                }
            }
        }
    }
    // The MIS BSDF sample is drawn (consuming two numbers) and its ray traced, but its result
    // is never used (path_tracing_integrator.cpp:117-135); only the draw is kept here.
    if (c.useMIS && !light.isDelta()) {
        const Frame frame(is.outward());
        V3 woL = frame.toLocal(is.wo);
        (void)bsdf.sample(woL, smp.get1D(), smp.get1D());
    }
    return Ld;
}

// VolumetricPathIntegrator::directLightingVolume, volumetric_path_integrator.cpp:91-119
Color directLightingVolume(const RenderScene& s, const V3& p, const V3& wo) {
    Color Ld(0.0f);
    for (const Light& l : s.lights) {
        if (l.type != RBR_LIGHT_POINT) continue;
        V3 toLight = l.position - p;
        float dist = toLight.length();
        V3 wi = toLight / dist;
        Ray sh;
        sh.o = p;
        sh.d = wi.normalized();
        sh.tmin = EPSILON;
        sh.tMax = dist;
        if (s.occluded(sh)) continue;
        Color tr = s.medium.transmittance(sh, dist);
        Color Li = l.color * l.intensity / (dist * dist);
        float ph = s.medium.hgEvaluate(wo, wi);
        Ld += Li * tr * ph;
    }
    return Ld;
}

// VolumetricPathIntegrator::directLightingSurface, volumetric_path_integrator.cpp:121-143
Color directLightingSurface(const RenderScene& s, const Isect& is, const Mixture& bsdf, Sampler& smp) {
    Color Ld(0.0f);
    for (const Light& l : s.lights) {
        auto ls = l.sample(is.p, smp.get1D(), smp.get1D());
        if (ls.pdf <= 0.0f || ls.Li.isBlack()) continue;
        Ray sh = spawnRay(is, ls.wi);
        sh.tMax = ls.distance - EPSILON;
        if (s.occluded(sh)) continue;
        const Frame frame(is.outward());
        V3 woL = frame.toLocal(is.wo);
        V3 wiL = frame.toLocal(ls.wi);
        Color f = bsdf.evaluate(woL, wiL);
        float cosTheta = std::abs(dot(ls.wi, is.n));
        Ld += f * ls.Li * cosTheta / ls.pdf;
    }
    return Ld;
}

}  // namespace

Color LiPath(const RenderScene& s, const Config& c, Ray ray, Sampler& smp) {
    Color L(0.0f), T(1.0f);
    bool specularBounce = true;
    for (int depth = 0; depth < c.maxDepth; ++depth) {
        Isect is;
        int mat;
        if (!s.intersect(ray, is, mat)) {
            L += T * s.background;
            break;
        }
        if (mat < 0) break;
        const Material& m = s.materials[mat];
        if (m.isEmissive() && (depth == 0 || specularBounce)) L += T * m.emission;
        Mixture bsdf = m.createBSDF();
        const Frame frame(is.outward());
        V3 woL = frame.toLocal(is.wo);
        if (c.useNEE && !bsdf.isDelta()) {
            Color Ld(0.0f);
            for (const Light& l : s.lights) Ld += sampleLightDirect(s, c, is, bsdf, l, smp);
            L += T * Ld;
        }
        auto bs = bsdf.sample(woL, smp.get1D(), smp.get1D());
        if (bs.pdf <= 0.0f || bs.f.isBlack()) break;
        V3 wi = frame.toWorld(bs.wi);
        float cosTheta = std::abs(dot(wi, is.n));
        T *= bs.f * (cosTheta / bs.pdf);
        if (depth > 3) {
            float q = std::max(0.05f, 1.0f - T.maxComponent());
            if (smp.get1D() < q) break;
            T /= (1.0f - q);
        }
        ray = spawnRay(is, wi);
        specularBounce = bs.isDelta;
    }
    return L;
}

Color LiVolPath(const RenderScene& s, const Config& c, Ray ray, Sampler& smp) {
    Color L(0.0f), T(1.0f);
    const bool inMedium = s.medium.type != RBR_MEDIUM_NONE;
    for (int depth = 0; depth < c.maxDepth; ++depth) {
        Isect is;
        int mat = -1;
        bool hit = s.intersect(ray, is, mat);
        if (inMedium) {
            float maxDist = hit ? is.t : INF;
            float t;
            Color w;
            bool scattered = s.medium.sampleDistance(ray, maxDist, smp.get1D(), t, w);
            T *= w;
            if (scattered) {
                V3 sp = ray.at(t);
                if (c.useNEE) L += T * directLightingVolume(s, sp, ray.d);
                V3 wi = s.medium.hgSample(ray.d, smp.get1D(), smp.get1D());
                Ray nr;
                nr.o = sp;
                nr.d = wi.normalized();
                nr.tmin = EPSILON;
                nr.tMax = INF;
                ray = nr;
                continue;
            }
        }
        if (!hit) {
            L += T * s.background;
            break;
        }
        if (mat < 0) break;
        const Material& m = s.materials[mat];
        if (m.isEmissive()) L += T * m.emission;
        Mixture bsdf = m.createBSDF();
        const Frame frame(is.outward());
        V3 woL = frame.toLocal(is.wo);
        if (c.useNEE && !bsdf.isDelta()) L += T * directLightingSurface(s, is, bsdf, smp);
        auto bs = bsdf.sample(woL, smp.get1D(), smp.get1D());
        if (bs.pdf <= 0.0f || bs.f.isBlack()) break;
        V3 wi = frame.toWorld(bs.wi);
        float cosTheta = std::abs(dot(wi, is.n));
        T *= bs.f * (cosTheta / bs.pdf);
        if (depth > 3) {
            float q = std::max(0.05f, 1.0f - T.maxComponent());
            if (smp.get1D() < q) break;
            T /= (1.0f - q);
        }
        ray = spawnRay(is, wi);
    }
    return L;
}

}  // namespace orc
