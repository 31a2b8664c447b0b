// CPU restatement of the reference's shading and integrators.  TEST INFRASTRUCTURE ONLY.
//
//   Material::createBSDF                     src/materials/material.cpp:69-108
//   MixtureBSDF                              src/materials/mixture_bsdf.cpp:5-104
//   LambertianBSDF                           src/materials/lambertian_bsdf.cpp:5-28
//   SpecularReflection/TransmissionBSDF      src/materials/specular_bsdf.cpp:5-57
//   GGXDistribution, Microfacet*BSDF         src/materials/microfacet_bsdf.cpp:9-179
//   fresnelDielectric / Schlick / Fresnel    include/render/materials/fresnel.h:16-80
//   Area/Point/Directional light sample      src/lighting/*_light.cpp
//   Homogeneous/ExponentialMedium            src/materials/medium.cpp:9-103
//   HenyeyGreensteinPhase                    src/materials/phase_function.cpp:17-47
//   sampling::*                              include/render/core/sampler.h:58-149
//   PinholeCamera::generateRay               src/camera/camera.cpp:5-9, perspective_intrinsics.h:43-45
//   Integrator::render                       src/renderers/integrator.cpp:7-50
//   PathTracingIntegrator::Li                src/renderers/path_tracing_integrator.cpp:5-138
//   VolumetricPathIntegrator::Li             src/renderers/volumetric_path_integrator.cpp:6-143
#pragma once

#include <random>
#include <vector>

#include "../include/rbr.h"
#include "orc_scene.h"

namespace orc {

using Color = V3;

struct Fresnel {
    enum Kind { None, Dielectric, Schlick } kind = None;
    float etaI = 1.0f, etaT = 1.0f;
    Color f0{1.0f};
    Color operator()(float cosThetaI) const;
};

float fresnelDielectric(float cosThetaI, float etaI, float etaT);

enum LobeType { L_LAMBERT, L_SPEC_REFL, L_SPEC_TRANS, L_MF_REFL, L_MF_TRANS };

struct BSDFSample {
    V3 wi{0.0f};
    Color f{0.0f};
    float pdf = 0.0f;
    bool isDelta = false;
};

struct Lobe {
    LobeType type;
    Color R;              // albedo / reflectance / transmittance
    Fresnel fresnel;      // reflection lobes
    float etaOut = 1.0f, etaIn = 1.0f;   // transmission lobes
    float alpha = 1.0f;   // GGX (already max(alpha, 1e-3))
    float weight = 0.0f;
    bool substrate = false;

    bool isDelta() const { return type == L_SPEC_REFL || type == L_SPEC_TRANS; }
    Color evaluate(const V3& wo, const V3& wi) const;
    BSDFSample sample(const V3& wo, float u1, float u2) const;
    float pdf(const V3& wo, const V3& wi) const;
    Color rho(const V3& wo) const;
};

struct Mixture {
    Lobe lobes[6];
    int n = 0;
    Fresnel substrateFresnel;
    void add(const Lobe& l, float weight, bool substrate = false);
    float effectiveWeight(const Lobe& l, const V3& wo) const;
    float selection(const V3& wo, float* w, float* s) const;
    Color evaluate(const V3& wo, const V3& wi) const;
    float pdf(const V3& wo, const V3& wi) const;
    BSDFSample sample(const V3& wo, float u1, float u2) const;
    bool isDelta() const;
};

struct Material {
    Color baseColor{0.8f};
    float metallic = 0.0f, roughness = 1.0f, ior = 1.5f, transmission = 0.0f;
    Color emission{0.0f};
    bool isEmissive() const { return emission.x > 0.0f || emission.y > 0.0f || emission.z > 0.0f; }
    Mixture createBSDF() const;
};

struct LightSample {
    V3 wi{0.0f};
    Color Li{0.0f};
    float pdf = 0.0f, distance = 0.0f;
    bool isDelta = false;
};

struct Light {
    int type;     // RBR_LIGHT_*
    Color color;
    float intensity;
    V3 position, u, v, normal, direction;
    bool isDelta() const { return type != RBR_LIGHT_AREA; }
    LightSample sample(const V3& p, float u1, float u2) const;
};

struct Medium {
    int type = RBR_MEDIUM_NONE;
    Color sigma_a, sigma_s;
    float falloff = 0.0f, baseHeight = 0.0f, g = 0.0f;
    bool sampleDistance(const Ray& r, float maxDist, float u, float& t, Color& weight) const;
    Color transmittance(const Ray& r, float dist) const;
    Color sigmaT(const V3& p) const;
    Color sigmaS(const V3& p) const;
    float densityScale(const V3& p) const;
    float hgEvaluate(const V3& wo, const V3& wi) const;
    V3 hgSample(const V3& wo, float u1, float u2) const;
};

struct Camera {
    V3 position, forward, right, up;
    float fx, fy, cx, cy;
    int width, height;
    Ray generateRay(float px, float py) const;
};

// One sampler per thread: the reference's UniformSampler (sampler.cpp:5-18).
struct Sampler {
    std::mt19937_64 rng;
    std::uniform_real_distribution<float> dist{0.0f, 1.0f};
    float get1D() { return dist(rng); }
    void setSeed(uint64_t s) { rng.seed(s); }
};

struct RenderScene {
    Scene geo;
    std::vector<Material> materials;
    std::vector<int32_t> primMaterial;   // by primitive id, -1 = none
    std::vector<Light> lights;
    Medium medium;
    Color background{0.0f};

    bool intersect(const Ray& r, Isect& is, int& mat) const;
    bool occluded(const Ray& r) const { return geo.anyBVH(r); }
};

struct Config {
    int integrator = RBR_INTEGRATOR_PATH;
    int spp = 1, maxDepth = 8;
    bool useNEE = true, useMIS = true;
};

Color LiPath(const RenderScene& s, const Config& c, Ray ray, Sampler& smp);
Color LiVolPath(const RenderScene& s, const Config& c, Ray ray, Sampler& smp);

// Spawned rays re-normalize their direction, as Ray's constructor does (ray.h:19-23).
Ray spawnRay(const Isect& is, const V3& dir);

}  // namespace orc
