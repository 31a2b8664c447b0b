// Device-side shading model: the reference's Material -> MixtureBSDF construction
// (material.cpp:69-108) evaluated without allocation, its lobes (lambertian_bsdf.cpp,
// specular_bsdf.cpp, microfacet_bsdf.cpp), Fresnel terms (fresnel.h), lights
// (area/point/directional_light.cpp), media + Henyey-Greenstein phase (medium.cpp,
// phase_function.cpp) and the sampling helpers of sampler.h.  The lobe list of a material
// depends only on the material, so the host precomputes it once (DMaterial) and a shading
// thread reads 16-word lobe records instead of building a heap mixture per bounce.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "prim_tests.h"

namespace rbr {

enum DLobeType : uint32_t { DL_LAMBERT = 0, DL_SPEC_REFL = 1, DL_SPEC_TRANS = 2, DL_MF_REFL = 3, DL_MF_TRANS = 4 };
enum DFresnel : uint32_t { DF_NONE = 0, DF_DIELECTRIC = 1, DF_SCHLICK = 2 };

struct DLobe {           // 64 bytes
    uint32_t type;
    float R[3];          // albedo / reflectance / transmittance
    uint32_t fkind;      // Fresnel of reflection lobes
    float fEtaI, fEtaT;  // dielectric Fresnel indices
    float f0[3];         // Schlick reflectance at normal incidence
    float etaOut, etaIn; // transmission lobes
    float alpha;         // GGX alpha, already max(roughness^2, 1e-3)
    float weight;
};

struct DMaterial {       // 6 lobes + header
    DLobe lobe[6];
    uint32_t n;
    uint32_t allDelta;   // MixtureBSDF::isDelta
    uint32_t emissive;
    uint32_t substrateMask;  // bit i: lobe i sits under the substrate Fresnel
    uint32_t sfKind;
    float sfEtaI, sfEtaT;
    float emission[3];
    // the material's own parameters, for the classic shading of the Whitted integrator
    float base[3], metallic, roughness, ior, transmission;
};

struct DLight {
    uint32_t type;
    float color[3];
    float intensity;
    float position[3];
    float u[3];
    float v[3];
    float normal[3];
    float direction[3];
    float area;          // |u x v|
    float pad[2];
};

struct DMedium {
    uint32_t type;
    float sa[3], ss[3];
    float falloff, base, g;
};

struct DCamera {
    float pos[3], fwd[3], right[3], up[3];
    float fx, fy, cx, cy;
    uint32_t width, height;
};

// One render job as the kernels see it.
struct DJob {
    DCamera cam;
    DMedium medium;
    const float4* primShade;     // 2 float4 per global primitive id
    const float4* smooth;        // 6 float4 per triangle with vertex normals: v0, e1, e2, n0, n1, n2
    const DMaterial* mats;
    const DLight* lights;
    uint32_t nLights;
    uint32_t integrator, maxDepth, useNEE, useMIS;
    uint32_t guided;             // bit 0: guided direction sampling, bit 1: record training samples
    uint32_t guideDepth;         // bounces below this draw from the guide
    uint32_t aoSamples;          // AO: rays per camera sample
    float aoDistance;            // AO: ray length
    float depthNear, depthFar;   // DEPTH: mapping range
    float ambient[3];            // WHITTED: Scene::ambientLight
    uint32_t root;               // BVH root node in the node arena
    uint32_t pathBase, pathCount;   // pool slots of this job in the current round
    uint32_t sampleBase;         // sample index of the job's first pool slot this round
    uint32_t pixels;             // width * height
    float bg[3];
    uint64_t seed;
    float* filmSum;              // 3 floats per pixel (device)
    double* momSum;              // optional: 3 doubles per pixel
    double* momSq;
};

// Outward surface normal at a hit from the shading records (2 float4 per primitive id):
// (normal, kind) (radius, smooth index, -, material).  Kinds 0 triangle, 2 quad and 3 plane
// hold the geometric normal; a sphere's (kind 1) is (P - centre) / radius; a triangle with
// vertex normals (kind 4) interpolates them at the hit (smoothNormal, Triangle::intersect,
// triangle.cpp:40-48), for the same ray o, d that found it.
__device__ __forceinline__ V hitOutward(const DJob& jb, uint32_t id, V o, V d, V P) {
    const float4 s0 = jb.primShade[2 * size_t(id)], s1 = jb.primShade[2 * size_t(id) + 1];
    const uint32_t kind = __float_as_uint(s0.w);
    if (kind == 1u) {
        const float inv = 1.0f / s1.x;
        return V{(P.x - s0.x) * inv, (P.y - s0.y) * inv, (P.z - s0.z) * inv};
    }
    if (kind == 4u) {
        const float4* r = jb.smooth + 6 * size_t(__float_as_uint(s1.y));
        const float4 a = r[0], b = r[1], c = r[2], n0 = r[3], n1 = r[4], n2 = r[5];
        return smoothNormal(o, d, V{a.x, a.y, a.z}, V{b.x, b.y, b.z}, V{c.x, c.y, c.z}, V{n0.x, n0.y, n0.z},
                            V{n1.x, n1.y, n1.z}, V{n2.x, n2.y, n2.z}, V{s0.x, s0.y, s0.z});
    }
    return V{s0.x, s0.y, s0.z};
}

}  // namespace rbr
