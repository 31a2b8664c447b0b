// CPU restatement of the reference's primitives and BVH.  TEST INFRASTRUCTURE ONLY.
//
//   Triangle::intersect (Moller-Trumbore)   src/geometry/triangle.cpp:6-65 (vertex normals 40-48)
//   Sphere::intersect                        src/geometry/sphere.cpp:6-43
//   Quad::intersect                          src/geometry/quad.cpp:12-42
//   Plane::intersect                         src/geometry/plane.cpp:12-35
//   SurfaceInteraction::setFaceNormal        include/render/core/interaction.h:74-83
//   BVH::build / findBestSplit               src/acceleration/bvh.cpp:10-31, 47-209
//   BVH::intersectNode / intersectNodeP      src/acceleration/bvh.cpp:211-267
//
// One deliberate difference from the reference: among hits with equal t the LOWEST
// primitive id wins (the reference keeps whichever equal-t primitive it tested last).
// The BVH traversal tests node boxes grown outward by a small margin so that rounding in
// the slab test can never hide a primitive whose Moller-Trumbore t is accepted; the BVH
// therefore returns exactly what the brute-force loop returns.
#pragma once

#include "orc_math.h"

#include <memory>
#include <vector>

namespace orc {

enum PrimType : uint32_t { PT_TRI = 0, PT_SPHERE = 1, PT_QUAD = 2, PT_PLANE = 3 };

struct Tri {
    V3 v0, v1, v2;
    uint32_t id;
    bool smooth = false;   // vertex normals n0..n2 (unit, as Triangle's constructor stores them)
    V3 n0, n1, n2;
};
struct Sph { V3 c; float r; uint32_t id; };
struct Quad { V3 origin, u, v, n; uint32_t id; };
struct Plane { V3 n; float d; uint32_t id; };

struct Isect {
    V3 p, n, gn, wo;
    float t = 0.0f;
    bool front = true;
    uint32_t id = 0xFFFFFFFFu;
    V3 outward() const { return front ? n : -n; }   // interaction.h:74-76
    void setFaceNormal(const Ray& r, const V3& out) {  // interaction.h:79-83
        front = dot(r.d, out) < 0.0f;
        n = front ? out : -out;
        gn = n;
    }
};

// Per-primitive tests.  They return the candidate t in [ray.tmin, ray.tMax] or false,
// without touching ray.tMax (the caller applies the lowest-id rule).
bool tri_t(const Tri& tr, const Ray& r, float& t);
bool sph_t(const Sph& s, const Ray& r, float& t);
bool quad_t(const Quad& q, const Ray& r, float& t);
bool plane_t(const Plane& p, const Ray& r, float& t);

struct PrimRef { uint32_t type, index; };

struct BNode {
    Box bounds, padded;
    std::unique_ptr<BNode> left, right;
    size_t start = 0, count = 0;
    bool leaf() const { return !left && !right; }
};

struct Scene {
    std::vector<Tri> tris;
    std::vector<Sph> sphs;
    std::vector<Quad> quads;
    std::vector<Plane> planes;

    // BVH over the bounded primitives (planes are tested linearly, as unbounded ones are in
    // accel_common.h:35-52).
    std::vector<PrimRef> ordered;
    std::unique_ptr<BNode> root;
    float pad = 0.0f;

    Box primBounds(const PrimRef& p) const;
    bool primT(const PrimRef& p, const Ray& r, float& t) const;
    uint32_t primId(const PrimRef& p) const;
    std::vector<PrimRef> allBounded() const;

    void buildBVH();
    // Closest hit with the lowest-id rule; returns id (0xFFFFFFFF on miss) and t.
    uint32_t closestBrute(const Ray& r, float& t) const;
    bool anyBrute(const Ray& r) const;
    uint32_t closestBVH(const Ray& r, float& t) const;
    bool anyBVH(const Ray& r) const;
    // Fills the surface interaction of primitive `id` hit at the given ray (re-running its
    // intersect with tMax = t, as the reference's fill does).
    bool fill(uint32_t id, const Ray& r, float t, Isect& is) const;

    // id -> (type, index)
    std::vector<PrimRef> byId;
    void indexIds();
};

}  // namespace orc
