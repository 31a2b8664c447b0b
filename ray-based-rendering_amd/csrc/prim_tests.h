// Exact primitive tests shared by the host traversal and the HIP kernels (__host__
// __device__).  Every operation keeps the reference's operand order; the translation
// units that include this file are compiled with FP contraction off, so host and device
// produce bit-identical t values.
//   Triangle::intersect  triangle.cpp:6-24   (e1, e2 precomputed with the same subtraction)
//   Sphere::intersect    sphere.cpp:6-21
//   Quad::intersect      quad.cpp:12-31
//   Plane::intersect     plane.cpp:12-17
#pragma once

#ifdef __HIPCC__
#define RBR_HD __host__ __device__ __forceinline__
#else
#define RBR_HD inline
#endif

#include <cmath>

namespace rbr {

constexpr float kEps = 1e-4f;   // EPSILON, common.h:24

struct V { float x, y, z; };

RBR_HD V vsub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RBR_HD float vdot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RBR_HD V vcross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// Vec3::normalized (vector.h:209-215): v * (1 / |v|), or 0 for a zero vector.
RBR_HD V vnormalized(V a) {
    const float len = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    if (len > 0.0f) {
        const float inv = 1.0f / len;
        return {a.x * inv, a.y * inv, a.z * inv};
    }
    return {0.0f, 0.0f, 0.0f};
}

// Shading normal of a triangle with vertex normals where the ray meets it (Triangle::intersect,
// triangle.cpp:6-48): the barycentrics u, v from the same Moller-Trumbore terms as the hit,
// normalized(n0 w + n1 u + n2 v) with w = 1 - u - v, turned into the hemisphere of the geometric
// normal geo.  n0..n2 are the unit corner normals (Triangle's constructor normalizes them).
RBR_HD V smoothNormal(V o, V d, V v0, V e1, V e2, V n0, V n1, V n2, V geo) {
    const V h = vcross(d, e2);
    const float a = vdot(e1, h);
    const float f = 1.0f / a;
    const V s = vsub(o, v0);
    const float u = f * vdot(s, h);
    const V q = vcross(s, e1);
    const float v = f * vdot(d, q);
    const float w = 1.0f - u - v;
    V n = vnormalized({n0.x * w + n1.x * u + n2.x * v, n0.y * w + n1.y * u + n2.y * v, n0.z * w + n1.z * u + n2.z * v});
    if (vdot(n, geo) < 0.0f) n = {-n.x, -n.y, -n.z};
    return n;
}

// Returns true and t when the triangle (v0, e1, e2) is hit with t in [tmin, tmax].
RBR_HD bool triTest(V o, V d, V v0, V e1, V e2, float tmin, float tmax, float& tout) {
    V h = vcross(d, e2);
    float a = vdot(e1, h);
    if (fabsf(a) < kEps) return false;
    float f = 1.0f / a;
    V s = vsub(o, v0);
    float u = f * vdot(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    V q = vcross(s, e1);
    float v = f * vdot(d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = f * vdot(e2, q);
    if (t < tmin || t > tmax) return false;
    tout = t;
    return true;
}

RBR_HD bool sphereTest(V o, V d, V c, float r, float tmin, float tmax, float& tout) {
    V oc = vsub(o, c);
    float a = vdot(d, d);
    float halfB = vdot(oc, d);
    float cc = vdot(oc, oc) - r * r;
    float disc = halfB * halfB - a * cc;
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    float t = (-halfB - sq) / a;
    if (t < tmin || t > tmax) {
        t = (-halfB + sq) / a;
        if (t < tmin || t > tmax) return false;
    }
    tout = t;
    return true;
}

RBR_HD bool quadTest(V o, V d, V org, V u, V vv, V n, float tmin, float tmax, float& tout) {
    float denom = vdot(n, d);
    if (fabsf(denom) < kEps) return false;
    float t = vdot(vsub(org, o), n) / denom;
    if (t < tmin || t > tmax) return false;
    V hp = {o.x + d.x * t, o.y + d.y * t, o.z + d.z * t};
    V lp = vsub(hp, org);
    float uLen2 = vdot(u, u);
    float vLen2 = vdot(vv, vv);
    float uc = vdot(lp, u) / uLen2;
    float vc = vdot(lp, vv) / vLen2;
    if (uc < 0.0f || uc > 1.0f || vc < 0.0f || vc > 1.0f) return false;
    tout = t;
    return true;
}

// Box of every point quadTest can accept for the quad (origin, u, v) = s[0..8].  The test
// projects the hit onto u and onto v separately, so when u is not orthogonal to v it accepts
// the parallelogram whose corners p = origin + a u + b v solve
//   [u.u u.v; u.v v.v] [a b]^T = [i u.u, j v.v]^T,  i, j in {0, 1},
// not the one spanned by u and v (Quad::getBounds, quad.cpp:44-51, bounds only the latter, so
// the reference's BVH misses hits its brute-force loop finds on such quads).  The box covers
// both parallelograms; for rectangles they coincide.  A degenerate quad (u parallel to v)
// keeps the box of its own corners.
RBR_HD void quadBounds(const float* s, float lo[3], float hi[3]) {
    for (int k = 0; k < 3; ++k) {
        const float c1 = s[k] + s[3 + k], c2 = s[k] + s[6 + k], c3 = s[k] + s[3 + k] + s[6 + k];
        lo[k] = fminf(fminf(fminf(s[k], c1), c2), c3);
        hi[k] = fmaxf(fmaxf(fmaxf(s[k], c1), c2), c3);
    }
    const double uu = double(s[3]) * s[3] + double(s[4]) * s[4] + double(s[5]) * s[5];
    const double vv = double(s[6]) * s[6] + double(s[7]) * s[7] + double(s[8]) * s[8];
    const double uv = double(s[3]) * s[6] + double(s[4]) * s[7] + double(s[5]) * s[8];
    const double det = uu * vv - uv * uv;
    if (!(det > 1e-12 * uu * vv)) return;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            const double a = vv * (i * uu - j * uv) / det, b = uu * (j * vv - i * uv) / det;
            for (int k = 0; k < 3; ++k) {
                const float p = float(double(s[k]) + a * s[3 + k] + b * s[6 + k]);
                lo[k] = fminf(lo[k], p);
                hi[k] = fmaxf(hi[k], p);
            }
        }
}

RBR_HD bool planeTest(V o, V d, V n, float pd, float tmin, float tmax, float& tout) {
    float denom = vdot(n, d);
    if (fabsf(denom) < kEps) return false;
    float t = -(vdot(n, o) + pd) / denom;
    if (t < tmin || t > tmax) return false;
    tout = t;
    return true;
}

}  // namespace rbr
