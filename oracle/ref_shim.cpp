// C wrapper around the REFERENCE's own classes, compiled from its sources where they lie
// (/root/reference/src, headers in /root/reference/include) into oracle/_ref/libref.so by
// oracle/Makefile.  TEST INFRASTRUCTURE ONLY: used to pin the oracle restatement and as the
// CPU baseline; the product never links it.  No reference source is copied here — this file
// only instantiates the reference's types through their public headers.
#include "render/acceleration/bvh.h"
#include "render/core/film.h"
#include "render/io/obj_importer.h"
#include "render/geometry/plane.h"
#include "render/geometry/quad.h"
#include "render/geometry/sphere.h"
#include "render/geometry/triangle.h"

#include <omp.h>

#include <cstdint>
#include <memory>
#include <unordered_map>
#include <vector>

#include "../include/rbr.h"

using namespace render;

namespace {

struct RefWorld {
    std::vector<std::unique_ptr<Primitive>> owned;
    std::vector<Primitive*> prims;
    std::unordered_map<const Primitive*, uint32_t> ids;
    BVH bvh;

    void add(std::unique_ptr<Primitive> p, uint32_t id) {
        ids[p.get()] = id;
        prims.push_back(p.get());
        owned.push_back(std::move(p));
    }
    uint32_t idOf(const Primitive* p) const {
        auto it = ids.find(p);
        return it == ids.end() ? RBR_MISS : it->second;
    }
};

inline Vec3 v3(const float* p) { return Vec3(p[0], p[1], p[2]); }

// A reference Ray whose direction is taken as given (the constructor would re-normalize).
inline Ray refRay(const rbr_ray& r) {
    Ray x;
    x.origin = v3(r.o);
    x.direction = v3(r.d);
    x.tMax = r.tmax;
    return x;
}

}  // namespace

extern "C" {

void* ref_world_create(const float* tris, uint32_t ntri, const float* spheres, uint32_t nsph,
                       const float* quads, uint32_t nquad, const float* planes, uint32_t nplane) {
    auto* w = new RefWorld();
    uint32_t id = 0;
    for (uint32_t i = 0; i < ntri; ++i, ++id) {
        const float* t = tris + 9 * size_t(i);
        w->add(std::make_unique<Triangle>(v3(t), v3(t + 3), v3(t + 6)), id);
    }
    for (uint32_t i = 0; i < nsph; ++i, ++id) {
        const float* t = spheres + 4 * size_t(i);
        w->add(std::make_unique<Sphere>(v3(t), t[3]), id);
    }
    for (uint32_t i = 0; i < nquad; ++i, ++id) {
        const float* t = quads + 12 * size_t(i);
        auto q = std::make_unique<Quad>(v3(t), v3(t + 3), v3(t + 6));
        q->normal = v3(t + 9);
        w->add(std::move(q), id);
    }
    for (uint32_t i = 0; i < nplane; ++i, ++id) {
        const float* t = planes + 4 * size_t(i);
        auto p = std::make_unique<Plane>(v3(t), t[3]);
        p->normal = v3(t);
        w->add(std::move(p), id);
    }
    return w;
}

void ref_world_destroy(void* h) { delete static_cast<RefWorld*>(h); }

// BVH::build (bvh.cpp:6-31); returns build milliseconds.
double ref_bvh_build(void* h) {
    auto* w = static_cast<RefWorld*>(h);
    double t0 = omp_get_wtime();
    w->bvh.build(w->prims);
    return (omp_get_wtime() - t0) * 1e3;
}

// BVH::intersect over a batch, OpenMP over rays (the reference's own loop is serial,
// accelerator_compare.cpp:115-119).  Returns seconds.
double ref_bvh_closest(void* h, const rbr_ray* rays, size_t n, rbr_hit* hits, int nthreads) {
    auto* w = static_cast<RefWorld*>(h);
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
    for (long long i = 0; i < (long long)n; ++i) {
        Ray r = refRay(rays[i]);
        SurfaceInteraction isect;
        if (w->bvh.intersect(r, isect)) {
            hits[i].id = w->idOf(isect.primitive);
            hits[i].t = isect.t;
        } else {
            hits[i].id = RBR_MISS;
            hits[i].t = 0.0f;
        }
    }
    return omp_get_wtime() - t0;
}

double ref_bvh_any(void* h, const rbr_ray* rays, size_t n, uint8_t* occ, int nthreads) {
    auto* w = static_cast<RefWorld*>(h);
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
    for (long long i = 0; i < (long long)n; ++i) occ[i] = w->bvh.intersectP(refRay(rays[i])) ? 1 : 0;
    return omp_get_wtime() - t0;
}

// The brute-force loop of test_accelerators.cpp:58-74 (the last equal-t primitive wins).
void ref_brute_closest(void* h, const rbr_ray* rays, size_t n, rbr_hit* hits, int nthreads) {
    auto* w = static_cast<RefWorld*>(h);
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (long long i = 0; i < (long long)n; ++i) {
        Ray r = refRay(rays[i]);
        SurfaceInteraction isect;
        bool hit = false;
        for (const Primitive* p : w->prims)
            if (p->intersect(r, isect)) hit = true;
        hits[i].id = hit ? w->idOf(isect.primitive) : RBR_MISS;
        hits[i].t = hit ? isect.t : 0.0f;
    }
}

void ref_brute_any(void* h, const rbr_ray* rays, size_t n, uint8_t* occ, int nthreads) {
    auto* w = static_cast<RefWorld*>(h);
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (long long i = 0; i < (long long)n; ++i) {
        Ray r = refRay(rays[i]);
        bool any = false;
        for (const Primitive* p : w->prims)
            if (p->intersectP(r)) { any = true; break; }
        occ[i] = any ? 1 : 0;
    }
}

int ref_max_threads(void) { return omp_get_max_threads(); }

// The reference's Film::toRGB8 / toRGB8ToneMapped on a film whose pixels are set by one
// addSample of weight 1 (getColor returns the value exactly).
void ref_film_to_rgb8(const float* rgb, int w, int h, float gamma, int tone_map, float exposure, unsigned char* out) {
    Film film(w, h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) film.addSample(x, y, v3(rgb + (size_t(y) * w + x) * 3), 1.0f);
    const std::vector<unsigned char> v = tone_map ? film.toRGB8ToneMapped(exposure, gamma) : film.toRGB8(gamma);
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
}

// ObjImporter::parse of `text`; the triangles as 9 floats each (positions[indices]).  Returns
// the triangle count, or -1 when the importer fails.
long ref_obj_parse(const char* text, size_t length, float* tris, size_t capacity) {
    TriangleMesh mesh;
    if (!ObjImporter::parse(std::string(text, length), mesh)) return -1;
    const size_t n = mesh.triangleCount();
    for (size_t t = 0; t < n && t < capacity; ++t)
        for (int c = 0; c < 3; ++c) {
            const Vec3& p = mesh.positions[size_t(mesh.indices[3 * t + c])];
            tris[9 * t + 3 * c] = p.x;
            tris[9 * t + 3 * c + 1] = p.y;
            tris[9 * t + 3 * c + 2] = p.z;
        }
    return long(n);
}

// The vertex normals ObjImporter::parse keeps (mesh.normals at each triangle corner, 9 floats per
// triangle).  Returns the triangle count when the mesh keeps normals, 0 when it does not, -1
// when the importer fails.
long ref_obj_parse_normals(const char* text, size_t length, float* normals, size_t capacity) {
    TriangleMesh mesh;
    if (!ObjImporter::parse(std::string(text, length), mesh)) return -1;
    if (mesh.normals.size() != mesh.positions.size()) return 0;
    const size_t n = mesh.triangleCount();
    for (size_t t = 0; t < n && t < capacity; ++t)
        for (int c = 0; c < 3; ++c) {
            const Vec3& v = mesh.normals[size_t(mesh.indices[3 * t + c])];
            normals[9 * t + 3 * c] = v.x;
            normals[9 * t + 3 * c + 1] = v.y;
            normals[9 * t + 3 * c + 2] = v.z;
        }
    return long(n);
}

// Triangle::intersect for a triangle with vertex normals (v, n: 9 floats each) and the ray
// Ray(o, d) (which normalizes d; dn receives the direction used).  On a hit writes the
// interaction's normal (faced towards the ray by setFaceNormal) and frontFace; returns 1.
int ref_triangle_normal(const float* v, const float* n, const float* o, const float* d, float* dn, float* normal,
                        int* front) {
    Triangle tri(v3(v), v3(v + 3), v3(v + 6), v3(n), v3(n + 3), v3(n + 6));
    Ray ray(v3(o), v3(d));
    dn[0] = ray.direction.x;
    dn[1] = ray.direction.y;
    dn[2] = ray.direction.z;
    SurfaceInteraction isect;
    if (!tri.intersect(ray, isect)) return 0;
    normal[0] = isect.normal.x;
    normal[1] = isect.normal.y;
    normal[2] = isect.normal.z;
    *front = isect.frontFace ? 1 : 0;
    return 1;
}

}  // extern "C"
