// C entry points of the CPU oracle.  TEST INFRASTRUCTURE ONLY.
#include "oracle.h"

#include "orc_scene.h"

#include <omp.h>

using namespace orc;

struct orc_scene {
    Scene s;
};

static inline V3 v3(const float* p) { return V3(p[0], p[1], p[2]); }

static inline Ray toRay(const rbr_ray& r) {
    Ray x;
    x.o = v3(r.o);
    x.d = v3(r.d);
    x.tmin = r.tmin;
    x.tMax = r.tmax;
    return x;
}

extern "C" {

orc_scene* orc_scene_create(const float* tris, uint32_t ntri, const uint32_t* tri_ids,
                            const float* spheres, uint32_t nsph, const uint32_t* sph_ids,
                            const float* quads, uint32_t nquad, const uint32_t* quad_ids,
                            const float* planes, uint32_t nplane, const uint32_t* plane_ids) {
    auto* o = new orc_scene();
    uint32_t next = 0;
    o->s.tris.resize(ntri);
    for (uint32_t i = 0; i < ntri; ++i) {
        const float* t = tris + 9 * size_t(i);
        o->s.tris[i] = {v3(t), v3(t + 3), v3(t + 6), tri_ids ? tri_ids[i] : next + i};
    }
    next += ntri;
    o->s.sphs.resize(nsph);
    for (uint32_t i = 0; i < nsph; ++i) {
        const float* t = spheres + 4 * size_t(i);
        o->s.sphs[i] = {v3(t), t[3], sph_ids ? sph_ids[i] : next + i};
    }
    next += nsph;
    o->s.quads.resize(nquad);
    for (uint32_t i = 0; i < nquad; ++i) {
        const float* t = quads + 12 * size_t(i);
        o->s.quads[i] = {v3(t), v3(t + 3), v3(t + 6), v3(t + 9), quad_ids ? quad_ids[i] : next + i};
    }
    next += nquad;
    o->s.planes.resize(nplane);
    for (uint32_t i = 0; i < nplane; ++i) {
        const float* t = planes + 4 * size_t(i);
        o->s.planes[i] = {v3(t), t[3], plane_ids ? plane_ids[i] : next + i};
    }
    o->s.indexIds();
    return o;
}

void orc_scene_destroy(orc_scene* s) { delete s; }

void orc_scene_build_bvh(orc_scene* s) { s->s.buildBVH(); }

static int threads(int n) { return n > 0 ? n : omp_get_max_threads(); }

void orc_closest_brute(const orc_scene* s, const rbr_ray* rays, size_t n, rbr_hit* hits, int nt) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads(nt))
    for (long long i = 0; i < (long long)n; ++i) {
        float t;
        uint32_t id = s->s.closestBrute(toRay(rays[i]), t);
        hits[i].id = id;
        hits[i].t = id == RBR_MISS ? 0.0f : t;
    }
}

void orc_any_brute(const orc_scene* s, const rbr_ray* rays, size_t n, uint8_t* occ, int nt) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads(nt))
    for (long long i = 0; i < (long long)n; ++i) occ[i] = s->s.anyBrute(toRay(rays[i])) ? 1 : 0;
}

void orc_closest_bvh(const orc_scene* s, const rbr_ray* rays, size_t n, rbr_hit* hits, int nt) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads(nt))
    for (long long i = 0; i < (long long)n; ++i) {
        float t;
        uint32_t id = s->s.closestBVH(toRay(rays[i]), t);
        hits[i].id = id;
        hits[i].t = id == RBR_MISS ? 0.0f : t;
    }
}

void orc_any_bvh(const orc_scene* s, const rbr_ray* rays, size_t n, uint8_t* occ, int nt) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads(nt))
    for (long long i = 0; i < (long long)n; ++i) occ[i] = s->s.anyBVH(toRay(rays[i])) ? 1 : 0;
}

int orc_triangle_t(const float* tri9, const rbr_ray* ray, float* t) {
    Tri tr{v3(tri9), v3(tri9 + 3), v3(tri9 + 6), 0};
    return tri_t(tr, toRay(*ray), *t) ? 1 : 0;
}

int orc_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
