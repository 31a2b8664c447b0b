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

// ---------------------------------------------------------------- rendering

#include "orc_render.h"

namespace {

struct RS {
    RenderScene s;
};

V3 f3(const float* p) { return V3(p[0], p[1], p[2]); }

orc::Camera toCam(const rbr_camera& c) {
    orc::Camera k;
    k.position = f3(c.position);
    k.forward = f3(c.forward);
    k.right = f3(c.right);
    k.up = f3(c.up);
    k.fx = c.fx;
    k.fy = c.fy;
    k.cx = c.cx;
    k.cy = c.cy;
    k.width = c.width;
    k.height = c.height;
    return k;
}

orc::Config toCfg(const rbr_render_config& c) {
    orc::Config k;
    k.integrator = c.integrator;
    k.spp = c.spp;
    k.maxDepth = c.max_depth;
    k.useNEE = c.use_nee != 0;
    k.useMIS = c.use_mis != 0;
    return k;
}

inline Color Li(const RenderScene& s, const orc::Config& c, const Ray& r, Sampler& smp) {
    return c.integrator == RBR_INTEGRATOR_VOLPATH ? LiVolPath(s, c, r, smp) : LiPath(s, c, r, smp);
}

}  // namespace

extern "C" {

void* orc_rscene_create(const float* tris, uint32_t ntri, const uint32_t* tri_ids, const float* spheres, uint32_t nsph,
                        const uint32_t* sph_ids, const float* quads, uint32_t nquad, const uint32_t* quad_ids,
                        const float* planes, uint32_t nplane, const uint32_t* plane_ids, const rbr_material* mats,
                        uint32_t nmat, const int32_t* primMat, uint32_t nPrimIds, const rbr_light* lights,
                        uint32_t nlights, const rbr_medium* medium, const float* background, const float* tri_normals) {
    orc_scene* g = orc_scene_create(tris, ntri, tri_ids, spheres, nsph, sph_ids, quads, nquad, quad_ids, planes, nplane,
                                    plane_ids);
    // vertex normals (9 per triangle; an all-zero row keeps the triangle flat), normalized as
    // Triangle's constructor does (triangle.h:29-33)
    for (uint32_t i = 0; tri_normals && i < ntri; ++i) {
        const float* n = tri_normals + 9 * size_t(i);
        bool any = false;
        for (int k = 0; k < 9; ++k) any = any || n[k] != 0.0f;
        if (!any) continue;
        orc::Tri& tr = g->s.tris[i];
        tr.smooth = true;
        tr.n0 = f3(n).normalized();
        tr.n1 = f3(n + 3).normalized();
        tr.n2 = f3(n + 6).normalized();
    }
    auto* r = new RS();
    r->s.geo = std::move(g->s);
    delete g;
    r->s.geo.buildBVH();
    for (uint32_t i = 0; i < nmat; ++i) {
        orc::Material m;
        m.baseColor = f3(mats[i].base_color);
        m.metallic = mats[i].metallic;
        m.roughness = mats[i].roughness;
        m.ior = mats[i].ior;
        m.transmission = mats[i].transmission;
        m.emission = f3(mats[i].emission);
        r->s.materials.push_back(m);
    }
    r->s.primMaterial.assign(primMat, primMat + nPrimIds);
    for (uint32_t i = 0; i < nlights; ++i) {
        const rbr_light& l = lights[i];
        orc::Light x;
        x.type = l.type;
        x.color = f3(l.color);
        x.intensity = l.intensity;
        x.position = f3(l.position);
        x.u = f3(l.u);
        x.v = f3(l.v);
        x.normal = f3(l.normal);
        x.direction = f3(l.direction);
        r->s.lights.push_back(x);
    }
    if (medium) {
        r->s.medium.type = medium->type;
        r->s.medium.sigma_a = f3(medium->sigma_a);
        r->s.medium.sigma_s = f3(medium->sigma_s);
        r->s.medium.falloff = medium->falloff;
        r->s.medium.baseHeight = medium->base_height;
        r->s.medium.g = medium->g;
    }
    if (background) r->s.background = f3(background);
    return r;
}

void orc_rscene_destroy(void* h) { delete static_cast<RS*>(h); }

// Integrator::render (integrator.cpp:7-50): rows in parallel, sampler reseeded with y*W per
// row, per sample jitter (get2D), two lens numbers, camera ray, Li; film = sum * (1/spp).
void orc_render(void* h, const rbr_camera* cam, const rbr_render_config* cfg, float* film, int nthreads) {
    const RenderScene& s = static_cast<RS*>(h)->s;
    const orc::Camera c = toCam(*cam);
    const orc::Config k = toCfg(*cfg);
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
    {
        Sampler smp;
#pragma omp for schedule(dynamic, 16)
        for (int y = 0; y < c.height; ++y) {
            smp.setSeed(uint64_t(y) * uint64_t(c.width));
            for (int x = 0; x < c.width; ++x) {
                Color sum(0.0f);
                for (int sIdx = 0; sIdx < k.spp; ++sIdx) {
                    float jx = smp.get1D();
                    float jy = smp.get1D();
                    float px = x + jx, py = y + jy;
                    (void)smp.get1D();   // lensU
                    (void)smp.get1D();   // lensV
                    sum += Li(s, k, c.generateRay(px, py), smp);
                }
                Color v = sum / float(k.spp);
                float* o = film + 3 * (size_t(y) * c.width + x);
                o[0] = v.x;
                o[1] = v.y;
                o[2] = v.z;
            }
        }
    }
}

// Per-pixel first and second moments of the single-sample radiance (double accumulation),
// each pixel with its own stream seeded from (seed, pixel index).  `pixels` (optional)
// restricts the work to a list of pixel indices; sums are written at those indices.
void orc_render_moments(void* h, const rbr_camera* cam, const rbr_render_config* cfg, const uint32_t* pixels,
                        size_t npix, double* sum, double* sumsq, int nthreads) {
    const RenderScene& s = static_cast<RS*>(h)->s;
    const orc::Camera c = toCam(*cam);
    const orc::Config k = toCfg(*cfg);
    const size_t total = pixels ? npix : size_t(c.width) * c.height;
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
    {
        Sampler smp;
#pragma omp for schedule(dynamic, 8)
        for (long long j = 0; j < (long long)total; ++j) {
            const uint32_t p = pixels ? pixels[j] : uint32_t(j);
            const int x = int(p % uint32_t(c.width)), y = int(p / uint32_t(c.width));
            smp.setSeed(cfg->seed * 0x9E3779B97F4A7C15ull + p);
            double s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
            for (int sIdx = 0; sIdx < k.spp; ++sIdx) {
                float jx = smp.get1D();
                float jy = smp.get1D();
                (void)smp.get1D();
                (void)smp.get1D();
                Color v = Li(s, k, c.generateRay(x + jx, y + jy), smp);
                s1[0] += v.x; s1[1] += v.y; s1[2] += v.z;
                s2[0] += double(v.x) * v.x; s2[1] += double(v.y) * v.y; s2[2] += double(v.z) * v.z;
            }
            for (int ch = 0; ch < 3; ++ch) {
                sum[3 * size_t(p) + ch] = s1[ch];
                sumsq[3 * size_t(p) + ch] = s2[ch];
            }
        }
    }
}

}  // extern "C"
