// C wrapper over the REFERENCE's scene, materials, lights, media, camera and integrators
// (compiled from its sources by oracle/Makefile into oracle/_ref/libref.so).  TEST
// INFRASTRUCTURE ONLY: it renders the same scene description the oracle and the GPU path
// take, through the reference's own Scene / PathTracingIntegrator / VolumetricPathIntegrator,
// so the oracle restatement can be compared against the reference itself.
#include "render/renderers/debug_integrators.h"
#include "render/camera/pinhole_camera.h"
#include "render/geometry/plane.h"
#include "render/geometry/quad.h"
#include "render/geometry/sphere.h"
#include "render/geometry/triangle.h"
#include "render/lighting/area_light.h"
#include "render/lighting/directional_light.h"
#include "render/lighting/point_light.h"
#include "render/renderers/path_tracing_integrator.h"
#include "render/renderers/volumetric_path_integrator.h"
#include "render/renderers/whitted_integrator.h"
#include "render/scene/scene.h"

#include <omp.h>

#include <memory>
#include <vector>

#include "../include/rbr.h"

using namespace render;

namespace {

inline Vec3 v3(const float* p) { return Vec3(p[0], p[1], p[2]); }

struct RefScene {
    Scene scene;
    std::vector<Material*> mats;
};

}  // namespace

extern "C" {

// Primitive ids follow the order tris, spheres, quads, planes (the default numbering of the
// C ABI); prim_material is indexed by that id.  triNormals (nullable): 9 floats per triangle,
// vertex normals of the triangles whose row is not all zero.
void* ref_scene_create(const float* tris, uint32_t ntri, const float* spheres, uint32_t nsph, const float* quads,
                       uint32_t nquad, const float* planes, uint32_t nplane, const rbr_material* mats, uint32_t nmat,
                       const int32_t* primMat, const rbr_light* lights, uint32_t nlights, const rbr_medium* medium,
                       const float* background, const float* triNormals) {
    auto* r = new RefScene();
    for (uint32_t i = 0; i < nmat; ++i) {
        auto m = std::make_unique<Material>();
        m->baseColor = v3(mats[i].base_color);
        m->metallic = mats[i].metallic;
        m->roughness = mats[i].roughness;
        m->ior = mats[i].ior;
        m->transmission = mats[i].transmission;
        m->emission = v3(mats[i].emission);
        r->mats.push_back(r->scene.addMaterial(std::move(m)));
    }
    uint32_t id = 0;
    auto matOf = [&](uint32_t pid) -> const Material* {
        int32_t m = primMat[pid];
        return m >= 0 ? r->mats[m] : nullptr;
    };
    for (uint32_t i = 0; i < ntri; ++i, ++id) {
        const float* t = tris + 9 * size_t(i);
        const float* n = triNormals ? triNormals + 9 * size_t(i) : nullptr;
        bool smooth = false;
        for (int k = 0; n && k < 9; ++k) smooth = smooth || n[k] != 0.0f;
        if (smooth)   // TriangleMesh::toPrimitives with normals (mesh.cpp:29-32)
            r->scene.addPrimitive(std::make_unique<Triangle>(v3(t), v3(t + 3), v3(t + 6), v3(n), v3(n + 3), v3(n + 6),
                                                             matOf(id)));
        else
            r->scene.addPrimitive(std::make_unique<Triangle>(v3(t), v3(t + 3), v3(t + 6), matOf(id)));
    }
    for (uint32_t i = 0; i < nsph; ++i, ++id) {
        const float* t = spheres + 4 * size_t(i);
        r->scene.addPrimitive(std::make_unique<Sphere>(v3(t), t[3], matOf(id)));
    }
    for (uint32_t i = 0; i < nquad; ++i, ++id) {
        const float* t = quads + 12 * size_t(i);
        auto q = std::make_unique<Quad>(v3(t), v3(t + 3), v3(t + 6), matOf(id));
        q->normal = v3(t + 9);
        r->scene.addPrimitive(std::move(q));
    }
    for (uint32_t i = 0; i < nplane; ++i, ++id) {
        const float* t = planes + 4 * size_t(i);
        auto p = std::make_unique<Plane>(v3(t), t[3], matOf(id));
        p->normal = v3(t);
        r->scene.addPrimitive(std::move(p));
    }
    for (uint32_t i = 0; i < nlights; ++i) {
        const rbr_light& l = lights[i];
        if (l.type == RBR_LIGHT_AREA) {
            auto a = std::make_unique<AreaLight>(v3(l.position), v3(l.u), v3(l.v), v3(l.color), l.intensity);
            a->normal = v3(l.normal);
            r->scene.addLight(std::move(a));
        } else if (l.type == RBR_LIGHT_POINT) {
            r->scene.addLight(std::make_unique<PointLight>(v3(l.position), v3(l.color), l.intensity));
        } else {
            auto d = std::make_unique<DirectionalLight>(v3(l.direction), v3(l.color), l.intensity);
            d->direction = v3(l.direction);
            r->scene.addLight(std::move(d));
        }
    }
    if (medium && medium->type == RBR_MEDIUM_HOMOGENEOUS)
        r->scene.setMedium(std::make_unique<HomogeneousMedium>(v3(medium->sigma_a), v3(medium->sigma_s), medium->g));
    else if (medium && medium->type == RBR_MEDIUM_EXPONENTIAL)
        r->scene.setMedium(std::make_unique<ExponentialMedium>(v3(medium->sigma_a), v3(medium->sigma_s), medium->falloff,
                                                               medium->base_height, medium->g));
    if (background) r->scene.backgroundColor = v3(background);
    r->scene.buildAccelerator();
    return r;
}

void ref_scene_destroy(void* h) { delete static_cast<RefScene*>(h); }

// Integrator::render of the reference on all OpenMP threads (or nthreads); returns seconds.
double ref_render(void* h, const rbr_camera* cam, const rbr_render_config* cfg, float* film_rgb, int nthreads) {
    auto* r = static_cast<RefScene*>(h);
    CameraExtrinsics ex;
    ex.position = v3(cam->position);
    ex.forward = v3(cam->forward);
    ex.right = v3(cam->right);
    ex.up = v3(cam->up);
    PinholeCamera camera(cam->width, cam->height, ex, PerspectiveIntrinsics(cam->fx, cam->fy, cam->cx, cam->cy));
    IntegratorConfig ic;
    ic.maxDepth = cfg->max_depth;
    ic.samplesPerPixel = cfg->spp;
    ic.useNEE = cfg->use_nee != 0;
    ic.useMIS = cfg->use_mis != 0;
    std::unique_ptr<Integrator> integ;
    if (cfg->integrator == RBR_INTEGRATOR_VOLPATH) integ = std::make_unique<VolumetricPathIntegrator>(ic, VolumeConfig{});
    else if (cfg->integrator == RBR_INTEGRATOR_NORMAL) integ = std::make_unique<NormalIntegrator>();
    else if (cfg->integrator == RBR_INTEGRATOR_DEPTH)
        integ = (cfg->depth_near != 0.0f || cfg->depth_far != 0.0f)
                    ? std::make_unique<DepthIntegrator>(cfg->depth_near, cfg->depth_far)
                    : std::make_unique<DepthIntegrator>();
    else if (cfg->integrator == RBR_INTEGRATOR_WHITTED) {
        integ = std::make_unique<WhittedIntegrator>(ic);
        r->scene.ambientLight = v3(cfg->ambient);
    } else if (cfg->integrator == RBR_INTEGRATOR_AO)
        integ = std::make_unique<AOIntegrator>(cfg->ao_samples ? cfg->ao_samples : 16,
                                               cfg->ao_distance > 0.0f ? cfg->ao_distance : 2.0f);
    else integ = std::make_unique<PathTracingIntegrator>(ic);
    integ->config = ic;   // render() reads samplesPerPixel from it
    Film film(cam->width, cam->height);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    double t0 = omp_get_wtime();
    integ->render(r->scene, camera, film);
    double sec = omp_get_wtime() - t0;
    for (int y = 0; y < cam->height; ++y)
        for (int x = 0; x < cam->width; ++x) {
            Color c = film.getPixel(x, y);
            float* o = film_rgb + 3 * (size_t(y) * cam->width + x);
            o[0] = c.x;
            o[1] = c.y;
            o[2] = c.z;
        }
    return sec;
}

}  // extern "C"
