// GpuPathTracingIntegrator / GpuVolumetricPathIntegrator (gpu_path_tracing_integrator.h).
#include "gpu_path_tracing_integrator.h"

#include <typeinfo>
#include <unordered_map>
#include <vector>

#include "rbr.h"
#include "render/camera/pinhole_camera.h"
#include "render/geometry/plane.h"
#include "render/geometry/quad.h"
#include "render/geometry/sphere.h"
#include "render/geometry/triangle.h"
#include "render/lighting/area_light.h"
#include "render/lighting/directional_light.h"
#include "render/lighting/point_light.h"
#include "render/materials/medium.h"
#include "render/materials/phase_function.h"

namespace render {

namespace {

inline void put3(std::vector<float>& v, const Vec3& a) {
    v.push_back(a.x);
    v.push_back(a.y);
    v.push_back(a.z);
}
inline void set3(float* d, const Vec3& a) {
    d[0] = a.x;
    d[1] = a.y;
    d[2] = a.z;
}

struct Handles {
    rbr_ctx* ctx = nullptr;
    rbr_scene* scene = nullptr;
    ~Handles() {
        if (scene) rbr_scene_destroy(scene);
        if (ctx) rbr_ctx_destroy(ctx);
    }
};

}  // namespace

bool gpuRender(const Scene& scene, const Camera& camera, Film& film, const IntegratorConfig& cfg, int integrator,
               int device, uint64_t seed, std::string& reason) {
    // ---- camera
    auto* pin = dynamic_cast<const PinholeCamera*>(&camera);
    if (!pin) { reason = "camera is not a PinholeCamera"; return false; }
    // ---- materials (the base six-parameter model only)
    std::unordered_map<const Material*, int32_t> matIndex;
    std::vector<rbr_material> mats;
    for (const auto& m : scene.materials()) {
        if (typeid(*m) != typeid(Material)) { reason = "a Material subclass overrides createBSDF"; return false; }
        rbr_material r;
        set3(r.base_color, m->baseColor);
        r.metallic = m->metallic;
        r.roughness = m->roughness;
        r.ior = m->ior;
        r.transmission = m->transmission;
        set3(r.emission, m->emission);
        matIndex[m.get()] = int32_t(mats.size());
        mats.push_back(r);
    }
    // ---- primitives (id = index in Scene::primitives, the order buildAccelerator passes)
    std::vector<float> tris, sphs, quads, planes;
    std::vector<uint32_t> triIds, sphIds, quadIds, planeIds;
    std::vector<int32_t> primMat;
    const auto& prims = scene.primitives();
    for (size_t i = 0; i < prims.size(); ++i) {
        const Primitive* p = prims[i].get();
        const uint32_t id = uint32_t(i);
        if (auto* t = dynamic_cast<const Triangle*>(p)) {
            if (t->hasNormals) { reason = "triangle with vertex normals"; return false; }
            put3(tris, t->v0); put3(tris, t->v1); put3(tris, t->v2);
            triIds.push_back(id);
        } else if (auto* s = dynamic_cast<const Sphere*>(p)) {
            put3(sphs, s->center);
            sphs.push_back(s->radius);
            sphIds.push_back(id);
        } else if (auto* q = dynamic_cast<const Quad*>(p)) {
            put3(quads, q->origin); put3(quads, q->u); put3(quads, q->v); put3(quads, q->normal);
            quadIds.push_back(id);
        } else if (auto* pl = dynamic_cast<const Plane*>(p)) {
            put3(planes, pl->normal);
            planes.push_back(pl->d);
            planeIds.push_back(id);
        } else {
            reason = "unsupported primitive type";
            return false;
        }
        auto it = p->material ? matIndex.find(p->material) : matIndex.end();
        if (p->material && it == matIndex.end()) { reason = "primitive material not owned by the scene"; return false; }
        primMat.push_back(p->material ? it->second : -1);
    }
    // ---- lights
    std::vector<rbr_light> lights;
    for (const auto& l : scene.lights()) {
        rbr_light r{};
        set3(r.color, l->color);
        r.intensity = l->intensity;
        if (auto* a = dynamic_cast<const AreaLight*>(l.get())) {
            r.type = RBR_LIGHT_AREA;
            set3(r.position, a->position); set3(r.u, a->u); set3(r.v, a->v); set3(r.normal, a->normal);
        } else if (auto* pt = dynamic_cast<const PointLight*>(l.get())) {
            r.type = RBR_LIGHT_POINT;
            set3(r.position, pt->position);
        } else if (auto* d = dynamic_cast<const DirectionalLight*>(l.get())) {
            r.type = RBR_LIGHT_DIRECTIONAL;
            set3(r.direction, d->direction);
        } else {
            reason = "unsupported light type";
            return false;
        }
        lights.push_back(r);
    }
    // ---- medium
    rbr_medium med{};
    if (const Medium* m = scene.medium()) {
        auto* hg = dynamic_cast<const HenyeyGreensteinPhase*>(m->getPhaseFunction());
        if (!hg) { reason = "medium phase function is not Henyey-Greenstein"; return false; }
        med.g = hg->g;
        if (auto* h = dynamic_cast<const HomogeneousMedium*>(m)) {
            med.type = RBR_MEDIUM_HOMOGENEOUS;
            set3(med.sigma_a, h->sigma_a); set3(med.sigma_s, h->sigma_s);
        } else if (auto* e = dynamic_cast<const ExponentialMedium*>(m)) {
            med.type = RBR_MEDIUM_EXPONENTIAL;
            set3(med.sigma_a, e->sigma_a); set3(med.sigma_s, e->sigma_s);
            med.falloff = e->falloff;
            med.base_height = e->baseHeight;
        } else {
            reason = "unsupported medium";
            return false;
        }
    }
    // ---- device
    Handles h;
    if (device < 0 || rbr_ctx_create(device, &h.ctx) != RBR_OK) { reason = "no GPU"; h.ctx = nullptr; return false; }
    auto check = [&](int rc) {
        if (rc != RBR_OK) reason = rbr_last_error(h.ctx);
        return rc == RBR_OK;
    };
    if (!check(rbr_scene_create(h.ctx, &h.scene))) return false;
    if (!check(rbr_scene_set_geometry(h.scene, tris.data(), uint32_t(triIds.size()), triIds.data(), sphs.data(),
                                      uint32_t(sphIds.size()), sphIds.data(), quads.data(), uint32_t(quadIds.size()),
                                      quadIds.data(), planes.data(), uint32_t(planeIds.size()), planeIds.data())))
        return false;
    if (!check(rbr_scene_build_bvh(h.scene, 4, 4))) return false;
    const float bg[3] = {scene.backgroundColor.x, scene.backgroundColor.y, scene.backgroundColor.z};
    if (!check(rbr_scene_set_shading(h.scene, mats.data(), uint32_t(mats.size()), primMat.data(),
                                     uint32_t(primMat.size()), lights.data(), uint32_t(lights.size()), &med, bg)))
        return false;
    rbr_job job{};
    job.scene = h.scene;
    const PerspectiveIntrinsics& K = pin->perspectiveIntrinsics();
    set3(job.camera.position, camera.extrinsics.position);
    set3(job.camera.forward, camera.extrinsics.forward);
    set3(job.camera.right, camera.extrinsics.right);
    set3(job.camera.up, camera.extrinsics.up);
    job.camera.fx = K.fx;
    job.camera.fy = K.fy;
    job.camera.cx = K.cx;
    job.camera.cy = K.cy;
    job.camera.width = film.width;
    job.camera.height = film.height;
    job.config.integrator = integrator;
    job.config.spp = cfg.samplesPerPixel;
    job.config.max_depth = cfg.maxDepth;
    job.config.use_nee = cfg.useNEE ? 1 : 0;
    job.config.use_mis = cfg.useMIS ? 1 : 0;
    job.config.seed = seed;
    std::vector<float> rgb(size_t(film.width) * film.height * 3);
    job.film_rgb = rgb.data();
    if (!check(rbr_render(h.ctx, &job, 1))) return false;
    for (int y = 0; y < film.height; ++y)
        for (int x = 0; x < film.width; ++x) {
            const float* c = &rgb[3 * (size_t(y) * film.width + x)];
            film.addSample(x, y, Color(c[0], c[1], c[2]));
        }
    return true;
}

void GpuPathTracingIntegrator::render(const Scene& scene, const Camera& camera, Film& film,
                                      std::function<void(float)> progress) const {
    reason_.clear();
    onGpu_ = gpuRender(scene, camera, film, config, RBR_INTEGRATOR_PATH, device_, seed_, reason_);
    if (!onGpu_) {
        PathTracingIntegrator::render(scene, camera, film, progress);
        return;
    }
    if (progress) progress(1.0f);
}

void GpuVolumetricPathIntegrator::render(const Scene& scene, const Camera& camera, Film& film,
                                         std::function<void(float)> progress) const {
    reason_.clear();
    onGpu_ = gpuRender(scene, camera, film, config, RBR_INTEGRATOR_VOLPATH, device_, seed_, reason_);
    if (!onGpu_) {
        VolumetricPathIntegrator::render(scene, camera, film, progress);
        return;
    }
    if (progress) progress(1.0f);
}

}  // namespace render
