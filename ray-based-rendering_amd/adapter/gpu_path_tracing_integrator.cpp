// GpuPathTracingIntegrator / GpuVolumetricPathIntegrator (gpu_path_tracing_integrator.h).
#include "gpu_path_tracing_integrator.h"

#include <cstring>
#include <mutex>
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

// The scene as the C ABI takes it.
struct Tables {
    std::vector<float> tris, sphs, quads, planes, normals;
    std::vector<uint32_t> triIds, sphIds, quadIds, planeIds;
    bool anyNormals = false;
    std::vector<rbr_material> mats;
    std::vector<int32_t> primMat;
    std::vector<rbr_light> lights;
    rbr_medium med{};
    float bg[3] = {0, 0, 0};

    bool sameGeometry(const Tables& o) const {
        return tris == o.tris && sphs == o.sphs && quads == o.quads && planes == o.planes && normals == o.normals &&
               triIds == o.triIds && sphIds == o.sphIds && quadIds == o.quadIds && planeIds == o.planeIds &&
               anyNormals == o.anyNormals;
    }
    bool sameShading(const Tables& o) const {
        auto eq = [](const void* a, const void* b, size_t n) { return n == 0 || std::memcmp(a, b, n) == 0; };
        return mats.size() == o.mats.size() && eq(mats.data(), o.mats.data(), mats.size() * sizeof(rbr_material)) &&
               primMat == o.primMat && lights.size() == o.lights.size() &&
               eq(lights.data(), o.lights.data(), lights.size() * sizeof(rbr_light)) &&
               std::memcmp(&med, &o.med, sizeof(med)) == 0 && std::memcmp(bg, o.bg, sizeof(bg)) == 0;
    }
};

// Fills t from the scene; false (with a reason) when the device path does not model it.
bool translate(const Scene& scene, Tables& t, std::string& reason) {
    std::unordered_map<const Material*, int32_t> matIndex;
    for (const auto& m : scene.materials()) {
        if (typeid(*m) != typeid(Material)) { reason = "a Material subclass overrides createBSDF"; return false; }
        rbr_material r;
        set3(r.base_color, m->baseColor);
        r.metallic = m->metallic;
        r.roughness = m->roughness;
        r.ior = m->ior;
        r.transmission = m->transmission;
        set3(r.emission, m->emission);
        matIndex[m.get()] = int32_t(t.mats.size());
        t.mats.push_back(r);
    }
    // Primitive id = index in Scene::primitives, the order buildAccelerator passes them.
    const auto& prims = scene.primitives();
    for (size_t i = 0; i < prims.size(); ++i) {
        const Primitive* p = prims[i].get();
        const uint32_t id = uint32_t(i);
        if (auto* tr = dynamic_cast<const Triangle*>(p)) {
            put3(t.tris, tr->v0); put3(t.tris, tr->v1); put3(t.tris, tr->v2);
            t.triIds.push_back(id);
            // Vertex normals as the triangle holds them (unit length already); zeros keep the
            // geometric normal for a triangle without them.
            if (tr->hasNormals) { put3(t.normals, tr->n0); put3(t.normals, tr->n1); put3(t.normals, tr->n2); }
            else t.normals.insert(t.normals.end(), 9, 0.0f);
            t.anyNormals = t.anyNormals || tr->hasNormals;
        } else if (auto* s = dynamic_cast<const Sphere*>(p)) {
            put3(t.sphs, s->center);
            t.sphs.push_back(s->radius);
            t.sphIds.push_back(id);
        } else if (auto* q = dynamic_cast<const Quad*>(p)) {
            put3(t.quads, q->origin); put3(t.quads, q->u); put3(t.quads, q->v); put3(t.quads, q->normal);
            t.quadIds.push_back(id);
        } else if (auto* pl = dynamic_cast<const Plane*>(p)) {
            put3(t.planes, pl->normal);
            t.planes.push_back(pl->d);
            t.planeIds.push_back(id);
        } else {
            reason = "unsupported primitive type";
            return false;
        }
        auto it = p->material ? matIndex.find(p->material) : matIndex.end();
        if (p->material && it == matIndex.end()) { reason = "primitive material not owned by the scene"; return false; }
        t.primMat.push_back(p->material ? it->second : -1);
    }
    if (!t.anyNormals) t.normals.clear();
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
        t.lights.push_back(r);
    }
    if (const Medium* m = scene.medium()) {
        auto* hg = dynamic_cast<const HenyeyGreensteinPhase*>(m->getPhaseFunction());
        if (!hg) { reason = "medium phase function is not Henyey-Greenstein"; return false; }
        t.med.g = hg->g;
        if (auto* h = dynamic_cast<const HomogeneousMedium*>(m)) {
            t.med.type = RBR_MEDIUM_HOMOGENEOUS;
            set3(t.med.sigma_a, h->sigma_a); set3(t.med.sigma_s, h->sigma_s);
        } else if (auto* e = dynamic_cast<const ExponentialMedium*>(m)) {
            t.med.type = RBR_MEDIUM_EXPONENTIAL;
            set3(t.med.sigma_a, e->sigma_a); set3(t.med.sigma_s, e->sigma_s);
            t.med.falloff = e->falloff;
            t.med.base_height = e->baseHeight;
        } else {
            reason = "unsupported medium";
            return false;
        }
    }
    set3(t.bg, scene.backgroundColor);
    return true;
}

void progressTrampoline(float fraction, void* user) { (*static_cast<std::function<void(float)>*>(user))(fraction); }

}  // namespace

// One context and scene per integrator object, reused while the scene is unchanged.  render()
// is const in the reference's interface; the mutex serializes calls on one object (the C ABI
// allows one call at a time per context).
class GpuRenderCache {
public:
    GpuRenderCache(int device, uint64_t seed) : device_(device), seed_(seed) {}
    ~GpuRenderCache() {
        if (scene_) rbr_scene_destroy(scene_);
        if (ctx_) rbr_ctx_destroy(ctx_);
    }

    bool render(const Scene& scene, const Camera& camera, Film& film, const IntegratorConfig& cfg, int integrator,
                std::function<void(float)>& progress);

    std::mutex mu;
    GpuTuning tuning;
    bool tuningDirty = true;
    bool onGpu = false;
    std::string reason;
    int builds = 0;

private:
    bool fail(const char* what) {
        reason = std::string(what) + ": " + rbr_last_error(ctx_);
        return false;
    }
    int device_;
    uint64_t seed_;
    rbr_ctx* ctx_ = nullptr;
    rbr_scene* scene_ = nullptr;
    Tables last_;
    bool haveGeometry_ = false, haveShading_ = false;
};

bool GpuRenderCache::render(const Scene& scene, const Camera& camera, Film& film, const IntegratorConfig& cfg,
                            int integrator, std::function<void(float)>& progress) {
    auto* pin = dynamic_cast<const PinholeCamera*>(&camera);
    if (!pin) { reason = "camera is not a PinholeCamera"; return false; }
    Tables t;
    if (!translate(scene, t, reason)) return false;
    if (!ctx_) {
        if (device_ < 0 || rbr_ctx_create(device_, &ctx_) != RBR_OK) { ctx_ = nullptr; reason = "no GPU"; return false; }
        if (rbr_scene_create(ctx_, &scene_) != RBR_OK) return fail("rbr_scene_create");
    }
    if (tuningDirty) {
        if (rbr_set_trace_tuning(ctx_, tuning.traceMode, tuning.ldsDepth, tuning.wavesPerCu) != RBR_OK)
            return fail("rbr_set_trace_tuning");
        if (rbr_set_render_tuning(ctx_, tuning.poolPaths, tuning.useGraphs, tuning.sortMaterials) != RBR_OK)
            return fail("rbr_set_render_tuning");
        tuningDirty = false;
        haveGeometry_ = false;   // the BVH layout may have changed
    }
    if (!haveGeometry_ || !t.sameGeometry(last_)) {
        haveGeometry_ = haveShading_ = false;
        if (rbr_scene_set_geometry(scene_, t.tris.data(), uint32_t(t.triIds.size()), t.triIds.data(), t.sphs.data(),
                                   uint32_t(t.sphIds.size()), t.sphIds.data(), t.quads.data(),
                                   uint32_t(t.quadIds.size()), t.quadIds.data(), t.planes.data(),
                                   uint32_t(t.planeIds.size()), t.planeIds.data()) != RBR_OK)
            return fail("rbr_scene_set_geometry");
        if (t.anyNormals && rbr_scene_set_tri_normals(scene_, t.normals.data(), t.triIds.size()) != RBR_OK)
            return fail("rbr_scene_set_tri_normals");
        if (rbr_scene_build_bvh(scene_, tuning.bvhWidth, tuning.leafMax) != RBR_OK) return fail("rbr_scene_build_bvh");
        ++builds;
        haveGeometry_ = true;
    }
    if (!haveShading_ || !t.sameShading(last_)) {
        haveShading_ = false;
        if (rbr_scene_set_shading(scene_, t.mats.data(), uint32_t(t.mats.size()), t.primMat.data(),
                                  uint32_t(t.primMat.size()), t.lights.data(), uint32_t(t.lights.size()), &t.med,
                                  t.bg) != RBR_OK)
            return fail("rbr_scene_set_shading");
        haveShading_ = true;
    }
    last_ = std::move(t);
    rbr_job job{};
    job.scene = scene_;
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
    job.config.seed = seed_;
    std::vector<float> rgb(size_t(film.width) * film.height * 3);
    job.film_rgb = rgb.data();
    const int rc = progress ? rbr_render_progress(ctx_, &job, 1, progressTrampoline, &progress)
                            : rbr_render(ctx_, &job, 1);
    if (rc != RBR_OK) return fail("rbr_render");
    for (int y = 0; y < film.height; ++y)
        for (int x = 0; x < film.width; ++x) {
            const float* c = &rgb[3 * (size_t(y) * film.width + x)];
            film.addSample(x, y, Color(c[0], c[1], c[2]));
        }
    return true;
}

// ---- GpuPathTracingIntegrator

GpuPathTracingIntegrator::GpuPathTracingIntegrator(const IntegratorConfig& cfg, int device, uint64_t seed)
    : PathTracingIntegrator(cfg), cache_(std::make_unique<GpuRenderCache>(device, seed)) {}
GpuPathTracingIntegrator::~GpuPathTracingIntegrator() = default;

void GpuPathTracingIntegrator::render(const Scene& scene, const Camera& camera, Film& film,
                                      std::function<void(float)> progress) const {
    {
        std::lock_guard<std::mutex> lock(cache_->mu);
        cache_->reason.clear();
        cache_->onGpu = cache_->render(scene, camera, film, config, RBR_INTEGRATOR_PATH, progress);
        if (cache_->onGpu) return;
    }
    PathTracingIntegrator::render(scene, camera, film, progress);
}

bool GpuPathTracingIntegrator::lastRenderOnGpu() const { return cache_->onGpu; }
const std::string& GpuPathTracingIntegrator::lastFallbackReason() const { return cache_->reason; }
int GpuPathTracingIntegrator::bvhBuilds() const { return cache_->builds; }
void GpuPathTracingIntegrator::setTuning(const GpuTuning& t) {
    std::lock_guard<std::mutex> lock(cache_->mu);
    cache_->tuning = t;
    cache_->tuningDirty = true;
}

// ---- GpuVolumetricPathIntegrator

GpuVolumetricPathIntegrator::GpuVolumetricPathIntegrator(const IntegratorConfig& cfg, int device, uint64_t seed)
    : VolumetricPathIntegrator(cfg, VolumeConfig{}), cache_(std::make_unique<GpuRenderCache>(device, seed)) {}
GpuVolumetricPathIntegrator::~GpuVolumetricPathIntegrator() = default;

void GpuVolumetricPathIntegrator::render(const Scene& scene, const Camera& camera, Film& film,
                                         std::function<void(float)> progress) const {
    {
        std::lock_guard<std::mutex> lock(cache_->mu);
        cache_->reason.clear();
        cache_->onGpu = cache_->render(scene, camera, film, config, RBR_INTEGRATOR_VOLPATH, progress);
        if (cache_->onGpu) return;
    }
    VolumetricPathIntegrator::render(scene, camera, film, progress);
}

bool GpuVolumetricPathIntegrator::lastRenderOnGpu() const { return cache_->onGpu; }
const std::string& GpuVolumetricPathIntegrator::lastFallbackReason() const { return cache_->reason; }
int GpuVolumetricPathIntegrator::bvhBuilds() const { return cache_->builds; }
void GpuVolumetricPathIntegrator::setTuning(const GpuTuning& t) {
    std::lock_guard<std::mutex> lock(cache_->mu);
    cache_->tuning = t;
    cache_->tuningDirty = true;
}

}  // namespace render
