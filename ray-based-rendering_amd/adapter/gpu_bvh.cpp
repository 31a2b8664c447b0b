// GpuBVH (gpu_bvh.h): Accelerator over the C ABI of the MI355X library.
#include "gpu_bvh.h"

#include <algorithm>

#include "rbr.h"
#include "render/geometry/plane.h"
#include "render/geometry/quad.h"
#include "render/geometry/sphere.h"
#include "render/geometry/triangle.h"

namespace render {

namespace {

inline void put3(std::vector<float>& v, const Vec3& a) {
    v.push_back(a.x);
    v.push_back(a.y);
    v.push_back(a.z);
}

inline rbr_ray toRay(const Ray& r) {
    rbr_ray x;
    x.o[0] = r.origin.x;
    x.o[1] = r.origin.y;
    x.o[2] = r.origin.z;
    x.d[0] = r.direction.x;
    x.d[1] = r.direction.y;
    x.d[2] = r.direction.z;
    x.tmin = EPSILON;
    x.tmax = r.tMax;
    return x;
}

}  // namespace

GpuBVH::GpuBVH(int device, int width, int leafMax) : device_(device), width_(width), leafMax_(leafMax) {
    if (device_ < 0 || rbr_ctx_create(device_, &ctx_) != RBR_OK) {
        device_ = -1;
        ctx_ = nullptr;
        rbr_ctx_create(-1, &ctx_);
    }
    rbr_scene_create(ctx_, &scene_);
}

GpuBVH::~GpuBVH() {
    rbr_scene_destroy(scene_);
    rbr_ctx_destroy(ctx_);
}

void GpuBVH::build(std::vector<Primitive*> primitives) {
    prims_ = std::move(primitives);
    other_.clear();
    std::vector<float> tris, sphs, quads, planes;
    std::vector<uint32_t> triIds, sphIds, quadIds, planeIds;
    for (size_t i = 0; i < prims_.size(); ++i) {
        const Primitive* p = prims_[i];
        const uint32_t id = uint32_t(i);
        if (auto* t = dynamic_cast<const Triangle*>(p)) {
            put3(tris, t->v0);
            put3(tris, t->v1);
            put3(tris, t->v2);
            triIds.push_back(id);
        } else if (auto* s = dynamic_cast<const Sphere*>(p)) {
            put3(sphs, s->center);
            sphs.push_back(s->radius);
            sphIds.push_back(id);
        } else if (auto* q = dynamic_cast<const Quad*>(p)) {
            put3(quads, q->origin);
            put3(quads, q->u);
            put3(quads, q->v);
            put3(quads, q->normal);
            quadIds.push_back(id);
        } else if (auto* pl = dynamic_cast<const Plane*>(p)) {
            put3(planes, pl->normal);
            planes.push_back(pl->d);
            planeIds.push_back(id);
        } else {
            other_.push_back(prims_[i]);
        }
    }
    rbr_scene_set_geometry(scene_, tris.data(), uint32_t(triIds.size()), triIds.data(), sphs.data(),
                           uint32_t(sphIds.size()), sphIds.data(), quads.data(), uint32_t(quadIds.size()),
                           quadIds.data(), planes.data(), uint32_t(planeIds.size()), planeIds.data());
    rbr_scene_build_bvh(scene_, width_, leafMax_);
}

bool GpuBVH::intersect(const Ray& ray, SurfaceInteraction& isect) const {
    bool hit = false;
    if (!prims_.empty()) {
        rbr_ray r = toRay(ray);
        rbr_hit h;
        if (rbr_host_trace_closest(scene_, &r, 1, &h, nullptr, nullptr) == RBR_OK && h.id != RBR_MISS)
            hit = prims_[h.id]->intersect(ray, isect);   // fills isect, shrinks ray.tMax to t
    }
    for (const Primitive* p : other_)
        if (p->intersect(ray, isect)) hit = true;
    return hit;
}

bool GpuBVH::intersectP(const Ray& ray) const {
    if (!prims_.empty()) {
        rbr_ray r = toRay(ray);
        uint8_t occ = 0;
        if (rbr_host_trace_any(scene_, &r, 1, &occ, nullptr, nullptr) == RBR_OK && occ) return true;
    }
    for (const Primitive* p : other_)
        if (p->intersectP(ray)) return true;
    return false;
}

BoundingBox GpuBVH::worldBounds() const {
    // Same as BVH::worldBounds (bvh.cpp:43-45): the union of every primitive's bounds, an
    // empty box when nothing was built.
    BoundingBox b;
    for (const Primitive* p : prims_) b.expand(p->getBounds());
    return b;
}

bool GpuBVH::traceClosest(const std::vector<Ray>& rays, std::vector<uint32_t>& ids, std::vector<float>& t) const {
    if (!onDevice()) return false;
    std::vector<rbr_ray> r(rays.size());
    for (size_t i = 0; i < rays.size(); ++i) r[i] = toRay(rays[i]);
    std::vector<rbr_hit> h(rays.size());
    if (rbr_trace_closest(scene_, r.data(), r.size(), h.data()) != RBR_OK) return false;
    ids.resize(rays.size());
    t.resize(rays.size());
    for (size_t i = 0; i < rays.size(); ++i) {
        ids[i] = h[i].id;
        t[i] = h[i].t;
    }
    return true;
}

bool GpuBVH::traceAny(const std::vector<Ray>& rays, std::vector<uint8_t>& occluded) const {
    if (!onDevice()) return false;
    std::vector<rbr_ray> r(rays.size());
    for (size_t i = 0; i < rays.size(); ++i) r[i] = toRay(rays[i]);
    occluded.resize(rays.size());
    return rbr_trace_any(scene_, r.data(), r.size(), occluded.data()) == RBR_OK;
}

}  // namespace render
