// GpuBVH: the reference's Accelerator interface (include/render/acceleration/accelerator.h:
// 19-34) backed by the MI355X library through its C ABI (include/rbr.h).  This is the class
// a maintainer adds to the reference as AcceleratorType::GpuBVH ("gpu_bvh"); it compiles
// against the reference's own headers.
//
//  * build() records every primitive by its index in the input vector (the primitive id),
//    converts Triangle / Sphere / Quad / Plane to the flat arrays of rbr_scene_set_geometry
//    and builds the wide BVH once for both the device and the host.  Primitives of other
//    types are kept in a host-side list and tested linearly, as unbounded primitives are in
//    accel_common.h:35-52.
//  * intersect()/intersectP() are called one ray at a time from the CPU integrators'
//    OpenMP threads (integrator.cpp:12-16), so they never launch a kernel: they walk the same
//    flat BVH on the host (rbr_host_trace_*), then re-run the winning primitive's own
//    intersect() with tMax = t so `isect` is filled exactly as the reference fills it.
//  * traceClosest()/traceAny() are the batched GPU entry points.
// Among hits with equal t the lowest primitive id wins (the reference's BVH keeps the last
// one it tested).
#pragma once

#include <cstdint>
#include <vector>

#include "render/acceleration/accelerator.h"

struct rbr_ctx;
struct rbr_scene;

namespace render {

class GpuBVH : public Accelerator {
public:
    // device >= 0: that HIP device; a context that cannot be created (no GPU) falls back to a
    // host-only context, so per-ray queries keep working on any machine.
    explicit GpuBVH(int device = 0, int width = 4, int leafMax = 4);
    ~GpuBVH() override;
    GpuBVH(const GpuBVH&) = delete;
    GpuBVH& operator=(const GpuBVH&) = delete;

    void build(std::vector<Primitive*> primitives) override;
    bool intersect(const Ray& ray, SurfaceInteraction& isect) const override;
    bool intersectP(const Ray& ray) const override;
    BoundingBox worldBounds() const override;

    // Batched queries on the GPU: ids[i] = primitive id (UINT32_MAX on a miss), t[i] = hit t.
    // Returns false when no GPU is available (the caller then uses the per-ray methods).
    bool traceClosest(const std::vector<Ray>& rays, std::vector<uint32_t>& ids, std::vector<float>& t) const;
    bool traceAny(const std::vector<Ray>& rays, std::vector<uint8_t>& occluded) const;

    bool onDevice() const { return device_ >= 0; }
    const std::vector<Primitive*>& primitives() const { return prims_; }

private:
    int device_;
    int width_, leafMax_;
    rbr_ctx* ctx_ = nullptr;
    rbr_scene* scene_ = nullptr;
    std::vector<Primitive*> prims_;      // by primitive id
    std::vector<Primitive*> other_;      // types the device does not know: tested linearly
};

}  // namespace render
