// Checks of the drop-in adapters through the reference's own interfaces, in the style of the
// reference's tests (tests/test_accelerators.cpp: a local check() counting failures, the exit
// code is the result).  Usage: rbr_adapter_tests [device]   (device -1: host only)
//
//  * GpuBVH must return exactly what a brute-force loop returns (hit/miss and t; here also
//    the primitive, with the lowest-id rule at exact ties), on the reference test's worlds
//    (random spheres + triangles + axis-aligned flat triangles + an infinite plane) and ray
//    mix (a third axis-aligned, a quarter with finite tMax); empty, single-sphere (t = 4) and
//    rebuild cases; batched GPU queries equal the per-ray ones.
//  * GpuPathTracingIntegrator renders a Cornell scene through Integrator::render; its image
//    mean matches the CPU PathTracingIntegrator's within the Monte Carlo error.
#include <cmath>
#include <cstdio>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "gpu_bvh.h"
#include "gpu_path_tracing_integrator.h"
#include "render/camera/pinhole_camera.h"
#include "render/geometry/plane.h"
#include "render/geometry/quad.h"
#include "render/geometry/sphere.h"
#include "render/geometry/triangle.h"
#include "render/lighting/area_light.h"
#include "render/scene/scene_builder.h"

using namespace render;

namespace {

int failures = 0;

void check(bool ok, const std::string& what) {
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s\n", what.c_str());
        ++failures;
    }
}

struct World {
    std::vector<std::unique_ptr<Primitive>> owned;
    std::vector<Primitive*> prims;
    void add(std::unique_ptr<Primitive> p) {
        prims.push_back(p.get());
        owned.push_back(std::move(p));
    }
};

World makeWorld(std::mt19937& rng, int nSpheres, int nTriangles, bool withPlane) {
    std::uniform_real_distribution<float> pos(-10.0f, 10.0f), size(0.1f, 1.5f);
    World w;
    for (int i = 0; i < nSpheres; ++i) w.add(std::make_unique<Sphere>(Vec3(pos(rng), pos(rng), pos(rng)), size(rng)));
    for (int i = 0; i < nTriangles; ++i) {
        Vec3 c(pos(rng), pos(rng), pos(rng));
        auto jitter = [&]() { return Vec3(size(rng), size(rng), size(rng)) - Vec3(0.8f); };
        w.add(std::make_unique<Triangle>(c + jitter(), c + jitter(), c + jitter()));
    }
    for (int i = 0; i < 20; ++i) {
        float z = pos(rng), x = pos(rng), y = pos(rng);
        w.add(std::make_unique<Triangle>(Vec3(x, y, z), Vec3(x + 2, y, z), Vec3(x, y + 2, z)));
    }
    if (withPlane) w.add(std::make_unique<Plane>(Vec3(0, 1, 0), 12.0f));
    return w;
}

// Brute force with the lowest-id rule at equal t.
bool bruteForce(const std::vector<Primitive*>& prims, const Ray& ray, float& tHit, int& id) {
    bool hit = false;
    tHit = ray.tMax;
    id = -1;
    for (size_t i = 0; i < prims.size(); ++i) {
        Ray r = ray;
        r.tMax = tHit;
        SurfaceInteraction is;
        if (prims[i]->intersect(r, is) && (!hit || is.t < tHit)) {
            hit = true;
            tHit = is.t;
            id = int(i);
        }
    }
    return hit;
}

void compare(GpuBVH& accel, const World& world, std::mt19937& rng, int nRays, const char* label) {
    accel.build(world.prims);
    std::uniform_real_distribution<float> pos(-14.0f, 14.0f), dir(-1.0f, 1.0f), len(0.5f, 40.0f);
    int mismatches = 0, mismatchesP = 0, idMismatch = 0, hits = 0, batchMismatch = 0;
    std::vector<Ray> rays;
    std::vector<int> refIds;
    for (int i = 0; i < nRays; ++i) {
        Vec3 o(pos(rng), pos(rng), pos(rng));
        Vec3 d(dir(rng), dir(rng), dir(rng));
        if (i % 3 == 0) d = Vec3(0.0f, 0.0f, (i % 2) ? 1.0f : -1.0f);
        else if (i % 7 == 0) d = Vec3(dir(rng), 0.0f, dir(rng));
        if (d.length() < 1e-3f) continue;
        Ray ray(o, d);
        if (i % 4 == 0) ray.tMax = len(rng);
        float tRef;
        int idRef;
        const bool refHit = bruteForce(world.prims, ray, tRef, idRef);
        hits += refHit;
        Ray r1 = ray;
        SurfaceInteraction isect;
        const bool hit = accel.intersect(r1, isect);
        if (hit != refHit || (hit && isect.t != tRef)) ++mismatches;
        if (hit && refHit && isect.primitive != world.prims[idRef]) ++idMismatch;
        bool anyRef = false;
        for (const Primitive* p : world.prims)
            if (p->intersectP(ray)) { anyRef = true; break; }
        if (accel.intersectP(ray) != anyRef) ++mismatchesP;
        rays.push_back(ray);
        refIds.push_back(idRef);
    }
    std::vector<uint32_t> ids;
    std::vector<float> ts;
    if (accel.traceClosest(rays, ids, ts)) {
        for (size_t i = 0; i < rays.size(); ++i)
            if (int64_t(ids[i] == 0xFFFFFFFFu ? -1 : int64_t(ids[i])) != refIds[i]) ++batchMismatch;
    }
    check(mismatches == 0, std::string("gpu_bvh closest hit (") + label + "): " + std::to_string(mismatches));
    check(idMismatch == 0, std::string("gpu_bvh primitive id (") + label + "): " + std::to_string(idMismatch));
    check(mismatchesP == 0, std::string("gpu_bvh any hit (") + label + "): " + std::to_string(mismatchesP));
    check(batchMismatch == 0, std::string("gpu_bvh batched ids (") + label + "): " + std::to_string(batchMismatch));
    check(hits > nRays / 20, std::string("test rays hit nothing (") + label + ")");
}

std::unique_ptr<Scene> cornell(SceneBuilder& b) {
    Material* white = b.diffuse(Color(0.73f));
    Material* red = b.diffuse(Color(0.65f, 0.05f, 0.05f));
    Material* green = b.diffuse(Color(0.12f, 0.45f, 0.15f));
    Material* light = b.emissive(Color(17.0f, 12.0f, 4.0f));
    b.quad(Vec3(0, 0, 0), Vec3(555, 0, 0), Vec3(0, 0, 555), white);
    b.quad(Vec3(0, 555, 0), Vec3(555, 0, 0), Vec3(0, 0, 555), white);
    b.quad(Vec3(0, 0, 555), Vec3(555, 0, 0), Vec3(0, 555, 0), white);
    b.quad(Vec3(555, 0, 0), Vec3(0, 555, 0), Vec3(0, 0, 555), red);
    b.quad(Vec3(0, 0, 0), Vec3(0, 555, 0), Vec3(0, 0, 555), green);
    b.quad(Vec3(213, 554, 227), Vec3(130, 0, 0), Vec3(0, 0, 105), light);
    b.sphere(Vec3(190, 90, 190), 90, white);
    b.sphere(Vec3(370, 90, 350), 90, b.glass(1.5f));
    b.areaLight(Vec3(278, 554, 279.5f), Vec3(130, 0, 0), Vec3(0, 0, 105), Color(17, 12, 4), 1.0f);
    return b.build();
}

double filmMean(const Film& f) {
    double s = 0.0;
    for (int y = 0; y < f.height; ++y)
        for (int x = 0; x < f.width; ++x) {
            Color c = f.getPixel(x, y);
            s += c.x + c.y + c.z;
        }
    return s / (3.0 * f.width * f.height);
}

}  // namespace

int main(int argc, char** argv) {
    const int device = argc > 1 ? std::atoi(argv[1]) : 0;
    std::mt19937 rng(12345);
    {
        GpuBVH empty(device);
        empty.build({});
        Ray ray(Vec3(0, 0, 5), Vec3(0, 0, -1));
        SurfaceInteraction isect;
        check(!empty.intersect(ray, isect), "gpu_bvh empty hit");
        check(!empty.intersectP(ray), "gpu_bvh empty hitP");
        check(empty.worldBounds().isEmpty(), "gpu_bvh empty bounds");
        World one;
        one.add(std::make_unique<Sphere>(Vec3(0, 0, 0), 1.0f));
        GpuBVH single(device);
        single.build(one.prims);
        Ray r2(Vec3(0, 0, 5), Vec3(0, 0, -1));
        check(single.intersect(r2, isect) && std::fabs(isect.t - 4.0f) < 1e-3f && std::fabs(r2.tMax - 4.0f) < 1e-3f,
              "gpu_bvh single sphere");
    }
    GpuBVH accel(device);
    std::printf("gpu_bvh on %s\n", accel.onDevice() ? "the GPU" : "the host");
    World small = makeWorld(rng, 30, 30, false);
    compare(accel, small, rng, 3000, "small");
    World big = makeWorld(rng, 400, 1500, false);
    compare(accel, big, rng, 6000, "large");
    World withPlane = makeWorld(rng, 60, 200, true);
    compare(accel, withPlane, rng, 3000, "with infinite plane");
    accel.build(big.prims);
    accel.build(small.prims);
    {
        Ray probe(Vec3(0, 0, 50), Vec3(0, 0, -1));
        SurfaceInteraction isect;
        float tRef;
        int idRef;
        const bool refHit = bruteForce(small.prims, probe, tRef, idRef);
        const bool hit = accel.intersect(probe, isect);
        check(hit == refHit && (!hit || isect.t == tRef), "gpu_bvh rebuild");
    }

    // Integrator drop-in: same image mean as the CPU integrator within the noise.
    SceneBuilder b1, b2;
    auto scene = cornell(b1);
    PinholeCamera cam(64, 64, Vec3(278, 278, -800), Vec3(278, 278, 0), Vec3(0, 1, 0), 40.0f);
    IntegratorConfig ic;
    ic.samplesPerPixel = 64;
    ic.maxDepth = 8;
    GpuPathTracingIntegrator gpu(ic, device, 7);
    PathTracingIntegrator cpu(ic);
    Film fg(64, 64), fc(64, 64);
    gpu.render(*scene, cam, fg);
    cpu.render(*scene, cam, fc);
    const double mg = filmMean(fg), mc = filmMean(fc);
    std::printf("integrator: %s, mean %.5f vs CPU %.5f%s%s\n", gpu.lastRenderOnGpu() ? "GPU" : "CPU fallback", mg, mc,
                gpu.lastRenderOnGpu() ? "" : ", reason: ", gpu.lastFallbackReason().c_str());
    check(std::fabs(mg - mc) < 0.02 * mc, "gpu_pathtracing image mean");
    if (device >= 0) check(gpu.lastRenderOnGpu(), "gpu_pathtracing ran on the GPU");
    else check(!gpu.lastRenderOnGpu(), "host-only run falls back to the CPU integrator");

    if (failures == 0) std::printf("All adapter tests passed\n");
    return failures == 0 ? 0 : 1;
}
