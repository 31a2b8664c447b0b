// Checks of the drop-in adapters through the reference's own interfaces, in the style of the
// reference's tests (tests/test_accelerators.cpp: a local check() counting failures, the exit
// code is the result).  Usage: rbr_adapter_tests [device [accel]]   (device -1: host only;
// accel: the accelerator checks only)
//
//  * GpuBVH must return exactly what a brute-force loop returns (hit/miss and t; here also
//    the primitive, with the lowest-id rule at exact ties), on the reference test's worlds
//    (random spheres + triangles + axis-aligned flat triangles + an infinite plane) and ray
//    mix (a third axis-aligned, a quarter with finite tMax); empty, single-sphere (t = 4) and
//    rebuild cases; batched GPU queries equal the per-ray ones.
//  * GpuBVH's per-ray queries from 8 threads at once equal the single-threaded answers
//    (integrator.cpp:12-16 calls them from OpenMP workers; std::threads here so that a
//    ThreadSanitizer build of this program sees every access of the adapter code).
//  * GpuPathTracingIntegrator renders through Integrator::render: on a Cornell scene with a
//    glass sphere and on a ball of triangles with vertex normals, the mean of every 16x16
//    block of the image matches the CPU PathTracingIntegrator's within the Monte Carlo error;
//    a second render of the same scene reuses the device BVH (a material change is picked up
//    without a rebuild); the progress callback reports increasing fractions ending at 1.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "gpu_bvh.h"
#include "gpu_path_tracing_integrator.h"
#include "render/camera/pinhole_camera.h"
#include "render/geometry/plane.h"
#include "render/geometry/quad.h"
#include "render/geometry/sphere.h"
#include "render/geometry/triangle.h"
#include "render/geometry/mesh.h"
#include "render/lighting/area_light.h"
#include "render/lighting/point_light.h"
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

// Block means of two films of one scene: every 16x16 block whose mean exceeds 0.02 agrees
// within `tol` relative (the blocks average 256 pixels x spp samples).
void compareBlocks(const Film& a, const Film& b, double tol, const std::string& label) {
    const int bs = 16;
    int bad = 0, used = 0;
    double worst = 0.0;
    for (int by = 0; by + bs <= a.height; by += bs)
        for (int bx = 0; bx + bs <= a.width; bx += bs) {
            double sa = 0.0, sb = 0.0;
            for (int y = by; y < by + bs; ++y)
                for (int x = bx; x < bx + bs; ++x) {
                    Color ca = a.getPixel(x, y), cb = b.getPixel(x, y);
                    sa += ca.x + ca.y + ca.z;
                    sb += cb.x + cb.y + cb.z;
                }
            sa /= 3.0 * bs * bs;
            sb /= 3.0 * bs * bs;
            if (sb < 0.02) continue;
            ++used;
            const double rel = std::fabs(sa - sb) / sb;
            worst = std::max(worst, rel);
            if (rel > tol) ++bad;
        }
    std::printf("  %s: %d blocks compared, worst relative difference %.4f\n", label.c_str(), used, worst);
    check(used > 0 && bad == 0, label + ": block means differ (" + std::to_string(bad) + " blocks)");
}

// A ball of triangles with vertex normals (the sphere's normals) on a floor, lit by a point
// light: TriangleMesh::toPrimitives keeps the normals (mesh.cpp:29-32).
std::unique_ptr<Scene> smoothBall(SceneBuilder& b) {
    TriangleMesh m;
    const int nu = 24, nv = 12;
    for (int j = 0; j <= nv; ++j)
        for (int i = 0; i <= nu; ++i) {
            const float th = 3.14159265f * float(j) / nv, ph = 2.0f * 3.14159265f * float(i) / nu;
            const Vec3 n(std::sin(th) * std::cos(ph), std::cos(th), std::sin(th) * std::sin(ph));
            m.positions.push_back(Vec3(0.0f, 1.0f, 0.0f) + n);
            m.normals.push_back(n);
        }
    for (int j = 0; j < nv; ++j)
        for (int i = 0; i < nu; ++i) {
            const int a = j * (nu + 1) + i, c = a + nu + 1;
            m.indices.insert(m.indices.end(), {a, c, a + 1, a + 1, c, c + 1});
        }
    b.mesh(m, b.diffuse(Color(0.8f, 0.5f, 0.3f)));
    b.quad(Vec3(-5, 0, -5), Vec3(10, 0, 0), Vec3(0, 0, 10), b.diffuse(Color(0.6f)));
    b.pointLight(Vec3(2, 5, -3), Color(1.0f), 40.0f);
    return b.build();
}

// Per-ray queries of one GpuBVH from 8 threads against the single-threaded answers.
void concurrentQueries(GpuBVH& accel, const World& world, std::mt19937& rng) {
    accel.build(world.prims);
    std::uniform_real_distribution<float> pos(-14.0f, 14.0f), dir(-1.0f, 1.0f);
    std::vector<Ray> rays;
    for (int i = 0; i < 4000; ++i) {
        Vec3 d(dir(rng), dir(rng), dir(rng));
        if (d.length() < 1e-3f) continue;
        rays.emplace_back(Vec3(pos(rng), pos(rng), pos(rng)), d);
    }
    std::vector<float> t0(rays.size());
    std::vector<int> h0(rays.size()), p0(rays.size());
    for (size_t i = 0; i < rays.size(); ++i) {
        Ray r = rays[i];
        SurfaceInteraction is;
        h0[i] = accel.intersect(r, is) ? 1 : 0;
        t0[i] = h0[i] ? is.t : 0.0f;
        p0[i] = accel.intersectP(rays[i]) ? 1 : 0;
    }
    std::atomic<int> mism{0};
    std::vector<std::thread> th;
    for (int k = 0; k < 8; ++k)
        th.emplace_back([&, k]() {
            for (size_t i = k; i < rays.size(); i += 8) {
                Ray r = rays[i];
                SurfaceInteraction is;
                const int h = accel.intersect(r, is) ? 1 : 0;
                if (h != h0[i] || (h && is.t != t0[i]) || (accel.intersectP(rays[i]) ? 1 : 0) != p0[i]) ++mism;
            }
        });
    for (auto& t : th) t.join();
    std::printf("  concurrent per-ray queries: %zu rays x 8 threads, %d mismatches\n", rays.size(), mism.load());
    check(mism.load() == 0, "gpu_bvh concurrent per-ray queries");
}

}  // namespace

int main(int argc, char** argv) {
    const int device = argc > 1 ? std::atoi(argv[1]) : 0;
    // "accel": the accelerator checks only (the ThreadSanitizer build runs these; its
    // renders would run single-threaded and take minutes).
    const bool accelOnly = argc > 2 && std::string(argv[2]) == "accel";
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

    concurrentQueries(accel, big, rng);
    if (accelOnly) {
        if (failures == 0) std::printf("All adapter tests passed\n");
        return failures == 0 ? 0 : 1;
    }

    // Integrator drop-in: the CPU integrator's image within the noise, block by block.
    SceneBuilder b1, b2;
    auto scene = cornell(b1);
    PinholeCamera cam(64, 64, Vec3(278, 278, -800), Vec3(278, 278, 0), Vec3(0, 1, 0), 40.0f);
    IntegratorConfig ic;
    ic.samplesPerPixel = 1024;   // the glass sphere's caustics are heavy-tailed
    ic.maxDepth = 8;
    GpuPathTracingIntegrator gpu(ic, device, 7);
    PathTracingIntegrator cpu(ic);
    Film fg(64, 64), fc(64, 64);
    std::mutex pm;
    std::vector<float> fractions;
    gpu.render(*scene, cam, fg, [&](float f) {
        std::lock_guard<std::mutex> lock(pm);
        fractions.push_back(f);
    });
    cpu.render(*scene, cam, fc);
    const double mg = filmMean(fg), mc = filmMean(fc);
    std::printf("integrator: %s, mean %.5f vs CPU %.5f%s%s\n", gpu.lastRenderOnGpu() ? "GPU" : "CPU fallback", mg, mc,
                gpu.lastRenderOnGpu() ? "" : ", reason: ", gpu.lastFallbackReason().c_str());
    check(std::fabs(mg - mc) < 0.02 * mc, "gpu_pathtracing image mean");
    compareBlocks(fg, fc, 0.08, "cornell blocks");
    check(!fractions.empty() && fractions.back() == 1.0f && std::is_sorted(fractions.begin(), fractions.end()),
          "progress fractions increase to 1");
    if (device >= 0) {
        check(gpu.lastRenderOnGpu(), "gpu_pathtracing ran on the GPU");
        // Same scene again: no rebuild.  A changed material is picked up without one.
        Film f2(64, 64);
        gpu.render(*scene, cam, f2);
        check(gpu.bvhBuilds() == 1, "second render of the same scene reuses the BVH");
        scene->materials()[0]->baseColor = Color(0.2f);
        Film f3(64, 64);
        gpu.render(*scene, cam, f3);
        check(gpu.bvhBuilds() == 1 && filmMean(f3) < 0.9 * filmMean(f2), "material change without a rebuild");
        scene->materials()[0]->baseColor = Color(0.73f);
    } else {
        check(!gpu.lastRenderOnGpu(), "host-only run falls back to the CPU integrator");
    }

    // Triangles with vertex normals render on the device with the interpolated normal.
    SceneBuilder b3;
    auto ball = smoothBall(b3);
    PinholeCamera cam2(64, 64, Vec3(0, 2.5f, -5), Vec3(0, 0.8f, 0), Vec3(0, 1, 0), 45.0f);
    GpuPathTracingIntegrator gpu2(ic, device, 9);
    Film fs(64, 64), fsc(64, 64);
    gpu2.render(*ball, cam2, fs);
    cpu.render(*ball, cam2, fsc);
    std::printf("smooth ball: %s, mean %.5f vs CPU %.5f\n", gpu2.lastRenderOnGpu() ? "GPU" : "CPU fallback",
                filmMean(fs), filmMean(fsc));
    if (device >= 0) check(gpu2.lastRenderOnGpu(), "vertex-normal triangles render on the GPU");
    compareBlocks(fs, fsc, 0.08, "smooth ball blocks");

    if (failures == 0) std::printf("All adapter tests passed\n");
    return failures == 0 ? 0 : 1;
}
