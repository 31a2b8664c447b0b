// Renders the reference's Cornell box through the reference's own selection points after
// reference_integration.patch: an accelerator parsed by parseAcceleratorType and built by
// Scene::setAcceleratorType / buildAccelerator, an integrator made by createIntegrator from a
// RenderConfig, the camera by createCamera, Integrator::render with a progress callback --
// the steps of apps/main.cpp:50-115.  apps/main.cpp itself reads its options through
// Config::parse (config.cpp), which this toolchain cannot compile (json_parser.h keeps
// std::unordered_map<std::string, ConfigValue> inside ConfigValue, an incomplete type that
// GCC 11's libstdc++ rejects), so the integrator names are mapped here with the same strings
// the patched config.cpp adds.
//
// Usage: rbr_reference_integration <accelerator> <integrator> <width> <height> <spp> <out.pfm>
//   accelerator: any name parseAcceleratorType accepts (bvh, grid, kdtree, octree, gpu_bvh, ...)
//   integrator : pathtracing | volumepath | gpu_pathtracing | gpu_volumepath
// Prints the accelerator and integrator names, whether the render ran on the GPU, and writes
// the film as PFM (RGB float, rows bottom-up).
#include <cstdio>
#include <cstdlib>
#include <string>

#include "cornell_box/cornell_box.h"
#include "render/acceleration/accelerator_factory.h"
#include "render/config/render_config.h"
#include "render/config/render_factory.h"
#include "render/core/film.h"
#include "render/renderers/gpu_path_tracing_integrator.h"

using namespace render;

int main(int argc, char** argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s accelerator integrator width height spp out.pfm\n", argv[0]);
        return 2;
    }
    RenderConfig config;
    if (!parseAcceleratorType(argv[1], config.accelerator)) {
        std::fprintf(stderr, "unknown accelerator %s\n", argv[1]);
        return 2;
    }
    const std::string type = argv[2];
    if (type == "pathtracing") config.integratorType = IntegratorType::PathTracing;
    else if (type == "volumepath") config.integratorType = IntegratorType::VolumetricPath;
    else if (type == "gpu_pathtracing") config.integratorType = IntegratorType::GpuPathTracing;
    else if (type == "gpu_volumepath") config.integratorType = IntegratorType::GpuVolumePath;
    else {
        std::fprintf(stderr, "unknown integrator %s\n", argv[2]);
        return 2;
    }
    config.image.width = std::atoi(argv[3]);
    config.image.height = std::atoi(argv[4]);
    config.samplesPerPixel = std::atoi(argv[5]);

    Scene scene = cornell_box::createScene();
    if (config.accelerator != AcceleratorType::BVH) {
        scene.setAcceleratorType(config.accelerator);
        scene.buildAccelerator();
    }
    auto integrator = createIntegrator(config);
    auto camera = createCamera(config);
    Film film(config.image.width, config.image.height);
    float last = -1.0f;
    int calls = 0;
    integrator->render(scene, *camera, film, [&](float f) {
        last = f;
        ++calls;
    });
    bool onGpu = false;
    if (auto* g = dynamic_cast<GpuPathTracingIntegrator*>(integrator.get())) onGpu = g->lastRenderOnGpu();
    if (auto* g = dynamic_cast<GpuVolumetricPathIntegrator*>(integrator.get())) onGpu = g->lastRenderOnGpu();
    std::printf("Accelerator: %s\nIntegrator: %s\nOn GPU: %s\nProgress: %d calls, last %.3f\n",
                acceleratorName(config.accelerator), integrator->name().c_str(), onGpu ? "yes" : "no", calls, last);

    FILE* f = std::fopen(argv[6], "wb");
    if (!f) return 1;
    std::fprintf(f, "PF\n%d %d\n-1.0\n", film.width, film.height);
    for (int y = film.height - 1; y >= 0; --y)
        for (int x = 0; x < film.width; ++x) {
            const Color c = film.getPixel(x, y);
            const float v[3] = {c.x, c.y, c.z};
            std::fwrite(v, sizeof(float), 3, f);
        }
    std::fclose(f);
    return 0;
}
