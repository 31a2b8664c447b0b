// GpuPathTracingIntegrator / GpuVolumetricPathIntegrator: the reference's Integrator
// interface (include/render/renderers/integrator.h:37-73) with render() replaced by the
// MI355X wavefront path tracer through the C ABI (rbr_render_progress).  They are the
// classes a maintainer adds to the reference as "gpu_pathtracing" / "gpu_volumepath".
//
// render() translates the Scene (Triangle with or without vertex normals, Sphere, Quad,
// Plane; the six-parameter Material; Area/Point/Directional lights; Homogeneous/Exponential
// medium with HG phase) and a PinholeCamera into device tables, renders
// config.samplesPerPixel samples per pixel and adds one sample per pixel equal to the mean, as
// Integrator::render does (integrator.cpp:36).  The device context, the uploaded geometry and
// its BVH are kept between render() calls and rebuilt only when the scene's geometry changes
// (shading tables are re-sent when materials, lights, medium or background change).
// `progress` receives the fraction of path samples done after each round of the launch, from
// a runtime thread, as the reference's workers call it (integrator.cpp:39-45).
// Anything the device does not model (other primitive or light types, a Material subclass with
// its own createBSDF, another camera model, no GPU) falls back to the CPU render of the base
// class; lastRenderOnGpu() tells which path ran.  Li() is the base class's CPU estimator
// (same semantics).
#pragma once

#include <cstdint>
#include <memory>
#include <string>

#include "render/renderers/path_tracing_integrator.h"
#include "render/renderers/volumetric_path_integrator.h"

namespace render {

// Launch settings of the device path (the traversal and render autotuners' choices for the
// benchmark workloads, tuning/traversal.json and tuning/render.json): quantized 4-wide BVH,
// deferred-step traversal (mode 11) with a 12-entry LDS stack and 32 waves per CU, a ray pool
// of 2^25 paths, hipGraph replay and material-ordered shading.
struct GpuTuning {
    int bvhWidth = 4 | 0x100;   // RBR_BVH_QUANTIZED
    int leafMax = 4;
    int traceMode = 11, ldsDepth = 12, wavesPerCu = 32;
    uint32_t poolPaths = 1u << 25;
    int useGraphs = 1, sortMaterials = 1;
};

class GpuRenderCache;   // device state kept between render() calls

class GpuPathTracingIntegrator : public PathTracingIntegrator {
public:
    explicit GpuPathTracingIntegrator(const IntegratorConfig& cfg = IntegratorConfig(), int device = 0,
                                      uint64_t seed = 1);
    ~GpuPathTracingIntegrator() override;
    void render(const Scene& scene, const Camera& camera, Film& film,
                std::function<void(float)> progress = nullptr) const override;
    std::string name() const override { return "GPU Path Tracing"; }
    bool lastRenderOnGpu() const;
    const std::string& lastFallbackReason() const;
    // Number of BVH builds so far (stays put while the same scene is rendered again).
    int bvhBuilds() const;
    void setTuning(const GpuTuning& t);

private:
    std::unique_ptr<GpuRenderCache> cache_;
};

class GpuVolumetricPathIntegrator : public VolumetricPathIntegrator {
public:
    explicit GpuVolumetricPathIntegrator(const IntegratorConfig& cfg = IntegratorConfig(), int device = 0,
                                         uint64_t seed = 1);
    ~GpuVolumetricPathIntegrator() override;
    void render(const Scene& scene, const Camera& camera, Film& film,
                std::function<void(float)> progress = nullptr) const override;
    std::string name() const override { return "GPU Volumetric Path Tracing"; }
    bool lastRenderOnGpu() const;
    const std::string& lastFallbackReason() const;
    int bvhBuilds() const;
    void setTuning(const GpuTuning& t);

private:
    std::unique_ptr<GpuRenderCache> cache_;
};

}  // namespace render
