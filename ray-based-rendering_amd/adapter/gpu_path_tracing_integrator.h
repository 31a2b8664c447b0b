// GpuPathTracingIntegrator / GpuVolumetricPathIntegrator: the reference's Integrator
// interface (include/render/renderers/integrator.h:37-73) with render() replaced by the
// MI355X wavefront path tracer through the C ABI (rbr_render).  They are the classes a
// maintainer adds to the reference as "gpu_pathtracing" / "gpu_volumepath".
//
// render() translates the Scene (Triangle/Sphere/Quad/Plane primitives, the six-parameter
// Material, Area/Point/Directional lights, Homogeneous/Exponential medium with HG phase) and
// a PinholeCamera into device tables, renders config.samplesPerPixel samples per pixel and
// adds one sample per pixel equal to the mean, as Integrator::render does
// (integrator.cpp:36).  Anything the device does not model (other primitive or light types,
// a Material subclass with its own createBSDF, another camera model, no GPU) falls back to
// the CPU render of the base class; lastRenderOnGpu() tells which path ran.  Li() is the
// base class's CPU estimator (same semantics).
#pragma once

#include <cstdint>
#include <string>

#include "render/renderers/path_tracing_integrator.h"
#include "render/renderers/volumetric_path_integrator.h"

namespace render {

class GpuPathTracingIntegrator : public PathTracingIntegrator {
public:
    explicit GpuPathTracingIntegrator(const IntegratorConfig& cfg = IntegratorConfig(), int device = 0,
                                      uint64_t seed = 1)
        : PathTracingIntegrator(cfg), device_(device), seed_(seed) {}
    void render(const Scene& scene, const Camera& camera, Film& film,
                std::function<void(float)> progress = nullptr) const override;
    std::string name() const override { return "GPU Path Tracing"; }
    bool lastRenderOnGpu() const { return onGpu_; }
    const std::string& lastFallbackReason() const { return reason_; }

private:
    int device_;
    uint64_t seed_;
    mutable bool onGpu_ = false;
    mutable std::string reason_;
};

class GpuVolumetricPathIntegrator : public VolumetricPathIntegrator {
public:
    explicit GpuVolumetricPathIntegrator(const IntegratorConfig& cfg = IntegratorConfig(), int device = 0,
                                         uint64_t seed = 1)
        : VolumetricPathIntegrator(cfg, VolumeConfig{}), device_(device), seed_(seed) {}
    void render(const Scene& scene, const Camera& camera, Film& film,
                std::function<void(float)> progress = nullptr) const override;
    std::string name() const override { return "GPU Volumetric Path Tracing"; }
    bool lastRenderOnGpu() const { return onGpu_; }
    const std::string& lastFallbackReason() const { return reason_; }

private:
    int device_;
    uint64_t seed_;
    mutable bool onGpu_ = false;
    mutable std::string reason_;
};

// Renders `scene` with the GPU path; returns false (and a reason) when the scene or camera
// uses something the device path does not model or no GPU is present.
bool gpuRender(const Scene& scene, const Camera& camera, Film& film, const IntegratorConfig& cfg, int integrator,
               int device, uint64_t seed, std::string& reason);

}  // namespace render
