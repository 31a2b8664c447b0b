// Internal definitions behind the opaque rbr_ctx / rbr_scene handles.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rbr.h"
#include "bvh.h"
#include "trace.h"

namespace rbr {

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct RenderState;   // render.hip
struct GuideState;    // guide_dev.h

}  // namespace rbr

struct rbr_ctx {
    int device = 0;
    int cuCount = 0;
    std::string name;
    hipStream_t stream = nullptr;      // the stream every launch goes to (ownStream or the caller's)
    hipStream_t ownStream = nullptr;   // created and destroyed by the context
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    rbr::RenderState* render = nullptr;
    rbr::GuideState* guide = nullptr;
    // Traversal tuning (rbr_set_trace_tuning) and persistent-launch state.
    int traceMode = 0;                 // 0 persistent waves, 1 one ray per lane, 2-4 per-lane fetch, 5-12 steps
    int traceModeAny = -1;             // mode of any-hit launches (-1: traceMode)
    int ldsDepth = rbr::kLdsDepthDefault;
    int wavesPerCu = rbr::kWavesPerCuDefault;
    rbr::DevBuf dCounter, dSpill;
    // Traversal watchdog (rbr_set_trace_step_cap): device flag, pinned host copy, cap override.
    rbr::DevBuf dFault;
    uint32_t* hFault = nullptr;
    uint32_t stepCap = 0;
    // Render tuning (rbr_set_render_tuning).
    uint32_t renderPool = 1u << 24;    // paths in flight (queue capacity)
    int renderGraphs = 1;              // replay each round's launch sequence as a hipGraph
    int renderSort = 1;                // shade the extension queue in material order
    uint32_t renderCensus = 0;         // rays sampled per queue and bounce for rbr_last_render_stats (0: off)
    // Temporaries of the device BVH build, kept (and grown) between builds so that a rebuild
    // allocates nothing once they fit.
    std::vector<rbr::DevBuf> buildTmp;
};

struct rbr_scene {
    rbr_ctx* ctx = nullptr;
    rbr::HostGeometry geom;
    std::vector<float> triNormals;   // unit corner normals, 9 per triangle of geom.tris, or empty
    rbr::WideBVH bvh;
    bool built = false;
    rbr::DevBuf dNodes, dPrims, dQuadN, dPlanes, dPlaneIds;
    rbr::DevBuf dRays, dOut;   // staging for the host-array trace entry points
    // Device build: the float nodes stay on the device, and the host copies of the nodes and
    // records (host traversal, render arena, exports) are made on first use (ensureHostBvh).
    rbr::DevBuf dFloatNodes;
    std::atomic<bool> hostStale{false};
    std::mutex hostMu;
    // Shading tables (render.hip).
    bool shaded = false;
    std::vector<rbr_material> materials;
    std::vector<int32_t> primMaterial;
    std::vector<rbr_light> lights;
    rbr_medium medium{};
    float background[3] = {0.0f, 0.0f, 0.0f};
    rbr::DevBuf dMaterials, dPrimMaterial, dLights;
    // Changes whenever geometry, BVH or shading change (unique across scenes), so render
    // state built from the scene can be reused while it is unchanged.
    uint64_t version = 0;
};

namespace rbr {

int fail(rbr_ctx* ctx, int code, const std::string& msg);
int failHip(rbr_ctx* ctx, hipError_t e, const char* what);
int ensureBuffer(rbr_ctx* ctx, DevBuf& b, size_t bytes);
void freeBuffer(DevBuf& b);
int uploadBuffer(rbr_ctx* ctx, DevBuf& b, const void* src, size_t bytes);
int useDevice(rbr_ctx* ctx);
TraceArgs sceneTraceArgs(const rbr_scene* s);
// Launches closest (any=false) or any-hit traversal of `a` (rays/outputs filled in) for a
// scene whose wide BVH has the given width and stack need, using the context's tuning.
int launchTraversal(rbr_ctx* ctx, TraceArgs a, int width, uint32_t maxStack, bool any);
// Fills in the watchdog fields of `a`: the context's cap override (else a.stepCap as the
// caller set it, the bound of the BVH traversed) and its fault flag.
void armWatchdog(rbr_ctx* ctx, TraceArgs& a);
// Enqueues the read of the fault flag on the context stream; after the stream is
// synchronized, checkFault returns RBR_ERR_WATCHDOG (and clears the flag) if it was raised.
int queueFaultRead(rbr_ctx* ctx);
int checkFault(rbr_ctx* ctx);
// Allocates what launchTraversal needs for a BVH of this width and stack need, so that the
// launches themselves allocate nothing (stream capture).
int prepareTraversal(rbr_ctx* ctx, uint32_t maxStack);

// Device build of the scene's wide BVH (bvh_gpu.hip): fills the scene's device node, float
// node and primitive buffers and s->bvh's counts and bounds; the host arrays of s->bvh stay
// empty until ensureHostBvh.
int buildWideBVHDevice(rbr_ctx* ctx, rbr_scene* s, int width, int leafMax, bool quant, bool optimal, bool ploc);

// Makes s->bvh's host arrays (float nodes, device-layout nodes, records) of a device-built
// tree from its device buffers, once; a no-op for host-built trees.  Safe to call from
// several threads at once.
int ensureHostBvh(const rbr_scene* s);

// Gives the scene a new version number (after any change to it).
void bumpVersion(rbr_scene* s);
void freeShading(rbr_scene* s);          // render.hip
void destroyRenderState(rbr_ctx* ctx);   // render.hip
void destroyGuideState(rbr_ctx* ctx);    // guide.hip

}  // namespace rbr
