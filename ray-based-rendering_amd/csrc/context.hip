// C ABI (include/rbr.h): contexts, scenes, geometry upload, BVH build and traversal.
#include <hip/hip_runtime.h>

#include <atomic>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "context.h"
#include "prim_tests.h"
#include "trace.h"

using namespace rbr;

// ---------------------------------------------------------------- helpers

int rbr::fail(rbr_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int rbr::failHip(rbr_ctx* ctx, hipError_t e, const char* what) {
    return fail(ctx, RBR_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int rbr::ensureBuffer(rbr_ctx* ctx, DevBuf& b, size_t bytes) {
    if (b.bytes >= bytes && b.ptr) return RBR_OK;
    if (b.ptr) {
        hipError_t e = hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
        if (e != hipSuccess) return failHip(ctx, e, "hipFree");
    }
    if (bytes == 0) return RBR_OK;
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        return failHip(ctx, e, "hipMalloc");
    }
    b.bytes = bytes;
    return RBR_OK;
}

void rbr::freeBuffer(DevBuf& b) {
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

int rbr::uploadBuffer(rbr_ctx* ctx, DevBuf& b, const void* src, size_t bytes) {
    int rc = ensureBuffer(ctx, b, bytes);
    if (rc) return rc;
    if (bytes == 0 || !src) return RBR_OK;   // no source (an empty table): allocation only
    hipError_t e = hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return failHip(ctx, e, "hipMemcpyAsync H2D");
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return failHip(ctx, e, "hipStreamSynchronize");
    return RBR_OK;
}

int rbr::ensureHostBvh(const rbr_scene* cs) {
    rbr_scene* s = const_cast<rbr_scene*>(cs);   // the host copies are a cache of the device tree
    if (!s->hostStale.load(std::memory_order_acquire)) return RBR_OK;
    std::lock_guard<std::mutex> lock(s->hostMu);
    if (!s->hostStale.load(std::memory_order_relaxed)) return RBR_OK;
    rbr_ctx* ctx = s->ctx;
    if (int rc = useDevice(ctx)) return rc;
    WideBVH& b = s->bvh;
    b.nodes.resize(size_t(b.nodeCount) * size_t(b.f4PerNode()));
    b.prims.resize(size_t(b.primCount) * 3);
    // float nodes: the scene's float copy of a quantized tree, else its device nodes
    const void* floatSrc = b.quantized ? s->dFloatNodes.ptr : s->dNodes.ptr;
    hipError_t e = hipMemcpyAsync(b.nodes.data(), floatSrc, b.nodes.size() * sizeof(F4), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && b.quantized) {
        b.qnodes.resize(size_t(b.nodeCount) * size_t(qNodeDwords(b.width)));
        e = hipMemcpyAsync(b.qnodes.data(), s->dNodes.ptr, b.qnodes.size() * 4, hipMemcpyDeviceToHost, ctx->stream);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(b.prims.data(), s->dPrims.ptr, b.primBytes(), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return failHip(ctx, e, "BVH host copy");
    s->hostStale.store(false, std::memory_order_release);
    return RBR_OK;
}

void rbr::bumpVersion(rbr_scene* s) {
    static std::atomic<uint64_t> counter{0};
    s->version = ++counter;
}

int rbr::useDevice(rbr_ctx* ctx) {
    if (ctx->device < 0) return fail(ctx, RBR_ERR_STATE, "host-only context (device -1) cannot run kernels");
    hipError_t e = hipSetDevice(ctx->device);
    return e == hipSuccess ? RBR_OK : failHip(ctx, e, "hipSetDevice");
}

static void freeSceneDevice(rbr_scene* s) {
    freeBuffer(s->dFloatNodes);
    freeBuffer(s->dNodes);
    freeBuffer(s->dPrims);
    freeBuffer(s->dQuadN);
    freeBuffer(s->dPlanes);
    freeBuffer(s->dPlaneIds);
    freeBuffer(s->dRays);
    freeBuffer(s->dOut);
    freeShading(s);
}

static TraceArgs traceArgs(const rbr_scene* s) {
    TraceArgs a{};
    // Watchdog bound: a ray enters each node at most once.
    a.stepCap = s->bvh.nodeCount + 1u;
    a.nodes = s->bvh.empty() ? nullptr : static_cast<const float4*>(s->dNodes.ptr);
    a.prims = static_cast<const float4*>(s->dPrims.ptr);
    a.quadN = static_cast<const float4*>(s->dQuadN.ptr);
    a.planes = static_cast<const float4*>(s->dPlanes.ptr);
    a.planeIds = static_cast<const uint32_t*>(s->dPlaneIds.ptr);
    a.nplanes = uint32_t(s->geom.planeIds.size());
    a.mixed = (s->geom.sphIds.size() + s->geom.quadIds.size()) > 0 ? 1u : 0u;
    return a;
}

TraceArgs rbr::sceneTraceArgs(const rbr_scene* s) { return traceArgs(s); }

void rbr::armWatchdog(rbr_ctx* ctx, TraceArgs& a) {
    if (ctx->stepCap) a.stepCap = ctx->stepCap;
    a.fault = static_cast<uint32_t*>(ctx->dFault.ptr);
}

int rbr::queueFaultRead(rbr_ctx* ctx) {
    if (!ctx->hFault) return RBR_OK;
    hipError_t e = hipMemcpyAsync(ctx->hFault, ctx->dFault.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
    return e == hipSuccess ? RBR_OK : failHip(ctx, e, "watchdog flag read");
}

int rbr::checkFault(rbr_ctx* ctx) {
    if (!ctx->hFault || *ctx->hFault == 0u) return RBR_OK;
    *ctx->hFault = 0u;
    hipError_t e = hipMemcpy(ctx->dFault.ptr, ctx->hFault, sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) return failHip(ctx, e, "watchdog flag clear");
    return fail(ctx, RBR_ERR_WATCHDOG,
                "traversal watchdog: a ray exceeded the step cap (corrupt BVH or rbr_set_trace_step_cap below the "
                "tree's node and primitive count); results since the last check are incomplete");
}

int rbr::launchTraversal(rbr_ctx* ctx, TraceArgs a, int width, uint32_t maxStack, bool any) {
    if (a.n == 0 && !a.nDev) return RBR_OK;
    armWatchdog(ctx, a);
    hipError_t e;
    const int m = any && ctx->traceModeAny >= 0 ? ctx->traceModeAny : ctx->traceMode;
    if (m == 1 && !a.nDev) {
        if (maxStack > uint32_t(kMaxStack)) return fail(ctx, RBR_ERR_UNSUPPORTED, "BVH too deep for the lane stack");
        e = launchTrace(a, width, any, ctx->ldsDepth, nullptr, 0, ctx->stream);
        return e == hipSuccess ? RBR_OK : failHip(ctx, e, "trace launch");
    }
    const unsigned grid = unsigned(std::max(1, ctx->cuCount) * ctx->wavesPerCu);
    const uint32_t spillDepth = std::max<uint32_t>(1u, maxStack > uint32_t(ctx->ldsDepth) ? maxStack - ctx->ldsDepth : 1u);
    if (int rc = ensureBuffer(ctx, ctx->dCounter, kCounterBytes)) return rc;
    if (int rc = ensureBuffer(ctx, ctx->dSpill, size_t(grid) * spillDepth * 64 * sizeof(uint2))) return rc;
    // modes 2, 3, 4: per-lane fetch once 8, 16, 32 lanes of a wave are idle (while-while rounds);
    // modes 5, 6, 7, 8: per-lane fetch once 4, 8, 16, 32 lanes are idle (single steps);
    // modes 9, 10, 11, 12: as 5-8 with primitive steps deferred until 32 lanes hold a leaf or
    // fewer than 48 lanes have an interior node
    const int ms = m >= 9 ? m - 4 : m;
    const uint32_t refill = ms >= 5 ? (2u << (ms - 4)) : ms >= 2 ? (4u << (ms - 1)) : 0u;
    PersistArgs p{static_cast<uint32_t*>(ctx->dCounter.ptr), static_cast<uint2*>(ctx->dSpill.ptr), spillDepth, refill,
                  ms >= 5 ? 1u : 0u, m >= 9 ? kDeferPrimMin : 0u, kDeferNodeMin};
    e = launchTrace(a, width, any, ctx->ldsDepth, &p, grid, ctx->stream);
    if (e != hipSuccess) return failHip(ctx, e, "trace launch");
    return RBR_OK;
}

int rbr::prepareTraversal(rbr_ctx* ctx, uint32_t maxStack) {
    const unsigned grid = unsigned(std::max(1, ctx->cuCount) * ctx->wavesPerCu);
    const uint32_t spillDepth = std::max<uint32_t>(1u, maxStack > uint32_t(ctx->ldsDepth) ? maxStack - ctx->ldsDepth : 1u);
    if (int rc = ensureBuffer(ctx, ctx->dCounter, kCounterBytes)) return rc;
    return ensureBuffer(ctx, ctx->dSpill, size_t(grid) * spillDepth * 64 * sizeof(uint2));
}

// ---------------------------------------------------------------- no-allocation scratch
// Device memory of rbr_trace_closest_device_noalloc: the batch counters (first kCounterBytes,
// rounded to 256 B), then the global stack spill rows of the persistent waves.  One copy per
// GPU, held until the process ends.  In host code `scratch` is only a stand-in for the symbol:
// its device address comes from hipGetSymbolAddress, once per device (noAllocScratch).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr size_t kNoAllocCounterBytes = (kCounterBytes + 255) / 256 * 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex gScratchMu;
void* gScratchPtr[kMaxDevices] = {};

// Device address of `scratch` on the context's device (which must be current); looked up on
// the first call for a device and kept.
int noAllocScratch(rbr_ctx* ctx, void** out) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return fail(ctx, RBR_ERR_UNSUPPORTED, "device index too large");
    std::lock_guard<std::mutex> lock(gScratchMu);
    if (!gScratchPtr[ctx->device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return failHip(ctx, e == hipSuccess ? hipErrorInvalidSymbol : e, "hipGetSymbolAddress(scratch)");
        gScratchPtr[ctx->device] = p;
    }
    *out = gScratchPtr[ctx->device];
    return RBR_OK;
}
}  // namespace

// ---------------------------------------------------------------- context

extern "C" {

int rbr_ctx_create(int device, rbr_ctx** out) {
    if (!out) return RBR_ERR_INVALID;
    *out = nullptr;
    if (device == -1) {   // host-only context: BVH build and host traversal, no kernels
        auto* ctx = new (std::nothrow) rbr_ctx();
        if (!ctx) return RBR_ERR_NOMEM;
        ctx->device = -1;
        ctx->name = "host";
        *out = ctx;
        return RBR_OK;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return RBR_ERR_INVALID;
    auto* ctx = new (std::nothrow) rbr_ctx();
    if (!ctx) return RBR_ERR_NOMEM;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        delete ctx;
        return RBR_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        ctx->cuCount = prop.multiProcessorCount;
        ctx->name = prop.gcnArchName;
        ctx->name += " / ";
        ctx->name += prop.name;
    }
    // Set-up of the no-allocation entry point: the scratch array's address on this device;
    // the traversal watchdog's flag (zeroed here, so no launch path allocates or clears it).
    void* scr = nullptr;
    bool ok = !(device < kMaxDevices && noAllocScratch(ctx, &scr) != RBR_OK);
    ok = ok && ensureBuffer(ctx, ctx->dFault, 256) == RBR_OK &&
         hipMemset(ctx->dFault.ptr, 0, 256) == hipSuccess &&
         hipHostMalloc(reinterpret_cast<void**>(&ctx->hFault), sizeof(uint32_t), hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        freeBuffer(ctx->dFault);
        if (ctx->hFault) (void)hipHostFree(ctx->hFault);
        (void)hipEventDestroy(ctx->ev0);
        (void)hipEventDestroy(ctx->ev1);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return RBR_ERR_HIP;
    }
    *ctx->hFault = 0u;
    ctx->ownStream = ctx->stream;
    *out = ctx;
    return RBR_OK;
}

int rbr_ctx_set_stream(rbr_ctx* ctx, void* stream) {
    if (!ctx) return RBR_ERR_INVALID;
    if (ctx->device < 0) return fail(ctx, RBR_ERR_STATE, "a host-only context has no stream");
    if (int rc = useDevice(ctx)) return rc;
    // Work already queued finishes before later launches go to the other stream, so the
    // context's buffers (batch counters, spill rows, render queues) are never used by two
    // streams at once.
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return failHip(ctx, e, "hipStreamSynchronize");
    ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->ownStream;
    return RBR_OK;
}

void rbr_ctx_destroy(rbr_ctx* ctx) {
    if (!ctx) return;
    if (ctx->device < 0) { delete ctx; return; }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    destroyRenderState(ctx);
    destroyGuideState(ctx);
    freeBuffer(ctx->dCounter);
    freeBuffer(ctx->dSpill);
    freeBuffer(ctx->dFault);
    for (DevBuf& b : ctx->buildTmp) freeBuffer(b);
    if (ctx->hFault) (void)hipHostFree(ctx->hFault);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipStreamDestroy(ctx->ownStream);
    delete ctx;
}

const char* rbr_last_error(const rbr_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int rbr_device_info(rbr_ctx* ctx, char* name, size_t name_len, int* cu_count) {
    if (!ctx) return RBR_ERR_INVALID;
    if (name && name_len) {
        std::snprintf(name, name_len, "%s", ctx->name.c_str());
    }
    if (cu_count) *cu_count = ctx->cuCount;
    return RBR_OK;
}

int rbr_device_alloc(rbr_ctx* ctx, size_t bytes, void** dptr) {
    if (!ctx || !dptr) return RBR_ERR_INVALID;
    if (int rc = useDevice(ctx)) return rc;
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
    if (e != hipSuccess) { *dptr = nullptr; return failHip(ctx, e, "hipMalloc"); }
    return RBR_OK;
}

int rbr_device_free(rbr_ctx* ctx, void* dptr) {
    if (!ctx) return RBR_ERR_INVALID;
    if (!dptr) return RBR_OK;
    if (int rc = useDevice(ctx)) return rc;
    hipError_t e = hipFree(dptr);
    return e == hipSuccess ? RBR_OK : failHip(ctx, e, "hipFree");
}

int rbr_copy_to_device(rbr_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return RBR_ERR_INVALID;
    if (int rc = useDevice(ctx)) return rc;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? RBR_OK : failHip(ctx, e, "copy to device");
}

int rbr_copy_to_host(rbr_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return RBR_ERR_INVALID;
    if (int rc = useDevice(ctx)) return rc;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? RBR_OK : failHip(ctx, e, "copy to host");
}

int rbr_synchronize(rbr_ctx* ctx) {
    if (!ctx) return RBR_ERR_INVALID;
    if (int rc = useDevice(ctx)) return rc;
    if (int rc = queueFaultRead(ctx)) return rc;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return failHip(ctx, e, "hipStreamSynchronize");
    return checkFault(ctx);
}

int rbr_set_trace_step_cap(rbr_ctx* ctx, uint32_t cap) {
    if (!ctx) return RBR_ERR_INVALID;
    ctx->stepCap = cap;
    return RBR_OK;
}

int rbr_trace_resident_waves(rbr_scene* s, int any, int* waves_per_cu) {
    if (!s || !waves_per_cu) return RBR_ERR_INVALID;
    rbr_ctx* ctx = s->ctx;
    if (int rc = useDevice(ctx)) return rc;
    if (!s->built) return fail(ctx, RBR_ERR_STATE, "rbr_trace_resident_waves before rbr_scene_build_bvh");
    const int m = any && ctx->traceModeAny >= 0 ? ctx->traceModeAny : ctx->traceMode;
    const bool mixed = (s->geom.sphIds.size() + s->geom.quadIds.size()) > 0;
    int blocks = 0;
    hipError_t e = traceResidentWaves(s->bvh.code(), any != 0, mixed, ctx->ldsDepth, m, &blocks);
    if (e != hipSuccess) return failHip(ctx, e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    *waves_per_cu = m == 1 ? blocks : std::min(blocks, ctx->wavesPerCu);
    return RBR_OK;
}

__global__ __launch_bounds__(256) void calibCopyKernel(float4* __restrict__ dst, const float4* __restrict__ src,
                                                       size_t n) {
    for (size_t i = blockIdx.x * size_t(256) + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) dst[i] = src[i];
}

int rbr_calib_copy(rbr_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || !dst || !src || bytes % 16) return RBR_ERR_INVALID;
    if (int rc = useDevice(ctx)) return rc;
    const size_t n = bytes / 16;
    const unsigned grid = unsigned(std::min<size_t>((n + 255) / 256, size_t(ctx->cuCount) * 16));
    if (grid == 0) return RBR_OK;
    hipLaunchKernelGGL(calibCopyKernel, dim3(grid), dim3(256), 0, ctx->stream, static_cast<float4*>(dst),
                       static_cast<const float4*>(src), n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? RBR_OK : failHip(ctx, e, "calib copy launch");
}

int rbr_set_trace_tuning(rbr_ctx* ctx, int mode, int lds_depth, int waves_per_cu) {
    if (!ctx) return RBR_ERR_INVALID;
    const int closest = mode & 0xFF, anyCode = (mode >> 8) & 0xFF;
    if (mode < 0 || (mode >> 16) != 0 || closest > 12 || anyCode > 13)
        return fail(ctx, RBR_ERR_INVALID,
                    "mode must be 0 (persistent), 1 (lane), 2-4 (per-lane fetch), 5-8 (steps) or 9-12 (deferred steps), "
                    "optionally | (any-hit mode + 1) << 8");
    if (lds_depth != 8 && lds_depth != 10 && lds_depth != 12 && lds_depth != 16 && lds_depth != 24 && lds_depth != 32)
        return fail(ctx, RBR_ERR_INVALID, "lds_depth must be 8, 10, 12, 16, 24 or 32");
    if (waves_per_cu < 1 || waves_per_cu > 32) return fail(ctx, RBR_ERR_INVALID, "waves_per_cu must be in 1..32");
    ctx->traceMode = closest;
    ctx->traceModeAny = anyCode ? anyCode - 1 : -1;
    ctx->ldsDepth = lds_depth;
    ctx->wavesPerCu = waves_per_cu;
    return RBR_OK;
}

// ---------------------------------------------------------------- scenes

int rbr_scene_create(rbr_ctx* ctx, rbr_scene** out) {
    if (!ctx || !out) return RBR_ERR_INVALID;
    auto* s = new (std::nothrow) rbr_scene();
    if (!s) return fail(ctx, RBR_ERR_NOMEM, "scene allocation");
    s->ctx = ctx;
    *out = s;
    return RBR_OK;
}

void rbr_scene_destroy(rbr_scene* s) {
    if (!s) return;
    if (s->ctx->device < 0) { delete s; return; }
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    freeSceneDevice(s);
    delete s;
}

int rbr_scene_set_geometry(rbr_scene* s, const float* tris, uint32_t ntri, const uint32_t* tri_ids,
                           const float* spheres, uint32_t nsph, const uint32_t* sph_ids, const float* quads,
                           uint32_t nquad, const uint32_t* quad_ids, const float* planes, uint32_t nplane,
                           const uint32_t* plane_ids) {
    if (!s) return RBR_ERR_INVALID;
    rbr_ctx* ctx = s->ctx;
    if ((ntri && !tris) || (nsph && !spheres) || (nquad && !quads) || (nplane && !planes))
        return fail(ctx, RBR_ERR_INVALID, "null primitive array with nonzero count");
    const uint64_t total = uint64_t(ntri) + nsph + nquad + nplane;
    if (total >= (1ull << 27)) return fail(ctx, RBR_ERR_INVALID, "too many primitives (limit 2^27)");
    HostGeometry g;
    uint32_t next = 0;
    auto ids = [&](std::vector<uint32_t>& dst, const uint32_t* src, uint32_t n) {
        dst.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            dst[i] = src ? src[i] : next + i;
            if (dst[i] == RBR_MISS) return false;
            g.maxId = std::max(g.maxId, dst[i]);
            g.anyId = true;
        }
        next += n;
        return true;
    };
    g.tris.assign(tris, tris + size_t(ntri) * 9);
    g.sphs.assign(spheres, spheres + size_t(nsph) * 4);
    g.quads.assign(quads, quads + size_t(nquad) * 12);
    g.planes.assign(planes, planes + size_t(nplane) * 4);
    if (!ids(g.triIds, tri_ids, ntri) || !ids(g.sphIds, sph_ids, nsph) || !ids(g.quadIds, quad_ids, nquad) ||
        !ids(g.planeIds, plane_ids, nplane))
        return fail(ctx, RBR_ERR_INVALID, "primitive id 0xFFFFFFFF is reserved for a miss");
    s->geom = std::move(g);
    s->triNormals.clear();
    s->built = false;
    s->bvh = WideBVH();
    bumpVersion(s);
    return RBR_OK;
}

int rbr_scene_set_tri_normals(rbr_scene* s, const float* normals, size_t ntri) {
    if (!s) return RBR_ERR_INVALID;
    if (ntri == 0 || !normals) {
        s->triNormals.clear();
        bumpVersion(s);
        return RBR_OK;
    }
    if (ntri != s->geom.tris.size() / 9)
        return fail(s->ctx, RBR_ERR_INVALID, "vertex normals must cover every triangle of the geometry");
    s->triNormals.resize(ntri * 9);
    for (size_t k = 0; k < ntri * 3; ++k) {
        const V n = vnormalized(V{normals[3 * k], normals[3 * k + 1], normals[3 * k + 2]});
        s->triNormals[3 * k] = n.x;
        s->triNormals[3 * k + 1] = n.y;
        s->triNormals[3 * k + 2] = n.z;
    }
    bumpVersion(s);
    return RBR_OK;
}

int rbr_triangle_shading_normal(const float* verts, const float* normals, const float* o, const float* d, float* out) {
    if (!verts || !normals || !o || !d || !out) return RBR_ERR_INVALID;
    const V v0{verts[0], verts[1], verts[2]};
    const V e1 = vsub(V{verts[3], verts[4], verts[5]}, v0), e2 = vsub(V{verts[6], verts[7], verts[8]}, v0);
    const V geo = vnormalized(vcross(e1, e2));
    const V n = smoothNormal(V{o[0], o[1], o[2]}, V{d[0], d[1], d[2]}, v0, e1, e2,
                             vnormalized(V{normals[0], normals[1], normals[2]}),
                             vnormalized(V{normals[3], normals[4], normals[5]}),
                             vnormalized(V{normals[6], normals[7], normals[8]}), geo);
    out[0] = n.x;
    out[1] = n.y;
    out[2] = n.z;
    return RBR_OK;
}

int rbr_scene_build_bvh(rbr_scene* s, int width, int leaf_max) {
    if (!s) return RBR_ERR_INVALID;
    rbr_ctx* ctx = s->ctx;
    const int W = width & 0xFF;
    if ((W != 4 && W != 8) || (width & ~(RBR_BVH_QUANTIZED | RBR_BVH_GREEDY | 0xFF)))
        return fail(ctx, RBR_ERR_INVALID, "BVH width must be 4 or 8, optionally | RBR_BVH_QUANTIZED | RBR_BVH_GREEDY");
    if (leaf_max < 1 || leaf_max > 8) return fail(ctx, RBR_ERR_INVALID, "leaf_max must be in 1..8");
    s->built = false;
    bumpVersion(s);
    s->hostStale = false;
    freeBuffer(s->dFloatNodes);
    buildWideBVH(s->geom, W, leaf_max, s->bvh, (width & RBR_BVH_GREEDY) == 0);
    if (width & kQuantFlag) quantizeWideBVH(s->bvh);
    if (!s->bvh.empty() && uint64_t(W - 1) * s->bvh.maxDepth + 1 > uint64_t(kMaxStack))
        return fail(ctx, RBR_ERR_UNSUPPORTED, "BVH too deep for the traversal stack");
    if (ctx->device < 0) { s->built = true; return RBR_OK; }
    if (int rc = useDevice(ctx)) return rc;
    int rc;
    if ((rc = uploadBuffer(ctx, s->dNodes, s->bvh.deviceNodes(), s->bvh.deviceNodeBytes()))) return rc;
    if ((rc = uploadBuffer(ctx, s->dPrims, s->bvh.prims.data(), s->bvh.prims.size() * sizeof(F4)))) return rc;
    if ((rc = uploadBuffer(ctx, s->dQuadN, s->bvh.quadNormals.data(), s->bvh.quadNormals.size() * sizeof(F4)))) return rc;
    if ((rc = uploadBuffer(ctx, s->dPlanes, s->geom.planes.data(), s->geom.planes.size() * sizeof(float)))) return rc;
    if ((rc = uploadBuffer(ctx, s->dPlaneIds, s->geom.planeIds.data(), s->geom.planeIds.size() * sizeof(uint32_t)))) return rc;
    s->built = true;
    return RBR_OK;
}

int rbr_scene_build_bvh_device(rbr_scene* s, int width, int leaf_max) {
    if (!s) return RBR_ERR_INVALID;
    rbr_ctx* ctx = s->ctx;
    const int W = width & 0xFF;
    if ((W != 4 && W != 8) || (width & ~(RBR_BVH_QUANTIZED | RBR_BVH_GREEDY | RBR_BVH_PLOC | 0xFF)))
        return fail(ctx, RBR_ERR_INVALID,
                    "BVH width must be 4 or 8, optionally | RBR_BVH_QUANTIZED | RBR_BVH_GREEDY | RBR_BVH_PLOC");
    if (leaf_max < 1 || leaf_max > 8) return fail(ctx, RBR_ERR_INVALID, "leaf_max must be in 1..8");
    if (ctx->device < 0) return fail(ctx, RBR_ERR_STATE, "the device build needs a GPU context");
    if (int rc = useDevice(ctx)) return rc;
    s->built = false;
    bumpVersion(s);
    s->hostStale = false;
    if (int rc = buildWideBVHDevice(ctx, s, W, leaf_max, (width & kQuantFlag) != 0, (width & RBR_BVH_GREEDY) == 0,
                                    (width & RBR_BVH_PLOC) != 0))
        return rc;
    s->hostStale = true;
    if (!s->bvh.empty() && uint64_t(W - 1) * s->bvh.maxDepth + 1 > uint64_t(kMaxStack))
        return fail(ctx, RBR_ERR_UNSUPPORTED, "BVH too deep for the traversal stack");
    int rc;
    if ((rc = uploadBuffer(ctx, s->dQuadN, s->bvh.quadNormals.data(), s->bvh.quadNormals.size() * sizeof(F4)))) return rc;
    if ((rc = uploadBuffer(ctx, s->dPlanes, s->geom.planes.data(), s->geom.planes.size() * sizeof(float)))) return rc;
    if ((rc = uploadBuffer(ctx, s->dPlaneIds, s->geom.planeIds.data(), s->geom.planeIds.size() * sizeof(uint32_t)))) return rc;
    s->built = true;
    return RBR_OK;
}

int rbr_scene_device_nodes(const rbr_scene* s, void* dst, size_t capacity, size_t* bytes) {
    if (!s) return RBR_ERR_INVALID;
    if (!s->built) return fail(s->ctx, RBR_ERR_STATE, "BVH not built");
    const size_t n = s->bvh.nodeBytes();
    if (bytes) *bytes = n;
    if (dst && capacity) {
        if (int rc = ensureHostBvh(s)) return rc;
        std::memcpy(dst, s->bvh.deviceNodes(), std::min(capacity, n));
    }
    return RBR_OK;
}

int rbr_scene_bvh_arrays(const rbr_scene* s, void* nodes, size_t node_capacity, void* prims, size_t prim_capacity,
                         size_t* node_bytes, size_t* prim_bytes) {
    if (!s) return RBR_ERR_INVALID;
    if (!s->built) return fail(s->ctx, RBR_ERR_STATE, "BVH not built");
    const size_t nb = size_t(s->bvh.nodeCount) * size_t(s->bvh.f4PerNode()) * sizeof(F4), pb = s->bvh.primBytes();
    if (node_bytes) *node_bytes = nb;
    if (prim_bytes) *prim_bytes = pb;
    if ((nodes && node_capacity) || (prims && prim_capacity))
        if (int rc = ensureHostBvh(s)) return rc;
    if (nodes && node_capacity) std::memcpy(nodes, s->bvh.nodes.data(), std::min(node_capacity, nb));
    if (prims && prim_capacity) std::memcpy(prims, s->bvh.prims.data(), std::min(prim_capacity, pb));
    return RBR_OK;
}

int rbr_scene_bvh_stats(const rbr_scene* s, rbr_bvh_stats* out) {
    if (!s || !out) return RBR_ERR_INVALID;
    if (!s->built) return fail(s->ctx, RBR_ERR_STATE, "BVH not built");
    out->width = uint32_t(s->bvh.code());
    out->node_count = s->bvh.nodeCount;
    out->leaf_count = s->bvh.leafCount;
    out->prim_count = s->bvh.primCount;
    out->unbounded = uint32_t(s->geom.planeIds.size());
    out->max_depth = s->bvh.maxDepth;
    out->node_bytes = uint64_t(s->bvh.nodeBytes());
    out->prim_bytes = uint64_t(s->bvh.primBytes());
    out->sah_cost = s->bvh.sahCost;
    out->build_ms = s->bvh.buildMs;
    return RBR_OK;
}

int rbr_scene_world_bounds(const rbr_scene* s, float lo[3], float hi[3]) {
    if (!s || !lo || !hi) return RBR_ERR_INVALID;
    for (int k = 0; k < 3; ++k) {
        lo[k] = s->built && !s->bvh.empty() ? s->bvh.worldLo[k] : 3.40282347e38f;
        hi[k] = s->built && !s->bvh.empty() ? s->bvh.worldHi[k] : -3.40282347e38f;
    }
    return RBR_OK;
}

// ---------------------------------------------------------------- traversal

static int traceDevice(rbr_scene* s, const rbr_ray* dRays, size_t n, void* dOut, bool any) {
    rbr_ctx* ctx = s->ctx;
    if (!s->built) return fail(ctx, RBR_ERR_STATE, "trace before rbr_scene_build_bvh");
    if (n > 0x80000000ull) return fail(ctx, RBR_ERR_INVALID, "ray batch too large (limit 2^31)");
    if (n && (!dRays || !dOut)) return fail(ctx, RBR_ERR_INVALID, "null ray or result pointer");
    TraceArgs a = traceArgs(s);
    a.rays = reinterpret_cast<const float4*>(dRays);
    a.n = uint32_t(n);
    a.hits = any ? nullptr : static_cast<uint2*>(dOut);
    a.occ = any ? static_cast<uint8_t*>(dOut) : nullptr;
    const uint32_t maxStack = s->bvh.empty() ? 1u : uint32_t(s->bvh.width - 1) * s->bvh.maxDepth + 1u;
    return launchTraversal(ctx, a, s->bvh.code(), maxStack, any);
}

static int traceHost(rbr_scene* s, const rbr_ray* rays, size_t n, void* out, bool any) {
    if (!s) return RBR_ERR_INVALID;
    rbr_ctx* ctx = s->ctx;
    if (n == 0) return RBR_OK;
    if (!rays || !out) return fail(ctx, RBR_ERR_INVALID, "null ray or result array");
    if (int rc = useDevice(ctx)) return rc;
    const size_t outBytes = n * (any ? sizeof(uint8_t) : sizeof(rbr_hit));
    int rc;
    if ((rc = ensureBuffer(ctx, s->dRays, n * sizeof(rbr_ray)))) return rc;
    if ((rc = ensureBuffer(ctx, s->dOut, outBytes))) return rc;
    hipError_t e = hipMemcpyAsync(s->dRays.ptr, rays, n * sizeof(rbr_ray), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return failHip(ctx, e, "ray upload");
    if ((rc = traceDevice(s, static_cast<const rbr_ray*>(s->dRays.ptr), n, s->dOut.ptr, any))) return rc;
    e = hipMemcpyAsync(out, s->dOut.ptr, outBytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e != hipSuccess) return failHip(ctx, e, "result download");
    if ((rc = queueFaultRead(ctx))) return rc;
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return failHip(ctx, e, "trace");
    return checkFault(ctx);
}

int rbr_trace_closest(rbr_scene* s, const rbr_ray* rays, size_t n, rbr_hit* hits) {
    return traceHost(s, rays, n, hits, false);
}

int rbr_trace_any(rbr_scene* s, const rbr_ray* rays, size_t n, uint8_t* occluded) {
    return traceHost(s, rays, n, occluded, true);
}

int rbr_trace_closest_device(rbr_scene* s, const rbr_ray* d_rays, size_t n, rbr_hit* d_hits) {
    if (!s) return RBR_ERR_INVALID;
    if (int rc = useDevice(s->ctx)) return rc;
    return traceDevice(s, d_rays, n, d_hits, false);
}

int rbr_trace_any_device(rbr_scene* s, const rbr_ray* d_rays, size_t n, uint8_t* d_occluded) {
    if (!s) return RBR_ERR_INVALID;
    if (int rc = useDevice(s->ctx)) return rc;
    return traceDevice(s, d_rays, n, d_occluded, true);
}

int rbr_trace_closest_device_noalloc(rbr_scene* s, const rbr_ray* d_rays, size_t n, rbr_hit* d_hits) {
    if (!s) return RBR_ERR_INVALID;
    rbr_ctx* ctx = s->ctx;
    if (int rc = useDevice(ctx)) return rc;
    if (!s->built) return fail(ctx, RBR_ERR_STATE, "trace before rbr_scene_build_bvh");
    if (n > 0x80000000ull) return fail(ctx, RBR_ERR_INVALID, "ray batch too large (limit 2^31)");
    if (n && (!d_rays || !d_hits)) return fail(ctx, RBR_ERR_INVALID, "null ray or result pointer");
    if (n == 0) return RBR_OK;
    void* base = nullptr;
    if (int rc = noAllocScratch(ctx, &base)) return rc;
    TraceArgs a = traceArgs(s);
    armWatchdog(ctx, a);
    a.rays = reinterpret_cast<const float4*>(d_rays);
    a.n = uint32_t(n);
    a.hits = reinterpret_cast<uint2*>(d_hits);
    const uint32_t maxStack = s->bvh.empty() ? 1u : uint32_t(s->bvh.width - 1) * s->bvh.maxDepth + 1u;
    const int m = ctx->traceMode;
    hipError_t e;
    if (m == 1) {   // one ray per lane: private stack, no scratch
        if (maxStack > uint32_t(kMaxStack)) return fail(ctx, RBR_ERR_UNSUPPORTED, "BVH too deep for the lane stack");
        e = launchTrace(a, s->bvh.code(), false, ctx->ldsDepth, nullptr, 0, ctx->stream);
        return e == hipSuccess ? RBR_OK : failHip(ctx, e, "trace launch");
    }
    // As many persistent waves as the tuning asks for and the spill rows of the scratch hold.
    const uint32_t spillDepth = std::max<uint32_t>(1u, maxStack > uint32_t(ctx->ldsDepth) ? maxStack - ctx->ldsDepth : 1u);
    const size_t perWave = size_t(spillDepth) * 64 * sizeof(uint2);
    const size_t fit = (kNoAllocScratchBytes - kNoAllocCounterBytes) / perWave;
    const size_t cus = size_t(std::max(1, ctx->cuCount));
    const size_t grid = std::min(cus * size_t(ctx->wavesPerCu), fit);
    if (grid < cus)
        return fail(ctx, RBR_ERR_UNSUPPORTED,
                    "BVH too deep for the no-allocation scratch: " + std::to_string(perWave) + " B of stack per wave, " +
                        std::to_string(kNoAllocScratchBytes) + " B in all, fewer than one wave per compute unit fits");
    const int ms = m >= 9 ? m - 4 : m;
    const uint32_t refill = ms >= 5 ? (2u << (ms - 4)) : ms >= 2 ? (4u << (ms - 1)) : 0u;
    unsigned char* b = static_cast<unsigned char*>(base);
    PersistArgs p{reinterpret_cast<uint32_t*>(b), reinterpret_cast<uint2*>(b + kNoAllocCounterBytes), spillDepth, refill,
                  ms >= 5 ? 1u : 0u, m >= 9 ? kDeferPrimMin : 0u, kDeferNodeMin};
    e = launchTrace(a, s->bvh.code(), false, ctx->ldsDepth, &p, unsigned(grid), ctx->stream);
    return e == hipSuccess ? RBR_OK : failHip(ctx, e, "trace launch");
}

int rbr_host_trace_closest(const rbr_scene* s, const rbr_ray* rays, size_t n, rbr_hit* hits, uint64_t* nodeVisits,
                           uint64_t* primTests) {
    if (!s) return RBR_ERR_INVALID;
    if (!s->built) return fail(s->ctx, RBR_ERR_STATE, "trace before rbr_scene_build_bvh");
    if (n && (!rays || !hits)) return fail(s->ctx, RBR_ERR_INVALID, "null ray or result array");
    if (int rc = ensureHostBvh(s)) return rc;
    TraceCounters total;
    for (size_t i = 0; i < n; ++i) {
        TraceCounters c;
        float t;
        uint32_t id = hostClosest(s->bvh, s->geom, rays[i].o, rays[i].d, rays[i].tmin, rays[i].tmax, &t, &c);
        hits[i].id = id;
        hits[i].t = id == RBR_MISS ? 0.0f : t;
        total.nodes += c.nodes;
        total.prims += c.prims;
    }
    if (nodeVisits) *nodeVisits = total.nodes;
    if (primTests) *primTests = total.prims;
    return RBR_OK;
}

int rbr_host_trace_any(const rbr_scene* s, const rbr_ray* rays, size_t n, uint8_t* occ, uint64_t* nodeVisits,
                       uint64_t* primTests) {
    if (!s) return RBR_ERR_INVALID;
    if (!s->built) return fail(s->ctx, RBR_ERR_STATE, "trace before rbr_scene_build_bvh");
    if (n && (!rays || !occ)) return fail(s->ctx, RBR_ERR_INVALID, "null ray or result array");
    if (int rc = ensureHostBvh(s)) return rc;
    TraceCounters total;
    for (size_t i = 0; i < n; ++i) {
        TraceCounters c;
        occ[i] = hostAny(s->bvh, s->geom, rays[i].o, rays[i].d, rays[i].tmin, rays[i].tmax, &c) ? 1 : 0;
        total.nodes += c.nodes;
        total.prims += c.prims;
    }
    if (nodeVisits) *nodeVisits = total.nodes;
    if (primTests) *primTests = total.prims;
    return RBR_OK;
}

int rbr_time_trace(rbr_scene* s, int kind, const rbr_ray* d_rays, size_t n, void* d_out, int iters, float* mean_ms) {
    if (!s || !mean_ms || iters < 1 || (kind != 0 && kind != 1)) return RBR_ERR_INVALID;
    rbr_ctx* ctx = s->ctx;
    if (int rc = useDevice(ctx)) return rc;
    hipError_t e = hipEventRecord(ctx->ev0, ctx->stream);
    if (e != hipSuccess) return failHip(ctx, e, "hipEventRecord");
    for (int it = 0; it < iters; ++it) {
        if (int rc = traceDevice(s, d_rays, n, d_out, kind == 1)) return rc;
    }
    e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess && queueFaultRead(ctx) != RBR_OK) return RBR_ERR_HIP;
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return failHip(ctx, e, "timing");
    float ms = 0.0f;
    e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    if (e != hipSuccess) return failHip(ctx, e, "hipEventElapsedTime");
    *mean_ms = ms / float(iters);
    return checkFault(ctx);
}

}  // extern "C"
