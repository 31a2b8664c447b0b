// Launch interface of the traversal kernels (trace.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rbr {

constexpr uint32_t RBR_MISS_ID = 0xFFFFFFFFu;
constexpr uint32_t kEmptyD = 0xFFFFFFFFu;
constexpr int kLdsDepthDefault = 8;    // LDS stack entries per lane (8 B each)
constexpr int kWavesPerCuDefault = 32;  // persistent waves per compute unit
constexpr int kMaxStack = 192;    // total stack entries per lane (LDS + private spill)

__host__ __device__ __forceinline__ bool isLeafD(uint32_t c) { return (c & 0x80000000u) != 0u; }
__host__ __device__ __forceinline__ uint32_t leafCountD(uint32_t c) { return ((c >> 27) & 7u) + 1u; }
__host__ __device__ __forceinline__ uint32_t leafStartD(uint32_t c) { return c & ((1u << 27) - 1u); }

struct TraceArgs {
    const float4* nodes;      // null when no bounded primitive exists
    const float4* prims;
    const float4* quadN;
    const float4* planes;
    const uint32_t* planeIds;
    uint32_t nplanes;
    uint32_t mixed;           // any non-triangle bounded primitive
    const float4* rays;       // 2 float4 per ray
    uint32_t n;               // ray count (ignored when nDev is set)
    const uint32_t* nDev;     // optional: ray count read from device memory at run time
    const uint32_t* rootOf;   // optional: per-ray root node (several BVHs in one node arena)
    uint2* hits;              // closest: (id, t bits)
    uint8_t* occ;             // any
    // Watchdog: a ray gives up after stepCap node visits plus primitive tests (a correct tree
    // needs at most its node and primitive counts) and sets *fault to 1.
    uint32_t stepCap;
    uint32_t* fault;
};

// Persistent-wave launch state: a batch counter (zeroed by a one-lane kernel launched just
// before the traversal on the same stream) and the per-wave spill rows of the stack.
// Batch counters of the persistent kernels: kSlices counters kSliceStride words apart.
constexpr uint32_t kSlices = 16;
constexpr uint32_t kSliceStride = 64;
constexpr size_t kCounterBytes = size_t(kSlices) * kSliceStride * sizeof(uint32_t);

// Step kernel with deferred primitive steps (trace modes 9-12): primitives are tested once this
// many lanes of a wave hold a leaf, or once fewer than kDeferNodeMin lanes have a node to test.
constexpr uint32_t kDeferPrimMin = 32;
constexpr uint32_t kDeferNodeMin = 48;

struct PersistArgs {
    uint32_t* counter;        // kSlices batch counters
    uint2* spill;             // [grid][spillDepth][64]
    uint32_t spillDepth;
    uint32_t refill;          // 0: waves take 64-ray batches; else per-lane fetch once this many lanes idle
    uint32_t step;            // with refill: single node / primitive steps instead of while-while rounds
    uint32_t primMin;         // step kernel: 0 tests primitives every iteration; else only once this many
    uint32_t nodeMin;         //   lanes hold a leaf or fewer than nodeMin lanes have an interior node
};

// ldsDepth in {8, 10, 12, 16, 24, 32}; p == nullptr selects the one-ray-per-lane kernel (which
// needs a.nDev == nullptr).
hipError_t launchTrace(const TraceArgs& a, int width, bool any, int ldsDepth, const PersistArgs* p, unsigned grid,
                       hipStream_t s);

// Waves of the kernel that launchTrace would run for (width, any, mixed, ldsDepth, tuning mode)
// that one compute unit holds at once (its registers and LDS; one wave per workgroup).
hipError_t traceResidentWaves(int width, bool any, bool mixed, int ldsDepth, int mode, int* wavesPerCu);

}  // namespace rbr
