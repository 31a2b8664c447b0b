// Wide-BVH closest-hit / any-hit traversal kernels for gfx950 (BVH::intersect / intersectP,
// bvh.cpp:33-41, 211-267): persistent waves over device-resident ray batches.  The per-ray
// traversal itself (node tests, stack, primitive tests) is in trace_dev.h.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bvh.h"
#include "prim_tests.h"
#include "trace.h"
#include "trace_dev.h"

#pragma clang fp contract(off)

namespace rbr {

namespace {

using namespace tdev;

// Result staging of the step kernel: 64-ray ranges held in LDS at once per wave (512 B each).
// With 6 slots 79 % of the closest-hit result lines reach HBM whole (WRITE_SIZE 1.28x the
// result bytes on the 2^22-ray batch; 1.68x with 4 slots, 4.5x without staging) at the same
// rate as 4 slots.
constexpr uint32_t kStageSlots = 6;

// Work distribution of the persistent kernels: the 64-ray batches are split into kSlices
// equal slices, each with its own counter (kSliceStride words apart), and a wave drains the
// slice of its workgroup index, then moves on to the following ones until every slice is
// drained.  Atomics on one address are serialized (about 11 ns apart on MI355X, so one
// counter caps a launch at ~90 M batch grabs/s, and that was the limit of the traversal of
// 2^22 rays); over 16 counters the grabs no longer queue.  The counters are zeroed by
// resetCounterKernel right before the launch on the same stream.
struct SliceWalk {
    uint32_t s;      // current slice
    uint32_t left;   // slices not yet found drained
    uint32_t per;    // rays per slice (a multiple of 64)
};

__device__ __forceinline__ SliceWalk sliceWalk(uint32_t n) {
    const uint32_t batches = (n + 63u) / 64u;
    return SliceWalk{blockIdx.x % kSlices, kSlices, ((batches + kSlices - 1u) / kSlices) * 64u};
}

// First ray of the next 64-ray batch, or n once every slice is drained.  Called by the whole
// wave with every lane active (the result is uniform); lane 0 does the atomic.
__device__ __forceinline__ uint32_t nextBatch(uint32_t* counters, uint32_t n, SliceWalk& w) {
    while (w.left > 0u) {
        uint32_t b = 0u;
        if (__lane_id() == 0) b = atomicAdd(counters + size_t(w.s) * kSliceStride, 64u);
        b = __builtin_amdgcn_readfirstlane(b);
        const uint32_t start = w.s * w.per + b;
        if (b < w.per && start < n) return start;
        --w.left;
        w.s = w.s + 1u == kSlices ? 0u : w.s + 1u;
    }
    return n;
}

// Persistent waves: each 64-lane workgroup takes batches of 64 consecutive rays (nextBatch)
// until every batch is taken, so slow rays do not idle a whole launch slot and every
// resident wave stays busy to the end.  The ray count can come from device memory (queue
// lengths written by an earlier kernel).
template <int W, bool Q, int LDS, bool ANY, bool MIXED>
__global__ __launch_bounds__(64) void traceKernelPersistent(TraceArgs a, PersistArgs p) {
    __shared__ uint2 ldsStack[LDS * 64];
    GlobalSpillStack<LDS> st;
    st.lds = (LdsU64*)ldsStack;
    st.spill = (GlobalU64*)(p.spill + size_t(blockIdx.x) * size_t(p.spillDepth) * 64);
    const uint32_t n = a.nDev ? *a.nDev : a.n;
    SliceWalk walk = sliceWalk(n);
    while (true) {
        const uint32_t base = nextBatch(p.counter, n, walk);
        if (base >= n) break;
        const uint32_t i = base + threadIdx.x;
        if (i < n) {
            const DRay r = loadRay(a.rays, i);
            float bestT;
            uint32_t bestId;
            bool occ;
            st.sp = 0;
            traceRay<W, Q, ANY, MIXED>(a, r, a.rootOf ? a.rootOf[i] : 0u, st, bestT, bestId, occ);
            storeResult<ANY>(a, i, bestT, bestId, occ);
        }
    }
}

// Persistent waves with per-lane ray fetch: a lane whose ray is done goes idle, and once at
// least p.refill lanes of the wave are idle the wave hands its idle lanes new rays, so a
// wave no longer waits for its slowest ray before taking work (incoherent rays vary
// several-fold in node count, and half of them may miss everything).  Rays come from a
// 64-ray range the wave reserves with nextBatch (the batch kernel's slices, one atomic per
// range); when the range runs dry mid-refill the rest comes from the next range.  Between
// refills each live lane runs one traversal round.  Every wave exits once every slice is
// drained and all of its lanes are idle.
template <int W, bool Q, int LDS, bool ANY, bool MIXED>
__global__ __launch_bounds__(64) void traceKernelDynamic(TraceArgs a, PersistArgs p) {
    __shared__ uint2 ldsStack[LDS * 64];
    GlobalSpillStack<LDS> st;
    st.lds = (LdsU64*)ldsStack;
    st.spill = (GlobalU64*)(p.spill + size_t(blockIdx.x) * size_t(p.spillDepth) * 64);
    const uint32_t n = a.nDev ? *a.nDev : a.n;
    const uint32_t lane = threadIdx.x;
    SliceWalk walk = sliceWalk(n);
    bool alive = false, more = true, occ = false;
    uint32_t i = 0, ref = kEmptyD, leaf = kEmptyD, bestId = RBR_MISS_ID, steps = 0u;
    uint32_t qBase = 0u, qEnd = 0u;   // reserved, not yet assigned rays (uniform)
    float bestT = 0.0f;
    DRay r{};
    while (true) {
        const uint64_t idle = __ballot(alive ? 0 : 1);
        const uint32_t nIdle = uint32_t(__popcll(idle));
        if (more && nIdle >= p.refill) {
            const uint32_t avail = qEnd - qBase;
            uint32_t base2 = 0u;
            if (avail < nIdle) base2 = nextBatch(p.counter, n, walk);
            if (!alive) {
                const uint32_t rank = uint32_t(__popcll(idle & ((1ull << lane) - 1ull)));
                const uint32_t idx = rank < avail ? qBase + rank : base2 + (rank - avail);
                if (idx < n) {
                    i = idx;
                    r = loadRay(a.rays, idx);
                    bestT = r.tmax;
                    bestId = RBR_MISS_ID;
                    occ = false;
                    st.sp = 0;
                    ref = a.nodes ? (a.rootOf ? a.rootOf[idx] : 0u) : kEmptyD;
                    leaf = kEmptyD;
                    steps = 0u;
                    alive = true;
                }
            }
            if (avail < nIdle) {
                qBase = base2 + (nIdle - avail);
                qEnd = base2 + 64u;
                if (base2 >= n) more = false;   // every slice drained
            } else {
                qBase += nIdle;
            }
        }
        if (!__any(alive ? 1 : 0)) {
            if (!more) break;
            continue;
        }
        if (alive) {
            if (ref != kEmptyD || leaf != kEmptyD)
                traceRound<W, Q, ANY, MIXED>(a, r, st, ref, leaf, bestT, bestId, occ, steps);
            if ((ANY && occ) || (ref == kEmptyD && leaf == kEmptyD)) {
                tracePlanes<ANY>(a, r, bestT, bestId, occ);
                storeResult<ANY>(a, i, bestT, bestId, occ);
                alive = false;
            }
        }
    }
}

// One step's fetches and tests: the primitive of the leaf in progress (lanes with primNow)
// and one interior node (lanes with nodeNow).  Both fetches are issued before either test
// waits for its data, so one memory latency covers the primitive and the node.  Only the lanes
// that test something fetch (the loads run under the lane mask): the address and data units
// then process those lanes alone, which measured 2-8 % faster than every lane fetching
// primitive 0 and node 0 in its place.
template <int W, bool Q, bool ANY, bool MIXED, typename Stack>
__device__ __forceinline__ void fetchTest(const TraceArgs& a, const DRay& r, Stack& st, bool primNow, bool nodeNow,
                                          uint32_t& prim, uint32_t& primLeft, uint32_t& ref, float& bestT,
                                          uint32_t& bestId, bool& occ, uint32_t& steps) {
    float4 p0, p1, p2;
    if (primNow) {
        const float4* rec = a.prims + size_t(prim) * 3;
        p0 = rec[0];
        p1 = rec[1];
        p2 = rec[2];
    }
    NodeRows<W, Q> nr;
    if (nodeNow) loadNode<W, Q>(a.nodes, ref, nr);
    if (primNow) {
        float t;
        uint32_t id;
        if (testPrimLoaded<MIXED>(a, p0, p1, p2, r, bestT, t, id)) {
            if (ANY) occ = true;
            else if (t < bestT || id < bestId) { bestT = t; bestId = id; }
        }
        ++prim;
        --primLeft;
        if (ANY && occ) {
            primLeft = 0u;
            ref = kEmptyD;
            st.sp = 0;
        }
    }
    if (nodeNow && !(ANY && occ)) {
        ++steps;   // node visits of the ray (watchdog)
        float tc[W];
        uint32_t rc[W];
        bool hit[W];
        children<W, Q>(nr, r, bestT, tc, rc, hit);
        ref = descend<W, ANY>(tc, rc, hit, st);
    }
}

// Settle: a leaf in ref starts when no leaf is in progress; an empty ref is refilled from the
// stack (entries farther than the closest hit are dropped).  One loop serves both, so a wave
// pays for the longest chain of pops of one lane, once per iteration.
template <bool ANY, typename Stack>
__device__ __forceinline__ void settle(Stack& st, bool alive, uint32_t& ref, uint32_t& prim, uint32_t& primLeft,
                                       float bestT) {
    if (!alive) return;
    while (true) {
        if (primLeft == 0u && ref != kEmptyD && isLeafD(ref)) {
            prim = leafStartD(ref);
            primLeft = leafCountD(ref);
            ref = kEmptyD;
        }
        if (ref != kEmptyD || st.sp == 0) break;
        const uint2 e = st.pop();
        if (ANY || u2f(e.y) <= bestT) ref = e.x;
    }
}

// Persistent waves in single steps with per-lane ray fetch.  Each iteration of a wave, every
// lane with an interior node to visit tests that node's children, and every lane with a leaf
// in progress tests one primitive of it, so no lane waits for the others to reach a leaf (the
// while-while rounds above make a lane wait until every lane holds one) and a lane whose ray
// is done is handed a new ray as soon as p.refill lanes are idle.  The fetches of the two steps
// are issued together (fetchTest).
// A lane holds at most one leaf in progress; a leaf met while one is in progress stays in ref
// until the current one is finished.  With DEFER the primitive step is skipped in iterations
// where fewer than p.primMin lanes hold a leaf and at least p.nodeMin lanes have an interior
// node, so the primitive tests run with more lanes at once; a ray waits for its leaves, never
// the other way round, so every wave still drains.  Every entry that can still hold a
// closer hit is visited and every primitive of every visited leaf is tested, so the hits are
// those of the other kernels.  Rays come from 64-ray ranges reserved with nextBatch, as in
// traceKernelDynamic, and every wave exits once every slice is drained and all of its lanes
// are idle.
template <int W, bool Q, int LDS, bool ANY, bool MIXED, bool DEFER>
__global__ __launch_bounds__(64) void traceKernelStep(TraceArgs a, PersistArgs p) {
    __shared__ uint2 ldsStack[LDS * 64];
    GlobalSpillStack<LDS> st;
    st.lds = (LdsU64*)ldsStack;
    st.spill = (GlobalU64*)(p.spill + size_t(blockIdx.x) * size_t(p.spillDepth) * 64);
    const uint32_t n = a.nDev ? *a.nDev : a.n;
    const uint32_t lane = threadIdx.x;
    SliceWalk walk = sliceWalk(n);
    if (!a.nodes) {   // no BVH: only the unbounded primitives, one ray per lane
        for (uint32_t base = nextBatch(p.counter, n, walk); base < n; base = nextBatch(p.counter, n, walk)) {
            const uint32_t idx = base + lane;
            if (idx >= n) continue;
            const DRay rr = loadRay(a.rays, idx);
            float bt = rr.tmax;
            uint32_t bid = RBR_MISS_ID;
            bool oc = false;
            tracePlanes<ANY>(a, rr, bt, bid, oc);
            storeResult<ANY>(a, idx, bt, bid, oc);
        }
        return;
    }
    bool alive = false, more = true, occ = false;
    uint32_t i = 0, ref = kEmptyD, prim = 0u, primLeft = 0u, bestId = RBR_MISS_ID, steps = 0u, iter = 0u;
    uint32_t qBase = 0u, qEnd = 0u;   // reserved, not yet assigned rays (uniform)
    float bestT = 0.0f;
    DRay r{};
    // Results staged in LDS per 64-ray range (kStageSlots ranges at a time) and written out
    // together once the range's last ray is done: the lanes of a wave finish rays of a range
    // at scattered times, and 8-byte (or 1-byte) writes spread over that time left the result
    // lines partly written in L2 and written back several times (4.5x the result bytes for
    // closest hit, 20x for any hit).  stageLeft[s] counts the rays of slot s not yet done; 0
    // means free.  A range that finds no free slot writes its results directly, so no lane
    // ever waits for a slot.
    __shared__ uint2 stage[kStageSlots * 64];
    __shared__ uint32_t stageLeft[kStageSlots];
    if (lane < kStageSlots) stageLeft[lane] = 0u;
    uint32_t slot = 0u, curSlot = 0u;   // slot of the lane's ray, of the current range
    while (true) {
        const uint64_t idle = __ballot(alive ? 0 : 1);
        const uint32_t nIdle = uint32_t(__popcll(idle));
        if (more && nIdle >= p.refill) {
            const uint32_t avail = qEnd - qBase;
            uint32_t base2 = 0u;
            // A new range takes the first free slot (the current range's is busy until its last
            // ray is done), or none: its rays then write their results directly.
            uint32_t newSlot = kStageSlots;
            if (avail < nIdle) {
                for (uint32_t s2 = 0; s2 < kStageSlots; ++s2)
                    if (__builtin_amdgcn_readfirstlane(stageLeft[s2]) == 0u) {
                        newSlot = s2;
                        break;
                    }
                base2 = nextBatch(p.counter, n, walk);
            }
            if (!alive) {
                const uint32_t rank = uint32_t(__popcll(idle & ((1ull << lane) - 1ull)));
                const uint32_t idx = rank < avail ? qBase + rank : base2 + (rank - avail);
                if (idx < n) {
                    i = idx;
                    slot = rank < avail ? curSlot : newSlot;
                    r = loadRay(a.rays, idx);
                    bestT = r.tmax;
                    bestId = RBR_MISS_ID;
                    occ = false;
                    st.sp = 0;
                    ref = a.rootOf ? a.rootOf[idx] : 0u;
                    primLeft = 0u;
                    steps = 0u;
                    alive = true;
                }
            }
            if (avail < nIdle) {
                qBase = base2 + (nIdle - avail);
                qEnd = base2 + 64u;
                if (base2 >= n) {
                    more = false;   // every slice drained
                } else {
                    // Ranges start at multiples of 64 (slices hold whole batches), so a ray's
                    // row in its slot is i & 63.
                    if (lane == 0 && newSlot < kStageSlots) stageLeft[newSlot] = min(64u, n - base2);
                    curSlot = newSlot;
                }
            } else {
                qBase += nIdle;
            }
        }
        if (!__any(alive ? 1 : 0)) {
            if (!more) break;
            continue;
        }
        // One primitive of the leaf in progress (DEFER: only once primMin lanes hold one, or
        // fewer than nodeMin lanes have an interior node to test), then one interior node.
        bool doPrim = true;
        if (DEFER) {
            const uint32_t nPrim = uint32_t(__popcll(__ballot(primLeft != 0u ? 1 : 0)));
            const uint32_t nNode = uint32_t(__popcll(__ballot(alive && ref != kEmptyD && !isLeafD(ref) ? 1 : 0)));
            doPrim = nPrim >= p.primMin || nNode < p.nodeMin;
        }
        const bool primNow = doPrim && primLeft != 0u;
        const bool nodeNow = alive && ref != kEmptyD && !isLeafD(ref);
        fetchTest<W, Q, ANY, MIXED>(a, r, st, primNow, nodeNow, prim, primLeft, ref, bestT, bestId, occ, steps);
        // Watchdog: a correct tree takes at most its node count of node visits per ray; past
        // the cap the ray is abandoned (checked every 32 iterations of the wave).
        // The ballot inside the wave-uniform test keeps this a branch taken once in 32 iterations
        // (without it the compiler turns the check into selects on every iteration).
        if ((++iter & 31u) == 0u && __any(alive && steps > a.stepCap ? 1 : 0)) {
            if (alive && steps > a.stepCap) {
                uint32_t none = kEmptyD;
                abandonRay(a, st, ref, none);
                primLeft = 0u;
            }
        }
        settle<ANY>(st, alive, ref, prim, primLeft, bestT);
        bool last = false;
        if (alive && ref == kEmptyD && primLeft == 0u) {
            tracePlanes<ANY>(a, r, bestT, bestId, occ);
            if (slot < kStageSlots) {
                stage[slot * 64u + (i & 63u)] = ANY ? make_uint2(occ ? 1u : 0u, 0u)
                                                    : make_uint2(bestId, bestId == RBR_MISS_ID ? 0u : f2u(bestT));
                last = atomicSub(&stageLeft[slot], 1u) == 1u;
            } else {
                storeResult<ANY>(a, i, bestT, bestId, occ);
            }
            alive = false;
        }
        // Write out every range whose last ray finished in this iteration (64 lanes, one row each).
        for (uint64_t fl = __ballot(last ? 1 : 0); fl; fl &= fl - 1ull) {
            const int l = __builtin_ctzll(fl);
            const uint32_t s = __builtin_amdgcn_readlane(slot, l);
            const uint32_t b = __builtin_amdgcn_readlane(i, l) & ~63u;
            __syncthreads();   // the workgroup is this one wave: orders the rows written above
            const uint2 e = stage[s * 64u + lane];
            if (b + lane < n) {
                if (ANY) a.occ[b + lane] = uint8_t(e.x);
                else a.hits[b + lane] = e;
            }
        }
    }
}

__global__ void resetCounterKernel(uint32_t* c) {
    if (threadIdx.x < kSlices) c[size_t(threadIdx.x) * kSliceStride] = 0u;
}

// One ray per lane, stack overflow in private memory (kept for comparison and for callers
// that cannot provide the spill buffer).
template <int W, bool Q, int LDS, bool ANY, bool MIXED>
__global__ __launch_bounds__(64) void traceKernelLane(TraceArgs a) {
    __shared__ uint2 ldsStack[LDS * 64];
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    if (i >= a.n) return;
    PrivateSpillStack<LDS> st;
    st.lds = ldsStack;
    const DRay r = loadRay(a.rays, i);
    float bestT;
    uint32_t bestId;
    bool occ;
    traceRay<W, Q, ANY, MIXED>(a, r, a.rootOf ? a.rootOf[i] : 0u, st, bestT, bestId, occ);
    storeResult<ANY>(a, i, bestT, bestId, occ);
}

template <int W, bool Q, int LDS, bool ANY>
hipError_t launchT(const TraceArgs& a, const PersistArgs* p, unsigned grid, hipStream_t s) {
    if (p && p->refill && p->step) {
        hipLaunchKernelGGL(resetCounterKernel, dim3(1), dim3(64), 0, s, p->counter);
        if (p->primMin) {
            if (a.mixed) hipLaunchKernelGGL((traceKernelStep<W, Q, LDS, ANY, true, true>), dim3(grid), dim3(64), 0, s, a, *p);
            else hipLaunchKernelGGL((traceKernelStep<W, Q, LDS, ANY, false, true>), dim3(grid), dim3(64), 0, s, a, *p);
        } else {
            if (a.mixed) hipLaunchKernelGGL((traceKernelStep<W, Q, LDS, ANY, true, false>), dim3(grid), dim3(64), 0, s, a, *p);
            else hipLaunchKernelGGL((traceKernelStep<W, Q, LDS, ANY, false, false>), dim3(grid), dim3(64), 0, s, a, *p);
        }
    } else if (p && p->refill) {
        hipLaunchKernelGGL(resetCounterKernel, dim3(1), dim3(64), 0, s, p->counter);
        if (a.mixed) hipLaunchKernelGGL((traceKernelDynamic<W, Q, LDS, ANY, true>), dim3(grid), dim3(64), 0, s, a, *p);
        else hipLaunchKernelGGL((traceKernelDynamic<W, Q, LDS, ANY, false>), dim3(grid), dim3(64), 0, s, a, *p);
    } else if (p) {
        hipLaunchKernelGGL(resetCounterKernel, dim3(1), dim3(64), 0, s, p->counter);
        if (a.mixed) hipLaunchKernelGGL((traceKernelPersistent<W, Q, LDS, ANY, true>), dim3(grid), dim3(64), 0, s, a, *p);
        else hipLaunchKernelGGL((traceKernelPersistent<W, Q, LDS, ANY, false>), dim3(grid), dim3(64), 0, s, a, *p);
    } else {
        dim3 g((a.n + 63) / 64);
        if (a.mixed) hipLaunchKernelGGL((traceKernelLane<W, Q, LDS, ANY, true>), g, dim3(64), 0, s, a);
        else hipLaunchKernelGGL((traceKernelLane<W, Q, LDS, ANY, false>), g, dim3(64), 0, s, a);
    }
    return hipGetLastError();
}

template <int W, bool Q, int LDS>
hipError_t launchW(const TraceArgs& a, const PersistArgs* p, unsigned grid, bool any, hipStream_t s) {
    return any ? launchT<W, Q, LDS, true>(a, p, grid, s) : launchT<W, Q, LDS, false>(a, p, grid, s);
}

// width: 4 or 8, | kQuantFlag for quantized nodes
template <int LDS>
hipError_t launchL(const TraceArgs& a, const PersistArgs* p, unsigned grid, int width, bool any, hipStream_t s) {
    switch (width) {
        case 4: return launchW<4, false, LDS>(a, p, grid, any, s);
        case 8: return launchW<8, false, LDS>(a, p, grid, any, s);
        case 4 | kQuantFlag: return launchW<4, true, LDS>(a, p, grid, any, s);
        case 8 | kQuantFlag: return launchW<8, true, LDS>(a, p, grid, any, s);
        default: return hipErrorInvalidValue;
    }
}

template <int W, bool Q, int LDS, bool ANY, bool MIXED>
hipError_t residentT(int mode, int* blocks) {
    const void* f;
    if (mode >= 9) f = reinterpret_cast<const void*>(&traceKernelStep<W, Q, LDS, ANY, MIXED, true>);
    else if (mode >= 5) f = reinterpret_cast<const void*>(&traceKernelStep<W, Q, LDS, ANY, MIXED, false>);
    else if (mode >= 2) f = reinterpret_cast<const void*>(&traceKernelDynamic<W, Q, LDS, ANY, MIXED>);
    else if (mode == 0) f = reinterpret_cast<const void*>(&traceKernelPersistent<W, Q, LDS, ANY, MIXED>);
    else f = reinterpret_cast<const void*>(&traceKernelLane<W, Q, LDS, ANY, MIXED>);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, f, 64, 0);
}

template <int LDS>
hipError_t residentL(int width, bool any, bool mixed, int mode, int* blocks) {
#define RBR_RES(W, Q)                                                                             \
    return any ? (mixed ? residentT<W, Q, LDS, true, true>(mode, blocks) : residentT<W, Q, LDS, true, false>(mode, blocks)) \
               : (mixed ? residentT<W, Q, LDS, false, true>(mode, blocks) : residentT<W, Q, LDS, false, false>(mode, blocks))
    switch (width) {
        case 4: RBR_RES(4, false);
        case 8: RBR_RES(8, false);
        case 4 | kQuantFlag: RBR_RES(4, true);
        case 8 | kQuantFlag: RBR_RES(8, true);
        default: return hipErrorInvalidValue;
    }
#undef RBR_RES
}

}  // namespace

hipError_t traceResidentWaves(int width, bool any, bool mixed, int ldsDepth, int mode, int* wavesPerCu) {
    switch (ldsDepth) {
        case 8: return residentL<8>(width, any, mixed, mode, wavesPerCu);
        case 10: return residentL<10>(width, any, mixed, mode, wavesPerCu);
        case 12: return residentL<12>(width, any, mixed, mode, wavesPerCu);
        case 24: return residentL<24>(width, any, mixed, mode, wavesPerCu);
        case 32: return residentL<32>(width, any, mixed, mode, wavesPerCu);
        default: return residentL<16>(width, any, mixed, mode, wavesPerCu);
    }
}

hipError_t launchTrace(const TraceArgs& args, int width, bool any, int ldsDepth, const PersistArgs* p, unsigned grid,
                       hipStream_t s) {
    if (args.n == 0 && !args.nDev) return hipSuccess;
    if (!p && args.nDev) return hipErrorInvalidValue;
    TraceArgs a = args;
    if (a.stepCap == 0u) a.stepCap = 0xFFFFFFFFu;   // no cap given: unbounded
    switch (ldsDepth) {
        case 8: return launchL<8>(a, p, grid, width, any, s);
        case 10: return launchL<10>(a, p, grid, width, any, s);
        case 12: return launchL<12>(a, p, grid, width, any, s);
        case 24: return launchL<24>(a, p, grid, width, any, s);
        case 32: return launchL<32>(a, p, grid, width, any, s);
        default: return launchL<16>(a, p, grid, width, any, s);
    }
}

}  // namespace rbr
