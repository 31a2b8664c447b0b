// Wide-BVH closest-hit / any-hit traversal kernels for gfx950.
//
// Replaces BVH::intersect / intersectP (bvh.cpp:33-41, 211-267): instead of a recursive,
// pointer-chasing walk per ray with a virtual call per primitive, one lane owns one ray and
// walks the flat wide BVH (bvh.h) with a short stack:
//   * node boxes are read as aligned float4 rows (one 128-byte line per BVH4 node);
//   * children are tested with the sign-selected FMA slab test against [0, t_best] and
//     visited nearest-first; pushed entries carry their entry distance so entries that
//     became farther than the current closest hit are dropped on pop without a fetch;
//   * the stack lives in LDS (kLdsDepth entries per lane, [depth][lane] so a wave's
//     pushes hit distinct banks) and spills to per-lane private memory beyond that;
//   * primitives are tested with the exact Moller-Trumbore arithmetic of the reference
//     (prim_tests.h, contraction off) and the lowest id wins among equal t.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bvh.h"
#include "prim_tests.h"
#include "trace.h"

#pragma clang fp contract(off)

namespace rbr {

namespace {

__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }

struct DRay {
    V o, d;
    float ix, iy, iz;      // 1/d with zero components replaced by +-1e-30
    float oix, oiy, oiz;   // o * inv
    bool nx, ny, nz;       // sign of inv
    float tmin, tmax;
};

__device__ __forceinline__ float safeInv(float x) {
    float y = fabsf(x) < 1e-30f ? copysignf(1e-30f, x) : x;
    return __builtin_amdgcn_rcpf(y);
}

__device__ __forceinline__ DRay loadRay(const float4* rays, uint32_t i) {
    float4 a = rays[2 * size_t(i)];
    float4 b = rays[2 * size_t(i) + 1];
    DRay r;
    r.o = V{a.x, a.y, a.z};
    r.d = V{a.w, b.x, b.y};
    r.tmin = b.z;
    r.tmax = b.w;
    r.ix = safeInv(r.d.x);
    r.iy = safeInv(r.d.y);
    r.iz = safeInv(r.d.z);
    r.oix = r.o.x * r.ix;
    r.oiy = r.o.y * r.iy;
    r.oiz = r.o.z * r.iz;
    r.nx = r.ix < 0.0f;
    r.ny = r.iy < 0.0f;
    r.nz = r.iz < 0.0f;
    return r;
}

// Slab test of one child; lo/hi per axis from the node rows.  Uses fused multiply-adds on
// purpose (the boxes are padded, so the rounding cannot cull a real hit).
__device__ __forceinline__ float slab(const DRay& r, float lx, float hx, float ly, float hy, float lz, float hz,
                                      float tfar, bool& hit) {
    float nx = r.nx ? hx : lx, fx = r.nx ? lx : hx;
    float ny = r.ny ? hy : ly, fy = r.ny ? ly : hy;
    float nz = r.nz ? hz : lz, fz = r.nz ? lz : hz;
    float t0x = __builtin_fmaf(nx, r.ix, -r.oix), t1x = __builtin_fmaf(fx, r.ix, -r.oix);
    float t0y = __builtin_fmaf(ny, r.iy, -r.oiy), t1y = __builtin_fmaf(fy, r.iy, -r.oiy);
    float t0z = __builtin_fmaf(nz, r.iz, -r.oiz), t1z = __builtin_fmaf(fz, r.iz, -r.oiz);
    float t0 = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, 0.0f));
    float t1 = fminf(fminf(t1x, t1y), fminf(t1z, tfar));
    hit = t0 <= t1;
    return t0;
}

// Children of a quantized node (bvh.h): the ray is moved into the node's grid once,
// t(q) = q * (s * 1/d) + (p - o) * 1/d, so each of the 6 planes of a child costs one FMA on
// its byte.  Fills entry distances (tc) and refs (rc) like the float path.
template <int W>
__device__ __forceinline__ void childrenQ(const float4* nodes, uint32_t ref, const DRay& r, float tfar, float* tc,
                                          uint32_t* rc) {
    constexpr int Q4 = W / 4;
    constexpr int NV = W == 4 ? 4 : 6;   // uint4 rows that hold data (64 B / 96 B)
    const uint4* nd = reinterpret_cast<const uint4*>(nodes) + size_t(ref) * (W == 4 ? 4 : 8);
    uint32_t w[4 * NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const uint4 v = nd[i];
        w[4 * i] = v.x;
        w[4 * i + 1] = v.y;
        w[4 * i + 2] = v.z;
        w[4 * i + 3] = v.w;
    }
    const float ax = u2f((w[3] & 0xFFu) << 23) * r.ix;
    const float ay = u2f(((w[3] >> 8) & 0xFFu) << 23) * r.iy;
    const float az = u2f(((w[3] >> 16) & 0xFFu) << 23) * r.iz;
    const float bx = __builtin_fmaf(u2f(w[0]), r.ix, -r.oix);
    const float by = __builtin_fmaf(u2f(w[1]), r.iy, -r.oiy);
    const float bz = __builtin_fmaf(u2f(w[2]), r.iz, -r.oiz);
    // near / far plane words per axis: one select covers 4 children
    uint32_t nw[3][Q4], fw[3][Q4];
#pragma unroll
    for (int g = 0; g < Q4; ++g) {
        nw[0][g] = r.nx ? w[4 + 1 * Q4 + g] : w[4 + 0 * Q4 + g];
        fw[0][g] = r.nx ? w[4 + 0 * Q4 + g] : w[4 + 1 * Q4 + g];
        nw[1][g] = r.ny ? w[4 + 3 * Q4 + g] : w[4 + 2 * Q4 + g];
        fw[1][g] = r.ny ? w[4 + 2 * Q4 + g] : w[4 + 3 * Q4 + g];
        nw[2][g] = r.nz ? w[4 + 5 * Q4 + g] : w[4 + 4 * Q4 + g];
        fw[2][g] = r.nz ? w[4 + 4 * Q4 + g] : w[4 + 5 * Q4 + g];
    }
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const int g = k >> 2, sh = 8 * (k & 3);
        const float t0x = __builtin_fmaf(float((nw[0][g] >> sh) & 0xFFu), ax, bx);
        const float t1x = __builtin_fmaf(float((fw[0][g] >> sh) & 0xFFu), ax, bx);
        const float t0y = __builtin_fmaf(float((nw[1][g] >> sh) & 0xFFu), ay, by);
        const float t1y = __builtin_fmaf(float((fw[1][g] >> sh) & 0xFFu), ay, by);
        const float t0z = __builtin_fmaf(float((nw[2][g] >> sh) & 0xFFu), az, bz);
        const float t1z = __builtin_fmaf(float((fw[2][g] >> sh) & 0xFFu), az, bz);
        const float t0 = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, 0.0f));
        const float t1 = fminf(fminf(t1x, t1y), fminf(t1z, tfar));
        const uint32_t ch = w[4 + 6 * Q4 + k];
        const bool h = t0 <= t1 && ch != kEmptyD;
        tc[k] = h ? t0 : INFINITY;
        rc[k] = h ? ch : kEmptyD;
    }
}

// Children of a float node (bvh.h), same outputs as childrenQ.
template <int W>
__device__ __forceinline__ void childrenF(const float4* nodes, uint32_t ref, const DRay& r, float tfar, float* tc,
                                          uint32_t* rc) {
    const float4* nd = nodes + size_t(ref) * (W == 4 ? 8 : 16);
#pragma unroll
    for (int half = 0; half < W / 4; ++half) {
        const int q = W / 4;
        float4 lx = nd[0 * q + half], hx = nd[1 * q + half], ly = nd[2 * q + half];
        float4 hy = nd[3 * q + half], lz = nd[4 * q + half], hz = nd[5 * q + half];
        uint4 ch = *reinterpret_cast<const uint4*>(nd + 6 * q + half);
        const float* LX = &lx.x; const float* HX = &hx.x; const float* LY = &ly.x;
        const float* HY = &hy.x; const float* LZ = &lz.x; const float* HZ = &hz.x;
        const uint32_t* CH = &ch.x;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            bool h;
            float t = slab(r, LX[k], HX[k], LY[k], HY[k], LZ[k], HZ[k], tfar, h);
            tc[half * 4 + k] = h ? t : INFINITY;
            rc[half * 4 + k] = h ? CH[k] : kEmptyD;
        }
    }
}

template <bool MIXED>
__device__ __forceinline__ bool testPrim(const TraceArgs& a, const float4* rec, const DRay& r, float tmax, float& t,
                                         uint32_t& id) {
    float4 p0 = rec[0], p1 = rec[1], p2 = rec[2];
    id = f2u(p0.w);
    if (!MIXED) return triTest(r.o, r.d, V{p0.x, p0.y, p0.z}, V{p1.x, p1.y, p1.z}, V{p2.x, p2.y, p2.z}, r.tmin, tmax, t);
    uint32_t kind = f2u(p1.w);
    if (kind == KIND_TRI)
        return triTest(r.o, r.d, V{p0.x, p0.y, p0.z}, V{p1.x, p1.y, p1.z}, V{p2.x, p2.y, p2.z}, r.tmin, tmax, t);
    if (kind == KIND_SPHERE) return sphereTest(r.o, r.d, V{p0.x, p0.y, p0.z}, p1.x, r.tmin, tmax, t);
    float4 n = a.quadN[f2u(p2.w)];
    return quadTest(r.o, r.d, V{p0.x, p0.y, p0.z}, V{p1.x, p1.y, p1.z}, V{p2.x, p2.y, p2.z}, V{n.x, n.y, n.z}, r.tmin,
                    tmax, t);
}

template <typename T>
__device__ __forceinline__ void cswap(float& ta, T& ra, float& tb, T& rb) {
    bool s = tb < ta;
    float tt = s ? tb : ta;
    tb = s ? ta : tb;
    ta = tt;
    T rr = s ? rb : ra;
    rb = s ? ra : rb;
    ra = rr;
}

// Traversal stack: LDS rows [depth][lane] for the first LDS entries of each lane; deeper
// entries spill to a per-wave region of a global buffer (persistent kernel) or to private
// memory (one-ray-per-lane kernel).
template <int LDS>
struct GlobalSpillStack {
    uint2* lds;     // this wave's LDS rows
    uint2* spill;   // this wave's spill rows [depth - LDS][64]
    int sp = 0;
    __device__ __forceinline__ void push(uint32_t ref, float t) {
        uint2 e = make_uint2(ref, f2u(t));
        if (sp < LDS) lds[sp * 64 + threadIdx.x] = e;
        else spill[(sp - LDS) * 64 + threadIdx.x] = e;
        ++sp;
    }
    __device__ __forceinline__ uint2 pop() {
        --sp;
        return sp < LDS ? lds[sp * 64 + threadIdx.x] : spill[(sp - LDS) * 64 + threadIdx.x];
    }
};

template <int LDS>
struct PrivateSpillStack {
    uint2* lds;
    uint2 spill[kMaxStack - LDS];
    int sp = 0;
    __device__ __forceinline__ void push(uint32_t ref, float t) {
        uint2 e = make_uint2(ref, f2u(t));
        if (sp < LDS) lds[sp * 64 + threadIdx.x] = e;
        else spill[sp - LDS] = e;
        ++sp;
    }
    __device__ __forceinline__ uint2 pop() {
        --sp;
        return sp < LDS ? lds[sp * 64 + threadIdx.x] : spill[sp - LDS];
    }
};

// One round of the traversal loop of one ray ("while-while" with postponed leaves, after Aila
// & Laine 2009): a lane walks interior nodes until it meets a leaf, parks it, and keeps
// walking until every lane of the wave that still traverses holds a parked leaf; then the
// wave tests leaves together.  Lanes in a wave so spend their iterations on the same kind of
// work.  Parking a leaf can only add node visits (culling by the closest hit comes later),
// never change the result: the closest hit is the minimum over every primitive tested, and
// every entry that can still hold a hit at least as close is visited.  On return, ref ==
// kEmptyD (or an any-hit occlusion) means the ray is done.
template <int W, bool Q, bool ANY, bool MIXED, typename Stack>
__device__ __forceinline__ void traceRound(const TraceArgs& a, const DRay& r, Stack& st, uint32_t& ref, uint32_t& leaf,
                                           float& bestT, uint32_t& bestId, bool& occluded) {
    // next stack entry that can still hold a hit (closest hit: entry distance <= best t)
    auto popNext = [&]() -> uint32_t {
        while (st.sp > 0) {
            const uint2 e = st.pop();
            if (ANY || u2f(e.y) <= bestT) return e.x;
        }
        return kEmptyD;
    };
    while (ref != kEmptyD && !isLeafD(ref)) {
        float tc[W];
        uint32_t rc[W];
        if (Q) childrenQ<W>(a.nodes, ref, r, bestT, tc, rc);
        else childrenF<W>(a.nodes, ref, r, bestT, tc, rc);
        if (!ANY) {
            // Sorting network, ascending entry distance (misses sort last as +inf).
            if (W == 4) {
                cswap(tc[0], rc[0], tc[1], rc[1]); cswap(tc[2], rc[2], tc[3], rc[3]);
                cswap(tc[0], rc[0], tc[2], rc[2]); cswap(tc[1], rc[1], tc[3], rc[3]);
                cswap(tc[1], rc[1], tc[2], rc[2]);
            } else {
                // Optimal 8-input network (19 comparators, depth 6).
                constexpr int net[19][2] = {{0, 2}, {1, 3}, {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6},
                                            {3, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}, {2, 4}, {3, 5},
                                            {1, 4}, {3, 6}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
                for (int c = 0; c < 19; ++c)
                    cswap(tc[net[c][0]], rc[net[c][0]], tc[net[c][1]], rc[net[c][1]]);
            }
        }
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < W; ++k) cnt += rc[k] != kEmptyD ? 1 : 0;
        if (cnt == 0) {
            ref = popNext();
        } else if (ANY) {
            uint32_t first = kEmptyD;
#pragma unroll
            for (int k = 0; k < W; ++k) {
                if (rc[k] == kEmptyD) continue;
                if (first == kEmptyD) first = rc[k];
                else st.push(rc[k], tc[k]);
            }
            ref = first;
        } else {
#pragma unroll
            for (int k = W - 1; k >= 1; --k)
                if (k < cnt) st.push(rc[k], tc[k]);
            ref = rc[0];
        }
        if (ref != kEmptyD && isLeafD(ref) && leaf == kEmptyD) {
            leaf = ref;
            ref = popNext();
        }
        if (__all(leaf != kEmptyD ? 1 : 0)) break;
    }
    // Leaves: the parked one, then ref for as long as it is a leaf.
    while (leaf != kEmptyD) {
        const uint32_t s = leafStartD(leaf), c = leafCountD(leaf);
        for (uint32_t k = 0; k < c; ++k) {
            float t;
            uint32_t id;
            if (testPrim<MIXED>(a, a.prims + size_t(s + k) * 3, r, bestT, t, id)) {
                if (ANY) { occluded = true; break; }
                if (t < bestT || id < bestId) { bestT = t; bestId = id; }
            }
        }
        if (ANY && occluded) break;
        leaf = kEmptyD;
        if (ref != kEmptyD && isLeafD(ref)) {
            leaf = ref;
            ref = popNext();
        }
    }
}

// Unbounded primitives (infinite planes), tested linearly after the BVH.
template <bool ANY>
__device__ __forceinline__ void tracePlanes(const TraceArgs& a, const DRay& r, float& bestT, uint32_t& bestId,
                                            bool& occluded) {
    for (uint32_t k = 0; k < a.nplanes && !(ANY && occluded); ++k) {
        float4 pl = a.planes[k];
        float t;
        if (planeTest(r.o, r.d, V{pl.x, pl.y, pl.z}, pl.w, r.tmin, bestT, t)) {
            if (ANY) { occluded = true; break; }
            uint32_t id = a.planeIds[k];
            if (t < bestT || id < bestId) { bestT = t; bestId = id; }
        }
    }
}

// One ray through the wide BVH and the unbounded list.  Returns the lowest-id closest hit
// (bestT/bestId) or, for ANY, whether any primitive is hit in [tmin, tmax].
template <int W, bool Q, bool ANY, bool MIXED, typename Stack>
__device__ __forceinline__ void traceRay(const TraceArgs& a, const DRay& r, uint32_t root, Stack& st, float& bestT,
                                         uint32_t& bestId, bool& occluded) {
    bestT = r.tmax;
    bestId = RBR_MISS_ID;
    occluded = false;
    if (a.nodes && root != kEmptyD) {
        uint32_t ref = root, leaf = kEmptyD;
        do {
            traceRound<W, Q, ANY, MIXED>(a, r, st, ref, leaf, bestT, bestId, occluded);
        } while (!((ANY && occluded) || ref == kEmptyD));
    }
    tracePlanes<ANY>(a, r, bestT, bestId, occluded);
}

template <bool ANY>
__device__ __forceinline__ void storeResult(const TraceArgs& a, uint32_t i, float bestT, uint32_t bestId, bool occ) {
    if (ANY) a.occ[i] = occ ? 1 : 0;
    else a.hits[i] = make_uint2(bestId, bestId == RBR_MISS_ID ? 0u : f2u(bestT));
}

// Persistent waves: each 64-lane workgroup takes batches of 64 consecutive rays from a
// device counter until the batch is exhausted, so slow rays do not idle a whole launch
// slot and every resident wave stays busy to the end.  The counter is zeroed by
// resetCounterKernel right before the launch on the same stream; the ray count can come
// from device memory (queue lengths written by an earlier kernel).
template <int W, bool Q, int LDS, bool ANY, bool MIXED>
__global__ __launch_bounds__(64) void traceKernelPersistent(TraceArgs a, PersistArgs p) {
    __shared__ uint2 ldsStack[LDS * 64];
    __shared__ uint32_t batch;
    GlobalSpillStack<LDS> st;
    st.lds = ldsStack;
    st.spill = p.spill + size_t(blockIdx.x) * size_t(p.spillDepth) * 64;
    const uint32_t n = a.nDev ? *a.nDev : a.n;
    while (true) {
        if (threadIdx.x == 0) batch = atomicAdd(p.counter, 64u);
        __syncthreads();
        const uint32_t base = batch;
        __syncthreads();
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
// 64-ray range the wave reserves with one atomic on the device counter (the rate of the
// batch kernel's atomics); when the range runs dry mid-refill the rest comes from the next
// range.  Between refills each live lane runs one traversal round.  Every wave exits once the
// counter has passed the ray count and all of its lanes are idle.
template <int W, bool Q, int LDS, bool ANY, bool MIXED>
__global__ __launch_bounds__(64) void traceKernelDynamic(TraceArgs a, PersistArgs p) {
    __shared__ uint2 ldsStack[LDS * 64];
    GlobalSpillStack<LDS> st;
    st.lds = ldsStack;
    st.spill = p.spill + size_t(blockIdx.x) * size_t(p.spillDepth) * 64;
    const uint32_t n = a.nDev ? *a.nDev : a.n;
    const uint32_t lane = threadIdx.x;
    bool alive = false, more = true, occ = false;
    uint32_t i = 0, ref = kEmptyD, leaf = kEmptyD, bestId = RBR_MISS_ID;
    uint32_t qBase = 0u, qEnd = 0u;   // reserved, not yet assigned rays (uniform)
    float bestT = 0.0f;
    DRay r{};
    while (true) {
        const uint64_t idle = __ballot(alive ? 0 : 1);
        const uint32_t nIdle = uint32_t(__popcll(idle));
        if (more && nIdle >= p.refill) {
            const uint32_t avail = qEnd - qBase;
            uint32_t base2 = 0u;
            if (avail < nIdle) {
                if (lane == 0) base2 = atomicAdd(p.counter, 64u);
                base2 = __shfl(base2, 0);
            }
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
                    alive = true;
                }
            }
            if (avail < nIdle) {
                qBase = base2 + (nIdle - avail);
                qEnd = base2 + 64u;
                if (base2 >= n) more = false;
            } else {
                qBase += nIdle;
            }
        }
        if (!__any(alive ? 1 : 0)) {
            if (!more) break;
            continue;
        }
        if (alive) {
            if (ref != kEmptyD || leaf != kEmptyD) traceRound<W, Q, ANY, MIXED>(a, r, st, ref, leaf, bestT, bestId, occ);
            if ((ANY && occ) || (ref == kEmptyD && leaf == kEmptyD)) {
                tracePlanes<ANY>(a, r, bestT, bestId, occ);
                storeResult<ANY>(a, i, bestT, bestId, occ);
                alive = false;
            }
        }
    }
}

__global__ void resetCounterKernel(uint32_t* c) { *c = 0u; }

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
    if (p && p->refill) {
        hipLaunchKernelGGL(resetCounterKernel, dim3(1), dim3(1), 0, s, p->counter);
        if (a.mixed) hipLaunchKernelGGL((traceKernelDynamic<W, Q, LDS, ANY, true>), dim3(grid), dim3(64), 0, s, a, *p);
        else hipLaunchKernelGGL((traceKernelDynamic<W, Q, LDS, ANY, false>), dim3(grid), dim3(64), 0, s, a, *p);
    } else if (p) {
        hipLaunchKernelGGL(resetCounterKernel, dim3(1), dim3(1), 0, s, p->counter);
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

}  // namespace

hipError_t launchTrace(const TraceArgs& a, int width, bool any, int ldsDepth, const PersistArgs* p, unsigned grid,
                       hipStream_t s) {
    if (a.n == 0 && !a.nDev) return hipSuccess;
    if (!p && a.nDev) return hipErrorInvalidValue;
    switch (ldsDepth) {
        case 8: return launchL<8>(a, p, grid, width, any, s);
        case 12: return launchL<12>(a, p, grid, width, any, s);
        case 24: return launchL<24>(a, p, grid, width, any, s);
        case 32: return launchL<32>(a, p, grid, width, any, s);
        default: return launchL<16>(a, p, grid, width, any, s);
    }
}

}  // namespace rbr
