// Device-side wide-BVH traversal shared by the traversal kernels (trace.hip) and the kernels
// that trace rays inline (the Whitted integrator in render.hip).
//
// Replaces BVH::intersect / intersectP (bvh.cpp:33-41, 211-267): one lane owns one ray and
// walks the flat wide BVH (bvh.h) with a short stack:
//   * node boxes are read as aligned 16-byte rows;
//   * children are tested with the sign-selected FMA slab test against [0, t_best] and
//     visited nearest-first; pushed entries carry their entry distance so entries that
//     became farther than the current closest hit are dropped on pop without a fetch;
//   * the stack lives in LDS rows ([depth][lane] so a wave's pushes hit distinct banks) and
//     spills to global or private memory beyond that;
//   * primitives are tested with the exact Moller-Trumbore arithmetic of the reference
//     (prim_tests.h, contraction off) and the lowest id wins among equal t.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bvh.h"
#include "prim_tests.h"
#include "trace.h"

#pragma clang fp contract(off)

namespace rbr {
namespace tdev {

__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ float4 u4f(uint4 v) { return make_float4(u2f(v.x), u2f(v.y), u2f(v.z), u2f(v.w)); }

struct DRay {
    V o, d;
    float ix, iy, iz;      // 1/d with zero components replaced by +-1e-30
    float oix, oiy, oiz;   // o * inv
    bool nx, ny, nz;       // sign of inv
    uint32_t selx, sely, selz;   // v_perm_b32 selectors of the near plane codes (childrenQ)
    float tmin, tmax;
};

// Selectors of v_perm_b32(plane dword, 0x64646464, sel) for the plane codes of a child pair
// (dword bytes q_lo[2g], q_lo[2g+1], q_hi[2g], q_hi[2g+1]; bvh.h): the result holds two
// half-precision numbers 1024 + q (byte q under the exponent byte 0x64), the lo codes for
// kSelLo and the hi codes for kSelHi = kSelLo ^ kSelSwap.
constexpr uint32_t kSelLo = 0x00050004u;
constexpr uint32_t kSelHi = 0x00070006u;
constexpr uint32_t kSelSwap = 0x00020002u;
constexpr uint32_t kHalfExp = 0x64646464u;

__device__ __forceinline__ float safeInv(float x) {
    float y = fabsf(x) < 1e-30f ? copysignf(1e-30f, x) : x;
    return __builtin_amdgcn_rcpf(y);
}

// The reciprocal direction and o/d the slab tests use, from o and d.
__device__ __forceinline__ void setupRay(DRay& r) {
    r.ix = safeInv(r.d.x);
    r.iy = safeInv(r.d.y);
    r.iz = safeInv(r.d.z);
    r.oix = r.o.x * r.ix;
    r.oiy = r.o.y * r.iy;
    r.oiz = r.o.z * r.iz;
    r.nx = r.ix < 0.0f;
    r.ny = r.iy < 0.0f;
    r.nz = r.iz < 0.0f;
    r.selx = r.nx ? kSelHi : kSelLo;
    r.sely = r.ny ? kSelHi : kSelLo;
    r.selz = r.nz ? kSelHi : kSelLo;
}

// A ray from its two 16-byte rows (o.xyz d.x | d.yz tmin tmax).
__device__ __forceinline__ DRay loadRay(const float4* rays, uint32_t i) {
    const float4 a = rays[2 * size_t(i)], b = rays[2 * size_t(i) + 1];
    DRay r;
    r.o = V{a.x, a.y, a.z};
    r.d = V{a.w, b.x, b.y};
    r.tmin = b.z;
    r.tmax = b.w;
    setupRay(r);
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

// The 16-byte rows of a node the child tests read: quantized BVH4 64 B, quantized BVH8 the
// first 96 B of its 128, float BVH4 128 B, float BVH8 256 B (bvh.h layouts).
template <int W, bool Q>
struct NodeRows {
    static constexpr int kRows = Q ? (W == 4 ? 4 : 6) : (W == 4 ? 8 : 16);
    static constexpr int kStride = Q ? (W == 4 ? 4 : 8) : (W == 4 ? 8 : 16);   // rows per node
    uint4 v[kRows];
};

template <int W, bool Q>
__device__ __forceinline__ void loadNode(const float4* nodes, uint32_t ref, NodeRows<W, Q>& nr) {
    const uint4* nd = reinterpret_cast<const uint4*>(nodes) + size_t(ref) * NodeRows<W, Q>::kStride;
#pragma unroll
    for (int i = 0; i < NodeRows<W, Q>::kRows; ++i) nr.v[i] = nd[i];
}

// Half-precision number j (0: low, 1: high) of a dword, widened to float.  Used as an FMA
// operand it becomes one v_fma_mix_f32 with op_sel.
__device__ __forceinline__ float halfOf(uint32_t w, int j) {
    return float(__builtin_bit_cast(_Float16, static_cast<unsigned short>(j ? w >> 16 : w & 0xFFFFu)));
}

// Children of a quantized node (bvh.h): the ray is moved into the node's grid once,
// t(c) = c * (s * 1/d) + (p - o) * 1/d with c = 1024 + q, so each of the 6 planes of a child
// costs one FMA.  One v_perm_b32 per axis and child pair picks the near (or far) plane codes
// by the ray's direction sign and widens them to the half-precision numbers 1024 + q, which
// v_fma_mix_f32 reads directly.  Fills entry distances (tc), refs (rc) and hit flags like the
// float path.
template <int W>
__device__ __forceinline__ void childrenQ(const NodeRows<W, true>& nr, const DRay& r, float tfar, float* tc,
                                          uint32_t* rc, bool* hit) {
    constexpr int H = W / 2;
    constexpr int NV = NodeRows<W, true>::kRows;
    uint32_t w[4 * NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        w[4 * i] = nr.v[i].x;
        w[4 * i + 1] = nr.v[i].y;
        w[4 * i + 2] = nr.v[i].z;
        w[4 * i + 3] = nr.v[i].w;
    }
    // per-axis scales: floats in dwords 3, 14, 15 (W = 4) or exponent bytes of dword 3 (W = 8)
    const float sx = W == 4 ? u2f(w[3]) : u2f((w[3] & 0xFFu) << 23);
    const float sy = W == 4 ? u2f(w[14]) : u2f(((w[3] >> 8) & 0xFFu) << 23);
    const float sz = W == 4 ? u2f(w[15]) : u2f(((w[3] >> 16) & 0xFFu) << 23);
    const float ax = sx * r.ix;
    const float ay = sy * r.iy;
    const float az = sz * r.iz;
    const float bx = __builtin_fmaf(u2f(w[0]), r.ix, -r.oix);
    const float by = __builtin_fmaf(u2f(w[1]), r.iy, -r.oiy);
    const float bz = __builtin_fmaf(u2f(w[2]), r.iz, -r.oiz);
    const uint32_t fselx = r.selx ^ kSelSwap, fsely = r.sely ^ kSelSwap, fselz = r.selz ^ kSelSwap;
#pragma unroll
    for (int g = 0; g < H; ++g) {
        const uint32_t nxw = __builtin_amdgcn_perm(w[4 + g], kHalfExp, r.selx);
        const uint32_t fxw = __builtin_amdgcn_perm(w[4 + g], kHalfExp, fselx);
        const uint32_t nyw = __builtin_amdgcn_perm(w[4 + H + g], kHalfExp, r.sely);
        const uint32_t fyw = __builtin_amdgcn_perm(w[4 + H + g], kHalfExp, fsely);
        const uint32_t nzw = __builtin_amdgcn_perm(w[4 + 2 * H + g], kHalfExp, r.selz);
        const uint32_t fzw = __builtin_amdgcn_perm(w[4 + 2 * H + g], kHalfExp, fselz);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int k = 2 * g + j;
            const float t0x = __builtin_fmaf(halfOf(nxw, j), ax, bx);
            const float t1x = __builtin_fmaf(halfOf(fxw, j), ax, bx);
            const float t0y = __builtin_fmaf(halfOf(nyw, j), ay, by);
            const float t1y = __builtin_fmaf(halfOf(fyw, j), ay, by);
            const float t0z = __builtin_fmaf(halfOf(nzw, j), az, bz);
            const float t1z = __builtin_fmaf(halfOf(fzw, j), az, bz);
            const float t0 = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, 0.0f));
            const float t1 = fminf(fminf(t1x, t1y), fminf(t1z, tfar));
            const uint32_t ch = w[4 + 3 * H + k];
            const bool h = t0 <= t1 && ch != kEmptyD;
            tc[k] = h ? t0 : INFINITY;
            rc[k] = ch;
            hit[k] = h;
        }
    }
}

// Children of a float node (bvh.h), same outputs as childrenQ.
template <int W>
__device__ __forceinline__ void childrenF(const NodeRows<W, false>& nr, const DRay& r, float tfar, float* tc,
                                          uint32_t* rc, bool* hit) {
    constexpr int q = W / 4;
#pragma unroll
    for (int half = 0; half < q; ++half) {
        const float4 lx = u4f(nr.v[0 * q + half]), hx = u4f(nr.v[1 * q + half]), ly = u4f(nr.v[2 * q + half]);
        const float4 hy = u4f(nr.v[3 * q + half]), lz = u4f(nr.v[4 * q + half]), hz = u4f(nr.v[5 * q + half]);
        const uint4 ch = nr.v[6 * q + half];
        const float* LX = &lx.x; const float* HX = &hx.x; const float* LY = &ly.x;
        const float* HY = &hy.x; const float* LZ = &lz.x; const float* HZ = &hz.x;
        const uint32_t* CH = &ch.x;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            bool h;
            float t = slab(r, LX[k], HX[k], LY[k], HY[k], LZ[k], HZ[k], tfar, h);
            h = h && CH[k] != kEmptyD;
            tc[half * 4 + k] = h ? t : INFINITY;
            rc[half * 4 + k] = CH[k];
            hit[half * 4 + k] = h;
        }
    }
}

// Entry distance (+inf when missed), ref (as stored, also for a missed child) and hit flag of
// every child slot; empty slots never hit.
template <int W, bool Q>
__device__ __forceinline__ void children(const NodeRows<W, Q>& nr, const DRay& r, float tfar, float* tc,
                                         uint32_t* rc, bool* hit) {
    if constexpr (Q) childrenQ<W>(nr, r, tfar, tc, rc, hit);
    else childrenF<W>(nr, r, tfar, tc, rc, hit);
}

// One primitive record (3 rows: v0 / centre / corner + id, e1 / radius + kind, e2 + quad
// normal index), already loaded, against the ray.
template <bool MIXED>
__device__ __forceinline__ bool testPrimLoaded(const TraceArgs& a, const float4& p0, const float4& p1,
                                               const float4& p2, const DRay& r, float tmax, float& t, uint32_t& id) {
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

template <bool MIXED>
__device__ __forceinline__ bool testPrim(const TraceArgs& a, const float4* rec, const DRay& r, float tmax, float& t,
                                         uint32_t& id) {
    const float4 p0 = rec[0], p1 = rec[1], p2 = rec[2];
    return testPrimLoaded<MIXED>(a, p0, p1, p2, r, tmax, t, id);
}

// Traversal stack: LDS rows [depth][lane] for the first LDS entries of each lane; deeper
// entries spill to a per-wave region of a global buffer (persistent kernel) or to private
// memory (one-ray-per-lane kernel).
// Entries are (ref, entry distance) packed in 64 bits.  The row pointers carry their address
// space: with plain pointers the compiler merges the LDS and global accesses of push/pop into
// one select of pointers and a flat access, which waits on both memory counters and costs
// several instructions to form the address.
typedef __attribute__((address_space(3))) unsigned long long LdsU64;
typedef __attribute__((address_space(1))) unsigned long long GlobalU64;

template <int LDS>
struct GlobalSpillStack {
    LdsU64* lds;        // this wave's LDS rows
    GlobalU64* spill;   // this wave's spill rows [depth - LDS][64]
    int sp = 0;
    __device__ __forceinline__ void push(uint32_t ref, float t) {
        const unsigned long long e = (static_cast<unsigned long long>(f2u(t)) << 32) | ref;
        if (sp < LDS) lds[sp * 64 + threadIdx.x] = e;
        else spill[(sp - LDS) * 64 + threadIdx.x] = e;
        ++sp;
    }
    __device__ __forceinline__ uint2 pop() {
        --sp;
        unsigned long long e;
        if (sp < LDS) e = lds[sp * 64 + threadIdx.x];
        else e = spill[(sp - LDS) * 64 + threadIdx.x];
        return make_uint2(static_cast<uint32_t>(e), static_cast<uint32_t>(e >> 32));
    }
};

// LDS rows of STRIDE lanes (the workgroup size), private spill beyond them.
template <int LDS, int STRIDE = 64>
struct PrivateSpillStack {
    uint2* lds;
    uint2 spill[kMaxStack - LDS];
    int sp = 0;
    __device__ __forceinline__ void push(uint32_t ref, float t) {
        uint2 e = make_uint2(ref, f2u(t));
        if (sp < LDS) lds[sp * STRIDE + threadIdx.x] = e;
        else spill[sp - LDS] = e;
        ++sp;
    }
    __device__ __forceinline__ uint2 pop() {
        --sp;
        return sp < LDS ? lds[sp * STRIDE + threadIdx.x] : spill[sp - LDS];
    }
};

// Next stack entry that can still hold a hit (closest hit: entry distance <= best t), or
// kEmptyD.
template <bool ANY, typename Stack>
__device__ __forceinline__ uint32_t popNext(Stack& st, float bestT) {
    while (st.sp > 0) {
        const uint2 e = st.pop();
        if (ANY || u2f(e.y) <= bestT) return e.x;
    }
    return kEmptyD;
}

template <typename T>
__device__ __forceinline__ void cswapKey(uint32_t& ka, T& ra, uint32_t& kb, T& rb) {
    bool s = kb < ka;
    uint32_t kk = s ? kb : ka;
    kb = s ? ka : kb;
    ka = kk;
    T rr = s ? rb : ra;
    rb = s ? ra : rb;
    ra = rr;
}

// The hit children of a node (entry distances tc; refs rc; hit flags): for closest hit they
// are ordered nearest first and all but the nearest are pushed; returns the child to visit
// next, or kEmptyD when no child is hit.  The closest-hit order sorts keys: the bits of the
// entry distance (non-negative, so integer order is float order) for a hit, 0xFFFFFFFF for a
// miss, so misses sort last and a key tells by itself whether its child is pushed.
template <int W, bool ANY, typename Stack>
__device__ __forceinline__ uint32_t descend(float* tc, uint32_t* rc, const bool* hit, Stack& st) {
    bool anyHit = false;
#pragma unroll
    for (int k = 0; k < W; ++k) anyHit = anyHit || hit[k];
    if (!anyHit) return kEmptyD;
    if (ANY) {
        uint32_t first = kEmptyD;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            if (!hit[k]) continue;
            if (first == kEmptyD) first = rc[k];
            else st.push(rc[k], tc[k]);
        }
        return first;
    }
    uint32_t key[W];
#pragma unroll
    for (int k = 0; k < W; ++k) key[k] = hit[k] ? f2u(tc[k]) : 0xFFFFFFFFu;
    // Sorting network, ascending key.
    if (W == 4) {
        cswapKey(key[0], rc[0], key[1], rc[1]); cswapKey(key[2], rc[2], key[3], rc[3]);
        cswapKey(key[0], rc[0], key[2], rc[2]); cswapKey(key[1], rc[1], key[3], rc[3]);
        cswapKey(key[1], rc[1], key[2], rc[2]);
    } else {
        // Optimal 8-input network (19 comparators, depth 6).
        constexpr int net[19][2] = {{0, 2}, {1, 3}, {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6},
                                    {3, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}, {2, 4}, {3, 5},
                                    {1, 4}, {3, 6}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
        for (int c = 0; c < 19; ++c) cswapKey(key[net[c][0]], rc[net[c][0]], key[net[c][1]], rc[net[c][1]]);
    }
#pragma unroll
    for (int k = W - 1; k >= 1; --k)
        if (key[k] != 0xFFFFFFFFu) st.push(rc[k], u2f(key[k]));
    return rc[0];
}

// Watchdog: gives up on the ray (it keeps the closest hit found so far) and raises the fault
// flag the host reads at its next synchronizing call.
template <typename Stack>
__device__ __forceinline__ void abandonRay(const TraceArgs& a, Stack& st, uint32_t& ref, uint32_t& leaf) {
    if (a.fault) atomicOr(a.fault, 1u);
    ref = kEmptyD;
    leaf = kEmptyD;
    st.sp = 0;
}

// One round of the traversal loop of one ray ("while-while" with postponed leaves, after Aila
// & Laine 2009): a lane walks interior nodes until it meets a leaf, parks it, and keeps
// walking until every lane of the wave that still traverses holds a parked leaf; then the
// wave tests leaves together.  Lanes in a wave so spend their iterations on the same kind of
// work.  Parking a leaf can only add node visits (culling by the closest hit comes later),
// never change the result: the closest hit is the minimum over every primitive tested, and
// every entry that can still hold a hit at least as close is visited.  On return, ref ==
// kEmptyD (or an any-hit occlusion) means the ray is done.  `steps` counts the ray's node
// visits across rounds; past a.stepCap the ray is abandoned (abandonRay), checked every 16
// iterations of the node loop and once at the end of the round.
template <int W, bool Q, bool ANY, bool MIXED, typename Stack>
__device__ __forceinline__ void traceRound(const TraceArgs& a, const DRay& r, Stack& st, uint32_t& ref, uint32_t& leaf,
                                           float& bestT, uint32_t& bestId, bool& occluded, uint32_t& steps) {
    uint32_t it = 0u;
    while (ref != kEmptyD && !isLeafD(ref)) {
        ++steps;
        if ((++it & 15u) == 0u && __any(steps > a.stepCap ? 1 : 0)) {
            if (steps > a.stepCap) {
                abandonRay(a, st, ref, leaf);
                return;
            }
        }
        float tc[W];
        uint32_t rc[W];
        bool hit[W];
        NodeRows<W, Q> nr;
        loadNode<W, Q>(a.nodes, ref, nr);
        children<W, Q>(nr, r, bestT, tc, rc, hit);
        ref = descend<W, ANY>(tc, rc, hit, st);
        if (ref == kEmptyD) ref = popNext<ANY>(st, bestT);
        if (ref != kEmptyD && isLeafD(ref) && leaf == kEmptyD) {
            leaf = ref;
            ref = popNext<ANY>(st, bestT);
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
            ref = popNext<ANY>(st, bestT);
        }
    }
    if (steps > a.stepCap) abandonRay(a, st, ref, leaf);
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
        uint32_t ref = root, leaf = kEmptyD, steps = 0;
        do {
            traceRound<W, Q, ANY, MIXED>(a, r, st, ref, leaf, bestT, bestId, occluded, steps);
        } while (!((ANY && occluded) || ref == kEmptyD));
    }
    tracePlanes<ANY>(a, r, bestT, bestId, occluded);
}

template <bool ANY>
__device__ __forceinline__ void storeResult(const TraceArgs& a, uint32_t i, float bestT, uint32_t bestId, bool occ) {
    if (ANY) a.occ[i] = occ ? 1 : 0;
    else a.hits[i] = make_uint2(bestId, bestId == RBR_MISS_ID ? 0u : f2u(bestT));
}

}  // namespace tdev
}  // namespace rbr
