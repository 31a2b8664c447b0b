// Device build of the wide BVH: the same node and primitive layouts as the host builder
// (bvh.h), so every traversal kernel, the render arena and the host traversal use its output
// unchanged.  The reference builds a binary SAH tree serially on the host (bvh.cpp:47-209);
// this build trades some tree quality for build time:
//   1. primitive boxes, world and centroid bounds (one pass, block reductions + atomics)
//   2. 30-bit Morton codes of the centroids, sorted with their primitive index (radix sort)
//   3. the binary radix tree over the sorted codes, every internal node in parallel (Karras
//      2012, "Maximizing parallelism in the construction of BVHs, octrees and k-d trees");
//      equal codes are ordered by index
//   4. internal boxes bottom-up: one thread per leaf climbs while it is the second child to
//      arrive at a node (per-node arrival counter)
//   5. collapse to W-wide nodes, one level per launch, by the host's SAH-optimal dynamic
//      program (its cost tables are filled in step 4's climb): a node's slots and which of
//      them are leaves (subtrees of at most leaf_max primitives, contiguous in Morton order)
//      come from the tables; or, with RBR_BVH_GREEDY, a node keeps opening its largest-area
//      child until it has W and every subtree of at most leaf_max primitives is a leaf; child
//      boxes are padded as the host pads them
//   6. leaf-ordered primitive records (bvh.h), and for quantized layouts the 8-bit nodes by
//      the host's own quantizer (bvh_quant.h), so both produce the same bits from the same
//      float nodes
// With RBR_BVH_PLOC, step 3 is locally-ordered clustering instead (gbPloc* below).  The host
// copies of nodes and records are made on their first use (ensureHostBvh).
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "bvh.h"
#include "bvh_quant.h"
#include "context.h"
#include "prim_tests.h"

namespace rbr {

namespace {

struct GGeom {
    const float* tris;
    const uint32_t* triIds;
    const float* sphs;
    const uint32_t* sphIds;
    const float* quads;
    const uint32_t* quadIds;
    uint32_t ntri, nsph, nquad;
};

constexpr uint32_t kLeafTag = 0x80000000u;   // binary child code: leaf (sorted position)

__device__ __forceinline__ float mn(float a, float b) { return b < a ? b : a; }   // std::min
__device__ __forceinline__ float mx(float a, float b) { return a < b ? b : a; }   // std::max

// Primitive box with the host builder's arithmetic (bvh_build.cpp).
__device__ void primBox(const GGeom& g, uint32_t i, float lo[3], float hi[3]) {
    if (i < g.ntri) {
        const float* t = g.tris + 9 * size_t(i);
        for (int k = 0; k < 3; ++k) {
            lo[k] = mn(mn(t[k], t[3 + k]), t[6 + k]);
            hi[k] = mx(mx(t[k], t[3 + k]), t[6 + k]);
        }
    } else if (i < g.ntri + g.nsph) {
        const float* s = g.sphs + 4 * size_t(i - g.ntri);
        for (int k = 0; k < 3; ++k) {
            lo[k] = s[k] - s[3];
            hi[k] = s[k] + s[3];
        }
    } else {
        quadBounds(g.quads + 12 * size_t(i - g.ntri - g.nsph), lo, hi);
    }
}

// Order-preserving float <-> uint32 (for atomicMin / atomicMax on floats).
__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unkey(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// red[0..2] world lo, [3..5] world hi, [6..8] centroid lo, [9..11] centroid hi (keys).
__global__ __launch_bounds__(256) void gbBoundsKernel(GGeom g, uint32_t n, float4* __restrict__ boxLo,
                                                      float4* __restrict__ boxHi, uint32_t* red) {
    __shared__ uint32_t sh[12];
    if (threadIdx.x < 12) sh[threadIdx.x] = (threadIdx.x % 6) < 3 ? 0xFFFFFFFFu : 0u;
    __syncthreads();
    uint32_t loc[12];
    for (int k = 0; k < 12; ++k) loc[k] = (k % 6) < 3 ? 0xFFFFFFFFu : 0u;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        float lo[3], hi[3];
        primBox(g, i, lo, hi);
        boxLo[i] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        boxHi[i] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        for (int k = 0; k < 3; ++k) {
            const float c = 0.5f * (lo[k] + hi[k]);
            loc[k] = min(loc[k], fkey(lo[k]));
            loc[3 + k] = max(loc[3 + k], fkey(hi[k]));
            loc[6 + k] = min(loc[6 + k], fkey(c));
            loc[9 + k] = max(loc[9 + k], fkey(c));
        }
    }
    for (int k = 0; k < 12; ++k) {
        if ((k % 6) < 3) atomicMin(&sh[k], loc[k]);
        else atomicMax(&sh[k], loc[k]);
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        if ((threadIdx.x % 6) < 3) atomicMin(&red[threadIdx.x], sh[threadIdx.x]);
        else atomicMax(&red[threadIdx.x], sh[threadIdx.x]);
    }
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {   // 10 bits -> every third bit
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ __launch_bounds__(256) void gbMortonKernel(uint32_t n, const float4* __restrict__ boxLo,
                                                      const float4* __restrict__ boxHi, const uint32_t* red,
                                                      uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 lo = boxLo[i], hi = boxHi[i];
    const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        const float a = unkey(red[6 + k]), b = unkey(red[9 + k]);
        const float ext = b - a;
        float u = ext > 0.0f ? (c[k] - a) / ext : 0.5f;
        u = fminf(fmaxf(u, 0.0f), 1.0f);
        q[k] = min(uint32_t(u * 1024.0f), 1023u);
    }
    keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    vals[i] = i;
}

// Common-prefix length of sorted keys i and j (j out of range: -1); equal keys compare
// their positions.
__device__ __forceinline__ int delta(const uint32_t* keys, uint32_t n, int i, int j) {
    if (j < 0 || j >= int(n)) return -1;
    const uint32_t a = keys[i], b = keys[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz(uint32_t(i) ^ uint32_t(j));
}

// Internal node i of the radix tree over n sorted keys: its range [first, last], children
// (internal index, or kLeafTag | sorted position) and the parent links of both.
__global__ __launch_bounds__(256) void gbKarrasKernel(uint32_t n, const uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ left, uint32_t* __restrict__ right,
                                                      uint32_t* __restrict__ first, uint32_t* __restrict__ last,
                                                      uint32_t* __restrict__ parentInt,
                                                      uint32_t* __restrict__ parentLeaf) {
    const int i = int(blockIdx.x * 256 + threadIdx.x);
    if (i >= int(n) - 1) return;
    const int d = delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    for (int t = (l + 1) >> 1;; t = (t + 1) >> 1) {
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
    }
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const uint32_t lc = lo == gamma ? (kLeafTag | uint32_t(gamma)) : uint32_t(gamma);
    const uint32_t rc = hi == gamma + 1 ? (kLeafTag | uint32_t(gamma + 1)) : uint32_t(gamma + 1);
    left[i] = lc;
    right[i] = rc;
    first[i] = uint32_t(lo);
    last[i] = uint32_t(hi);
    if (lc & kLeafTag) parentLeaf[lc & ~kLeafTag] = uint32_t(i);
    else parentInt[lc] = uint32_t(i);
    if (rc & kLeafTag) parentLeaf[rc & ~kLeafTag] = uint32_t(i);
    else parentInt[rc] = uint32_t(i);
}

// ---- Locally-ordered clustering (after Meister & Bittner 2018, "Parallel Locally-Ordered
// Clustering for Bounding Volume Hierarchy Construction"): the sorted primitives start as
// clusters; every round, each cluster finds the cluster within kPlocRadius positions whose
// union box has the least surface area, mutual nearest neighbours merge (the left one keeps
// the place of the pair) and the array is compacted, until one cluster is left.  The result
// replaces the radix tree: the same arrays (children, parent links, leaf ranges) in the same
// encoding, with leaves renumbered to the depth-first order of the new tree so that every
// subtree still covers a contiguous range of positions.
constexpr int kPlocRadius = 16;

__device__ __forceinline__ float unionArea(float4 alo, float4 ahi, float4 blo, float4 bhi) {
    const float dx = mx(ahi.x, bhi.x) - mn(alo.x, blo.x);
    const float dy = mx(ahi.y, bhi.y) - mn(alo.y, blo.y);
    const float dz = mx(ahi.z, bhi.z) - mn(alo.z, blo.z);
    return dx * dy + dy * dz + dz * dx;
}

// Nearest neighbour of every cluster by the key (union area, lower index, higher index): a
// strict order on pairs, so the least pair of a round is mutual and every round merges.
__global__ __launch_bounds__(256) void gbPlocNearestKernel(uint32_t m, const float4* __restrict__ lo,
                                                           const float4* __restrict__ hi, int32_t* __restrict__ nn) {
    __shared__ float4 slo[256 + 2 * kPlocRadius], shi[256 + 2 * kPlocRadius];
    const int base = int(blockIdx.x * 256) - kPlocRadius;
    for (int t = threadIdx.x; t < 256 + 2 * kPlocRadius; t += 256) {
        const int k = base + t;
        if (k >= 0 && k < int(m)) {
            slo[t] = lo[k];
            shi[t] = hi[k];
        }
    }
    __syncthreads();
    const int i = int(blockIdx.x * 256 + threadIdx.x);
    if (i >= int(m)) return;
    const int c = int(threadIdx.x) + kPlocRadius;
    const float4 a = slo[c], b = shi[c];
    float best = INFINITY;
    int bj = -1;
    for (int d = -kPlocRadius; d <= kPlocRadius; ++d) {
        const int j = i + d;
        if (d == 0 || j < 0 || j >= int(m)) continue;
        const float ar = unionArea(a, b, slo[c + d], shi[c + d]);
        // for equal areas the pair with the lower (then higher) index wins; i is common to both
        const bool better = bj < 0 || ar < best ||
                            (ar == best && (min(i, j) < min(i, bj) || (min(i, j) == min(i, bj) && max(i, j) < max(i, bj))));
        if (better) {
            best = ar;
            bj = j;
        }
    }
    nn[i] = bj;
}

// Pairs (2k, 2k + 1): used for a round in which no pair was mutual (only possible with
// non-finite boxes), so the clustering always ends.
__global__ __launch_bounds__(256) void gbPlocPairKernel(uint32_t m, int32_t* __restrict__ nn) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const uint32_t j = i ^ 1u;
    nn[i] = j < m ? int32_t(j) : -1;
}

__device__ __forceinline__ uint32_t clusterCount(uint32_t code, const uint32_t* count) {
    return (code & kLeafTag) ? 1u : count[code];
}

// A mutual pair (i < j) becomes internal node (n - 2 - merge number), so the last merge, the
// root, is node 0; keep[i] says whether position i survives the round.
__global__ __launch_bounds__(256) void gbPlocMergeKernel(uint32_t m, uint32_t n, const uint32_t* __restrict__ C,
                                                         const float4* __restrict__ lo, const float4* __restrict__ hi,
                                                         const int32_t* __restrict__ nn, uint32_t* merges,
                                                         uint32_t* __restrict__ left, uint32_t* __restrict__ right,
                                                         uint32_t* __restrict__ parentInt, uint32_t* __restrict__ parentLeaf,
                                                         uint32_t* __restrict__ count, uint32_t* __restrict__ outC,
                                                         float4* __restrict__ outLo, float4* __restrict__ outHi,
                                                         uint32_t* __restrict__ keep) {
    const int i = int(blockIdx.x * 256 + threadIdx.x);
    if (i >= int(m)) return;
    const int j = nn[i];
    const bool mutual = j >= 0 && nn[j] == i;
    if (mutual && i > j) {
        keep[i] = 0u;
        return;
    }
    keep[i] = 1u;
    if (!mutual) {
        outC[i] = C[i];
        outLo[i] = lo[i];
        outHi[i] = hi[i];
        return;
    }
    const uint32_t a = C[i], b = C[j];
    const uint32_t id = (n - 2u) - atomicAdd(merges, 1u);
    left[id] = a;
    right[id] = b;
    if (a & kLeafTag) parentLeaf[a & ~kLeafTag] = id;
    else parentInt[a] = id;
    if (b & kLeafTag) parentLeaf[b & ~kLeafTag] = id;
    else parentInt[b] = id;
    count[id] = clusterCount(a, count) + clusterCount(b, count);
    outC[i] = id;
    const float4 la = lo[i], lb = lo[j], ha = hi[i], hb = hi[j];
    outLo[i] = make_float4(mn(la.x, lb.x), mn(la.y, lb.y), mn(la.z, lb.z), 0.0f);
    outHi[i] = make_float4(mx(ha.x, hb.x), mx(ha.y, hb.y), mx(ha.z, hb.z), 0.0f);
}

// Compaction of the survivors (pos: exclusive prefix sum of keep); the new count goes to
// *mOut.
__global__ __launch_bounds__(256) void gbPlocCompactKernel(uint32_t m, const uint32_t* __restrict__ keep,
                                                           const uint32_t* __restrict__ pos, const uint32_t* __restrict__ inC,
                                                           const float4* __restrict__ inLo, const float4* __restrict__ inHi,
                                                           uint32_t* __restrict__ C, float4* __restrict__ lo,
                                                           float4* __restrict__ hi, uint32_t* mOut) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    if (i == m - 1u) *mOut = pos[i] + keep[i];
    if (!keep[i]) return;
    const uint32_t k = pos[i];
    C[k] = inC[i];
    lo[k] = inLo[i];
    hi[k] = inHi[i];
}

// Depth-first position of the first leaf under a node: the sum, over the ancestors where the
// path goes right, of the left sibling's primitive count.
__device__ __forceinline__ uint32_t dfsOffset(uint32_t code, uint32_t p, const uint32_t* left, const uint32_t* right,
                                              const uint32_t* parentInt, const uint32_t* count) {
    uint32_t off = 0u;
    while (true) {
        if (right[p] == code) off += clusterCount(left[p], count);
        if (p == 0u) break;
        code = p;
        p = parentInt[p];
    }
    return off;
}

// Leaves: depth-first position of every sorted position; internal nodes: their leaf range.
__global__ __launch_bounds__(256) void gbPlocRangesKernel(uint32_t n, const uint32_t* __restrict__ left,
                                                          const uint32_t* __restrict__ right,
                                                          const uint32_t* __restrict__ parentInt,
                                                          const uint32_t* __restrict__ parentLeaf,
                                                          const uint32_t* __restrict__ count, uint32_t* __restrict__ leafPos,
                                                          uint32_t* __restrict__ first, uint32_t* __restrict__ last) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) leafPos[k] = dfsOffset(kLeafTag | k, parentLeaf[k], left, right, parentInt, count);
    if (k < n - 1u) {
        const uint32_t off = k == 0u ? 0u : dfsOffset(k, parentInt[k], left, right, parentInt, count);
        first[k] = off;
        last[k] = off + count[k] - 1u;
    }
}

// Renumbers the leaves to their depth-first positions: the primitive of each position, the
// parent links of the leaves and the leaf codes of the children.
__global__ __launch_bounds__(256) void gbPlocRenumberKernel(uint32_t n, const uint32_t* __restrict__ leafPos,
                                                            const uint32_t* __restrict__ valsIn,
                                                            const uint32_t* __restrict__ parentLeafIn,
                                                            uint32_t* __restrict__ valsOut,
                                                            uint32_t* __restrict__ parentLeafOut, uint32_t* left,
                                                            uint32_t* right) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) {
        const uint32_t q = leafPos[k];
        valsOut[q] = valsIn[k];
        parentLeafOut[q] = parentLeafIn[k];
    }
    if (k < n - 1u) {
        const uint32_t a = left[k], b = right[k];
        if (a & kLeafTag) left[k] = kLeafTag | leafPos[a & ~kLeafTag];
        if (b & kLeafTag) right[k] = kLeafTag | leafPos[b & ~kLeafTag];
    }
}

// Cluster boxes of the sorted primitives.
__global__ __launch_bounds__(256) void gbPlocInitKernel(uint32_t n, const uint32_t* __restrict__ vals,
                                                        const float4* __restrict__ boxLo, const float4* __restrict__ boxHi,
                                                        uint32_t* __restrict__ C, float4* __restrict__ lo,
                                                        float4* __restrict__ hi) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    C[k] = kLeafTag | k;
    lo[k] = boxLo[vals[k]];
    hi[k] = boxHi[vals[k]];
}

__device__ __forceinline__ float ld(const float* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void st(float* p, float v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Box of a binary child: a leaf is its primitive's box (sorted position -> primitive).
__device__ __forceinline__ void childBox(uint32_t c, const uint32_t* vals, const float4* boxLo, const float4* boxHi,
                                         const float* nodeBox, float lo[3], float hi[3]) {
    if (c & kLeafTag) {
        const uint32_t p = vals[c & ~kLeafTag];
        const float4 a = boxLo[p], b = boxHi[p];
        lo[0] = a.x; lo[1] = a.y; lo[2] = a.z;
        hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
    } else {
        const float* q = nodeBox + 6 * size_t(c);
        for (int k = 0; k < 3; ++k) {
            lo[k] = ld(q + k);
            hi[k] = ld(q + 3 + k);
        }
    }
}

__device__ __forceinline__ float boxArea(const float lo[3], const float hi[3]) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0.0f || dy < 0.0f || dz < 0.0f) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// SAH-optimal collapse (the host builder's dynamic program, bvh_build.cpp), evaluated in the
// same bottom-up pass as the boxes.  Per internal node: cost[W] (least SAH cost of the
// subtree when it may fill at most i slots of its wide parent) and one decision word:
//   bit i-1 (i = 1): 1 = one leaf, 0 = one wide node;  (i > 1): 1 = split over both
//   children, 0 = use at most i - 1 slots;  bits 8 + 3 (j - 2): slots of the left child
//   when j slots are split.
// A binary leaf (one primitive) costs A c_prim for any number of slots.
struct DpArgs {
    uint32_t W, leafMax;
    float cPrim;
    const uint32_t *first, *last;
    float* cost;   // [internal node][W]
    float* dec;    // [internal node]: decision word (bits in a float, stored like the boxes)
};

__device__ __forceinline__ void childCosts(uint32_t c, const uint32_t* vals, const float4* boxLo,
                                           const float4* boxHi, const float* nodeBox, const DpArgs& dp, float* out) {
    if (c & kLeafTag) {
        float lo[3], hi[3];
        childBox(c, vals, boxLo, boxHi, nodeBox, lo, hi);
        const float v = boxArea(lo, hi) * dp.cPrim;
        for (uint32_t i = 0; i < dp.W; ++i) out[i] = v;
    } else {
        for (uint32_t i = 0; i < dp.W; ++i) out[i] = ld(dp.cost + size_t(c) * dp.W + i);
    }
}

// Internal boxes (and the collapse's cost tables) bottom-up.  Boxes and tables are written and
// read only with device-scope atomic stores and loads (coherent across the XCDs' L2 caches by
// themselves), so ordering them around the arrival counter needs no cache maintenance: a wait
// for this thread's earlier stores before its arrival, and a wait for the arrival's result
// before any read, so the second arrival at a node sees what its sibling's thread wrote.  (An
// acquire-release arrival adds a whole-L2 writeback and invalidation per node visit, for data
// no plain access touches: the pass took 3.8 ms on a 1M-triangle mesh.)
__device__ __forceinline__ uint32_t arrive(uint32_t* counter) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    const uint32_t prior = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    return prior;
}

__global__ __launch_bounds__(256) void gbRefitKernel(uint32_t n, const uint32_t* __restrict__ vals,
                                                     const float4* __restrict__ boxLo, const float4* __restrict__ boxHi,
                                                     const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
                                                     const uint32_t* __restrict__ parentInt,
                                                     const uint32_t* __restrict__ parentLeaf, uint32_t* arrivals,
                                                     float* nodeBox, DpArgs dp) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n || n < 2) return;
    uint32_t p = parentLeaf[i];
    while (true) {
        const uint32_t prior = arrive(&arrivals[p]);
        if (prior == 0u) return;   // the sibling's thread finishes this node
        float alo[3], ahi[3], blo[3], bhi[3];
        childBox(left[p], vals, boxLo, boxHi, nodeBox, alo, ahi);
        childBox(right[p], vals, boxLo, boxHi, nodeBox, blo, bhi);
        float* q = nodeBox + 6 * size_t(p);
        float lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = mn(alo[k], blo[k]);
            hi[k] = mx(ahi[k], bhi[k]);
            st(q + k, lo[k]);
            st(q + 3 + k, hi[k]);
        }
        float cl[8], cr[8], dist[9];
        childCosts(left[p], vals, boxLo, boxHi, nodeBox, dp, cl);
        childCosts(right[p], vals, boxLo, boxHi, nodeBox, dp, cr);
        uint32_t dec = 0u;
        for (uint32_t j = 2; j <= dp.W; ++j) {
            float best = INFINITY;
            uint32_t bk = 1;
            for (uint32_t k = 1; k < j; ++k) {
                const float v = cl[k - 1] + cr[j - k - 1];
                if (v < best) { best = v; bk = k; }
            }
            dist[j] = best;
            dec |= bk << (8 + 3 * (j - 2));
        }
        const float a = boxArea(lo, hi);
        const uint32_t cnt = dp.last[p] - dp.first[p] + 1u;
        const float leafC = cnt <= dp.leafMax ? a * float(cnt) * dp.cPrim : INFINITY;
        const float wideC = a + dist[dp.W];   // one node visit costs 1
        float c = leafC <= wideC ? leafC : wideC;
        if (leafC <= wideC) dec |= 1u;
        float* co = dp.cost + size_t(p) * dp.W;
        st(co, c);
        for (uint32_t j = 2; j <= dp.W; ++j) {
            if (dist[j] < c) { c = dist[j]; dec |= 1u << (j - 1); }
            st(co + j - 1, c);
        }
        st(dp.dec + p, __uint_as_float(dec));
        if (p == 0u) return;
        p = parentInt[p];
    }
}

struct CollapseArgs {
    uint32_t n, W, leafMax;
    const uint32_t *vals, *left, *right, *first, *last;
    const float4 *boxLo, *boxHi;
    const float* nodeBox;
    const float* dec;       // decision words of the optimal collapse (nullptr: greedy)
    const uint32_t* red;
    float* nodes;           // float wide nodes (bvh.h), f4PerNode * 4 floats each
    uint32_t* counters;     // [0] wide nodes, [1] leaves, [2] next-level items
};

__device__ __forceinline__ uint32_t codeCount(const CollapseArgs& a, uint32_t c) {
    return (c & kLeafTag) ? 1u : a.last[c] - a.first[c] + 1u;
}
__device__ __forceinline__ uint32_t codeFirst(const CollapseArgs& a, uint32_t c) {
    return (c & kLeafTag) ? (c & ~kLeafTag) : a.first[c];
}

// One level of the collapse: item k = (binary node code, wide node index).
__global__ __launch_bounds__(256) void gbCollapseKernel(CollapseArgs a, const uint2* __restrict__ in, uint32_t nIn,
                                                        uint2* __restrict__ out) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nIn) return;
    const uint2 item = in[k];
    const uint32_t W = a.W;
    uint32_t kids[8];
    bool kidLeaf[8];
    uint32_t nk = 0;
    const auto decOf = [&](uint32_t c) { return __float_as_uint(a.dec[c]); };
    if (item.x & kLeafTag) {
        kids[nk] = item.x;   // a single primitive
        kidLeaf[nk++] = true;
    } else if (a.dec) {
        // slots from the decision words: (binary node, slots) pairs, at most W pending
        if (decOf(item.x) & 1u) {
            kids[nk] = item.x;   // a root best kept as one leaf
            kidLeaf[nk++] = true;
        } else {
            uint32_t sc[8], ss[8], ns = 0;
            const uint32_t k0 = (decOf(item.x) >> (8 + 3 * (W - 2))) & 7u;
            sc[ns] = a.left[item.x]; ss[ns++] = k0;
            sc[ns] = a.right[item.x]; ss[ns++] = W - k0;
            while (ns > 0 && nk < W) {
                --ns;
                const uint32_t c = sc[ns];
                uint32_t slots = ss[ns];
                if (c & kLeafTag) {
                    kids[nk] = c;
                    kidLeaf[nk++] = true;
                    continue;
                }
                const uint32_t d = decOf(c);
                while (slots > 1u && !((d >> (slots - 1u)) & 1u)) --slots;
                if (slots == 1u) {
                    kids[nk] = c;
                    kidLeaf[nk++] = (d & 1u) != 0u;
                    continue;
                }
                const uint32_t k = (d >> (8 + 3 * (slots - 2u))) & 7u;
                sc[ns] = a.left[c]; ss[ns++] = k;
                sc[ns] = a.right[c]; ss[ns++] = slots - k;
            }
        }
    } else if (codeCount(a, item.x) <= a.leafMax) {
        kids[nk] = item.x;   // a root small enough to be one leaf
        kidLeaf[nk++] = true;
    } else {
        kids[nk++] = a.left[item.x];
        kids[nk++] = a.right[item.x];
        while (nk < W) {
            int best = -1;
            float bestA = -1.0f;
            for (uint32_t i = 0; i < nk; ++i) {
                const uint32_t c = kids[i];
                if ((c & kLeafTag) || codeCount(a, c) <= a.leafMax) continue;
                float lo[3], hi[3];
                childBox(c, a.vals, a.boxLo, a.boxHi, a.nodeBox, lo, hi);
                const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
                const float ar = 2.0f * (dx * dy + dy * dz + dz * dx);
                if (ar > bestA) {
                    bestA = ar;
                    best = int(i);
                }
            }
            if (best < 0) break;
            const uint32_t open = kids[best];
            kids[best] = a.left[open];
            kids[nk++] = a.right[open];
        }
        for (uint32_t i = 0; i < nk; ++i) kidLeaf[i] = (kids[i] & kLeafTag) || codeCount(a, kids[i]) <= a.leafMax;
    }
    // padding of the host collapse: pad + |x| * 1e-6 outward, pad = 1e-5 * extent + 1e-6
    float ext = 0.0f;
    for (int c = 0; c < 3; ++c) ext = fmaxf(ext, unkey(a.red[3 + c]) - unkey(a.red[c]));
    const float pad = 1e-5f * ext + 1e-6f;
    float* f = a.nodes + size_t(item.y) * (W == 4 ? 32 : 64);
    for (uint32_t i = 0; i < W; ++i) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        uint32_t ref = kEmpty;
        if (i < nk) {
            const uint32_t c = kids[i];
            float blo[3], bhi[3];
            childBox(c, a.vals, a.boxLo, a.boxHi, a.nodeBox, blo, bhi);
            for (int q = 0; q < 3; ++q) {
                lo[q] = blo[q] - (pad + fabsf(blo[q]) * 1e-6f);
                hi[q] = bhi[q] + (pad + fabsf(bhi[q]) * 1e-6f);
            }
            const uint32_t cnt = codeCount(a, c);
            if (kidLeaf[i]) {
                ref = kLeafBit | ((cnt - 1u) << kCountShift) | codeFirst(a, c);   // bvh.h leaf encoding
                atomicAdd(&a.counters[1], 1u);
            } else {
                ref = atomicAdd(&a.counters[0], 1u);
                out[atomicAdd(&a.counters[2], 1u)] = make_uint2(c, ref);
            }
        }
        f[0 * W + i] = lo[0];
        f[1 * W + i] = hi[0];
        f[2 * W + i] = lo[1];
        f[3 * W + i] = hi[1];
        f[4 * W + i] = lo[2];
        f[5 * W + i] = hi[2];
        f[6 * W + i] = __uint_as_float(ref);
    }
    for (uint32_t i = 7 * W; i < (W == 4 ? 32u : 64u); ++i) f[i] = 0.0f;
}

// Leaf-ordered primitive records (bvh.h), as the host builder writes them.
__global__ __launch_bounds__(256) void gbRecordsKernel(GGeom g, uint32_t n, const uint32_t* __restrict__ vals,
                                                       float4* __restrict__ rec) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = vals[i];
    float4* r = rec + 3 * size_t(i);
    if (p < g.ntri) {
        const float* t = g.tris + 9 * size_t(p);
        r[0] = make_float4(t[0], t[1], t[2], __uint_as_float(g.triIds[p]));
        r[1] = make_float4(t[3] - t[0], t[4] - t[1], t[5] - t[2], __uint_as_float(KIND_TRI));
        r[2] = make_float4(t[6] - t[0], t[7] - t[1], t[8] - t[2], 0.0f);
    } else if (p < g.ntri + g.nsph) {
        const uint32_t q = p - g.ntri;
        const float* s = g.sphs + 4 * size_t(q);
        r[0] = make_float4(s[0], s[1], s[2], __uint_as_float(g.sphIds[q]));
        r[1] = make_float4(s[3], 0.0f, 0.0f, __uint_as_float(KIND_SPHERE));
        r[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else {
        const uint32_t q = p - g.ntri - g.nsph;
        const float* s = g.quads + 12 * size_t(q);
        r[0] = make_float4(s[0], s[1], s[2], __uint_as_float(g.quadIds[q]));
        r[1] = make_float4(s[3], s[4], s[5], __uint_as_float(KIND_QUAD));
        r[2] = make_float4(s[6], s[7], s[8], __uint_as_float(q));
    }
}

__global__ __launch_bounds__(256) void gbQuantizeKernel(uint32_t count, uint32_t W, const float* __restrict__ nodes,
                                                        uint32_t* __restrict__ q) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const uint32_t dw = W == 4 ? 16u : 32u;
    uint32_t tmp[32];
    for (uint32_t k = 0; k < dw; ++k) tmp[k] = 0u;
    quantizeNode(nodes + size_t(i) * (W == 4 ? 32 : 64), int(W), tmp);
    for (uint32_t k = 0; k < dw; ++k) q[size_t(i) * dw + k] = tmp[k];
}

// Reduction seeds, counters (wide node 0 is the root) and the root item: binary node 0, or
// the single primitive when there is only one.
__global__ void gbInitKernel(uint32_t* red, uint32_t* counters, uint2* items, uint32_t rootCode) {
    const uint32_t t = threadIdx.x;
    if (t < 12) red[t] = (t % 6) < 3 ? 0xFFFFFFFFu : 0u;
    if (t < 3) counters[t] = t == 0 ? 1u : 0u;
    if (t == 0) items[0] = make_uint2(rootCode, 0u);
}

__global__ __launch_bounds__(256) void gbZeroKernel(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0u;
}

float hostUnkey(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

}  // namespace

// The binary tree by locally-ordered clustering (gbPloc* kernels) over the n >= 2 sorted
// primitives (vals: sorted position -> primitive): fills the same arrays as the radix tree
// (children, leaf ranges, parent links, with leaves at their depth-first positions) and
// dVals3, the primitive of each depth-first position.  One round per launch group; the host
// reads the cluster count after each.
static int buildClusterTree(rbr_ctx* ctx, uint32_t n, const uint32_t* vals, const float4* boxLo, const float4* boxHi,
                            DevBuf& dC0, DevBuf& dC1, DevBuf& dL0, DevBuf& dL1, DevBuf& dH0, DevBuf& dH1, DevBuf& dNn,
                            DevBuf& dKeep, DevBuf& dPos, DevBuf& dCount, DevBuf& dScanTmp, DevBuf& dLeafPos,
                            DevBuf& dVals3, DevBuf& dPLeaf2, DevBuf& dCnt, DevBuf& dLeft, DevBuf& dRight, DevBuf& dFirst,
                            DevBuf& dLast, DevBuf& dPInt, DevBuf& dPLeaf) {
    hipStream_t st = ctx->stream;
    int rc = RBR_OK;
    auto chk = [&](int r) {
        if (r && !rc) rc = r;
        return rc == RBR_OK;
    };
    size_t scanBytes = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, static_cast<uint32_t*>(nullptr),
                                                    static_cast<uint32_t*>(nullptr), int(n), st);
    if (e != hipSuccess) return failHip(ctx, e, "cluster scan size");
    const size_t u = size_t(n) * 4, f = size_t(n) * 16;
    if (!chk(ensureBuffer(ctx, dC0, u)) || !chk(ensureBuffer(ctx, dC1, u)) || !chk(ensureBuffer(ctx, dL0, f)) ||
        !chk(ensureBuffer(ctx, dL1, f)) || !chk(ensureBuffer(ctx, dH0, f)) || !chk(ensureBuffer(ctx, dH1, f)) ||
        !chk(ensureBuffer(ctx, dNn, u)) || !chk(ensureBuffer(ctx, dKeep, u)) || !chk(ensureBuffer(ctx, dPos, u)) ||
        !chk(ensureBuffer(ctx, dCount, u)) || !chk(ensureBuffer(ctx, dScanTmp, std::max<size_t>(scanBytes, 16))) ||
        !chk(ensureBuffer(ctx, dLeafPos, u)) || !chk(ensureBuffer(ctx, dVals3, u)) || !chk(ensureBuffer(ctx, dPLeaf2, u)))
        return rc;
    auto U = [](DevBuf& b) { return static_cast<uint32_t*>(b.ptr); };
    auto F = [](DevBuf& b) { return static_cast<float4*>(b.ptr); };
    const unsigned blocks = (n + 255) / 256;
    hipLaunchKernelGGL(gbPlocInitKernel, dim3(blocks), dim3(256), 0, st, n, vals, boxLo, boxHi, U(dC0), F(dL0), F(dH0));
    // dCnt[3]: merges so far, dCnt[2]: cluster count after a round (the collapse resets both)
    hipLaunchKernelGGL(gbZeroKernel, dim3(1), dim3(256), 0, st, U(dCnt) + 2, 2u);
    uint32_t m = n, stalled = 0, rounds = 0;
    while (m > 1) {
        if (++rounds > 4096) return fail(ctx, RBR_ERR_UNSUPPORTED, "cluster build did not converge");
        const unsigned b = (m + 255) / 256;
        if (stalled) hipLaunchKernelGGL(gbPlocPairKernel, dim3(b), dim3(256), 0, st, m, reinterpret_cast<int32_t*>(U(dNn)));
        else hipLaunchKernelGGL(gbPlocNearestKernel, dim3(b), dim3(256), 0, st, m, F(dL0), F(dH0),
                                reinterpret_cast<int32_t*>(U(dNn)));
        hipLaunchKernelGGL(gbPlocMergeKernel, dim3(b), dim3(256), 0, st, m, n, U(dC0), F(dL0), F(dH0),
                           reinterpret_cast<const int32_t*>(U(dNn)), U(dCnt) + 3, U(dLeft), U(dRight), U(dPInt),
                           U(dPLeaf2), U(dCount), U(dC1), F(dL1), F(dH1), U(dKeep));
        e = hipcub::DeviceScan::ExclusiveSum(dScanTmp.ptr, scanBytes, U(dKeep), U(dPos), int(m), st);
        if (e != hipSuccess) return failHip(ctx, e, "cluster scan");
        hipLaunchKernelGGL(gbPlocCompactKernel, dim3(b), dim3(256), 0, st, m, U(dKeep), U(dPos), U(dC1), F(dL1), F(dH1),
                           U(dC0), F(dL0), F(dH0), U(dCnt) + 2);
        uint32_t next = 0;
        e = hipMemcpyAsync(&next, U(dCnt) + 2, 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return failHip(ctx, e, "cluster round");
        stalled = next == m ? 1u : 0u;
        m = next;
    }
    hipLaunchKernelGGL(gbPlocRangesKernel, dim3(blocks), dim3(256), 0, st, n, U(dLeft), U(dRight), U(dPInt), U(dPLeaf2),
                       U(dCount), U(dLeafPos), U(dFirst), U(dLast));
    hipLaunchKernelGGL(gbPlocRenumberKernel, dim3(blocks), dim3(256), 0, st, n, U(dLeafPos), vals, U(dPLeaf2),
                       U(dVals3), U(dPLeaf), U(dLeft), U(dRight));
    hipLaunchKernelGGL(gbZeroKernel, dim3(1), dim3(256), 0, st, U(dCnt) + 2, 2u);
    e = hipGetLastError();
    return e == hipSuccess ? RBR_OK : failHip(ctx, e, "cluster build");
}

int buildWideBVHDevice(rbr_ctx* ctx, rbr_scene* s, int width, int leafMax, bool quant, bool optimal, bool ploc) {
    using Clock = std::chrono::steady_clock;
    const auto t0 = Clock::now();
    const HostGeometry& geo = s->geom;
    WideBVH& out = s->bvh;
    out = WideBVH();
    out.width = width;
    const uint32_t ntri = uint32_t(geo.triIds.size()), nsph = uint32_t(geo.sphIds.size()),
                   nquad = uint32_t(geo.quadIds.size());
    const uint32_t n = ntri + nsph + nquad;
    out.primCount = n;
    int rc = RBR_OK;
    if (n == 0) return rc;
    if (n >= kStartMask) return fail(ctx, RBR_ERR_UNSUPPORTED, "too many primitives for the leaf encoding");
    hipStream_t st0 = ctx->stream;
    // Device temporaries: the context's, grown as needed and kept for the next build.
    std::vector<DevBuf>& tmp = ctx->buildTmp;
    if (tmp.size() < 42) tmp.resize(42);
    DevBuf &dTri = tmp[0], &dTriId = tmp[1], &dSph = tmp[2], &dSphId = tmp[3], &dQuad = tmp[4], &dQuadId = tmp[5],
           &dLo = tmp[6], &dHi = tmp[7], &dRed = tmp[8], &dKeys = tmp[9], &dVals = tmp[10], &dKeys2 = tmp[11],
           &dVals2 = tmp[12], &dSortTmp = tmp[13], &dLeft = tmp[14], &dRight = tmp[15], &dFirst = tmp[16],
           &dLast = tmp[17], &dPInt = tmp[18], &dPLeaf = tmp[19], &dArr = tmp[20], &dNodeBox = tmp[21],
           &dItemsA = tmp[22], &dItemsB = tmp[23], &dCnt = tmp[24], &dCost = tmp[25], &dDec = tmp[26],
           &dFloatNodes = tmp[27];
    // clustering (ploc): two cluster arrays, nearest neighbours, survivors, ranks, counts
    DevBuf &dC0 = tmp[28], &dC1 = tmp[29], &dL0 = tmp[30], &dL1 = tmp[31], &dH0 = tmp[32], &dH1 = tmp[33],
           &dNn = tmp[34], &dKeep = tmp[35], &dPos = tmp[36], &dCount = tmp[37], &dScanTmp = tmp[38],
           &dLeafPos = tmp[39], &dVals3 = tmp[40], &dPLeaf2 = tmp[41];
    auto chk = [&](int r) {
        if (r && !rc) rc = r;
        return rc == RBR_OK;
    };
    const uint32_t ni = std::max(1u, n - 1);
    const int W = width;
    const size_t f4n = size_t(W == 4 ? 8 : 16);
    const size_t maxNodes = std::max<size_t>(1, n);
    if (!chk(uploadBuffer(ctx, dTri, geo.tris.data(), geo.tris.size() * 4)) ||
        !chk(uploadBuffer(ctx, dTriId, geo.triIds.data(), geo.triIds.size() * 4)) ||
        !chk(uploadBuffer(ctx, dSph, geo.sphs.data(), geo.sphs.size() * 4)) ||
        !chk(uploadBuffer(ctx, dSphId, geo.sphIds.data(), geo.sphIds.size() * 4)) ||
        !chk(uploadBuffer(ctx, dQuad, geo.quads.data(), geo.quads.size() * 4)) ||
        !chk(uploadBuffer(ctx, dQuadId, geo.quadIds.data(), geo.quadIds.size() * 4)) ||
        !chk(ensureBuffer(ctx, dLo, size_t(n) * 16)) || !chk(ensureBuffer(ctx, dHi, size_t(n) * 16)) ||
        !chk(ensureBuffer(ctx, dRed, 12 * 4)) || !chk(ensureBuffer(ctx, dKeys, size_t(n) * 4)) ||
        !chk(ensureBuffer(ctx, dVals, size_t(n) * 4)) || !chk(ensureBuffer(ctx, dKeys2, size_t(n) * 4)) ||
        !chk(ensureBuffer(ctx, dVals2, size_t(n) * 4)) || !chk(ensureBuffer(ctx, dLeft, size_t(ni) * 4)) ||
        !chk(ensureBuffer(ctx, dRight, size_t(ni) * 4)) || !chk(ensureBuffer(ctx, dFirst, size_t(ni) * 4)) ||
        !chk(ensureBuffer(ctx, dLast, size_t(ni) * 4)) || !chk(ensureBuffer(ctx, dPInt, size_t(ni) * 4)) ||
        !chk(ensureBuffer(ctx, dPLeaf, size_t(n) * 4)) || !chk(ensureBuffer(ctx, dArr, size_t(ni) * 4)) ||
        !chk(ensureBuffer(ctx, dNodeBox, size_t(ni) * 24)) || !chk(ensureBuffer(ctx, dItemsA, maxNodes * 8)) ||
        !chk(ensureBuffer(ctx, dItemsB, maxNodes * 8)) || !chk(ensureBuffer(ctx, dCnt, 16)) ||
        !chk(ensureBuffer(ctx, dFloatNodes, maxNodes * f4n * 16)) ||
        !chk(ensureBuffer(ctx, dCost, size_t(ni) * W * 4)) || !chk(ensureBuffer(ctx, dDec, size_t(ni) * 4)) ||
        !chk(ensureBuffer(ctx, s->dPrims, size_t(n) * 48))) {
        return rc;
    }
    GGeom g{static_cast<const float*>(dTri.ptr), static_cast<const uint32_t*>(dTriId.ptr),
            static_cast<const float*>(dSph.ptr), static_cast<const uint32_t*>(dSphId.ptr),
            static_cast<const float*>(dQuad.ptr), static_cast<const uint32_t*>(dQuadId.ptr), ntri, nsph, nquad};
    auto u32 = [](DevBuf& b) { return static_cast<uint32_t*>(b.ptr); };
    auto f4 = [](DevBuf& b) { return static_cast<float4*>(b.ptr); };
    const unsigned blocks = (n + 255) / 256;
    hipLaunchKernelGGL(gbInitKernel, dim3(1), dim3(64), 0, st0, u32(dRed), u32(dCnt), static_cast<uint2*>(dItemsA.ptr),
                       n == 1 ? kLeafTag : 0u);
    hipLaunchKernelGGL(gbBoundsKernel, dim3(std::min(blocks, 2048u)), dim3(256), 0, st0, g, n, f4(dLo), f4(dHi),
                       u32(dRed));
    hipLaunchKernelGGL(gbMortonKernel, dim3(blocks), dim3(256), 0, st0, n, f4(dLo), f4(dHi), u32(dRed), u32(dKeys),
                       u32(dVals));
    size_t tmpBytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, tmpBytes, u32(dKeys), u32(dKeys2), u32(dVals),
                                                      u32(dVals2), int(n), 0, 30, st0);
    if (e != hipSuccess || !chk(ensureBuffer(ctx, dSortTmp, std::max<size_t>(tmpBytes, 16)))) {
        return rc ? rc : failHip(ctx, e, "radix sort size");
    }
    e = hipcub::DeviceRadixSort::SortPairs(dSortTmp.ptr, tmpBytes, u32(dKeys), u32(dKeys2), u32(dVals), u32(dVals2),
                                           int(n), 0, 30, st0);
    if (e != hipSuccess) {
        return failHip(ctx, e, "radix sort");
    }
    const uint32_t* keys = u32(dKeys2);
    const uint32_t* vals = u32(dVals2);
    if (n >= 2 && ploc) {
        if ((rc = buildClusterTree(ctx, n, u32(dVals2), f4(dLo), f4(dHi), dC0, dC1, dL0, dL1, dH0, dH1, dNn, dKeep, dPos,
                                   dCount, dScanTmp, dLeafPos, dVals3, dPLeaf2, dCnt, dLeft, dRight, dFirst, dLast, dPInt,
                                   dPLeaf)))
            return rc;
        vals = u32(dVals3);
    } else if (n >= 2) {
        hipLaunchKernelGGL(gbKarrasKernel, dim3((n - 1 + 255) / 256), dim3(256), 0, st0, n, keys, u32(dLeft), u32(dRight),
                           u32(dFirst), u32(dLast), u32(dPInt), u32(dPLeaf));
    }
    if (n >= 2) {
        hipLaunchKernelGGL(gbZeroKernel, dim3((n - 1 + 255) / 256), dim3(256), 0, st0, u32(dArr), n - 1);
        const DpArgs dp{uint32_t(W), uint32_t(leafMax), kCollapsePrimCost, u32(dFirst), u32(dLast),
                        static_cast<float*>(dCost.ptr), static_cast<float*>(dDec.ptr)};
        hipLaunchKernelGGL(gbRefitKernel, dim3(blocks), dim3(256), 0, st0, n, vals, f4(dLo), f4(dHi), u32(dLeft),
                           u32(dRight), u32(dPInt), u32(dPLeaf), u32(dArr), static_cast<float*>(dNodeBox.ptr), dp);
    }
    CollapseArgs ca{n, uint32_t(W), uint32_t(leafMax), vals, u32(dLeft), u32(dRight), u32(dFirst), u32(dLast),
                    f4(dLo), f4(dHi), static_cast<const float*>(dNodeBox.ptr),
                    optimal && n >= 2 ? static_cast<const float*>(dDec.ptr) : nullptr, u32(dRed),
                    static_cast<float*>(dFloatNodes.ptr), u32(dCnt)};
    uint32_t nIn = 1, depth = 0;
    uint2* itIn = static_cast<uint2*>(dItemsA.ptr);
    uint2* itOut = static_cast<uint2*>(dItemsB.ptr);
    uint32_t cnt[3] = {1, 0, 0};
    while (nIn > 0) {
        ++depth;
        hipLaunchKernelGGL(gbCollapseKernel, dim3((nIn + 255) / 256), dim3(256), 0, st0, ca, itIn, nIn, itOut);
        e = hipMemcpyAsync(cnt, dCnt.ptr, 12, hipMemcpyDeviceToHost, st0);
        if (e == hipSuccess) e = hipStreamSynchronize(st0);
        if (e != hipSuccess) {
            return failHip(ctx, e, "collapse level");
        }
        nIn = cnt[2];
        hipLaunchKernelGGL(gbZeroKernel, dim3(1), dim3(256), 0, st0, u32(dCnt) + 2, 1u);
        std::swap(itIn, itOut);
    }
    out.nodeCount = cnt[0];
    out.leafCount = cnt[1];
    out.maxDepth = depth;
    hipLaunchKernelGGL(gbRecordsKernel, dim3(blocks), dim3(256), 0, st0, g, n, vals, static_cast<float4*>(s->dPrims.ptr));
    const size_t floatBytes = size_t(out.nodeCount) * f4n * 16;
    // The scene keeps its device nodes and, for a quantized tree, the float nodes too (the
    // source of its host copies, made on first use).
    if (quant) {
        const size_t qBytes = size_t(out.nodeCount) * qNodeDwords(W) * 4;
        if (!chk(ensureBuffer(ctx, s->dNodes, qBytes)) || !chk(ensureBuffer(ctx, s->dFloatNodes, floatBytes))) return rc;
        hipLaunchKernelGGL(gbQuantizeKernel, dim3((out.nodeCount + 255) / 256), dim3(256), 0, st0, out.nodeCount,
                           uint32_t(W), static_cast<const float*>(dFloatNodes.ptr), u32(s->dNodes));
        e = hipMemcpyAsync(s->dFloatNodes.ptr, dFloatNodes.ptr, floatBytes, hipMemcpyDeviceToDevice, st0);
    } else {
        if (!chk(ensureBuffer(ctx, s->dNodes, floatBytes))) return rc;
        freeBuffer(s->dFloatNodes);
        e = hipMemcpyAsync(s->dNodes.ptr, dFloatNodes.ptr, floatBytes, hipMemcpyDeviceToDevice, st0);
    }
    // Only the bounds come back now; the host copies of nodes and records are made on first use.
    uint32_t red[12];
    if (e == hipSuccess) e = hipMemcpyAsync(red, dRed.ptr, 48, hipMemcpyDeviceToHost, st0);
    if (e == hipSuccess) e = hipStreamSynchronize(st0);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return failHip(ctx, e, "device BVH build");
    out.quantized = quant;
    for (int k = 0; k < 3; ++k) {
        out.worldLo[k] = hostUnkey(red[k]);
        out.worldHi[k] = hostUnkey(red[3 + k]);
    }
    out.quadNormals.resize(nquad);
    for (uint32_t i = 0; i < nquad; ++i) {
        const float* q = &geo.quads[12 * size_t(i)];
        out.quadNormals[i] = {q[9], q[10], q[11], 0.0f};
    }
    out.sahCost = 0.0;   // the binary tree stays on the device
    out.buildMs = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    return RBR_OK;
}

}  // namespace rbr
