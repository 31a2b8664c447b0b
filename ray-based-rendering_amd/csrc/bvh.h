// Wide BVH layout shared by the host builder, the host traversal and the HIP kernels.
//
// Replaces the reference's pointer-linked binary BVH (include/render/acceleration/bvh.h:
// 16-25, unique_ptr children, one virtual Primitive per leaf slot) with flat arrays that
// live in HBM:
//
//   nodes  : W-wide nodes, child boxes stored SoA inside the node so a node is one
//            (W=4: 128 B) or two (W=8: 256 B) aligned cache lines:
//              float lo_x[W], hi_x[W], lo_y[W], hi_y[W], lo_z[W], hi_z[W];
//              uint32 child[W];  (+ padding to 128 B / 256 B)
//            child encodings:
//              interior : node index                                (bit 31 = 0)
//              leaf     : 0x80000000 | (count-1) << 27 | first_prim (count 1..8)
//              empty    : 0xFFFFFFFF, with an inverted box (lo=+inf, hi=-inf)
//   prims  : leaf-ordered 48-byte records (3 x float4):
//              triangle : (v0, id) (e1, kind=0) (e2, 0)      e1 = v1-v0, e2 = v2-v0
//              sphere   : (center, id) (radius, 0, 0, kind=1) (0)
//              quad     : (origin, id) (u, kind=2) (v, quad_index)
//            e1/e2 are computed on the host with the same FP32 subtraction the reference
//            does per test (triangle.cpp:7-8), so the device sees identical values.
//   quad_n : per-quad normal (float4), planes : (normal, d) + ids, tested linearly.
//
// Child boxes are grown outward by a margin (scene extent * 1e-5 plus 1e-6 relative per
// coordinate) so that the approximate slab test can never cull a primitive whose exact
// Moller-Trumbore test accepts the ray.
//
// Quantized nodes (device layout when built with kQuantFlag; W = 4: 64 B, W = 8: 128 B):
//   dword 0-2  node origin p (biased, see below; float)
//   dword 3    W = 8: biased exponents ex | ey << 8 | ez << 16: per-axis scale s = 2^(e - 127);
//              W = 4: the x scale s_x as a float (bits ex << 23)
//   then for axis x, y, z: W/2 dwords, dword g holding the plane codes of children 2g and
//   2g + 1 as bytes (q_lo[2g], q_lo[2g+1], q_hi[2g], q_hi[2g+1])
//   then W child refs (same encodings as above), zero padding to the node size, except that
//   W = 4 keeps s_y and s_z as floats in dwords 14 and 15 (the traversal then multiplies
//   them with 1/d directly instead of decoding exponent bytes on every node visit)
// The stored origin is biased (bvh_quant.h): p = the children's lower corner - 1024 s,
// rounded down, and a code q stands for the plane p + (1024 + q) s.  The decoded box lo =
// fma(1024 + q_lo, s, p), hi = fma(1024 + q_hi, s, p) (one rounding, as the device evaluates
// it) contains the padded float box of the child; empty children have q_lo = 255, q_hi = 0
// and ref kEmpty.  The host keeps the float nodes for its own traversal.
#pragma once

#include <cstdint>
#include <cstddef>
#include <vector>

namespace rbr {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kCountShift = 27;
constexpr uint32_t kStartMask = (1u << 27) - 1u;
constexpr uint32_t kMaxLeaf = 8;
constexpr int kQuantFlag = 0x100;   // width code flag: quantized child boxes

inline int qNodeDwords(int W) { return W == 4 ? 16 : 32; }
inline int qRefOffset(int W) { return 4 + 6 * (W / 4); }

enum PrimKind : uint32_t { KIND_TRI = 0, KIND_SPHERE = 1, KIND_QUAD = 2 };

inline bool isLeaf(uint32_t c) { return (c & kLeafBit) != 0u; }
inline uint32_t leafCount(uint32_t c) { return ((c >> kCountShift) & 7u) + 1u; }
inline uint32_t leafStart(uint32_t c) { return c & kStartMask; }
inline uint32_t makeLeaf(uint32_t start, uint32_t count) {
    return kLeafBit | ((count - 1u) << kCountShift) | start;
}

struct F4 { float x, y, z, w; };

struct HostGeometry {
    // Bounded primitives in caller order.
    std::vector<float> tris;      // 9 per tri
    std::vector<uint32_t> triIds;
    std::vector<float> sphs;      // 4 per sphere
    std::vector<uint32_t> sphIds;
    std::vector<float> quads;     // 12 per quad
    std::vector<uint32_t> quadIds;
    std::vector<float> planes;    // 4 per plane
    std::vector<uint32_t> planeIds;
    uint32_t maxId = 0;
    bool anyId = false;
};

struct WideBVH {
    int width = 4;
    bool quantized = false;
    std::vector<F4> nodes;        // width==4: 8 F4 per node; width==8: 16 F4 per node
    std::vector<uint32_t> qnodes; // quantized device nodes (qNodeDwords(width) per node)
    std::vector<F4> prims;        // 3 F4 per record
    std::vector<F4> quadNormals;
    uint32_t nodeCount = 0, leafCount = 0, primCount = 0, maxDepth = 0;
    double sahCost = 0.0, buildMs = 0.0;
    float worldLo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, worldHi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    bool empty() const { return primCount == 0; }
    int f4PerNode() const { return width == 4 ? 8 : 16; }
    // What the kernels are given: width code, dwords per node, child-ref dword offset.
    int code() const { return width | (quantized ? kQuantFlag : 0); }
    int deviceNodeDwords() const { return quantized ? qNodeDwords(width) : 4 * f4PerNode(); }
    int deviceRefOffset() const { return quantized ? qRefOffset(width) : 6 * width; }
    const void* deviceNodes() const { return quantized ? static_cast<const void*>(qnodes.data()) : nodes.data(); }
    size_t deviceNodeBytes() const { return quantized ? qnodes.size() * 4 : nodes.size() * sizeof(F4); }
    // The same sizes from the counts (also while the host arrays of a device build are empty).
    size_t nodeBytes() const { return size_t(nodeCount) * size_t(deviceNodeDwords()) * 4; }
    size_t primBytes() const { return size_t(primCount) * 3 * sizeof(F4); }
};

// Fills b.qnodes from the float nodes (conservative 8-bit child boxes) and sets b.quantized.
void quantizeWideBVH(WideBVH& b);

// Builds the binary binned-SAH tree and collapses it to `width` (4 or 8); leaves hold at
// most leaf_max (1..8) primitives.  optimal: the SAH-optimal collapse with one primitive test
// costing primCost node visits; else the greedy largest-area collapse.
constexpr float kCollapsePrimCost = 0.8f;
void buildWideBVH(const HostGeometry& g, int width, int leafMax, WideBVH& out, bool optimal = true,
                  float primCost = kCollapsePrimCost);

// Host traversal of the same flat layout (the per-ray path of the Accelerator adapter and
// the node/primitive visit counter).  Same acceptance rule as the kernels.
struct TraceCounters { uint64_t nodes = 0, prims = 0; };
uint32_t hostClosest(const WideBVH& b, const HostGeometry& g, const float o[3], const float d[3],
                     float tmin, float tmax, float* tHit, TraceCounters* cnt);
bool hostAny(const WideBVH& b, const HostGeometry& g, const float o[3], const float d[3], float tmin,
             float tmax, TraceCounters* cnt);
// Node visits and primitive tests summed over a batch of rays (32 B each: o, d, tmin, tmax)
// traced by the host traversal, closest hit or any hit; OpenMP over the rays.
TraceCounters hostCount(const WideBVH& b, const HostGeometry& g, const float* rays, size_t n, bool any);

}  // namespace rbr
