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
#pragma once

#include <cstdint>
#include <vector>

namespace rbr {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kCountShift = 27;
constexpr uint32_t kStartMask = (1u << 27) - 1u;
constexpr uint32_t kMaxLeaf = 8;

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
    std::vector<F4> nodes;        // width==4: 8 F4 per node; width==8: 16 F4 per node
    std::vector<F4> prims;        // 3 F4 per record
    std::vector<F4> quadNormals;
    uint32_t nodeCount = 0, leafCount = 0, primCount = 0, maxDepth = 0;
    double sahCost = 0.0, buildMs = 0.0;
    float worldLo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, worldHi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    bool empty() const { return primCount == 0; }
    int f4PerNode() const { return width == 4 ? 8 : 16; }
};

// Builds the binary binned-SAH tree and collapses it to `width` (4 or 8); leaves hold at
// most leaf_max (1..8) primitives.
void buildWideBVH(const HostGeometry& g, int width, int leafMax, WideBVH& out);

// Host traversal of the same flat layout (the per-ray path of the Accelerator adapter and
// the node/primitive visit counter).  Same acceptance rule as the kernels.
struct TraceCounters { uint64_t nodes = 0, prims = 0; };
uint32_t hostClosest(const WideBVH& b, const HostGeometry& g, const float o[3], const float d[3],
                     float tmin, float tmax, float* tHit, TraceCounters* cnt);
bool hostAny(const WideBVH& b, const HostGeometry& g, const float o[3], const float d[3], float tmin,
             float tmax, TraceCounters* cnt);

}  // namespace rbr
