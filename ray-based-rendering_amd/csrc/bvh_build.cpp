// Host construction of the wide BVH: binned SAH over all three axes (32 bins), built in
// parallel with OpenMP tasks, then collapsed to `width` children per node, either by the
// SAH-optimal dynamic program (default) or greedily (open the largest-area interior child
// until the node has `width` children).  Replaces BVH::build / buildRecursive /
// findBestSplit (bvh.cpp:10-31, 47-209): same input (every bounded primitive, id = caller's
// index), different tree; results do not depend on the tree because every traversal
// returns the lowest-id closest hit.
#include "bvh.h"
#include "bvh_quant.h"
#include "prim_tests.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <utility>
#include <vector>

namespace rbr {

namespace {

struct BPrim {
    float lo[3], hi[3], c[3];
    uint32_t kind, index;
};

struct BNode {
    float lo[3], hi[3];
    int32_t left, right;   // -1: leaf
    uint32_t start, count;
    bool leaf() const { return left < 0; }
};

// Vectors whose elements are left uninitialised on allocation: the builder fills its large
// arrays in parallel loops (which also spreads their first touch over the threads) instead of
// having them zeroed serially first.
template <typename T>
struct NoInitAlloc : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <typename U>
    NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
    template <typename U>
    void construct(U* q) noexcept {
        ::new (static_cast<void*>(q)) U;
    }
    template <typename U, typename... A>
    void construct(U* q, A&&... a) {
        ::new (static_cast<void*>(q)) U(std::forward<A>(a)...);
    }
};
using PrimVec = std::vector<BPrim, NoInitAlloc<BPrim>>;
using NodeVec = std::vector<BNode, NoInitAlloc<BNode>>;

inline float area(const float lo[3], const float hi[3]) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Node indices are laid out by range, so no counter is shared between threads: the node of a
// range of k primitives owns a block of 2k - 1 indices (itself first, then its left child's
// block, then its right child's), which holds the subtree down to single primitives; where a
// leaf keeps several primitives the rest of its block stays unused (an empty leaf).
struct Builder {
    PrimVec& p;
    NodeVec& nodes;
    int leafMax;
    static constexpr int kBins = 32;

    void leaf(BNode& nd, uint32_t s, uint32_t e) {
        nd.left = nd.right = -1;
        nd.start = s;
        nd.count = e - s;
    }

    // Bounds of a range: node box and centroid box.
    struct Bounds {
        float lo[3], hi[3], clo[3], chi[3];
        void clear() {
            for (int k = 0; k < 3; ++k) { lo[k] = clo[k] = 3.4e38f; hi[k] = chi[k] = -3.4e38f; }
        }
        void add(const BPrim& q) {
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], q.lo[k]);
                hi[k] = std::max(hi[k], q.hi[k]);
                clo[k] = std::min(clo[k], q.c[k]);
                chi[k] = std::max(chi[k], q.c[k]);
            }
        }
        void merge(const Bounds& o) {
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], o.lo[k]);
                hi[k] = std::max(hi[k], o.hi[k]);
                clo[k] = std::min(clo[k], o.clo[k]);
                chi[k] = std::max(chi[k], o.chi[k]);
            }
        }
    };
    // The 32 bins of the three axes, filled in one pass over a range.
    struct Bins {
        int cnt[3][kBins];
        float lo[3][kBins][3], hi[3][kBins][3];
        void clear() {
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < kBins; ++b) {
                    cnt[a][b] = 0;
                    for (int k = 0; k < 3; ++k) { lo[a][b][k] = 3.4e38f; hi[a][b][k] = -3.4e38f; }
                }
        }
        void merge(const Bins& o) {
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < kBins; ++b) {
                    cnt[a][b] += o.cnt[a][b];
                    for (int k = 0; k < 3; ++k) {
                        lo[a][b][k] = std::min(lo[a][b][k], o.lo[a][b][k]);
                        hi[a][b][k] = std::max(hi[a][b][k], o.hi[a][b][k]);
                    }
                }
        }
    };
    // Ranges above this size are scanned in parallel chunks (min, max and counts merge exactly,
    // so the split found does not depend on the chunking).
    static constexpr uint32_t kParallelScan = 1u << 17, kScanChunk = 1u << 15;

    void bounds(uint32_t s, uint32_t e, Bounds& out) const {
        out.clear();
        if (e - s <= kParallelScan) {
            for (uint32_t i = s; i < e; ++i) out.add(p[i]);
            return;
        }
        const uint32_t chunks = (e - s + kScanChunk - 1) / kScanChunk;
        std::vector<Bounds> part(chunks);
#pragma omp taskloop default(shared) grainsize(1)
        for (uint32_t c = 0; c < chunks; ++c) {
            part[c].clear();
            const uint32_t a = s + c * kScanChunk, b = std::min(e, a + kScanChunk);
            for (uint32_t i = a; i < b; ++i) part[c].add(p[i]);
        }
        for (const Bounds& b : part) out.merge(b);
    }

    void binRange(uint32_t s, uint32_t e, const float cmin[3], const float scale[3], const bool use[3], Bins& out) const {
        for (uint32_t i = s; i < e; ++i) {
            const BPrim& q = p[i];
            for (int ax = 0; ax < 3; ++ax) {
                if (!use[ax]) continue;
                const int b = std::min(kBins - 1, int((q.c[ax] - cmin[ax]) * scale[ax]));
                out.cnt[ax][b]++;
                for (int k = 0; k < 3; ++k) {
                    out.lo[ax][b][k] = std::min(out.lo[ax][b][k], q.lo[k]);
                    out.hi[ax][b][k] = std::max(out.hi[ax][b][k], q.hi[k]);
                }
            }
        }
    }

    void bins(uint32_t s, uint32_t e, const float cmin[3], const float scale[3], const bool use[3], Bins& out) const {
        out.clear();
        if (e - s <= kParallelScan) {
            binRange(s, e, cmin, scale, use, out);
            return;
        }
        const uint32_t chunks = (e - s + kScanChunk - 1) / kScanChunk;
        std::vector<Bins> part(chunks);
#pragma omp taskloop default(shared) grainsize(1)
        for (uint32_t c = 0; c < chunks; ++c) {
            part[c].clear();
            const uint32_t a = s + c * kScanChunk;
            binRange(a, std::min(e, a + kScanChunk), cmin, scale, use, part[c]);
        }
        for (const Bins& b : part) out.merge(b);
    }

    void build(int32_t ni, uint32_t s, uint32_t e, int depth) {
        BNode& nd = nodes[ni];
        Bounds bd;
        bounds(s, e, bd);
        const float* clo = bd.clo;
        const float* chi = bd.chi;
        for (int k = 0; k < 3; ++k) { nd.lo[k] = bd.lo[k]; nd.hi[k] = bd.hi[k]; }
        const uint32_t n = e - s;
        if (n <= 1 || depth > 60) { leaf(nd, s, e); if (n > (uint32_t)leafMax) forceSplit(ni, s, e, depth); return; }

        // Binned SAH on every axis (one pass fills the bins of all three).
        float bestCost = 3.4e38f;
        int bestAxis = -1, bestBin = -1;
        const float parentArea = area(nd.lo, nd.hi);
        float scales[3];
        bool use[3];
        for (int ax = 0; ax < 3; ++ax) {
            const float ext = chi[ax] - clo[ax];
            use[ax] = ext > 0.0f;
            scales[ax] = use[ax] ? float(kBins) * (1.0f - 1e-6f) / ext : 0.0f;
        }
        std::unique_ptr<Bins> dense;
        if (n > kSparse) {
            dense.reset(new Bins);
            bins(s, e, clo, scales, use, *dense);
        }
        for (int ax = 0; ax < 3; ++ax) {
            if (!use[ax]) continue;
            if (n <= kSparse) {
                sparseBins(s, e, ax, clo[ax], scales[ax], parentArea, bestCost, bestAxis, bestBin);
                continue;
            }
            const int* cnt = dense->cnt[ax];
            const float(*blo)[3] = dense->lo[ax];
            const float(*bhi)[3] = dense->hi[ax];
            float rightArea[kBins];
            int rightCnt[kBins];
            float rlo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, rhi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
            int rc = 0;
            for (int b = kBins - 1; b > 0; --b) {
                rc += cnt[b];
                for (int k = 0; k < 3; ++k) { rlo[k] = std::min(rlo[k], blo[b][k]); rhi[k] = std::max(rhi[k], bhi[b][k]); }
                rightArea[b] = area(rlo, rhi);
                rightCnt[b] = rc;
            }
            float llo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, lhi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
            int lc = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                lc += cnt[b];
                for (int k = 0; k < 3; ++k) { llo[k] = std::min(llo[k], blo[b][k]); lhi[k] = std::max(lhi[k], bhi[b][k]); }
                if (lc == 0 || rightCnt[b + 1] == 0) continue;
                float cost = 1.0f + (area(llo, lhi) * lc + rightArea[b + 1] * rightCnt[b + 1]) / std::max(parentArea, 1e-30f);
                if (cost < bestCost) { bestCost = cost; bestAxis = ax; bestBin = b; }
            }
        }
        if (bestAxis < 0) {   // all centroids coincide
            leaf(nd, s, e);
            if (n > (uint32_t)leafMax) forceSplit(ni, s, e, depth);
            return;
        }
        if (n <= (uint32_t)leafMax && float(n) <= bestCost) { leaf(nd, s, e); return; }

        const float ext = chi[bestAxis] - clo[bestAxis];
        const float scale = float(kBins) * (1.0f - 1e-6f) / ext;
        const float cmin = clo[bestAxis];
        auto mid = std::partition(p.begin() + s, p.begin() + e, [&](const BPrim& q) {
            return std::min(kBins - 1, int((q.c[bestAxis] - cmin) * scale)) <= bestBin;
        });
        uint32_t m = uint32_t(mid - p.begin());
        if (m == s || m == e) m = s + n / 2;
        split(ni, s, m, e, depth);
    }

    // The binned SAH sweep of one axis over the occupied bins only, for small ranges: the same
    // candidates in the same order with the same float operations as the 32-bin sweep (an
    // empty bin changes neither box nor count, and of equal costs the first is kept), so it
    // picks the same split; it just skips the empty bins, which dominate below a few dozen
    // primitives.
    static constexpr uint32_t kSparse = 24;
    void sparseBins(uint32_t s, uint32_t e, int ax, float cmin, float scale, float parentArea, float& bestCost,
                    int& bestAxis, int& bestBin) const {
        const uint32_t n = e - s;
        uint8_t bin[kSparse];
        uint8_t ord[kSparse];
        for (uint32_t i = 0; i < n; ++i) {
            bin[i] = uint8_t(std::min(kBins - 1, int((p[s + i].c[ax] - cmin) * scale)));
            uint32_t j = i;   // insertion sort of the primitives by bin
            while (j > 0 && bin[ord[j - 1]] > bin[i]) { ord[j] = ord[j - 1]; --j; }
            ord[j] = uint8_t(i);
        }
        // occupied bins in ascending order: bin, count, box
        int ob[kSparse], oc[kSparse], m = 0;
        float olo[kSparse][3], ohi[kSparse][3];
        for (uint32_t k = 0; k < n; ++k) {
            const uint32_t i = ord[k];
            if (m == 0 || ob[m - 1] != bin[i]) {
                ob[m] = bin[i];
                oc[m] = 0;
                for (int q = 0; q < 3; ++q) { olo[m][q] = 3.4e38f; ohi[m][q] = -3.4e38f; }
                ++m;
            }
            ++oc[m - 1];
            for (int q = 0; q < 3; ++q) {
                olo[m - 1][q] = std::min(olo[m - 1][q], p[s + i].lo[q]);
                ohi[m - 1][q] = std::max(ohi[m - 1][q], p[s + i].hi[q]);
            }
        }
        float rightArea[kSparse];
        int rightCnt[kSparse];
        float rlo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, rhi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
        int rc = 0;
        for (int j = m - 1; j > 0; --j) {
            rc += oc[j];
            for (int q = 0; q < 3; ++q) { rlo[q] = std::min(rlo[q], olo[j][q]); rhi[q] = std::max(rhi[q], ohi[j][q]); }
            rightArea[j] = area(rlo, rhi);
            rightCnt[j] = rc;
        }
        float llo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, lhi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
        int lc = 0;
        for (int j = 0; j + 1 < m; ++j) {
            lc += oc[j];
            for (int q = 0; q < 3; ++q) { llo[q] = std::min(llo[q], olo[j][q]); lhi[q] = std::max(lhi[q], ohi[j][q]); }
            if (ob[j] > kBins - 2) continue;
            float cost = 1.0f + (area(llo, lhi) * lc + rightArea[j + 1] * rightCnt[j + 1]) / std::max(parentArea, 1e-30f);
            if (cost < bestCost) { bestCost = cost; bestAxis = ax; bestBin = ob[j]; }
        }
    }

    void forceSplit(int32_t ni, uint32_t s, uint32_t e, int depth) {
        split(ni, s, s + (e - s) / 2, e, depth);
    }

    void split(int32_t ni, uint32_t s, uint32_t m, uint32_t e, int depth) {
        const int32_t l = ni + 1, r = ni + 2 * int32_t(m - s);
        nodes[ni].left = l;
        nodes[ni].right = r;
        nodes[ni].start = s;
        nodes[ni].count = e - s;
        if (e - s > 16384) {
#pragma omp task default(shared) firstprivate(l, s, m, depth)
            build(l, s, m, depth + 1);
#pragma omp task default(shared) firstprivate(r, m, e, depth)
            build(r, m, e, depth + 1);
#pragma omp taskwait
        } else {
            build(l, s, m, depth + 1);
            build(r, m, e, depth + 1);
        }
    }
};

struct Collapser {
    const NodeVec& bn;
    WideBVH& out;
    int W;
    float pad;

    void padBox(const float lo[3], const float hi[3], float plo[3], float phi[3]) const {
        for (int k = 0; k < 3; ++k) {
            plo[k] = lo[k] - (pad + std::fabs(lo[k]) * 1e-6f);
            phi[k] = hi[k] + (pad + std::fabs(hi[k]) * 1e-6f);
        }
    }

    // Emits the wide node whose children are drawn from binary node `b` (interior, or a
    // leaf root) and returns its index.
    uint32_t emit(int32_t b, uint32_t depth) {
        out.maxDepth = std::max(out.maxDepth, depth);
        int32_t kids[8];
        int nk = 0;
        if (bn[b].leaf()) {
            kids[nk++] = b;
        } else {
            kids[nk++] = bn[b].left;
            kids[nk++] = bn[b].right;
            while (nk < W) {
                int best = -1;
                float bestA = -1.0f;
                for (int i = 0; i < nk; ++i) {
                    const BNode& c = bn[kids[i]];
                    if (c.leaf()) continue;
                    float a = area(c.lo, c.hi);
                    if (a > bestA) { bestA = a; best = i; }
                }
                if (best < 0) break;
                int32_t open = kids[best];
                kids[best] = bn[open].left;
                kids[nk++] = bn[open].right;
            }
        }
        const uint32_t idx = out.nodeCount++;
        const int per = out.f4PerNode();
        out.nodes.resize(size_t(out.nodeCount) * per);
        // Fill children (recursion may reallocate out.nodes, so write through indices).
        uint32_t child[8];
        float blo[8][3], bhi[8][3];
        for (int i = 0; i < W; ++i) {
            child[i] = kEmpty;
            for (int k = 0; k < 3; ++k) { blo[i][k] = INFINITY; bhi[i][k] = -INFINITY; }
        }
        for (int i = 0; i < nk; ++i) {
            const BNode& c = bn[kids[i]];
            padBox(c.lo, c.hi, blo[i], bhi[i]);
            if (c.leaf()) {
                child[i] = makeLeaf(c.start, c.count);
                out.leafCount++;
            } else {
                child[i] = emit(kids[i], depth + 1);
            }
        }
        float* f = reinterpret_cast<float*>(&out.nodes[size_t(idx) * per]);
        for (int i = 0; i < W; ++i) {
            f[0 * W + i] = blo[i][0];
            f[1 * W + i] = bhi[i][0];
            f[2 * W + i] = blo[i][1];
            f[3 * W + i] = bhi[i][1];
            f[4 * W + i] = blo[i][2];
            f[5 * W + i] = bhi[i][2];
            std::memcpy(&f[6 * W + i], &child[i], 4);
        }
        return idx;
    }
};

// SAH-optimal collapse of the binary tree into `width`-wide nodes (after Ylitie, Karras &
// Laine 2017, "Efficient incoherent ray traversal on GPUs through compressed wide BVHs").
// Bottom-up over the binary nodes, cost(n, i) is the least SAH cost of the binary subtree n
// when it may fill at most i child slots of its wide parent.  A slot is either a leaf (any
// binary subtree of at most leafMax primitives: its primitives are contiguous) or a wide node
// whose slots are drawn from n's two children:
//   cost(n, 1) = min(leaf(n), A(n) c_node + dist(n, W)),   leaf(n) = A(n) count(n) c_prim
//   cost(n, i) = min(dist(n, i), cost(n, i - 1)),           dist(n, j) = min_k cost(l, k) + cost(r, j - k)
// so leaf sizes and which binary levels become wide nodes are chosen together.  The greedy
// collapse above fixes the leaves first and fills each node by area alone.
struct OptimalCollapser {
    static constexpr int8_t kLeaf = -1, kWide = -2, kFewer = -3, kDist = 0;
    const NodeVec& bn;
    WideBVH& out;
    int W, leafMax;
    float pad, cNode, cPrim;
    std::vector<float> cost;     // [node * W + i - 1]
    std::vector<int8_t> how;     // same index: kLeaf / kWide (i = 1), kDist / kFewer (i > 1)
    std::vector<int8_t> split;   // [node * W + j - 1]: slots given to the left child in dist(n, j)

    void solve() {
        const size_t n = bn.size();
        cost.assign(n * W, INFINITY);
        how.assign(n * W, kFewer);
        split.assign(n * W, 1);
#pragma omp parallel
#pragma omp single
        solveSubtree(0);
    }

    // Post-order over binary subtree m (a node's table needs its children's); subtrees of more
    // than kTaskPrims primitives run as tasks.  Each table is a function of the children's
    // alone, so the result does not depend on the schedule.
    static constexpr uint32_t kTaskPrims = 16384;
    void solveSubtree(int32_t m) {
        const BNode& b = bn[m];
        if (!b.leaf()) {
            if (b.count > kTaskPrims) {
#pragma omp task default(shared)
                solveSubtree(b.left);
#pragma omp task default(shared)
                solveSubtree(b.right);
#pragma omp taskwait
            } else {
                solveSubtree(b.left);
                solveSubtree(b.right);
            }
        }
        solveNode(size_t(m));
    }

    void solveNode(size_t m) {
        const BNode& b = bn[m];
        const float a = area(b.lo, b.hi);
        const float leafC = b.count <= uint32_t(leafMax) ? a * float(b.count) * cPrim : INFINITY;
        float* c = &cost[m * W];
        int8_t* h = &how[m * W];
        if (b.leaf()) {
            c[0] = leafC;
            h[0] = kLeaf;
            for (int i = 1; i < W; ++i) c[i] = c[0];
            return;
        }
        const float* cl = &cost[size_t(b.left) * W];
        const float* cr = &cost[size_t(b.right) * W];
        float dist[9];
        for (int j = 2; j <= W; ++j) {
            float best = INFINITY;
            int bk = 1;
            for (int k = 1; k < j; ++k) {
                const float v = cl[k - 1] + cr[j - k - 1];
                if (v < best) { best = v; bk = k; }
            }
            dist[j] = best;
            split[m * W + j - 1] = int8_t(bk);
        }
        const float wideC = a * cNode + dist[W];
        if (leafC <= wideC) { c[0] = leafC; h[0] = kLeaf; }
        else { c[0] = wideC; h[0] = kWide; }
        for (int i = 2; i <= W; ++i) {
            if (dist[i] < c[i - 2]) { c[i - 1] = dist[i]; h[i - 1] = kDist; }
            else { c[i - 1] = c[i - 2]; h[i - 1] = kFewer; }
        }
    }

    // Slots of binary subtree m given at most i of them: (node, is leaf).
    void gather(int32_t m, int i, int32_t* kids, bool* leaf, int& nk) const {
        while (how[size_t(m) * W + i - 1] == kFewer) --i;
        const int8_t h = how[size_t(m) * W + i - 1];
        if (i == 1) {
            kids[nk] = m;
            leaf[nk++] = h == kLeaf;
            return;
        }
        const int k = split[size_t(m) * W + i - 1];
        gather(bn[m].left, k, kids, leaf, nk);
        gather(bn[m].right, i - k, kids, leaf, nk);
    }

    void padBox(const float lo[3], const float hi[3], float plo[3], float phi[3]) const {
        for (int k = 0; k < 3; ++k) {
            plo[k] = lo[k] - (pad + std::fabs(lo[k]) * 1e-6f);
            phi[k] = hi[k] + (pad + std::fabs(hi[k]) * 1e-6f);
        }
    }

    // The slots of the wide node drawn from binary node b; a root that is best kept as one
    // leaf becomes a node with that single leaf.
    void slots(int32_t b, int32_t* kids, bool* leaf, int& nk) const {
        nk = 0;
        if (bn[b].leaf() || how[size_t(b) * W] == kLeaf) {
            kids[nk] = b;
            leaf[nk++] = true;
        } else {
            const int k = split[size_t(b) * W + W - 1];
            gather(bn[b].left, k, kids, leaf, nk);
            gather(bn[b].right, W - k, kids, leaf, nk);
        }
    }

    // Wide nodes in the subtree of the wide node drawn from b (itself included), kept per
    // binary node for the emission.
    std::vector<uint32_t> wideSize;
    uint32_t sizeSubtree(int32_t b) {
        int32_t kids[8];
        bool leaf[8];
        int nk;
        slots(b, kids, leaf, nk);
        uint32_t part[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const bool par = bn[b].count > kTaskPrims;
        for (int i = 0; i < nk; ++i) {
            if (leaf[i]) continue;
            if (par) {
#pragma omp task default(shared) firstprivate(i)
                part[i] = sizeSubtree(kids[i]);
            } else {
                part[i] = sizeSubtree(kids[i]);
            }
        }
        if (par) {
#pragma omp taskwait
        }
        uint32_t total = 1;
        for (int i = 0; i < nk; ++i) total += part[i];
        wideSize[b] = total;
        return total;
    }

    // Emits the wide nodes in depth-first pre-order: the node drawn from b at index idx, its
    // interior children at idx + 1 + the sizes of the subtrees before them, so subtrees can be
    // written in parallel and still land where a serial depth-first emission puts them.
    std::atomic<uint32_t> leaves{0}, deepest{0};
    void emitAt(int32_t b, uint32_t idx, uint32_t depth) {
        uint32_t d = deepest.load(std::memory_order_relaxed);
        while (depth > d && !deepest.compare_exchange_weak(d, depth, std::memory_order_relaxed)) {}
        int32_t kids[8];
        bool leaf[8];
        int nk;
        slots(b, kids, leaf, nk);
        const int per = out.f4PerNode();
        uint32_t child[8];
        float blo[8][3], bhi[8][3];
        for (int i = 0; i < W; ++i) {
            child[i] = kEmpty;
            for (int k = 0; k < 3; ++k) { blo[i][k] = INFINITY; bhi[i][k] = -INFINITY; }
        }
        const bool par = bn[b].count > kTaskPrims;
        uint32_t next = idx + 1, nleaf = 0;
        for (int i = 0; i < nk; ++i) {
            const BNode& c = bn[kids[i]];
            padBox(c.lo, c.hi, blo[i], bhi[i]);
            if (leaf[i]) {
                child[i] = makeLeaf(c.start, c.count);
                ++nleaf;
                continue;
            }
            child[i] = next;
            const int32_t kb = kids[i];
            const uint32_t at = next;
            next += wideSize[kb];
            if (par) {
#pragma omp task default(shared) firstprivate(kb, at)
                emitAt(kb, at, depth + 1);
            } else {
                emitAt(kb, at, depth + 1);
            }
        }
        leaves.fetch_add(nleaf, std::memory_order_relaxed);
        float* f = reinterpret_cast<float*>(&out.nodes[size_t(idx) * per]);
        for (int i = 0; i < W; ++i) {
            f[0 * W + i] = blo[i][0];
            f[1 * W + i] = bhi[i][0];
            f[2 * W + i] = blo[i][1];
            f[3 * W + i] = bhi[i][1];
            f[4 * W + i] = blo[i][2];
            f[5 * W + i] = bhi[i][2];
            std::memcpy(&f[6 * W + i], &child[i], 4);
        }
        if (par) {
#pragma omp taskwait
        }
    }

    // The whole wide tree from binary root 0.
    void emitAll() {
        wideSize.assign(bn.size(), 0u);
        uint32_t total = 0;
#pragma omp parallel
#pragma omp single
        total = sizeSubtree(0);
        out.nodeCount = total;
        out.nodes.assign(size_t(total) * out.f4PerNode(), F4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma omp parallel
#pragma omp single
        emitAt(0, 0, 1);
        out.leafCount = leaves.load();
        out.maxDepth = deepest.load();
    }
};

inline void minmax3(const float* a, const float* b, const float* c, float lo[3], float hi[3]) {
    for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(std::min(a[k], b[k]), c[k]);
        hi[k] = std::max(std::max(a[k], b[k]), c[k]);
    }
}

}  // namespace

void buildWideBVH(const HostGeometry& g, int width, int leafMax, WideBVH& out, bool optimal, float primCost) {
    auto t0 = std::chrono::steady_clock::now();
    out = WideBVH();
    out.width = width;
    const uint32_t ntri = uint32_t(g.triIds.size()), nsph = uint32_t(g.sphIds.size()), nquad = uint32_t(g.quadIds.size());
    const uint32_t n = ntri + nsph + nquad;
    out.primCount = n;
    if (n == 0) return;

    PrimVec p(n);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)ntri; ++i) {
        const float* t = &g.tris[9 * size_t(i)];
        BPrim& q = p[i];
        minmax3(t, t + 3, t + 6, q.lo, q.hi);
        q.kind = KIND_TRI;
        q.index = uint32_t(i);
    }
    for (uint32_t i = 0; i < nsph; ++i) {
        const float* s = &g.sphs[4 * size_t(i)];
        BPrim& q = p[ntri + i];
        for (int k = 0; k < 3; ++k) { q.lo[k] = s[k] - s[3]; q.hi[k] = s[k] + s[3]; }
        q.kind = KIND_SPHERE;
        q.index = i;
    }
    for (uint32_t i = 0; i < nquad; ++i) {
        BPrim& q = p[ntri + nsph + i];
        quadBounds(&g.quads[12 * size_t(i)], q.lo, q.hi);
        q.kind = KIND_QUAD;
        q.index = i;
    }
    float wlo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, whi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
#pragma omp parallel for schedule(static) reduction(min : wlo[:3]) reduction(max : whi[:3])
    for (long long i = 0; i < (long long)n; ++i) {
        for (int k = 0; k < 3; ++k) {
            p[i].c[k] = 0.5f * (p[i].lo[k] + p[i].hi[k]);
            wlo[k] = std::min(wlo[k], p[i].lo[k]);
            whi[k] = std::max(whi[k], p[i].hi[k]);
        }
    }
    for (int k = 0; k < 3; ++k) { out.worldLo[k] = wlo[k]; out.worldHi[k] = whi[k]; }

    NodeVec bn(size_t(2) * n - 1);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)bn.size(); ++i) {   // unused blocks read as empty leaves
        BNode& x = bn[i];
        for (int k = 0; k < 3; ++k) x.lo[k] = x.hi[k] = 0.0f;
        x.left = x.right = -1;
        x.start = x.count = 0;
    }
    // The optimal collapse picks the leaves itself, so its binary tree goes down to single
    // primitives.
    Builder b{p, bn, optimal ? 1 : std::max(1, std::min(leafMax, int(kMaxLeaf)))};
#pragma omp parallel
#pragma omp single
    b.build(0, 0, n, 0);

    // SAH cost of the binary tree (Ct = 1, Ci = 1), relative to the root area.
    const float rootA = std::max(area(bn[0].lo, bn[0].hi), 1e-30f);
    double sah = 0.0;
    for (const BNode& x : bn) sah += double(area(x.lo, x.hi)) / rootA * (x.leaf() ? double(x.count) : 1.0);
    out.sahCost = sah;

    const float ext = std::max({whi[0] - wlo[0], whi[1] - wlo[1], whi[2] - wlo[2]});
    out.nodes.reserve(size_t(n / 2 + 8) * out.f4PerNode());
    if (optimal) {
        OptimalCollapser c{bn, out, width, std::max(1, std::min(leafMax, int(kMaxLeaf))), 1e-5f * ext + 1e-6f, 1.0f,
                           primCost, {}, {}, {}};
        c.solve();
        c.emitAll();
    } else {
        Collapser c{bn, out, width, 1e-5f * ext + 1e-6f};
        c.emit(0, 1);
    }

    // Leaf-ordered primitive records.
    out.prims.resize(size_t(n) * 3);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        const BPrim& q = p[i];
        F4* r = &out.prims[size_t(i) * 3];
        float idf, kf;
        uint32_t kind = q.kind;
        std::memcpy(&kf, &kind, 4);
        if (q.kind == KIND_TRI) {
            const float* t = &g.tris[9 * size_t(q.index)];
            uint32_t id = g.triIds[q.index];
            std::memcpy(&idf, &id, 4);
            r[0] = {t[0], t[1], t[2], idf};
            r[1] = {t[3] - t[0], t[4] - t[1], t[5] - t[2], kf};
            r[2] = {t[6] - t[0], t[7] - t[1], t[8] - t[2], 0.0f};
        } else if (q.kind == KIND_SPHERE) {
            const float* s = &g.sphs[4 * size_t(q.index)];
            uint32_t id = g.sphIds[q.index];
            std::memcpy(&idf, &id, 4);
            r[0] = {s[0], s[1], s[2], idf};
            r[1] = {s[3], 0.0f, 0.0f, kf};
            r[2] = {0.0f, 0.0f, 0.0f, 0.0f};
        } else {
            const float* s = &g.quads[12 * size_t(q.index)];
            uint32_t id = g.quadIds[q.index];
            std::memcpy(&idf, &id, 4);
            float qi;
            uint32_t qidx = q.index;
            std::memcpy(&qi, &qidx, 4);
            r[0] = {s[0], s[1], s[2], idf};
            r[1] = {s[3], s[4], s[5], kf};
            r[2] = {s[6], s[7], s[8], qi};
        }
    }
    out.quadNormals.resize(nquad);
    for (uint32_t i = 0; i < nquad; ++i) {
        const float* s = &g.quads[12 * size_t(i)];
        out.quadNormals[i] = {s[9], s[10], s[11], 0.0f};
    }
    out.buildMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ---------------------------------------------------------------- host traversal

namespace {

inline uint32_t asU(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

inline bool testRecord(const F4* r, const WideBVH& b, V o, V d, float tmin, float tmax, float& t) {
    const uint32_t kind = asU(r[1].w);
    if (kind == KIND_TRI)
        return triTest(o, d, V{r[0].x, r[0].y, r[0].z}, V{r[1].x, r[1].y, r[1].z}, V{r[2].x, r[2].y, r[2].z}, tmin, tmax, t);
    if (kind == KIND_SPHERE) return sphereTest(o, d, V{r[0].x, r[0].y, r[0].z}, r[1].x, tmin, tmax, t);
    const F4& n = b.quadNormals[asU(r[2].w)];
    return quadTest(o, d, V{r[0].x, r[0].y, r[0].z}, V{r[1].x, r[1].y, r[1].z}, V{r[2].x, r[2].y, r[2].z},
                    V{n.x, n.y, n.z}, tmin, tmax, t);
}

struct HostRay {
    V o, d, inv;
    int neg[3];
};

inline HostRay makeRay(const float o[3], const float d[3]) {
    HostRay r;
    r.o = {o[0], o[1], o[2]};
    r.d = {d[0], d[1], d[2]};
    float dd[3] = {d[0], d[1], d[2]};
    float inv[3];
    for (int k = 0; k < 3; ++k) {
        float x = dd[k];
        if (std::fabs(x) < 1e-30f) x = std::copysign(1e-30f, x);
        inv[k] = 1.0f / x;
        r.neg[k] = inv[k] < 0.0f;
    }
    r.inv = {inv[0], inv[1], inv[2]};
    return r;
}

// Slab test of child i of a node against [0, tfar].
inline bool childHit(const float* f, int W, int i, const HostRay& r, float tfar, float& tn) {
    const float ox = r.o.x, oy = r.o.y, oz = r.o.z;
    float lx = f[0 * W + i], hx = f[1 * W + i], ly = f[2 * W + i], hy = f[3 * W + i], lz = f[4 * W + i], hz = f[5 * W + i];
    float nx = r.neg[0] ? hx : lx, fx = r.neg[0] ? lx : hx;
    float ny = r.neg[1] ? hy : ly, fy = r.neg[1] ? ly : hy;
    float nz = r.neg[2] ? hz : lz, fz = r.neg[2] ? lz : hz;
    float t0 = std::max(std::max((nx - ox) * r.inv.x, (ny - oy) * r.inv.y), std::max((nz - oz) * r.inv.z, 0.0f));
    float t1 = std::min(std::min((fx - ox) * r.inv.x, (fy - oy) * r.inv.y), std::min((fz - oz) * r.inv.z, tfar));
    tn = t0;
    return t0 <= t1;
}

}  // namespace

uint32_t hostClosest(const WideBVH& b, const HostGeometry& g, const float o[3], const float d[3], float tmin,
                     float tmax, float* tHit, TraceCounters* cnt) {
    uint32_t bestId = 0xFFFFFFFFu;
    float bestT = tmax;
    HostRay r = makeRay(o, d);
    if (!b.empty()) {
        const int W = b.width, per = b.f4PerNode();
        uint32_t stk[256];
        float stt[256];
        int sp = 0;
        stk[sp] = 0;
        stt[sp] = 0.0f;
        sp++;
        while (sp > 0) {
            --sp;
            uint32_t ref = stk[sp];
            if (stt[sp] > bestT) continue;
            if (isLeaf(ref)) {
                const uint32_t s = leafStart(ref), c = leafCount(ref);
                for (uint32_t k = 0; k < c; ++k) {
                    const F4* rec = &b.prims[size_t(s + k) * 3];
                    float t;
                    if (cnt) cnt->prims++;
                    if (testRecord(rec, b, r.o, r.d, tmin, bestT, t)) {
                        uint32_t id = asU(rec[0].w);
                        if (t < bestT || id < bestId) { bestT = t; bestId = id; }
                    }
                }
                continue;
            }
            if (cnt) cnt->nodes++;
            const float* f = reinterpret_cast<const float*>(&b.nodes[size_t(ref) * per]);
            const uint32_t* ch = reinterpret_cast<const uint32_t*>(f + 6 * W);
            uint32_t hr[8];
            float ht[8];
            int nh = 0;
            for (int i = 0; i < W; ++i) {
                float tn;
                if (ch[i] != kEmpty && childHit(f, W, i, r, bestT, tn)) { hr[nh] = ch[i]; ht[nh] = tn; nh++; }
            }
            // Push farthest first so the nearest is popped next.
            for (int i = 1; i < nh; ++i)
                for (int j = i; j > 0 && ht[j] > ht[j - 1]; --j) { std::swap(ht[j], ht[j - 1]); std::swap(hr[j], hr[j - 1]); }
            for (int i = 0; i < nh; ++i) { stk[sp] = hr[i]; stt[sp] = ht[i]; sp++; }
        }
    }
    const size_t np = g.planeIds.size();
    for (size_t k = 0; k < np; ++k) {
        const float* pl = &g.planes[4 * k];
        float t;
        if (planeTest(r.o, r.d, V{pl[0], pl[1], pl[2]}, pl[3], tmin, bestT, t)) {
            uint32_t id = g.planeIds[k];
            if (t < bestT || id < bestId) { bestT = t; bestId = id; }
        }
    }
    if (tHit) *tHit = bestT;
    return bestId;
}

bool hostAny(const WideBVH& b, const HostGeometry& g, const float o[3], const float d[3], float tmin, float tmax,
             TraceCounters* cnt) {
    HostRay r = makeRay(o, d);
    if (!b.empty()) {
        const int W = b.width, per = b.f4PerNode();
        uint32_t stk[256];
        int sp = 0;
        stk[sp++] = 0;
        while (sp > 0) {
            uint32_t ref = stk[--sp];
            if (isLeaf(ref)) {
                const uint32_t s = leafStart(ref), c = leafCount(ref);
                for (uint32_t k = 0; k < c; ++k) {
                    float t;
                    if (cnt) cnt->prims++;
                    if (testRecord(&b.prims[size_t(s + k) * 3], b, r.o, r.d, tmin, tmax, t)) return true;
                }
                continue;
            }
            if (cnt) cnt->nodes++;
            const float* f = reinterpret_cast<const float*>(&b.nodes[size_t(ref) * per]);
            const uint32_t* ch = reinterpret_cast<const uint32_t*>(f + 6 * W);
            for (int i = 0; i < W; ++i) {
                float tn;
                if (ch[i] != kEmpty && childHit(f, W, i, r, tmax, tn)) stk[sp++] = ch[i];
            }
        }
    }
    const size_t np = g.planeIds.size();
    for (size_t k = 0; k < np; ++k) {
        const float* pl = &g.planes[4 * k];
        float t;
        if (planeTest(r.o, r.d, V{pl[0], pl[1], pl[2]}, pl[3], tmin, tmax, t)) return true;
    }
    return false;
}

}  // namespace rbr

namespace rbr {

void quantizeWideBVH(WideBVH& b) {
    const int W = b.width, per = b.f4PerNode(), dw = qNodeDwords(W);
    b.qnodes.assign(size_t(b.nodeCount) * dw, 0u);
#pragma omp parallel for schedule(static)
    for (long long n = 0; n < (long long)b.nodeCount; ++n)
        quantizeNode(reinterpret_cast<const float*>(&b.nodes[size_t(n) * per]), W, &b.qnodes[size_t(n) * dw]);
    b.quantized = true;
}

TraceCounters hostCount(const WideBVH& b, const HostGeometry& g, const float* rays, size_t n, bool any) {
    uint64_t nodes = 0, prims = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : nodes, prims)
    for (long long i = 0; i < (long long)n; ++i) {
        const float* r = rays + 8 * size_t(i);
        TraceCounters c;
        float t;
        if (any) hostAny(b, g, r, r + 3, r[6], r[7], &c);
        else hostClosest(b, g, r, r + 3, r[6], r[7], &t, &c);
        nodes += c.nodes;
        prims += c.prims;
    }
    TraceCounters out;
    out.nodes = nodes;
    out.prims = prims;
    return out;
}

}  // namespace rbr
