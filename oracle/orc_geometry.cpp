// CPU restatement of the reference's primitive tests, brute-force loop and BVH.
// TEST INFRASTRUCTURE ONLY (see orc_scene.h for the file:line map).
#include "orc_scene.h"

#include <algorithm>

namespace orc {

// Triangle::intersect, triangle.cpp:6-24.  Same operand order: e1 = v1 - v0,
// e2 = v2 - v0, h = d x e2, a = e1.h, |a| < EPSILON rejects, f = 1/a, ...
bool tri_t(const Tri& tr, const Ray& r, float& tout) {
    V3 e1 = tr.v1 - tr.v0;
    V3 e2 = tr.v2 - tr.v0;
    V3 h = cross(r.d, e2);
    float a = dot(e1, h);
    if (std::abs(a) < EPSILON) return false;
    float f = 1.0f / a;
    V3 s = r.o - tr.v0;
    float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    V3 q = cross(s, e1);
    float v = f * dot(r.d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = f * dot(e2, q);
    if (t < r.tmin || t > r.tMax) return false;
    tout = t;
    return true;
}

// Sphere::intersect, sphere.cpp:6-21 (near root first, then far root).
bool sph_t(const Sph& sp, const Ray& r, float& tout) {
    V3 oc = r.o - sp.c;
    float a = dot(r.d, r.d);
    float halfB = dot(oc, r.d);
    float c = dot(oc, oc) - sp.r * sp.r;
    float disc = halfB * halfB - a * c;
    if (disc < 0.0f) return false;
    float sq = std::sqrt(disc);
    float t = (-halfB - sq) / a;
    if (t < r.tmin || t > r.tMax) {
        t = (-halfB + sq) / a;
        if (t < r.tmin || t > r.tMax) return false;
    }
    tout = t;
    return true;
}

// Quad::intersect, quad.cpp:12-31.
bool quad_t(const Quad& q, const Ray& r, float& tout) {
    float denom = dot(q.n, r.d);
    if (std::abs(denom) < EPSILON) return false;
    float t = dot(q.origin - r.o, q.n) / denom;
    if (t < r.tmin || t > r.tMax) return false;
    V3 hp = r.at(t);
    V3 lp = hp - q.origin;
    float uLen2 = dot(q.u, q.u);
    float vLen2 = dot(q.v, q.v);
    float uc = dot(lp, q.u) / uLen2;
    float vc = dot(lp, q.v) / vLen2;
    if (uc < 0.0f || uc > 1.0f || vc < 0.0f || vc > 1.0f) return false;
    tout = t;
    return true;
}

// Plane::intersect, plane.cpp:12-17.
bool plane_t(const Plane& p, const Ray& r, float& tout) {
    float denom = dot(p.n, r.d);
    if (std::abs(denom) < EPSILON) return false;
    float t = -(dot(p.n, r.o) + p.d) / denom;
    if (t < r.tmin || t > r.tMax) return false;
    tout = t;
    return true;
}

Box Scene::primBounds(const PrimRef& p) const {
    switch (p.type) {
        case PT_TRI: {  // Triangle::getBounds, triangle.cpp:68-73
            const Tri& t = tris[p.index];
            return Box(V3::min(V3::min(t.v0, t.v1), t.v2), V3::max(V3::max(t.v0, t.v1), t.v2));
        }
        case PT_SPHERE: {  // Sphere::getBounds, sphere.cpp:45-50
            const Sph& s = sphs[p.index];
            return Box(s.c - V3(s.r), s.c + V3(s.r));
        }
        case PT_QUAD: {  // Quad::getBounds, quad.cpp:44-51
            const Quad& q = quads[p.index];
            Box b;
            b.expand(q.origin);
            b.expand(q.origin + q.u);
            b.expand(q.origin + q.v);
            b.expand(q.origin + q.u + q.v);
            return b;
        }
        default:
            return Box(V3(-MAXF), V3(MAXF));
    }
}

bool Scene::primT(const PrimRef& p, const Ray& r, float& t) const {
    switch (p.type) {
        case PT_TRI: return tri_t(tris[p.index], r, t);
        case PT_SPHERE: return sph_t(sphs[p.index], r, t);
        case PT_QUAD: return quad_t(quads[p.index], r, t);
        default: return plane_t(planes[p.index], r, t);
    }
}

uint32_t Scene::primId(const PrimRef& p) const {
    switch (p.type) {
        case PT_TRI: return tris[p.index].id;
        case PT_SPHERE: return sphs[p.index].id;
        case PT_QUAD: return quads[p.index].id;
        default: return planes[p.index].id;
    }
}

std::vector<PrimRef> Scene::allBounded() const {
    std::vector<PrimRef> v;
    v.reserve(tris.size() + sphs.size() + quads.size());
    for (uint32_t i = 0; i < tris.size(); ++i) v.push_back({PT_TRI, i});
    for (uint32_t i = 0; i < sphs.size(); ++i) v.push_back({PT_SPHERE, i});
    for (uint32_t i = 0; i < quads.size(); ++i) v.push_back({PT_QUAD, i});
    return v;
}

void Scene::indexIds() {
    uint32_t maxId = 0;
    bool any = false;
    auto upd = [&](uint32_t id) { maxId = std::max(maxId, id); any = true; };
    for (auto& t : tris) upd(t.id);
    for (auto& s : sphs) upd(s.id);
    for (auto& q : quads) upd(q.id);
    for (auto& p : planes) upd(p.id);
    byId.assign(any ? maxId + 1 : 0, PrimRef{0xFFFFFFFFu, 0});
    for (uint32_t i = 0; i < tris.size(); ++i) byId[tris[i].id] = {PT_TRI, i};
    for (uint32_t i = 0; i < sphs.size(); ++i) byId[sphs[i].id] = {PT_SPHERE, i};
    for (uint32_t i = 0; i < quads.size(); ++i) byId[quads[i].id] = {PT_QUAD, i};
    for (uint32_t i = 0; i < planes.size(); ++i) byId[planes[i].id] = {PT_PLANE, i};
}

// ---------------------------------------------------------------- brute force

static inline void acceptLowest(uint32_t id, float t, uint32_t& bestId, float& bestT) {
    if (t < bestT || (t == bestT && id < bestId)) {
        bestT = t;
        bestId = id;
    }
}

uint32_t Scene::closestBrute(const Ray& r, float& tout) const {
    Ray probe = r;
    uint32_t bestId = 0xFFFFFFFFu;
    float bestT = r.tMax;
    float t;
    for (const Tri& x : tris) { probe.tMax = bestT; if (tri_t(x, probe, t)) acceptLowest(x.id, t, bestId, bestT); }
    for (const Sph& x : sphs) { probe.tMax = bestT; if (sph_t(x, probe, t)) acceptLowest(x.id, t, bestId, bestT); }
    for (const Quad& x : quads) { probe.tMax = bestT; if (quad_t(x, probe, t)) acceptLowest(x.id, t, bestId, bestT); }
    for (const Plane& x : planes) { probe.tMax = bestT; if (plane_t(x, probe, t)) acceptLowest(x.id, t, bestId, bestT); }
    tout = bestT;
    return bestId;
}

bool Scene::anyBrute(const Ray& r) const {
    float t;
    for (const Tri& x : tris) if (tri_t(x, r, t)) return true;
    for (const Sph& x : sphs) if (sph_t(x, r, t)) return true;
    for (const Quad& x : quads) if (quad_t(x, r, t)) return true;
    for (const Plane& x : planes) if (plane_t(x, r, t)) return true;
    return false;
}

// ---------------------------------------------------------------- BVH build (bvh.cpp)

namespace {

struct PrimInfo {
    size_t index;
    Box bounds;
    V3 centroid() const { return bounds.center(); }
};

constexpr size_t MAX_PRIMS_IN_NODE = 4;   // bvh.h:41
constexpr int MAX_DEPTH = 32;             // bvh.h:42

size_t findBestSplit(std::vector<PrimInfo>& pi, size_t start, size_t end, const Box& bounds,
                     const Box& cb, int dim) {
    constexpr int nBuckets = 12;
    struct Bucket { int count = 0; Box bounds; };
    Bucket buckets[nBuckets];
    for (size_t i = start; i < end; ++i) {
        int b = nBuckets * ((pi[i].centroid()[dim] - cb.lo[dim]) / (cb.hi[dim] - cb.lo[dim]));
        if (b == nBuckets) b = nBuckets - 1;
        buckets[b].count++;
        buckets[b].bounds.expand(pi[i].bounds);
    }
    float cost[nBuckets - 1];
    for (int i = 0; i < nBuckets - 1; ++i) {
        Box b0, b1;
        int count0 = 0, count1 = 0;
        for (int j = 0; j <= i; ++j) { b0.expand(buckets[j].bounds); count0 += buckets[j].count; }
        for (int j = i + 1; j < nBuckets; ++j) { b1.expand(buckets[j].bounds); count1 += buckets[j].count; }
        cost[i] = 1.0f + (count0 * b0.surfaceArea() + count1 * b1.surfaceArea()) / bounds.surfaceArea();
    }
    float minCost = cost[0];
    int minCostBucket = 0;
    for (int i = 1; i < nBuckets - 1; ++i) {
        if (cost[i] < minCost) { minCost = cost[i]; minCostBucket = i; }
    }
    float leafCost = static_cast<float>(end - start);
    if (minCost >= leafCost) return (start + end) / 2;   // bvh.cpp:193-196 (range left unpartitioned)
    auto midIter = std::partition(pi.begin() + start, pi.begin() + end, [=](const PrimInfo& p) {
        int b = nBuckets * ((p.centroid()[dim] - cb.lo[dim]) / (cb.hi[dim] - cb.lo[dim]));
        if (b == nBuckets) b = nBuckets - 1;
        return b <= minCostBucket;
    });
    return midIter - pi.begin();
}

struct Builder {
    const std::vector<PrimRef>& src;
    std::vector<PrimRef>& ordered;
    float pad;

    Box padBox(const Box& b) const {
        Box p = b;
        for (int i = 0; i < 3; ++i) {
            p.lo[i] = b.lo[i] - (pad + std::abs(b.lo[i]) * 1e-6f);
            p.hi[i] = b.hi[i] + (pad + std::abs(b.hi[i]) * 1e-6f);
        }
        return p;
    }

    std::unique_ptr<BNode> build(std::vector<PrimInfo>& pi, size_t start, size_t end, int depth) {
        auto node = std::make_unique<BNode>();
        Box bounds;
        for (size_t i = start; i < end; ++i) bounds.expand(pi[i].bounds);
        node->bounds = bounds;
        node->padded = padBox(bounds);
        size_t n = end - start;
        auto makeLeaf = [&]() {
            node->start = ordered.size();
            node->count = n;
            for (size_t i = start; i < end; ++i) ordered.push_back(src[pi[i].index]);
        };
        if (n <= MAX_PRIMS_IN_NODE || depth >= MAX_DEPTH) { makeLeaf(); return node; }
        Box cb;
        for (size_t i = start; i < end; ++i) cb.expand(pi[i].centroid());
        int dim = cb.maximumExtent();
        if (cb.hi[dim] == cb.lo[dim]) { makeLeaf(); return node; }
        size_t mid;
        if (n <= 2) {
            mid = (start + end) / 2;
            std::nth_element(pi.begin() + start, pi.begin() + mid, pi.begin() + end,
                             [dim](const PrimInfo& a, const PrimInfo& b) { return a.centroid()[dim] < b.centroid()[dim]; });
        } else {
            mid = findBestSplit(pi, start, end, bounds, cb, dim);
        }
        node->left = build(pi, start, mid, depth + 1);
        node->right = build(pi, mid, end, depth + 1);
        return node;
    }
};

}  // namespace

void Scene::buildBVH() {
    std::vector<PrimRef> src = allBounded();
    ordered.clear();
    root.reset();
    if (src.empty()) return;
    std::vector<PrimInfo> pi(src.size());
    Box world;
    for (size_t i = 0; i < src.size(); ++i) {
        pi[i] = {i, primBounds(src[i])};
        world.expand(pi[i].bounds);
    }
    V3 ext = world.diagonal();
    pad = 1e-5f * std::max({ext.x, ext.y, ext.z}) + 1e-6f;
    ordered.reserve(src.size());
    Builder b{src, ordered, pad};
    root = b.build(pi, 0, src.size(), 0);
}

// ---------------------------------------------------------------- BVH traversal

namespace {

struct Walker {
    const Scene& s;
    const Ray& ray;
    uint32_t bestId = 0xFFFFFFFFu;
    float bestT;

    // bvh.cpp:211-248: box test, leaf loop, then both children nearest-first.
    void visit(const BNode* node) {
        Ray probe = ray;
        probe.tMax = bestT;
        float a, b;
        if (!node->padded.intersect(probe, a, b)) return;
        if (node->leaf()) {
            for (size_t i = 0; i < node->count; ++i) {
                const PrimRef& p = s.ordered[node->start + i];
                probe.tMax = bestT;
                float t;
                if (s.primT(p, probe, t)) acceptLowest(s.primId(p), t, bestId, bestT);
            }
            return;
        }
        float tl, tr, x;
        bool hl = node->left->padded.intersect(probe, tl, x);
        bool hr = node->right->padded.intersect(probe, tr, x);
        const BNode* first = node->left.get();
        const BNode* second = node->right.get();
        if (hl && hr && tl > tr) std::swap(first, second);
        if (hl || hr) {
            visit(first);
            visit(second);
        }
    }

    bool visitP(const BNode* node) {
        float a, b;
        if (!node->padded.intersect(ray, a, b)) return false;
        if (node->leaf()) {
            for (size_t i = 0; i < node->count; ++i) {
                float t;
                if (s.primT(s.ordered[node->start + i], ray, t)) return true;
            }
            return false;
        }
        return visitP(node->left.get()) || visitP(node->right.get());
    }
};

}  // namespace

uint32_t Scene::closestBVH(const Ray& r, float& tout) const {
    Walker w{*this, r};
    w.bestT = r.tMax;
    if (root) w.visit(root.get());
    Ray probe = r;
    float t;
    for (const Plane& p : planes) {
        probe.tMax = w.bestT;
        if (plane_t(p, probe, t)) acceptLowest(p.id, t, w.bestId, w.bestT);
    }
    tout = w.bestT;
    return w.bestId;
}

bool Scene::anyBVH(const Ray& r) const {
    Walker w{*this, r};
    if (root && w.visitP(root.get())) return true;
    float t;
    for (const Plane& p : planes) if (plane_t(p, r, t)) return true;
    return false;
}

// ---------------------------------------------------------------- interaction fill

bool Scene::fill(uint32_t id, const Ray& r, float t, Isect& is) const {
    if (id >= byId.size()) return false;
    const PrimRef& p = byId[id];
    is.t = t;
    is.wo = -r.d;
    is.id = id;
    switch (p.type) {
        case PT_TRI: {  // triangle.cpp:30-51
            const Tri& tr = tris[p.index];
            is.p = r.at(t);
            V3 e1 = tr.v1 - tr.v0, e2 = tr.v2 - tr.v0;
            V3 gn = cross(e1, e2).normalized();
            if (!tr.smooth) {
                is.setFaceNormal(r, gn);
                return true;
            }
            // barycentrics of the Moller-Trumbore terms (triangle.cpp:7-28), then the
            // interpolated normal turned into the geometric hemisphere (triangle.cpp:40-48)
            V3 h = cross(r.d, e2);
            float a = dot(e1, h);
            float f = 1.0f / a;
            V3 sv = r.o - tr.v0;
            float u = f * dot(sv, h);
            V3 q = cross(sv, e1);
            float v = f * dot(r.d, q);
            float w = 1.0f - u - v;
            V3 n = (tr.n0 * w + tr.n1 * u + tr.n2 * v).normalized();
            if (dot(n, gn) < 0.0f) n = -n;
            is.setFaceNormal(r, n);
            return true;
        }
        case PT_SPHERE: {  // sphere.cpp:24-35
            const Sph& s = sphs[p.index];
            V3 hp = r.at(t);
            V3 nn = (hp - s.c) / s.r;
            is.p = hp;
            is.setFaceNormal(r, nn);
            return true;
        }
        case PT_QUAD: {  // quad.cpp:25-37
            const Quad& q = quads[p.index];
            is.p = r.at(t);
            is.setFaceNormal(r, q.n);
            return true;
        }
        case PT_PLANE: {  // plane.cpp:19-26
            const Plane& pl = planes[p.index];
            is.p = r.at(t);
            is.setFaceNormal(r, pl.n);
            return true;
        }
        default:
            return false;
    }
}

}  // namespace orc
