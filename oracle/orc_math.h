// CPU restatement of the reference's FP32 vector math.  TEST INFRASTRUCTURE ONLY: this
// directory is the oracle that checks the HIP path; nothing in the product links it.
//
// Every operation keeps the reference's operand order (compile with -ffp-contract=off so
// no multiply-add is fused), because closest-hit ids and t must come out bit-identical.
//   Vec3 ops, dot, cross, normalized:     include/render/core/vector.h:112-289
//   division by a scalar = multiply by 1/s: vector.h:153-156
//   coordinateSystem:                      vector.h:282-289
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>

namespace orc {

constexpr float PI = 3.14159265358979323846f;
constexpr float TWO_PI = 2.0f * PI;
constexpr float INV_PI = 1.0f / PI;
constexpr float INV_TWO_PI = 1.0f / TWO_PI;
constexpr float EPSILON = 1e-4f;                       // common.h:24
constexpr float INF = std::numeric_limits<float>::infinity();
constexpr float MAXF = std::numeric_limits<float>::max();

struct V3 {
    float x, y, z;
    V3() : x(0.0f), y(0.0f), z(0.0f) {}
    explicit V3(float s) : x(s), y(s), z(s) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    V3 operator-() const { return V3(-x, -y, -z); }
    V3 operator+(const V3& v) const { return V3(x + v.x, y + v.y, z + v.z); }
    V3 operator-(const V3& v) const { return V3(x - v.x, y - v.y, z - v.z); }
    V3 operator*(float s) const { return V3(x * s, y * s, z * s); }
    V3 operator*(const V3& v) const { return V3(x * v.x, y * v.y, z * v.z); }
    V3 operator/(float s) const { float inv = 1.0f / s; return V3(x * inv, y * inv, z * inv); }
    V3 operator/(const V3& v) const { return V3(x / v.x, y / v.y, z / v.z); }
    V3& operator+=(const V3& v) { x += v.x; y += v.y; z += v.z; return *this; }
    V3& operator*=(float s) { x *= s; y *= s; z *= s; return *this; }
    V3& operator*=(const V3& v) { x *= v.x; y *= v.y; z *= v.z; return *this; }
    V3& operator/=(float s) { float inv = 1.0f / s; x *= inv; y *= inv; z *= inv; return *this; }
    float lengthSquared() const { return x * x + y * y + z * z; }
    float length() const { return std::sqrt(lengthSquared()); }
    V3 normalized() const { float len = length(); return len > 0.0f ? *this / len : V3(0.0f); }
    float dot(const V3& v) const { return x * v.x + y * v.y + z * v.z; }
    V3 cross(const V3& v) const { return V3(y * v.z - z * v.y, z * v.x - x * v.z, x * v.y - y * v.x); }
    bool isBlack() const { return x <= 0.0f && y <= 0.0f && z <= 0.0f; }
    float maxComponent() const { return std::max({x, y, z}); }
    static V3 min(const V3& a, const V3& b) { return V3(std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)); }
    static V3 max(const V3& a, const V3& b) { return V3(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)); }
};

inline V3 operator*(float s, const V3& v) { return v * s; }
inline float dot(const V3& a, const V3& b) { return a.dot(b); }
inline V3 cross(const V3& a, const V3& b) { return a.cross(b); }
inline V3 normalize(const V3& v) { return v.normalized(); }
inline float clampf(float v, float lo, float hi) { return std::max(lo, std::min(hi, v)); }

inline void coordinateSystem(const V3& n, V3& t, V3& b) {
    if (std::abs(n.x) > std::abs(n.y)) t = V3(-n.z, 0.0f, n.x).normalized();
    else t = V3(0.0f, n.z, -n.y).normalized();
    b = cross(n, t);
}

// Ray with a mutable far limit (ray.h:10-35).  The oracle never re-normalizes directions:
// the C1 ray sets are normalized once when generated.
struct Ray {
    V3 o, d;
    float tmin = EPSILON;
    mutable float tMax = INF;
    V3 at(float t) const { return o + d * t; }
};

// BoundingBox (bbox.h:19-120).
struct Box {
    V3 lo{MAXF}, hi{-MAXF};
    Box() = default;
    Box(const V3& a, const V3& b) : lo(V3::min(a, b)), hi(V3::max(a, b)) {}
    V3 center() const { return (lo + hi) * 0.5f; }
    V3 diagonal() const { return hi - lo; }
    float surfaceArea() const { V3 d = diagonal(); return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x); }
    int maximumExtent() const {
        V3 d = diagonal();
        if (d.x > d.y && d.x > d.z) return 0;
        if (d.y > d.z) return 1;
        return 2;
    }
    bool isEmpty() const { return lo.x > hi.x || lo.y > hi.y || lo.z > hi.z; }
    void expand(const V3& p) { lo = V3::min(lo, p); hi = V3::max(hi, p); }
    void expand(const Box& b) { lo = V3::min(lo, b.lo); hi = V3::max(hi, b.hi); }
    // Slab test on [0, tMax] (bbox.h:92-120).
    bool intersect(const Ray& r, float& t0o, float& t1o) const {
        float t0 = 0.0f, t1 = r.tMax;
        for (int i = 0; i < 3; ++i) {
            const float dir = r.d[i];
            if (dir == 0.0f) {
                if (r.o[i] < lo[i] || r.o[i] > hi[i]) return false;
                continue;
            }
            const float inv = 1.0f / dir;
            float tn = (lo[i] - r.o[i]) * inv;
            float tf = (hi[i] - r.o[i]) * inv;
            if (tn > tf) std::swap(tn, tf);
            tf *= 1.0f + 2.0f * std::numeric_limits<float>::epsilon();
            t0 = std::max(t0, tn);
            t1 = std::min(t1, tf);
            if (t0 > t1) return false;
        }
        t0o = t0;
        t1o = t1;
        return true;
    }
};

}  // namespace orc
