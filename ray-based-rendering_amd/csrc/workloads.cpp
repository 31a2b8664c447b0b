// Seeded synthetic workloads (host C++): the displaced icosphere mesh and the ray batches
// of the closest-hit / any-hit benchmark, plus the counter-based uniform stream they use.
#include "rbr_workloads.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline float u01(uint64_t seed, uint64_t idx) {
    uint64_t h = mix64(seed * 0xD1342543DE82EF95ull ^ mix64(idx));
    return float(h >> 40) * (1.0f / 16777216.0f);
}

inline float lattice(int x, int y, int z, uint64_t seed) {
    uint64_t h = mix64(seed ^ mix64((uint64_t(uint32_t(x)) << 42) ^ (uint64_t(uint32_t(y)) << 21) ^ uint64_t(uint32_t(z))));
    return float(h >> 40) * (2.0f / 16777216.0f) - 1.0f;
}

inline float fade(float t) { return t * t * (3.0f - 2.0f * t); }

float valueNoise(float x, float y, float z, uint64_t seed) {
    float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int ix = int(fx), iy = int(fy), iz = int(fz);
    float tx = fade(x - fx), ty = fade(y - fy), tz = fade(z - fz);
    float c[8];
    for (int k = 0; k < 8; ++k) c[k] = lattice(ix + (k & 1), iy + ((k >> 1) & 1), iz + ((k >> 2) & 1), seed);
    float x00 = c[0] + (c[1] - c[0]) * tx, x10 = c[2] + (c[3] - c[2]) * tx;
    float x01 = c[4] + (c[5] - c[4]) * tx, x11 = c[6] + (c[7] - c[6]) * tx;
    float y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}

struct P3 { float x, y, z; };

const float kT = 1.6180339887498949f;
const P3 kIcoV[12] = {{-1, kT, 0}, {1, kT, 0}, {-1, -kT, 0}, {1, -kT, 0}, {0, -1, kT}, {0, 1, kT},
                      {0, -1, -kT}, {0, 1, -kT}, {kT, 0, -1}, {kT, 0, 1}, {-kT, 0, -1}, {-kT, 0, 1}};
const int kIcoF[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                          {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                          {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};

}  // namespace

extern "C" {

float rbrw_u01(uint64_t seed, uint64_t index) { return u01(seed, index); }

float rbrw_noise3(float x, float y, float z, uint64_t seed) {
    float n = 1.0f * valueNoise(4.0f * x, 4.0f * y, 4.0f * z, seed) +
              0.5f * valueNoise(8.0f * x, 8.0f * y, 8.0f * z, seed + 1) +
              0.25f * valueNoise(16.0f * x, 16.0f * y, 16.0f * z, seed + 2);
    return n * (1.0f / 1.75f);
}

size_t rbrw_icosphere(int n, float radius, float amp, uint64_t seed, float* tris) {
    if (n < 1) return 0;
    const size_t perFace = size_t(n) * size_t(n);
    if (!tris) return 20 * perFace;
    P3 unit[12];
    for (int i = 0; i < 12; ++i) {
        float l = std::sqrt(kIcoV[i].x * kIcoV[i].x + kIcoV[i].y * kIcoV[i].y + kIcoV[i].z * kIcoV[i].z);
        unit[i] = {kIcoV[i].x / l, kIcoV[i].y / l, kIcoV[i].z / l};
    }
    const float fn = float(n);
#pragma omp parallel for schedule(dynamic, 1)
    for (int f = 0; f < 20; ++f) {
        const int a = kIcoF[f][0], b = kIcoF[f][1], c = kIcoF[f][2];
        // Vertex from integer barycentrics, summed in ascending corner-id order so a vertex
        // shared by two faces comes out bit-identical on both.
        auto vertex = [&](int i, int j) -> P3 {
            std::pair<int, int> w[3] = {{a, n - i - j}, {b, i}, {c, j}};
            std::sort(w, w + 3);
            float px = 0.0f, py = 0.0f, pz = 0.0f;
            for (int k = 0; k < 3; ++k) {
                float s = float(w[k].second) / fn;
                px += unit[w[k].first].x * s;
                py += unit[w[k].first].y * s;
                pz += unit[w[k].first].z * s;
            }
            float l = std::sqrt(px * px + py * py + pz * pz);
            px /= l; py /= l; pz /= l;
            float r = radius * (1.0f + amp * rbrw_noise3(px, py, pz, seed));
            return {px * r, py * r, pz * r};
        };
        float* out = tris + size_t(f) * perFace * 9;
        size_t k = 0;
        auto emit = [&](const P3& p0, const P3& p1, const P3& p2) {
            float* o = out + 9 * k++;
            o[0] = p0.x; o[1] = p0.y; o[2] = p0.z;
            o[3] = p1.x; o[4] = p1.y; o[5] = p1.z;
            o[6] = p2.x; o[7] = p2.y; o[8] = p2.z;
        };
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j + i < n; ++j) {
                P3 p00 = vertex(i, j), p10 = vertex(i + 1, j), p01 = vertex(i, j + 1);
                emit(p00, p10, p01);
                if (i + j < n - 1) emit(p10, vertex(i + 1, j + 1), p01);
            }
        }
    }
    return 20 * perFace;
}

void rbrw_rays_uniform(const float lo[3], const float hi[3], float scale, size_t n, uint64_t seed,
                       float tmin, float tmax, rbr_ray* rays) {
    float c[3], h[3];
    for (int k = 0; k < 3; ++k) {
        c[k] = 0.5f * (lo[k] + hi[k]);
        h[k] = 0.5f * (hi[k] - lo[k]) * scale;
    }
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        uint64_t base = uint64_t(i) * 5;
        rbr_ray& r = rays[i];
        for (int k = 0; k < 3; ++k) r.o[k] = c[k] + h[k] * (2.0f * u01(seed, base + k) - 1.0f);
        float z = 1.0f - 2.0f * u01(seed, base + 3);
        float phi = 6.28318530717958647692f * u01(seed, base + 4);
        float s = std::sqrt(std::max(0.0f, 1.0f - z * z));
        float d[3] = {s * std::cos(phi), s * std::sin(phi), z};
        float inv = 1.0f / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int k = 0; k < 3; ++k) r.d[k] = d[k] * inv;
        r.tmin = tmin;
        r.tmax = tmax;
    }
}

void rbrw_rays_shadow(const float lo[3], const float hi[3], size_t n, uint64_t seed, float tmin,
                      rbr_ray* rays) {
    float c[3], diag2 = 0.0f;
    for (int k = 0; k < 3; ++k) {
        c[k] = 0.5f * (lo[k] + hi[k]);
        diag2 += (hi[k] - lo[k]) * (hi[k] - lo[k]);
    }
    const float R = 0.5f * std::sqrt(diag2);
    const float light[3] = {c[0] + 0.3f * 3.0f * R, c[1] + 1.0f * 3.0f * R, c[2] + 0.2f * 3.0f * R};
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        uint64_t base = uint64_t(i) * 3;
        rbr_ray& r = rays[i];
        float to[3];
        for (int k = 0; k < 3; ++k) {
            r.o[k] = lo[k] + u01(seed, base + k) * (hi[k] - lo[k]);
            to[k] = light[k] - r.o[k];
        }
        float dist = std::sqrt(to[0] * to[0] + to[1] * to[1] + to[2] * to[2]);
        float inv = 1.0f / dist;
        for (int k = 0; k < 3; ++k) r.d[k] = to[k] * inv;
        r.tmin = tmin;
        r.tmax = dist * 0.999f;
    }
}

}  // extern "C"
