/*
 * rbr_workloads.h — seeded synthetic inputs shared by the tests, the benchmark and the
 * harness programs (librbr_workloads.so, host-only C++).  Every generator is deterministic
 * for a given seed and independent of the thread count.
 */
#ifndef RBR_WORKLOADS_H
#define RBR_WORKLOADS_H

#include <stddef.h>
#include <stdint.h>

#include "rbr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Geodesic icosphere: each of the 20 icosahedron faces split into n*n triangles (20*n*n
 * triangles), vertices on the sphere of `radius` about the origin, displaced radially by
 * radius*(1 + amp*noise(dir)) where noise is 3-octave seeded value noise (frequencies
 * 4/8/16, amplitudes 1/0.5/0.25, normalized to [-1,1]).  Shared vertices are computed
 * identically on both faces, so the mesh is watertight.  `tris` receives 9 floats per
 * triangle (v0 v1 v2).  Returns the triangle count. */
size_t rbrw_icosphere(int n, float radius, float amp, uint64_t seed, float* tris);

/* 3-octave value noise on a unit direction, in [-1, 1]. */
float rbrw_noise3(float x, float y, float z, uint64_t seed);

/* Closest-hit batch: origins uniform in the box [lo,hi] scaled by `scale` about its centre,
 * directions uniform on the sphere (z = 1-2u1, phi = 2*pi*u2) normalized once; tmin, tmax
 * as given. */
void rbrw_rays_uniform(const float lo[3], const float hi[3], float scale, size_t n, uint64_t seed,
                       float tmin, float tmax, rbr_ray* rays);

/* Shadow batch: origins uniform in [lo,hi], aimed at c + (0.3,1,0.2)*3R (c = box centre,
 * R = half the box diagonal, as apps/accelerator_compare.cpp:80-96), tmax = 0.999*dist. */
void rbrw_rays_shadow(const float lo[3], const float hi[3], size_t n, uint64_t seed, float tmin,
                      rbr_ray* rays);

/* Uniform float in [0,1) from (seed, index): the generators' counter-based stream. */
float rbrw_u01(uint64_t seed, uint64_t index);

#ifdef __cplusplus
}
#endif

#endif
