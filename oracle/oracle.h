/*
 * oracle.h — C entry points of the CPU oracle (liboracle.so).
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle restates the reference renderer's algorithms on the
 * CPU (FP32, no fused multiply-add) and is used by the tests, the smoke check and the CPU
 * baseline leg of the benchmark as the checker.  The product library never links it.
 *
 * Ray / hit layouts are those of include/rbr.h (32-byte ray, 8-byte hit).
 */
#ifndef RBR_ORACLE_H
#define RBR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/rbr.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

/* Same argument layout as rbr_scene_set_geometry. */
orc_scene* orc_scene_create(const float* tris, uint32_t ntri, const uint32_t* tri_ids,
                            const float* spheres, uint32_t nsph, const uint32_t* sph_ids,
                            const float* quads, uint32_t nquad, const uint32_t* quad_ids,
                            const float* planes, uint32_t nplane, const uint32_t* plane_ids);
void orc_scene_destroy(orc_scene* s);
void orc_scene_build_bvh(orc_scene* s);

/* Lowest-id closest hit / any hit, by brute force or through the oracle BVH.  nthreads <= 0
 * uses every OpenMP thread. */
void orc_closest_brute(const orc_scene* s, const rbr_ray* rays, size_t n, rbr_hit* hits, int nthreads);
void orc_any_brute(const orc_scene* s, const rbr_ray* rays, size_t n, uint8_t* occ, int nthreads);
void orc_closest_bvh(const orc_scene* s, const rbr_ray* rays, size_t n, rbr_hit* hits, int nthreads);
void orc_any_bvh(const orc_scene* s, const rbr_ray* rays, size_t n, uint8_t* occ, int nthreads);

/* Moller-Trumbore on one triangle (9 floats) exactly as the oracle evaluates it; returns 1 and
 * t on a hit in [tmin, tmax]. */
int orc_triangle_t(const float* tri9, const rbr_ray* ray, float* t);

int orc_max_threads(void);

#ifdef __cplusplus
}
#endif

#endif
