/*
 * rbr.h — C ABI of the MI355X (gfx950) ray-traversal and path-tracing library.
 *
 * Everything here is plain C: opaque handles, caller-owned host arrays, int status codes
 * (RBR_OK == 0) and a per-context error string.  No C++ types, exceptions or torch types
 * cross this boundary.  One context per host thread (calls on one context are not
 * reentrant); several contexts may share a device.
 *
 * Each entry point names the reference interface it stands behind (HuynhNguyenPhuc/
 * Ray-Based-Rendering, C++17).  The C++ adapter classes that a maintainer adds to the
 * reference (GpuBVH : Accelerator, GpuPathTracingIntegrator : Integrator) call only these.
 */
#ifndef RBR_H
#define RBR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
enum {
    RBR_OK = 0,
    RBR_ERR_INVALID = 1,   /* bad argument (null pointer, size out of range, bad enum) */
    RBR_ERR_HIP = 2,       /* a HIP runtime call failed; rbr_last_error() has the text   */
    RBR_ERR_STATE = 3,     /* call out of order (e.g. trace before a BVH was built)      */
    RBR_ERR_NOMEM = 4,     /* device or host allocation failed                           */
    RBR_ERR_UNSUPPORTED = 5,/* feature not handled by the GPU path (caller falls back)   */
    RBR_ERR_WATCHDOG = 6,  /* a ray exceeded the traversal step cap (rbr_set_trace_step_cap):
                              the BVH is corrupt or the cap below its size; the results of the
                              launches since the last check are incomplete                  */
    RBR_ERR_IO = 7         /* a file could not be opened, read or written, or is not of the
                              expected format (rbr_guide_save / rbr_guide_load)             */
};

#define RBR_MISS 0xFFFFFFFFu

/* ---------------------------------------------------------------- plain data */

/* One ray, 32 bytes.  The direction is used as given: it is never re-normalized on the
 * device (the reference's Ray constructor normalizes, ray.h:19-23; callers that build
 * rays that way pass the normalized direction).  A hit is accepted for t in [tmin, tmax],
 * both ends inclusive, as Triangle::intersect does (triangle.cpp:24, with tmin = EPSILON). */
typedef struct rbr_ray {
    float o[3];
    float d[3];
    float tmin;
    float tmax;
} rbr_ray;

/* Closest hit, 8 bytes: id is the primitive id given at upload (RBR_MISS on a miss);
 * among hits with equal t the lowest id wins. */
typedef struct rbr_hit {
    uint32_t id;
    float t;
} rbr_hit;

typedef struct rbr_bvh_stats {
    uint32_t width;        /* 4 or 8, | RBR_BVH_QUANTIZED                */
    uint32_t node_count;   /* wide nodes                                  */
    uint32_t leaf_count;   /* leaf references                              */
    uint32_t prim_count;   /* bounded primitives in the tree               */
    uint32_t unbounded;    /* infinite planes tested linearly              */
    uint32_t max_depth;    /* wide-tree depth                              */
    uint64_t node_bytes;   /* device bytes of the node array               */
    uint64_t prim_bytes;   /* device bytes of the leaf primitive records   */
    double sah_cost;       /* SAH cost of the binary tree (Ct = 1, Ci = 1) */
    double build_ms;       /* host build time                              */
} rbr_bvh_stats;

/* Material with the reference's six parameters (material.h:30-35). */
typedef struct rbr_material {
    float base_color[3];
    float metallic;
    float roughness;
    float ior;
    float transmission;
    float emission[3];
} rbr_material;

enum { RBR_LIGHT_AREA = 0, RBR_LIGHT_POINT = 1, RBR_LIGHT_DIRECTIONAL = 2 };

/* Light (light.h:48-63).  Area: quad centred at position with edges u, v; normal is
 * normalize(cross(u, v)) as AreaLight's constructor computes it (area_light.cpp:10).
 * Directional: direction is the normalized direction FROM the light. */
typedef struct rbr_light {
    int32_t type;
    float color[3];
    float intensity;
    float position[3];
    float u[3];
    float v[3];
    float normal[3];
    float direction[3];
} rbr_light;

enum { RBR_MEDIUM_NONE = 0, RBR_MEDIUM_HOMOGENEOUS = 1, RBR_MEDIUM_EXPONENTIAL = 2 };

/* Participating medium (medium.h:52-108) with a Henyey-Greenstein phase of parameter g. */
typedef struct rbr_medium {
    int32_t type;
    float sigma_a[3];
    float sigma_s[3];
    float falloff;
    float base_height;
    float g;
} rbr_medium;

/* Pinhole camera: CameraExtrinsics (camera_extrinsics.h:18-45) + PerspectiveIntrinsics
 * (perspective_intrinsics.h:20-45).  Raster y points down. */
typedef struct rbr_camera {
    float position[3];
    float forward[3];
    float right[3];
    float up[3];
    float fx, fy, cx, cy;
    int32_t width, height;
} rbr_camera;

/* PATH / VOLPATH: PathTracingIntegrator / VolumetricPathIntegrator.  NORMAL, DEPTH, AO: the
 * reference's debug integrators (debug_integrators.cpp:5-60): shading normal mapped to
 * [0, 1], linear depth (white near), ambient occlusion by cosine-weighted rays weighted by
 * 1 - t / distance.  One rbr_render call takes either debug jobs only or path jobs only. */
enum {
    RBR_INTEGRATOR_PATH = 0,
    RBR_INTEGRATOR_VOLPATH = 1,
    RBR_INTEGRATOR_NORMAL = 2,
    RBR_INTEGRATOR_DEPTH = 3,
    RBR_INTEGRATOR_AO = 4,
    RBR_INTEGRATOR_WHITTED = 5   /* WhittedIntegrator (whitted_integrator.cpp) */
};

/* IntegratorConfig (integrator.h:22-28) plus the GPU stream seed. */
typedef struct rbr_render_config {
    int32_t integrator;     /* RBR_INTEGRATOR_*                  */
    int32_t spp;
    int32_t max_depth;
    int32_t use_nee;
    int32_t use_mis;
    int32_t guided;         /* bit 0: MLP path guiding; bit 1: record training samples (PATH only) */
    uint64_t seed;
    uint32_t sample_offset; /* first sample index (progressive)  */
    uint32_t guide_depth;   /* guided jobs: bounces 0..guide_depth-1 draw from the guide (0 = all) */
    /* debug integrators (0 = the reference's defaults, debug_integrators.h:24-47) */
    int32_t ao_samples;     /* AO rays per camera sample (16)        */
    float ao_distance;      /* AO ray length (2.0)                   */
    float depth_near;       /* DEPTH near plane (0.1)                */
    float depth_far;        /* DEPTH far plane (100.0)               */
    /* WHITTED: the scene's ambient light (Scene::ambientLight, scene.h:28; the reference's
     * default is 0.1 in every channel, here it is what the caller sets) */
    float ambient[3];
} rbr_render_config;

typedef struct rbr_ctx rbr_ctx;
typedef struct rbr_scene rbr_scene;

/* One render job: its scene, camera, settings and the caller's film (RGB float, row-major
 * y*width+x, 3 floats per pixel).  The film receives the per-pixel mean over the job's
 * samples, as Integrator::render's film.addSample(x, y, sum / spp) does (integrator.cpp:36). */
typedef struct rbr_job {
    rbr_scene* scene;
    rbr_camera camera;
    rbr_render_config config;
    float* film_rgb;
} rbr_job;

/* ---------------------------------------------------------------- context */

/* Creates a context on HIP device `device` with its own stream.  device == -1 gives a
 * host-only context: geometry, BVH build and the host traversal work; kernels do not. */
int rbr_ctx_create(int device, rbr_ctx** out);
void rbr_ctx_destroy(rbr_ctx* ctx);
/* Text of the last error on this context ("" if none).  Valid until the next call. */
const char* rbr_last_error(const rbr_ctx* ctx);
/* Name of the device the context runs on, and its compute-unit count. */
int rbr_device_info(rbr_ctx* ctx, char* name, size_t name_len, int* cu_count);

/* Device buffers for callers that keep rays/hits resident in HBM. */
int rbr_device_alloc(rbr_ctx* ctx, size_t bytes, void** dptr);
int rbr_device_free(rbr_ctx* ctx, void* dptr);
int rbr_copy_to_device(rbr_ctx* ctx, void* dst, const void* src, size_t bytes);
int rbr_copy_to_host(rbr_ctx* ctx, void* dst, const void* src, size_t bytes);
int rbr_synchronize(rbr_ctx* ctx);
/* Profiling helper with no counterpart in the reference: a plain streaming copy kernel of
 * `bytes` (multiple of 16) on the context stream, used to calibrate hardware byte counters
 * (FETCH_SIZE / WRITE_SIZE) against a kernel whose traffic is known. */
int rbr_calib_copy(rbr_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);

/* ---------------------------------------------------------------- geometry + BVH
 * Replaces Accelerator::build (accelerator.h:25) for the BVH (bvh.cpp:10-31).  The
 * primitive types are those of geometry/: Triangle (triangle.h:13-22), Sphere, Quad
 * (quad.cpp:6-10) and the unbounded Plane (plane.cpp:6-10).  Ids are the caller's
 * primitive ids (the index in the vector passed to build()); a NULL id array numbers the
 * primitives of that type consecutively after the previous types (tris, spheres, quads,
 * planes).  Rebuilding replaces the previous contents (test_accelerators.cpp:148-158). */
int rbr_scene_create(rbr_ctx* ctx, rbr_scene** out);
void rbr_scene_destroy(rbr_scene* scene);

int rbr_scene_set_geometry(rbr_scene* scene,
                           const float* tris /* 9*n: v0 v1 v2 */, uint32_t ntri, const uint32_t* tri_ids,
                           const float* spheres /* 4*n: center radius */, uint32_t nsph, const uint32_t* sph_ids,
                           const float* quads /* 12*n: origin u v normal */, uint32_t nquad, const uint32_t* quad_ids,
                           const float* planes /* 4*n: normal d */, uint32_t nplane, const uint32_t* plane_ids);
/* Vertex normals of the triangles (Triangle's constructor with n0, n1, n2, triangle.h:29-33):
 * 9 floats per triangle of the last rbr_scene_set_geometry, in its order, each normalized here
 * as the constructor does.  Such triangles shade with the normal interpolated at the hit and
 * turned towards the geometric normal (Triangle::intersect, triangle.cpp:40-48); their hits are
 * unchanged.  A triangle whose nine values are all zero keeps its geometric normal (a mesh
 * without normals in a scene with others).  NULL or 0 removes them; rbr_scene_set_geometry
 * removes them too. */
int rbr_scene_set_tri_normals(rbr_scene* scene, const float* normals, size_t ntri);
/* The shading normal Triangle::intersect gives a triangle (v0 v1 v2, 9 floats) with vertex
 * normals (n0 n1 n2, 9 floats, normalized here) where the ray o + t d meets it (the ray must
 * hit it); what the renderer uses for such triangles. */
int rbr_triangle_shading_normal(const float* verts, const float* normals, const float* o, const float* d, float* out);

/* Builds the wide BVH (width 4 or 8) over the bounded primitives, leaves of at most
 * leaf_max primitives (1..8), and uploads it (replaces BVH::build, bvh.cpp:10-31, 47-209).
 * The binary SAH tree is collapsed to `width` by the SAH-optimal dynamic program, which also
 * chooses the leaf sizes; width | RBR_BVH_GREEDY collapses it greedily instead (largest-area
 * child opened first, leaves fixed by the binary build).  width | RBR_BVH_QUANTIZED stores the
 * device nodes with 8-bit child boxes relative to each node (64 B per 4-wide node, 128 B per
 * 8-wide node instead of 128 / 256 B); the boxes are rounded outward, so hits are unchanged. */
#define RBR_BVH_QUANTIZED 0x100
#define RBR_BVH_GREEDY 0x200
/* Device build only: the binary tree by locally-ordered clustering (mutual nearest neighbours
 * within 16 positions of the Morton order, merged round by round) instead of the radix tree. */
#define RBR_BVH_PLOC 0x400
int rbr_scene_build_bvh(rbr_scene* scene, int width, int leaf_max);
/* The same layouts built on the GPU (no reference counterpart; the reference builds serially
 * on the host, bvh.cpp:47-209): Morton-ordered radix tree (LBVH), collapsed to `width` by the
 * host's SAH-optimal program (| RBR_BVH_GREEDY: the host's largest-area rule), leaves of up to
 * leaf_max Morton-contiguous primitives, quantized by the host's quantizer.  Builds in milliseconds where the host SAH build takes hundreds; the
 * tree costs more node visits per ray.  Hits are the same (they do not depend on the tree).
 * Needs a GPU context.  The build's device temporaries stay with the context (grown as needed,
 * released by rbr_ctx_destroy), so rebuilding allocates nothing once they fit; the host copies of
 * the nodes and records (host traversal, rbr_scene_device_nodes / rbr_scene_bvh_arrays, renders
 * of several scenes at once) are made on their first use, which may come from several threads. */
int rbr_scene_build_bvh_device(rbr_scene* scene, int width, int leaf_max);
/* Copies the node array as the kernels read it (layout of bvh.h; up to capacity bytes) and
 * reports its size: for tooling and layout tests.  Works on host-only contexts. */
int rbr_scene_device_nodes(const rbr_scene* scene, void* dst, size_t capacity, size_t* bytes);
/* The float wide nodes (bvh.h layout; the host traversal's copy, also for quantized builds)
 * and the leaf-ordered primitive records (48 B each), for inspection and tests; no reference
 * counterpart.  Sizes are returned in *node_bytes / *prim_bytes. */
int rbr_scene_bvh_arrays(const rbr_scene* scene, void* nodes, size_t node_capacity, void* prims, size_t prim_capacity,
                         size_t* node_bytes, size_t* prim_bytes);
int rbr_scene_bvh_stats(const rbr_scene* scene, rbr_bvh_stats* out);
/* Accelerator::worldBounds (accelerator.h:34): empty box (lo=+FLT_MAX, hi=-FLT_MAX) when
 * nothing bounded was built. */
int rbr_scene_world_bounds(const rbr_scene* scene, float lo[3], float hi[3]);

/* ---------------------------------------------------------------- traversal
 * Batched Accelerator::intersect / intersectP (accelerator.h:28-31; bvh.cpp:33-41).
 * Host-array variants copy in, trace and copy out synchronously. */
int rbr_trace_closest(rbr_scene* scene, const rbr_ray* rays, size_t n, rbr_hit* hits);
int rbr_trace_any(rbr_scene* scene, const rbr_ray* rays, size_t n, uint8_t* occluded);
/* Device-pointer variants: asynchronous on the context stream.  rbr_ctx_set_stream makes a
 * caller's hipStream_t (passed as void*) the context stream: the device-pointer entry points are
 * then the asynchronous variants on that stream (the synchronous ones wait on it).  The call
 * first waits for the work already queued; NULL returns to the context's own stream, which the
 * caller does before destroying its stream.  No counterpart in the reference (its traversal is
 * synchronous host code). */
int rbr_ctx_set_stream(rbr_ctx* ctx, void* stream);
int rbr_trace_closest_device(rbr_scene* scene, const rbr_ray* d_rays, size_t n, rbr_hit* d_hits);
int rbr_trace_any_device(rbr_scene* scene, const rbr_ray* d_rays, size_t n, uint8_t* d_occluded);
/* rbr_trace_closest_device without any allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same hits, bit for bit.  Its batch counters and
 * stack spill rows live in a fixed 64 MiB scratch array that each GPU the library runs on
 * holds until the process ends; the context's waves per CU are lowered to what that scratch
 * holds for the BVH's depth (RBR_ERR_UNSUPPORTED if not even one wave per compute unit
 * fits), n is limited to 2^31 as in the other entry points.  All calls on one GPU share that
 * scratch, so calls that overlap on one GPU (other streams, other contexts, other host
 * threads) race; serialize them. */
int rbr_trace_closest_device_noalloc(rbr_scene* scene, const rbr_ray* d_rays, size_t n, rbr_hit* d_hits);
/* Launches `iters` back-to-back traversals (kind 0 closest, 1 any) over device-resident rays
 * and returns the mean kernel time measured with HIP events on the launch stream. */
int rbr_time_trace(rbr_scene* scene, int kind, const rbr_ray* d_rays, size_t n, void* d_out,
                   int iters, float* mean_ms);

/* Traversal launch tuning (the autotuner's knobs): mode 0 = persistent waves that fetch
 * 64-ray batches from a device counter (default), 1 = one ray per lane, 2 / 3 / 4 =
 * persistent waves whose idle lanes take new rays once 8 / 16 / 32 of them are idle,
 * 5 / 6 / 7 / 8 = persistent waves that advance every lane by one node and one primitive per
 * iteration and refill once 4 / 8 / 16 / 32 lanes are idle, 9 / 10 / 11 / 12 = the same with
 * the primitive step deferred until 32 lanes hold a leaf or fewer than 48 lanes have a node
 * to test; mode | (m + 1) << 8 runs any-hit launches (shadow rays) in mode m and closest-hit
 * launches in mode & 0xFF; lds_depth = LDS
 * stack entries per lane (8, 10, 12, 16, 24, 32; deeper entries spill to a per-wave global
 * buffer); waves_per_cu = persistent waves launched per compute unit (1..32; a CU holds as
 * many at once as the kernel's LDS (lds_depth * 512 B of stack, plus 3 KiB of result staging
 * in modes 5-12) and registers allow, rbr_trace_resident_waves; a larger grid runs its extra
 * waves as others end).  Mode 1 needs the ray count on the host: launches whose count stays on the
 * device (the rays of rbr_render) run mode 0 instead, for closest and for any hit. */
int rbr_set_trace_tuning(rbr_ctx* ctx, int mode, int lds_depth, int waves_per_cu);

/* Waves per compute unit that the traversal launch of the current tuning (closest hit, or any
 * hit when `any`) keeps resident at once on this scene's BVH layout: the persistent grid of
 * waves_per_cu capped by what the kernel's registers and LDS allow (hipOccupancy...), or the
 * per-CU limit alone for mode 1.  A tuning query with no counterpart in the reference. */
int rbr_trace_resident_waves(rbr_scene* scene, int any, int* waves_per_cu);

/* Watchdog of the traversal kernels (no counterpart in the reference, whose loops cannot
 * fail but for ExponentialMedium::sampleDistance, medium.cpp:68-84).  A ray's node visits
 * cannot exceed its BVH's node count (each node has one parent, so it is entered at most
 * once, and a leaf holds at most 8 primitives), so every kernel counts them per ray and gives
 * up on a ray past `cap` (checked every 16-32 loop iterations): the kernel still ends
 * normally, the ray keeps the closest hit found so far, and a device flag is raised.  The
 * synchronous entry points (rbr_trace_closest / _any, rbr_render, rbr_time_trace,
 * rbr_synchronize) read and clear that flag and then return RBR_ERR_WATCHDOG; device-pointer
 * launches report at the next rbr_synchronize.  cap = 0 (the default) uses that bound, the
 * node count + 1 of the BVH (or node arena) being traversed; a smaller cap forces the error
 * path (tests). */
int rbr_set_trace_step_cap(rbr_ctx* ctx, uint32_t cap);

/* Per-ray host traversal of the same flat wide BVH with the same acceptance rule — the path
 * the Accelerator adapter's per-ray intersect()/intersectP() take, since those are called
 * one ray at a time from the CPU integrators (integrator.cpp:12-16) and must not launch a
 * kernel.  Optional counters return the wide-node visits and primitive tests summed over
 * the batch (the traversal-inclusive byte count of the bandwidth roofline uses them). */
int rbr_host_trace_closest(const rbr_scene* scene, const rbr_ray* rays, size_t n, rbr_hit* hits,
                           uint64_t* node_visits, uint64_t* prim_tests);
int rbr_host_trace_any(const rbr_scene* scene, const rbr_ray* rays, size_t n, uint8_t* occluded,
                       uint64_t* node_visits, uint64_t* prim_tests);

/* ---------------------------------------------------------------- shading + rendering
 * Scene shading tables used by the wavefront integrators: one material per primitive id
 * (Primitive::material, primitive.h:21; -1 = no material, which ends the path as
 * PathTracingIntegrator::Li does, path_tracing_integrator.cpp:19-20), the lights
 * (Scene::lights, scene.h:92), the medium (Scene::medium) and the background colour. */
int rbr_scene_set_shading(rbr_scene* scene,
                          const rbr_material* materials, uint32_t nmat,
                          const int32_t* prim_material, uint32_t nprim_ids,
                          const rbr_light* lights, uint32_t nlights,
                          const rbr_medium* medium,
                          const float background[3]);

/* Integrator::render (integrator.cpp:7-50) for PathTracingIntegrator::Li
 * (path_tracing_integrator.cpp:5-68) or VolumetricPathIntegrator::Li
 * (volumetric_path_integrator.cpp:6-89), for one or many jobs sharing one device ray pool. */
int rbr_render(rbr_ctx* ctx, const rbr_job* jobs, int njobs);
/* rbr_render with progress reports, for Integrator::render's progress callback
 * (integrator.cpp:39-45; the reference calls it from its worker threads): progress(fraction,
 * user) is called after each round of the launch (a round renders as many whole samples of
 * every job as the ray pool holds) with the fraction of all path samples done, from a HIP
 * runtime thread, in increasing order, the last time with 1; every call has returned when
 * rbr_render_progress returns.  The function must not call into this library. */
typedef void (*rbr_progress_fn)(float fraction, void* user);
int rbr_render_progress(rbr_ctx* ctx, const rbr_job* jobs, int njobs, rbr_progress_fn progress, void* user);

/* Per-pixel sums for variance estimates: adds each sample's radiance and its square to
 * caller arrays (3 doubles per pixel each); used by the convergence tooling. */
int rbr_render_moments(rbr_ctx* ctx, const rbr_job* job, double* sum_rgb, double* sumsq_rgb);

/* Render launch tuning (the autotuner's knobs; no counterpart in the reference):
 * pool_paths = paths in flight, i.e. the queue capacity (2^16..2^26, default 2^24; a job
 * never gets fewer than one sample's worth); use_graphs = 1 records each round's launch
 * sequence as a hipGraph and replays it while nothing it depends on changes (default 1);
 * sort_materials = 1 shades each bounce's queue in material order, misses grouped apart
 * (default 1).  Images do
 * not depend on use_graphs or sort_materials; a pool that holds fewer of a job's samples
 * only changes the float summation order of its per-pixel means. */
int rbr_set_render_tuning(rbr_ctx* ctx, uint32_t pool_paths, int use_graphs, int sort_materials);

/* Time of the last rbr_render call: total ms and per-stage ms (extend, shade, shadow). */
int rbr_last_render_timing(rbr_ctx* ctx, double* total_ms, double* stage_ms /* 4 */);

/* Traversal census of path-tracing renders (profiling helper, no counterpart in the
 * reference): with rays_per_queue > 0, every later rbr_render / rbr_render_moments of PATH or
 * VOLPATH jobs also copies an evenly strided sample of up to rays_per_queue rays of each
 * bounce's extension queue and shadow queue (with their BVH roots) and, once the render is
 * done, traces the samples with the host traversal of the same wide BVH to count node visits
 * and primitive tests; hipGraph replay is off while it is on.  Images are unchanged.  0 turns
 * it off (the default).  rbr_last_render_stats returns the counts of the last such render. */
typedef struct rbr_render_stats {
    uint64_t extension_rays;      /* closest-hit rays traced: camera rays and continuations */
    uint64_t shadow_rays;         /* any-hit rays traced: NEE shadow rays                  */
    uint64_t shadow_empty;        /* shadow-queue slots left empty by the chunked appends  */
    uint64_t sampled_extension;   /* rays traced by the host census                        */
    uint64_t sampled_shadow;
    double ext_nodes_per_ray;     /* wide-node visits per extension ray (count-weighted)   */
    double ext_prims_per_ray;     /* primitive tests per extension ray                     */
    double shadow_nodes_per_ray;  /* per non-empty shadow ray                              */
    double shadow_prims_per_ray;
} rbr_render_stats;
int rbr_set_render_census(rbr_ctx* ctx, uint32_t rays_per_queue);
int rbr_last_render_stats(rbr_ctx* ctx, rbr_render_stats* out);

/* ---------------------------------------------------------------- path guiding
 * Extension without a counterpart in the reference, whose integrators sample the BSDF only
 * (path_tracing_integrator.cpp:40-60).  A per-context MLP maps a shading point (position
 * scaled by the box lo..hi, outward normal, wo) to a mixture of 8 von Mises-Fisher lobes over
 * incident directions.  Layout of the network and its parameters: rbr_guide_param_count()
 * floats, W0[64][224] W1[64][64] W2[64][64] W3[64][64] W4[48][64] b0[64] b1[64] b2[64] b3[64]
 * b4[48] (W[out][in]; input features 216..223 and outputs 41..47 are padding; output 40 is the
 * logit of the probability s with which a guided vertex samples the guide rather than the
 * BSDF, zero at initialisation).  Inference and training run on the matrix cores in BF16 with
 * FP32 accumulation. */
int rbr_guide_param_count(void);
int rbr_guide_init(rbr_ctx* ctx, uint64_t seed, const float lo[3], const float hi[3]);
int rbr_guide_get_params(rbr_ctx* ctx, float* params, size_t n);
int rbr_guide_set_params(rbr_ctx* ctx, const float* params, size_t n);
/* The trained guide as a file: FP32 parameters, scene box, and the Adam moments and step when
 * training has started, so that rbr_guide_load (on any context of this library) continues
 * training exactly where the saved one stood.  RBR_ERR_IO for a file that cannot be written or
 * read or is not a guide file of this network. */
int rbr_guide_save(rbr_ctx* ctx, const char* path);
int rbr_guide_load(rbr_ctx* ctx, const char* path);
/* queries: 9 floats each (position, normal, wo); lobes: 40 floats each, 8 x (mu 3, kappa, weight),
 * at the FP16 precision the renderer samples them with; the weights are the mixture weights
 * times s and sum to s. */
int rbr_guide_eval(rbr_ctx* ctx, const float* queries, size_t n, float* lobes);
/* Draws per_query directions from each query's mixture with the renderer's sampler (u: 3
 * uniforms per direction; lobe choice, polar angle, azimuth) and returns them with their
 * density: dirs 3 floats and pdfs 1 float per direction, query-major. */
int rbr_guide_sample(rbr_ctx* ctx, const float* queries, size_t nq, size_t per_query, const float* u, float* dirs,
                     float* pdfs);
/* samples: 14 floats each (position 3, normal 3, wo 3, direction 3, weight c >= 0, BSDF density
 * pb of the direction, <= 0 for none; the direction counts as drawn from the BSDF).
 * loss = (1/n) sum -c log q(direction) with q = s p + (1 - s) pb where pb > 0 and q = p
 * elsewhere (p the guide's mixture density).  grad (rbr_guide_param_count floats): the lobe
 * outputs learn -c log p, the selection row (output 40) the second moment c^2 pb / q of the
 * one-sample MIS estimator, which the hidden layers do not see. */
int rbr_guide_loss_grad(rbr_ctx* ctx, const float* samples, size_t n, double* loss, float* grad);
/* One Adam step (beta 0.9 / 0.999) on the loss above; loss (optional) is the pre-step value. */
int rbr_guide_train_step(rbr_ctx* ctx, const float* samples, size_t n, float lr, double* loss);

/* Mean device time of one launch (HIP events on the context stream) over resident synthetic
 * inputs: kind 0 = inference of n queries, kind 1 = one training step (forward, backward,
 * weight gradients, Adam with a zero learning rate, so the weights do not change) on n
 * samples.  For profiling the matrix-core kernels. */
int rbr_guide_time(rbr_ctx* ctx, int kind, size_t n, int iters, float* mean_ms);

/* Guided rendering: rbr_render jobs with config.guided = 1 draw each non-delta continuation
 * direction from 0.5 x guide + 0.5 x BSDF (one-sample MIS; NEE and its weights unchanged, so
 * the image converges to the plain integrator's); guided = 3 also records one training
 * sample per guided vertex; guided = 2 records the same samples from the plain integrator's
 * BSDF-sampled paths (no guide inference, the image is the plain one bit for bit).  Records
 * stay on the device until cleared. */
int rbr_guide_reserve(rbr_ctx* ctx, uint32_t capacity);      /* record capacity, default 2^22 */
int rbr_guide_collected(rbr_ctx* ctx, uint64_t* count, double* weight_sum);
/* steps Adam steps, each on `batch` records drawn without repetition (random permutation). */
int rbr_guide_train_collected(rbr_ctx* ctx, int steps, uint32_t batch, float lr, double* loss);
int rbr_guide_clear_collected(rbr_ctx* ctx);

/* ---- Scene ingestion (host only) ---------------------------------------------------------- */
/* Triangles of Wavefront OBJ text as ObjImporter::parse builds them (obj_importer.cpp:59-150):
 * v / vt / vn / f records, 1-based and negative indices, polygons fanned around their first
 * corner; 9 floats per triangle (positions).  tris == NULL returns the count only; *normals /
 * *uvs report the vn / vt records.  RBR_ERR_INVALID when a face names a missing vertex or no
 * triangle results (the reference's failures). */
int rbr_obj_parse(const char* text, size_t length, float* tris, size_t capacity, size_t* count, size_t* normals,
                  size_t* uvs);
/* The vertex normals of the same triangles, 9 floats per triangle (its corners' "vn" vectors as
 * written), when the reference's importer keeps them: some corner names a normal and every
 * corner's normal is nonzero (obj_importer.cpp:135-143); *count = 0 when it does not.  For
 * rbr_scene_set_tri_normals.  normals == NULL returns the count only. */
int rbr_obj_parse_normals(const char* text, size_t length, float* normals, size_t capacity, size_t* count);

/* ---- Film output (host only; films are float RGB, row-major from the top row) ----------- */
/* 8-bit conversion of a mean-radiance film, as Film::toRGB8 / Film::toRGB8ToneMapped
 * (film.cpp:27-63): per channel v (times exposure, then v / (v + 1) when tone_map),
 * pow(v, 1/gamma) clamped to [0, 1] (a NaN becomes 1, as the reference's clamp does),
 * * 255 + 0.5, truncated.  out holds width * height * 3 bytes. */
int rbr_film_to_rgb8(const float* rgb, int width, int height, float gamma, int tone_map, float exposure,
                     unsigned char* out);
/* PNG of an 8-bit RGB image, the default format of ImageExporter::writeRGB8
 * (image_exporter.cpp:13-39); stored deflate blocks, any PNG reader decodes it. */
int rbr_film_write_png(const char* path, const unsigned char* rgb8, int width, int height);
/* HDR film as PFM (little-endian FP32, rows bottom-up); no reference counterpart, the
 * reference writes 8-bit images only.  read: rgb == NULL returns the size only. */
int rbr_film_write_pfm(const char* path, const float* rgb, int width, int height);
int rbr_film_read_pfm(const char* path, int* width, int* height, float* rgb, size_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* RBR_H */
