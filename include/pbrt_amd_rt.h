/*
 * pbrt_amd_rt.h -- C ABI of the MI355X (gfx950) ray-space accelerators.
 *
 * This is the drop-in boundary below pbrt's batched Aggregate interface.  It replaces the
 * device side of the reference's three OpenCL accelerators:
 *   Accelerator "naive"        src/accelerators/naive.cpp:204-437        (NaiveAccel)
 *   Accelerator "rayhierarchy" src/accelerators/rayhierarchy.cpp:602-1326 (RayHieararchy)
 *   Accelerator "ray-bvh"      src/accelerators/rayBVH.cpp:615-1323      (RayBVH)
 * and their OpenCL plumbing (src/core/GPUparallel.cpp:479-712).  Plain pointers and sizes
 * only; every entry point returns an rt_status (0 = success).  No C++ exception crosses it.
 *
 * Memory: scenes (triangles, UVs, the ray-bvh BVH) are copied once into device memory and
 * stay resident until rt_scene_destroy.  Ray and hit arrays passed to the trace calls are
 * DEVICE pointers owned by the caller; scratch (the ray hierarchy, reduction keys) is owned
 * by the scene and grows on demand.  Calls on one scene must not overlap in time.
 */
#ifndef PBRT_AMD_RT_H
#define PBRT_AMD_RT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rt_status {
  RT_OK = 0,
  RT_ERR_INVALID = 1,     /* bad argument (null pointer, size, shape, parameter) */
  RT_ERR_HIP = 2,         /* a HIP runtime call failed */
  RT_ERR_OOM = 3,         /* device allocation failed */
  RT_ERR_UNSUPPORTED = 4, /* parameter combination this build does not implement */
  RT_ERR_NO_DEVICE = 5,   /* no usable gfx950 device */
  RT_ERR_STATE = 6,       /* call order violated (e.g. ray-bvh trace before rt_scene_set_bvh) */
  RT_ERR_TOO_LARGE = 7    /* input exceeds the static scratch of rt_trace_closest_noalloc */
} rt_status;

/* Accelerator names of MakeAccelerator (src/core/api.cpp:576-596). */
enum { RT_ACCEL_NAIVE = 0, RT_ACCEL_RAYHIERARCHY = 1, RT_ACCEL_RAYBVH = 2 };
/* Ray-hierarchy node types (rayhierarchy.cpp:68-123, rayBVH.cpp:62-100): "box_dir" (6-D boxes,
 * default), "ia" (interval arithmetic), "box_uv" (origin box + intervals of d.y/d.x and d.z,
 * culled by intersectsNode5DB.cl as C parses it). */
enum { RT_NODE_BOX_DIR = 0, RT_NODE_IA = 1, RT_NODE_BOX_UV = 2 };
/* Closest-hit modes: primary camera rays (Intersect without bounce, rayhierarchy.cpp:816) and
 * secondary rays with validity (Intersect(..., bounce), rayhierarchy.cpp:602). */
enum { RT_CLOSEST_PRIMARY = 0, RT_CLOSEST_SECONDARY = 1 };

/* Accelerator parameters of CreateNaiveAccelerator (naive.cpp:440-445), CreateRayHieararchy
 * (rayhierarchy.cpp:1329-1348) and CreateRayBVH (rayBVH.cpp:1326-1344). */
typedef struct rt_accel_desc {
  int32_t kind;         /* RT_ACCEL_* */
  int32_t node;         /* RT_NODE_*            ("node", default box_dir) */
  int32_t chunk_x;      /* leaf tile width       ("chunkXSize", default 4) */
  int32_t chunk_y;      /* leaf tile height      ("chunkYSize", default 4) */
  int32_t height;       /* hierarchy height      ("height", default 3) */
  int32_t repair_run;   /* repair passes         ("repairRun", default 1; rayhierarchy only) */
  int32_t bvh_height;   /* BVH levels kept below the frontier ("BVHheight", default 3) */
  int32_t max_bvh_prim; /* BVH leaf size         ("maxBVHPrim", default 1) */
} rt_accel_desc;

/* BVH node of the ray-bvh handoff, the reference's GPUNode (src/accelerators/bvh.h:75-79):
 * interior nodes hold the index of child A in prim_offset (child B = A+1), leaves hold the
 * first primitive position (into prim_order) and n_primitives > 0. */
typedef struct rt_gpu_node {
  float ax, ay, az, bx, by, bz;
  uint32_t prim_offset;
  uint32_t n_primitives;
} rt_gpu_node;

/* A ray batch in DEVICE memory as structure of arrays, laid out as a row-major ray image of
 * width x height rays (the sampler order the reference's hierarchy assumes,
 * rayhierarchy.cpp:258-265,377).  Use height = 1 for unstructured sets.
 * valid: one byte per ray, nonzero = valid; NULL = all valid.  Ignored for
 * RT_CLOSEST_PRIMARY, where every ray is traced (rayhierarchy.cpp:833-853). */
typedef struct rt_rays {
  const float *ox, *oy, *oz;
  const float *dx, *dy, *dz;
  const float *tmin, *tmax;
  const uint8_t *valid;
  int32_t width, height;
} rt_rays;

/* Closest-hit outputs in DEVICE memory, one entry per ray.  id is the ORIGINAL triangle index
 * (ray-bvh ids are mapped back from BVH order) or -1 for a miss; t = INFINITY and u = v = 0 on
 * a miss.  (u, v) are the interpolated surface coordinates of computeDpTuTv.cl:87-89.
 * dpdu/dpdv are optional (NULL skips them): 3 floats per ray each, x y z of ray i at
 * [3i, 3i+3). */
typedef struct rt_hits {
  int32_t *id;
  float *t, *u, *v;
  float *dpdu, *dpdv;
} rt_hits;

typedef struct rt_ctx rt_ctx;
typedef struct rt_scene rt_scene;

/* Fills desc with the reference defaults for the accelerator kind. */
int rt_accel_desc_default(int32_t kind, rt_accel_desc *desc);

/* Opens the HIP device (replaces OpenCL::OpenCL, GPUparallel.cpp:39-124). */
int rt_ctx_create(int32_t device, rt_ctx **out);
int rt_ctx_destroy(rt_ctx *ctx);
/* Conditions the traversal kernels record instead of failing (no counterpart in the
 * reference, whose kernels have none of these bounded queues).  rt_ctx_check waits for the
 * context's device, returns the RT_FLAG_* bits recorded since the last check in *flags and
 * clears them.  0 means every trace since then completed its traversal; a set bit means
 * some ray-bvh results may be incomplete (the host-side subtree limit of rt_scene_set_bvh
 * keeps both within their bounds).
 *   RT_FLAG_BVH_STACK: a block's lockstep traversal stack would have overflowed;
 *   RT_FLAG_PAIR_CAP: a block produced more (leaf, BVH leaf) pairs in one round than its
 *   pair queue holds. */
#define RT_FLAG_BVH_STACK 1u
#define RT_FLAG_PAIR_CAP 2u
int rt_ctx_check(rt_ctx *ctx, uint32_t *flags);

/* Uploads the flattened triangles once (replaces the per-call vertex uploads of
 * naive.cpp:266-278, rayhierarchy.cpp:933-948, rayBVH.cpp:705-706).
 * v9: HOST array of 9 floats per triangle (p1, p2, p3); uv6: HOST array of 6 floats per
 * triangle or NULL for the default (0,0),(1,0),(1,1) (naive.cpp:91-97). */
int rt_scene_create(rt_ctx *ctx, const float *v9, const float *uv6, uint32_t ntri,
                    const rt_accel_desc *desc, rt_scene **out);
/* ray-bvh only: the CPU-built BVH (BVHAccel::GPUbufferWide, bvh.cpp:154-206, one or more
 * 60k-triangle parts concatenated).  roots are the frontier nodes (topLevelNodes of each part),
 * prim_order maps BVH position -> original triangle index.  All HOST arrays. */
int rt_scene_set_bvh(rt_scene *scene, const rt_gpu_node *nodes, uint32_t nnodes,
                     const uint32_t *roots, uint32_t nroots, const uint32_t *prim_order);
int rt_scene_destroy(rt_scene *scene);

/* BVHAccel split methods ("splitmethod", bvh.cpp:214-221): "sah" (default), "middle",
 * "equal". */
enum { RT_SPLIT_SAH = 0, RT_SPLIT_MIDDLE = 1, RT_SPLIT_EQUAL = 2 };

/* Host-side CPU BVH of ray-bvh, as RayBVH builds it before its first trace
 * (rayBVH.cpp:144-146,204-276; BVHAccel bvh.cpp:209-434; GPUbufferWide bvh.cpp:154-206):
 * global BVH order, parts of at most 60,000 triangles, per-part BVH (split_method RT_SPLIT_*,
 * max_prims <= 4 per leaf) flattened breadth first keeping the lowest bvh_height levels.
 * Returns malloc'ed HOST arrays for rt_scene_set_bvh; release them with
 * rt_raybvh_free_handoff. */
int rt_raybvh_build_handoff(const float *v9, uint32_t ntri, uint32_t max_prims,
                            int32_t split_method, uint32_t bvh_height, rt_gpu_node **nodes,
                            uint32_t *nnodes, uint32_t **roots, uint32_t *nroots,
                            uint32_t **prim_order);
int rt_raybvh_free_handoff(rt_gpu_node *nodes, uint32_t *roots, uint32_t *prim_order);
/* Primitive order of pbrt's BVHAccel over the triangles (split_method RT_SPLIT_*, leaves of at
 * most min(255, max_prims) primitives; bvh.cpp:209-434): order[i] = input index of the
 * triangle at BVH position i.  RayHieararchy's "bool sort" parameter swaps its primitives into
 * this order before flattening them (rayhierarchy.cpp:152-155, parameters at :1332,1337-1338).
 * Where the reference's SAH build lists a leaf's primitives twice (bvh.cpp:417-428, more than 4
 * primitives per leaf) the first occurrences are kept: the later copies never win a tie.
 * HOST arrays. */
int rt_bvh_order(const float *v9, uint32_t ntri, uint32_t max_prims, int32_t split_method,
                 uint32_t *order);

/* Batched closest hit: RayHieararchy::Intersect (rayhierarchy.cpp:602-813,816-1140),
 * RayBVH::Intersect (rayBVH.cpp:615-836,839-1138), NaiveAccel::Intersect (naive.cpp:204-338).
 * mode: RT_CLOSEST_PRIMARY or RT_CLOSEST_SECONDARY.  stream: hipStream_t or NULL. */
int rt_trace_closest(rt_scene *scene, const rt_rays *rays, int32_t mode, const rt_hits *hits,
                     void *stream);
/* Batched occlusion: IntersectP (rayhierarchy.cpp:1170-1326, rayBVH.cpp:1170-1323,
 * naive.cpp:357-437).  occluded[i] = 1 when ray i is valid and blocked in (tmin, tmax). */
int rt_trace_occluded(rt_scene *scene, const rt_rays *rays, uint8_t *occluded, void *stream);

/* rt_trace_closest without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments and the same results as rt_trace_closest; its
 * scratch (hierarchy, reduction keys, work queue, job table) lives in a static device array of
 * 64 MiB per device, set up by rt_ctx_create and held until the process ends.  A call needs
 * at most about 22 bytes per ray, its hierarchy included (naive: 8 bytes per ray), so up to
 * about 3 million rays; a larger one returns RT_ERR_TOO_LARGE.  Every call on a device shares that array, which keeps
 * the last call's contents: calls that overlap in time on one device race (serialize them,
 * e.g. on one stream). */
int rt_trace_closest_noalloc(rt_scene *scene, const rt_rays *rays, int32_t mode,
                             const rt_hits *hits, void *stream);

/* Shadow rays toward a point light from closest hits, as the directlighting integrator builds
 * them (VisibilityTester::SetSegment, src/core/light.h:79-84, PointLight::Sample_L,
 * src/lights/point.cpp:42-49): o = p = o_r + d_r*t, d = (L-p)/|L-p|, tmin = 1e-3*t,
 * tmax = |L-p|, valid = (id >= 0).  All DEVICE arrays of n rays. */
int rt_make_shadow_rays(const rt_rays *primary, const int32_t *id, const float *t,
                        const float light[3], float *ox, float *oy, float *oz, float *dx,
                        float *dy, float *dz, float *tmin, float *tmax, uint8_t *valid,
                        void *stream);

/* Bounce rays of the path integrator from closest hits (PathIntegrator::Li's BSDF sampling
 * step, src/integrators/path.cpp:43-157, with a cosine-weighted Lambertian lobe): o = p,
 * a direction about the hit triangle's geometric normal turned toward the incoming ray,
 * tmin = 1e-3*t (rayEpsilon), tmax = INFINITY, valid = (id >= 0 and the input ray valid).
 * Random numbers are hashed from (seed, ray index, depth), so the rays do not depend on the
 * launch.  All DEVICE arrays of the input image's size. */
int rt_make_bounce_rays(rt_scene *scene, const rt_rays *in, const int32_t *id, const float *t,
                        uint32_t seed, int32_t depth, float *ox, float *oy, float *oz, float *dx,
                        float *dy, float *dz, float *tmin, float *tmax, uint8_t *valid,
                        void *stream);

/* Launch-shape parameters (set by the autotuner; 0 restores the built-in choice).  None of
 * them changes any result.
 *   RT_PARAM_CLUSTERS_CLOSEST / RT_PARAM_CLUSTERS_OCCLUDED: passing triangle clusters (64
 *   triangles each) per rayhierarchy work unit for closest-hit / occlusion traces; smaller
 *   values split a block's work over more units of the persistent traversal kernel.
 *   RT_PARAM_GRID_CLOSEST / RT_PARAM_GRID_OCCLUDED: work-groups of the persistent traversal
 *   kernel (batches only); the built-in choice (0) is one work-group per resident slot of the
 *   kernel on the device (CUs x work-groups per CU).
 *   RT_PARAM_GRAPH: 1 = replay the batch's launch sequence as a HIP graph while its jobs are
 *   unchanged (batches only; needs a non-null stream).
 *   RT_PARAM_RAY_CLUSTERS_CLOSEST / RT_PARAM_RAY_CLUSTERS_OCCLUDED: a block (top-level
 *   hierarchy node) whose root passes at least this many triangle clusters (ray-bvh: clusters
 *   of frontier roots) of a scene holding at least four times as many, or whose leaves' ray
 *   directions spread by more than 0.25 on average
 *   (largest component range per leaf: an incoherent group such as random shadow segments,
 *   diffuse bounces or shadow rays from scattered points), is traced ray by ray --
 *   each ray walks a conservative BVH of the triangles and the reference's exact cull chain
 *   is evaluated for the triangles that pass its strict leaf test -- in place of enumerating
 *   the block's (leaf, triangle) candidate pairs.  1 sends every block with work that way;
 *   2^29 keeps only the direction test; a value above 2^29 (at most 2^30) sends no block.
 *   Built-in (0): 2^29 for closest hits, 256 (rayhierarchy) or 64 (ray-bvh) for occlusion.
 *   RT_PARAM_RAY_GRID_CLOSEST / RT_PARAM_RAY_GRID_OCCLUDED: work-groups (one wavefront each)
 *   of the per-ray kernel (batches only); 0 = one per resident slot.
 *   RT_PARAM_WALK_CLOSEST / RT_PARAM_WALK_OCCLUDED: v in 1..7: a round of the per-ray kernel
 *   walks the search BVH until at most 1/2^v of its lanes still look for a leaf (7: until
 *   every lane holds one); 0 = the built-in 3.
 *   RT_PARAM_TILE_CLOSEST / RT_PARAM_TILE_OCCLUDED: LDS tile of the block path's unit kernel:
 *   0 = by block size (256 rays and three levels where every block fits, else 1,024 rays and
 *   four levels); 1 = the 1,024-ray tile for every block (more LDS per work-group, fewer
 *   work-groups per CU). */
enum {
  RT_PARAM_CLUSTERS_CLOSEST = 1,
  RT_PARAM_CLUSTERS_OCCLUDED = 2,
  RT_PARAM_GRID_CLOSEST = 3,
  RT_PARAM_GRID_OCCLUDED = 4,
  RT_PARAM_GRAPH = 5,
  RT_PARAM_RAY_CLUSTERS_CLOSEST = 6,
  RT_PARAM_RAY_CLUSTERS_OCCLUDED = 7,
  RT_PARAM_RAY_GRID_CLOSEST = 8,
  RT_PARAM_RAY_GRID_OCCLUDED = 9,
  RT_PARAM_WALK_CLOSEST = 10,
  RT_PARAM_WALK_OCCLUDED = 11,
  RT_PARAM_TILE_CLOSEST = 12,
  RT_PARAM_TILE_OCCLUDED = 13
};
int rt_scene_set_param(rt_scene *scene, int32_t param, int32_t value);

/* HIP-event timing of the traversal kernels.  rt_scene_set_timing(scene, 1) starts recording
 * events around every traversal launch (and clears earlier records); rt_scene_kernel_time
 * sums the recorded launches of one mode (RT_CLOSEST_PRIMARY, RT_CLOSEST_SECONDARY, or 2 for
 * occlusion), waiting for them to finish.  mode + 3 gives only the per-ray kernel of those
 * launches, mode + 6 only the unit (block) kernel.  This is the counterpart of the reference's
 * GPU_TIMES accounting (GPUparallel.h:18-27,528-539). */
int rt_scene_set_timing(rt_scene *scene, int32_t enable);
int rt_scene_kernel_time(rt_scene *scene, int32_t mode, float *total_ms, int32_t *launches);

/* Batched form: many rayhierarchy jobs -- scenes and ray images of any size, closest-hit and
 * occlusion mixed -- run as one launch sequence per (mode, node type) group.  Every group
 * builds all its hierarchies in one launch and walks one shared work queue with one
 * persistent, work-stealing traversal kernel.  Results are identical to running each job
 * alone with rt_trace_closest / rt_trace_occluded.  No counterpart in the reference, which
 * serialises every accelerator call (rayhierarchy.cpp:208,636,855,1212); it serves the
 * breadth-first renderer's per-task calls of many scenes and the request loop of
 * rt_server_*.
 * The jobs' ray and output arrays are DEVICE arrays that must stay valid until the stream
 * has run the batch.  Calls on one batch must not overlap in time. */
typedef struct rt_batch rt_batch;
int rt_batch_create(rt_ctx *ctx, rt_batch **out);
int rt_batch_destroy(rt_batch *batch);
/* Removes every job (keeps the device scratch and tuned parameters). */
int rt_batch_clear(rt_batch *batch);
int rt_batch_add_closest(rt_batch *batch, rt_scene *scene, const rt_rays *rays, int32_t mode,
                         const rt_hits *hits);
int rt_batch_add_occluded(rt_batch *batch, rt_scene *scene, const rt_rays *rays,
                          uint8_t *occluded);
int rt_batch_run(rt_batch *batch, void *stream);
int rt_batch_set_param(rt_batch *batch, int32_t param, int32_t value);
int rt_batch_get_param(rt_batch *batch, int32_t param, int32_t *value);
/* Autotuner: times the batch's current jobs over the launch-shape candidates, `reps` runs
 * each, and keeps the fastest per mode: first the per-ray routing threshold x per-ray grid,
 * then clusters per work unit x persistent grid of the unit kernel.  No candidate changes a
 * result. */
int rt_batch_autotune(rt_batch *batch, void *stream, int32_t reps);
/* Event timing of the batch's traversal kernels, as rt_scene_set_timing / kernel_time. */
int rt_batch_set_timing(rt_batch *batch, int32_t enable);
int rt_batch_kernel_time(rt_batch *batch, int32_t mode, float *total_ms, int32_t *launches);

/* Serving form: a request loop over resident scenes.  Each request renders one pinhole
 * camera image (width*spp x height rays, the spp samples of a pixel adjacent in x; pixel
 * centres for spp = 1, hashed jitter from `seed` otherwise) with one shadow ray per hit toward
 * a point light.  Requests are admitted in arrival order while their device memory fits the
 * budget and fewer than max_active are in flight; every tick of the loop launches ONE batch
 * that holds the primary rays of newly admitted requests and the shadow rays of requests whose
 * primaries ran in the previous tick (continuous batching), so work of requests that arrive
 * at different times overlaps on the device.  The loop prepares the next tick (admission,
 * device blocks from a pool allocated at creation, camera rays, job tables) while the current
 * one runs, and allocates no device memory for a request.  Each request's result is the same
 * alone and in any mix.  No counterpart in the reference (the breadth-first renderer traces
 * one image task at a time, breadth-first.cpp:134-153); it is for render services. */
typedef struct rt_server rt_server;
typedef struct rt_request {
  int32_t scene;           /* index into the server's scene list */
  int32_t width, height, spp;
  float origin[3];         /* camera position; the camera looks along +z */
  float fov_deg;           /* vertical field of view */
  float light[3];          /* point light of the shadow rays */
  uint32_t seed;           /* jitter seed (spp > 1) */
  /* optional HOST outputs, width*spp*height entries each (NULL = not copied back) */
  int32_t *host_id;
  float *host_t;
  uint8_t *host_occluded;
} rt_request;
typedef struct rt_request_stats {
  double submit_s, admit_s, done_s; /* steady-clock seconds */
  uint64_t primary_rays;             /* camera rays; one shadow ray slot per camera ray */
  int32_t status;                   /* rt_status of the request */
  uint64_t shadow_rays;              /* valid shadow rays traced (camera rays that hit) */
} rt_request_stats;
/* scenes: rayhierarchy or ray-bvh scenes (with their BVH) that stay resident for the
 * server's lifetime; any other scene gives RT_ERR_UNSUPPORTED.
 * mem_budget: device bytes the in-flight requests may hold (0 = half the free memory, at most
 * 32 GiB); the server allocates a pool of that size for the requests' ray and result
 * buffers. */
int rt_server_create(rt_ctx *ctx, rt_scene *const *scenes, int32_t nscenes, uint64_t mem_budget,
                     int32_t max_active, rt_server **out);
/* Queues a request (thread-safe, returns at once).  A request its scene cannot trace (the
 * ray image's hierarchy shape) fails here with that status, and so does one whose buffers
 * exceed the pool (RT_ERR_OOM); a queued request cannot fail because of another. */
int rt_server_submit(rt_server *server, const rt_request *req, uint64_t *ticket);
/* Waits for a request (timeout_s < 0: no limit); RT_ERR_STATE on time-out. */
int rt_server_wait(rt_server *server, uint64_t ticket, double timeout_s, rt_request_stats *stats);
/* Ticks run so far and the largest number of requests one tick carried. */
int rt_server_counters(rt_server *server, uint64_t *ticks, int32_t *max_tick_requests);
/* Finishes every queued request, then stops the loop. */
int rt_server_destroy(rt_server *server);

const char *rt_error_string(int status);

#ifdef __cplusplus
}
#endif
#endif
