/*
 * rt_oracle.h -- CPU oracle for the ray-space accelerators (naive, rayhierarchy, ray-bvh).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is compared
 * against; it is never linked into, loaded by or called from the product library.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Every function restates the reference's algorithm from its source (file:line cited at
 * each definition in rt_oracle.cpp) with a fixed FP32 operation order:
 *   - no FMA contraction (built with -ffp-contract=off),
 *   - IEEE correctly rounded division and square root,
 *   - dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z, cross() in FP32 (the OpenCL kernels),
 *   - normalize(v) = v * (1/sqrt(dot(v,v))), zero vectors returned unchanged,
 *   - min/max as selects (b<a?b:a), the OpenCL double literals of the kernels taken in FP32.
 * The nearest-hit reduction is min t, ties -> smallest triangle id (the race-free reading of
 * the reference's non-atomic min in intersectAllLeaves.cl:58-61 / naive.cl:49-52).
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Ray batch as structure of arrays (host memory).  W*H rays in row-major ray-image order. */
typedef struct orc_rays {
  const float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
  const uint8_t *valid; /* NULL = all valid */
  int32_t W, H;
} orc_rays;

/* Accelerator parameters (reference defaults: chunk 4x4, height 3, box_dir, repairRun 1). */
typedef struct orc_params {
  int32_t kind;       /* 0 naive, 1 rayhierarchy, 2 ray-bvh */
  int32_t node;       /* 0 box_dir (6DB), 1 ia, 2 box_uv (5DB) */
  int32_t cx, cy;     /* leaf tile size in rays */
  int32_t height;     /* hierarchy height */
  int32_t repair_run; /* rayhierarchy primary repair passes */
} orc_params;

/* modes */
#define ORC_CLOSEST_PRIMARY 0
#define ORC_CLOSEST_SECONDARY 1
#define ORC_OCCLUDED 2

/* BVH handoff node, layout of the reference's GPUNode (bvh.h:75-79). */
typedef struct orc_gpu_node {
  float ax, ay, az, bx, by, bz;
  uint32_t primOffset; /* interior: index of child A (child B = A+1); leaf: first prim position */
  uint32_t nPrimitives;
} orc_gpu_node;

/* Traversal statistics (O-ref counts, used for the algorithmic-byte model). */
typedef struct orc_stats {
  uint64_t cull_tests;  /* (RH node, triangle-or-BVH-node) cull evaluations */
  uint64_t leaf_pairs;  /* (leaf, triangle) pairs reaching the leaf test */
  uint64_t ray_tests;   /* ray-triangle tests executed at leaves */
  uint64_t repair_rays; /* rays whose result the repair pass changed */
} orc_stats;

/* ---- single predicates (closed-form tests) ---- */
/* returns 1 when the barycentric/determinant test passes; out = {t, b1, b2} */
int orc_mt_strict(const float o[3], const float d[3], const float v9[9], float out[3]);
int orc_mt_naive(const float o[3], const float d[3], const float v9[9], float out[3]);
int orc_mt_naive_p(const float o[3], const float d[3], const float v9[9], float out[3]);
int orc_mt_repair(const float o[3], const float d[3], const float v9[9], float out[3]);
int orc_mt_pbrt(const float o[3], const float d[3], const float v9[9], float out[3]);
/* node16 = {dmin xyz, valid, dmax xyz, 0, omin xyz, 0, omax xyz, 0} */
int orc_cull_tri(const float node16[16], const float v9[9]);
int orc_cull_node6(const float node16[16], const float bmin[3], const float bmax[3]);
int orc_cull_ia(const float node16[16], const float bmin[3], const float bmax[3]);
/* 5DB: node16 = {min slope, min dz, 0, valid, max slope, max dz, 0, 0, omin, 0, omax, 0} */
int orc_cull_5db(const float node16[16], const float bmin[3], const float bmax[3]);

/* ---- ray hierarchy ---- */
/* number of levels and per-level grid dims (dims[2*l], dims[2*l+1]; l=0 is the leaf level) */
int orc_rh_levels(int32_t W, int32_t H, int32_t cx, int32_t cy, int32_t height, int32_t *dims);
/* 1 when the reference's image indexing is well formed for this shape */
int orc_rh_reference_well_formed(int32_t W, int32_t H, int32_t cx, int32_t cy, int32_t height);
/* builds every level (leaf first) into nodes (16 floats each); returns the node count.
 * node 2 (box_uv) stores the slope and dz intervals in place of the direction box. */
int64_t orc_rh_build(const orc_rays *r, int32_t cx, int32_t cy, int32_t height, int32_t use_valid,
                     int32_t node, float *nodes);

/* ---- traces ---- */
/* closest modes write id/t/b1/b2 (miss: id -1, t INF); occluded writes occ (0/1).
 * flags (optional, one byte per ray): bit 0 near tie -- the two nearest candidates lie
 * within 1e-6*t; bit 1 ulp-marginal -- a candidate whose leaf test or chain of culls rests on
 * an operand within 4 ulp of its threshold could change the ray's result (rayhierarchy and
 * ray-bvh; naive writes 0). */
int orc_trace(const float *v9, uint32_t ntri, const orc_rays *r, const orc_params *p, int32_t mode,
              int32_t *id, float *t, float *b1, float *b2, uint8_t *occ, orc_stats *st,
              uint8_t *flags);
/* ray-bvh: triangles in original order, BVH nodes of all parts concatenated, roots = frontier */
int orc_trace_raybvh(const float *v9, uint32_t ntri, const orc_gpu_node *nodes, uint32_t nnodes,
                     const uint32_t *roots, uint32_t nroots, const uint32_t *prim_order,
                     const orc_rays *r, const orc_params *p, int32_t mode, int32_t *id, float *t,
                     float *b1, float *b2, uint8_t *occ, orc_stats *st, uint8_t *flags);

/* ---- the library's per-ray path, restated ---- */
typedef struct orc_search_stats {
  uint64_t rays;   /* rays traced */
  uint64_t visits; /* interior search-BVH nodes loaded */
  uint64_t tests;  /* strict leaf tests */
  uint64_t culls;  /* chain cull evaluations (hierarchy node x triangle or BVH node) */
  uint64_t nodes;  /* interior nodes of the search BVH */
  uint64_t chains; /* cull chains evaluated (strict-test hits that could change the ray) */
} orc_search_stats;
/* Per-ray traversal of the search BVH (binned SAH, one triangle per leaf, boxes grown by
 * 2^-11 max(1, max |coordinate|)) with the strict leaf test and the reference's exact cull
 * chain for passing triangles.  kind 1 (rayhierarchy) or 2 (ray-bvh, with its handoff).
 * ray_mask (optional) selects the rays to trace; the others keep the miss / 0 outputs.
 * Results must equal orc_trace / orc_trace_raybvh without repair; counts feed the work model
 * of the per-ray kernel. */
int orc_search_trace(const float *v9, uint32_t ntri, const orc_gpu_node *nodes, uint32_t nnodes,
                     const uint32_t *roots, uint32_t nroots, const uint32_t *prim_order,
                     const orc_rays *r, const orc_params *p, int32_t mode, const uint8_t *ray_mask,
                     int32_t nthreads, int32_t *id, float *t, float *b1, float *b2, uint8_t *occ,
                     orc_search_stats *st);

/* shading record of computeDpTuTv.cl:8-91 for hits (id>=0): uv (2/ray), dpdu, dpdv (3/ray);
 * naive_rule: the naive kernels' fallback for |det| < EPS (naive.cl:60-61) */
void orc_shading(const float *v9, const float *uv6, uint32_t ntri, const orc_rays *r,
                 const int32_t *id, int32_t naive_rule, float *u, float *v, float *dpdu,
                 float *dpdv);

/* ---- reference BVH build (bvh.cpp:209-455) ---- */
/* Builds the SAH BVH (12 buckets) over triangle bounds and returns the BVH primitive order
 * (position -> input index) in order_out.  gpu=1 also produces the BFS GPUNode layout
 * (GPUbufferWide, bvh.cpp:154-206): nodes_out (capacity 2*n), returns node count, top-level
 * count in *top_out.  gpu=0 produces the depth-first LinearBVHNode layout in an internal
 * handle used by orc_cpu_bvh_* (returned via *handle_out). */
int64_t orc_bvh_build(const float *v9, uint32_t ntri, uint32_t max_prims, int32_t split,
                      int32_t gpu, uint32_t bvh_height, uint32_t *order_out,
                      orc_gpu_node *nodes_out, uint32_t *top_out);
/* orderedPrims of BVHAccel (split 0 sah, 1 middle, 2 equal; max_prims up to 255), with the
 * repeats that the SAH leaf branch without a return adds (bvh.cpp:417-428).  Returns the
 * length, or -1 when it exceeds cap. */
int64_t orc_bvh_order(const float *v9, uint32_t ntri, uint32_t max_prims, int32_t split,
                      uint32_t *order_out, uint64_t cap);
/* ray-bvh handoff exactly as RayBVH::Preprocess (rayBVH.cpp:144-146,204-276): global BVH
 * order, 60k-triangle parts, per-part BVH + GPUbufferWide.  Returns node count. */
int64_t orc_raybvh_handoff(const float *v9, uint32_t ntri, uint32_t max_prims, int32_t split,
                           uint32_t bvh_height,
                           orc_gpu_node *nodes_out, uint32_t *roots_out, uint32_t *nroots_out,
                           uint32_t *prim_order_out);

/* ---- pbrt CPU accelerator (the timed CPU baseline): BVHAccel maxnodeprims + Triangle ---- */
void *orc_cpu_bvh_create(const float *v9, uint32_t ntri, uint32_t max_prims);
void orc_cpu_bvh_destroy(void *h);
/* nthreads <= 0 -> hardware_concurrency */
void orc_cpu_bvh_closest(void *h, const orc_rays *r, int32_t nthreads, int32_t *id, float *t,
                         float *b1, float *b2);
void orc_cpu_bvh_occluded(void *h, const orc_rays *r, int32_t nthreads, uint8_t *occ);
/* exhaustive pbrt-predicate closest hit (O-exact) */
void orc_exact_closest(const float *v9, uint32_t ntri, const orc_rays *r, int32_t *id, float *t);

#ifdef __cplusplus
}
#endif
#endif
