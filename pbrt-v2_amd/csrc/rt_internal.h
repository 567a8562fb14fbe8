// rt_internal.h -- host-side declarations shared by the C ABI and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

struct rt_ctx {
  int device;
  uint32_t* d_flags;  // RT_FLAG_* recorded by the traversal kernels (rt_ctx_check)
};
struct rt_scene;

namespace rti {

// Clusters a block's cluster pass tests: Morton clusters of triangles (rayhierarchy) or of
// frontier roots (ray-bvh); 0 for naive scenes.
uint32_t scene_nclusters(const rt_scene* s);
// RT_OK when `s` can trace a W x H ray image through the batched form (a rayhierarchy or
// ray-bvh scene with its BVH, a hierarchy shape the kernels support), else the status the
// batch would return.
int scene_check_image(const rt_scene* s, int W, int H);

constexpr int kMaxLevels = 12;  // hierarchy levels of a ray image
constexpr int kUnitLevels = 4;  // levels of the blocks the unit kernels and the LDS build take
                                // (leaves per block <= 64, uint64 masks); deeper blocks are
                                // built in global memory and take the per-ray path
constexpr int kMaxSub = 85;     // 1 + 4 + 16 + 64 nodes in such a block's subtree
constexpr int kThreads = 256;   // work-group size of every traversal kernel
constexpr int kMaxBlockRays = 1024;      // blocks the unit kernels stage in LDS; larger
                                         // blocks always take the per-ray path
constexpr int64_t kMaxRaysPerBlock = 1 << 24;  // top-level blocks of at most 2^24 rays
constexpr int kClusterSize = 64;  // triangles per Morton cluster
// The unit queue and the per-ray chunk slots are claimed through kClaimQueues pairs of
// counters, 128 bytes apart after the first 64 words of a group's counter block (kCounterBytes
// in all): range q of the queue is taken first by the work-groups with blockIdx % kClaimQueues
// == q (those the dispatcher places on one XCD), so each counter is hit from one XCD until its
// range runs out.
constexpr int kClaimQueues = 8;
constexpr int kClaimBase = 64, kClaimStride = 32;
constexpr int kCounterBytes = 4 * (kClaimBase + kClaimStride * kClaimQueues);

// Ray batch in device memory (structure of arrays, row-major ray image).
struct RaysDev {
  const float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
  const uint8_t* valid;
  int W, H;
};

// Shape of a ray hierarchy: levels, grid dims per level, node offsets, block geometry.
// Level 0 is the leaf level; the top level L-1 has one work-group per node ("block").
struct RHGeom {
  int W, H, cx, cy, L;
  int gw[kMaxLevels], gh[kMaxLevels], off[kMaxLevels];
  int nlx;     // leaves per block side = 1 << (L-1)
  int bw, bh;  // block size in rays
  int rb;      // rays per block
  int g;       // rays per leaf
  int nsub;    // nodes in a block subtree
  int ntop;    // number of blocks
  int nnodes;  // total nodes over all levels
};

// Returns 0 on success, nonzero when the shape is not supported.
int rh_geometry(int W, int H, int cx, int cy, int height, RHGeom* g);

struct HitOut {
  int32_t* id;
  float *t, *u, *v, *dpdu, *dpdv;
};

// One rayhierarchy traversal job: a resident scene, a ray image and its outputs, with the
// job's scratch (hierarchy nodes, passing-cluster lists, reduction keys) in device memory.
// A launch group holds any number of jobs of one mode and node type; the kernels find their
// job from a global block or ray index (block_base / ray_base are prefix sums).
struct RhItem {
  RHGeom g;
  RaysDev r;
  const float* v9;       // original triangle order (finalize, repair)
  const float* uv6;      // or null
  const float* v9s;      // Morton-sorted traversal copy, clusters of 64
  const uint32_t* perm;  // sorted slot -> original triangle id
  const uint32_t* sperm; // search-BVH position -> original triangle id (per-ray path)
  const float4* cbox;    // cluster AABBs (min, max)
  const float4* tbox;    // AABB (min, max) of every sorted triangle
  const float4* tmt;     // ray-bvh: leaf-test operands (v1, e1, e2) of every triangle, original
                         // order (null for rayhierarchy)
  // per-ray path: padded search BVH over the sorted triangles (4 float4 per node), their
  // leaf-test operands in sorted order, and for ray-bvh the reference BVH leaf of every
  // original triangle (bit 31 set: not the first primitive of its leaf; ~0u: in no leaf)
  const float4* snode;
  const float4* tmts;
  const uint32_t* tleaf;
  uint32_t ntri;
  uint32_t nclusters;    // clusters of 64 triangles (rayhierarchy) or 64 frontier roots
  // ray-bvh: BVH nodes (2 float4 each), frontier roots (Morton-ordered; cbox holds the boxes
  // of their clusters), BVH position -> triangle id, and per node its parent and its depth
  // below its frontier root
  uint32_t nroots;
  const float4* bvh;
  const uint32_t* roots;
  const uint32_t* prim_order;
  const uint32_t* bvh_parent;
  const uint8_t* bvh_depth;
  float4* nodes;         // 4 float4 per hierarchy node
  uint32_t* clist;       // ntop x nclusters passing clusters
  uint64_t* keys;        // closest: one (t, id) key per ray
  uint32_t* repkeys;     // primary with repair: nearest repair-region t key per ray, else null
  uint8_t* occ;          // occlusion flags
  HitOut out;            // closest-hit outputs
  uint32_t* flags;       // the context's RT_FLAG_* word
  uint32_t block_base, ray_base;
  int32_t repair_run;
  uint32_t chunk_base;   // first per-ray chunk slot of the job (its blocks' slots follow)
  uint32_t unit_base;    // first staging slot of the job's work units (fixed per block)
  uint32_t unit_stride;  // staging slots per block: the most units a block can make
};

// A launch group: `items` is the device copy of the job table; units, counters and the
// repair list are shared by the group's jobs.
struct RhGroup {
  const RhItem* items;
  int nitems;
  int mode;  // 0 closest primary, 1 closest secondary, 2 occluded
  int node;  // 0 box_dir, 1 ia
  int big;   // some job has more than 256 rays per block or 4 levels
  int bvh;   // ray-bvh jobs (product traversal with the scene's BVH)
  uint32_t nblocks, nrays;
  uint4* units;       // {item, block, first cluster, count}, compacted from ustage
  uint4* ustage;      // the units of every block at fixed slots (job unit_base + block * stride)
  uint4* bunits;      // per block: {units made, first staging slot, queue slot within its
                      // scan tile, the tile's unit total (first block of a tile only)}
  uint2* rq;          // per-ray path: one slot per 64-ray chunk of every block, at the job's
                      // chunk_base + block * chunks per block + chunk: {item, block * chunks
                      // per block + chunk}, or item 0xffffffff for a block of the block path
  uint32_t* ctr;      // [0] units, [2] rays flagged for repair, [4] set when any block goes
                      // per-ray; [64 + 32 q] and [64 + 32 q + 1]: next unit and next per-ray
                      // chunk slot of range q (kClaimQueues ranges)
  uint32_t* replist;  // global ray indices flagged for repair (primary with repair only)
  int per_unit;       // passing clusters per work unit
  int grid;           // work-groups of the persistent unit kernel; 0 = one per resident slot
  uint32_t max_units;  // upper bound of the group's work units
  uint32_t max_chunks;  // the group's per-ray chunk slots
  // blocks whose root passes at least this many clusters are traced ray by ray; 0 sends every
  // block with a traced ray that way without the cluster scan (job cluster lists unused)
  uint32_t ray_clusters;
  int ray_grid;  // work-groups (one wavefront each) of the per-ray kernel; 0 = resident slots
  int walk;      // per-ray walk rounds stop at <= 1/2^walk searching lanes (1..7)
};
constexpr int kDefaultWalk = 3;

// Search BVH of the per-ray path over triangles in its own search order (see
// rt_search_bvh.cpp).  Node = 4 float4: child 0 box (lo.xyz, hi.x), (hi.yz, child 1 lo.xy),
// (lo.z, hi.xyz), (ref0, ref1 as bits, 0, 0).  A ref with bit 31 clear is an interior node
// index; with bit 31 set it is a leaf of (bits 27..30) + 1 triangles starting at search
// position (bits 0..26).
// Boxes are grown outward by the scene's padding.  The depth is at most kSearchDepth.
#ifndef RT_SEARCH_DEPTH
#define RT_SEARCH_DEPTH 26
#endif
constexpr int kSearchDepth = RT_SEARCH_DEPTH;
// Most triangles a leaf holds (the ref's count field allows 16).  One: a leaf test costs
// about as much as a node visit (52 against 64 bytes, 45 against 24 FLOP), and a lane's leaf
// is then a single test, so the lanes of a wavefront leave the leaf phase together.  Measured
// on a batch of 64 cluster scenes against caps of 2, 3 and 4: 8.00-8.03 ms per step against
// 8.18, 8.31-8.43 and 8.44.
constexpr int kSearchLeaf = 1;
// Builds the search BVH of `ntri` triangles (v9 in the caller's order): the node array
// (4 float4 per node) and the search order (position -> triangle index) whose consecutive
// positions the leaves reference.
void build_search_bvh(const float* v9, uint32_t ntri, std::vector<float4>& nodes,
                      std::vector<uint32_t>& order);

// Allocation-free trace: a static device array of kNoAllocScratch bytes per device (defined in
// rt_kernels.hip).  noalloc_setup looks up its address on `device`, which must be current
// (once per device; 0 on success); noalloc_scratch returns it (null before setup).
constexpr size_t kNoAllocScratch = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
int noalloc_setup(int device);
void* noalloc_scratch(int device);
// Writes a job table into device memory through the kernel arguments (no copy, no allocation).
hipError_t launch_write_table(const RhItem& item, RhItem* dst, hipStream_t s);

// Launchers (all asynchronous on `s`).
hipError_t launch_fill_u64(uint64_t* p, uint64_t v, size_t n, hipStream_t s);
hipError_t launch_fill_u8(uint8_t* p, uint8_t v, size_t n, hipStream_t s);
// rayhierarchy / ray-bvh group: init, fused build, cluster (root) pass, persistent unit
// kernel, per-ray kernel and, for closest hits, finalize and (rayhierarchy) repair.  When
// given, t0 is recorded before the unit kernel, tm between the unit and per-ray kernels and
// t1 after the per-ray kernel.
hipError_t launch_rh_group(const RhGroup& G, hipStream_t s, hipEvent_t t0, hipEvent_t tm,
                           hipEvent_t t1);
hipError_t launch_naive_trace(const RaysDev& r, int mode, const float* v9, uint32_t ntri,
                              int tchunks, uint64_t* keys, uint8_t* occ, hipStream_t s);
// Hit records of the naive path from the reduction keys.
hipError_t launch_finalize(const RaysDev& r, const float* v9, const float* uv6,
                           const uint64_t* keys, HitOut out, bool naive_rule, hipStream_t s);
hipError_t launch_camera_rays(int W, int H, int spp, const float origin[3], float tan_half,
                              float aspect, uint32_t seed, float* ox, float* oy, float* oz,
                              float* dx, float* dy, float* dz, float* tmin, float* tmax,
                              hipStream_t s);
hipError_t launch_bounce_rays(const RaysDev& in, const float* v9, const int32_t* id,
                              const float* t, uint32_t seed, int depth, float* ox, float* oy,
                              float* oz, float* dx, float* dy, float* dz, float* tmin,
                              float* tmax, uint8_t* valid, hipStream_t s);
// count (optional): incremented by the number of valid shadow rays written.
hipError_t launch_shadow_rays(const RaysDev& pr, const int32_t* id, const float* t, float lx,
                              float ly, float lz, float* ox, float* oy, float* oz, float* dx,
                              float* dy, float* dz, float* tmin, float* tmax, uint8_t* valid,
                              hipStream_t s, uint32_t* count = nullptr);

}  // namespace rti
