// rt_internal.h -- host-side declarations shared by the C ABI and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rti {

constexpr int kMaxLevels = 4;   // leaves per block <= 64 (uint64 masks)
constexpr int kMaxSub = 85;     // 1 + 4 + 16 + 64 nodes in a block subtree
constexpr int kThreads = 256;   // work-group size of every traversal kernel
constexpr int kMaxRpt = 16;     // rays per thread (block of <= 4096 rays)

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

// Triangles Morton-sorted into clusters of 64 (traversal copy of a scene): v9s[9*slot],
// perm[slot] = original triangle id, cbox[2*c] / cbox[2*c+1] = cluster AABB min / max.
struct SortedTris {
  const float* v9s;
  const uint32_t* perm;
  const float4* cbox;
  uint32_t ntri, nclusters;
};
constexpr int kClusterSize = 64;

// Work queue of the rayhierarchy traversal: passing clusters per block (clist, ntop x
// nclusters), units {block, first, count, 0} of at most per_unit passing clusters, the unit
// count (zeroed before the cluster pass) and the grid of the persistent unit kernel.
struct RhWork {
  uint32_t* clist;
  uint4* units;
  uint32_t* unit_count;
  int per_unit;
  int grid;
};

struct HitOut {
  int32_t* id;
  float *t, *u, *v, *dpdu, *dpdv;
};

// Launchers (all asynchronous on `s`).
hipError_t launch_fill_u64(uint64_t* p, uint64_t v, size_t n, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
hipError_t launch_fill_u8(uint8_t* p, uint8_t v, size_t n, hipStream_t s);
hipError_t launch_rh_build(const RHGeom& g, const RaysDev& r, bool use_valid, float4* nodes,
                           hipStream_t s);
// mode 0 closest primary, 1 closest secondary, 2 occluded
// rayhierarchy traversal: the cluster pass fills the work queue, the persistent unit kernel
// walks it (keys / repkeys / occlusion flags must be initialised before).
hipError_t launch_rh_clusters(const RHGeom& g, int mode, int node_kind, const SortedTris& st,
                              const RhWork& w, const float4* nodes, hipStream_t s);
hipError_t launch_rh_units(const RHGeom& g, const RaysDev& r, int mode, int node_kind,
                           const SortedTris& st, const RhWork& w, const float4* nodes,
                           uint64_t* keys, uint32_t* repkeys, uint8_t* occ, hipStream_t s);
hipError_t launch_bvh_trace(const RHGeom& g, const RaysDev& r, int mode, int node_kind,
                            const float* v9, const float4* bvh, const uint32_t* roots,
                            uint32_t nroots, const uint32_t* prim_order, int tile_roots,
                            int tchunks, const float4* nodes, uint64_t* keys, uint8_t* occ,
                            hipStream_t s);
hipError_t launch_naive_trace(const RaysDev& r, int mode, const float* v9, uint32_t ntri,
                              int tchunks, uint64_t* keys, uint8_t* occ, hipStream_t s);
hipError_t launch_finalize(const RaysDev& r, const float* v9, const float* uv6,
                           const uint64_t* keys, const uint32_t* repkeys, uint32_t* replist,
                           HitOut out, hipStream_t s);
hipError_t launch_repair(const RHGeom& g, const RaysDev& r, int node_kind, const float* v9,
                         const float* uv6, uint32_t ntri, int repair_run, const float4* nodes,
                         const uint32_t* replist, HitOut out, hipStream_t s);
hipError_t launch_shadow_rays(const RaysDev& pr, const int32_t* id, const float* t, float lx,
                              float ly, float lz, float* ox, float* oy, float* oz, float* dx,
                              float* dy, float* dz, float* tmin, float* tmax, uint8_t* valid,
                              hipStream_t s);

}  // namespace rti
