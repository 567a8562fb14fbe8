// rt_kernels.hip -- gfx950 kernels of the ray-space accelerators.
//
// Data layout in HBM (all FP32 unless noted):
//   triangles  v9[9*ntri] (p1, p2, p3), resident per scene; uv6[6*ntri] optional
//   rays       structure of arrays ox..dz, tmin, tmax (+ valid bytes), row-major ray image
//   hierarchy  4 float4 per node {dmin|valid, dmax, omin, omax}, level-major, row-major
//   keys       one uint64 per ray: (order-preserving t bits << 32) | triangle id
//   bvh        2 float4 per node {ax ay az bx}{by bz primOffset nPrimitives}
//
// Work decomposition of a rayhierarchy / ray-bvh trace (a launch group of one or more jobs):
//   init      clears the per-ray keys / occlusion flags and the group counters;
//   build     one 256-thread work-group per top-level hierarchy node ("block") builds the
//             block's leaves and levels from its rays;
//   cluster   one work-group per block bounds every cluster of 64 triangles (or of 64 ray-bvh
//             frontier roots) against the block root and queues work units of a few passing
//             clusters;
//   unit      a persistent grid (one work-group per resident slot) takes units from the queue
//             with an atomic counter, stages the unit's block of rays and subtree in LDS and
//             traverses (phases below); per-ray results are reduced in LDS and merged with
//             global atomics;
//   finalize  turns each ray's (t, id) key into id / t / u / v (+ dpdu, dpdv);
//   repair    the reference's repairRun pass for the rays finalize flagged.
// The nearest hit is reduced as a (t, id) key with atomic min, so the result is
// deterministic (ties -> smallest id).
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

#include "../../include/pbrt_amd_rt.h"
#include "rt_device.h"
#include "rt_internal.h"

using namespace rtd;

namespace rti {

namespace {

constexpr uint32_t kNoKey32 = 0xffffffffu;

__device__ __forceinline__ int pow4_off(int s) { return ((1 << (2 * s)) - 1) / 3; }

// Pointers known to address global memory.  Loads and stores through them compile to global_*
// instructions; a generic pointer read from a job table compiles to flat_*, whose completion
// is also counted with the LDS operations, so LDS and memory waits serialize.
#define RT_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ RT_GLOBAL T* glb(T* p) {
  return (RT_GLOBAL T*)p;
}

// float4 element i through a generic or a global pointer (the global one is read as the
// native 4-vector, which has no address-space restriction on its copy).
typedef float rt_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld4(const float4* p, size_t i) { return p[i]; }
__device__ __forceinline__ float4 ld4(RT_GLOBAL const float4* p, size_t i) {
  const rt_f4 v = ((RT_GLOBAL const rt_f4*)p)[i];
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 ld4(RT_GLOBAL float4* p, size_t i) {
  return ld4((RT_GLOBAL const float4*)p, i);
}

template <class P>
__device__ __forceinline__ V3 ld3(P p) { return {p[0], p[1], p[2]}; }

template <class P>
__device__ __forceinline__ Tri load_tri(P v9, uint32_t i) {
  const P p = v9 + 9 * (size_t)i;
  return make_tri(ld3(p), ld3(p + 3), ld3(p + 6));
}

// Leaf-test operands of a triangle (v1, e1 = v2 - v1, e2 = v3 - v1, precomputed in FP32 on the
// host) from 3 float4 per triangle; v2/v3 are not filled (the leaf test does not read them).
// Used where triangles are reached through an index (ray-bvh leaves): three 16-byte loads in
// place of nine scalar ones.
template <class P>
__device__ __forceinline__ Tri load_mt(P tmt, uint32_t i) {
  const float4 a = ld4(tmt, 3 * (size_t)i), b = ld4(tmt, 3 * (size_t)i + 1),
               c = ld4(tmt, 3 * (size_t)i + 2);
  Tri T;
  T.v1 = {a.x, a.y, a.z};
  T.e1 = {a.w, b.x, b.y};
  T.e2 = {b.z, b.w, c.x};
  T.v2 = T.v3 = T.v1;
  return T;
}

__device__ __forceinline__ void node_store(float4* __restrict__ nodes, size_t idx, const float* b,
                                           bool valid) {
  float4* p = nodes + 4 * idx;
  p[0] = make_float4(b[0], b[1], b[2], valid ? 1.f : 0.f);
  p[1] = make_float4(b[3], b[4], b[5], 0.f);
  p[2] = make_float4(b[6], b[7], b[8], 0.f);
  p[3] = make_float4(b[9], b[10], b[11], 0.f);
}

// Leaf-major mapping of a block-local ray slot: the rays of one leaf are consecutive slots,
// so a wavefront covers whole leaves.
__device__ __forceinline__ bool slot_xy(const RHGeom& g, int K, int J, int rl, int& x, int& y,
                                        int& leaf) {
  if (rl >= g.rb) return false;
  leaf = rl / g.g;
  int q = rl - leaf * g.g;
  int lk = leaf % g.nlx, lj = leaf / g.nlx;
  x = K * g.bw + lk * g.cx + q % g.cx;
  y = J * g.bh + lj * g.cy + q / g.cx;
  return x < g.W && y < g.H;
}

__device__ __forceinline__ void box_combine(float* a, bool& ha, const float* b, bool hb) {
  if (!hb) return;
  if (!ha) {
    for (int k = 0; k < 12; ++k) a[k] = b[k];
    ha = true;
    return;
  }
  for (int k = 0; k < 3; ++k) a[k] = fmn(a[k], b[k]);
  for (int k = 3; k < 6; ++k) a[k] = fmx(a[k], b[k]);
  for (int k = 6; k < 9; ++k) a[k] = fmn(a[k], b[k]);
  for (int k = 9; k < 12; ++k) a[k] = fmx(a[k], b[k]);
}

// Direction part of ray i: 6DB / IA keep d.  Node "box_uv" keeps (slope, d.z, 0) with the
// slope d.y/d.x, 0 when d.x == 0 (rayhconstruct5DB.cl:23,32-33); the validity-aware build
// (rayhconstructP5DB.cl:25-29,43-47,52-57) gives a ray with d.x == 0 the slope interval
// [-1, 1] and marks it as a restart: the kernel assigns that interval, dropping what the
// leaf's earlier rays gave, and later rays widen it again.
template <bool USE_VALID, bool UV>
__device__ __forceinline__ void ray_box(const RaysDev& r, size_t i, float* v, bool& restart) {
  const float dx = r.dx[i], dy = r.dy[i], dz = r.dz[i];
  restart = false;
  if constexpr (UV) {
    const bool flat = dx == 0.f;
    v[0] = v[3] = flat ? 0.f : dy / dx;
    if (USE_VALID && flat) {
      v[0] = -1.f;
      v[3] = 1.f;
      restart = true;
    }
    v[1] = v[4] = dz;
    v[2] = v[5] = 0.f;
  } else {
    v[0] = v[3] = dx;
    v[1] = v[4] = dy;
    v[2] = v[5] = dz;
  }
  v[6] = v[9] = r.ox[i];
  v[7] = v[10] = r.oy[i];
  v[8] = v[11] = r.oz[i];
}

// Merge of the rays of one leaf in the reference's read order, a before b: the component-wise
// min/max, except that a restart in b (validity-aware 5DB) replaces a's slope interval.
__device__ __forceinline__ void leaf_merge(float* a, bool& ha, bool& ra, const float* b, bool hb,
                                           bool rb) {
  if (!hb) return;
  if (!ha) {
    for (int k = 0; k < 12; ++k) a[k] = b[k];
    ha = true;
    ra = rb;
    return;
  }
  box_combine(a, ha, b, true);
  if (rb) {
    a[0] = b[0];
    a[3] = b[3];
  }
  ra = ra || rb;
}

// Box of the leaf whose first ray is (x0, y0): its rays merged one by one in the reference's
// read order (row by row, left to right); zeros and false when no ray is traced.
template <bool USE_VALID, bool UV>
__device__ bool leaf_box(const RHGeom& g, const RaysDev& r, int x0, int y0, float* v) {
  bool have = false, restart = false;
  for (int y = y0; y < min(g.H, y0 + g.cy); ++y)
    for (int x = x0; x < min(g.W, x0 + g.cx); ++x) {
      const size_t i = (size_t)y * g.W + x;
      if (USE_VALID && r.valid != nullptr && r.valid[i] == 0) continue;
      float b[12];
      bool rb;
      ray_box<USE_VALID, UV>(r, i, b, rb);
      leaf_merge(v, have, restart, b, true, rb);
    }
  if (!have)
    for (int k = 0; k < 12; ++k) v[k] = 0.f;
  return have;
}

// ---------------------------------------------------------------------------------------
// Fused hierarchy build: one work-group builds the leaves and every level of one block
// (rayhconstruct6DB.cl / rayhconstructP6DB.cl / rayhconstruct(P)5DB.cl + levelConstruct6DB.cl /
// levelConstructP6DB.cl / levelConstruct(P)5DB.cl folded into one launch instead of
// 1 + (height-1) launches with host waits, rayhierarchy.cpp:375-549).  A leaf's rays are
// merged in the reference's read order (row by row, left to right): lane pairs at distance
// 1, 2, 4, ... combine the lower lane's part first, so every lane ends with the in-order
// merge.
template <bool USE_VALID, bool UV>
__device__ void build_block(const RHGeom& g, const RaysDev& r, float4* __restrict__ nodes, int b,
                            float (*sBox)[12], int* sVal) {
  const int tid = threadIdx.x;
  const int top = g.L - 1;
  const int K = b % g.gw[top], J = b / g.gw[top];
  const int nleaf = g.nlx * g.nlx;
  const int leaf_off = g.nsub - nleaf;

  const bool pow2 = g.g <= 64 && (g.g & (g.g - 1)) == 0;
  if (pow2) {
    for (int base = 0; base < g.rb; base += kThreads) {
      int rl = base + tid;
      int x = 0, y = 0, leaf = 0;
      float v[12];
      bool have = false, restart = false;
      bool in_block = rl < g.rb;
      if (slot_xy(g, K, J, rl, x, y, leaf)) {
        size_t i = (size_t)y * g.W + x;
        have = !USE_VALID || r.valid == nullptr || r.valid[i] != 0;
        if (have) ray_box<USE_VALID, UV>(r, i, v, restart);
      }
      if (!have)
        for (int k = 0; k < 12; ++k) v[k] = 0.f;
      for (int m = 1; m < g.g; m <<= 1) {
        float o[12];
        for (int k = 0; k < 12; ++k) o[k] = __shfl_xor(v[k], m);
        bool ho = __shfl_xor((int)have, m) != 0;
        bool ro = __shfl_xor((int)restart, m) != 0;
        if (tid & m) {  // the partner's rays come first: ours are merged after them
          leaf_merge(o, ho, ro, v, have, restart);
          for (int k = 0; k < 12; ++k) v[k] = o[k];
          have = ho;
          restart = ro;
        } else {
          leaf_merge(v, have, restart, o, ho, ro);
        }
      }
      if (in_block && (rl % g.g) == 0) {
        int li = rl / g.g;
        for (int k = 0; k < 12; ++k) sBox[leaf_off + li][k] = v[k];
        sVal[leaf_off + li] = have ? 1 : 0;
      }
    }
  } else {
    for (int li = tid; li < nleaf; li += kThreads) {
      float v[12];
      const bool have = leaf_box<USE_VALID, UV>(g, r, K * g.bw + (li % g.nlx) * g.cx,
                                                J * g.bh + (li / g.nlx) * g.cy, v);
      for (int k = 0; k < 12; ++k) sBox[leaf_off + li][k] = v[k];
      sVal[leaf_off + li] = have ? 1 : 0;
    }
  }
  __syncthreads();
  for (int li = tid; li < nleaf; li += kThreads) {
    int gk = K * g.nlx + li % g.nlx, gj = J * g.nlx + li / g.nlx;
    if (gk < g.gw[0] && gj < g.gh[0])
      node_store(nodes, (size_t)g.off[0] + (size_t)gj * g.gw[0] + gk, sBox[leaf_off + li],
                 sVal[leaf_off + li] != 0);
  }
  for (int l = 1; l < g.L; ++l) {
    const int s = top - l, nl = 1 << s, cnl = nl * 2;
    const int soff = pow4_off(s), coff = pow4_off(s + 1);
    for (int i = tid; i < nl * nl; i += kThreads) {
      int dk = i % nl, dj = i / nl;
      float v[12];
      bool have = false;
      for (int a = 0; a < 4; ++a) {
        int c = (2 * dj + (a >> 1)) * cnl + 2 * dk + (a & 1);
        box_combine(v, have, sBox[coff + c], sVal[coff + c] != 0);
      }
      if (!have)
        for (int k = 0; k < 12; ++k) v[k] = 0.f;
      for (int k = 0; k < 12; ++k) sBox[soff + i][k] = v[k];
      sVal[soff + i] = have ? 1 : 0;
      int gk = K * nl + dk, gj = J * nl + dj;
      if (gk < g.gw[l] && gj < g.gh[l])
        node_store(nodes, (size_t)g.off[l] + (size_t)gj * g.gw[l] + gk, v, have);
    }
    __syncthreads();
  }
}

// Job of a global block index in a launch group, for an index that is uniform over the
// wavefront: the lanes read 64 jobs' block prefixes at once and count those at or below the
// index (the prefixes increase with the job), one round trip per 64 jobs instead of a binary
// search's chain of dependent loads.
__device__ __forceinline__ int find_block_item(const RhItem* __restrict__ items, int n,
                                               uint32_t gb) {
  int k = 0;
  for (int c = 0; c < n; c += 64) {
    const int j = c + (int)__lane_id();
    const int m = __popcll(__ballot(j < n && items[j].block_base <= gb));
    k += m;
    if (m < 64) break;
  }
  return k - 1;
}
// Job of a global ray index (per lane).
__device__ __forceinline__ int find_ray_item(const RhItem* __restrict__ items, int n,
                                             uint32_t gr) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].ray_base <= gr) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Build of a block of more than kUnitLevels levels (heights of 5 and more where the image
// allows them): its subtree does not fit the LDS of build_block, so the leaves and every level
// go straight to the node array, and each level reads its four children back from it (the
// work-group's own writes, ordered by the barrier between levels).  Children outside a level's
// grid are skipped, as build_block skips the invalid ones; the merge order is the same.
template <bool USE_VALID, bool UV>
__device__ void build_block_deep(const RHGeom& g, const RaysDev& r, float4* __restrict__ nodes,
                                 int b) {
  const int tid = threadIdx.x;
  const int top = g.L - 1;
  const int K = b % g.gw[top], J = b / g.gw[top];
  for (int li = tid; li < g.nlx * g.nlx; li += kThreads) {
    const int gk = K * g.nlx + li % g.nlx, gj = J * g.nlx + li / g.nlx;
    if (gk >= g.gw[0] || gj >= g.gh[0]) continue;
    float v[12];
    const bool have = leaf_box<USE_VALID, UV>(g, r, gk * g.cx, gj * g.cy, v);
    node_store(nodes, (size_t)g.off[0] + (size_t)gj * g.gw[0] + gk, v, have);
  }
  __syncthreads();
  for (int l = 1; l < g.L; ++l) {
    const int nl = g.nlx >> l;
    for (int i = tid; i < nl * nl; i += kThreads) {
      const int gk = K * nl + i % nl, gj = J * nl + i / nl;
      if (gk >= g.gw[l] || gj >= g.gh[l]) continue;
      float v[12];
      bool have = false;
      for (int a = 0; a < 4; ++a) {
        const int ck = 2 * gk + (a & 1), cj = 2 * gj + (a >> 1);
        if (ck >= g.gw[l - 1] || cj >= g.gh[l - 1]) continue;
        const float4* p = nodes + 4 * ((size_t)g.off[l - 1] + (size_t)cj * g.gw[l - 1] + ck);
        const float4 p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3];
        const float c[12] = {p0.x, p0.y, p0.z, p1.x, p1.y, p1.z,
                             p2.x, p2.y, p2.z, p3.x, p3.y, p3.z};
        box_combine(v, have, c, p0.w != 0.f);
      }
      if (!have)
        for (int k = 0; k < 12; ++k) v[k] = 0.f;
      node_store(nodes, (size_t)g.off[l] + (size_t)gj * g.gw[l] + gk, v, have);
    }
    __syncthreads();
  }
}

template <bool USE_VALID, bool UV>
__global__ __launch_bounds__(256) void rh_group_build_kernel(const RhItem* __restrict__ items,
                                                             int nitems) {
  __shared__ float sBox[kMaxSub][12];
  __shared__ int sVal[kMaxSub];
  const int k = find_block_item(items, nitems, blockIdx.x);
  const RhItem& it = items[k];
  if (it.g.L > kUnitLevels)
    build_block_deep<USE_VALID, UV>(it.g, it.r, it.nodes, (int)(blockIdx.x - it.block_base));
  else
    build_block<USE_VALID, UV>(it.g, it.r, it.nodes, (int)(blockIdx.x - it.block_base), sBox,
                               sVal);
}

// Clears the reduction state of every job: keys, repair keys, occlusion flags, counters.
template <int MODE>
__global__ __launch_bounds__(256) void rh_init_kernel(const RhItem* __restrict__ items, int nitems,
                                                      uint32_t nrays, uint32_t* __restrict__ ctr) {
  const uint32_t gi = blockIdx.x * kThreads + threadIdx.x;
  if (gi < 8) ctr[gi] = 0u;
  if (gi < kClaimQueues) {
    ctr[kClaimBase + kClaimStride * gi] = 0u;      // units
    ctr[kClaimBase + kClaimStride * gi + 1] = 0u;  // per-ray chunk slots
  }
  if (gi >= nrays) return;
  const RhItem& it = items[find_ray_item(items, nitems, gi)];
  const uint32_t i = gi - it.ray_base;
  if (MODE == 2) {
    it.occ[i] = 0;
  } else {
    it.keys[i] = kMissKey;
    if (MODE == 0 && it.repkeys) it.repkeys[i] = 0xffffffffu;
  }
}

// Loads the subtree of block (K, J) into LDS: sub-level s = L-1-l, row-major within it.
template <class P>
__device__ __forceinline__ void load_subtree(const RHGeom& g, int K, int J, P nodes, Node* sN,
                                             int* sV) {
  const int top = g.L - 1;
  for (int i = threadIdx.x; i < g.nsub; i += kThreads) {
    int s = 0;
    while (pow4_off(s + 1) <= i) ++s;
    int local = i - pow4_off(s), nl = 1 << s, l = top - s;
    int gk = K * nl + local % nl, gj = J * nl + local / nl;
    // the level tables are read with constant indices (scalar loads issued together) and
    // selected per lane, not loaded per lane one after another
    int gw = 0, gh = 0, off = 0;
#pragma unroll
    for (int q = 0; q < kUnitLevels; ++q) {
      if (q == l) {
        gw = g.gw[q];
        gh = g.gh[q];
        off = g.off[q];
      }
    }
    int valid = 0;
    Node n{};
    if (gk < gw && gj < gh) {
      const P p = nodes + 4 * ((size_t)off + (size_t)gj * gw + gk);
      float4 a = ld4(p, 0), b = ld4(p, 1), c = ld4(p, 2), d = ld4(p, 3);
      n.dmin = {a.x, a.y, a.z};
      n.dmax = {b.x, b.y, b.z};
      n.omin = {c.x, c.y, c.z};
      n.omax = {d.x, d.y, d.z};
      valid = a.w != 0.f;
    }
    sN[i] = n;
    sV[i] = valid;
  }
}

// ---------------------------------------------------------------------------------------
// Hierarchy x triangle traversal (intersection6DB.cl, intersection6DB2.cl, intersectionP6DB.cl,
// intersectionIA.cl), breadth first with compaction at every level, one work unit (a block and
// a few of its passing clusters) at a time.  Triangles are stored Morton-sorted in clusters of
// 64 with AABBs.
//
// The result of a ray depends only on triangles that pass the whole cull chain (the
// reference's candidate set) AND pass the leaf test.  The traversal therefore descends with
// cheap exact-conservative bounds of the culls (never false where the cull is true) and
// evaluates the exact cull chain only for (triangle, leaf) pairs in which some ray passes the
// leaf test:
//   phase A  each thread takes one triangle of the unit's clusters through the per-triangle
//            bound against the block root; survivors queue in LDS (ballot compaction);
//   phase B  queued triangles are expanded kExpand at a time into 4 (triangle, child) items
//            each, bounded in parallel and compacted again, level by level down to the leaves
//            (depth first over rounds, so each level needs one queue);
//   phase C  (triangle, leaf) pairs are expanded into (pair, ray) leaf tests over the rays
//            staged in LDS; the pairs with a passing test are compacted and get the exact cull
//            chain; the tests of verified pairs update the ray (LDS atomic min of the (t, id)
//            key / occlusion bit).
constexpr int kCluster = 64;     // triangles per cluster
constexpr uint32_t kNoUnit = 0xffffffffu;  // no unit left in the queue
constexpr int kQ0Cap = 1024;     // root survivors held before expansion
constexpr int kQCap = 1024;      // survivors of one expansion round
constexpr int kExpand = 256;     // parents expanded per round (4 children each, 4 per thread)

// Cluster bound: conservative for every triangle (rayhierarchy) or BVH node box (ray-bvh, whose
// cull is intersectsNode in every mode) inside [cmin, cmax].
template <int MODE, int NODE, bool BVH = false>
__device__ __forceinline__ bool cluster_bound(const Node& n, V3 cmin, V3 cmax) {
  if (NODE == 1) return cull_ia(n, cmin, cmax);
  if (NODE == 2) return bound_5db(n, cmin, cmax);
  if (MODE == 0 && !BVH) return cluster_tri(n, cmin, cmax);
  return cluster_node6(n, cmin, cmax);
}
__device__ __forceinline__ V3 tri_bmin(const Tri& T) { return vmin(vmin(T.v1, T.v2), T.v3); }
__device__ __forceinline__ V3 tri_bmax(const Tri& T) { return vmax(vmax(T.v1, T.v2), T.v3); }
// Exact-conservative bound of the cull (IA is cheap enough to be its own bound).
template <int MODE, int NODE>
__device__ __forceinline__ bool cull_bound(const Node& n, const Tri& T) {
  if (NODE == 1) return cull_ia(n, tri_bmin(T), tri_bmax(T));
  if (NODE == 2) return bound_5db(n, tri_bmin(T), tri_bmax(T));
  if (MODE == 0) return prefilter_tri(n, T);
  return prefilter_node6(n, tri_bmin(T), tri_bmax(T));
}
// Bound of sorted triangle `slot`: intersectsTri's bound needs the vertices; the others
// depend on the triangle's AABB only, which is stored per sorted triangle.
template <int MODE, int NODE, class PF, class P4>
__device__ __forceinline__ bool cull_bound_slot(const Node& n, PF v9s, P4 tbox, uint32_t slot) {
  if (MODE == 0 && NODE == 0) return prefilter_tri(n, load_tri(v9s, slot));
  const float4 lo = ld4(tbox, 2 * (size_t)slot), hi = ld4(tbox, 2 * (size_t)slot + 1);
  const V3 bmin = {lo.x, lo.y, lo.z}, bmax = {hi.x, hi.y, hi.z};
  if (NODE == 1) return cull_ia(n, bmin, bmax);
  if (NODE == 2) return bound_5db(n, bmin, bmax);
  return prefilter_node6(n, bmin, bmax);
}
// The reference's cull: intersectsTri for 6DB primary rays, else intersectsNode of the node
// type on the triangle's bounds (the 5DB kernels use it in every mode, the repair pass's
// yetAnotherIntersection5DB.cl:78 included).
template <int MODE, int NODE>
__device__ __forceinline__ bool cull_exact(const Node& n, const Tri& T) {
  if (NODE == 1) return cull_ia(n, tri_bmin(T), tri_bmax(T));
  if (NODE == 2) return cull_5db(n, tri_bmin(T), tri_bmax(T));
  if (MODE == 0) return cull_tri(n, T);
  return cull_node6(n, tri_bmin(T), tri_bmax(T));
}

// Appends `item` for the lanes with `pass` set; one LDS atomic per wavefront.
__device__ __forceinline__ void ballot_append(uint32_t* q, int* cnt, bool pass, uint32_t item) {
  const uint64_t bal = __ballot(pass);
  const int lane = __lane_id();
  const int prefix = __popcll(bal & ((1ull << lane) - 1ull));
  int base = 0;
  if (lane == 0 && bal) base = atomicAdd(cnt, __popcll(bal));
  base = __shfl(base, 0);
  if (pass) q[base + prefix] = item;
}

// Chunk cursors of the expansion levels (kept in registers, identical in every thread).
struct Cursors {
  int c0 = 0, c1 = 0, c2 = 0, c3 = 0;
  __device__ int get(int s) const { return s == 0 ? c0 : s == 1 ? c1 : s == 2 ? c2 : c3; }
  __device__ void set(int s, int v) {
    if (s == 0) c0 = v;
    else if (s == 1) c1 = v;
    else if (s == 2) c2 = v;
    else c3 = v;
  }
};

// Phase C pair state.  Closest hits: 1 = some ray of the pair passes its leaf test, | kPairOk =
// the exact cull chain holds.  Occlusion with at most 31 rays per leaf: bit r = ray r of the
// leaf passes its leaf test, kPairOk as above; the verified bits are applied without running
// the leaf test again (it is the same deterministic test on the same operands).
template <int MODE>
using PairState = typename std::conditional<MODE == 2, uint32_t, uint8_t>::type;
template <int MODE>
constexpr PairState<MODE> kPairOk = MODE == 2 ? (PairState<MODE>)0x80000000u : (PairState<MODE>)2;

template <int LMAX>
struct SubSize {
  static constexpr int value = ((1 << (2 * LMAX)) - 1) / 3;
};

template <int MODE, int RBMAX, int LMAX>
struct RhShared {
  Node sN[SubSize<LMAX>::value];
  int sV[SubSize<LMAX>::value];
  float4 sRo[RBMAX];  // origin, tmin
  float4 sRd[RBMAX];  // direction, tmax
  unsigned long long sKey[MODE == 2 ? 1 : RBMAX];
  // repair key per ray (primary), or occlusion bit per ray (shadow); a bitmask keeps the
  // occlusion kernel within the LDS of six work-groups per CU
  uint32_t sAux[MODE == 0 ? RBMAX : MODE == 2 ? (RBMAX + 31) / 32 : 1];
  uint8_t sRv[RBMAX];    // ray exists and is traced
  PairState<MODE> sPair[kQ0Cap];  // phase C pair state
  uint32_t sQ0[kQ0Cap];
  uint32_t sQ[LMAX > 1 ? LMAX - 1 : 1][kQCap];  // level s >= 1 survivors live in sQ[s - 1]
  int sCnt[kUnitLevels];
  int sSeg[kThreads / 64];  // phase C: flagged pairs of each wavefront's quarter of the list
  int sTraced, sOccluded;
};

// Marks ray `slot` occluded; true when it was not marked before.
template <class SH>
__device__ __forceinline__ bool mark_occluded(SH& sh, int slot) {
  const uint32_t bit = 1u << (slot & 31);
  return (atomicOr(&sh.sAux[slot >> 5], bit) & bit) == 0u;
}

// The exact cull chain from the block root to leaf `li` (reference candidate test).
template <int MODE, int NODE, int RBMAX, int LMAX>
__device__ __forceinline__ bool exact_chain(const RHGeom& g, const RhShared<MODE, RBMAX, LMAX>& sh,
                                            const Tri& T, int li) {
  const int top = g.L - 1;
  const int lx = li & (g.nlx - 1), ly = li >> top;
  for (int s = 0; s <= top; ++s) {
    const int sh_ = top - s;
    const int idx = pow4_off(s) + (ly >> sh_) * (1 << s) + (lx >> sh_);
    if (!sh.sV[idx] || !cull_exact<MODE, NODE>(sh.sN[idx], T)) return false;
  }
  return true;
}

// Leaf test of ray slot `slot` against T; returns whether it would update the ray.
template <int MODE, class SH>
__device__ __forceinline__ bool leaf_hit(const SH& sh, int slot, const Tri& T, float& t, float& b1,
                                         float& b2) {
  const float4 o4 = sh.sRo[slot], d4 = sh.sRd[slot];
  if (!mt_strict(V3{o4.x, o4.y, o4.z}, V3{d4.x, d4.y, d4.z}, T, t, b1, b2)) return false;
  if (MODE == 2) return t > o4.w && t < d4.w;  // intersectAllLeaves.cl:177-179
  const bool main = t > o4.w && t < __int_as_float(0x7f800000);  // intersectAllLeaves.cl:57-61
  return main || (MODE == 0 && repair_region(b1, b2) && t >= o4.w);
}

// Phase C over `n` (triangle slot, leaf) pairs.  Item = (pair, ray of its leaf): consecutive
// lanes take consecutive ray slots of one pair, so the LDS ray reads are conflict-free and
// the triangle load is a broadcast.
//
// Between C1 and C2 every wavefront moves the flagged pairs of its quarter of the list to the
// front of that quarter (in place, in list order), so that the exact cull chain of C2 runs
// with every lane on a flagged pair and C3 sees flagged pairs only.  Pairs without a passing
// leaf test cannot change a ray, so dropping them changes nothing.
template <int MODE, int NODE, int RBMAX, int LMAX, class PF, class PU>
__device__ __forceinline__ void leaf_rays(const RHGeom& g, PF v9s, PU perm, uint32_t* q, int n,
                                          RhShared<MODE, RBMAX, LMAX>& sh) {
  const int G = g.g;
  const int total = n * G;
  const bool pow2 = (G & (G - 1)) == 0;
  const int lg = __ffs(G) - 1;
  const bool masked = MODE == 2 && G <= 31;
  for (int e = threadIdx.x; e < n; e += kThreads) sh.sPair[e] = 0;
  __syncthreads();
  // C1: leaf tests; flag pairs with a passing test (occlusion: the passing rays).  A ray found
  // occluded has its sRv cleared, so it takes no further tests.
  for (int item = threadIdx.x; item < total; item += kThreads) {
    const int e = pow2 ? item >> lg : item / G;
    const uint32_t P = q[e];
    const int r = item - e * G;
    const int slot = (int)(P & 63u) * G + r;
    if (!sh.sRv[slot]) continue;
    float t, b1, b2;
    if (leaf_hit<MODE>(sh, slot, load_tri(v9s, P >> 6), t, b1, b2)) {
      if constexpr (MODE == 2) {
        if (masked) atomicOr(&sh.sPair[e], 1u << r);
        else sh.sPair[e] = 1;
      } else {
        sh.sPair[e] = 1;
      }
    }
  }
  __syncthreads();
  // Compaction: wavefront w owns list positions [w * S, min(n, (w + 1) * S)), S a multiple of
  // 64.  A flagged entry moves to a position at or below its own, and each 64-entry chunk is
  // read in full before it is written, so no entry is overwritten before it has been read.
  static_assert(kThreads == 256, "four wavefronts of 64 per work-group");
  const int lane = __lane_id();
  const int wv = threadIdx.x >> 6;
  const int S = ((n + kThreads - 1) / kThreads) * 64;
  const int s0 = wv * S, s1 = min(n, s0 + S);
  int cw = 0;
  for (int base = s0; base < s1; base += 64) {
    const int e = base + lane;
    PairState<MODE> ps = 0;
    uint32_t P = 0;
    if (e < s1) {
      ps = sh.sPair[e];
      P = q[e];
    }
    const uint64_t bal = __ballot(ps != 0);
    if (ps != 0) {
      const int d = s0 + cw + __popcll(bal & ((1ull << lane) - 1ull));
      q[d] = P;
      sh.sPair[d] = ps;
    }
    cw += __popcll(bal);
  }
  if (lane == 0) sh.sSeg[wv] = cw;
  // C2: exact cull chain of the wavefront's flagged pairs (written by this wavefront above)
  int verified = 0;
  for (int i = lane; i < cw; i += 64) {
    const int e = s0 + i;
    const PairState<MODE> ps = sh.sPair[e];
    const uint32_t P = q[e];
    const bool ok = exact_chain<MODE, NODE, RBMAX, LMAX>(g, sh, load_tri(v9s, P >> 6), (int)(P & 63u));
    sh.sPair[e] = ok ? (PairState<MODE>)(ps | kPairOk<MODE>) : (PairState<MODE>)0;
    verified |= ok;
  }
  if (!__syncthreads_or(verified)) return;
  // C3: apply the leaf tests of verified pairs.  Flagged pair j (over the four quarters in
  // order) sits at position w * S + (j - p[w]), p the prefix sums of the quarters' counts.
  const int p1 = sh.sSeg[0], p2 = p1 + sh.sSeg[1], p3 = p2 + sh.sSeg[2], nf = p3 + sh.sSeg[3];
  const int ftotal = nf * G;
  for (int item = threadIdx.x; item < ftotal; item += kThreads) {
    const int j = pow2 ? item >> lg : item / G;
    const int r = item - j * G;
    const int w = (j >= p1) + (j >= p2) + (j >= p3);
    const int e = w * S + j - (w == 0 ? 0 : w == 1 ? p1 : w == 2 ? p2 : p3);
    const PairState<MODE> ps = sh.sPair[e];
    if (!(ps & kPairOk<MODE>)) continue;
    const uint32_t P = q[e];
    const uint32_t slot_t = P >> 6;
    const int slot = (int)(P & 63u) * G + r;
    if (masked) {  // occlusion: the rays that passed in C1
      if (!((ps >> r) & 1u)) continue;
      if (mark_occluded(sh, slot)) atomicAdd(&sh.sOccluded, 1);
      sh.sRv[slot] = 0;
      continue;
    }
    if (!sh.sRv[slot]) continue;
    float t, b1, b2;
    if (!leaf_hit<MODE>(sh, slot, load_tri(v9s, slot_t), t, b1, b2)) continue;
    const float tmin = sh.sRo[slot].w;
    if (MODE == 2) {
      if (mark_occluded(sh, slot)) atomicAdd(&sh.sOccluded, 1);
      sh.sRv[slot] = 0;
    } else {
      if (t > tmin && t < __int_as_float(0x7f800000))
        atomicMin(&sh.sKey[slot], (unsigned long long)hit_key(t, (int32_t)perm[slot_t]));
      if constexpr (MODE == 0) {
        if (repair_region(b1, b2) && t >= tmin) atomicMin(&sh.sAux[slot], fkey(t));
      }
    }
  }
  __syncthreads();
}

// Phases B and C for the triangles queued in sQ0; empties the queue.
template <int MODE, int NODE, int RBMAX, int LMAX, class PF, class P4, class PU>
__device__ __forceinline__ void expand_queue(const RHGeom& g, PF v9s, P4 tbox, PU perm,
                                             RhShared<MODE, RBMAX, LMAX>& sh) {
  const int tid = threadIdx.x;
  const int top = g.L - 1;
  if (top == 0) {
    leaf_rays<MODE, NODE, RBMAX, LMAX>(g, v9s, perm, sh.sQ0, sh.sCnt[0], sh);
  } else {
    Cursors cur;
    int s = 0;
    while (true) {
      const int cnt = sh.sCnt[s];
      if (s == top) {
        leaf_rays<MODE, NODE, RBMAX, LMAX>(g, v9s, perm, sh.sQ[s - 1], cnt, sh);
        --s;
        continue;
      }
      const int c = cur.get(s);
      if (c >= cnt) {
        cur.set(s, 0);
        if (s == 0) break;
        --s;
        continue;
      }
      cur.set(s, c + kExpand);
      const int np = min(kExpand, cnt - c);
      const uint32_t* src = s == 0 ? sh.sQ0 : sh.sQ[s - 1];
      __syncthreads();
      if (tid == 0) sh.sCnt[s + 1] = 0;
      __syncthreads();
      // 4 children per parent, every thread takes children until all np * 4 are bounded (a
      // uniform trip count: the append is a wavefront ballot)
      for (int cb = 0; cb < np * 4; cb += kThreads) {
        const int ct = cb + tid;
        bool ok = false;
        uint32_t item = 0;
        if (ct < np * 4) {
          const uint32_t P = src[c + (ct >> 2)];
          const uint32_t slot_t = P >> 6;
          const int pl = (int)(P & 63u), a = ct & 3;
          const int pnl = 1 << s;
          const int dk = pl & (pnl - 1), dj = pl >> s;
          const int ch = (2 * dj + (a >> 1)) * (2 * pnl) + 2 * dk + (a & 1);
          const int ni = pow4_off(s + 1) + ch;
          item = (slot_t << 6) | (uint32_t)ch;
          if (sh.sV[ni]) ok = cull_bound_slot<MODE, NODE>(sh.sN[ni], v9s, tbox, slot_t);
        }
        ballot_append(sh.sQ[s], &sh.sCnt[s + 1], ok, item);
      }
      __syncthreads();
      ++s;
    }
  }
  if (tid == 0) sh.sCnt[0] = 0;
  __syncthreads();
}

// Cluster pass: one work-group per block (of any job of the group) tests every cluster (of
// triangles, or of ray-bvh frontier roots) against the block root (cluster bound), stores the
// passing cluster ids in the job's clist[block * nclusters + ...] and appends work units
// {job, block, first, count} of at most `per_unit` passing clusters.
//
// A block whose root passes at least `ray_clusters` clusters of a scene of at least four times
// as many, or whose leaves' ray directions spread by more than kIncoherent on average, is an
// incoherent ray group: its candidate set
// holds most (leaf, triangle) pairs, so it is queued for the per-ray path instead, as 64-ray
// chunks {job, block * chunks per block + chunk} in the block's fixed slots of the queue (no
// shared counter: thousands of blocks appending through one atomic serialise on it); the
// slots of every other block are marked empty.  ray_clusters == 0 queues every block with a
// traced ray that way without scanning the clusters, and so does a block of more than
// kMaxBlockRays rays or kUnitLevels levels (the unit kernels stage at most that much; the
// reference's own scenes use such blocks, e.g. 100x4 leaves at height 3 or 4).
constexpr float kIncoherent = 0.25f;
template <int MODE, int NODE, bool BVH>
__global__ __launch_bounds__(256) void rh_cluster_kernel(const RhItem* __restrict__ items,
                                                         int nitems, uint4* __restrict__ ustage,
                                                         uint4* __restrict__ bunits,
                                                         uint2* __restrict__ rq,
                                                         uint32_t* __restrict__ ctr,
                                                         int per_unit, uint32_t ray_clusters) {
  __shared__ int sCnt;
  __shared__ float sSpread;
  __shared__ int sLeaves;
  const int tid = threadIdx.x;
  const int k = find_block_item(items, nitems, blockIdx.x);
  const RhItem& it = items[k];
  const RHGeom& g = it.g;
  const int top = g.L - 1;
  const int b = (int)(blockIdx.x - it.block_base);
  const int K = b % g.gw[top], J = b / g.gw[top];
  const float4* p = it.nodes + 4 * ((size_t)g.off[top] + (size_t)J * g.gw[top] + K);
  const float4 a0 = p[0], a1 = p[1], a2 = p[2], a3 = p[3];
  // this block's chunk slots: filled for the per-ray path, marked empty otherwise
  const int nch = (g.rb + 63) / 64;
  uint2* const slots = rq + it.chunk_base + (size_t)b * (size_t)nch;
  auto queue_chunks = [&](bool per_ray) {
    for (int q = tid; q < nch; q += kThreads)
      slots[q] = make_uint2(per_ray ? (uint32_t)k : 0xffffffffu,
                            (uint32_t)b * (uint32_t)nch + (uint32_t)q);
    if (per_ray && tid == 0) ctr[4] = 1u;  // the per-ray kernel has work
    if (tid == 0) bunits[blockIdx.x] = make_uint4(0u, 0u, 0u, 0u);  // block-path blocks rewrite it
  };
  if (a0.w == 0.f) {  // no traced ray below this root (uniform)
    queue_chunks(false);
    return;
  }
  Node root;
  root.dmin = {a0.x, a0.y, a0.z};
  root.dmax = {a1.x, a1.y, a1.z};
  root.omin = {a2.x, a2.y, a2.z};
  root.omax = {a3.x, a3.y, a3.z};
  if (ray_clusters == 0 || g.rb > kMaxBlockRays || g.L > kUnitLevels) {
    // to the per-ray path without the cluster scan (no cluster list needed)
    queue_chunks(true);
    return;
  }
  if (tid == 0) {
    sCnt = 0;
    sSpread = 0.f;
    sLeaves = 0;
  }
  __syncthreads();
  // mean direction spread of the block's leaves (largest component range of each leaf's ray
  // directions): near 0 for camera rays and shadow rays toward a light from nearby points,
  // large for random segments, diffuse bounces and shadow rays from scattered points
  if (tid < g.nlx * g.nlx) {
    const int gk = K * g.nlx + tid % g.nlx, gj = J * g.nlx + tid / g.nlx;
    if (gk < g.gw[0] && gj < g.gh[0]) {
      const float4* q = it.nodes + 4 * ((size_t)g.off[0] + (size_t)gj * g.gw[0] + gk);
      const float4 lo = q[0], hi = q[1];
      if (lo.w != 0.f) {
        // box_uv leaves hold slope and d.z intervals: the d.z range is the spread
        const float sp = NODE == 2 ? hi.y - lo.y
                                   : fmaxf(fmaxf(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z);
        atomicAdd(&sSpread, sp);
        atomicAdd(&sLeaves, 1);
      }
    }
  }
  const uint32_t nclusters = it.nclusters;
  const float4* __restrict__ cbox = it.cbox;
  uint32_t* out = it.clist + (size_t)b * nclusters;
  __syncthreads();
  // incoherent: at least `ray_clusters` clusters pass in a scene of at least 4 * ray_clusters
  // clusters (in smaller scenes the block path's work is bounded by the scene), or the block's
  // leaves spread their ray directions by more than kIncoherent on average (ray_clusters above
  // 2^29: never; 1: every block with a passing cluster).  The scan stops as soon as the
  // passing count decides that (a per-ray block needs no cluster list).
  const bool wide = sLeaves > 0 && sSpread > kIncoherent * (float)sLeaves;
  const bool may_heavy = ray_clusters == 1u || (uint64_t)nclusters >= 4ull * ray_clusters;
  const uint32_t enough = (ray_clusters <= (1u << 29) && wide) ? 1u
                          : may_heavy ? ray_clusters : 0xffffffffu;
  for (uint32_t cb = 0; cb < nclusters; cb += kThreads) {
    const uint32_t ci = cb + tid;
    bool ok = false;
    if (ci < nclusters) {
      const float4 lo = cbox[2 * (size_t)ci], hi = cbox[2 * (size_t)ci + 1];
      ok = cluster_bound<MODE, NODE, BVH>(root, V3{lo.x, lo.y, lo.z}, V3{hi.x, hi.y, hi.z});
    }
    ballot_append(out, &sCnt, ok, ci);
    if (enough != 0xffffffffu) {
      __syncthreads();
      if ((uint32_t)sCnt >= enough) break;  // uniform: every thread reads the same count
      __syncthreads();
    }
  }
  __syncthreads();
  const int cnt = sCnt;
  const bool heavy = (uint32_t)cnt >= ray_clusters && may_heavy;
  const bool per_ray = heavy || (ray_clusters <= (1u << 29) && cnt > 0 && wide);
  queue_chunks(per_ray);
  if (per_ray) return;
  // the block's units go to its fixed staging slots (no shared append counter: thousands of
  // blocks appending through one atomic serialise on it); rh_unit_scan_kernel and
  // rh_unit_pack_kernel pack them
  const int nu = (cnt + per_unit - 1) / per_unit;
  const uint32_t first = it.unit_base + (uint32_t)b * it.unit_stride;
  for (int u = tid; u < nu; u += kThreads)
    ustage[first + u] = make_uint4((uint32_t)k, (uint32_t)b, (uint32_t)(u * per_unit),
                                   (uint32_t)min(per_unit, cnt - u * per_unit));
  if (tid == 0) bunits[blockIdx.x] = make_uint4((uint32_t)nu, first, 0u, 0u);
}

// Packs the staged units of every block of a group into one queue in block order.
// - The scan takes tiles of 1,024 consecutive blocks, one work-group per tile and one block
//   per thread (coalesced). It scans within each wavefront by shuffles and across the 16
//   wavefronts through LDS, leaving each block's exclusive prefix within its tile in .z and
//   the tile's total in .w of the tile's first block.
// - The pack copies each block's units to their queue slots, one wavefront per block (four to
//   a work-group), so the copies of all blocks are in flight at once. Each wavefront first
//   adds up the totals of the tiles before its block's tile, one tile per lane. The wavefront
//   of the last block sets the queue's length, ctr[0].
// A single work-group walking the tiles in turn took ~37 us for the 58,000 blocks of a
// 64-scene batch.
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void rh_unit_scan_kernel(uint4* __restrict__ bunits,
                                                                    uint32_t nblocks) {
  __shared__ uint32_t sWave[kScanThreads / 64];
  const uint32_t tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  const uint32_t b = blockIdx.x * kScanThreads + tid;
  const uint32_t n = b < nblocks ? bunits[b].x : 0u;
  uint32_t incl = n;  // inclusive prefix within the wavefront
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t v = (uint32_t)__shfl_up((int)incl, d);
    if (lane >= d) incl += v;
  }
  if (lane == 63) sWave[wave] = incl;
  __syncthreads();
  uint32_t before = 0;  // units of the earlier wavefronts of the tile
  for (uint32_t w = 0; w < wave; ++w) before += sWave[w];
  if (b < nblocks) bunits[b].z = before + incl - n;
  if (tid == 0) {  // b is the tile's first block, inside the grid's range
    uint32_t total = 0;
    for (uint32_t w = 0; w < kScanThreads / 64; ++w) total += sWave[w];
    bunits[b].w = total;
  }
}
constexpr int kPackThreads = 256;
__global__ __launch_bounds__(kPackThreads) void rh_unit_pack_kernel(
    const uint4* __restrict__ bunits, uint32_t nblocks, const uint4* __restrict__ ustage,
    uint4* __restrict__ units, uint32_t* __restrict__ ctr) {
  const uint32_t b = blockIdx.x * (kPackThreads / 64u) + (threadIdx.x >> 6);
  if (b >= nblocks) return;  // uniform within the wavefront
  const uint32_t lane = __lane_id();
  const uint4 bu = bunits[b];
  uint32_t off = 0;  // units of the earlier tiles
  for (uint32_t t = lane; t < b / (uint32_t)kScanThreads; t += 64u)
    off += bunits[(size_t)t * kScanThreads].w;
  for (int d = 32; d > 0; d >>= 1) off += (uint32_t)__shfl_xor((int)off, d);
  const uint32_t to = off + bu.z;
  for (uint32_t u = lane; u < bu.x; u += 64u) units[to + u] = ustage[bu.y + u];
  if (b == nblocks - 1 && lane == 0) ctr[0] = to + bu.x;
}

// Claims the next unit of a queue of nunits split into kClaimQueues contiguous ranges with a
// counter each.  Work-group w starts in range w % kClaimQueues: the hardware deals
// work-groups to the 8 XCDs in turn, so each XCD works through its own range of blocks in its
// own L2, and one counter serves an eighth of the claims (claims of one counter serialise).
// A work-group whose range is used up moves on to the next.  Returns nunits when none is
// left.  One thread of the work-group calls it and keeps the state.  word 0 selects the unit
// counters, word 1 those of the per-ray chunk slots.
struct UnitClaim {
  uint32_t q, tried;
};
__device__ __forceinline__ uint32_t claim_unit(uint32_t* ctr, uint32_t nunits, UnitClaim& c,
                                               int word = 0) {
  while (c.tried < (uint32_t)kClaimQueues) {
    const uint32_t lo = (uint32_t)((uint64_t)nunits * c.q / kClaimQueues);
    const uint32_t hi = (uint32_t)((uint64_t)nunits * (c.q + 1) / kClaimQueues);
    const uint32_t u = lo + atomicAdd(&ctr[kClaimBase + kClaimStride * c.q + word], 1u);
    if (u < hi) return u;
    c.q = (c.q + 1) % kClaimQueues;
    ++c.tried;
  }
  return nunits;
}

// Persistent unit kernel: a fixed grid takes units from the group's queue with an atomic
// counter (work stealing across every job of the group).  A unit stages its block's rays and
// subtree, takes the triangles of its clusters through the per-triangle bound (phase A),
// expands the survivors (phase B), runs the verified leaf tests (phase C) and merges the block
// results into the job's per-ray keys / occlusion flags with atomics.
template <int MODE, int NODE, int RBMAX, int LMAX>
__global__ __launch_bounds__(256) void rh_unit_kernel(const RhItem* __restrict__ items,
                                                      const uint4* __restrict__ units,
                                                      uint32_t* __restrict__ ctr) {
  __shared__ RhShared<MODE, RBMAX, LMAX> sh;
  __shared__ uint4 sUnit;  // the unit taken next; x = kNoUnit when the queue is empty
  const int tid = threadIdx.x;
  const uint32_t nunits = __builtin_amdgcn_readfirstlane(ctr[0]);
  if (nunits == 0) return;  // every block went to the per-ray path (or had no work)
  UnitClaim claim{blockIdx.x % kClaimQueues, 0u};
  if (tid == 0) {
    const uint32_t u = claim_unit(ctr, nunits, claim);
    sUnit = u < nunits ? units[u] : make_uint4(kNoUnit, 0u, 0u, 0u);
  }
  if (tid < kUnitLevels) sh.sCnt[tid] = 0;
  if (tid == 0) {
    sh.sTraced = 0;
    sh.sOccluded = 0;
  }
  __syncthreads();
  for (;;) {
    const uint4 unit = make_uint4(__builtin_amdgcn_readfirstlane(sUnit.x),
                                  __builtin_amdgcn_readfirstlane(sUnit.y),
                                  __builtin_amdgcn_readfirstlane(sUnit.z),
                                  __builtin_amdgcn_readfirstlane(sUnit.w));
    if (unit.x == kNoUnit) break;
    const RhItem& it = items[unit.x];
    const RHGeom& g = it.g;
    const RaysDev& r = it.r;
    const auto v9s = glb(it.v9s);
    const auto perm = glb(it.perm);
    const auto tbox = glb(it.tbox);
    const uint32_t ntri = it.ntri;
    const int top = g.L - 1;
    const int b = (int)unit.y;
    const int K = b % g.gw[top], J = b / g.gw[top];
    const auto cl = glb(it.clist) + (size_t)b * it.nclusters + unit.z;
    const int ncl = (int)unit.w;
    // first cluster of this thread, loaded with the staging loads
    const uint32_t cfirst = tid / kCluster < ncl ? cl[tid / kCluster] : 0u;
    // subtree, rays, validity and occlusion flags: independent loads, issued together
    load_subtree(g, K, J, glb(it.nodes), sh.sN, sh.sV);
    for (int slot = tid; slot < RBMAX; slot += kThreads) {
      int x = 0, y = 0, leaf = 0;
      bool traced = false;
      if (slot_xy(g, K, J, slot, x, y, leaf)) {
        const size_t idx = (size_t)y * g.W + x;
        const bool valid = MODE == 0 || r.valid == nullptr || glb(r.valid)[idx] != 0;
        // a ray another unit already found occluded needs no more tests
        const bool found = MODE == 2 && glb(it.occ)[idx] != 0;
        traced = valid && !found;
        sh.sRo[slot] = make_float4(glb(r.ox)[idx], glb(r.oy)[idx], glb(r.oz)[idx],
                                   glb(r.tmin)[idx]);
        sh.sRd[slot] = make_float4(glb(r.dx)[idx], glb(r.dy)[idx], glb(r.dz)[idx],
                                   glb(r.tmax)[idx]);
      }
      sh.sRv[slot] = traced ? 1 : 0;
      if (MODE != 2) sh.sKey[slot] = kMissKey;
      if constexpr (MODE == 0) sh.sAux[slot] = kNoKey32;
      if constexpr (MODE == 2) {
        if ((slot & 31) == 0) sh.sAux[slot >> 5] = 0u;
      }
      const uint64_t bal = __ballot(traced);
      if (__lane_id() == 0 && bal) atomicAdd(&sh.sTraced, __popcll(bal));
    }
    __syncthreads();
    if (sh.sV[0] && sh.sTraced > 0) {
      const Node& root = sh.sN[0];
      bool done = false;
      for (int k0 = 0; k0 < ncl && !done; k0 += kThreads / kCluster) {
        const int k = k0 + tid / kCluster;
        const uint32_t c = k0 == 0 ? cfirst : (k < ncl ? cl[k] : 0u);
        const uint32_t slot = k < ncl ? c * kCluster + (uint32_t)(tid % kCluster) : ntri;
        bool pass = false;
        if (slot < ntri) pass = cull_bound_slot<MODE, NODE>(root, v9s, tbox, slot);
        ballot_append(sh.sQ0, &sh.sCnt[0], pass, slot << 6);
        __syncthreads();
        const int cnt0 = sh.sCnt[0];
        __syncthreads();  // every thread has read the count before the next round appends
        if (cnt0 > kQ0Cap - kThreads) {
          expand_queue<MODE, NODE, RBMAX, LMAX>(g, v9s, tbox, perm, sh);
          if (MODE == 2 && sh.sOccluded >= sh.sTraced) done = true;
        }
      }
      if (!done && sh.sCnt[0] > 0) expand_queue<MODE, NODE, RBMAX, LMAX>(g, v9s, tbox, perm, sh);
    }
    // The next unit is claimed once this one's traversal is over (a claim made earlier would
    // hold a second unit per work-group when the queue runs dry); the atomic and the
    // descriptor load overlap the merge below.
    uint4 ndesc = make_uint4(kNoUnit, 0u, 0u, 0u);
    if (tid == 0) {
      const uint32_t next = claim_unit(ctr, nunits, claim);
      if (next < nunits) ndesc = units[next];
    }
    for (int slot = tid; slot < g.rb; slot += kThreads) {
      int x, y, leaf;
      if (!slot_xy(g, K, J, slot, x, y, leaf)) continue;
      const size_t idx = (size_t)y * g.W + x;
      if (MODE == 2) {
        if ((sh.sAux[slot >> 5] >> (slot & 31)) & 1u) glb(it.occ)[idx] = 1;
      } else {
        const unsigned long long key = sh.sKey[slot];
        if (key != kMissKey) atomicMin((unsigned long long*)&it.keys[idx], key);
        if constexpr (MODE == 0) {
          if (it.repkeys) {
            const uint32_t rk = sh.sAux[slot];
            if (rk != kNoKey32) atomicMin(&it.repkeys[idx], rk);
          }
        }
      }
    }
    __syncthreads();  // LDS is rewritten by the next unit
    if (tid == 0) {
      sUnit = ndesc;
      sh.sTraced = 0;
      sh.sOccluded = 0;
    }
    if (tid < kUnitLevels) sh.sCnt[tid] = 0;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------
// ray-bvh as a launch group (intersection6DB_BVH.cl, intersection6DB_BVH2.cl,
// intersectionP6DB_BVH.cl).  The reference walks, for every block root and every frontier
// root of the CPU-built BVH, the product of the block's hierarchy and the root's BVH subtree
// in lockstep (both descend while interior; a leaf waits for the other side), culling each
// (hierarchy node, BVH node) pair with intersectsNode / the IA test, and tests the triangles
// of every (hierarchy leaf, BVH leaf) pair it reaches (only the first one for secondary rays).
// A (leaf, BVH leaf) pair is reached iff every pair on its unique lockstep path passes, so the
// same hit-first scheme as the rayhierarchy path applies: the traversal descends with the
// exact-conservative bound of the cull, runs the leaf tests, and evaluates the exact cull
// chain only for pairs in which some ray passes its leaf test.
constexpr int kBvhQCap = 3072;   // traversal stack: 256 + 224 * depth <= 3072 for depth <= 12
constexpr int kPairCap = 512;    // (BVH leaf, hierarchy leaf) pairs awaiting leaf tests
constexpr int kBvhPop = 32;      // traversal items expanded per round (8 children each)

template <int MODE, int RBMAX, int LMAX>
struct BvhShared {
  Node sN[SubSize<LMAX>::value];
  int sV[SubSize<LMAX>::value];
  float4 sRo[RBMAX];
  float4 sRd[RBMAX];
  unsigned long long sKey[MODE == 2 ? 1 : RBMAX];
  uint32_t sAux[MODE == 2 ? RBMAX : 1];  // occlusion flag (shadow)
  uint8_t sRv[RBMAX];
  uint32_t sT[kBvhQCap];   // traversal items: BVH node << 7 | subtree node index
  uint32_t sPn[kPairCap];  // pair: BVH leaf << 6 | hierarchy leaf
  // Occlusion keeps each pair's primitive range (first position, count) in LDS: its leaf
  // tests are the bulk of the work and an extra dependent load each costs more than the
  // occupancy they take.  Closest hits read the range from the leaf node, which keeps the
  // kernel within six work-groups per CU.
  uint32_t sPt[MODE == 2 ? kPairCap : 1];
  uint32_t sPc[MODE == 2 ? kPairCap : 1];
  PairState<MODE> sPair[kPairCap];  // phase C pair state
  int sSeg[kThreads / 64];  // phase C: flagged pairs of each wavefront's quarter of the list
  int sTq, sPq;
  int sTraced, sOccluded;
};

template <class P>
__device__ __forceinline__ void bvh_box(P bvh, uint32_t n, V3& bmin, V3& bmax, uint32_t& off,
                                        uint32_t& np) {
  const float4 a = ld4(bvh, 2 * (size_t)n), b = ld4(bvh, 2 * (size_t)n + 1);
  bmin = {a.x, a.y, a.z};
  bmax = {a.w, b.x, b.y};
  off = __float_as_uint(b.z);
  np = __float_as_uint(b.w);
}


template <int NODE>
__device__ __forceinline__ bool bvh_bound(const Node& n, V3 bmin, V3 bmax) {
  if (NODE == 1) return cull_ia(n, bmin, bmax);
  if (NODE == 2) return bound_5db(n, bmin, bmax);
  return prefilter_node6(n, bmin, bmax);
}
template <int NODE>
__device__ __forceinline__ bool bvh_cull(const Node& n, V3 bmin, V3 bmax) {
  if (NODE == 1) return cull_ia(n, bmin, bmax);
  if (NODE == 2) return cull_5db(n, bmin, bmax);
  return cull_node6(n, bmin, bmax);
}
// The 5DB BVH kernels skip a node with no primitives and primitive offset 0 with its subtree
// (intersection5DB_BVH.cl:18,56 and variants); the 6DB ones do not test it.
template <int NODE>
__device__ __forceinline__ bool bvh_live(uint32_t off, uint32_t np) {
  return NODE != 2 || np != 0 || off != 0;
}

// Sub-level of subtree node index i (0 root, then 4, 16, 64 nodes).
__device__ __forceinline__ int sub_level(int i) { return i == 0 ? 0 : i < 5 ? 1 : i < 21 ? 2 : 3; }

// The exact lockstep chain of the pair (BVH leaf b, hierarchy leaf li): step s visits the
// hierarchy node of level max(top - s, 0) above li and the BVH node of depth min(s, d) above b.
template <int NODE, class SH>
__device__ __forceinline__ bool bvh_chain(const RHGeom& g, const SH& sh, const RhItem& it,
                                          uint32_t b, int li) {
  const int top = g.L - 1;
  const int lx = li & (g.nlx - 1), ly = li >> top;
  const int d = (int)glb(it.bvh_depth)[b];
  uint32_t node = b;
  int dn = d;
  for (int st = max(top, d); st >= 0; --st) {
    const int bd = min(st, d);
    while (dn > bd) {
      node = glb(it.bvh_parent)[node];
      --dn;
    }
    const int ss = min(st, top);  // subtree sub-level of the hierarchy node
    const int sh_ = top - ss;
    const int idx = pow4_off(ss) + (ly >> sh_) * (1 << ss) + (lx >> sh_);
    V3 bmin, bmax;
    uint32_t off, np;
    bvh_box(glb(it.bvh), node, bmin, bmax, off, np);
    if (!bvh_live<NODE>(off, np)) return false;
    if (!sh.sV[idx] || !bvh_cull<NODE>(sh.sN[idx], bmin, bmax)) return false;
  }
  return true;
}

// First primitive position and number of primitives that pair e tests: its BVH leaf's, or
// only the first one for 6DB / IA secondary rays (intersection6DB_BVH2.cl:88; the 5DB kernel
// tests them all, intersection5DB_BVH2.cl:84).
template <int MODE, int NODE, class SH>
__device__ __forceinline__ void pair_prims(const SH& sh, const RhItem& it, int e, uint32_t& off,
                                           uint32_t& cnt) {
  if constexpr (MODE == 2) {
    off = sh.sPt[e];
    cnt = sh.sPc[e];
  } else {
    const float4 b = ld4(glb(it.bvh), 2 * (size_t)(sh.sPn[e] >> 6) + 1);
    off = __float_as_uint(b.z);
    cnt = MODE == 1 && NODE != 2 ? 1u : __float_as_uint(b.w);
  }
}

// Leaf tests of the queued pairs (test, exact chain of pairs with a passing test, apply).
template <int MODE, int NODE, int RBMAX, int LMAX>
__device__ __forceinline__ void bvh_leaf_rays(const RHGeom& g, const RhItem& it,
                                              BvhShared<MODE, RBMAX, LMAX>& sh) {
  constexpr int TM = MODE == 0 ? 1 : MODE;  // ray-bvh has no repair pass
  const int n = sh.sPq;
  const int G = g.g;
  const int total = n * G;
  const bool pow2 = (G & (G - 1)) == 0;
  const int lg = __ffs(G) - 1;
  const bool masked = MODE == 2 && G <= 31;
  for (int e = threadIdx.x; e < n; e += kThreads) sh.sPair[e] = 0;
  __syncthreads();
  for (int item = threadIdx.x; item < total; item += kThreads) {
    const int e = pow2 ? item >> lg : item / G;
    const int r = item - e * G;
    const int slot = (int)(sh.sPn[e] & 63u) * G + r;
    if (!sh.sRv[slot]) continue;  // cleared once the ray is found occluded
    uint32_t off, cnt;
    pair_prims<MODE, NODE>(sh, it, e, off, cnt);
    for (uint32_t f = 0; f < cnt; ++f) {
      float t, b1, b2;
      if (leaf_hit<TM>(sh, slot, load_mt(glb(it.tmt), glb(it.prim_order)[off + f]), t, b1,
                       b2)) {
        if constexpr (MODE == 2) {
          if (masked) atomicOr(&sh.sPair[e], 1u << r);
          else sh.sPair[e] = 1;
        } else {
          sh.sPair[e] = 1;
        }
        break;
      }
    }
  }
  __syncthreads();
  // Compaction of the flagged pairs, per wavefront and in place, as in leaf_rays.
  const int lane = __lane_id();
  const int wv = threadIdx.x >> 6;
  const int S = ((n + kThreads - 1) / kThreads) * 64;
  const int s0 = wv * S, s1 = min(n, s0 + S);
  int cw = 0;
  for (int base = s0; base < s1; base += 64) {
    const int e = base + lane;
    PairState<MODE> ps = 0;
    uint32_t pn = 0, pt = 0, pc = 0;
    if (e < s1) {
      ps = sh.sPair[e];
      pn = sh.sPn[e];
      if constexpr (MODE == 2) {
        pt = sh.sPt[e];
        pc = sh.sPc[e];
      }
    }
    const uint64_t bal = __ballot(ps != 0);
    if (ps != 0) {
      const int d = s0 + cw + __popcll(bal & ((1ull << lane) - 1ull));
      sh.sPair[d] = ps;
      sh.sPn[d] = pn;
      if constexpr (MODE == 2) {
        sh.sPt[d] = pt;
        sh.sPc[d] = pc;
      }
    }
    cw += __popcll(bal);
  }
  if (lane == 0) sh.sSeg[wv] = cw;
  int verified = 0;
  for (int i = lane; i < cw; i += 64) {
    const int e = s0 + i;
    const PairState<MODE> ps = sh.sPair[e];
    const bool ok = bvh_chain<NODE>(g, sh, it, sh.sPn[e] >> 6, (int)(sh.sPn[e] & 63u));
    sh.sPair[e] = ok ? (PairState<MODE>)(ps | kPairOk<MODE>) : (PairState<MODE>)0;
    verified |= ok;
  }
  if (__syncthreads_or(verified)) {
    const int p1 = sh.sSeg[0], p2 = p1 + sh.sSeg[1], p3 = p2 + sh.sSeg[2], nf = p3 + sh.sSeg[3];
    const int ftotal = nf * G;
    for (int item = threadIdx.x; item < ftotal; item += kThreads) {
      const int j = pow2 ? item >> lg : item / G;
      const int r = item - j * G;
      const int w = (j >= p1) + (j >= p2) + (j >= p3);
      const int e = w * S + j - (w == 0 ? 0 : w == 1 ? p1 : w == 2 ? p2 : p3);
      const PairState<MODE> ps = sh.sPair[e];
      if (!(ps & kPairOk<MODE>)) continue;
      const int slot = (int)(sh.sPn[e] & 63u) * G + r;
      if (masked) {  // occlusion: the rays that passed in C1
        if (!((ps >> r) & 1u)) continue;
        if (atomicOr(&sh.sAux[slot], 1u) == 0u) atomicAdd(&sh.sOccluded, 1);
        sh.sRv[slot] = 0;
        continue;
      }
      if (!sh.sRv[slot]) continue;
      uint32_t off, cnt;
      pair_prims<MODE, NODE>(sh, it, e, off, cnt);
      for (uint32_t f = 0; f < cnt; ++f) {
        const uint32_t tri = glb(it.prim_order)[off + f];
        float t, b1, b2;
        if (!leaf_hit<TM>(sh, slot, load_mt(glb(it.tmt), tri), t, b1, b2)) continue;
        if (MODE == 2) {
          if (atomicOr(&sh.sAux[slot], 1u) == 0u) atomicAdd(&sh.sOccluded, 1);
          sh.sRv[slot] = 0;
          break;
        }
        atomicMin(&sh.sKey[slot], (unsigned long long)hit_key(t, (int32_t)tri));
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) sh.sPq = 0;
  __syncthreads();
}

// Appends the (BVH leaf, hierarchy leaf) pair of a reached leaf x leaf pair; its primitives
// are tested together (only the first one for secondary rays, intersection6DB_BVH2.cl).
template <int MODE, class SH>
__device__ __forceinline__ void bvh_emit(SH& sh, uint32_t b, uint32_t off, uint32_t np, int li,
                                         uint32_t* flags) {
  const int e = atomicAdd(&sh.sPq, 1);
  if (e < kPairCap) {  // room for every emission of one round is kept free
    sh.sPn[e] = (b << 6) | (uint32_t)li;
    if constexpr (MODE == 2) {
      sh.sPt[e] = off;
      sh.sPc[e] = np;
    }
  } else {
    atomicOr(flags, RT_FLAG_PAIR_CAP);  // reported by rt_ctx_check
  }
}

template <int MODE, int NODE, int RBMAX, int LMAX>
__global__ __launch_bounds__(256) void bvh_unit_kernel(const RhItem* __restrict__ items,
                                                       const uint4* __restrict__ units,
                                                       uint32_t* __restrict__ ctr) {
  __shared__ BvhShared<MODE, RBMAX, LMAX> sh;
  __shared__ uint32_t sUnit;
  const int tid = threadIdx.x;
  const uint32_t nunits = __builtin_amdgcn_readfirstlane(ctr[0]);
  if (nunits == 0) return;
  UnitClaim claim{blockIdx.x % kClaimQueues, 0u};
  for (;;) {
    if (tid == 0) sUnit = claim_unit(ctr, nunits, claim);
    __syncthreads();
    const uint32_t u = __builtin_amdgcn_readfirstlane(sUnit);
    if (u >= nunits) break;
    const uint4 unit = units[u];
    const RhItem& it = items[__builtin_amdgcn_readfirstlane(unit.x)];
    const RHGeom& g = it.g;
    const RaysDev& r = it.r;
    const int top = g.L - 1;
    const int b = (int)__builtin_amdgcn_readfirstlane(unit.y);
    const int K = b % g.gw[top], J = b / g.gw[top];
    if (tid == 0) {
      sh.sTq = 0;
      sh.sPq = 0;
      sh.sTraced = 0;
      sh.sOccluded = 0;
    }
    load_subtree(g, K, J, glb(it.nodes), sh.sN, sh.sV);
    __syncthreads();
    for (int slot = tid; slot < RBMAX; slot += kThreads) {
      int x = 0, y = 0, leaf = 0;
      bool traced = false;
      if (slot_xy(g, K, J, slot, x, y, leaf)) {
        const size_t idx = (size_t)y * g.W + x;
        traced = MODE == 0 || r.valid == nullptr || glb(r.valid)[idx] != 0;
        if (MODE == 2 && traced && glb(it.occ)[idx]) traced = false;
        sh.sRo[slot] = make_float4(glb(r.ox)[idx], glb(r.oy)[idx], glb(r.oz)[idx],
                                   glb(r.tmin)[idx]);
        sh.sRd[slot] = make_float4(glb(r.dx)[idx], glb(r.dy)[idx], glb(r.dz)[idx],
                                   glb(r.tmax)[idx]);
      }
      sh.sRv[slot] = traced ? 1 : 0;
      if (MODE != 2) sh.sKey[slot] = kMissKey;
      if (MODE == 2) sh.sAux[slot] = 0u;
      const uint64_t bal = __ballot(traced);
      if (__lane_id() == 0 && bal) atomicAdd(&sh.sTraced, __popcll(bal));
    }
    __syncthreads();
    if (sh.sV[0] && sh.sTraced > 0) {
      // The unit's clusters of frontier roots passed the cluster bound against the block
      // root.  Four clusters (256 roots) at a time: the roots that pass their own bound seed
      // the traversal stack, which is drained before the next four.  Pairs are tested when
      // their queue could not take another round's emissions, and at the end.
      const auto cl = glb(it.clist) + (size_t)b * it.nclusters + unit.z;
      const int ncl = (int)unit.w;
      bool stop = false;
      for (int k0 = 0; k0 < ncl && !stop; k0 += kThreads / kCluster) {
        if (sh.sPq > kPairCap - kThreads) {
          bvh_leaf_rays<MODE, NODE, RBMAX, LMAX>(g, it, sh);
          if (MODE == 2 && sh.sOccluded >= sh.sTraced) break;
        }
        const int k = k0 + tid / kCluster;
        const uint32_t ri = k < ncl ? cl[k] * kCluster + (uint32_t)(tid % kCluster) : it.nroots;
        if (ri < it.nroots) {
          const uint32_t node = glb(it.roots)[ri];
          V3 bmin, bmax;
          uint32_t off, np;
          bvh_box(glb(it.bvh), node, bmin, bmax, off, np);
          if (bvh_live<NODE>(off, np) && bvh_bound<NODE>(sh.sN[0], bmin, bmax)) {
            if (top == 0 && np > 0) {
              bvh_emit<MODE>(sh, node, off, np, 0, it.flags);
            } else {
              const int slot = atomicAdd(&sh.sTq, 1);
              sh.sT[slot] = node << 7;  // at most 256 seeds on an empty stack
            }
          }
        }
        __syncthreads();
        for (;;) {
          const int q = min(sh.sTq, kBvhQCap);
          if (sh.sPq > kPairCap - kThreads) {
            bvh_leaf_rays<MODE, NODE, RBMAX, LMAX>(g, it, sh);
            if (MODE == 2 && sh.sOccluded >= sh.sTraced) {
              stop = true;
              break;
            }
          }
          if (q == 0) break;
          // pop the newest items (bounded stack), room for 8 children each
          int npop = min(kBvhPop, q);
          while (npop > 0 && q - npop + 8 * npop > kBvhQCap) --npop;
          if (npop == 0) {  // cannot happen for subtrees of depth <= 12 (checked on the host)
            if (tid == 0) atomicOr(it.flags, RT_FLAG_BVH_STACK);  // reported by rt_ctx_check
            break;
          }
          uint32_t item = 0;
          const int pi = tid >> 3, c = tid & 7;
          if (pi < npop) item = sh.sT[q - npop + pi];
          __syncthreads();
          if (tid == 0) sh.sTq = q - npop;
          __syncthreads();
          bool push = false;
          uint32_t child = 0;
          if (pi < npop) {
            const uint32_t B = item >> 7;
            const int R = (int)(item & 127u);
            const int sl = sub_level(R);
            V3 bmin, bmax;
            uint32_t off, np;
            bvh_box(glb(it.bvh), B, bmin, bmax, off, np);
            const bool bleaf = np > 0, rleaf = sl == top;
            int rc = -1;
            uint32_t bc = B;
            if (!rleaf && !bleaf) {
              rc = c >> 1;
              bc = off + (uint32_t)(c & 1);
            } else if (!rleaf) {
              rc = c < 4 ? c : -2;
            } else if (!bleaf) {
              rc = c < 2 ? -1 : -2;
              bc = off + (uint32_t)c;
            } else {
              rc = -2;  // leaf x leaf items are emitted, never queued
            }
            if (rc != -2) {
              int Rn = R;
              if (rc >= 0) {
                const int pl = R - pow4_off(sl), pnl = 1 << sl;
                const int dk = pl & (pnl - 1), dj = pl >> sl;
                Rn = pow4_off(sl + 1) + (2 * dj + (rc >> 1)) * (2 * pnl) + 2 * dk + (rc & 1);
              }
              if (sh.sV[Rn]) {
                V3 cmin, cmax;
                uint32_t coff, cnp;
                bvh_box(glb(it.bvh), bc, cmin, cmax, coff, cnp);
                if (bvh_live<NODE>(coff, cnp) && bvh_bound<NODE>(sh.sN[Rn], cmin, cmax)) {
                  if (sub_level(Rn) == top && cnp > 0)
                    bvh_emit<MODE>(sh, bc, coff, cnp, Rn - pow4_off(top), it.flags);
                  else {
                    push = true;
                    child = (bc << 7) | (uint32_t)Rn;
                  }
                }
              }
            }
          }
          ballot_append(sh.sT, &sh.sTq, push, child);
          __syncthreads();
        }
        if (tid == 0) sh.sTq = 0;  // drained, or abandoned once every ray is occluded
        __syncthreads();
      }
      if (sh.sPq > 0 && !(MODE == 2 && sh.sOccluded >= sh.sTraced))
        bvh_leaf_rays<MODE, NODE, RBMAX, LMAX>(g, it, sh);
    }
    __syncthreads();
    for (int slot = tid; slot < g.rb; slot += kThreads) {
      int x, y, leaf;
      if (!slot_xy(g, K, J, slot, x, y, leaf)) continue;
      const size_t idx = (size_t)y * g.W + x;
      if (MODE == 2) {
        if (sh.sAux[slot]) glb(it.occ)[idx] = 1;
      } else {
        const unsigned long long key = sh.sKey[slot];
        if (key != kMissKey) atomicMin((unsigned long long*)&it.keys[idx], key);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------
// Per-ray path (blocks the cluster pass queued as incoherent).  A ray's result is the argmin
// over the triangles that pass its strict leaf test AND the exact cull chain of its leaf (for
// ray-bvh: that are tested at a (BVH leaf, hierarchy leaf) pair whose lockstep chain passes).
// Every such triangle passes the strict leaf test, so each ray walks the scene's search BVH
// (padded, conservative boxes), runs the strict leaf test on the triangles of the leaves it
// reaches, and evaluates the exact chain only for those that pass it and could change the
// ray.  The (t, id) key, the repair-region key and the occlusion flag are the block path's,
// so the outcome is the same; every ray of a block is handled by exactly one lane, which
// writes its results directly.
constexpr uint32_t kSearchLeafBit = 0x80000000u;
constexpr float kSlack = 0x1p-16f;  // relative slack of the slab-test parameters

// Slab test of the ray against a padded box: passes when [tn, tf] (tf widened by kSlack)
// meets [tlo, thi]; tn is the entry parameter used to order the children.
__device__ __forceinline__ bool slab(V3 o, V3 inv, V3 lo, V3 hi, float tlo, float thi, float& tn) {
  const float ax = (lo.x - o.x) * inv.x, bx = (hi.x - o.x) * inv.x;
  const float ay = (lo.y - o.y) * inv.y, by = (hi.y - o.y) * inv.y;
  const float az = (lo.z - o.z) * inv.z, bz = (hi.z - o.z) * inv.z;
  tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
  float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
  tf = tf + fabsf(tf) * kSlack;
  return tn <= tf && tn <= thi && tf >= tlo;
}

// Hierarchy node of level l above image position (x, y), from the job's node array.  The
// job (and so g) may differ between the lanes of a wavefront: the level's table entries are
// read with the lane's own index, two loads instead of a select over every level.
template <class P>
__device__ __forceinline__ bool chain_node(const RHGeom& g, P nodes, int x, int y, int l, Node& n) {
  const int gw = g.gw[l], off = g.off[l];
  const int cxl = g.cx << l, cyl = g.cy << l;
  const P p = nodes + 4 * ((size_t)off + (size_t)(y / cyl) * gw + (x / cxl));
  const float4 a = ld4(p, 0), b = ld4(p, 1), c = ld4(p, 2), d = ld4(p, 3);
  n.dmin = {a.x, a.y, a.z};
  n.dmax = {b.x, b.y, b.z};
  n.omin = {c.x, c.y, c.z};
  n.omax = {d.x, d.y, d.z};
  return a.w != 0.f;
}

// rayhierarchy: the exact cull chain of the ray's leaf for triangle `id`
// (intersection6DB.cl:51-119 and variants: every level from the block root to the leaf).
template <int MODE, int NODE>
__device__ __forceinline__ bool ray_chain_rh(const RhItem& it, int x, int y, uint32_t id) {
  const Tri T = load_tri(glb(it.v9), id);
  for (int l = it.g.L - 1; l >= 0; --l) {
    Node n;
    if (!chain_node(it.g, glb(it.nodes), x, y, l, n)) return false;
    if (!cull_exact<MODE, NODE>(n, T)) return false;
  }
  return true;
}

// ray-bvh: the exact lockstep chain of (BVH leaf b, the ray's hierarchy leaf): step s visits
// the hierarchy node of level max(top - s, 0) and the BVH node of depth min(s, d) above b.
template <int NODE>
__device__ __forceinline__ bool ray_chain_bvh(const RhItem& it, int x, int y, uint32_t b) {
  const int top = it.g.L - 1;
  const int d = (int)glb(it.bvh_depth)[b];
  uint32_t node = b;
  int dn = d;
  for (int st = max(top, d); st >= 0; --st) {
    const int bd = min(st, d);
    while (dn > bd) {
      node = glb(it.bvh_parent)[node];
      --dn;
    }
    Node n;
    if (!chain_node(it.g, glb(it.nodes), x, y, top - min(st, top), n)) return false;
    V3 bmin, bmax;
    uint32_t off, np;
    bvh_box(glb(it.bvh), node, bmin, bmax, off, np);
    if (!bvh_live<NODE>(off, np) || !bvh_cull<NODE>(n, bmin, bmax)) return false;
  }
  return true;
}

// Per-ray kernel: one wavefront per work-group, every lane one ray at a time.  The queue holds
// 64-ray chunks; a wavefront takes a chunk and deals its rays to lanes without one, and when a
// lane's ray is done it takes the next ray of the chunk (or of the next chunk) at the next
// round, so lanes of long traversals do not leave the others idle.  A round walks interior
// nodes until every lane holds a leaf or is done (Aila and Laine's while-while), then tests
// the postponed leaves.
//
// Occlusion defers the cull chains: a candidate (strict hit inside the ray's interval) goes to
// a wavefront queue in LDS as {job, ray, triangle or BVH leaf, lane and ray serial}, and the
// queue is evaluated 64 candidates at a time, one per lane, once it holds kChainFlush.  A
// passing chain sets the ray's occlusion flag itself and ends the traversal of its lane if the
// lane still holds that ray.  The ray is occluded when any of its candidates passes its chain,
// whatever the order, so the result is the same; the lanes no longer run a chain one or two at
// a time, and a lane does not wait for its candidates before it takes the next ray.  The queue
// is also evaluated at least every kChainAge rounds (a lane takes at most one ray per round,
// so the 10-bit serial cannot wrap while an entry waits) and at the end.
constexpr uint32_t kChainFlush = 32;
constexpr uint32_t kChainQ = kChainFlush + 64;  // below kChainFlush, plus one per lane a round
constexpr uint32_t kChainAge = 512;
// Closest hits defer their chains the same way.  A candidate (a strict hit that beats the
// lane's bound) is queued with its t, and the lane goes on pruning with it as a provisional
// bound.  A passing chain merges the candidate into the ray's global key (atomic min, so the
// order does not matter); the ray's lane waits at its end until its candidates are evaluated.
// When a candidate that bounded the search fails its chain, the provisional bound may have
// pruned the answer, so the lane traces the ray again from the root, evaluating each chain at
// once (chains pass for most hits, so this is rare).  The queue (four words an entry) is four
// entries shorter than occlusion's so that five work-groups per SIMD still fit the LDS.
constexpr uint32_t kChainFlushC = 28;
constexpr uint32_t kChainQC = kChainFlushC + 64;
// Five wavefronts per SIMD (at most 96 VGPRs; the LDS of 20 work-groups per CU): the walk is
// latency-bound and gains more from resident wavefronts than it loses to a few spilled
// registers of the chain evaluation.
template <int MODE, int NODE, bool BVH>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8))) void ray_path_kernel(
    const RhItem* __restrict__ items, const uint2* __restrict__ rq, uint32_t* __restrict__ ctr,
    uint32_t nchunks, uint32_t walk) {
  __shared__ uint32_t sStk[kSearchDepth][64];
  constexpr uint32_t kQ = MODE == 2 ? kChainQ : kChainQC;
  __shared__ uint32_t sQIdx[kQ], sQRef[kQ], sQMeta[kQ];
  __shared__ float sQT[MODE == 2 ? 1 : kQ];
  __shared__ unsigned long long sQHit;  // occlusion: rays found occluded; closest: failed bounds
  __shared__ uint32_t sQN;
  constexpr uint32_t kDone = 0xffffffffu;  // not a node reference (leaves hold < 16 triangles)
  // Rayhierarchy occlusion pauses a lane that queued a candidate: its chain usually passes,
  // which ends the ray, so walking on is mostly wasted (7% faster on a batch of 64 scenes).
  // ray-bvh occlusion walks on (pausing made a 3-bounce frame's passes 1.4% slower).
  constexpr bool kPause = MODE == 2 && !BVH;
  const int lane = threadIdx.x;
  const uint64_t lanes_below = (1ull << lane) - 1ull;
  // wavefront-uniform: the chunk being dealt and its next slot
  uint32_t cjob = 0, cblk = 0, cpart = 0, cslot = 64;
  bool more = true;
  // lane state
  bool act = false;
  uint32_t jk = 0;
  int x = 0, y = 0;
  size_t idx = 0;
  V3 o = {0.f, 0.f, 0.f}, d = {0.f, 0.f, 1.f}, inv = {1.f, 1.f, 1.f};
  float tmin = 0.f, tmax = 0.f, tlo = 0.f;
  uint64_t best = kMissKey;
  uint32_t rep = kNoKey32;
  bool occluded = false, track_rep = false;
  // whether the lane has a candidate queued (rayhierarchy occlusion: the lane pauses until it
  // is evaluated); closest: the provisional bound of queued candidates, and whether the ray is
  // traced again with its chains evaluated at once
  uint64_t spec = kMissKey;
  bool pend = false, nospec = false;
  int sp = 0;
  uint32_t cur = kDone, pl = kDone;
  auto pop = [&]() { return sp > 0 ? sStk[--sp][lane] : kDone; };
  // occlusion: deferred chain candidates (count sQN), the serial of this lane's ray, and the
  // rounds since the queue was last evaluated (wavefront-uniform)
  uint32_t serial = 0, age = 0;
  // the cluster pass sets ctr[4] when it routed any block here
  if (__builtin_amdgcn_readfirstlane(ctr[4]) == 0u) return;
  UnitClaim claim{blockIdx.x % kClaimQueues, 0u};
  if (lane == 0) {
    sQN = 0;
    sQHit = 0ull;
  }
  __syncthreads();
  // Evaluates the cq queued candidates, 64 per pass (one per lane).  Wavefront-uniform, after
  // a barrier that made the entries visible.
  auto run_chains = [&](uint32_t cq) {
    for (uint32_t base = 0; base < cq; base += 64) {
      const uint32_t e = base + (uint32_t)lane;
      const bool have = e < cq;
      const uint32_t meta = have ? sQMeta[e] : 0u;
      const uint32_t owner = (meta >> 10) & 63u;
      const uint32_t owner_serial = (uint32_t)__shfl((int)serial, (int)owner);
      if (have) {
        const RhItem& oit = items[meta >> 16];
        const uint32_t ri = sQIdx[e];
        const int ox = (int)(ri % (uint32_t)oit.g.W), oy = (int)(ri / (uint32_t)oit.g.W);
        if constexpr (MODE == 2) {
          bool ok;
          if constexpr (BVH) ok = ray_chain_bvh<NODE>(oit, ox, oy, sQRef[e]);
          else ok = ray_chain_rh<MODE, NODE>(oit, ox, oy, sQRef[e]);
          if (ok) {
            glb(oit.occ)[ri] = 1;
            if ((meta & 1023u) == owner_serial) atomicOr(&sQHit, 1ull << owner);
          }
        } else {
          // closest: sQRef holds the triangle id, meta bit 1 a main candidate, bit 0 a
          // repair-region one
          const uint32_t id = sQRef[e];
          const float t = sQT[e];
          bool ok;
          if constexpr (BVH) ok = ray_chain_bvh<NODE>(oit, ox, oy, glb(oit.tleaf)[id] & 0x7fffffffu);
          else ok = ray_chain_rh<MODE, NODE>(oit, ox, oy, id);
          if (ok) {
            if (meta & 2u) atomicMin((unsigned long long*)&oit.keys[ri], hit_key(t, (int32_t)id));
            if (meta & 1u) atomicMin(&oit.repkeys[ri], fkey(t));
          } else if (meta & 2u) {
            atomicOr(&sQHit, 1ull << owner);
          }
        }
      }
    }
    __syncthreads();
    if constexpr (MODE == 2) {
      if (act && ((sQHit >> lane) & 1ull)) occluded = true;
      pend = false;  // every queued candidate has been evaluated
    } else if (act && pend) {
      pend = false;
      if ((sQHit >> lane) & 1ull) {  // a bound failed: trace the ray again
        nospec = true;
        spec = kMissKey;
        sp = 0;
        cur = 0u;
        pl = kDone;
      }
    }
    __syncthreads();
    if (lane == 0) {
      sQN = 0;
      sQHit = 0ull;
    }
    __syncthreads();
    age = 0;
  };
  for (;;) {
    // deal rays to the lanes without one
    for (;;) {
      const uint64_t need = __ballot(!act);
      if (need == 0 || !more) break;
      if (cslot >= 64) {
        uint32_t e = 0;
        if (lane == 0) e = claim_unit(ctr, nchunks, claim, 1);
        e = __builtin_amdgcn_readfirstlane(e);
        if (e >= nchunks) {
          more = false;
          break;
        }
        const uint2 w = rq[e];
        if (__builtin_amdgcn_readfirstlane(w.x) == 0xffffffffu) continue;  // a block-path block
        cjob = __builtin_amdgcn_readfirstlane(w.x);
        const uint32_t c = __builtin_amdgcn_readfirstlane(w.y);
        const uint32_t nch = ((uint32_t)items[cjob].g.rb + 63u) / 64u;
        cblk = c / nch;
        cpart = c - cblk * nch;
        cslot = 0;
      }
      const uint32_t avail = 64u - cslot;
      const uint32_t rank = (uint32_t)__popcll(need & lanes_below);
      if (!act && rank < avail) {
        const RhItem& it = items[cjob];
        const RHGeom& g = it.g;
        const int top = g.L - 1;
        const int K = (int)cblk % g.gw[top], J = (int)cblk / g.gw[top];
        int leaf = 0;
        bool have = slot_xy(g, K, J, (int)(cpart * 64u + cslot + rank), x, y, leaf) && it.ntri > 0;
        if (have) {
          idx = (size_t)y * g.W + x;
          const RaysDev& r = it.r;
          if (MODE != 0 && r.valid != nullptr && glb(r.valid)[idx] == 0) have = false;
        }
        if (have) {
          const RaysDev& r = it.r;
          jk = cjob;
          o = {glb(r.ox)[idx], glb(r.oy)[idx], glb(r.oz)[idx]};
          d = {glb(r.dx)[idx], glb(r.dy)[idx], glb(r.dz)[idx]};
          tmin = glb(r.tmin)[idx];
          tmax = glb(r.tmax)[idx];
          // inverse direction for the slab test; zero components become tiny ones, so the
          // slab of that axis is (almost) the whole line when the origin is inside it and
          // empty otherwise
          auto safe = [](float v) {
            return fabsf(v) < 0x1p-80f ? (v < 0.f ? -0x1p-80f : 0x1p-80f) : v;
          };
          inv = {1.0f / safe(d.x), 1.0f / safe(d.y), 1.0f / safe(d.z)};
          tlo = tmin - fabsf(tmin) * kSlack;
          best = kMissKey;
          rep = kNoKey32;
          occluded = false;
          spec = kMissKey;
          pend = false;
          nospec = false;
          track_rep = MODE == 0 && !BVH && it.repkeys != nullptr;
          sp = 0;
          cur = 0u;  // interior node 0 (the root)
          pl = kDone;
          act = true;
          serial = (serial + 1u) & 1023u;
        }
      }
      cslot += min(avail, (uint32_t)__popcll(need));
    }
    if (!__any(act)) {
      const uint32_t cq = __builtin_amdgcn_readfirstlane(sQN);  // the candidates still queued
      if (cq > 0u) run_chains(cq);
      break;
    }
    const RhItem& it = items[jk];
    const auto snode = glb(it.snode);
    const auto tmts = glb(it.tmts);
    const auto perm = glb(it.sperm);
    // node phase
    for (;;) {
      if (act && cur != kDone && !(cur & kSearchLeafBit) && !(kPause && pend)) {
        const float4 n0 = ld4(snode, 4 * (size_t)cur), n1 = ld4(snode, 4 * (size_t)cur + 1),
                     n2 = ld4(snode, 4 * (size_t)cur + 2), n3 = ld4(snode, 4 * (size_t)cur + 3);
        const uint64_t bound = spec < best ? spec : best;
        float thi = MODE == 2 ? tmax : (bound == kMissKey ? __int_as_float(0x7f800000)
                                                          : fkey_inv((uint32_t)(bound >> 32)));
        thi = thi + fabsf(thi) * kSlack;
        float tn0, tn1;
        const bool h0 = slab(o, inv, V3{n0.x, n0.y, n0.z}, V3{n0.w, n1.x, n1.y}, tlo, thi, tn0);
        const bool h1 = slab(o, inv, V3{n1.z, n1.w, n2.x}, V3{n2.y, n2.z, n2.w}, tlo, thi, tn1);
        const uint32_t r0 = __float_as_uint(n3.x), r1 = __float_as_uint(n3.y);
        if (h0 && h1) {
          const bool first1 = tn1 < tn0;
          sStk[sp++][lane] = first1 ? r0 : r1;
          cur = first1 ? r1 : r0;
        } else if (h0 || h1) {
          cur = h0 ? r0 : r1;
        } else {
          cur = pop();
        }
        if (pl == kDone && cur != kDone && (cur & kSearchLeafBit)) {
          pl = cur;
          cur = pop();
        }
      }
      // walk on while more than 1/2^walk of the lanes holding a ray have neither a postponed
      // leaf nor finished (fewer lanes idle behind the longest walk); the others go on from
      // their node in the next round (every round walks at least one step, so each lane
      // advances)
      const uint32_t searching = (uint32_t)__popcll(__ballot(
          act && pl == kDone && cur != kDone && !(cur & kSearchLeafBit) && !(kPause && pend)));
      if (searching <= ((uint32_t)__popcll(__ballot(act)) >> walk)) break;
    }
    // leaf phase
    if constexpr (MODE == 2) {
      // wavefront-uniform rounds over the leaf's (at most kSearchLeaf) triangles, so the queue
      // can be evaluated between them
      const bool haveleaf = act && pl != kDone && !(kPause && pend);
      const uint32_t first = pl & ((1u << 27) - 1u);
      const uint32_t cnt = haveleaf ? ((pl >> 27) & 15u) + 1u : 0u;
      for (uint32_t f = 0; f < (uint32_t)kSearchLeaf; ++f) {
        if (!__any(f < cnt)) break;
        if (f < cnt) {
          const uint32_t slot = first + f;
          uint32_t ref = perm[slot];
          bool ok = true;
          if constexpr (BVH) {
            const uint32_t bl = glb(it.tleaf)[ref];
            ok = bl != 0xffffffffu;
            ref = bl & 0x7fffffffu;
          }
          float t, b1, b2;
          // intersectAllLeaves.cl:177-179; the chain waits in the queue
          if (ok && mt_strict(o, d, load_mt(tmts, slot), t, b1, b2) && t > tmin && t < tmax) {
            const uint32_t e = atomicAdd(&sQN, 1u);  // below kChainQ, see kChainQ
            sQIdx[e] = (uint32_t)idx;
            sQRef[e] = ref;
            sQMeta[e] = (jk << 16) | ((uint32_t)lane << 10) | serial;
            pend = kPause;
          }
        }
        __syncthreads();
        const uint32_t cq = __builtin_amdgcn_readfirstlane(sQN);
        if (cq >= kChainFlush) run_chains(cq);
      }
      if (haveleaf) pl = kDone;
    } else {
      // wavefront-uniform rounds over the leaf's triangles, as for occlusion
      const bool haveleaf = act && pl != kDone;
      const uint32_t first = pl & ((1u << 27) - 1u);
      const uint32_t cnt = haveleaf ? ((pl >> 27) & 15u) + 1u : 0u;
      for (uint32_t f = 0; f < (uint32_t)kSearchLeaf; ++f) {
        if (!__any(f < cnt)) break;
        if (f < cnt) {
          const uint32_t slot = first + f;
          const uint32_t id = perm[slot];
          bool ok = true;
          uint32_t bl = 0;
          if constexpr (BVH) {
            bl = glb(it.tleaf)[id];
            // first primitive only (6DB / IA secondary rays)
            ok = bl != 0xffffffffu && !(MODE == 1 && NODE != 2 && (bl & 0x80000000u));
            bl &= 0x7fffffffu;
          }
          float t, b1, b2;
          if (ok && mt_strict(o, d, load_mt(tmts, slot), t, b1, b2)) {
            const bool main = t > tmin && t < __int_as_float(0x7f800000);
            const bool rp = track_rep && repair_region(b1, b2) && t >= tmin;
            const uint64_t key = hit_key(t, (int32_t)id);
            const bool um = main && key < best && key < spec, ur = rp && fkey(t) < rep;
            if (um || ur) {
              if (nospec) {
                bool pass;
                if constexpr (BVH) pass = ray_chain_bvh<NODE>(it, x, y, bl);
                else pass = ray_chain_rh<MODE, NODE>(it, x, y, id);
                if (pass) {
                  if (um) best = key;
                  if (ur) rep = fkey(t);
                }
              } else {
                const uint32_t e = atomicAdd(&sQN, 1u);  // below kChainQC, see kChainQC
                sQIdx[e] = (uint32_t)idx;
                sQRef[e] = id;
                sQT[e] = t;
                sQMeta[e] = (jk << 16) | ((uint32_t)lane << 10) | (um ? 2u : 0u) | (ur ? 1u : 0u);
                if (um) spec = key;
                pend = true;
              }
            }
          }
        }
        __syncthreads();
        const uint32_t cq = __builtin_amdgcn_readfirstlane(sQN);
        if (cq >= kChainFlushC) run_chains(cq);
      }
      if (haveleaf) pl = kDone;
    }
    if (act && pl == kDone && cur != kDone && (cur & kSearchLeafBit)) {  // a second leaf waits
      pl = cur;
      cur = pop();
    }
    {
      // evaluate the queue every kChainAge rounds, and every round once no chunk is left to
      // deal (idle lanes cannot take new rays, and a ray found occluded or bounded early ends
      // sooner); also once a quarter of the lanes holding a ray wait: closest-hit lanes at the
      // ray's end for their candidates, paused occlusion lanes as soon as they queue one (if
      // the chain fails, the lane walks on from where it paused)
      bool flush = ++age >= kChainAge || !more;
      const uint32_t nact = (uint32_t)__popcll(__ballot(act));
      const uint32_t nwait = (uint32_t)__popcll(
          __ballot(act && pend && (kPause || (cur == kDone && pl == kDone))));
      flush = flush || (nwait > 0u && 4u * nwait >= nact);
      if (flush) {
        const uint32_t cq = __builtin_amdgcn_readfirstlane(sQN);
        if (cq > 0u) run_chains(cq);
        age = 0;
      }
    }
    if (act && occluded) cur = pl = kDone;
    if (act && cur == kDone && pl == kDone && !pend) {  // the ray is done: its result, then a new ray
      if constexpr (MODE == 2) {
        if (occluded) glb(it.occ)[idx] = 1;
      } else {
        // passing queued candidates were merged by the evaluation; chains evaluated at once
        // (a ray traced again) leave their result here
        if (best != kMissKey) atomicMin((unsigned long long*)&it.keys[idx], best);
        if (track_rep && rep != kNoKey32) atomicMin(&it.repkeys[idx], rep);
      }
      act = false;
    }
  }
}

// Triangle staged in LDS as v1 and the two edges (the leaf test's operands).
__device__ __forceinline__ void stage_tri(float4 (*sT)[3], int slot, const Tri& T) {
  sT[slot][0] = make_float4(T.v1.x, T.v1.y, T.v1.z, T.e1.x);
  sT[slot][1] = make_float4(T.e1.y, T.e1.z, T.e2.x, T.e2.y);
  sT[slot][2] = make_float4(T.e2.z, 0.f, 0.f, 0.f);
}
__device__ __forceinline__ Tri staged_tri(const float4 (*sT)[3], int slot) {
  float4 a = sT[slot][0], b = sT[slot][1], c = sT[slot][2];
  Tri T;
  T.v1 = {a.x, a.y, a.z};
  T.e1 = {a.w, b.x, b.y};
  T.e2 = {b.z, b.w, c.x};
  T.v2 = T.v3 = T.v1;  // unused by the leaf test
  return T;
}

// ---------------------------------------------------------------------------------------
// Naive brute force (naive.cl "TriRay" semantics, ray-parallel): every ray of the block tests
// every triangle of its chunk, triangles staged through LDS.
template <int MODE>
__global__ __launch_bounds__(256) void naive_kernel(RaysDev r, const float* __restrict__ v9,
                                                    uint32_t ntri, uint32_t chunk,
                                                    uint64_t* __restrict__ keys,
                                                    uint8_t* __restrict__ occ_out, int tchunks) {
  __shared__ float4 sT[kThreads][3];
  const int tid = threadIdx.x;
  const size_t n = (size_t)r.W * r.H;
  const size_t i = (size_t)blockIdx.x * kThreads + tid;
  const bool exists = i < n;
  const bool valid = exists && (MODE == 0 || r.valid == nullptr || r.valid[i] != 0);
  V3 o{}, d{};
  float tmin = 0.f, tmax = 0.f;
  if (valid) {
    o = {r.ox[i], r.oy[i], r.oz[i]};
    d = {r.dx[i], r.dy[i], r.dz[i]};
    tmin = r.tmin[i];
    tmax = r.tmax[i];
  }
  uint64_t best = kMissKey;
  bool occ = false;
  const uint32_t t0 = blockIdx.y * chunk;
  const uint32_t t1 = min(ntri, t0 + chunk);
  for (uint32_t base = t0; base < t1; base += kThreads) {
    const uint32_t ti = base + tid;
    if (ti < t1) stage_tri(sT, tid, load_tri(v9, ti));
    __syncthreads();
    const int cnt = (int)min((uint32_t)kThreads, t1 - base);
    if (valid && !(MODE == 2 && occ)) {
      for (int k = 0; k < cnt; ++k) {
        const Tri T = staged_tri(sT, k);
        float t, b1, b2;
        if (MODE == 2) {
          if (!mt_naive_p(o, d, T, t)) continue;
          if (t < tmin || t > tmax) continue;
          occ = true;
          break;
        } else {
          if (!mt_naive(o, d, T, t, b1, b2)) continue;
          if (t < tmin) continue;
          uint64_t kk = hit_key(t, (int32_t)(base + k));
          if (kk < best) best = kk;
        }
      }
    }
    if (MODE == 2) {
      if (__syncthreads_and(!valid || occ)) break;
    } else {
      __syncthreads();
    }
  }
  if (!exists) return;
  if (MODE == 2) {
    if (tchunks == 1) occ_out[i] = occ ? 1 : 0;
    else if (occ) occ_out[i] = 1;
  } else {
    if (tchunks == 1) keys[i] = best;
    else if (best != kMissKey) atomicMin((unsigned long long*)&keys[i], (unsigned long long)best);
  }
}

// ---------------------------------------------------------------------------------------
// Hit record: id, t, (u, v) and optionally dpdu/dpdv of computeDpTuTv.cl:8-91.
// Hit record of ray i (computeDpTuTv.cl:8-91).  naive_rule: the naive kernels' inline copy of
// that code takes the coordinate-system fallback for |det| < EPS (naive.cl:60-61,153-154)
// where computeDpTuTv.cl:57 takes it for det == 0 only.
__device__ void write_hit(const RaysDev& r, const float* __restrict__ v9,
                          const float* __restrict__ uv6, size_t i, int32_t id, float t,
                          HitOut out, bool naive_rule = false) {
  if (id < 0) {
    out.id[i] = -1;
    out.t[i] = __int_as_float(0x7f800000);
    out.u[i] = 0.f;
    out.v[i] = 0.f;
    if (out.dpdu)
      for (int c = 0; c < 3; ++c) out.dpdu[3 * i + c] = 0.f;
    if (out.dpdv)
      for (int c = 0; c < 3; ++c) out.dpdv[3 * i + c] = 0.f;
    return;
  }
  const Tri T = load_tri(v9, (uint32_t)id);
  const V3 o = {r.ox[i], r.oy[i], r.oz[i]};
  const V3 d = {r.dx[i], r.dy[i], r.dz[i]};
  float b1, b2;
  bary(o, d, T, b1, b2);
  float uv[6] = {0.f, 0.f, 1.f, 0.f, 1.f, 1.f};
  if (uv6)
    for (int k = 0; k < 6; ++k) uv[k] = uv6[6 * (size_t)id + k];
  const float b0 = 1 - b1 - b2;
  out.id[i] = id;
  out.t[i] = t;
  out.u[i] = (b0 * uv[0] + b1 * uv[2]) + b2 * uv[4];
  out.v[i] = (b0 * uv[1] + b1 * uv[3]) + b2 * uv[5];
  if (out.dpdu || out.dpdv) {
    const float du1 = uv[0] - uv[4], du2 = uv[2] - uv[4];
    const float dv1 = uv[1] - uv[5], dv2 = uv[3] - uv[5];
    const V3 dp1 = sub(T.v1, T.v3), dp2 = sub(T.v2, T.v3);
    const float det = du1 * dv2 - dv1 * du2;
    V3 pu, pv;
    if (naive_rule ? (det > -kEps && det < kEps) : det == 0.f) {
      const V3 tn = normalize(cross(T.e2, T.e1));
      if (fabsf(tn.x) > fabsf(tn.y)) {
        const float il = 1.0f / sqrtf(tn.x * tn.x + tn.z * tn.z);
        pu = {-tn.z * il, 0.f, tn.x * il};
      } else {
        const float il = 1.0f / sqrtf(tn.y * tn.y + tn.z * tn.z);
        pu = {0.f, tn.z * il, -tn.y * il};
      }
      pv = cross(tn, pu);
    } else {
      const float id2 = 1.f / det;
      pu = {(dv2 * dp1.x - dv1 * dp2.x) * id2, (dv2 * dp1.y - dv1 * dp2.y) * id2,
            (dv2 * dp1.z - dv1 * dp2.z) * id2};
      pv = {(-du2 * dp1.x + du1 * dp2.x) * id2, (-du2 * dp1.y + du1 * dp2.y) * id2,
            (-du2 * dp1.z + du1 * dp2.z) * id2};
    }
    if (out.dpdu) {
      out.dpdu[3 * i] = pu.x; out.dpdu[3 * i + 1] = pu.y; out.dpdu[3 * i + 2] = pu.z;
    }
    if (out.dpdv) {
      out.dpdv[3 * i] = pv.x; out.dpdv[3 * i + 1] = pv.y; out.dpdv[3 * i + 2] = pv.z;
    }
  }
}

// Hit record of one ray from its reduction key; returns t (INFINITY on a miss).
__device__ __forceinline__ float finalize_ray(const RaysDev& r, const float* __restrict__ v9,
                                              const float* __restrict__ uv6, uint64_t key,
                                              size_t i, const HitOut& out,
                                              bool naive_rule = false) {
  if (key == kMissKey) {
    write_hit(r, v9, uv6, i, -1, 0.f, out);
    return __int_as_float(0x7f800000);
  }
  const float t = fkey_inv((uint32_t)(key >> 32));
  write_hit(r, v9, uv6, i, (int32_t)(uint32_t)key, t, out, naive_rule);
  return t;
}

// Hit records of the naive accelerator (and of an empty scene's misses).
__global__ __launch_bounds__(256) void finalize_kernel(RaysDev r, const float* __restrict__ v9,
                                                       const float* __restrict__ uv6,
                                                       const uint64_t* __restrict__ keys,
                                                       HitOut out, int naive_rule) {
  const size_t n = (size_t)r.W * r.H;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  finalize_ray(r, v9, uv6, keys[i], i, out, naive_rule != 0);
}

// Hit records of every job of a closest-hit group.  Rays whose nearest repair-region
// candidate lies at or below tHit - EPS are listed for the repair pass (only those can
// change, see rh_repair_kernel).
__global__ __launch_bounds__(256) void rh_finalize_kernel(const RhItem* __restrict__ items,
                                                          int nitems, uint32_t nrays,
                                                          uint32_t* __restrict__ ctr,
                                                          uint32_t* __restrict__ replist) {
  const uint32_t gi = blockIdx.x * kThreads + threadIdx.x;
  if (gi >= nrays) return;
  const RhItem& it = items[find_ray_item(items, nitems, gi)];
  const uint32_t i = gi - it.ray_base;
  const float tstar = finalize_ray(it.r, it.v9, it.uv6, it.keys[i], i, it.out);
  if (it.repkeys) {
    const uint32_t rk = it.repkeys[i];
    if (rk != kNoKey32 && !(fkey_inv(rk) > tstar - kEps)) replist[atomicAdd(&ctr[2], 1u)] = gi;
  }
}

// Exact sequential repair pass (yetAnotherIntersection6DB.cl with repairRun passes, in
// increasing triangle id order) for the rays whose outcome it can change: the nearest
// repair-region candidate lies at or below tHit - EPS (hits at exactly tmin, or ties at t
// large enough that t - EPS rounds to t).  One work-group per flagged ray.
template <int NODE>
__global__ __launch_bounds__(256) void rh_repair_kernel(const RhItem* __restrict__ items,
                                                        int nitems,
                                                        const uint32_t* __restrict__ ctr,
                                                        const uint32_t* __restrict__ replist) {
  __shared__ Node sChain[kMaxLevels];
  __shared__ int sWc[kThreads / 64];
  __shared__ float sEt[kThreads];
  __shared__ uint32_t sEi[kThreads];
  __shared__ float sTH;
  __shared__ int sIdx;
  const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  const uint32_t count = ctr[2];
  for (uint32_t w = blockIdx.x; w < count; w += gridDim.x) {
    const uint32_t gi = __builtin_amdgcn_readfirstlane(replist[w]);
    const RhItem& it = items[find_ray_item(items, nitems, gi)];
    const RHGeom& g = it.g;
    const RaysDev& r = it.r;
    const float* __restrict__ v9 = it.v9;
    const uint32_t ntri = it.ntri;
    const size_t i = gi - it.ray_base;
    const int x = (int)(i % (size_t)g.W), y = (int)(i / (size_t)g.W);
    if (tid < g.L) {
      const int l = tid;
      const int gk = x / (g.cx << l), gj = y / (g.cy << l);
      const float4* p = it.nodes + 4 * ((size_t)g.off[l] + (size_t)gj * g.gw[l] + gk);
      float4 a = ld4(p, 0), b = ld4(p, 1), c = ld4(p, 2), d = ld4(p, 3);
      sChain[l].dmin = {a.x, a.y, a.z};
      sChain[l].dmax = {b.x, b.y, b.z};
      sChain[l].omin = {c.x, c.y, c.z};
      sChain[l].omax = {d.x, d.y, d.z};
    }
    if (tid == 0) {
      sTH = it.out.t[i];
      sIdx = it.out.id[i];
    }
    __syncthreads();
    const V3 o = {r.ox[i], r.oy[i], r.oz[i]};
    const V3 d = {r.dx[i], r.dy[i], r.dz[i]};
    const float tmin = r.tmin[i];
    for (int run = 0; run < it.repair_run; ++run) {
      for (uint32_t base = 0; base < ntri; base += kThreads) {
        const uint32_t ti = base + tid;
        bool have = false;
        float t = 0.f;
        if (ti < ntri) {
          const Tri T = load_tri(v9, ti);
          bool pass = true;
          if (NODE != 0) {
            const V3 bmin = vmin(vmin(T.v1, T.v2), T.v3), bmax = vmax(vmax(T.v1, T.v2), T.v3);
            for (int l = g.L - 1; l >= 0 && pass; --l)
              pass = NODE == 1 ? cull_ia(sChain[l], bmin, bmax) : cull_5db(sChain[l], bmin, bmax);
          } else {
            for (int l = g.L - 1; l >= 0 && pass; --l) pass = cull_tri(sChain[l], T);
          }
          float b1, b2;
          if (pass && mt_strict(o, d, T, t, b1, b2) && repair_region(b1, b2) && t >= tmin)
            have = true;
        }
        const uint64_t bal = __ballot(have);
        const int prefix = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) sWc[wave] = __popcll(bal);
        __syncthreads();
        int woff = 0, total = 0;
        for (int k = 0; k < kThreads / 64; ++k) {
          if (k < wave) woff += sWc[k];
          total += sWc[k];
        }
        if (have) {
          sEt[woff + prefix] = t;
          sEi[woff + prefix] = ti;
        }
        __syncthreads();
        if (tid == 0) {
          for (int e = 0; e < total; ++e) {
            if (sEt[e] > sTH - kEps || (int)sEi[e] == sIdx) continue;
            sTH = sEt[e];
            sIdx = (int)sEi[e];
          }
        }
        __syncthreads();
      }
    }
    if (tid == 0) write_hit(r, v9, it.uv6, i, sIdx, sTH, it.out);
    __syncthreads();
  }
}

// Shadow rays toward a point light (light.h:79-84, point.cpp:42-49).
// When `count` is given, the number of valid shadow rays is added to it.
__global__ __launch_bounds__(256) void shadow_kernel(RaysDev pr, const int32_t* __restrict__ id,
                                                     const float* __restrict__ tt, float lx,
                                                     float ly, float lz, float* ox, float* oy,
                                                     float* oz, float* dx, float* dy, float* dz,
                                                     float* tmin, float* tmax, uint8_t* valid,
                                                     uint32_t* count) {
  const size_t n = (size_t)pr.W * pr.H;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (count) {
    const uint64_t bal = __ballot(i < n && id[i] >= 0);
    if (__lane_id() == 0 && bal) atomicAdd(count, (uint32_t)__popcll(bal));
  }
  if (i >= n) return;
  if (id[i] < 0) {
    ox[i] = oy[i] = oz[i] = 0.f;
    dx[i] = dy[i] = dz[i] = 0.f;
    tmin[i] = tmax[i] = 0.f;
    valid[i] = 0;
    return;
  }
  const float t = tt[i];
  const float px = pr.ox[i] + pr.dx[i] * t;
  const float py = pr.oy[i] + pr.dy[i] * t;
  const float pz = pr.oz[i] + pr.dz[i] * t;
  const float vx = lx - px, vy = ly - py, vz = lz - pz;
  const float dist = sqrtf((vx * vx + vy * vy) + vz * vz);
  const float inv = 1.f / dist;
  ox[i] = px;
  oy[i] = py;
  oz[i] = pz;
  dx[i] = vx * inv;
  dy[i] = vy * inv;
  dz[i] = vz * inv;
  tmin[i] = 1e-3f * t;
  tmax[i] = dist * (1.f - 0.f);
  valid[i] = 1;
}

// Pinhole camera rays of a render request: image (W*spp) x H, the spp samples of a pixel
// adjacent in x (the reference sampler's order); pixel centres for spp == 1, otherwise a
// counter-based hash of (seed, x, y, sample) gives the jitter.  The FP32 operation order is
// that of the workloads' numpy camera, so spp == 1 images are identical to it.
__device__ __forceinline__ uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}
__global__ __launch_bounds__(256) void camera_kernel(int W, int H, int spp, float cx, float cy,
                                                     float cz, float tan_half, float aspect,
                                                     uint32_t seed, float* ox, float* oy,
                                                     float* oz, float* dx, float* dy, float* dz,
                                                     float* tmin, float* tmax) {
  const size_t n = (size_t)W * spp * H;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int row = W * spp;
  const int y = (int)(i / (size_t)row), c = (int)(i % (size_t)row);
  const int x = c / spp, k = c % spp;
  float jx = 0.5f, jy = 0.5f;
  if (spp > 1) {
    const uint32_t h = mix32(seed ^ mix32((uint32_t)x * 0x9e3779b9u ^ mix32((uint32_t)y ^
                                                                          ((uint32_t)k << 20))));
    jx = (float)(h >> 8) * (1.0f / 16777216.0f);
    jy = (float)(mix32(h) >> 8) * (1.0f / 16777216.0f);
  }
  const float px = ((((float)x + jx) / (float)W) * 2.0f - 1.0f) * tan_half * aspect;
  const float py = (1.0f - (((float)y + jy) / (float)H) * 2.0f) * tan_half;
  const float pz = 1.0f;
  const float l = sqrtf((px * px + py * py) + pz * pz);
  const float inv = 1.0f / l;
  ox[i] = cx;
  oy[i] = cy;
  oz[i] = cz;
  dx[i] = px * inv;
  dy[i] = py * inv;
  dz[i] = pz * inv;
  tmin[i] = 0.f;
  tmax[i] = __int_as_float(0x7f800000);
}

// Path-tracing bounce rays from closest hits: origin p = o + d*t, a cosine-weighted direction
// about the geometric normal of the hit triangle turned toward the incoming ray, tmin =
// 1e-3*t (pbrt's rayEpsilon), tmax = INFINITY; random numbers hashed from (seed, ray, depth)
// so the rays do not depend on the launch order.
__global__ __launch_bounds__(256) void bounce_kernel(RaysDev in, const float* __restrict__ v9,
                                                     const int32_t* __restrict__ id,
                                                     const float* __restrict__ tt, uint32_t seed,
                                                     int depth, float* ox, float* oy, float* oz,
                                                     float* dx, float* dy, float* dz,
                                                     float* tmin, float* tmax, uint8_t* valid) {
  const size_t n = (size_t)in.W * in.H;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t tri = id[i];
  if (tri < 0 || (in.valid != nullptr && in.valid[i] == 0)) {
    ox[i] = oy[i] = oz[i] = 0.f;
    dx[i] = dy[i] = dz[i] = 0.f;
    tmin[i] = tmax[i] = 0.f;
    valid[i] = 0;
    return;
  }
  const float t = tt[i];
  const V3 d = {in.dx[i], in.dy[i], in.dz[i]};
  const V3 p = {in.ox[i] + d.x * t, in.oy[i] + d.y * t, in.oz[i] + d.z * t};
  const Tri T = load_tri(v9, (uint32_t)tri);
  V3 nn = normalize(cross(T.e1, T.e2));
  if (dot(nn, d) > 0.f) nn = {-nn.x, -nn.y, -nn.z};
  const uint32_t h = mix32(seed ^ mix32((uint32_t)i * 0x9e3779b9u ^ mix32((uint32_t)depth)));
  const float u1 = (float)(h >> 8) * (1.0f / 16777216.0f);
  const float u2 = (float)(mix32(h) >> 8) * (1.0f / 16777216.0f);
  const float r = sqrtf(u1), phi = 6.28318530717958647692f * u2;
  const float lx = r * cosf(phi), ly = r * sinf(phi), lz = sqrtf(fmaxf(0.f, 1.f - u1));
  // orthonormal basis about nn (Duff et al. 2017)
  const float sgn = nn.z >= 0.f ? 1.f : -1.f;
  const float a = -1.f / (sgn + nn.z);
  const float b = nn.x * nn.y * a;
  const V3 t1 = {1.f + sgn * nn.x * nn.x * a, sgn * b, -sgn * nn.x};
  const V3 t2 = {b, sgn + nn.y * nn.y * a, -nn.y};
  const V3 w = {t1.x * lx + t2.x * ly + nn.x * lz, t1.y * lx + t2.y * ly + nn.y * lz,
                t1.z * lx + t2.z * ly + nn.z * lz};
  ox[i] = p.x;
  oy[i] = p.y;
  oz[i] = p.z;
  dx[i] = w.x;
  dy[i] = w.y;
  dz[i] = w.z;
  tmin[i] = 1e-3f * t;
  tmax[i] = __int_as_float(0x7f800000);
  valid[i] = 1;
}

template <class T>
__global__ void fill_kernel(T* p, T v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

inline dim3 grid1(size_t n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

// Work-groups of a persistent kernel that are resident at once on the current device: CUs x
// work-groups per CU (limited by registers and LDS) for `threads`-wide work-groups, at most
// `cap`.  A larger grid only adds work-groups that start when the queue is nearly drained.
// The count is kept per (device, kernel) once known (2048 until then).
template <class F>
unsigned resident_grid(F kernel, uint32_t cap, int threads = kThreads) {
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, unsigned> known;
  int dev = 0;
  unsigned n = 0;
  const bool have_dev = hipGetDevice(&dev) == hipSuccess;
  const std::pair<int, const void*> key{dev, reinterpret_cast<const void*>(kernel)};
  if (have_dev) {
    std::lock_guard<std::mutex> lk(mu);
    auto f = known.find(key);
    if (f != known.end()) n = f->second;
  }
  if (n == 0) {
    int cus = 0, per_cu = 0;
    if (have_dev &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) ==
            hipSuccess &&
        cus > 0 && per_cu > 0) {
      n = (unsigned)cus * (unsigned)per_cu;
      std::lock_guard<std::mutex> lk(mu);
      known[key] = n;
    } else {
      n = 2048;
    }
  }
  return std::max(1u, std::min(n, std::max(cap, 1u)));
}

// Launch sequence of one rayhierarchy group with the mode and node type fixed.
template <int MODE, int NODE>
hipError_t rh_group_dispatch(const RhGroup& G, hipStream_t s, hipEvent_t t0, hipEvent_t tm,
                             hipEvent_t t1) {
  hipLaunchKernelGGL((rh_init_kernel<MODE>), grid1(std::max<uint32_t>(G.nrays, 4)),
                     dim3(kThreads), 0, s, G.items, G.nitems, G.nrays, G.ctr);
  hipLaunchKernelGGL((rh_group_build_kernel<MODE != 0, NODE == 2>), dim3(G.nblocks),
                     dim3(kThreads), 0, s, G.items, G.nitems);
  if (G.bvh)
    hipLaunchKernelGGL((rh_cluster_kernel<MODE, NODE, true>), dim3(G.nblocks), dim3(kThreads), 0,
                       s, G.items, G.nitems, G.ustage, G.bunits, G.rq, G.ctr, G.per_unit,
                       G.ray_clusters);
  else
    hipLaunchKernelGGL((rh_cluster_kernel<MODE, NODE, false>), dim3(G.nblocks), dim3(kThreads),
                       0, s, G.items, G.nitems, G.ustage, G.bunits, G.rq, G.ctr, G.per_unit,
                       G.ray_clusters);
  if (G.ray_clusters > 1) {  // 0 and 1 send every block with work per-ray: no unit
    hipLaunchKernelGGL(rh_unit_scan_kernel, dim3((G.nblocks + kScanThreads - 1) / kScanThreads),
                       dim3(kScanThreads), 0, s, G.bunits, G.nblocks);
    hipLaunchKernelGGL(rh_unit_pack_kernel, dim3((G.nblocks + kPackThreads / 64 - 1) /
                                                 (kPackThreads / 64)),
                       dim3(kPackThreads), 0, s, G.bunits, G.nblocks, G.ustage, G.units, G.ctr);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (t0) {
    e = hipEventRecord(t0, s);
    if (e != hipSuccess) return e;
  }
  auto launch = [&](auto kernel) {
    const dim3 grid(G.grid > 0 ? (unsigned)G.grid : resident_grid(kernel, G.max_units));
    hipLaunchKernelGGL(kernel, grid, dim3(kThreads), 0, s, G.items, G.units, G.ctr);
  };
  if (G.bvh && G.big)
    launch(bvh_unit_kernel<MODE, NODE, 1024, 4>);
  else if (G.bvh)
    launch(bvh_unit_kernel<MODE, NODE, 256, 3>);
  else if (G.big)
    launch(rh_unit_kernel<MODE, NODE, 1024, 4>);
  else
    launch(rh_unit_kernel<MODE, NODE, 256, 3>);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (tm) {
    e = hipEventRecord(tm, s);
    if (e != hipSuccess) return e;
  }
  auto launch_rays = [&](auto kernel) {
    const dim3 grid(G.ray_grid > 0 ? std::max(1u, std::min((unsigned)G.ray_grid, G.max_chunks))
                                   : resident_grid(kernel, G.max_chunks, 64));
    hipLaunchKernelGGL(kernel, grid, dim3(64), 0, s, G.items, G.rq, G.ctr, G.max_chunks,
                       (uint32_t)std::min(std::max(G.walk, 1), 7));
  };
  if (G.bvh)
    launch_rays(ray_path_kernel<MODE, NODE, true>);
  else
    launch_rays(ray_path_kernel<MODE, NODE, false>);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (t1) {
    e = hipEventRecord(t1, s);
    if (e != hipSuccess) return e;
  }
  if (MODE != 2) {
    hipLaunchKernelGGL(rh_finalize_kernel, grid1(G.nrays), dim3(kThreads), 0, s, G.items,
                       G.nitems, G.nrays, G.ctr, G.replist);
    if (MODE == 0 && G.replist && !G.bvh)
      hipLaunchKernelGGL((rh_repair_kernel<NODE>), dim3(64), dim3(kThreads), 0, s, G.items,
                         G.nitems, G.ctr, G.replist);
  }
  return hipGetLastError();
}

// Copies a job table passed by value (in the kernel arguments) into device memory, so that a
// launch group can be set up without a host-to-device copy.
__global__ __launch_bounds__(64) void write_table_kernel(RhItem item, RhItem* __restrict__ dst) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&item);
  uint32_t* out = reinterpret_cast<uint32_t*>(dst);
  for (uint32_t i = threadIdx.x; i < sizeof(RhItem) / 4; i += 64) out[i] = src[i];
}

// Scratch of the allocation-free trace (rt_trace_closest_noalloc): one array per device, held
// until the process ends.  Its device address is looked up in this file, once per device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

}  // namespace

// Device address of `scratch` on each device, looked up by noalloc_setup.
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return 1;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != device) return 1;  // the device must be current
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return 1;
  g_scratch_ptr[device] = p;
  return 0;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

hipError_t launch_write_table(const RhItem& item, RhItem* dst, hipStream_t s) {
  hipLaunchKernelGGL(write_table_kernel, dim3(1), dim3(64), 0, s, item, dst);
  return hipGetLastError();
}

int rh_geometry(int W, int H, int cx, int cy, int height, RHGeom* g) {
  if (W <= 0 || H <= 0 || cx <= 0 || cy <= 0 || height <= 0) return 1;
  if (cx > (1 << 24) || cy > (1 << 24)) return 3;
  int gw = (int)(((int64_t)W + cx - 1) / cx), gh = (int)(((int64_t)H + cy - 1) / cy);
  g->W = W;
  g->H = H;
  g->cx = cx;
  g->cy = cy;
  g->gw[0] = gw;
  g->gh[0] = gh;
  int L = 1;
  // The level count is the reference's: it halves with floor and stops after a level whose
  // floor width or height is odd (rayhierarchy.cpp:412-442).  The grids themselves halve with
  // ceil, so a parent also covers the last child column (row) that the reference's floor
  // leaves out where it halves an odd dimension.
  int fw = gw, fh = gh;
  for (int i = 1; i < height; ++i) {
    if (L >= kMaxLevels) return 2;
    gw = (gw + 1) / 2;
    gh = (gh + 1) / 2;
    fw /= 2;
    fh /= 2;
    g->gw[L] = gw;
    g->gh[L] = gh;
    ++L;
    // a zero floor dimension leaves the reference no top-level node (it traces nothing):
    // stop there too
    if ((fw & 1) || (fh & 1) || fw == 0 || fh == 0) break;
  }
  g->L = L;
  int off = 0;
  for (int l = 0; l < L; ++l) {
    g->off[l] = off;
    off += g->gw[l] * g->gh[l];
  }
  g->nnodes = off;
  g->nlx = 1 << (L - 1);
  if ((int64_t)cx * g->nlx * cy * g->nlx > kMaxRaysPerBlock) return 3;
  g->bw = cx * g->nlx;
  g->bh = cy * g->nlx;
  g->rb = g->bw * g->bh;
  g->g = cx * cy;
  g->nsub = ((1 << (2 * L)) - 1) / 3;
  g->ntop = g->gw[L - 1] * g->gh[L - 1];
  return 0;
}

hipError_t launch_fill_u64(uint64_t* p, uint64_t v, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel<uint64_t>, grid1(n), dim3(kThreads), 0, s, p, v, n);
  return hipGetLastError();
}
hipError_t launch_fill_u8(uint8_t* p, uint8_t v, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel<uint8_t>, grid1(n), dim3(kThreads), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_rh_group(const RhGroup& G, hipStream_t s, hipEvent_t t0, hipEvent_t tm,
                           hipEvent_t t1) {
  if (G.nitems <= 0 || G.nrays == 0) return hipSuccess;
#define RT_CALL(M, N) rh_group_dispatch<M, N>(G, s, t0, tm, t1)
  if (G.node == 1) {
    if (G.mode == 0) return RT_CALL(0, 1);
    if (G.mode == 1) return RT_CALL(1, 1);
    return RT_CALL(2, 1);
  }
  if (G.node == 2) {
    if (G.mode == 0) return RT_CALL(0, 2);
    if (G.mode == 1) return RT_CALL(1, 2);
    return RT_CALL(2, 2);
  }
  if (G.mode == 0) return RT_CALL(0, 0);
  if (G.mode == 1) return RT_CALL(1, 0);
  return RT_CALL(2, 0);
#undef RT_CALL
}

hipError_t launch_naive_trace(const RaysDev& r, int mode, const float* v9, uint32_t ntri,
                              int tchunks, uint64_t* keys, uint8_t* occ, hipStream_t s) {
  const size_t n = (size_t)r.W * r.H;
  const uint32_t chunk = (uint32_t)((ntri + tchunks - 1) / tchunks);
  dim3 grid((unsigned)((n + kThreads - 1) / kThreads), tchunks);
  if (mode == 0)
    hipLaunchKernelGGL(naive_kernel<0>, grid, dim3(kThreads), 0, s, r, v9, ntri, chunk, keys, occ, tchunks);
  else if (mode == 1)
    hipLaunchKernelGGL(naive_kernel<1>, grid, dim3(kThreads), 0, s, r, v9, ntri, chunk, keys, occ, tchunks);
  else
    hipLaunchKernelGGL(naive_kernel<2>, grid, dim3(kThreads), 0, s, r, v9, ntri, chunk, keys, occ, tchunks);
  return hipGetLastError();
}

hipError_t launch_finalize(const RaysDev& r, const float* v9, const float* uv6,
                           const uint64_t* keys, HitOut out, bool naive_rule, hipStream_t s) {
  const size_t n = (size_t)r.W * r.H;
  hipLaunchKernelGGL(finalize_kernel, grid1(n), dim3(kThreads), 0, s, r, v9, uv6, keys, out,
                     naive_rule ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_camera_rays(int W, int H, int spp, const float origin[3], float tan_half,
                              float aspect, uint32_t seed, float* ox, float* oy, float* oz,
                              float* dx, float* dy, float* dz, float* tmin, float* tmax,
                              hipStream_t s) {
  const size_t n = (size_t)W * spp * H;
  hipLaunchKernelGGL(camera_kernel, grid1(n), dim3(kThreads), 0, s, W, H, spp, origin[0],
                     origin[1], origin[2], tan_half, aspect, seed, ox, oy, oz, dx, dy, dz, tmin,
                     tmax);
  return hipGetLastError();
}

hipError_t launch_bounce_rays(const RaysDev& in, const float* v9, const int32_t* id,
                              const float* t, uint32_t seed, int depth, float* ox, float* oy,
                              float* oz, float* dx, float* dy, float* dz, float* tmin,
                              float* tmax, uint8_t* valid, hipStream_t s) {
  const size_t n = (size_t)in.W * in.H;
  hipLaunchKernelGGL(bounce_kernel, grid1(n), dim3(kThreads), 0, s, in, v9, id, t, seed, depth,
                     ox, oy, oz, dx, dy, dz, tmin, tmax, valid);
  return hipGetLastError();
}

hipError_t launch_shadow_rays(const RaysDev& pr, const int32_t* id, const float* t, float lx,
                              float ly, float lz, float* ox, float* oy, float* oz, float* dx,
                              float* dy, float* dz, float* tmin, float* tmax, uint8_t* valid,
                              hipStream_t s, uint32_t* count) {
  const size_t n = (size_t)pr.W * pr.H;
  hipLaunchKernelGGL(shadow_kernel, grid1(n), dim3(kThreads), 0, s, pr, id, t, lx, ly, lz, ox,
                     oy, oz, dx, dy, dz, tmin, tmax, valid, count);
  return hipGetLastError();
}

}  // namespace rti
