// rt_capi.cpp -- C ABI of the ray-space accelerators (include/pbrt_amd_rt.h).
//
// Owns device residency (scene triangles, UVs, BVH handoff, scratch) and sequences the
// kernels of rt_kernels.hip on the caller's stream.  No host synchronisation happens inside
// a trace call, so a caller can pipeline calls on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/pbrt_amd_rt.h"
#include "rt_internal.h"

using namespace rti;

// HIP-event timing of the traversal kernels: one event triple per launch while enabled
// (before the unit kernel, between it and the per-ray kernel, after the per-ray kernel).
struct Timer {
  struct Slot {
    int mode;
    bool split;  // m was recorded (rayhierarchy / ray-bvh launches)
    hipEvent_t a, m, b;
  };
  bool on = false;
  std::vector<Slot> slots;
  size_t used = 0;
  ~Timer() {
    for (auto& t : slots) {
      (void)hipEventDestroy(t.a);
      (void)hipEventDestroy(t.m);
      (void)hipEventDestroy(t.b);
    }
  }
  // The events for the next launch of `mode`, or null when timing is off.
  Slot* next(int mode, bool split) {
    if (!on) return nullptr;
    if (used == slots.size()) {
      Slot t{mode, split, nullptr, nullptr, nullptr};
      if (hipEventCreate(&t.a) != hipSuccess) return nullptr;
      if (hipEventCreate(&t.m) != hipSuccess) {
        (void)hipEventDestroy(t.a);
        return nullptr;
      }
      if (hipEventCreate(&t.b) != hipSuccess) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.m);
        return nullptr;
      }
      slots.push_back(t);
    }
    Slot& t = slots[used++];
    t.mode = mode;
    t.split = split;
    return &t;
  }
  // part 0: whole traversal, 1: the unit (block) kernel, 2: the per-ray kernel.
  int total(int mode, int part, float* ms_out, int32_t* n_out) {
    float sum = 0.f;
    int32_t n = 0;
    for (size_t i = 0; i < used; ++i) {
      const Slot& t = slots[i];
      if (t.mode != mode) continue;
      if (hipEventSynchronize(t.b) != hipSuccess) return RT_ERR_HIP;
      float ms = 0.f;
      hipEvent_t from = t.a, to = t.b;
      if (part == 1 && t.split) to = t.m;
      if (part == 2) {
        if (!t.split) continue;
        from = t.m;
      }
      if (hipEventElapsedTime(&ms, from, to) != hipSuccess) return RT_ERR_HIP;
      sum += ms;
      ++n;
    }
    *ms_out = sum;
    *n_out = n;
    return RT_OK;
  }
  int total(int mode, float* ms_out, int32_t* n_out) { return total(mode, 0, ms_out, n_out); }
};

// Scratch of rayhierarchy launch groups: one device arena (job tables, per-job hierarchy,
// cluster lists and keys, per-group work queue) and a pinned host copy of the job tables.
// The tables are uploaded only when they change, so repeated calls on the same buffers cost
// no host-to-device copy.
struct RhWorkspace {
  char* d_arena = nullptr;
  size_t arena_cap = 0;
  RhItem* h_table = nullptr;  // pinned
  size_t table_cap = 0;
  std::vector<RhItem> last;  // tables as last uploaded
  const char* last_arena = nullptr;
  hipEvent_t uploaded = nullptr;
  ~RhWorkspace() {
    if (uploaded) (void)hipEventSynchronize(uploaded);
    (void)hipFree(d_arena);
    (void)hipHostFree(h_table);
    if (uploaded) (void)hipEventDestroy(uploaded);
  }
};

// One rayhierarchy job of a plain call or a batch.
struct RhJob {
  rt_scene* s;
  RaysDev r;
  int mode;  // 0 primary, 1 secondary, 2 occluded
  HitOut out;
  uint8_t* occ;
};

// Launch-shape parameters: passing clusters per work unit and persistent grid, per mode
// (index 0 closest, 1 occluded); 0 = built-in choice.
struct RhParams {
  int per[2] = {0, 0};
  int grid[2] = {0, 0};
  int ray[2] = {0, 0};  // passing clusters from which a block takes the per-ray path
  int rgrid[2] = {0, 0};  // work-groups of the per-ray kernel
  int walk[2] = {0, 0};   // per-ray walk rounds (RT_PARAM_WALK_*)
  int tile[2] = {0, 0};   // 1: the 1024-ray LDS tile for every block (RT_PARAM_TILE_*)
};

struct rt_scene {
  rt_ctx* ctx = nullptr;
  rt_accel_desc desc{};
  uint32_t ntri = 0;
  float* d_v9 = nullptr;
  float* d_uv6 = nullptr;
  // traversal copy for rayhierarchy: Morton-sorted triangles in clusters of 64
  float* d_v9s = nullptr;
  uint32_t* d_perm = nullptr;
  uint32_t* d_sperm = nullptr;  // search-BVH position -> triangle id
  float4* d_cbox = nullptr;
  float4* d_tbox = nullptr;  // AABB of every sorted triangle
  float4* d_tmt = nullptr;   // ray-bvh: leaf-test operands v1, e1, e2 (3 float4 per triangle)
  // per-ray path: search BVH over the sorted triangles and their leaf-test operands
  float4* d_snode = nullptr;
  float4* d_tmts = nullptr;
  uint32_t* d_tleaf = nullptr;  // ray-bvh: reference BVH leaf of every original triangle
  uint32_t nclusters = 0;
  int clusters_closest = 0, clusters_occluded = 0;  // 0 = built-in choice
  int ray_closest = 0, ray_occluded = 0;            // per-ray path thresholds; 0 = built-in
  int walk_closest = 0, walk_occluded = 0;          // per-ray walk rounds; 0 = built-in
  int tile_closest = 0, tile_occluded = 0;          // unit-kernel LDS tile; 0 = by block size
  // ray-bvh handoff
  float4* d_bvh = nullptr;
  uint32_t* d_roots = nullptr;
  uint32_t* d_prim_order = nullptr;
  float4* d_rcbox = nullptr;          // boxes of the clusters of 64 Morton-ordered roots
  uint32_t nrclusters = 0;
  uint32_t* d_bvh_parent = nullptr;  // parent of every node below a frontier root
  uint8_t* d_bvh_depth = nullptr;    // its depth below that root
  uint32_t nroots = 0;
  int roots_per_unit = 0;  // root clusters per ray-bvh unit; 0 = built-in choice
  bool has_bvh = false;
  // scratch of the naive path, grown on demand
  uint64_t* d_keys = nullptr;
  size_t keys_cap = 0;
  // scratch of the rayhierarchy and ray-bvh paths (a launch group of one job), one per trace
  // mode, so that a renderer alternating camera, shadow and bounce calls on the same buffers
  // finds its job table already on the device
  RhWorkspace ws[3];
  Timer timer;
  std::mutex mu;
};

struct rt_batch {
  rt_ctx* ctx = nullptr;
  std::vector<RhJob> jobs;
  RhWorkspace ws;
  RhParams prm;
  bool use_graph = false;
  // captured launch sequence, valid while the jobs, parameters and arena are unchanged
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;
  bool graph_valid = false;
  // side streams (default priority) and events that fork a captured sequence's groups into
  // parallel branches: forks[0] is recorded on the caller's stream, forks[k + 1] on side[k]
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> forks;
  Timer timer;
  std::mutex mu;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    graph_valid = false;
  }
  ~rt_batch() {
    drop_graph();
    for (hipStream_t s : side) (void)hipStreamDestroy(s);
    for (hipEvent_t e : forks) (void)hipEventDestroy(e);
  }
};

namespace {

int hip_status(hipError_t e) {
  if (e == hipSuccess) return RT_OK;
  if (e == hipErrorOutOfMemory) return RT_ERR_OOM;
  return RT_ERR_HIP;
}

#define RT_TRY(expr)                      \
  do {                                    \
    hipError_t _e = (expr);               \
    if (_e != hipSuccess) return hip_status(_e); \
  } while (0)

template <class T>
int grow(T*& p, size_t& cap, size_t need) {
  if (need <= cap && p) return RT_OK;
  if (p) {
    hipError_t e = hipFree(p);
    p = nullptr;
    cap = 0;
    if (e != hipSuccess) return hip_status(e);
  }
  size_t n = std::max<size_t>(need, 1);
  hipError_t e = hipMalloc((void**)&p, n * sizeof(T));
  if (e != hipSuccess) {
    p = nullptr;
    return hip_status(e);
  }
  cap = n;
  return RT_OK;
}

int check_rays(const rt_rays* r, bool need_valid_ok) {
  (void)need_valid_ok;
  if (!r || !r->ox || !r->oy || !r->oz || !r->dx || !r->dy || !r->dz || !r->tmin || !r->tmax)
    return RT_ERR_INVALID;
  if (r->width <= 0 || r->height <= 0) return RT_ERR_INVALID;
  if ((int64_t)r->width * r->height > (int64_t)0x7fffffff) return RT_ERR_INVALID;
  return RT_OK;
}

RaysDev to_dev(const rt_rays* r) {
  return RaysDev{r->ox, r->oy, r->oz, r->dx, r->dy, r->dz, r->tmin, r->tmax, r->valid,
                 r->width, r->height};
}

// Triangle chunks per block so that a launch has >= ~1024 work-groups (4 per CU).
int pick_tchunks(int nblocks, uint32_t nwork, uint32_t per_tile) {
  int want = (2048 + nblocks - 1) / std::max(nblocks, 1);  // ~8 work-groups per CU
  int maxc = (int)std::max<uint32_t>(1, nwork / std::max<uint32_t>(per_tile, 1));
  return std::max(1, std::min(want, maxc));
}

int desc_ok(const rt_accel_desc* d) {
  if (d->kind < RT_ACCEL_NAIVE || d->kind > RT_ACCEL_RAYBVH) return RT_ERR_INVALID;
  if (d->kind == RT_ACCEL_NAIVE) return RT_OK;
  if (d->node != RT_NODE_BOX_DIR && d->node != RT_NODE_IA && d->node != RT_NODE_BOX_UV)
    return RT_ERR_INVALID;
  if (d->chunk_x <= 0 || d->chunk_y <= 0 || d->height <= 0) return RT_ERR_INVALID;
  if (d->repair_run < 0) return RT_ERR_INVALID;
  if (d->kind == RT_ACCEL_RAYBVH) {
    if (d->bvh_height < 0 || d->bvh_height > 24) return RT_ERR_UNSUPPORTED;
    if (d->max_bvh_prim <= 0 || d->max_bvh_prim > 4) return RT_ERR_UNSUPPORTED;
  }
  return RT_OK;
}

// 30-bit Morton code of point p quantised to 10 bits per axis over [lo, hi].
uint32_t morton3(const float* p, const float* lo, const float* hi) {
  auto spread = [](uint32_t x) {
    x &= 1023u;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
  };
  uint32_t q[3];
  for (int a = 0; a < 3; ++a) {
    const float ext = hi[a] - lo[a];
    float f = ext > 0.f ? (p[a] - lo[a]) / ext : 0.f;
    f = std::min(std::max(f, 0.f), 1.f);
    q[a] = (uint32_t)(f * 1023.f);
  }
  return (spread(q[0]) << 2) | (spread(q[1]) << 1) | spread(q[2]);
}

// Morton order of triangle centroids (10 bits per axis), ties by triangle id, cut into clusters
// of kClusterSize with their vertex AABBs.  Only the traversal order changes: hits carry the
// original ids through perm.
void sort_clusters(const float* v9, uint32_t ntri, std::vector<float>& v9s,
                   std::vector<uint32_t>& perm, std::vector<float4>& cbox) {
  std::vector<float> c(3 * (size_t)ntri);
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < ntri; ++i)
    for (int a = 0; a < 3; ++a) {
      const float* p = v9 + 9 * (size_t)i;
      const float v = (p[a] + p[3 + a] + p[6 + a]) * (1.0f / 3.0f);
      c[3 * (size_t)i + a] = v;
      lo[a] = std::min(lo[a], v);
      hi[a] = std::max(hi[a], v);
    }
  std::vector<uint32_t> code(ntri);
  for (uint32_t i = 0; i < ntri; ++i) code[i] = morton3(c.data() + 3 * (size_t)i, lo, hi);
  perm.resize(ntri);
  for (uint32_t i = 0; i < ntri; ++i) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(),
                   [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
  v9s.resize(9 * (size_t)ntri);
  for (uint32_t i = 0; i < ntri; ++i)
    std::memcpy(&v9s[9 * (size_t)i], v9 + 9 * (size_t)perm[i], 9 * sizeof(float));
  const uint32_t nc = (ntri + kClusterSize - 1) / kClusterSize;
  cbox.assign(2 * (size_t)nc, make_float4(0.f, 0.f, 0.f, 0.f));
  for (uint32_t k = 0; k < nc; ++k) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = k * kClusterSize; i < std::min(ntri, (k + 1) * kClusterSize); ++i)
      for (int v = 0; v < 3; ++v)
        for (int a = 0; a < 3; ++a) {
          const float x = v9s[9 * (size_t)i + 3 * v + a];
          mn[a] = std::min(mn[a], x);
          mx[a] = std::max(mx[a], x);
        }
    cbox[2 * (size_t)k] = make_float4(mn[0], mn[1], mn[2], 0.f);
    cbox[2 * (size_t)k + 1] = make_float4(mx[0], mx[1], mx[2], 0.f);
  }
}

// Leaf-test operands v1, e1 = v2 - v1, e2 = v3 - v1 (FP32, as the device's make_tri) of the
// triangles of v9, 3 float4 each.
std::vector<float4> mt_operands(const float* v9, uint32_t ntri) {
  std::vector<float4> m(3 * (size_t)ntri);
  for (uint32_t i = 0; i < ntri; ++i) {
    const float* p = v9 + 9 * (size_t)i;
    const float e1[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]};
    const float e2[3] = {p[6] - p[0], p[7] - p[1], p[8] - p[2]};
    m[3 * (size_t)i] = make_float4(p[0], p[1], p[2], e1[0]);
    m[3 * (size_t)i + 1] = make_float4(e1[1], e1[2], e2[0], e2[1]);
    m[3 * (size_t)i + 2] = make_float4(e2[2], 0.f, 0.f, 0.f);
  }
  return m;
}

template <class T>
hipError_t upload(T*& d, const T* h, size_t n) {
  hipError_t e = hipMalloc((void**)&d, sizeof(T) * std::max<size_t>(n, 1));
  if (e == hipSuccess && n) e = hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice);
  return e;
}

// Launches `fn` bracketed by the timer's next event pair for `mode` (if timing is on).
template <class F>
hipError_t timed_launch(Timer& tm, int mode, hipStream_t st, F fn) {
  Timer::Slot* t = tm.next(mode, false);
  if (t) {
    hipError_t e = hipEventRecord(t->a, st);
    if (e != hipSuccess) return e;
  }
  hipError_t e = fn();
  if (e == hipSuccess && t) e = hipEventRecord(t->b, st);
  return e;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr int kDefaultPer[2] = {8, 4};  // passing clusters per unit: closest, occluded
// Blocks whose root passes at least this many clusters (64 triangles or 64 frontier roots
// each) in a scene of four times as many, or whose leaves' ray directions spread widely, are
// traced ray by ray.  2^29 keeps only the direction test.  Built-in choice per accelerator and
// mode (closest, occluded), measured on the torus, room, uneven-batch and serving workloads:
// coherent camera blocks are fastest on the block path; shadow blocks that pass many clusters
// of a large scene are faster ray by ray, more so behind ray-bvh's lockstep block path.
constexpr int kDefaultRay[2][2] = {{1 << 29, 256}, {1 << 29, 64}};
int default_ray(bool bvh, int mi) { return kDefaultRay[bvh ? 1 : 0][mi]; }
constexpr int kDefaultRootClustersPerUnit = 8;  // passing root clusters per ray-bvh unit

// The job-table entry of one job: geometry of its ray image and the scene's resident arrays
// (scratch pointers and prefix sums are filled by the caller).
int fill_item(const RhJob& jb, RhItem& it) {
  const bool bvh = jb.s->desc.kind == RT_ACCEL_RAYBVH;
  std::memset(&it, 0, sizeof(it));
  if (rh_geometry(jb.r.W, jb.r.H, jb.s->desc.chunk_x, jb.s->desc.chunk_y, jb.s->desc.height,
                  &it.g))
    return RT_ERR_UNSUPPORTED;
  // queue entries pack the triangle slot in 26 bits
  if (jb.s->ntri >= (1u << 26)) return RT_ERR_UNSUPPORTED;
  if (bvh && jb.s->ntri > 0 && !jb.s->has_bvh) return RT_ERR_STATE;
  it.r = jb.r;
  it.v9 = jb.s->d_v9;
  it.uv6 = jb.s->d_uv6;
  it.v9s = jb.s->d_v9s;
  it.perm = jb.s->d_perm;
  it.sperm = jb.s->d_sperm;
  it.cbox = jb.s->d_cbox;
  it.tbox = jb.s->d_tbox;
  it.tmt = jb.s->d_tmt;
  it.snode = jb.s->d_snode;
  it.tmts = jb.s->d_tmts;
  it.tleaf = jb.s->d_tleaf;
  it.ntri = jb.s->ntri;
  it.flags = jb.s->ctx->d_flags;
  it.nclusters = bvh ? (jb.s->ntri ? jb.s->nrclusters : 0) : jb.s->nclusters;
  if (bvh) {
    it.cbox = jb.s->d_rcbox;
    it.nroots = jb.s->nroots;
    it.bvh = jb.s->d_bvh;
    it.roots = jb.s->d_roots;
    it.prim_order = jb.s->d_prim_order;
    it.bvh_parent = jb.s->d_bvh_parent;
    it.bvh_depth = jb.s->d_bvh_depth;
  }
  it.out = jb.out;
  it.occ = jb.occ;
  it.repair_run = (jb.mode == 0 && !bvh) ? jb.s->desc.repair_run : 0;
  return RT_OK;
}

// Lays out the jobs as launch groups (one per mode and node type, in job order) in the
// workspace arena, fills the job tables and uploads them when they changed.  `changed` tells
// whether the device tables or arena were rewritten.
int rh_prepare(RhWorkspace& ws, const std::vector<RhJob>& jobs, const RhParams& prm,
               hipStream_t st, std::vector<RhGroup>& groups, bool* changed) {
  groups.clear();
  *changed = false;
  // group membership in job order
  std::vector<std::pair<int, int>> keys;  // (mode, node + 4 * is_raybvh), node 0..2
  std::vector<std::vector<int>> members;
  for (size_t j = 0; j < jobs.size(); ++j) {
    const int bvh = jobs[j].s->desc.kind == RT_ACCEL_RAYBVH ? 1 : 0;
    const std::pair<int, int> k{jobs[j].mode, jobs[j].s->desc.node + 4 * bvh};
    size_t gi = 0;
    while (gi < keys.size() && keys[gi] != k) ++gi;
    if (gi == keys.size()) {
      keys.push_back(k);
      members.emplace_back();
    }
    members[gi].push_back((int)j);
  }
  // the per-ray kernel's deferred-chain queue names a job in 16 bits
  for (const auto& m : members)
    if (m.size() > 65535u) return RT_ERR_UNSUPPORTED;
  // geometry and sizes
  std::vector<RhItem> items(jobs.size());
  std::vector<size_t> off_nodes(jobs.size()), off_clist(jobs.size()), off_keys(jobs.size()),
      off_rep(jobs.size());
  size_t total = 0;
  const size_t tables = align256(sizeof(RhItem) * std::max<size_t>(jobs.size(), 1));
  total += tables;
  struct GroupLayout {
    size_t units, ustage, bunits, rq, ctr, replist, max_units, max_chunks;
    uint32_t ray;
    int rgrid, walk;
    uint32_t nblocks, nrays;
    bool big, repair;
    int per, grid;
  };
  std::vector<GroupLayout> gl(keys.size());
  for (size_t gi = 0; gi < keys.size(); ++gi) {
    const int mode = keys[gi].first;
    const bool bvh = (keys[gi].second >> 2) != 0;
    const int mi = mode == 2 ? 1 : 0;
    GroupLayout& L = gl[gi];
    // work units: passing clusters of triangles (rayhierarchy) or of frontier roots (ray-bvh)
    L.per = prm.per[mi] > 0 ? prm.per[mi] : (bvh ? kDefaultRootClustersPerUnit : kDefaultPer[mi]);
    L.ray = (uint32_t)(prm.ray[mi] > 0 ? prm.ray[mi] : default_ray(bvh, mi));
    L.rgrid = prm.rgrid[mi];
    L.walk = prm.walk[mi] > 0 ? prm.walk[mi] : kDefaultWalk;
    L.max_units = 0;
    L.max_chunks = 0;
    L.nblocks = 0;
    L.nrays = 0;
    L.big = prm.tile[mi] == 1;
    L.repair = false;
    uint64_t blocks = 0, rays = 0;
    for (int j : members[gi]) {
      const RhJob& jb = jobs[j];
      RhItem& it = items[j];
      const int fst = fill_item(jb, it);
      if (fst) return fst;
      const size_t n = (size_t)jb.r.W * jb.r.H;
      it.block_base = (uint32_t)blocks;
      it.ray_base = (uint32_t)rays;
      it.chunk_base = (uint32_t)L.max_chunks;
      it.unit_base = (uint32_t)L.max_units;
      it.unit_stride = (it.nclusters + L.per - 1) / L.per;
      blocks += (uint64_t)it.g.ntop;
      rays += n;
      // the 1024-ray unit kernels, for blocks beyond 256 rays or 3 levels that the unit
      // kernels can take (larger blocks take the per-ray path)
      if (it.g.rb <= kMaxBlockRays && it.g.L <= kUnitLevels && (it.g.rb > 256 || it.g.L > 3))
        L.big = true;
      if (it.repair_run > 0) L.repair = true;
      L.max_units += (size_t)it.g.ntop * ((it.nclusters + L.per - 1) / L.per);
      L.max_chunks += (size_t)it.g.ntop * (size_t)((it.g.rb + 63) / 64);
      off_nodes[j] = total;
      total += align256(sizeof(float4) * 4 * (size_t)it.g.nnodes);
      off_clist[j] = total;
      total += align256(sizeof(uint32_t) * (size_t)it.g.ntop * std::max<uint32_t>(it.nclusters, 1));
      off_keys[j] = total;
      if (mode != 2) total += align256(sizeof(uint64_t) * n);
      off_rep[j] = total;
      if (it.repair_run > 0) total += align256(sizeof(uint32_t) * n);
    }
    if (blocks >= (1ull << 31) || rays >= (1ull << 31)) return RT_ERR_UNSUPPORTED;
    L.nblocks = (uint32_t)blocks;
    L.nrays = (uint32_t)rays;
    // 0: the launch takes one work-group per resident slot of the traversal kernel
    L.grid = prm.grid[mi] > 0
                 ? (int)std::max<size_t>(1, std::min<size_t>(L.max_units, (size_t)prm.grid[mi]))
                 : 0;
    L.units = total;
    total += align256(sizeof(uint4) * std::max<size_t>(L.max_units, 1));
    L.ustage = total;
    total += align256(sizeof(uint4) * std::max<size_t>(L.max_units, 1));
    L.bunits = total;
    total += align256(sizeof(uint4) * std::max<size_t>(L.nblocks, 1));
    L.rq = total;
    total += align256(sizeof(uint2) * std::max<size_t>(L.max_chunks, 1));
    L.ctr = total;
    total += align256(kCounterBytes);
    L.replist = total;
    if (L.repair) total += align256(sizeof(uint32_t) * L.nrays);
  }
  // arena
  if (total > ws.arena_cap || !ws.d_arena) {
    if (ws.d_arena) {
      hipError_t e = hipFree(ws.d_arena);  // waits for work still using it
      ws.d_arena = nullptr;
      ws.arena_cap = 0;
      if (e != hipSuccess) return hip_status(e);
    }
    ws.last_arena = nullptr;  // the new arena holds no table yet
    const size_t cap = total + total / 4;
    hipError_t e = hipMalloc((void**)&ws.d_arena, cap);
    if (e != hipSuccess) {
      ws.d_arena = nullptr;
      return hip_status(e);
    }
    ws.arena_cap = cap;
  }
  char* A = ws.d_arena;
  for (size_t j = 0; j < jobs.size(); ++j) {
    RhItem& it = items[j];
    it.nodes = (float4*)(A + off_nodes[j]);
    it.clist = (uint32_t*)(A + off_clist[j]);
    it.keys = jobs[j].mode != 2 ? (uint64_t*)(A + off_keys[j]) : nullptr;
    it.repkeys = it.repair_run > 0 ? (uint32_t*)(A + off_rep[j]) : nullptr;
  }
  // tables in group order
  std::vector<RhItem> table;
  table.reserve(jobs.size());
  for (size_t gi = 0; gi < keys.size(); ++gi) {
    RhGroup G{};
    G.items = (const RhItem*)A + table.size();
    for (int j : members[gi]) table.push_back(items[j]);
    G.nitems = (int)members[gi].size();
    G.mode = keys[gi].first;
    G.node = keys[gi].second & 3;
    G.bvh = keys[gi].second >> 2;
    G.big = gl[gi].big ? 1 : 0;
    G.nblocks = gl[gi].nblocks;
    G.nrays = gl[gi].nrays;
    G.units = (uint4*)(A + gl[gi].units);
    G.ustage = (uint4*)(A + gl[gi].ustage);
    G.bunits = (uint4*)(A + gl[gi].bunits);
    G.rq = (uint2*)(A + gl[gi].rq);
    G.max_chunks = (uint32_t)std::min<size_t>(gl[gi].max_chunks, 0xffffffffu);
    G.ray_clusters = gl[gi].ray;
    G.ray_grid = gl[gi].rgrid;
    G.walk = gl[gi].walk;
    G.ctr = (uint32_t*)(A + gl[gi].ctr);
    G.replist = gl[gi].repair ? (uint32_t*)(A + gl[gi].replist) : nullptr;
    G.per_unit = gl[gi].per;
    G.grid = gl[gi].grid;
    G.max_units = (uint32_t)std::min<size_t>(gl[gi].max_units, 0xffffffffu);
    groups.push_back(G);
  }
  const bool same = ws.last_arena == A && ws.last.size() == table.size() &&
                    (table.empty() ||
                     std::memcmp(ws.last.data(), table.data(), sizeof(RhItem) * table.size()) == 0);
  if (!same && !table.empty()) {
    if (!ws.uploaded) RT_TRY(hipEventCreateWithFlags(&ws.uploaded, hipEventDisableTiming));
    else RT_TRY(hipEventSynchronize(ws.uploaded));  // the pinned copy is free again
    if (table.size() > ws.table_cap) {
      (void)hipHostFree(ws.h_table);
      ws.h_table = nullptr;
      ws.table_cap = 0;
      RT_TRY(hipHostMalloc((void**)&ws.h_table, sizeof(RhItem) * table.size(), 0));
      ws.table_cap = table.size();
    }
    std::memcpy(ws.h_table, table.data(), sizeof(RhItem) * table.size());
    RT_TRY(hipMemcpyAsync(A, ws.h_table, sizeof(RhItem) * table.size(), hipMemcpyHostToDevice,
                          st));
    RT_TRY(hipEventRecord(ws.uploaded, st));
    ws.last = table;
    ws.last_arena = A;
    *changed = true;
  }
  return RT_OK;
}

}  // namespace

namespace rti {
uint32_t scene_nclusters(const rt_scene* s) {
  return s->desc.kind == RT_ACCEL_RAYBVH ? s->nrclusters : s->nclusters;
}
int scene_check_image(const rt_scene* s, int W, int H) {
  if (s->desc.kind != RT_ACCEL_RAYHIERARCHY && s->desc.kind != RT_ACCEL_RAYBVH)
    return RT_ERR_UNSUPPORTED;
  if (s->desc.kind == RT_ACCEL_RAYBVH && s->ntri > 0 && !s->has_bvh) return RT_ERR_STATE;
  RHGeom g;
  if (rh_geometry(W, H, s->desc.chunk_x, s->desc.chunk_y, s->desc.height, &g))
    return RT_ERR_UNSUPPORTED;
  if (s->ntri >= (1u << 26)) return RT_ERR_UNSUPPORTED;
  return RT_OK;
}
}  // namespace rti

extern "C" {

int rt_accel_desc_default(int32_t kind, rt_accel_desc* d) {
  if (!d || kind < RT_ACCEL_NAIVE || kind > RT_ACCEL_RAYBVH) return RT_ERR_INVALID;
  d->kind = kind;
  d->node = RT_NODE_BOX_DIR;
  d->chunk_x = 4;
  d->chunk_y = 4;
  d->height = 3;
  d->repair_run = kind == RT_ACCEL_RAYHIERARCHY ? 1 : 0;
  d->bvh_height = 3;
  d->max_bvh_prim = 1;
  return RT_OK;
}

int rt_ctx_create(int32_t device, rt_ctx** out) {
  if (!out) return RT_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return RT_ERR_NO_DEVICE;
  if (device < 0 || device >= n) return RT_ERR_INVALID;
  RT_TRY(hipSetDevice(device));
  // the allocation-free trace's scratch: its address on this device, looked up once
  if (noalloc_setup(device)) return RT_ERR_HIP;
  rt_ctx* c = new (std::nothrow) rt_ctx{device, nullptr};
  if (!c) return RT_ERR_OOM;
  hipError_t e = hipMalloc((void**)&c->d_flags, 256);
  if (e == hipSuccess) e = hipMemset(c->d_flags, 0, 256);
  if (e != hipSuccess) {
    (void)hipFree(c->d_flags);
    delete c;
    return e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;
  }
  *out = c;
  return RT_OK;
}

int rt_ctx_destroy(rt_ctx* c) {
  if (c) {
    (void)hipSetDevice(c->device);
    (void)hipFree(c->d_flags);
  }
  delete c;
  return RT_OK;
}

int rt_ctx_check(rt_ctx* c, uint32_t* flags) {
  if (!c || !flags) return RT_ERR_INVALID;
  RT_TRY(hipSetDevice(c->device));
  RT_TRY(hipDeviceSynchronize());
  RT_TRY(hipMemcpy(flags, c->d_flags, sizeof(uint32_t), hipMemcpyDeviceToHost));
  RT_TRY(hipMemset(c->d_flags, 0, sizeof(uint32_t)));
  return RT_OK;
}

int rt_scene_create(rt_ctx* ctx, const float* v9, const float* uv6, uint32_t ntri,
                    const rt_accel_desc* desc, rt_scene** out) {
  if (!ctx || !out || !desc || (ntri && !v9)) return RT_ERR_INVALID;
  *out = nullptr;
  int st = desc_ok(desc);
  if (st) return st;
  RT_TRY(hipSetDevice(ctx->device));
  rt_scene* s = new (std::nothrow) rt_scene();
  if (!s) return RT_ERR_OOM;
  s->ctx = ctx;
  s->desc = *desc;
  s->ntri = ntri;
  hipError_t e = hipSuccess;
  if (ntri) {
    e = hipMalloc((void**)&s->d_v9, sizeof(float) * 9 * (size_t)ntri);
    if (e == hipSuccess)
      e = hipMemcpy(s->d_v9, v9, sizeof(float) * 9 * (size_t)ntri, hipMemcpyHostToDevice);
    if (e == hipSuccess && uv6) {
      e = hipMalloc((void**)&s->d_uv6, sizeof(float) * 6 * (size_t)ntri);
      if (e == hipSuccess)
        e = hipMemcpy(s->d_uv6, uv6, sizeof(float) * 6 * (size_t)ntri, hipMemcpyHostToDevice);
    }
  }
  if (e == hipSuccess && desc->kind != RT_ACCEL_NAIVE && ntri) {
    // Morton-sorted traversal copy: clusters of 64 with AABBs (rayhierarchy block path); the
    // search BVH with its leaf-test operands in search order (per-ray path of both
    // accelerators)
    std::vector<float> v9s;
    std::vector<uint32_t> perm;
    std::vector<float4> cbox;
    sort_clusters(v9, ntri, v9s, perm, cbox);
    e = upload(s->d_v9s, v9s.data(), v9s.size());
    if (e == hipSuccess) e = upload(s->d_perm, perm.data(), perm.size());
    if (e == hipSuccess) {
      std::vector<float4> snode;
      std::vector<uint32_t> sorder;
      build_search_bvh(v9, ntri, snode, sorder);
      e = upload(s->d_snode, snode.data(), snode.size());
      if (e == hipSuccess) e = upload(s->d_sperm, sorder.data(), sorder.size());
      if (e == hipSuccess) {
        std::vector<float> v9o(9 * (size_t)ntri);
        for (uint32_t i = 0; i < ntri; ++i)
          std::memcpy(&v9o[9 * (size_t)i], v9 + 9 * (size_t)sorder[i], 9 * sizeof(float));
        const std::vector<float4> mts = mt_operands(v9o.data(), ntri);
        e = upload(s->d_tmts, mts.data(), mts.size());
      }
    }
    if (e == hipSuccess && desc->kind == RT_ACCEL_RAYHIERARCHY) {
      e = upload(s->d_cbox, cbox.data(), cbox.size());
      if (e == hipSuccess) {
        // per-triangle AABBs with the device's min/max selects, vertex order v1, v2, v3
        std::vector<float4> tbox(2 * (size_t)ntri);
        for (uint32_t i = 0; i < ntri; ++i) {
          const float* p = &v9s[9 * (size_t)i];
          float mn[3], mx[3];
          for (int a = 0; a < 3; ++a) {
            mn[a] = p[3 + a] < p[a] ? p[3 + a] : p[a];
            mn[a] = p[6 + a] < mn[a] ? p[6 + a] : mn[a];
            mx[a] = p[3 + a] > p[a] ? p[3 + a] : p[a];
            mx[a] = p[6 + a] > mx[a] ? p[6 + a] : mx[a];
          }
          tbox[2 * (size_t)i] = make_float4(mn[0], mn[1], mn[2], 0.f);
          tbox[2 * (size_t)i + 1] = make_float4(mx[0], mx[1], mx[2], 0.f);
        }
        e = upload(s->d_tbox, tbox.data(), tbox.size());
      }
      s->nclusters = (uint32_t)(cbox.size() / 2);
    }
  }
  if (e == hipSuccess && desc->kind == RT_ACCEL_RAYBVH && ntri) {
    const std::vector<float4> mt = mt_operands(v9, ntri);  // original order
    e = upload(s->d_tmt, mt.data(), mt.size());
  }
  if (e != hipSuccess) {
    rt_scene_destroy(s);
    return hip_status(e);
  }
  *out = s;
  return RT_OK;
}

int rt_scene_set_bvh(rt_scene* s, const rt_gpu_node* nodes, uint32_t nnodes, const uint32_t* roots,
                     uint32_t nroots, const uint32_t* prim_order) {
  if (!s || !nodes || !roots || !prim_order || nnodes == 0) return RT_ERR_INVALID;
  if (s->desc.kind != RT_ACCEL_RAYBVH) return RT_ERR_STATE;
  std::lock_guard<std::mutex> lk(s->mu);
  // Validate the handoff on the host so the device traversal never leaves its arrays.
  std::vector<float4> packed(2 * (size_t)nnodes);
  for (uint32_t i = 0; i < nnodes; ++i) {
    const rt_gpu_node& n = nodes[i];
    if (n.n_primitives) {
      if ((uint64_t)n.prim_offset + n.n_primitives > s->ntri) return RT_ERR_INVALID;
    } else if ((uint64_t)n.prim_offset + 1 >= nnodes) {
      return RT_ERR_INVALID;
    }
    uint32_t po = n.prim_offset, np = n.n_primitives;
    float fpo, fnp;
    std::memcpy(&fpo, &po, 4);
    std::memcpy(&fnp, &np, 4);
    packed[2 * (size_t)i] = make_float4(n.ax, n.ay, n.az, n.bx);
    packed[2 * (size_t)i + 1] = make_float4(n.by, n.bz, fpo, fnp);
  }
  // prim_order must be a permutation: every triangle sits at exactly one BVH position
  {
    std::vector<uint8_t> used(s->ntri, 0);
    for (uint32_t i = 0; i < s->ntri; ++i) {
      if (prim_order[i] >= s->ntri || used[prim_order[i]]) return RT_ERR_INVALID;
      used[prim_order[i]] = 1;
    }
  }
  // Every node below a frontier root gets its parent and its depth below that root (the
  // device rebuilds the exact lockstep cull chain of a hit from them).  Subtrees deeper than
  // 12 levels, nodes shared by two roots and cycles are rejected: the device traversal's
  // bounded stack holds 256 + 224 * depth items.
  if (nnodes >= (1u << 25)) return RT_ERR_UNSUPPORTED;
  std::vector<uint32_t> parent(nnodes, 0xffffffffu);
  std::vector<uint8_t> depth(nnodes, 0), seen(nnodes, 0);
  for (uint32_t r = 0; r < nroots; ++r) {
    if (roots[r] >= nnodes || seen[roots[r]]) return RT_ERR_INVALID;
    std::vector<uint32_t> stack{roots[r]};
    seen[roots[r]] = 1;
    parent[roots[r]] = roots[r];
    while (!stack.empty()) {
      const uint32_t n = stack.back();
      stack.pop_back();
      if (nodes[n].n_primitives) continue;
      // the 5DB kernels skip a node with no primitives and offset 0 with its subtree
      // (intersection5DB_BVH.cl:18,56); its children are not reached through it
      if (s->desc.node == RT_NODE_BOX_UV && nodes[n].prim_offset == 0) continue;
      if (depth[n] >= 12) return RT_ERR_UNSUPPORTED;
      for (uint32_t c = nodes[n].prim_offset; c <= nodes[n].prim_offset + 1; ++c) {
        if (seen[c]) return RT_ERR_INVALID;
        seen[c] = 1;
        parent[c] = n;
        depth[c] = (uint8_t)(depth[n] + 1);
        stack.push_back(c);
      }
    }
  }
  // Reference BVH leaf of every triangle below a frontier root, for the per-ray path: the leaf
  // node index, bit 31 set when the triangle is not the leaf's first primitive (6DB / IA
  // secondary rays test only the first, intersection6DB_BVH2.cl:88); ~0u for a triangle in no
  // reachable leaf.
  std::vector<uint32_t> tleaf(s->ntri, 0xffffffffu);
  for (uint32_t n = 0; n < nnodes; ++n) {
    if (!seen[n] || !nodes[n].n_primitives) continue;
    for (uint32_t f = 0; f < nodes[n].n_primitives; ++f)
      tleaf[prim_order[nodes[n].prim_offset + f]] = n | (f > 0 ? 0x80000000u : 0u);
  }
  // Frontier roots in Morton order of their box centres, cut into clusters of 64 with the
  // union of their boxes (the traversal order does not change any result).
  std::vector<uint32_t> rsorted(roots, roots + nroots);
  std::vector<float4> rcbox;
  {
    std::vector<float> c(3 * (size_t)nroots);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t r = 0; r < nroots; ++r) {
      const rt_gpu_node& n = nodes[roots[r]];
      const float ctr[3] = {0.5f * (n.ax + n.bx), 0.5f * (n.ay + n.by), 0.5f * (n.az + n.bz)};
      for (int a = 0; a < 3; ++a) {
        c[3 * (size_t)r + a] = ctr[a];
        lo[a] = std::min(lo[a], ctr[a]);
        hi[a] = std::max(hi[a], ctr[a]);
      }
    }
    std::vector<uint32_t> code(nroots), idx(nroots);
    for (uint32_t r = 0; r < nroots; ++r) {
      code[r] = morton3(c.data() + 3 * (size_t)r, lo, hi);
      idx[r] = r;
    }
    std::stable_sort(idx.begin(), idx.end(),
                     [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
    for (uint32_t r = 0; r < nroots; ++r) rsorted[r] = roots[idx[r]];
    const uint32_t nc = (nroots + kClusterSize - 1) / kClusterSize;
    rcbox.assign(2 * (size_t)std::max<uint32_t>(nc, 1), make_float4(0.f, 0.f, 0.f, 0.f));
    for (uint32_t k = 0; k < nc; ++k) {
      float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (uint32_t r = k * kClusterSize; r < std::min(nroots, (k + 1) * kClusterSize); ++r) {
        const rt_gpu_node& n = nodes[rsorted[r]];
        mn[0] = std::min(mn[0], n.ax);
        mn[1] = std::min(mn[1], n.ay);
        mn[2] = std::min(mn[2], n.az);
        mx[0] = std::max(mx[0], n.bx);
        mx[1] = std::max(mx[1], n.by);
        mx[2] = std::max(mx[2], n.bz);
      }
      rcbox[2 * (size_t)k] = make_float4(mn[0], mn[1], mn[2], 0.f);
      rcbox[2 * (size_t)k + 1] = make_float4(mx[0], mx[1], mx[2], 0.f);
    }
    s->nrclusters = nc;
  }
  RT_TRY(hipSetDevice(s->ctx->device));
  if (s->d_rcbox) (void)hipFree(s->d_rcbox);
  s->d_rcbox = nullptr;
  if (s->d_bvh) (void)hipFree(s->d_bvh);
  if (s->d_roots) (void)hipFree(s->d_roots);
  if (s->d_prim_order) (void)hipFree(s->d_prim_order);
  if (s->d_bvh_parent) (void)hipFree(s->d_bvh_parent);
  if (s->d_bvh_depth) (void)hipFree(s->d_bvh_depth);
  if (s->d_tleaf) (void)hipFree(s->d_tleaf);
  s->d_tleaf = nullptr;
  s->d_bvh = nullptr;
  s->d_roots = s->d_prim_order = s->d_bvh_parent = nullptr;
  s->d_bvh_depth = nullptr;
  s->has_bvh = false;
  RT_TRY(upload(s->d_bvh, packed.data(), packed.size()));
  RT_TRY(upload(s->d_roots, rsorted.data(), nroots));
  RT_TRY(upload(s->d_rcbox, rcbox.data(), rcbox.size()));
  RT_TRY(upload(s->d_prim_order, prim_order, s->ntri));
  RT_TRY(upload(s->d_bvh_parent, parent.data(), parent.size()));
  RT_TRY(upload(s->d_bvh_depth, depth.data(), depth.size()));
  RT_TRY(upload(s->d_tleaf, tleaf.data(), tleaf.size()));
  s->nroots = nroots;
  s->has_bvh = true;
  return RT_OK;
}

int rt_scene_destroy(rt_scene* s) {
  if (!s) return RT_OK;
  (void)hipSetDevice(s->ctx->device);
  (void)hipFree(s->d_v9);
  (void)hipFree(s->d_uv6);
  (void)hipFree(s->d_v9s);
  (void)hipFree(s->d_perm);
  (void)hipFree(s->d_sperm);
  (void)hipFree(s->d_cbox);
  (void)hipFree(s->d_tbox);
  (void)hipFree(s->d_tmt);
  (void)hipFree(s->d_snode);
  (void)hipFree(s->d_tmts);
  (void)hipFree(s->d_tleaf);
  (void)hipFree(s->d_bvh);
  (void)hipFree(s->d_roots);
  (void)hipFree(s->d_prim_order);
  (void)hipFree(s->d_bvh_parent);
  (void)hipFree(s->d_bvh_depth);
  (void)hipFree(s->d_rcbox);
  (void)hipFree(s->d_keys);
  delete s;
  return RT_OK;
}

static int trace_impl(rt_scene* s, const rt_rays* rays, int mode, const rt_hits* hits,
                      uint8_t* occ, void* stream) {
  int st = check_rays(rays, true);
  if (st) return st;
  std::lock_guard<std::mutex> lk(s->mu);
  RT_TRY(hipSetDevice(s->ctx->device));
  hipStream_t hs = (hipStream_t)stream;
  const size_t n = (size_t)rays->width * rays->height;
  const RaysDev r = to_dev(rays);
  const bool closest = mode != 2;
  const HitOut out = closest ? HitOut{hits->id, hits->t, hits->u, hits->v, hits->dpdu, hits->dpdv}
                             : HitOut{};
  const uint32_t ntri = s->ntri;
  const int kind = s->desc.kind;
  if (kind == RT_ACCEL_RAYHIERARCHY || kind == RT_ACCEL_RAYBVH) {
    // a launch group of one job (an empty scene yields no work units: every ray misses)
    if (kind == RT_ACCEL_RAYBVH && !s->has_bvh && ntri > 0) return RT_ERR_STATE;
    const std::vector<RhJob> jobs{RhJob{s, r, mode, out, occ}};
    std::vector<RhGroup> groups;
    bool changed = false;
    RhParams prm;
    prm.per[0] = kind == RT_ACCEL_RAYBVH ? s->roots_per_unit : s->clusters_closest;
    prm.per[1] = kind == RT_ACCEL_RAYBVH ? s->roots_per_unit : s->clusters_occluded;
    prm.ray[0] = s->ray_closest;
    prm.ray[1] = s->ray_occluded;
    prm.walk[0] = s->walk_closest;
    prm.walk[1] = s->walk_occluded;
    prm.tile[0] = s->tile_closest;
    prm.tile[1] = s->tile_occluded;
    st = rh_prepare(s->ws[mode], jobs, prm, hs, groups, &changed);
    if (st) return st;
    Timer::Slot* t = s->timer.next(mode, true);
    RT_TRY(launch_rh_group(groups[0], hs, t ? t->a : nullptr, t ? t->m : nullptr,
                           t ? t->b : nullptr));
    return RT_OK;
  }
  if (closest) {
    st = grow(s->d_keys, s->keys_cap, n);
    if (st) return st;
  }
  if (ntri == 0) {
    // empty scene: everything misses
    if (closest) RT_TRY(launch_fill_u64(s->d_keys, ~0ull, n, hs));
    else RT_TRY(launch_fill_u8(occ, 0, n, hs));
  } else if (kind == RT_ACCEL_NAIVE) {
    const int nblk = (int)((n + 255) / 256);
    const int tchunks = pick_tchunks(nblk, ntri, 64);
    if (tchunks > 1) {
      if (closest) RT_TRY(launch_fill_u64(s->d_keys, ~0ull, n, hs));
      else RT_TRY(launch_fill_u8(occ, 0, n, hs));
    }
    RT_TRY(timed_launch(s->timer, mode, hs, [&] {
      return launch_naive_trace(r, mode, s->d_v9, ntri, tchunks, s->d_keys, occ, hs);
    }));
  }
  if (closest)
    RT_TRY(launch_finalize(r, s->d_v9, s->d_uv6, s->d_keys, out, kind == RT_ACCEL_NAIVE, hs));
  return RT_OK;
}

int rt_trace_closest(rt_scene* s, const rt_rays* rays, int32_t mode, const rt_hits* hits,
                     void* stream) {
  if (!s || !hits || !hits->id || !hits->t || !hits->u || !hits->v) return RT_ERR_INVALID;
  if (mode != RT_CLOSEST_PRIMARY && mode != RT_CLOSEST_SECONDARY) return RT_ERR_INVALID;
  return trace_impl(s, rays, mode, hits, nullptr, stream);
}

int rt_trace_occluded(rt_scene* s, const rt_rays* rays, uint8_t* occluded, void* stream) {
  if (!s || !occluded) return RT_ERR_INVALID;
  return trace_impl(s, rays, 2, nullptr, occluded, stream);
}

int rt_trace_closest_noalloc(rt_scene* s, const rt_rays* rays, int32_t mode, const rt_hits* hits,
                             void* stream) {
  if (!s || !hits || !hits->id || !hits->t || !hits->u || !hits->v) return RT_ERR_INVALID;
  if (mode != RT_CLOSEST_PRIMARY && mode != RT_CLOSEST_SECONDARY) return RT_ERR_INVALID;
  int st = check_rays(rays, true);
  if (st) return st;
  char* base = (char*)noalloc_scratch(s->ctx->device);
  if (!base) return RT_ERR_STATE;  // the context's set-up found no scratch on this device
  RT_TRY(hipSetDevice(s->ctx->device));
  hipStream_t hs = (hipStream_t)stream;
  const size_t n = (size_t)rays->width * rays->height;
  const RaysDev r = to_dev(rays);
  const HitOut out{hits->id, hits->t, hits->u, hits->v, hits->dpdu, hits->dpdv};
  const int kind = s->desc.kind;
  if (kind == RT_ACCEL_NAIVE || s->ntri == 0) {
    // naive brute force (or an empty scene): the reduction keys are the only scratch
    if (align256(sizeof(uint64_t) * n) > kNoAllocScratch) return RT_ERR_TOO_LARGE;
    uint64_t* keys = (uint64_t*)base;
    RT_TRY(launch_fill_u64(keys, ~0ull, n, hs));
    if (s->ntri > 0) {
      const int nblk = (int)((n + 255) / 256);
      const int tchunks = pick_tchunks(nblk, s->ntri, 64);
      RT_TRY(launch_naive_trace(r, mode, s->d_v9, s->ntri, tchunks, keys, nullptr, hs));
    }
    RT_TRY(launch_finalize(r, s->d_v9, s->d_uv6, keys, out, kind == RT_ACCEL_NAIVE, hs));
    return RT_OK;
  }
  // rayhierarchy / ray-bvh: one job whose blocks all take the per-ray path (the same result as
  // the block path), so no cluster lists are needed
  RhItem it;
  st = fill_item(RhJob{s, r, mode, out, nullptr}, it);
  if (st) return st;
  const uint32_t nch = (uint32_t)((it.g.rb + 63) / 64);
  const bool repair = it.repair_run > 0;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += align256(std::max<size_t>(bytes, 1));
    return base + at;
  };
  RhItem* table = (RhItem*)take(sizeof(RhItem));
  it.nodes = (float4*)take(sizeof(float4) * 4 * (size_t)it.g.nnodes);
  it.keys = (uint64_t*)take(sizeof(uint64_t) * n);
  it.repkeys = repair ? (uint32_t*)take(sizeof(uint32_t) * n) : nullptr;
  it.clist = nullptr;
  RhGroup G{};
  G.units = (uint4*)take(sizeof(uint4));
  G.ustage = G.units;  // every block takes the per-ray path: no unit is staged
  G.bunits = (uint4*)take(sizeof(uint4) * (size_t)it.g.ntop);
  G.rq = (uint2*)take(sizeof(uint2) * (size_t)it.g.ntop * nch);
  G.ctr = (uint32_t*)take(kCounterBytes);
  G.replist = repair ? (uint32_t*)take(sizeof(uint32_t) * n) : nullptr;
  if (off > kNoAllocScratch || n >= (1ull << 31)) return RT_ERR_TOO_LARGE;
  it.block_base = 0;
  it.ray_base = 0;
  it.chunk_base = 0;
  G.items = table;
  G.nitems = 1;
  G.mode = mode;
  G.node = s->desc.node;
  G.bvh = kind == RT_ACCEL_RAYBVH ? 1 : 0;
  G.big = (it.g.rb > 256 || it.g.L > 3) ? 1 : 0;
  G.nblocks = (uint32_t)it.g.ntop;
  G.nrays = (uint32_t)n;
  G.per_unit = 1;
  G.grid = 1;
  G.max_units = 0;
  G.max_chunks = (uint32_t)it.g.ntop * nch;
  G.ray_clusters = 0;
  G.ray_grid = 0;
  G.walk = kDefaultWalk;
  RT_TRY(launch_write_table(it, table, hs));
  RT_TRY(launch_rh_group(G, hs, nullptr, nullptr, nullptr));
  return RT_OK;
}

int rt_make_shadow_rays(const rt_rays* primary, const int32_t* id, const float* t,
                        const float light[3], float* ox, float* oy, float* oz, float* dx,
                        float* dy, float* dz, float* tmin, float* tmax, uint8_t* valid,
                        void* stream) {
  int st = check_rays(primary, false);
  if (st) return st;
  if (!id || !t || !light || !ox || !oy || !oz || !dx || !dy || !dz || !tmin || !tmax || !valid)
    return RT_ERR_INVALID;
  RT_TRY(launch_shadow_rays(to_dev(primary), id, t, light[0], light[1], light[2], ox, oy, oz, dx,
                            dy, dz, tmin, tmax, valid, (hipStream_t)stream));
  return RT_OK;
}

int rt_make_bounce_rays(rt_scene* s, const rt_rays* in, const int32_t* id, const float* t,
                        uint32_t seed, int32_t depth, float* ox, float* oy, float* oz, float* dx,
                        float* dy, float* dz, float* tmin, float* tmax, uint8_t* valid,
                        void* stream) {
  if (!s) return RT_ERR_INVALID;
  int st = check_rays(in, true);
  if (st) return st;
  if (!id || !t || !ox || !oy || !oz || !dx || !dy || !dz || !tmin || !tmax || !valid)
    return RT_ERR_INVALID;
  if (s->ntri == 0) {
    RT_TRY(launch_fill_u8(valid, 0, (size_t)in->width * in->height, (hipStream_t)stream));
    return RT_OK;
  }
  RT_TRY(launch_bounce_rays(to_dev(in), s->d_v9, id, t, seed, depth, ox, oy, oz, dx, dy, dz,
                            tmin, tmax, valid, (hipStream_t)stream));
  return RT_OK;
}

int rt_scene_set_param(rt_scene* s, int32_t param, int32_t value) {
  if (!s || value < 0) return RT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->desc.kind == RT_ACCEL_RAYBVH &&
      (param == RT_PARAM_CLUSTERS_CLOSEST || param == RT_PARAM_CLUSTERS_OCCLUDED))
    s->roots_per_unit = value;
  else if (param == RT_PARAM_CLUSTERS_CLOSEST) s->clusters_closest = value;
  else if (param == RT_PARAM_CLUSTERS_OCCLUDED) s->clusters_occluded = value;
  else if (param == RT_PARAM_RAY_CLUSTERS_CLOSEST) s->ray_closest = value;
  else if (param == RT_PARAM_RAY_CLUSTERS_OCCLUDED) s->ray_occluded = value;
  else if (param == RT_PARAM_WALK_CLOSEST && value <= 7) s->walk_closest = value;
  else if (param == RT_PARAM_WALK_OCCLUDED && value <= 7) s->walk_occluded = value;
  else if (param == RT_PARAM_TILE_CLOSEST && value <= 1) s->tile_closest = value;
  else if (param == RT_PARAM_TILE_OCCLUDED && value <= 1) s->tile_occluded = value;
  else return RT_ERR_INVALID;
  return RT_OK;
}

int rt_scene_set_timing(rt_scene* s, int32_t enable) {
  if (!s) return RT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  s->timer.on = enable != 0;
  s->timer.used = 0;
  return RT_OK;
}

int rt_scene_kernel_time(rt_scene* s, int32_t mode, float* total_ms, int32_t* launches) {
  if (!s || !total_ms || !launches || mode < 0 || mode > 8) return RT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  return s->timer.total(mode % 3, mode < 3 ? 0 : mode < 6 ? 2 : 1, total_ms, launches);
}

int rt_batch_create(rt_ctx* ctx, rt_batch** out) {
  if (!ctx || !out) return RT_ERR_INVALID;
  *out = nullptr;
  rt_batch* b = new (std::nothrow) rt_batch();
  if (!b) return RT_ERR_OOM;
  b->ctx = ctx;
  *out = b;
  return RT_OK;
}

int rt_batch_destroy(rt_batch* b) {
  if (!b) return RT_OK;
  (void)hipSetDevice(b->ctx->device);
  delete b;
  return RT_OK;
}

int rt_batch_clear(rt_batch* b) {
  if (!b) return RT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(b->mu);
  b->jobs.clear();
  return RT_OK;
}

static int batch_add(rt_batch* b, rt_scene* s, const rt_rays* rays, int mode, const rt_hits* hits,
                     uint8_t* occ) {
  if (!b || !s) return RT_ERR_INVALID;
  if (s->ctx->device != b->ctx->device) return RT_ERR_INVALID;
  int st = check_rays(rays, true);
  if (st) return st;
  st = rti::scene_check_image(s, rays->width, rays->height);
  if (st) return st;
  std::lock_guard<std::mutex> lk(b->mu);
  const HitOut out = hits ? HitOut{hits->id, hits->t, hits->u, hits->v, hits->dpdu, hits->dpdv}
                          : HitOut{};
  b->jobs.push_back(RhJob{s, to_dev(rays), mode, out, occ});
  return RT_OK;
}

int rt_batch_add_closest(rt_batch* b, rt_scene* s, const rt_rays* rays, int32_t mode,
                         const rt_hits* hits) {
  if (!hits || !hits->id || !hits->t || !hits->u || !hits->v) return RT_ERR_INVALID;
  if (mode != RT_CLOSEST_PRIMARY && mode != RT_CLOSEST_SECONDARY) return RT_ERR_INVALID;
  return batch_add(b, s, rays, mode, hits, nullptr);
}

int rt_batch_add_occluded(rt_batch* b, rt_scene* s, const rt_rays* rays, uint8_t* occluded) {
  if (!occluded) return RT_ERR_INVALID;
  return batch_add(b, s, rays, 2, nullptr, occluded);
}

// Launches every group of the batch on `hs` (eagerly).
static int batch_launch(rt_batch* b, const std::vector<RhGroup>& groups, hipStream_t hs,
                        bool timed) {
  for (const RhGroup& G : groups) {
    Timer::Slot* t = timed ? b->timer.next(G.mode, true) : nullptr;
    RT_TRY(launch_rh_group(G, hs, t ? t->a : nullptr, t ? t->m : nullptr, t ? t->b : nullptr));
  }
  return RT_OK;
}

// Captured form of batch_launch.  The groups of a batch are independent (disjoint jobs, arena
// regions and counters), so every group after the first runs on a side stream forked from hs
// and joined back: the graph holds the groups as parallel branches, and one group's small
// kernels (build, cluster pass, finalize) overlap another's traversal.  The side streams and
// events exist before the capture begins (see ensure_side).
constexpr size_t kSideStreams = 3;  // with the caller's stream, the device's 4 hardware queues
static int batch_launch_forked(rt_batch* b, const std::vector<RhGroup>& groups, hipStream_t hs) {
  const size_t nside = std::min(groups.size() - 1, b->side.size());
  if (groups.size() < 2 || nside == 0) return batch_launch(b, groups, hs, false);
  RT_TRY(hipEventRecord(b->forks[0], hs));
  for (size_t k = 0; k < nside; ++k) RT_TRY(hipStreamWaitEvent(b->side[k], b->forks[0], 0));
  for (size_t gi = 0; gi < groups.size(); ++gi)
    RT_TRY(launch_rh_group(groups[gi], gi == 0 ? hs : b->side[(gi - 1) % nside], nullptr,
                           nullptr, nullptr));
  for (size_t k = 0; k < nside; ++k) {
    RT_TRY(hipEventRecord(b->forks[k + 1], b->side[k]));
    RT_TRY(hipStreamWaitEvent(hs, b->forks[k + 1], 0));
  }
  return RT_OK;
}

// Creates the side streams and events that batch_launch_forked needs for `ngroups` groups.
static int ensure_side(rt_batch* b, size_t ngroups) {
  const size_t want = std::min(ngroups > 0 ? ngroups - 1 : 0, kSideStreams);
  while (b->side.size() < want) {
    hipStream_t s = nullptr;
    RT_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    b->side.push_back(s);
  }
  while (b->forks.size() < b->side.size() + 1) {
    hipEvent_t e = nullptr;
    RT_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->forks.push_back(e);
  }
  return RT_OK;
}

static int batch_run_locked(rt_batch* b, hipStream_t hs) {
  RT_TRY(hipSetDevice(b->ctx->device));
  if (b->jobs.empty()) return RT_OK;
  std::vector<RhGroup> groups;
  bool changed = false;
  int st = rh_prepare(b->ws, b->jobs, b->prm, hs, groups, &changed);
  if (st) return st;
  // Graph replay needs a capturable stream and no per-launch timing events.
  if (!b->use_graph || !hs || b->timer.on) return batch_launch(b, groups, hs, true);
  if (changed || !b->graph_valid || b->graph_stream != hs) {
    b->drop_graph();
    st = ensure_side(b, groups.size());
    if (st) return st;
    RT_TRY(hipStreamBeginCapture(hs, hipStreamCaptureModeThreadLocal));
    st = batch_launch_forked(b, groups, hs);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(hs, &g);
    if (st) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    RT_TRY(e);
    b->graph = g;
    RT_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    b->graph_stream = hs;
    b->graph_valid = true;
  }
  RT_TRY(hipGraphLaunch(b->exec, hs));
  return RT_OK;
}

int rt_batch_run(rt_batch* b, void* stream) {
  if (!b) return RT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(b->mu);
  return batch_run_locked(b, (hipStream_t)stream);
}

static int* batch_param(rt_batch* b, int32_t param) {
  switch (param) {
    case RT_PARAM_CLUSTERS_CLOSEST: return &b->prm.per[0];
    case RT_PARAM_CLUSTERS_OCCLUDED: return &b->prm.per[1];
    case RT_PARAM_GRID_CLOSEST: return &b->prm.grid[0];
    case RT_PARAM_GRID_OCCLUDED: return &b->prm.grid[1];
    case RT_PARAM_RAY_CLUSTERS_CLOSEST: return &b->prm.ray[0];
    case RT_PARAM_RAY_CLUSTERS_OCCLUDED: return &b->prm.ray[1];
    case RT_PARAM_RAY_GRID_CLOSEST: return &b->prm.rgrid[0];
    case RT_PARAM_RAY_GRID_OCCLUDED: return &b->prm.rgrid[1];
    case RT_PARAM_WALK_CLOSEST: return &b->prm.walk[0];
    case RT_PARAM_WALK_OCCLUDED: return &b->prm.walk[1];
    case RT_PARAM_TILE_CLOSEST: return &b->prm.tile[0];
    case RT_PARAM_TILE_OCCLUDED: return &b->prm.tile[1];
    default: return nullptr;
  }
}

int rt_batch_set_param(rt_batch* b, int32_t param, int32_t value) {
  if (!b || value < 0) return RT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(b->mu);
  if (param == RT_PARAM_GRAPH) {
    b->use_graph = value != 0;
    if (!b->use_graph) b->drop_graph();
    return RT_OK;
  }
  int* p = batch_param(b, param);
  if (!p) return RT_ERR_INVALID;
  if (param == RT_PARAM_GRID_CLOSEST || param == RT_PARAM_GRID_OCCLUDED ||
      param == RT_PARAM_RAY_CLUSTERS_CLOSEST || param == RT_PARAM_RAY_CLUSTERS_OCCLUDED ||
      param == RT_PARAM_RAY_GRID_CLOSEST || param == RT_PARAM_RAY_GRID_OCCLUDED) {
    if (value > (1 << 30)) return RT_ERR_INVALID;
  } else if (param == RT_PARAM_WALK_CLOSEST || param == RT_PARAM_WALK_OCCLUDED) {
    if (value > 7) return RT_ERR_INVALID;
  } else if (param == RT_PARAM_TILE_CLOSEST || param == RT_PARAM_TILE_OCCLUDED) {
    if (value > 1) return RT_ERR_INVALID;
  } else if (value > (1 << 16)) {
    return RT_ERR_INVALID;
  }
  *p = value;
  b->graph_valid = false;
  return RT_OK;
}

int rt_batch_get_param(rt_batch* b, int32_t param, int32_t* value) {
  if (!b || !value) return RT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(b->mu);
  if (param == RT_PARAM_GRAPH) {
    *value = b->use_graph ? 1 : 0;
    return RT_OK;
  }
  int* p = batch_param(b, param);
  if (!p) return RT_ERR_INVALID;
  const int mi = (param == RT_PARAM_CLUSTERS_OCCLUDED || param == RT_PARAM_GRID_OCCLUDED ||
                  param == RT_PARAM_RAY_CLUSTERS_OCCLUDED || param == RT_PARAM_RAY_GRID_OCCLUDED ||
                  param == RT_PARAM_WALK_OCCLUDED)
                     ? 1
                     : 0;
  if (param == RT_PARAM_WALK_CLOSEST || param == RT_PARAM_WALK_OCCLUDED) {
    *value = *p > 0 ? *p : kDefaultWalk;
    return RT_OK;
  }
  if (param == RT_PARAM_TILE_CLOSEST || param == RT_PARAM_TILE_OCCLUDED) {
    *value = *p;
    return RT_OK;
  }
  const bool grid = param == RT_PARAM_GRID_CLOSEST || param == RT_PARAM_GRID_OCCLUDED ||
                    param == RT_PARAM_RAY_GRID_CLOSEST || param == RT_PARAM_RAY_GRID_OCCLUDED;
  const bool ray = param == RT_PARAM_RAY_CLUSTERS_CLOSEST || param == RT_PARAM_RAY_CLUSTERS_OCCLUDED;
  // 0: the built-in choice (which for the per-ray thresholds depends on the accelerator)
  *value = *p > 0 ? *p : ((grid || ray) ? 0 : kDefaultPer[mi]);
  return RT_OK;
}

}  // extern "C"

// Times the batch over `cands` launch-shape candidates (each applies itself to b->prm for
// both modes) and keeps, per mode, the fields `keep` copies from the fastest candidate.
template <class Apply, class Keep>
static int autotune_stage(rt_batch* b, hipStream_t hs, int reps, int ncand, Apply apply,
                          Keep keep) {
  float best[2] = {INFINITY, INFINITY};
  RhParams out = b->prm;
  const RhParams base = b->prm;
  int st = RT_OK;
  for (int c = 0; c < ncand && !st; ++c) {
    b->prm = base;
    apply(c, b->prm);
    st = batch_run_locked(b, hs);  // warm-up (allocates)
    b->timer.used = 0;
    for (int r = 0; r < reps && !st; ++r) st = batch_run_locked(b, hs);
    if (st) break;
    for (int m = 0; m < 2 && !st; ++m) {
      float ms = 0.f, ms2 = 0.f;
      int32_t n = 0, n2 = 0;
      if (m == 0) {
        st = b->timer.total(0, &ms, &n);
        if (!st) st = b->timer.total(1, &ms2, &n2);
        ms += ms2;
        n += n2;
      } else {
        st = b->timer.total(2, &ms, &n);
      }
      if (!st && n > 0 && ms < best[m]) {
        best[m] = ms;
        keep(m, b->prm, out);
      }
    }
    b->timer.used = 0;
  }
  b->prm = out;
  return st;
}

extern "C" {

int rt_batch_autotune(rt_batch* b, void* stream, int32_t reps) {
  if (!b || reps <= 0) return RT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(b->mu);
  RT_TRY(hipSetDevice(b->ctx->device));
  hipStream_t hs = (hipStream_t)stream;
  if (b->jobs.empty()) return RT_OK;
  const bool was_on = b->timer.on;
  const bool was_graph = b->use_graph;
  b->use_graph = false;
  b->timer.on = true;
  // stage 1: which blocks take the per-ray path (1: all, 2^30: none) and the per-ray grid
  // (one wavefront per work-group: 8 or 16 per CU, or every resident slot)
  static const int kRay[] = {1, 64, 256, 1 << 29, 1 << 30};
  static const int kRGrid[] = {0, 2048, 4096};
  auto routing = [&]() {
    return autotune_stage(
        b, hs, reps, 5 * 3,
        [&](int c, RhParams& p) {
          p.ray[0] = p.ray[1] = kRay[c / 3];
          p.rgrid[0] = p.rgrid[1] = kRGrid[c % 3];
        },
        [](int m, const RhParams& p, RhParams& o) {
          o.ray[m] = p.ray[m];
          o.rgrid[m] = p.rgrid[m];
        });
  };
  int st = routing();
  // stage 2: how far a round of the per-ray kernel walks before its leaf tests
  static const int kWalk[] = {1, 2, 3, 7};
  if (!st)
    st = autotune_stage(
        b, hs, reps, 4, [&](int c, RhParams& p) { p.walk[0] = p.walk[1] = kWalk[c]; },
        [](int m, const RhParams& p, RhParams& o) { o.walk[m] = p.walk[m]; });
  // stage 1 again with the walk just chosen: a faster per-ray kernel can move the best split
  // between the two paths (on a batch of 64 cluster scenes it then sent every camera block
  // per-ray: 7.65 ms per step against 7.73)
  if (!st) st = routing();
  // stage 3: passing clusters per work unit x persistent grid of the unit kernel
  static const int kPer[] = {2, 4, 8, 16, 32};
  static const int kGrid[] = {0, 1024, 2048};  // 0: one work-group per resident slot
  if (!st)
    st = autotune_stage(
        b, hs, reps, 5 * 3,
        [&](int c, RhParams& p) {
          p.per[0] = p.per[1] = kPer[c % 5];
          p.grid[0] = p.grid[1] = kGrid[c / 5];
        },
        [](int m, const RhParams& p, RhParams& o) {
          o.per[m] = p.per[m];
          o.grid[m] = p.grid[m];
        });
  // stage 4: the unit kernel's LDS tile (by block size, or 1024 rays for every block)
  if (!st)
    st = autotune_stage(
        b, hs, reps, 2, [&](int c, RhParams& p) { p.tile[0] = p.tile[1] = c; },
        [](int m, const RhParams& p, RhParams& o) { o.tile[m] = p.tile[m]; });
  b->timer.on = was_on;
  b->timer.used = 0;
  b->use_graph = was_graph;
  b->graph_valid = false;
  return st;
}

int rt_batch_set_timing(rt_batch* b, int32_t enable) {
  if (!b) return RT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(b->mu);
  b->timer.on = enable != 0;
  b->timer.used = 0;
  return RT_OK;
}

int rt_batch_kernel_time(rt_batch* b, int32_t mode, float* total_ms, int32_t* launches) {
  if (!b || !total_ms || !launches || mode < 0 || mode > 8) return RT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(b->mu);
  return b->timer.total(mode % 3, mode < 3 ? 0 : mode < 6 ? 2 : 1, total_ms, launches);
}

const char* rt_error_string(int status) {
  switch (status) {
    case RT_OK: return "ok";
    case RT_ERR_INVALID: return "invalid argument";
    case RT_ERR_HIP: return "HIP runtime error";
    case RT_ERR_OOM: return "device out of memory";
    case RT_ERR_UNSUPPORTED: return "unsupported parameter combination";
    case RT_ERR_NO_DEVICE: return "no HIP device";
    case RT_ERR_STATE: return "call order violated";
    case RT_ERR_TOO_LARGE: return "input exceeds the static scratch of the allocation-free trace";
    default: return "unknown status";
  }
}

}  // extern "C"
