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

struct rt_ctx {
  int device;
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
  float4* d_cbox = nullptr;
  uint32_t nclusters = 0;
  int clusters_closest = 0, clusters_occluded = 0;  // 0 = built-in choice
  // ray-bvh handoff
  float4* d_bvh = nullptr;
  uint32_t* d_roots = nullptr;
  uint32_t* d_prim_order = nullptr;
  uint32_t nroots = 0;
  int tile_roots = 256;
  bool has_bvh = false;
  // scratch, grown on demand
  float4* d_nodes = nullptr;
  size_t nodes_cap = 0;
  uint64_t* d_keys = nullptr;
  size_t keys_cap = 0;
  uint32_t* d_repkeys = nullptr;
  size_t rep_cap = 0;
  uint32_t* d_replist = nullptr;
  size_t replist_cap = 0;
  uint32_t* d_clist = nullptr;  // passing clusters per block
  size_t clist_cap = 0;
  uint4* d_units = nullptr;  // work units {block, first, count}
  size_t units_cap = 0;
  uint32_t* d_ucount = nullptr;
  size_t ucount_cap = 0;
  // HIP-event timing of the traversal kernels (one event pair per launch while enabled)
  struct Timed {
    int mode;
    hipEvent_t a, b;
  };
  bool timing = false;
  std::vector<Timed> timed;
  size_t timed_used = 0;
  std::mutex mu;
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
  int want = (1024 + nblocks - 1) / std::max(nblocks, 1);
  int maxc = (int)std::max<uint32_t>(1, nwork / std::max<uint32_t>(per_tile, 1));
  return std::max(1, std::min(want, maxc));
}

int desc_ok(const rt_accel_desc* d) {
  if (d->kind < RT_ACCEL_NAIVE || d->kind > RT_ACCEL_RAYBVH) return RT_ERR_INVALID;
  if (d->kind == RT_ACCEL_NAIVE) return RT_OK;
  if (d->node != RT_NODE_BOX_DIR && d->node != RT_NODE_IA) return RT_ERR_UNSUPPORTED;
  if (d->chunk_x <= 0 || d->chunk_y <= 0 || d->height <= 0) return RT_ERR_INVALID;
  if (d->repair_run < 0) return RT_ERR_INVALID;
  if (d->kind == RT_ACCEL_RAYBVH) {
    if (d->bvh_height < 0 || d->bvh_height > 24) return RT_ERR_UNSUPPORTED;
    if (d->max_bvh_prim <= 0 || d->max_bvh_prim > 4) return RT_ERR_UNSUPPORTED;
  }
  return RT_OK;
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
  auto spread = [](uint32_t x) {
    x &= 1023u;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
  };
  std::vector<uint32_t> code(ntri);
  for (uint32_t i = 0; i < ntri; ++i) {
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
      const float ext = hi[a] - lo[a];
      float f = ext > 0.f ? (c[3 * (size_t)i + a] - lo[a]) / ext : 0.f;
      f = std::min(std::max(f, 0.f), 1.f);
      q[a] = (uint32_t)(f * 1023.f);
    }
    code[i] = (spread(q[0]) << 2) | (spread(q[1]) << 1) | spread(q[2]);
  }
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

template <class T>
hipError_t upload(T*& d, const T* h, size_t n) {
  hipError_t e = hipMalloc((void**)&d, sizeof(T) * std::max<size_t>(n, 1));
  if (e == hipSuccess && n) e = hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice);
  return e;
}

// Records the start event of a traversal launch; returns the slot or -1.
int timing_begin(rt_scene* s, int mode, hipStream_t st) {
  if (!s->timing) return -1;
  if (s->timed_used == s->timed.size()) {
    rt_scene::Timed t{mode, nullptr, nullptr};
    if (hipEventCreate(&t.a) != hipSuccess) return -1;
    if (hipEventCreate(&t.b) != hipSuccess) {
      (void)hipEventDestroy(t.a);
      return -1;
    }
    s->timed.push_back(t);
  }
  rt_scene::Timed& t = s->timed[s->timed_used];
  t.mode = mode;
  if (hipEventRecord(t.a, st) != hipSuccess) return -1;
  return (int)s->timed_used++;
}
void timing_end(rt_scene* s, int slot, hipStream_t st) {
  if (slot >= 0 && hipEventRecord(s->timed[slot].b, st) != hipSuccess) s->timed[slot].mode = -1;
}

}  // namespace

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
  rt_ctx* c = new (std::nothrow) rt_ctx{device};
  if (!c) return RT_ERR_OOM;
  *out = c;
  return RT_OK;
}

int rt_ctx_destroy(rt_ctx* c) {
  delete c;
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
  if (e == hipSuccess && desc->kind == RT_ACCEL_RAYHIERARCHY && ntri) {
    std::vector<float> v9s;
    std::vector<uint32_t> perm;
    std::vector<float4> cbox;
    sort_clusters(v9, ntri, v9s, perm, cbox);
    e = upload(s->d_v9s, v9s.data(), v9s.size());
    if (e == hipSuccess) e = upload(s->d_perm, perm.data(), perm.size());
    if (e == hipSuccess) e = upload(s->d_cbox, cbox.data(), cbox.size());
    s->nclusters = (uint32_t)(cbox.size() / 2);
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
  for (uint32_t i = 0; i < s->ntri; ++i)
    if (prim_order[i] >= s->ntri) return RT_ERR_INVALID;
  // Entries one frontier root can emit (primitives below it) bound the per-tile LDS list.
  // The device walks each root's subtree with a 32-entry stack: subtrees deeper than 24
  // levels (the BVHheight limit) are rejected here, as are cycles.
  uint32_t max_emit = 1;
  for (uint32_t r = 0; r < nroots; ++r) {
    if (roots[r] >= nnodes) return RT_ERR_INVALID;
    std::vector<std::pair<uint32_t, uint32_t>> stack{{roots[r], 0u}};
    uint32_t emit = 0;
    while (!stack.empty()) {
      auto [n, depth] = stack.back();
      stack.pop_back();
      if (depth > 24) return RT_ERR_UNSUPPORTED;
      if (nodes[n].n_primitives) emit += nodes[n].n_primitives;
      else {
        stack.push_back({nodes[n].prim_offset, depth + 1});
        stack.push_back({nodes[n].prim_offset + 1, depth + 1});
      }
    }
    max_emit = std::max(max_emit, emit);
  }
  if (max_emit > 2048) return RT_ERR_UNSUPPORTED;
  RT_TRY(hipSetDevice(s->ctx->device));
  if (s->d_bvh) (void)hipFree(s->d_bvh);
  if (s->d_roots) (void)hipFree(s->d_roots);
  if (s->d_prim_order) (void)hipFree(s->d_prim_order);
  s->d_bvh = nullptr;
  s->d_roots = s->d_prim_order = nullptr;
  s->has_bvh = false;
  RT_TRY(hipMalloc((void**)&s->d_bvh, sizeof(float4) * packed.size()));
  RT_TRY(hipMemcpy(s->d_bvh, packed.data(), sizeof(float4) * packed.size(), hipMemcpyHostToDevice));
  RT_TRY(hipMalloc((void**)&s->d_roots, sizeof(uint32_t) * std::max<uint32_t>(nroots, 1)));
  if (nroots)
    RT_TRY(hipMemcpy(s->d_roots, roots, sizeof(uint32_t) * nroots, hipMemcpyHostToDevice));
  RT_TRY(hipMalloc((void**)&s->d_prim_order, sizeof(uint32_t) * std::max<uint32_t>(s->ntri, 1)));
  if (s->ntri)
    RT_TRY(hipMemcpy(s->d_prim_order, prim_order, sizeof(uint32_t) * s->ntri, hipMemcpyHostToDevice));
  s->nroots = nroots;
  s->tile_roots = (int)std::max<uint32_t>(1, std::min<uint32_t>(256, 2048 / max_emit));
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
  (void)hipFree(s->d_cbox);
  (void)hipFree(s->d_bvh);
  (void)hipFree(s->d_roots);
  (void)hipFree(s->d_prim_order);
  (void)hipFree(s->d_nodes);
  (void)hipFree(s->d_keys);
  (void)hipFree(s->d_repkeys);
  (void)hipFree(s->d_replist);
  (void)hipFree(s->d_clist);
  (void)hipFree(s->d_units);
  (void)hipFree(s->d_ucount);
  for (auto& t : s->timed) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
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
  if (closest) {
    st = grow(s->d_keys, s->keys_cap, n);
    if (st) return st;
  }
  const uint32_t ntri = s->ntri;
  const int kind = s->desc.kind;
  if (ntri == 0 || (kind == RT_ACCEL_RAYBVH && s->nroots == 0)) {
    // empty scene: everything misses
    if (closest) RT_TRY(launch_fill_u64(s->d_keys, ~0ull, n, hs));
    else RT_TRY(launch_fill_u8(occ, 0, n, hs));
  } else if (kind == RT_ACCEL_NAIVE) {
    const int nblk = (int)((n + 255) / 256);
    const int tchunks = pick_tchunks(nblk, ntri, 256);
    if (tchunks > 1) {
      if (closest) RT_TRY(launch_fill_u64(s->d_keys, ~0ull, n, hs));
      else RT_TRY(launch_fill_u8(occ, 0, n, hs));
    }
    const int slot = timing_begin(s, mode, hs);
    RT_TRY(launch_naive_trace(r, mode, s->d_v9, ntri, tchunks, s->d_keys, occ, hs));
    timing_end(s, slot, hs);
  } else {
    RHGeom g;
    int ge = rh_geometry(rays->width, rays->height, s->desc.chunk_x, s->desc.chunk_y,
                         s->desc.height, &g);
    if (ge) return RT_ERR_UNSUPPORTED;
    st = grow(s->d_nodes, s->nodes_cap, 4 * (size_t)g.nnodes);
    if (st) return st;
    RT_TRY(launch_rh_build(g, r, mode != RT_CLOSEST_PRIMARY, s->d_nodes, hs));
    if (kind == RT_ACCEL_RAYHIERARCHY) {
      // LDS-staged blocks hold at most 1024 rays; queue entries pack the triangle id in 26 bits.
      if (g.rb > 1024 || ntri >= (1u << 26)) return RT_ERR_UNSUPPORTED;
      const bool repair = mode == RT_CLOSEST_PRIMARY && s->desc.repair_run > 0;
      // Work units of at most `per` passing clusters (16 for closest hits, 4 for occlusion by
      // default): blocks whose culls pass many clusters (wide origin boxes at silhouettes)
      // become many units that the persistent kernel spreads over the whole GPU.
      const int per_set = closest ? s->clusters_closest : s->clusters_occluded;
      const int per = per_set > 0 ? per_set : (closest ? 16 : 4);
      const size_t max_units = (size_t)g.ntop * ((s->nclusters + per - 1) / per);
      st = grow(s->d_clist, s->clist_cap, (size_t)g.ntop * s->nclusters);
      if (st) return st;
      st = grow(s->d_units, s->units_cap, std::max<size_t>(max_units, 1));
      if (st) return st;
      st = grow(s->d_ucount, s->ucount_cap, 1);
      if (st) return st;
      RT_TRY(launch_fill_u32(s->d_ucount, 0u, 1, hs));
      if (repair) {
        st = grow(s->d_repkeys, s->rep_cap, n);
        if (st) return st;
        st = grow(s->d_replist, s->replist_cap, n + 1);
        if (st) return st;
        RT_TRY(launch_fill_u32(s->d_replist, 0u, 1, hs));
        RT_TRY(launch_fill_u32(s->d_repkeys, 0xffffffffu, n, hs));
      }
      if (closest) RT_TRY(launch_fill_u64(s->d_keys, ~0ull, n, hs));
      else RT_TRY(launch_fill_u8(occ, 0, n, hs));
      const SortedTris st{s->d_v9s, s->d_perm, s->d_cbox, ntri, s->nclusters};
      const RhWork w{s->d_clist, s->d_units, s->d_ucount, per,
                     (int)std::min<size_t>(std::max<size_t>(max_units, 1), 2048)};
      RT_TRY(launch_rh_clusters(g, mode, s->desc.node, st, w, s->d_nodes, hs));
      const int slot = timing_begin(s, mode, hs);  // the dominant kernel
      RT_TRY(launch_rh_units(g, r, mode, s->desc.node, st, w, s->d_nodes, s->d_keys,
                             repair ? s->d_repkeys : nullptr, occ, hs));
      timing_end(s, slot, hs);
      if (closest) {
        HitOut out{hits->id, hits->t, hits->u, hits->v, hits->dpdu, hits->dpdv};
        RT_TRY(launch_finalize(r, s->d_v9, s->d_uv6, s->d_keys, repair ? s->d_repkeys : nullptr,
                               s->d_replist, out, hs));
        if (repair)
          RT_TRY(launch_repair(g, r, s->desc.node, s->d_v9, s->d_uv6, ntri, s->desc.repair_run,
                               s->d_nodes, s->d_replist, out, hs));
        return RT_OK;
      }
      return RT_OK;
    }
    // ray-bvh
    if (!s->has_bvh) return RT_ERR_STATE;
    const int tchunks = pick_tchunks(g.ntop, s->nroots, (uint32_t)s->tile_roots);
    if (tchunks > 1) {
      if (closest) RT_TRY(launch_fill_u64(s->d_keys, ~0ull, n, hs));
      else RT_TRY(launch_fill_u8(occ, 0, n, hs));
    }
    const int slot = timing_begin(s, mode, hs);
    RT_TRY(launch_bvh_trace(g, r, mode, s->desc.node, s->d_v9, s->d_bvh, s->d_roots, s->nroots,
                            s->d_prim_order, s->tile_roots, tchunks, s->d_nodes, s->d_keys, occ,
                            hs));
    timing_end(s, slot, hs);
  }
  if (closest) {
    HitOut out{hits->id, hits->t, hits->u, hits->v, hits->dpdu, hits->dpdv};
    RT_TRY(launch_finalize(r, s->d_v9, s->d_uv6, s->d_keys, nullptr, nullptr, out, hs));
  }
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

int rt_scene_set_param(rt_scene* s, int32_t param, int32_t value) {
  if (!s || value < 0) return RT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (param == RT_PARAM_CLUSTERS_CLOSEST) s->clusters_closest = value;
  else if (param == RT_PARAM_CLUSTERS_OCCLUDED) s->clusters_occluded = value;
  else return RT_ERR_INVALID;
  return RT_OK;
}

int rt_scene_set_timing(rt_scene* s, int32_t enable) {
  if (!s) return RT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  s->timing = enable != 0;
  s->timed_used = 0;
  return RT_OK;
}

int rt_scene_kernel_time(rt_scene* s, int32_t mode, float* total_ms, int32_t* launches) {
  if (!s || !total_ms || !launches) return RT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  float sum = 0.f;
  int32_t n = 0;
  for (size_t i = 0; i < s->timed_used; ++i) {
    const rt_scene::Timed& t = s->timed[i];
    if (t.mode != mode) continue;
    RT_TRY(hipEventSynchronize(t.b));
    float ms = 0.f;
    RT_TRY(hipEventElapsedTime(&ms, t.a, t.b));
    sum += ms;
    ++n;
  }
  *total_ms = sum;
  *launches = n;
  return RT_OK;
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
    default: return "unknown status";
  }
}

}  // extern "C"
