// pbrt_accel.cpp -- NaiveAccel / RayHieararchy / RayBVH over the C ABI (see pbrt_accel.h).
//
// Per batched call the reference packs the rays on the host, uploads them, runs its OpenCL
// kernels and unpacks tHit / index / (tu,tv) / (dpdu,dpdv) into pbrt Intersections
// (rayhierarchy.cpp:823-853 and 1100-1131, naive.cpp:234-330).  Here the triangles stay
// resident on the device; the rays go through pinned staging buffers into a structure-of-arrays
// ray image and one ordered HIP stream carries upload, traversal and readback.
#include "pbrt_accel.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace pbrt_amd {

namespace {

SevereHandler g_severe = nullptr;

void check(int st, const char* what) {
  if (st != RT_OK) Severe(std::string(what) + ": " + rt_error_string(st));
}
void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) Severe(std::string(what) + ": " + hipGetErrorString(e));
}

// One device context per process, shared by every accelerator (the reference keeps one
// OpenCL context per accelerator; a HIP context is per process already).
rt_ctx* shared_ctx() {
  static std::mutex mu;
  static rt_ctx* ctx = nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!ctx) check(rt_ctx_create(0, &ctx), "rt_ctx_create");
  return ctx;
}

// Flattens the meshes into the AoS triangle arrays of the reference (rayhierarchy.cpp:162-199,
// naive.cpp:59-101): v9 = p1,p2,p3; uv6 = per-vertex UVs or (0,0),(1,0),(1,1).
void flatten(const std::vector<TriangleMesh>& meshes, std::vector<float>* v9,
             std::vector<float>* uv6, std::vector<uint32_t>* prim, std::vector<int32_t>* shape) {
  for (const TriangleMesh& m : meshes) {
    if (m.vi.size() % 3 != 0) Severe("triangle mesh index count is not a multiple of 3");
    if (!m.uv.empty() && m.uv.size() != 2 * m.p.size())
      Severe("triangle mesh UV count does not match its vertex count");
    const size_t nt = m.vi.size() / 3;
    for (size_t t = 0; t < nt; ++t) {
      for (int k = 0; k < 3; ++k) {
        const int i = m.vi[3 * t + k];
        if (i < 0 || (size_t)i >= m.p.size()) Severe("triangle mesh vertex index out of range");
        v9->push_back(m.p[i].x);
        v9->push_back(m.p[i].y);
        v9->push_back(m.p[i].z);
        if (m.uv.empty()) {
          static const float kDefault[6] = {0.f, 0.f, 1.f, 0.f, 1.f, 1.f};
          uv6->push_back(kDefault[2 * k]);
          uv6->push_back(kDefault[2 * k + 1]);
        } else {
          uv6->push_back(m.uv[2 * i]);
          uv6->push_back(m.uv[2 * i + 1]);
        }
      }
      prim->push_back(m.primitiveId);
      shape->push_back(m.shapeId);
    }
  }
}

rt_accel_desc make_desc(int32_t kind) {
  rt_accel_desc d;
  check(rt_accel_desc_default(kind, &d), "rt_accel_desc_default");
  return d;
}

int32_t node_kind(const std::string& node) {
  if (node == "box_dir") return RT_NODE_BOX_DIR;
  if (node == "ia") return RT_NODE_IA;
  if (node == "box_uv") return RT_NODE_BOX_UV;
  Severe("Unknown accelarator node type " + node);
}

}  // namespace

void SetSevereHandler(SevereHandler h) { g_severe = h; }

void Severe(const std::string& msg) {
  if (g_severe) g_severe(msg);
  std::fprintf(stderr, "Fatal error: %s\n", msg.c_str());
  std::abort();
}

// Device and pinned host buffers of one accelerator, grown to the largest batch seen.
struct Aggregate::Staging {
  hipStream_t stream = nullptr;
  size_t cap = 0;
  // device: 8 ray floats + valid, hits id/t/u/v/dpdu/dpdv, occlusion flags
  float* d_ray = nullptr;
  uint8_t* d_valid = nullptr;
  int32_t* d_id = nullptr;
  float* d_hit = nullptr;  // planes of n floats: t, u, v; then dpdu, dpdv (3n each, xyz per ray)
  uint8_t* d_occ = nullptr;
  // pinned host twins
  float* h_ray = nullptr;
  uint8_t* h_valid = nullptr;
  int32_t* h_id = nullptr;
  float* h_hit = nullptr;
  uint8_t* h_occ = nullptr;

  ~Staging() { release(); }
  void release() {
    (void)hipFree(d_ray);
    (void)hipFree(d_valid);
    (void)hipFree(d_id);
    (void)hipFree(d_hit);
    (void)hipFree(d_occ);
    (void)hipHostFree(h_ray);
    (void)hipHostFree(h_valid);
    (void)hipHostFree(h_id);
    (void)hipHostFree(h_hit);
    (void)hipHostFree(h_occ);
    d_ray = h_ray = d_hit = h_hit = nullptr;
    d_valid = h_valid = d_occ = h_occ = nullptr;
    d_id = h_id = nullptr;
    cap = 0;
  }
  void reserve(size_t n) {
    if (!stream) check_hip(hipStreamCreate(&stream), "hipStreamCreate");
    if (n <= cap) return;
    release();
    check_hip(hipMalloc(&d_ray, 8 * n * sizeof(float)), "hipMalloc rays");
    check_hip(hipMalloc(&d_valid, n), "hipMalloc valid");
    check_hip(hipMalloc(&d_id, n * sizeof(int32_t)), "hipMalloc hit ids");
    check_hip(hipMalloc(&d_hit, 9 * n * sizeof(float)), "hipMalloc hits");
    check_hip(hipMalloc(&d_occ, n), "hipMalloc occlusion");
    check_hip(hipHostMalloc(&h_ray, 8 * n * sizeof(float), 0), "hipHostMalloc rays");
    check_hip(hipHostMalloc(&h_valid, n, 0), "hipHostMalloc valid");
    check_hip(hipHostMalloc(&h_id, n * sizeof(int32_t), 0), "hipHostMalloc hit ids");
    check_hip(hipHostMalloc(&h_hit, 9 * n * sizeof(float), 0), "hipHostMalloc hits");
    check_hip(hipHostMalloc(&h_occ, n, 0), "hipHostMalloc occlusion");
    cap = n;
  }
  rt_rays rays(int w, int h) const {
    const size_t n = cap;
    rt_rays r;
    r.ox = d_ray;
    r.oy = d_ray + n;
    r.oz = d_ray + 2 * n;
    r.dx = d_ray + 3 * n;
    r.dy = d_ray + 4 * n;
    r.dz = d_ray + 5 * n;
    r.tmin = d_ray + 6 * n;
    r.tmax = d_ray + 7 * n;
    r.valid = d_valid;
    r.width = w;
    r.height = h;
    return r;
  }
  void upload(size_t n) {
    for (int c = 0; c < 8; ++c)
      check_hip(hipMemcpyAsync(d_ray + c * cap, h_ray + c * cap, n * sizeof(float),
                               hipMemcpyHostToDevice, stream),
                "ray upload");
    check_hip(hipMemcpyAsync(d_valid, h_valid, n, hipMemcpyHostToDevice, stream), "valid upload");
  }
};

Aggregate::Aggregate(const std::vector<TriangleMesh>& meshes, const rt_accel_desc& desc,
                     uint32_t sort_max_prims, int32_t split_method)
    : desc_(desc), split_(split_method), st_(new Staging) {
  std::vector<float> v9, uv6;
  flatten(meshes, &v9, &uv6, &tri_prim_, &tri_shape_);
  ntri_ = (uint32_t)tri_prim_.size();
  if (ntri_ == 0) Severe("accelerator built over a scene without triangles");
  if (sort_max_prims > 0) {
    // RayHieararchy "sort": the primitives are swapped into BVHAccel order before they are
    // flattened (rayhierarchy.cpp:152-155), so ids, and the smallest-id rule among hits at
    // equal t, follow that order.
    tri_orig_.resize(ntri_);
    check(rt_bvh_order(v9.data(), ntri_, sort_max_prims, split_, tri_orig_.data()),
          "rt_bvh_order");
    std::vector<float> v9s(v9.size()), uv6s(uv6.size());
    std::vector<uint32_t> prim(ntri_);
    std::vector<int32_t> shape(ntri_);
    for (uint32_t k = 0; k < ntri_; ++k) {
      const uint32_t o = tri_orig_[k];
      std::copy(v9.begin() + 9 * (size_t)o, v9.begin() + 9 * (size_t)o + 9, v9s.begin() + 9 * (size_t)k);
      std::copy(uv6.begin() + 6 * (size_t)o, uv6.begin() + 6 * (size_t)o + 6,
                uv6s.begin() + 6 * (size_t)k);
      prim[k] = tri_prim_[o];
      shape[k] = tri_shape_[o];
    }
    v9.swap(v9s);
    uv6.swap(uv6s);
    tri_prim_.swap(prim);
    tri_shape_.swap(shape);
  }
  check(rt_scene_create(shared_ctx(), v9.data(), uv6.data(), ntri_, &desc_, &scene_),
        "rt_scene_create");
  if (desc_.kind == RT_ACCEL_RAYBVH) {
    // RayBVH builds BVHAccel over its triangles on the host before the first trace
    // (rayBVH.cpp:204-277) and hands the flattened nodes to the device.
    rt_gpu_node* nodes = nullptr;
    uint32_t nnodes = 0, nroots = 0;
    uint32_t *roots = nullptr, *order = nullptr;
    check(rt_raybvh_build_handoff(v9.data(), ntri_, (uint32_t)desc_.max_bvh_prim, split_,
                                  (uint32_t)desc_.bvh_height, &nodes, &nnodes, &roots, &nroots,
                                  &order),
          "rt_raybvh_build_handoff");
    const int st = rt_scene_set_bvh(scene_, nodes, nnodes, roots, nroots, order);
    rt_raybvh_free_handoff(nodes, roots, order);
    check(st, "rt_scene_set_bvh");
  }
}

Aggregate::~Aggregate() {
  if (st_ && st_->stream) (void)hipStreamSynchronize(st_->stream);
  if (scene_) rt_scene_destroy(scene_);
  if (st_ && st_->stream) (void)hipStreamDestroy(st_->stream);
}

void Aggregate::Preprocess(int xResolution, int yResolution, unsigned samplesPerPixel, int nx,
                           int ny) {
  if (nx <= 0 || ny <= 0) Severe("Preprocess: task counts must be positive");
  xres_ = xResolution / nx;
  yres_ = yResolution / ny;
  spp_ = samplesPerPixel;
}

void Aggregate::ImageShape(unsigned count, int* w, int* h) const {
  const long long W = (long long)xres_ * spp_;
  if (desc_.kind == RT_ACCEL_NAIVE || W <= 0 || yres_ <= 0) {
    *w = (int)count;  // no ray image: one row
    *h = 1;
  } else if ((long long)count <= W * yres_) {
    *w = (int)W;  // the task image of the renderer (rayhierarchy.cpp:377-389)
    *h = yres_;
  } else {
    *w = (int)W;
    *h = (int)((count + W - 1) / W);
  }
}

// Shared closest-hit path.  bounce == false: every ray traced (primary camera rays);
// bounce == true: hit[i] selects the rays.  A batch that does not fill the ray image is traced
// padded, the padding rays marked invalid; with every real ray valid the secondary mode gives
// the primary mode's hits.
void Aggregate::Closest(const RayDifferential* r, Intersection* in, bool* hit, unsigned count,
                        bool bounce) {
  if (count == 0) return;
  int w = 0, h = 0;
  ImageShape(count, &w, &h);
  const size_t n = (size_t)w * h;
  const int32_t mode = (bounce || n != count) ? RT_CLOSEST_SECONDARY : RT_CLOSEST_PRIMARY;
  const bool naive = desc_.kind == RT_ACCEL_NAIVE;
  st_->reserve(n);
  Staging& s = *st_;
  const size_t c = s.cap;
  for (size_t k = 0; k < n; ++k) {
    const bool real = k < count;
    const Ray* ray = real ? &r[k] : nullptr;
    s.h_ray[k] = real ? ray->o.x : 0.f;
    s.h_ray[c + k] = real ? ray->o.y : 0.f;
    s.h_ray[2 * c + k] = real ? ray->o.z : 0.f;
    s.h_ray[3 * c + k] = real ? ray->d.x : 0.f;
    s.h_ray[4 * c + k] = real ? ray->d.y : 0.f;
    s.h_ray[5 * c + k] = real ? ray->d.z : 0.f;
    s.h_ray[6 * c + k] = real ? ray->mint : 0.f;
    s.h_ray[7 * c + k] = INFINITY;  // closest-hit rays are traced to infinity (:845)
    bool valid = real && (!bounce || hit[k]);
    // naive.cl marks invalid bounce rays with dir.x = -10 and skips every ray whose dir.x is
    // below -5 (naive.cpp:244-245, naive.cl:96-187)
    if (naive && bounce && real && ray->d.x < -5.f) valid = false;
    s.h_valid[k] = valid ? 1 : 0;
  }
  s.upload(n);
  const rt_rays rays = s.rays(w, h);
  rt_hits hits;
  hits.id = s.d_id;
  hits.t = s.d_hit;
  hits.u = s.d_hit + c;
  hits.v = s.d_hit + 2 * c;
  hits.dpdu = s.d_hit + 3 * c;  // x, y, z of ray i at [3i, 3i+3)
  hits.dpdv = s.d_hit + 6 * c;
  check(rt_trace_closest(scene_, &rays, mode, &hits, s.stream), "rt_trace_closest");
  check_hip(hipMemcpyAsync(s.h_id, s.d_id, count * sizeof(int32_t), hipMemcpyDeviceToHost,
                           s.stream),
            "hit id readback");
  for (int f = 0; f < 3; ++f)
    check_hip(hipMemcpyAsync(s.h_hit + f * c, s.d_hit + f * c, count * sizeof(float),
                             hipMemcpyDeviceToHost, s.stream),
              "hit readback");
  for (int f = 3; f < 9; f += 3)
    check_hip(hipMemcpyAsync(s.h_hit + f * c, s.d_hit + f * c, 3 * count * sizeof(float),
                             hipMemcpyDeviceToHost, s.stream),
              "hit readback");
  check_hip(hipStreamSynchronize(s.stream), "trace");

  // unpack (rayhierarchy.cpp:1100-1131, naive.cpp:290-327)
  for (unsigned i = 0; i < count; ++i) {
    if (bounce && !hit[i]) continue;
    hit[i] = false;
    const int32_t id = s.h_id[i];
    if (id < 0) continue;
    // naive.cpp:298 takes index 0 for "no hit", so a hit on the first triangle is dropped
    if (naive && id == 0) continue;
    const float t = s.h_hit[i];
    Intersection& is = in[i];
    is.dg.p = r[i](t);
    is.dg.u = s.h_hit[c + i];
    is.dg.v = s.h_hit[2 * c + i];
    const float* du = s.h_hit + 3 * c + 3 * (size_t)i;
    const float* dv = s.h_hit + 6 * c + 3 * (size_t)i;
    is.dg.dpdu = Vector(du[0], du[1], du[2]);
    is.dg.dpdv = Vector(dv[0], dv[1], dv[2]);
    is.triangle = tri_orig_.empty() ? id : (int32_t)tri_orig_[id];
    is.shapeId = tri_shape_[id];
    is.primitiveId = tri_prim_[id];
    is.rayEpsilon = 1e-3f * t;
    r[i].maxt = t;
    hit[i] = true;
  }
}

void Aggregate::Occluded(const Ray* r, char* occluded, size_t count, const bool* hit) {
  if (count == 0) return;
  int w = 0, h = 0;
  ImageShape((unsigned)count, &w, &h);
  const size_t n = (size_t)w * h;
  st_->reserve(n);
  Staging& s = *st_;
  const size_t c = s.cap;
  for (size_t k = 0; k < n; ++k) {
    const bool real = k < count;
    const Ray* ray = real ? &r[k] : nullptr;
    s.h_ray[k] = real ? ray->o.x : 0.f;
    s.h_ray[c + k] = real ? ray->o.y : 0.f;
    s.h_ray[2 * c + k] = real ? ray->o.z : 0.f;
    s.h_ray[3 * c + k] = real ? ray->d.x : 0.f;
    s.h_ray[4 * c + k] = real ? ray->d.y : 0.f;
    s.h_ray[5 * c + k] = real ? ray->d.z : 0.f;
    s.h_ray[6 * c + k] = real ? ray->mint : 0.f;
    s.h_ray[7 * c + k] = real ? ray->maxt : 0.f;  // shadow rays keep their extent (:1200)
    s.h_valid[k] = (real && hit[k]) ? 1 : 0;
  }
  s.upload(n);
  const rt_rays rays = s.rays(w, h);
  check(rt_trace_occluded(scene_, &rays, s.d_occ, s.stream), "rt_trace_occluded");
  check_hip(hipMemcpyAsync(s.h_occ, s.d_occ, count, hipMemcpyDeviceToHost, s.stream),
            "occlusion readback");
  check_hip(hipStreamSynchronize(s.stream), "trace");
  for (size_t k = 0; k < count; ++k) occluded[k] = (hit[k] && s.h_occ[k]) ? '1' : '0';
}

// --- NaiveAccel (naive.cpp:18-445) -------------------------------------------------------------

static rt_accel_desc naive_desc(bool onGPU, const std::string& method) {
  if (!onGPU) Severe("naive: onGPU=false (CPU traversal) is not provided; use a CPU aggregate");
  if (method != "TriRay")
    Severe("naive: method \"" + method + "\" is not supported; only \"TriRay\" is");
  return make_desc(RT_ACCEL_NAIVE);
}

NaiveAccel::NaiveAccel(const std::vector<TriangleMesh>& meshes, bool onGPU,
                       const std::string& method)
    : Aggregate(meshes, naive_desc(onGPU, method)) {}
void NaiveAccel::Intersect(const RayDifferential* r, Intersection* in, bool* hit, unsigned count) {
  Closest(r, in, hit, count, false);
}
void NaiveAccel::Intersect(const RayDifferential* r, Intersection* in, bool* hit, unsigned count,
                           int bounce) {
  Closest(r, in, hit, count, bounce != 0);  // naive.cpp:215: bounce 0 is the camera kernel
}
void NaiveAccel::IntersectP(const Ray* r, char* occluded, size_t count, const bool* hit) {
  Occluded(r, occluded, count, hit);
}

// --- RayHieararchy (rayhierarchy.cpp:35-1348) --------------------------------------------------

// BVHAccel's split method names (bvh.cpp:214-221): an unknown name warns and uses "sah".
static int32_t split_kind(const std::string& sm) {
  if (sm == "sah") return RT_SPLIT_SAH;
  if (sm == "middle") return RT_SPLIT_MIDDLE;
  if (sm == "equal") return RT_SPLIT_EQUAL;
  std::fprintf(stderr, "Warning: BVH split method \"%s\" unknown.  Using \"sah\".\n",
               sm.c_str());
  return RT_SPLIT_SAH;
}

static rt_accel_desc rh_desc(int32_t kind, bool onGPU, int chunkX, int chunkY, int height,
                             const std::string& node, int maxBVHPrim, int BVHheight,
                             unsigned repairRun) {
  if (!onGPU) Severe("ray-space accelerators need the GPU (onGPU=false is not provided)");
  rt_accel_desc d = make_desc(kind);
  d.node = node_kind(node);
  d.chunk_x = chunkX;
  d.chunk_y = chunkY;
  d.height = height;
  d.repair_run = (int32_t)repairRun;
  d.max_bvh_prim = maxBVHPrim;
  if (kind == RT_ACCEL_RAYBVH) d.bvh_height = BVHheight;
  return d;
}

RayHieararchy::RayHieararchy(const std::vector<TriangleMesh>& meshes, bool onGPU, int chunkX,
                             int chunkY, int height, const std::string& node, bool sortVert,
                             const std::string& sm, int maxBVHPrim, unsigned repairRun)
    : Aggregate(meshes, rh_desc(RT_ACCEL_RAYHIERARCHY, onGPU, chunkX, chunkY, height, node,
                                maxBVHPrim, 0, repairRun),
                sortVert ? (uint32_t)std::max(maxBVHPrim, 1) : 0u, split_kind(sm)) {}
void RayHieararchy::Intersect(const RayDifferential* r, Intersection* in, bool* hit,
                              unsigned count) {
  Closest(r, in, hit, count, false);
}
void RayHieararchy::Intersect(const RayDifferential* r, Intersection* in, bool* hit,
                              unsigned count, int) {
  Closest(r, in, hit, count, true);
}
void RayHieararchy::IntersectP(const Ray* r, char* occluded, size_t count, const bool* hit) {
  Occluded(r, occluded, count, hit);
}

// --- RayBVH (rayBVH.cpp:144-1344) --------------------------------------------------------------

RayBVH::RayBVH(const std::vector<TriangleMesh>& meshes, bool onGPU, int chunkX, int chunkY,
               int height, const std::string& node, const std::string& sm, int maxBVHPrim,
               int BVHheight)
    : Aggregate(meshes, rh_desc(RT_ACCEL_RAYBVH, onGPU, chunkX, chunkY, height, node,
                                maxBVHPrim, BVHheight, 0),
                0u, split_kind(sm)) {}
void RayBVH::Intersect(const RayDifferential* r, Intersection* in, bool* hit, unsigned count) {
  Closest(r, in, hit, count, false);
}
void RayBVH::Intersect(const RayDifferential* r, Intersection* in, bool* hit, unsigned count,
                       int) {
  Closest(r, in, hit, count, true);
}
void RayBVH::IntersectP(const Ray* r, char* occluded, size_t count, const bool* hit) {
  Occluded(r, occluded, count, hit);
}

// --- factories (naive.cpp:440-445, rayhierarchy.cpp:1329-1348, rayBVH.cpp:1326-1344) ----------

NaiveAccel* CreateNaiveAccelerator(const std::vector<TriangleMesh>& meshes, const ParamSet& ps) {
  const bool onGPU = ps.FindOneBool("onGPU", true);
  const std::string method = ps.FindOneString("method", "TriRay");
  return new NaiveAccel(meshes, onGPU, method);
}

RayHieararchy* CreateRayHieararchy(const std::vector<TriangleMesh>& meshes, const ParamSet& ps) {
  const bool onGPU = ps.FindOneBool("onGPU", true);
  const bool sortVert = ps.FindOneBool("sort", false);
  const int chunkX = ps.FindOneInt("chunkXSize", 4);
  const int chunkY = ps.FindOneInt("chunkYSize", 4);
  const int height = ps.FindOneInt("height", 3);
  const std::string node = ps.FindOneString("node", "box_dir");
  const std::string sm = ps.FindOneString("splitmethod", "sah");
  const int maxBVHPrim = ps.FindOneInt("maxBVHPrim", 1);
  const unsigned repairRun = (unsigned)ps.FindOneInt("repairRun", 1);
  return new RayHieararchy(meshes, onGPU, chunkX, chunkY, height, node, sortVert, sm, maxBVHPrim,
                           repairRun);
}

RayBVH* CreateRayBVH(const std::vector<TriangleMesh>& meshes, const ParamSet& ps) {
  const bool onGPU = ps.FindOneBool("onGPU", true);
  const int chunkX = ps.FindOneInt("chunkXSize", 4);
  const int chunkY = ps.FindOneInt("chunkYSize", 4);
  const int height = ps.FindOneInt("height", 3);
  const int BVHheight = ps.FindOneInt("BVHheight", 3);
  const int maxBVHPrim = ps.FindOneInt("maxBVHPrim", 1);
  const std::string node = ps.FindOneString("node", "box_dir");
  const std::string sm = ps.FindOneString("splitmethod", "sah");
  return new RayBVH(meshes, onGPU, chunkX, chunkY, height, node, sm, maxBVHPrim, BVHheight);
}

Aggregate* MakeAccelerator(const std::string& name, const std::vector<TriangleMesh>& meshes,
                           const ParamSet& ps) {
  if (name == "naive") return CreateNaiveAccelerator(meshes, ps);
  if (name == "rayhierarchy") return CreateRayHieararchy(meshes, ps);
  if (name == "ray-bvh") return CreateRayBVH(meshes, ps);
  Severe("Accelerator \"" + name + "\" unknown.");
}

// --- Scene (scene.h:52-99, scene.cpp:55-121) ------------------------------------------------------

void Scene::Intersect(const RayDifferential* ray, Intersection* isect, bool* hit, int count,
                      unsigned) const {
  if (auto* rh = dynamic_cast<RayHieararchy*>(aggregate)) return rh->Intersect(ray, isect, hit, count);
  if (auto* rb = dynamic_cast<RayBVH*>(aggregate)) return rb->Intersect(ray, isect, hit, count);
  if (auto* na = dynamic_cast<NaiveAccel*>(aggregate)) return na->Intersect(ray, isect, hit, count);
  Severe("Called Intersect with unsupported aggregate!");
}

void Scene::Intersect(const RayDifferential* ray, Intersection* isect, bool* hit, int count,
                      unsigned, int bounce) const {
  if (auto* rh = dynamic_cast<RayHieararchy*>(aggregate))
    return rh->Intersect(ray, isect, hit, count, bounce);
  if (auto* rb = dynamic_cast<RayBVH*>(aggregate))
    return rb->Intersect(ray, isect, hit, count, bounce);
  if (auto* na = dynamic_cast<NaiveAccel*>(aggregate))
    return na->Intersect(ray, isect, hit, count, bounce + 1);
  Severe("Called Intersect with unsupported aggregate!");
}

void Scene::IntersectP(const Ray* ray, char* occluded, size_t count, const bool* hit) const {
  if (auto* rh = dynamic_cast<RayHieararchy*>(aggregate))
    return rh->IntersectP(ray, occluded, count, hit);
  if (auto* rb = dynamic_cast<RayBVH*>(aggregate)) return rb->IntersectP(ray, occluded, count, hit);
  if (auto* na = dynamic_cast<NaiveAccel*>(aggregate))
    return na->IntersectP(ray, occluded, count, hit);
  Severe("Called IntersectP with unsupported aggregate!");
}

unsigned Scene::MaxRaysPerCall() const {
  if (dynamic_cast<RayHieararchy*>(aggregate) || dynamic_cast<RayBVH*>(aggregate) ||
      dynamic_cast<NaiveAccel*>(aggregate))
    return aggregate->MaxRaysPerCall();
  Severe("Called MaxRaysPerCall with unsoppurted aggregate!");
}

void Scene::Preprocess(int xResolution, int yResolution, unsigned samplesPerPixel, int nx,
                       int ny) const {
  if (!aggregate) Severe("Called Preprocess with unsupported aggregate!");
  aggregate->Preprocess(xResolution, yResolution, samplesPerPixel, nx, ny);
}

}  // namespace pbrt_amd
