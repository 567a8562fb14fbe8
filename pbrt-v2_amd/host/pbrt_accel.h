// pbrt_accel.h -- pbrt's batched Aggregate surface on top of the MI355X C ABI.
//
// Host-side mirror of the three OpenCL accelerators of the reference, with the same names,
// parameters, argument meaning and unpack conventions:
//   NaiveAccel     src/accelerators/naive.h:16-109, naive.cpp:18-445
//   RayHieararchy  src/accelerators/rayhierarchy.h:18-212, rayhierarchy.cpp:35-1348
//   RayBVH         src/accelerators/rayBVH.h:20-212, rayBVH.cpp:144-1344
//   MakeAccelerator src/core/api.cpp:576-596
// Only the geometry types the accelerators touch are restated (Ray, RayDifferential,
// Intersection with its DifferentialGeometry, triangle meshes, ParamSet lookups).  Device
// scenes stay resident; every batched call copies the rays in, runs the HIP traversal through
// include/pbrt_amd_rt.h and unpacks the hits on the host as the reference does.
#pragma once
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pbrt_amd_rt.h"

namespace pbrt_amd {

struct Vector {
  float x = 0, y = 0, z = 0;
  Vector() = default;
  Vector(float a, float b, float c) : x(a), y(b), z(c) {}
};
struct Point {
  float x = 0, y = 0, z = 0;
  Point() = default;
  Point(float a, float b, float c) : x(a), y(b), z(c) {}
  Point operator+(const Vector& v) const { return Point(x + v.x, y + v.y, z + v.z); }
};
inline Vector operator*(const Vector& v, float f) { return Vector(v.x * f, v.y * f, v.z * f); }

// geometry.h Ray: mint/maxt are mutable, the accelerators shrink maxt to the hit distance.
struct Ray {
  Point o;
  Vector d;
  mutable float mint = 0.f, maxt = INFINITY;
  float time = 0.f;
  int depth = 0;
  Ray() = default;
  Ray(const Point& o_, const Vector& d_, float start, float end = INFINITY, float t = 0.f,
      int dep = 0)
      : o(o_), d(d_), mint(start), maxt(end), time(t), depth(dep) {}
  Point operator()(float t) const { return o + d * t; }
};
struct RayDifferential : Ray {
  bool hasDifferentials = false;
  Point rxOrigin, ryOrigin;
  Vector rxDirection, ryDirection;
  RayDifferential() = default;
  RayDifferential(const Point& o_, const Vector& d_, float start, float end = INFINITY)
      : Ray(o_, d_, start, end) {}
};

struct DifferentialGeometry {
  Point p;
  Vector dpdu, dpdv;
  float u = 0.f, v = 0.f;
};
// intersection.h Intersection, reduced to what the accelerators fill (rayhierarchy.cpp:1118-1126).
struct Intersection {
  DifferentialGeometry dg;
  int32_t triangle = -1;     // index of the hit triangle in the accelerator's input order
  uint32_t primitiveId = 0;  // id of the owning primitive (mesh)
  int32_t shapeId = 0;
  float rayEpsilon = 0.f;
};

// A world-space triangle mesh (trianglemesh.h): vertex positions, index triples, optional UVs.
struct TriangleMesh {
  std::vector<Point> p;
  std::vector<int> vi;  // 3 per triangle
  std::vector<float> uv;  // 2 per vertex or empty (default (0,0),(1,0),(1,1) per triangle)
  uint32_t primitiveId = 0;
  int32_t shapeId = 0;
};

// paramset.h lookups used by the accelerator factories.
class ParamSet {
 public:
  void AddInt(const std::string& n, int v) { ints_[n] = v; }
  void AddBool(const std::string& n, bool v) { bools_[n] = v; }
  void AddString(const std::string& n, const std::string& v) { strings_[n] = v; }
  int FindOneInt(const std::string& n, int d) const {
    auto it = ints_.find(n);
    return it == ints_.end() ? d : it->second;
  }
  bool FindOneBool(const std::string& n, bool d) const {
    auto it = bools_.find(n);
    return it == bools_.end() ? d : it->second;
  }
  std::string FindOneString(const std::string& n, const std::string& d) const {
    auto it = strings_.find(n);
    return it == strings_.end() ? d : it->second;
  }
  // "type name=value" entries separated by ';', by type then name (for reports and tests).
  std::string Describe() const {
    std::string s;
    for (const auto& kv : ints_) s += "integer " + kv.first + "=" + std::to_string(kv.second) + ";";
    for (const auto& kv : bools_) s += "bool " + kv.first + "=" + (kv.second ? "true;" : "false;");
    for (const auto& kv : strings_) s += "string " + kv.first + "=" + kv.second + ";";
    return s;
  }

 private:
  std::map<std::string, int> ints_;
  std::map<std::string, bool> bools_;
  std::map<std::string, std::string> strings_;
};

// Error reporting of error.cpp: Severe() aborts the process by default, as the reference does;
// an embedding may install a handler that throws instead.
using SevereHandler = void (*)(const std::string& msg);
void SetSevereHandler(SevereHandler h);
[[noreturn]] void Severe(const std::string& msg);

// The batched part of pbrt's Aggregate interface that the breadth-first renderer and the
// multi-ray integrators call (primitive.h:125-133, scene.h:52-99).
class Aggregate {
 public:
  virtual ~Aggregate();
  // Closest hits of primary camera rays (every ray traced).
  virtual void Intersect(const RayDifferential* r, Intersection* in, bool* hit,
                         unsigned count) = 0;
  // Closest hits of bounce rays: on input hit[i] marks the valid rays.
  virtual void Intersect(const RayDifferential* r, Intersection* in, bool* hit, unsigned count,
                         int bounce) = 0;
  // Shadow rays: occluded[i] = '1' when ray i is valid (hit[i]) and blocked, else '0'.
  virtual void IntersectP(const Ray* r, char* occluded, size_t count, const bool* hit) = 0;
  // Largest ray batch one call takes (rayhierarchy.cpp:309).
  virtual unsigned MaxRaysPerCall() { return 409600; }
  // Ray-image shape of one renderer task: (xResolution / nx * spp) x (yResolution / ny)
  // (rayhierarchy.cpp:258-265).
  virtual void Preprocess(int xResolution, int yResolution, unsigned samplesPerPixel, int nx,
                          int ny);
  uint32_t TriangleCount() const { return ntri_; }

 protected:
  // sort_max_prims > 0 reorders the triangles into BVHAccel order with leaves of that many
  // primitives (RayHieararchy's "sort"); Intersection::triangle stays the input index.
  // split_method (RT_SPLIT_*) is BVHAccel's for that order and for ray-bvh's CPU BVH.
  Aggregate(const std::vector<TriangleMesh>& meshes, const rt_accel_desc& desc,
            uint32_t sort_max_prims = 0, int32_t split_method = RT_SPLIT_SAH);
  // Ray image for `count` rays: the preprocessed task image (xResolution * spp) x yResolution,
  // grown by whole rows when the batch is larger; one row for naive or without Preprocess.
  void ImageShape(unsigned count, int* w, int* h) const;
  void Closest(const RayDifferential* r, Intersection* in, bool* hit, unsigned count, bool bounce);
  void Occluded(const Ray* r, char* occluded, size_t count, const bool* hit);

  rt_scene* scene_ = nullptr;
  rt_accel_desc desc_{};
  uint32_t ntri_ = 0;
  std::vector<uint32_t> tri_prim_;   // triangle -> primitiveId
  std::vector<int32_t> tri_shape_;   // triangle -> shapeId
  std::vector<uint32_t> tri_orig_;   // with "sort": device triangle -> input index
  int32_t split_ = RT_SPLIT_SAH;     // BVHAccel split method (sort order, ray-bvh BVH)
  int xres_ = 0, yres_ = 0;
  unsigned spp_ = 1;

 private:
  struct Staging;
  std::unique_ptr<Staging> st_;
};

class NaiveAccel : public Aggregate {
 public:
  NaiveAccel(const std::vector<TriangleMesh>& meshes, bool onGPU, const std::string& method);
  void Intersect(const RayDifferential* r, Intersection* in, bool* hit, unsigned count) override;
  // naive.h:53 has one Intersect with bounce = 0 by default: bounce 0 traces camera rays (every
  // ray), any other value bounce rays (the rays with hit[i] set).
  void Intersect(const RayDifferential* r, Intersection* in, bool* hit, unsigned count,
                 int bounce) override;
  void IntersectP(const Ray* r, char* occluded, size_t count, const bool* hit) override;
};

class RayHieararchy : public Aggregate {
 public:
  RayHieararchy(const std::vector<TriangleMesh>& meshes, bool onGPU, int chunkX, int chunkY,
                int height, const std::string& node, bool sortVert, const std::string& sm,
                int maxBVHPrim, unsigned repairRun);
  void Intersect(const RayDifferential* r, Intersection* in, bool* hit, unsigned count) override;
  void Intersect(const RayDifferential* r, Intersection* in, bool* hit, unsigned count,
                 int bounce) override;
  void IntersectP(const Ray* r, char* occluded, size_t count, const bool* hit) override;
};

class RayBVH : public Aggregate {
 public:
  RayBVH(const std::vector<TriangleMesh>& meshes, bool onGPU, int chunkX, int chunkY,
         int height, const std::string& node, const std::string& sm, int maxBVHPrim,
         int BVHheight);
  void Intersect(const RayDifferential* r, Intersection* in, bool* hit, unsigned count) override;
  void Intersect(const RayDifferential* r, Intersection* in, bool* hit, unsigned count,
                 int bounce) override;
  void IntersectP(const Ray* r, char* occluded, size_t count, const bool* hit) override;
};

NaiveAccel* CreateNaiveAccelerator(const std::vector<TriangleMesh>& meshes, const ParamSet& ps);
RayHieararchy* CreateRayHieararchy(const std::vector<TriangleMesh>& meshes, const ParamSet& ps);
RayBVH* CreateRayBVH(const std::vector<TriangleMesh>& meshes, const ParamSet& ps);
// "naive", "rayhierarchy", "ray-bvh"; anything else is Severe (api.cpp:576-596).
Aggregate* MakeAccelerator(const std::string& name, const std::vector<TriangleMesh>& meshes,
                           const ParamSet& ps);

// The batched entry points of pbrt's Scene (scene.h:52-99, scene.cpp:55-121): every call is
// routed to the accelerator by its type (dynamic_cast to RayHieararchy, RayBVH, NaiveAccel);
// any other aggregate is Severe.  Scene::Intersect(..., bounce) hands NaiveAccel bounce + 1,
// because NaiveAccel takes bounce 0 for camera rays (scene.cpp:114).
class Scene {
 public:
  explicit Scene(Aggregate* accel) : aggregate(accel) {}
  void Intersect(const RayDifferential* ray, Intersection* isect, bool* hit, int count,
                 unsigned samplesPerPixel) const;
  void Intersect(const RayDifferential* ray, Intersection* isect, bool* hit, int count,
                 unsigned samplesPerPixel, int bounce) const;
  void IntersectP(const Ray* ray, char* occluded, size_t count, const bool* hit) const;
  unsigned MaxRaysPerCall() const;
  // Scene::Preprocess(camera, spp, nx, ny) with the camera's film resolution.
  void Preprocess(int xResolution, int yResolution, unsigned samplesPerPixel, int nx,
                  int ny) const;

  Aggregate* aggregate;
};

}  // namespace pbrt_amd
