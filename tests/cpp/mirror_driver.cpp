// mirror_driver.cpp -- test-only C entry point that drives the pbrt accelerator layer
// (pbrt-v2_amd/libpbrt_accel.so) the way pbrt's breadth-first renderer does: MakeAccelerator
// from a ParamSet, Preprocess with the task image, then batched Intersect / IntersectP on
// pbrt rays.  The GPU tests call it through ctypes and compare with the oracle.
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../pbrt-v2_amd/host/breadth_first.h"
#include "../../pbrt-v2_amd/host/pbrt_accel.h"

using namespace pbrt_amd;

namespace {
std::string g_err;
[[noreturn]] void throw_severe(const std::string& msg) { throw std::runtime_error(msg); }
}  // namespace

extern "C" {

const char* mirror_last_error() { return g_err.c_str(); }

// accel: "naive", "rayhierarchy" or "ray-bvh"; params as "name=value" pairs separated by ';'
// (ints, or strings for node/method/splitmethod).  The triangles become two meshes: the first
// split triangles with primitiveId 7 / shapeId 70, the rest with 8 / 80.
// mode 0: primary Intersect, 1: bounce Intersect (hit_io selects rays), 2: IntersectP.
// ray8: n x 8 floats (ox oy oz dx dy dz mint maxt).  Outputs per ray: tri (-1 for no hit),
// prim, maxt after the call, uv (2), dpdu/dpdv (6), occ ('0'/'1').  calls > 1 repeats the
// whole batch on fresh copies of the rays (staging reuse), the outputs come from the last.
int mirror_run(const char* accel, const char* params, const float* v9, const float* uv6,
               uint32_t ntri, uint32_t split, int xres, int yres, int spp, const float* ray8,
               uint32_t n, int mode, int calls, uint8_t* hit_io, int32_t* tri, int32_t* prim,
               float* maxt, float* uv, float* dp, char* occ) {
  SetSevereHandler(throw_severe);
  try {
    ParamSet ps;
    std::string p = params ? params : "";
    size_t pos = 0;
    while (pos < p.size()) {
      size_t end = p.find(';', pos);
      if (end == std::string::npos) end = p.size();
      const std::string kv = p.substr(pos, end - pos);
      pos = end + 1;
      const size_t eq = kv.find('=');
      if (eq == std::string::npos) continue;
      const std::string k = kv.substr(0, eq), v = kv.substr(eq + 1);
      if (k == "node" || k == "method" || k == "splitmethod")
        ps.AddString(k, v);
      else if (k == "onGPU" || k == "sort")
        ps.AddBool(k, v == "true" || v == "1");
      else
        ps.AddInt(k, std::stoi(v));
    }
    std::vector<TriangleMesh> meshes(split > 0 && split < ntri ? 2 : 1);
    for (uint32_t t = 0; t < ntri; ++t) {
      TriangleMesh& m = meshes[(meshes.size() == 2 && t >= split) ? 1 : 0];
      for (int k = 0; k < 3; ++k) {
        m.vi.push_back((int)m.p.size());
        m.p.emplace_back(v9[9 * t + 3 * k], v9[9 * t + 3 * k + 1], v9[9 * t + 3 * k + 2]);
        if (uv6) {
          m.uv.push_back(uv6[6 * t + 2 * k]);
          m.uv.push_back(uv6[6 * t + 2 * k + 1]);
        }
      }
    }
    meshes[0].primitiveId = 7;
    meshes[0].shapeId = 70;
    if (meshes.size() == 2) {
      meshes[1].primitiveId = 8;
      meshes[1].shapeId = 80;
    }
    std::unique_ptr<Aggregate> a(MakeAccelerator(accel, meshes, ps));
    if (xres > 0) a->Preprocess(xres, yres, (unsigned)spp, 1, 1);
    std::vector<RayDifferential> rays(n);
    std::vector<Ray> shadow(mode == 2 ? n : 0);  // IntersectP takes a Ray array
    std::vector<Intersection> isect(n);
    std::unique_ptr<bool[]> hit(new bool[n]);
    for (int c = 0; c < calls; ++c) {
      for (uint32_t i = 0; i < n; ++i) {
        const float* q = ray8 + 8 * (size_t)i;
        rays[i] = RayDifferential(Point(q[0], q[1], q[2]), Vector(q[3], q[4], q[5]), q[6], q[7]);
        if (mode == 2) shadow[i] = Ray(Point(q[0], q[1], q[2]), Vector(q[3], q[4], q[5]), q[6], q[7]);
        isect[i] = Intersection();
        hit[i] = hit_io[i] != 0;
      }
      if (mode == 0)
        a->Intersect(rays.data(), isect.data(), hit.get(), n);
      else if (mode == 1)
        a->Intersect(rays.data(), isect.data(), hit.get(), n, 1);
      else
        a->IntersectP(shadow.data(), occ, n, hit.get());
    }
    for (uint32_t i = 0; i < n; ++i) {
      if (mode == 2) continue;
      hit_io[i] = hit[i] ? 1 : 0;
      tri[i] = hit[i] ? isect[i].triangle : -1;
      prim[i] = hit[i] ? (int32_t)isect[i].primitiveId : -1;
      maxt[i] = rays[i].maxt;
      uv[2 * i] = isect[i].dg.u;
      uv[2 * i + 1] = isect[i].dg.v;
      const Vector& du = isect[i].dg.dpdu;
      const Vector& dv = isect[i].dg.dpdv;
      const float d6[6] = {du.x, du.y, du.z, dv.x, dv.y, dv.z};
      std::memcpy(dp + 6 * (size_t)i, d6, sizeof d6);
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// Breadth-first render of the triangles (one mesh) with accelerator `accel`: image into rgb
// (3 floats per pixel); stats = {tasks, nx, ny, camera rays, shadow rays, bounce rays}.
int mirror_render(const char* accel, const char* params, const float* v9, uint32_t ntri,
                  int xres, int yres, int spp, const float* origin, float fov_deg,
                  const float* light, float intensity, float albedo, int max_depth,
                  float* rgb, int64_t* stats, const char* pfm_path) {
  SetSevereHandler(throw_severe);
  try {
    ParamSet ps;
    std::string p = params ? params : "";
    if (p.rfind("node=", 0) == 0) ps.AddString("node", p.substr(5));
    std::vector<TriangleMesh> meshes(1);
    for (uint32_t t = 0; t < ntri; ++t)
      for (int k = 0; k < 3; ++k) {
        meshes[0].vi.push_back((int)meshes[0].p.size());
        meshes[0].p.emplace_back(v9[9 * t + 3 * k], v9[9 * t + 3 * k + 1], v9[9 * t + 3 * k + 2]);
      }
    std::unique_ptr<Aggregate> a(MakeAccelerator(accel, meshes, ps));
    RenderSettings rs;
    rs.xres = xres;
    rs.yres = yres;
    rs.spp = spp;
    rs.origin = Point(origin[0], origin[1], origin[2]);
    rs.fov_deg = fov_deg;
    rs.light = Point(light[0], light[1], light[2]);
    rs.intensity = intensity;
    rs.albedo = albedo;
    rs.max_depth = max_depth;
    std::vector<float> img;
    const RenderStats st = RenderBreadthFirst(*a, meshes, rs, &img);
    std::memcpy(rgb, img.data(), sizeof(float) * img.size());
    stats[0] = st.tasks;
    stats[1] = st.nx;
    stats[2] = st.ny;
    stats[3] = (int64_t)st.camera_rays;
    stats[4] = (int64_t)st.shadow_rays;
    stats[5] = (int64_t)st.bounce_rays;
    if (pfm_path && !WritePFM(pfm_path, xres, yres, img)) throw std::runtime_error("PFM write");
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

}  // extern "C"
