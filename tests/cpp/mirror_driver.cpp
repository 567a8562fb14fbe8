// mirror_driver.cpp -- test-only C entry point that drives the pbrt accelerator layer
// (pbrt-v2_amd/libpbrt_accel.so) the way pbrt's breadth-first renderer does: MakeAccelerator
// from a ParamSet, Preprocess with the task image, then batched Intersect / IntersectP on
// pbrt rays.  The GPU tests call it through ctypes and compare with the oracle.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../pbrt-v2_amd/host/breadth_first.h"
#include "../../pbrt-v2_amd/host/pbrt_accel.h"
#include "../../pbrt-v2_amd/host/pbrt_scene.h"

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
// mode 0: primary Intersect, 1: bounce Intersect (hit_io selects rays), 2: IntersectP,
// 3: Scene::Intersect, 4: Scene::Intersect with bounce 0, 5: Scene::IntersectP,
// 6: the accelerator's Intersect(..., bounce) with bounce 0.
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
    const bool occl = mode == 2 || mode == 5;
    std::vector<Ray> shadow(occl ? n : 0);  // IntersectP takes a Ray array
    std::vector<Intersection> isect(n);
    std::unique_ptr<bool[]> hit(new bool[n]);
    for (int c = 0; c < calls; ++c) {
      for (uint32_t i = 0; i < n; ++i) {
        const float* q = ray8 + 8 * (size_t)i;
        rays[i] = RayDifferential(Point(q[0], q[1], q[2]), Vector(q[3], q[4], q[5]), q[6], q[7]);
        if (occl) shadow[i] = Ray(Point(q[0], q[1], q[2]), Vector(q[3], q[4], q[5]), q[6], q[7]);
        isect[i] = Intersection();
        hit[i] = hit_io[i] != 0;
      }
      const Scene scene(a.get());
      if (mode == 0)
        a->Intersect(rays.data(), isect.data(), hit.get(), n);
      else if (mode == 1)
        a->Intersect(rays.data(), isect.data(), hit.get(), n, 1);
      else if (mode == 2)
        a->IntersectP(shadow.data(), occ, n, hit.get());
      else if (mode == 3)  // through pbrt's Scene
        scene.Intersect(rays.data(), isect.data(), hit.get(), (int)n, (unsigned)spp);
      else if (mode == 4)  // the path integrator's first bounce: Scene bounce 0
        scene.Intersect(rays.data(), isect.data(), hit.get(), (int)n, (unsigned)spp, 0);
      else if (mode == 5)
        scene.IntersectP(shadow.data(), occ, n, hit.get());
      else  // the accelerator's bounce overload with bounce 0
        a->Intersect(rays.data(), isect.data(), hit.get(), n, 0);
    }
    for (uint32_t i = 0; i < n; ++i) {
      if (occl) continue;
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

// Scene over an aggregate it does not know (none): which = 0 Intersect, 1 Intersect(bounce),
// 2 IntersectP, 3 MaxRaysPerCall.  Returns 1 with the Severe message, 0 if nothing failed.
int mirror_scene_unsupported(int which) {
  SetSevereHandler(throw_severe);
  try {
    const Scene scene(nullptr);
    RayDifferential r;
    Ray sr;
    Intersection is;
    bool hit = true;
    char occ = '0';
    if (which == 0) scene.Intersect(&r, &is, &hit, 1, 1);
    else if (which == 1) scene.Intersect(&r, &is, &hit, 1, 1, 0);
    else if (which == 2) scene.IntersectP(&sr, &occ, 1, &hit);
    else (void)scene.MaxRaysPerCall();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// Parses a pbrt scene file.  v9: room for `cap` triangles, filled with the world-space
// triangles in the order the accelerators flatten them.  info: "key=value" lines (accelerator,
// params, integrator, renderer, sampler, film, spp, fov, camera_to_world, screen, light,
// intensity, max_depth, has_point_light, meshes, skipped, then one "message=" line each).
// Returns the triangle count (which may exceed cap), or -1 with mirror_last_error().
int64_t mirror_parse(const char* path, float* v9, uint64_t cap, char* info, uint64_t info_cap) {
  SetSevereHandler(throw_severe);
  try {
    SceneFile sc;
    ParseSceneFile(path, &sc);
    uint64_t n = 0;
    for (const TriangleMesh& m : sc.meshes)
      for (size_t t = 0; t + 2 < m.vi.size(); t += 3, ++n) {
        if (n >= cap || !v9) continue;
        for (int k = 0; k < 3; ++k) {
          const Point& q = m.p[m.vi[t + k]];
          v9[9 * n + 3 * k] = q.x;
          v9[9 * n + 3 * k + 1] = q.y;
          v9[9 * n + 3 * k + 2] = q.z;
        }
      }
    const RenderSettings& rs = sc.render;
    std::string s;
    auto f = [](float x) {
      char b[32];
      std::snprintf(b, sizeof b, "%.9g", x);
      return std::string(b);
    };
    s += "accelerator=" + sc.accelerator + "\n";
    s += "params=" + sc.acceleratorParams.Describe() + "\n";
    s += "integrator=" + sc.surfaceIntegrator + "\nrenderer=" + sc.renderer + "\n";
    s += "sampler=" + sc.sampler + "\n";
    s += "film=" + std::to_string(rs.xres) + "," + std::to_string(rs.yres) + "\n";
    s += "spp=" + std::to_string(rs.spp) + "\nfov=" + f(rs.fov_deg) + "\n";
    s += "camera_to_world=";
    for (int i = 0; i < 12; ++i) s += f(rs.camera_to_world[i]) + (i < 11 ? "," : "\n");
    s += "screen=";
    for (int i = 0; i < 4; ++i) s += f(rs.screen[i]) + (i < 3 ? "," : "\n");
    s += "light=" + f(rs.light.x) + "," + f(rs.light.y) + "," + f(rs.light.z) + "\n";
    s += "intensity=" + f(rs.intensity) + "\nmax_depth=" + std::to_string(rs.max_depth) + "\n";
    s += std::string("has_point_light=") + (sc.hasPointLight ? "1" : "0") + "\n";
    s += "meshes=" + std::to_string(sc.meshes.size()) + "\n";
    s += "skipped=" + std::to_string(sc.skippedShapes) + "\n";
    for (const std::string& m : sc.messages) s += "message=" + m + "\n";
    if (info && info_cap) {
      const size_t k = std::min<size_t>(s.size(), info_cap - 1);
      std::memcpy(info, s.data(), k);
      info[k] = 0;
    }
    return (int64_t)n;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Renders a pbrt scene file with RenderScene (accel_override: "" for the file's Accelerator).
// rgb: room for rgb_cap floats (3 per pixel of the film).  stats as mirror_render.
int mirror_render_scene(const char* path, const char* accel_override, float* rgb,
                        uint64_t rgb_cap, int64_t* stats, const char* pfm_path) {
  SetSevereHandler(throw_severe);
  try {
    SceneFile sc;
    ParseSceneFile(path, &sc);
    std::vector<float> img;
    const RenderStats st = RenderScene(sc, accel_override ? accel_override : "", &img);
    if (img.size() > rgb_cap) throw std::runtime_error("rgb buffer too small");
    std::memcpy(rgb, img.data(), sizeof(float) * img.size());
    stats[0] = st.tasks;
    stats[1] = st.nx;
    stats[2] = st.ny;
    stats[3] = (int64_t)st.camera_rays;
    stats[4] = (int64_t)st.shadow_rays;
    stats[5] = (int64_t)st.bounce_rays;
    if (pfm_path && !WritePFM(pfm_path, sc.render.xres, sc.render.yres, img))
      throw std::runtime_error("PFM write");
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// WritePFM then ReadPFM of an image: rgb_out gets what was read back; returns 0 when both
// succeeded with the same size.
int mirror_pfm_roundtrip(const char* path, int xres, int yres, const float* rgb_in,
                         float* rgb_out) {
  SetSevereHandler(throw_severe);
  try {
    std::vector<float> img(rgb_in, rgb_in + 3 * (size_t)xres * yres), back;
    if (!WritePFM(path, xres, yres, img)) throw std::runtime_error("WritePFM failed");
    int x = 0, y = 0;
    if (!ReadPFM(path, &x, &y, &back)) throw std::runtime_error("ReadPFM failed");
    if (x != xres || y != yres) throw std::runtime_error("PFM size changed");
    std::memcpy(rgb_out, back.data(), sizeof(float) * back.size());
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// ReadPFM into rgb (room for cap floats); returns 0 and the size, or 1.
int mirror_read_pfm(const char* path, int* xres, int* yres, float* rgb, uint64_t cap) {
  std::vector<float> img;
  if (!ReadPFM(path, xres, yres, &img) || img.size() > cap) return 1;
  std::memcpy(rgb, img.data(), sizeof(float) * img.size());
  return 0;
}

// DiffImages: out = {small, big, avg1, avg2, avg_delta_percent, mse, differ}.
int mirror_diff_images(const float* a, const float* b, int xres, int yres, float tol,
                       double* out) {
  SetSevereHandler(throw_severe);
  try {
    const size_t n = 3 * (size_t)xres * yres;
    const ImageDiff d = DiffImages(std::vector<float>(a, a + n), std::vector<float>(b, b + n),
                                   xres, yres, tol);
    const double o[7] = {(double)d.small, (double)d.big, d.avg1, d.avg2, d.avg_delta_percent,
                         d.mse, d.differ ? 1.0 : 0.0};
    std::memcpy(out, o, sizeof o);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

}  // extern "C"
