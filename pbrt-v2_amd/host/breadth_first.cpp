// breadth_first.cpp -- see breadth_first.h.
#include "breadth_first.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace pbrt_amd {

namespace {

constexpr float kInvPi = 0.31830988618379067154f;

int RoundUpPow2(int v) {
  v--;
  v |= v >> 1;
  v |= v >> 2;
  v |= v >> 4;
  v |= v >> 8;
  v |= v >> 16;
  return v + 1;
}

uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}

struct V {
  float x, y, z;
};
V sub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float dot(V a, V b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
V normalize(V v) {
  const float l = std::sqrt(dot(v, v));
  if (l == 0.f) return v;
  const float inv = 1.f / l;
  return {v.x * inv, v.y * inv, v.z * inv};
}

// Geometric normal of every triangle, in the accelerator's triangle order.
std::vector<V> triangle_normals(const std::vector<TriangleMesh>& meshes) {
  std::vector<V> n;
  for (const TriangleMesh& m : meshes)
    for (size_t t = 0; t + 2 < m.vi.size(); t += 3) {
      const Point& a = m.p[m.vi[t]];
      const Point& b = m.p[m.vi[t + 1]];
      const Point& c = m.p[m.vi[t + 2]];
      const V e1 = {b.x - a.x, b.y - a.y, b.z - a.z}, e2 = {c.x - a.x, c.y - a.y, c.z - a.z};
      n.push_back(normalize(cross(e1, e2)));
    }
  return n;
}

// Pinhole camera ray through image-plane point (fx, fy) in pixels: the FP32 steps of the
// workloads' camera, so one centred sample per pixel reproduces it exactly.
RayDifferential camera_ray(const RenderSettings& rs, float tan_half, float aspect, float fx,
                           float fy) {
  const float px = ((fx / (float)rs.xres) * 2.0f - 1.0f) * tan_half * aspect;
  const float py = (1.0f - (fy / (float)rs.yres) * 2.0f) * tan_half;
  const float pz = 1.0f;
  const float l = std::sqrt((px * px + py * py) + pz * pz);
  const float inv = 1.0f / l;
  return RayDifferential(rs.origin, Vector(px * inv, py * inv, pz * inv), 0.f, INFINITY);
}

// pbrt's perspective camera ray through raster point (fx, fy): screen-window point at z = 1
// scaled by tan(fov / 2), normalized in camera space, then CameraToWorld (the direction is
// rotated, not renormalized, as Transform::operator()(Ray) does).
RayDifferential camera_ray_pbrt(const RenderSettings& rs, float tan_half, float fx, float fy) {
  const float* sw = rs.screen;
  const float sx = sw[0] + (fx / (float)rs.xres) * (sw[1] - sw[0]);
  const float sy = sw[3] - (fy / (float)rs.yres) * (sw[3] - sw[2]);
  const V d = normalize(V{sx * tan_half, sy * tan_half, 1.0f});
  const float* M = rs.camera_to_world;
  const Vector w((M[0] * d.x + M[1] * d.y) + M[2] * d.z, (M[4] * d.x + M[5] * d.y) + M[6] * d.z,
                 (M[8] * d.x + M[9] * d.y) + M[10] * d.z);
  return RayDifferential(Point(M[3], M[7], M[11]), w, 0.f, INFINITY);
}

}  // namespace

void TilesInXY(int count, int dx, int dy, int* nx, int* ny) {
  *nx = count;
  *ny = 1;
  while ((*nx & 0x1) == 0 && 2 * dx * *ny < dy * *nx) {
    *nx >>= 1;
    *ny <<= 1;
  }
}

void ComputeSubWindow(int num, int count, int xres, int yres, int* x0, int* x1, int* y0,
                      int* y1) {
  int nx, ny;
  TilesInXY(count, xres, yres, &nx, &ny);
  const int xo = num % nx, yo = num / nx;
  const float tx0 = float(xo) / float(nx), tx1 = float(xo + 1) / float(nx);
  const float ty0 = float(yo) / float(ny), ty1 = float(yo + 1) / float(ny);
  auto lerp = [](float t, float a, float b) { return (1.f - t) * a + t * b; };
  *x0 = (int)std::floor(lerp(tx0, 0.f, (float)xres));
  *x1 = (int)std::floor(lerp(tx1, 0.f, (float)xres));
  *y0 = (int)std::floor(lerp(ty0, 0.f, (float)yres));
  *y1 = (int)std::floor(lerp(ty1, 0.f, (float)yres));
}

RenderStats RenderBreadthFirst(Aggregate& accel, const std::vector<TriangleMesh>& meshes,
                               const RenderSettings& rs, std::vector<float>* rgb) {
  RenderStats st;
  const std::vector<V> normals = triangle_normals(meshes);
  rgb->assign(3 * (size_t)rs.xres * rs.yres, 0.f);
  const long long nPixels = (long long)rs.xres * rs.yres;
  const unsigned maxRays = accel.MaxRaysPerCall();
  int nTasks = (int)((nPixels * rs.spp + maxRays - 1) / maxRays);
  nTasks = RoundUpPow2(std::max(nTasks, 1));
  TilesInXY(nTasks, rs.xres, rs.yres, &st.nx, &st.ny);
  st.tasks = nTasks;
  accel.Preprocess(rs.xres, rs.yres, (unsigned)rs.spp, st.nx, st.ny);
  const float tan_half =
      (float)std::tan((double)rs.fov_deg * (3.14159265358979323846 / 180.0) / 2.0);
  const float aspect = (float)((double)rs.xres / (double)rs.yres);
  const V light = {rs.light.x, rs.light.y, rs.light.z};

  for (int i = 0; i < nTasks; ++i) {
    const int num = nTasks - 1 - i;  // breadth-first.cpp:150 runs the tasks in reverse
    int x0, x1, y0, y1;
    ComputeSubWindow(num, nTasks, rs.xres, rs.yres, &x0, &x1, &y0, &y1);
    const int w = x1 - x0, h = y1 - y0;
    if (w <= 0 || h <= 0) continue;
    const unsigned count = (unsigned)(w * rs.spp * h);
    // camera rays in sampler order: pixels row-major, the spp samples of a pixel adjacent
    std::vector<RayDifferential> rays(count);
    std::vector<int> pixel(count);
    for (int y = y0; y < y1; ++y)
      for (int x = x0; x < x1; ++x)
        for (int k = 0; k < rs.spp; ++k) {
          const size_t j = ((size_t)(y - y0) * w + (x - x0)) * rs.spp + k;
          const float jx = ((float)k + 0.5f) / (float)rs.spp;
          rays[j] = rs.has_camera_to_world
                        ? camera_ray_pbrt(rs, tan_half, (float)x + jx, (float)y + 0.5f)
                        : camera_ray(rs, tan_half, aspect, (float)x + jx, (float)y + 0.5f);
          pixel[j] = y * rs.xres + x;
        }
    st.camera_rays += count;
    std::vector<Intersection> isect(count);
    std::unique_ptr<bool[]> hit(new bool[count]);
    accel.Intersect(rays.data(), isect.data(), hit.get(), count);
    std::vector<float> L(count, 0.f), beta(count, 1.f);
    std::vector<Ray> shadow(count);
    std::unique_ptr<bool[]> shoot(new bool[count]);
    std::vector<char> occluded(count);
    for (int depth = 0;; ++depth) {
      // direct lighting at the current hits: one shadow ray each (SetSegment)
      for (unsigned j = 0; j < count; ++j) {
        shoot[j] = hit[j];
        if (!hit[j]) continue;
        const Point& p = isect[j].dg.p;
        const V v = sub(light, V{p.x, p.y, p.z});
        const float dist = std::sqrt(dot(v, v));
        const float inv = 1.f / dist;
        shadow[j] = Ray(p, Vector(v.x * inv, v.y * inv, v.z * inv), isect[j].rayEpsilon,
                        dist * (1.f - 0.f));
      }
      accel.IntersectP(shadow.data(), occluded.data(), count, shoot.get());
      for (unsigned j = 0; j < count; ++j) {
        if (!shoot[j]) continue;
        ++st.shadow_rays;
        if (occluded[j] == '1') continue;
        V n = normals[(size_t)isect[j].triangle];
        const V d = {rays[j].d.x, rays[j].d.y, rays[j].d.z};
        if (dot(n, d) > 0.f) n = {-n.x, -n.y, -n.z};
        const Vector& wi = shadow[j].d;
        const float c = std::max(0.f, dot(n, V{wi.x, wi.y, wi.z}));
        const float d2 = shadow[j].maxt * shadow[j].maxt;
        L[j] += beta[j] * (rs.albedo * kInvPi) * rs.intensity * c / d2;
      }
      if (depth >= rs.max_depth) break;
      // path: cosine-weighted bounce from every live hit; the Lambertian weight is albedo
      bool any = false;
      for (unsigned j = 0; j < count; ++j) {
        if (!hit[j]) continue;
        V n = normals[(size_t)isect[j].triangle];
        const V d = {rays[j].d.x, rays[j].d.y, rays[j].d.z};
        if (dot(n, d) > 0.f) n = {-n.x, -n.y, -n.z};
        const uint32_t hh = mix32(rs.seed ^ mix32(j * 0x9e3779b9u ^ mix32((uint32_t)depth +
                                                                         (uint32_t)num * 977u)));
        const float u1 = (float)(hh >> 8) * (1.0f / 16777216.0f);
        const float u2 = (float)(mix32(hh) >> 8) * (1.0f / 16777216.0f);
        const float r = std::sqrt(u1), phi = 6.28318530717958647692f * u2;
        const float lx = r * std::cos(phi), ly = r * std::sin(phi);
        const float lz = std::sqrt(std::max(0.f, 1.f - u1));
        const float sgn = n.z >= 0.f ? 1.f : -1.f;
        const float a = -1.f / (sgn + n.z);
        const float b = n.x * n.y * a;
        const V t1 = {1.f + sgn * n.x * n.x * a, sgn * b, -sgn * n.x};
        const V t2 = {b, sgn + n.y * n.y * a, -n.y};
        const Vector wo(t1.x * lx + t2.x * ly + n.x * lz, t1.y * lx + t2.y * ly + n.y * lz,
                        t1.z * lx + t2.z * ly + n.z * lz);
        rays[j] = RayDifferential(isect[j].dg.p, wo, isect[j].rayEpsilon, INFINITY);
        beta[j] *= rs.albedo;
        any = true;
        ++st.bounce_rays;
      }
      if (!any) break;
      accel.Intersect(rays.data(), isect.data(), hit.get(), count, depth + 1);
    }
    for (unsigned j = 0; j < count; ++j) {
      const float v = L[j] / (float)rs.spp;
      for (int c = 0; c < 3; ++c) (*rgb)[3 * (size_t)pixel[j] + c] += v;
    }
  }
  return st;
}

bool WritePFM(const char* path, int xres, int yres, const std::vector<float>& rgb) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  std::fprintf(f, "PF\n%d %d\n-1\n", xres, yres);
  // PFM stores the bottom row first
  bool ok = true;
  for (int y = yres - 1; y >= 0 && ok; --y)
    ok = std::fwrite(&rgb[3 * (size_t)y * xres], sizeof(float), 3 * (size_t)xres, f) ==
         3 * (size_t)xres;
  return std::fclose(f) == 0 && ok;
}

bool ReadPFM(const char* path, int* xres, int* yres, std::vector<float>* rgb) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  char tag[3] = {0, 0, 0};
  double scale = 0.0;
  bool ok = std::fscanf(f, "%2s %d %d %lf", tag, xres, yres, &scale) == 4 &&
            std::strcmp(tag, "PF") == 0 && *xres > 0 && *yres > 0 && scale != 0.0 &&
            std::fgetc(f) != EOF;  // the single whitespace byte after the scale
  if (ok) {
    const size_t row = 3 * (size_t)*xres;
    rgb->assign(row * (size_t)*yres, 0.f);
    // PFM stores the bottom row first
    for (int y = *yres - 1; y >= 0 && ok; --y)
      ok = std::fread(&(*rgb)[row * (size_t)y], sizeof(float), row, f) == row;
    const uint16_t probe = 1;
    const bool host_little = *(const uint8_t*)&probe == 1;
    if (ok && (scale < 0.0) != host_little)  // negative scale: little-endian data
      for (float& v : *rgb) {
        uint32_t u;
        std::memcpy(&u, &v, 4);
        u = __builtin_bswap32(u);
        std::memcpy(&v, &u, 4);
      }
  }
  std::fclose(f);
  return ok;
}

ImageDiff DiffImages(const std::vector<float>& a, const std::vector<float>& b, int xres,
                     int yres, float tol_percent) {
  ImageDiff d;
  const size_t n = 3 * (size_t)xres * (size_t)yres;
  if (a.size() < n || b.size() < n) Severe("DiffImages: image smaller than xres x yres");
  double sum1 = 0.0, sum2 = 0.0;
  for (size_t i = 0; i < n; ++i) {
    if (a[i] == 0.f && b[i] == 0.f) continue;
    sum1 += a[i];
    sum2 += b[i];
    const float rel = std::fabs(a[i] - b[i]) / a[i];
    d.mse += (double)(a[i] - b[i]) * (a[i] - b[i]);
    if (rel > .005f) ++d.small;
    if (rel > .05f) ++d.big;
  }
  const double count = 3.0 * xres * yres;
  d.avg1 = sum1 / count;
  d.avg2 = sum2 / count;
  const double lo = std::min(d.avg1, d.avg2);
  d.avg_delta_percent = lo != 0.0 ? 100.0 * (d.avg1 - d.avg2) / lo : (d.avg1 == d.avg2 ? 0.0 : INFINITY);
  d.mse /= count;
  d.differ = tol_percent == 0.f ? (d.big > 0 || d.small > 0)
                                : std::fabs(d.avg_delta_percent) > tol_percent;
  return d;
}

}  // namespace pbrt_amd
