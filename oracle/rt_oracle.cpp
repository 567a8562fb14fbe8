// rt_oracle.cpp -- CPU oracle for the ray-space accelerators.
//
// TEST INFRASTRUCTURE ONLY (the checker).  Never linked into the product library.
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off, no fast-math).
//
// Reference files restated here (paths relative to the reference root):
//   leaf predicates        src/bin/cl/intersectAllLeaves.cl:12-238
//   naive predicates       src/bin/cl/naive.cl:4-231
//   pbrt predicate         src/shapes/trianglemesh.cpp:119-240, src/core/geometry.h:481-488
//   6DB culls              src/bin/cl/intersectsNode6DB.cl:6-99
//   IA cull                src/bin/cl/intersectsNodeIA.cl:6-65
//   RH leaf/level build    src/bin/cl/rayhconstruct6DB.cl, rayhconstructP6DB.cl,
//                          levelConstruct6DB.cl, levelConstructP6DB.cl,
//                          host loop src/accelerators/rayhierarchy.cpp:375-549
//   RH traversal           src/bin/cl/intersection6DB.cl, intersection6DB2.cl,
//                          intersectionP6DB.cl, intersectionIA.cl
//   repair pass            src/bin/cl/yetAnotherIntersection6DB.cl, host rayhierarchy.cpp:998-1045
//   BVH x RH traversal     src/bin/cl/intersection6DB_BVH.cl, intersection6DB_BVH2.cl,
//                          intersectionP6DB_BVH.cl
//   shading record         src/bin/cl/computeDpTuTv.cl:8-91
//   BVH build / handoff    src/accelerators/bvh.cpp:33-455, rayBVH.cpp:144-146,204-276
//   CPU BVH traversal      src/accelerators/bvh.cpp:119-141,464-565
#include "rt_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

const float kInf = std::numeric_limits<float>::infinity();
// EPS of intersectAllLeaves.cl:8 / naive.cl:2, taken in FP32 (the kernels do not enable
// cl_khr_fp64, so their double literals are single precision).
const float kEps = 0.002f;

struct V3 {
  float x, y, z;
};
inline float fmn(float a, float b) { return b < a ? b : a; }
inline float fmx(float a, float b) { return b > a ? b : a; }
inline V3 vmin(V3 a, V3 b) { return {fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 normalize(V3 v) {
  float l2 = dot(v, v);
  if (l2 == 0.f) return v;
  float inv = 1.0f / std::sqrt(l2);
  return {v.x * inv, v.y * inv, v.z * inv};
}
// pbrt's Cross computes in double and rounds to float (geometry.h:481-488).
inline V3 cross_d(V3 a, V3 b) {
  double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
  return {(float)((ay * bz) - (az * by)), (float)((az * bx) - (ax * bz)),
          (float)((ax * by) - (ay * bx))};
}

// ---- decision margins (parity bookkeeping) ---------------------------------------------
// A decision is ulp-marginal when its operand lies within 4 ulp of the threshold it is
// compared with.  For a comparison with 0 of a sum of products (det, the barycentric
// numerators) the ulp is that of the largest product; for normalised direction components
// it is that of 1.  Tracking is on only while a trace asks for per-ray flags.
thread_local bool g_track = false;
thread_local bool g_marg = false;
inline float ulp_of(float x) {
  x = std::fabs(x);
  if (!(x >= std::numeric_limits<float>::min())) return std::numeric_limits<float>::denorm_min();
  int e;
  (void)std::frexp(x, &e);  // x = m * 2^e, m in [0.5, 1)
  return std::ldexp(1.0f, e - 24);
}
inline bool near_thr(float a, float b) {
  if (!std::isfinite(a) || !std::isfinite(b)) return false;
  return std::fabs(a - b) <= 4.f * ulp_of(std::fmax(std::fabs(a), std::fabs(b)));
}
inline bool near_zero(float x, float scale) { return std::fabs(x) <= 4.f * ulp_of(scale); }
inline float dot_scale(V3 a, V3 b) {
  return std::fmax(std::fmax(std::fabs(a.x * b.x), std::fabs(a.y * b.y)), std::fabs(a.z * b.z));
}
inline void mark(bool m) {
  if (g_track && m) g_marg = true;
}

struct Tri {
  V3 v1, v2, v3, e1, e2, bmin, bmax;
};
inline Tri load_tri(const float* v9, uint32_t i) {
  const float* p = v9 + 9 * (size_t)i;
  Tri t;
  t.v1 = {p[0], p[1], p[2]};
  t.v2 = {p[3], p[4], p[5]};
  t.v3 = {p[6], p[7], p[8]};
  t.e1 = sub(t.v2, t.v1);
  t.e2 = sub(t.v3, t.v1);
  t.bmin = vmin(vmin(t.v1, t.v2), t.v3);
  t.bmax = vmax(vmax(t.v1, t.v2), t.v3);
  return t;
}

// ---- leaf predicates -------------------------------------------------------------------
// intersectAllLeaves.cl:40-55 (strict: det != 0, b1 in (0,1), b2 > 0, b1+b2 < 1)
inline bool mt_strict(V3 o, V3 d, const Tri& T, float& t, float& b1, float& b2) {
  V3 s = cross(d, T.e2);
  float det = dot(s, T.e1);
  if (det == 0.0f) return false;
  float inv = 1.0f / det;
  V3 ov = sub(o, T.v1);
  b1 = dot(ov, s) * inv;
  if (!(b1 > 0.f && b1 < 1.f)) return false;
  V3 s2 = cross(ov, T.e1);
  b2 = dot(d, s2) * inv;
  if (!(b2 > 0.f && (b1 + b2) < 1.f)) return false;
  t = dot(T.e2, s2) * inv;
  return true;
}
// mt_strict, and in `marg` whether its outcome rests on a marginal decision: an accepted test
// with any marginal comparison, or a rejection made by a marginal comparison.  Unless det is
// 0, t is computed as the accepted test would, also for a rejection (so that a marginal
// rejection can be weighed against the ray's result); `t_known` tells.
inline bool mt_strict_m(V3 o, V3 d, const Tri& T, float& t, float& b1, float& b2, bool& marg,
                        bool& t_known) {
  bool m = false;
  t_known = false;
  V3 s = cross(d, T.e2);
  float det = dot(s, T.e1);
  const bool mdet = near_zero(det, dot_scale(s, T.e1));
  if (det == 0.0f) {
    marg = mdet;
    return false;
  }
  m |= mdet;
  float inv = 1.0f / det;
  V3 ov = sub(o, T.v1);
  const float n1 = dot(ov, s);
  b1 = n1 * inv;
  V3 s2 = cross(ov, T.e1);
  const float n2 = dot(d, s2);
  b2 = n2 * inv;
  t = dot(T.e2, s2) * inv;
  t_known = true;
  const bool m1 = near_zero(n1, dot_scale(ov, s)) || near_thr(b1, 1.f);
  if (!(b1 > 0.f && b1 < 1.f)) {
    marg = m1;
    return false;
  }
  m |= m1;
  const bool m2 = near_zero(n2, dot_scale(d, s2)) || near_thr(b1 + b2, 1.f);
  if (!(b2 > 0.f && (b1 + b2) < 1.f)) {
    marg = m2;
    return false;
  }
  m |= m2;
  marg = m;
  return true;
}
// naive.cl:30-46 (TriRay closest: |det| >= EPS, loose barycentric bounds)
const float kNaiveLo = -1e-3f - kEps;
const float kNaiveHi = (1.0f + 1e-3f) + kEps;
inline bool mt_naive(V3 o, V3 d, const Tri& T, float& t, float& b1, float& b2) {
  V3 s1 = cross(d, T.e2);
  float div = dot(s1, T.e1);
  if (div < kEps && div > -kEps) return false;
  float inv = 1.0f / div;
  V3 dd = sub(o, T.v1);
  b1 = dot(dd, s1) * inv;
  if (b1 < kNaiveLo || b1 > kNaiveHi) return false;
  V3 s2 = cross(dd, T.e1);
  b2 = dot(d, s2) * inv;
  if (b2 < kNaiveLo || (b1 + b2) > kNaiveHi) return false;
  t = dot(T.e2, s2) * inv;
  return true;
}
// naive.cl:211-227 (IntersectionP: det != 0, bounds +-1e-3)
const float kNaivePLo = -1e-3f;
const float kNaivePHi = 1.0f + 1e-3f;
inline bool mt_naive_p(V3 o, V3 d, const Tri& T, float& t, float& b1, float& b2) {
  V3 s1 = cross(d, T.e2);
  float div = dot(s1, T.e1);
  if (div == 0.0f) return false;
  float inv = 1.0f / div;
  V3 dd = sub(o, T.v1);
  b1 = dot(dd, s1) * inv;
  if (b1 < kNaivePLo || b1 > kNaivePHi) return false;
  V3 s2 = cross(dd, T.e1);
  b2 = dot(d, s2) * inv;
  if (b2 < kNaivePLo || (b1 + b2) > kNaivePHi) return false;
  t = dot(T.e2, s2) * inv;
  return true;
}
// intersectAllLeaves.cl:210-225 (yetAnotherIntersectAllLeaves: shrunk barycentric bounds)
const float kRepHi1 = 1.0f + kEps;
const float kRepHi2 = 1.0f - kEps;
inline bool mt_repair(V3 o, V3 d, const Tri& T, float& t, float& b1, float& b2) {
  V3 s = cross(d, T.e2);
  float det = dot(s, T.e1);
  if (det == 0.0f) return false;
  float inv = 1.0f / det;
  V3 ov = sub(o, T.v1);
  b1 = dot(ov, s) * inv;
  if (b1 < kEps || b1 > kRepHi1) return false;
  V3 s2 = cross(ov, T.e1);
  b2 = dot(d, s2) * inv;
  if (b2 < kEps || (b1 + b2) > kRepHi2) return false;
  t = dot(T.e2, s2) * inv;
  return true;
}
// trianglemesh.cpp:119-152 (pbrt inclusive predicate, Cross in double)
inline bool mt_pbrt(V3 o, V3 d, const Tri& T, float& t, float& b1, float& b2) {
  V3 s1 = cross_d(d, T.e2);
  float div = (s1.x * T.e1.x + s1.y * T.e1.y) + s1.z * T.e1.z;
  if (div == 0.0f) return false;
  float inv = 1.f / div;
  V3 dd = sub(o, T.v1);
  b1 = dot(dd, s1) * inv;
  if (b1 < 0.f || b1 > 1.f) return false;
  V3 s2 = cross_d(dd, T.e1);
  b2 = dot(d, s2) * inv;
  if (b2 < 0.f || b1 + b2 > 1.f) return false;
  t = dot(T.e2, s2) * inv;
  return true;
}

// ---- culls ------------------------------------------------------------------------------
struct Node {
  V3 dmin, dmax, omin, omax;
  bool valid;
};

inline bool box_logic(V3 tmin, V3 tmax, const Node& n) {
  // intersectsNode6DB.cl:53-63 / 88-98; the uninitialised `ret` is taken as false.
  V3 f0 = vmax(tmin, n.dmin);
  V3 f1 = vmin(tmax, n.dmax);
  if (g_track) {
    mark(near_zero(tmin.x, 1.f) || near_zero(tmax.x, 1.f) || near_zero(tmin.y, 1.f) ||
         near_zero(tmax.y, 1.f) || near_zero(tmin.z, 1.f) || near_zero(tmax.z, 1.f));
    mark(near_thr(f0.x, f1.x + kEps) || near_thr(f0.y, f1.y + kEps) ||
         near_thr(f0.z, f1.z + kEps));
  }
  bool ret = false;
  if (tmin.x < 0.f && tmax.x > 0.f) ret = (f0.x < f1.x + kEps);
  if (tmin.y < 0.f && tmax.y > 0.f) ret |= (f0.y < f1.y + kEps);
  if (tmin.z < 0.f && tmax.z > 0.f) ret |= (f0.z < f1.z + kEps);
  return ret || (f0.x < f1.x + kEps && f0.y < f1.y + kEps && f0.z < f1.z + kEps);
}

// intersectsNode6DB.cl:66-99
inline bool cull_tri(const Node& n, const Tri& T) {
  V3 oc = {(n.omin.x + n.omax.x) / 2, (n.omin.y + n.omax.y) / 2, (n.omin.z + n.omax.z) / 2};
  V3 a = normalize(sub(T.v1, oc));
  V3 tmin = a, tmax = a;
  V3 b = normalize(sub(T.v2, oc));
  tmin = vmin(tmin, b);
  tmax = vmax(tmax, b);
  V3 c = normalize(sub(T.v3, oc));
  tmin = vmin(tmin, c);
  tmax = vmax(tmax, c);
  return box_logic(tmin, tmax, n);
}

// intersectsNode6DB.cl:6-64
inline bool cull_node6(const Node& n, V3 bmin, V3 bmax) {
  V3 h = {(n.omax.x - n.omin.x) / 2, (n.omax.y - n.omin.y) / 2, (n.omax.z - n.omin.z) / 2};
  bmin = sub(bmin, h);
  bmax = add(bmax, h);
  V3 oc = add(h, n.omin);
  const V3 corners[8] = {{bmin.x, bmin.y, bmin.z}, {bmin.x, bmin.y, bmax.z},
                         {bmin.x, bmax.y, bmin.z}, {bmin.x, bmax.y, bmax.z},
                         {bmax.x, bmin.y, bmin.z}, {bmax.x, bmax.y, bmin.z},
                         {bmax.x, bmin.y, bmax.z}, {bmax.x, bmax.y, bmax.z}};
  V3 r = normalize(sub(corners[0], oc));
  V3 tmin = r, tmax = r;
  for (int i = 1; i < 8; ++i) {
    r = normalize(sub(corners[i], oc));
    tmin = vmin(tmin, r);
    tmax = vmax(tmax, r);
  }
  return box_logic(tmin, tmax, n);
}

// intersectsNodeIA.cl:6-65
inline void ia_axis(float lo, float hi, float dlo, float dhi, float& smin, float& smax) {
  float a = lo * dlo, b = hi * dlo, c = hi * dhi, d = lo * dhi;
  smin = fmn(fmn(fmn(a, b), c), d);
  smax = fmx(fmx(fmx(a, b), c), d);
}
inline bool cull_ia(const Node& n, V3 bmin, V3 bmax) {
  V3 lo = sub(bmin, n.omax);
  V3 hi = sub(bmax, n.omin);
  if (g_track)
    mark(near_zero(n.dmin.x, 1.f) || near_zero(n.dmax.x, 1.f) || near_zero(n.dmin.y, 1.f) ||
         near_zero(n.dmax.y, 1.f) || near_zero(n.dmin.z, 1.f) || near_zero(n.dmax.z, 1.f));
  if ((n.dmin.x <= 0.f && n.dmax.x >= 0.f) || (n.dmin.y <= 0.f && n.dmax.y >= 0.f) ||
      (n.dmin.z <= 0.f && n.dmax.z >= 0.f))
    return true;
  float ix0 = 1 / n.dmax.x, ix1 = 1 / n.dmin.x;
  float iy0 = 1 / n.dmax.y, iy1 = 1 / n.dmin.y;
  float iz0 = 1 / n.dmax.z, iz1 = 1 / n.dmin.z;
  float sx, sy, tx, ty, ux, uy;
  ia_axis(lo.x, hi.x, ix0, ix1, sx, sy);
  ia_axis(lo.y, hi.y, iy0, iy1, tx, ty);
  ia_axis(lo.z, hi.z, iz0, iz1, ux, uy);
  sx = fmx(fmx(sx, tx), ux);
  sy = fmn(fmn(sy, ty), uy);
  if (g_track) mark(near_thr(sx, sy + kEps));
  return sx < sy + kEps;
}

// intersectsNode5DB.cl:1-83 (node "box_uv"; the node's dmin/dmax hold (slope, dz) intervals:
// dmin = (min d.y/d.x, min d.z, 0), dmax = (max d.y/d.x, max d.z, 0)).  As C parses it,
// line 79 tests range & (0x1100 == 0x1100) || range & (0x0011 == 0x0011), i.e. range & 1:
// some corner direction has y not below zero (:18-73 set bit 0 for y >= 0 or NaN).  On line
// 81 the first conjunct is (max(..) < min(..)) + EPS, an int 0 or 1 plus EPS, never zero, so
// the slope intervals never decide.  What remains: the node passes when a normalised corner
// direction of the origin-grown box has y >= 0, else when the corners' z range meets the
// node's dz range within EPS.  Corners in the kernel's order (:17, 23, 31, 39, 47, 55, 63, 71).
inline bool cull_5db(const Node& n, V3 bmin, V3 bmax) {
  V3 h = {(n.omax.x - n.omin.x) / 2, (n.omax.y - n.omin.y) / 2, (n.omax.z - n.omin.z) / 2};
  bmin = sub(bmin, h);
  bmax = add(bmax, h);
  V3 oc = add(h, n.omin);
  const V3 corners[8] = {{bmin.x, bmin.y, bmin.z}, {bmax.x, bmax.y, bmax.z},
                         {bmin.x, bmax.y, bmin.z}, {bmin.x, bmax.y, bmax.z},
                         {bmin.x, bmin.y, bmax.z}, {bmax.x, bmin.y, bmin.z},
                         {bmax.x, bmax.y, bmin.z}, {bmax.x, bmin.y, bmax.z}};
  bool up = false, ymarg = false;
  float zlo = 0.f, zhi = 0.f;
  for (int i = 0; i < 8; ++i) {
    const V3 r = normalize(sub(corners[i], oc));
    if (!(r.y < 0.f)) up = true;
    ymarg |= near_zero(r.y, 1.f);
    if (i == 0) {
      zlo = zhi = r.z;
    } else {
      zlo = fmn(zlo, r.z);
      zhi = fmx(zhi, r.z);
    }
  }
  if (g_track) mark(ymarg);
  if (up) return true;
  const float f0 = fmx(zlo, n.dmin.y), f1 = fmn(zhi, n.dmax.y);
  if (g_track) mark(near_thr(f0, f1 + kEps));
  return f0 < f1 + kEps;
}

// The cull of a hierarchy node against an axis-aligned box (a triangle's bounds or a BVH
// node): intersectsNode of the node type's header.
inline bool cull_box(int node_kind, const Node& n, V3 bmin, V3 bmax) {
  if (node_kind == 1) return cull_ia(n, bmin, bmax);
  if (node_kind == 2) return cull_5db(n, bmin, bmax);
  return cull_node6(n, bmin, bmax);
}
// The 5DB BVH kernels skip a node with no primitives and primitive offset 0
// (intersection5DB_BVH.cl:18,56 and the P / 2 variants); the 6DB ones have no such test.
inline bool bvh_dead(int node_kind, const orc_gpu_node& b) {
  return node_kind == 2 && b.nPrimitives == 0 && b.primOffset == 0;
}

inline Node node_from16(const float* p) {
  Node n;
  n.dmin = {p[0], p[1], p[2]};
  n.valid = p[3] != 0.f;
  n.dmax = {p[4], p[5], p[6]};
  n.omin = {p[8], p[9], p[10]};
  n.omax = {p[12], p[13], p[14]};
  return n;
}
inline void node_to16(const Node& n, float* p) {
  p[0] = n.dmin.x; p[1] = n.dmin.y; p[2] = n.dmin.z; p[3] = n.valid ? 1.f : 0.f;
  p[4] = n.dmax.x; p[5] = n.dmax.y; p[6] = n.dmax.z; p[7] = 0.f;
  p[8] = n.omin.x; p[9] = n.omin.y; p[10] = n.omin.z; p[11] = 0.f;
  p[12] = n.omax.x; p[13] = n.omax.y; p[14] = n.omax.z; p[15] = 0.f;
}

// ---- ray hierarchy -----------------------------------------------------------------------
struct RH {
  int L = 0, W = 0, H = 0, cx = 0, cy = 0;
  std::vector<int> gw, gh;
  std::vector<size_t> off;
  std::vector<Node> nodes;
  const Node& at(int l, int k, int j) const { return nodes[off[l] + (size_t)j * gw[l] + k]; }
};

int rh_levels(int W, int H, int cx, int cy, int height, int* dims) {
  // Leaf grid ceil(W/cx) x ceil(H/cy) (rayhierarchy.cpp:377).  The reference halves the grid
  // with floor at every level and stops after a level whose floor width or height is odd
  // (:412-442); the level count follows that rule exactly.  Where it halves an odd dimension
  // its parents leave the last child column (row) uncovered and those rays are never tested;
  // here the parents cover every child (ceil), a superset of the reference's coverage with the
  // same chain length for the rays it does cover.
  int gw = (W + cx - 1) / cx, gh = (H + cy - 1) / cy;
  int fw = gw, fh = gh;
  dims[0] = gw;
  dims[1] = gh;
  int L = 1;
  for (int i = 1; i < height; ++i) {
    gw = (gw + 1) / 2;
    gh = (gh + 1) / 2;
    fw /= 2;
    fh /= 2;
    dims[2 * L] = gw;
    dims[2 * L + 1] = gh;
    ++L;
    // a zero floor dimension leaves the reference no top-level node (it traces nothing):
    // stop there too
    if ((fw & 1) || (fh & 1) || fw == 0 || fh == 0) break;
  }
  return L;
}

inline bool ray_valid(const orc_rays* r, size_t i) { return r->valid == nullptr || r->valid[i]; }
inline V3 ray_o(const orc_rays* r, size_t i) { return {r->ox[i], r->oy[i], r->oz[i]}; }
inline V3 ray_d(const orc_rays* r, size_t i) { return {r->dx[i], r->dy[i], r->dz[i]}; }

// Direction part of a leaf for node "box_uv": the slope d.y/d.x and d.z.
//   rayhconstruct5DB.cl:21-40 (primary rays): slope 0 when d.x == 0, the slope and d.z
//     intervals are the min/max over the tile;
//   rayhconstructP5DB.cl:21-67 (validity-aware): a valid ray with d.x == 0 sets the slope
//     interval to [-1, 1], dropping what the tile's earlier rays gave (the kernel assigns, it
//     does not merge), later rays widen it again.  The tile is read row by row, left to right;
//     the reads the kernel clamps to the image edge repeat an edge ray right after itself,
//     which changes nothing.
inline void uv_leaf_ray(V3 d, bool first, bool validity_aware, Node& n) {
  const bool flat = d.x == 0.f;
  const float s = flat ? 0.f : d.y / d.x;
  if (validity_aware && flat) {
    n.dmin.x = -1.f;
    n.dmax.x = 1.f;
  } else if (first) {
    n.dmin.x = n.dmax.x = s;
  } else {
    n.dmin.x = fmn(n.dmin.x, s);
    n.dmax.x = fmx(n.dmax.x, s);
  }
  if (first) {
    n.dmin.y = n.dmax.y = d.z;
  } else {
    n.dmin.y = fmn(n.dmin.y, d.z);
    n.dmax.y = fmx(n.dmax.y, d.z);
  }
  n.dmin.z = n.dmax.z = 0.f;
}

// rayhconstruct6DB.cl:24-38 / rayhconstructP6DB.cl:71-94 (5DB: uv_leaf_ray) and the 2x2
// merges of levelConstruct6DB.cl / levelConstructP6DB.cl / levelConstruct5DB.cl:32-59 /
// levelConstructP5DB.cl:25-94: every node is the component-wise min/max of the direction part
// and of o over the (valid) rays below it.
RH rh_build(const orc_rays* r, int cx, int cy, int height, bool use_valid, int node_kind) {
  RH h;
  h.W = r->W;
  h.H = r->H;
  h.cx = cx;
  h.cy = cy;
  int dims[128];
  h.L = rh_levels(r->W, r->H, cx, cy, height, dims);
  size_t total = 0;
  for (int l = 0; l < h.L; ++l) {
    h.gw.push_back(dims[2 * l]);
    h.gh.push_back(dims[2 * l + 1]);
    h.off.push_back(total);
    total += (size_t)dims[2 * l] * dims[2 * l + 1];
  }
  h.nodes.assign(total, Node{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, false});
  for (int j = 0; j < h.gh[0]; ++j)
    for (int k = 0; k < h.gw[0]; ++k) {
      Node n{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, false};
      for (int y = j * cy; y < std::min(r->H, (j + 1) * cy); ++y)
        for (int x = k * cx; x < std::min(r->W, (k + 1) * cx); ++x) {
          size_t i = (size_t)y * r->W + x;
          if (use_valid && !ray_valid(r, i)) continue;
          V3 d = ray_d(r, i), o = ray_o(r, i);
          if (node_kind == 2) {
            uv_leaf_ray(d, !n.valid, use_valid, n);
            n.omin = n.valid ? vmin(n.omin, o) : o;
            n.omax = n.valid ? vmax(n.omax, o) : o;
            n.valid = true;
          } else if (!n.valid) {
            n.dmin = n.dmax = d;
            n.omin = n.omax = o;
            n.valid = true;
          } else {
            n.dmin = vmin(n.dmin, d);
            n.dmax = vmax(n.dmax, d);
            n.omin = vmin(n.omin, o);
            n.omax = vmax(n.omax, o);
          }
        }
      h.nodes[h.off[0] + (size_t)j * h.gw[0] + k] = n;
    }
  for (int l = 1; l < h.L; ++l)
    for (int j = 0; j < h.gh[l]; ++j)
      for (int k = 0; k < h.gw[l]; ++k) {
        Node n{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, false};
        for (int cj = 2 * j; cj <= 2 * j + 1; ++cj)
          for (int ck = 2 * k; ck <= 2 * k + 1; ++ck) {
            if (ck >= h.gw[l - 1] || cj >= h.gh[l - 1]) continue;
            const Node& c = h.at(l - 1, ck, cj);
            if (!c.valid) continue;
            if (!n.valid) {
              n = c;
            } else {
              n.dmin = vmin(n.dmin, c.dmin);
              n.dmax = vmax(n.dmax, c.dmax);
              n.omin = vmin(n.omin, c.omin);
              n.omax = vmax(n.omax, c.omax);
            }
          }
        h.nodes[h.off[l] + (size_t)j * h.gw[l] + k] = n;
      }
  return h;
}

struct Out {
  int32_t* id;
  float *t, *b1, *b2;
  uint8_t* occ;
};

// Per-ray parity bookkeeping, kept while a trace asks for flags:
//   t1 / t2   the two smallest t of accepted candidates (closest modes), for near ties;
//   tmarg     the smallest t of a candidate whose outcome rests on a marginal decision (a
//             marginal leaf-test comparison, or a marginal cull on its chain); -INF when that
//             t is unknown or (occlusion) the candidate lies in the ray's interval;
//   robust    occlusion found by a candidate with no marginal decision.
struct Margins {
  std::vector<float> t1, t2, tmarg;
  std::vector<uint8_t> robust;
  void init(size_t n) {
    t1.assign(n, kInf);
    t2.assign(n, kInf);
    tmarg.assign(n, kInf);
    robust.assign(n, 0);
  }
};

// Leaf-level work shared by the RH and ray-bvh traversals.
struct LeafCtx {
  const orc_rays* r;
  const RH* h;
  int mode;
  Out out;
  orc_stats* st;
  Margins* mg;  // null unless flags are asked for
};

// Leaf tests of triangle T for the rays of leaf (k, j).  `shadow`: the pair is not a candidate
// (a cull on its chain rejected it) but that cull was marginal, so the tests only record
// marginal candidates; `chain_marg`: some cull on the chain was marginal.
inline void leaf_test(LeafCtx& c, int k, int j, const Tri& T, int32_t tid, bool shadow = false,
                      bool chain_marg = false) {
  const orc_rays* r = c.r;
  Margins* mg = c.mg;
  if (!shadow) c.st->leaf_pairs++;
  int x0 = k * c.h->cx, y0 = j * c.h->cy;
  int x1 = std::min(r->W, x0 + c.h->cx), y1 = std::min(r->H, y0 + c.h->cy);
  for (int y = y0; y < y1; ++y)
    for (int x = x0; x < x1; ++x) {
      size_t i = (size_t)y * r->W + x;
      if (c.mode != ORC_CLOSEST_PRIMARY && !ray_valid(r, i)) continue;
      if (c.mode == ORC_OCCLUDED && c.out.occ[i] && (!mg || mg->robust[i])) continue;
      if (!shadow) c.st->ray_tests++;
      float t, b1, b2;
      bool pm = false, t_known = true;
      const bool ok = mg ? mt_strict_m(ray_o(r, i), ray_d(r, i), T, t, b1, b2, pm, t_known)
                         : mt_strict(ray_o(r, i), ray_d(r, i), T, t, b1, b2);
      if (!ok) {
        if (!pm) continue;
        // a marginal rejection counts where the accepted test would have counted
        if (!t_known) {
          mg->tmarg[i] = -kInf;
        } else if (c.mode == ORC_OCCLUDED) {
          if (!(t <= r->tmin[i]) && !(t >= r->tmax[i])) mg->tmarg[i] = -kInf;
        } else if (t > r->tmin[i]) {
          mg->tmarg[i] = std::fmin(mg->tmarg[i], t);
        }
        continue;
      }
      bool m = pm || chain_marg || shadow;
      if (c.mode == ORC_OCCLUDED) {
        // intersectAllLeaves.cl:177-179
        const bool in = t > r->tmin[i] && t < r->tmax[i];
        if (mg) m |= near_thr(t, r->tmin[i]) || near_thr(t, r->tmax[i]);
        if (in && !shadow) {
          c.out.occ[i] = 1;
          if (mg && !m) mg->robust[i] = 1;
        }
        if (mg && m && (in || near_thr(t, r->tmin[i]) || near_thr(t, r->tmax[i])))
          mg->tmarg[i] = -kInf;
      } else {
        // intersectAllLeaves.cl:57-61 (triangles visited in increasing id order, so the
        // strict compare keeps the smallest id among equal t)
        const bool main = t > r->tmin[i];
        if (mg) m |= near_thr(t, r->tmin[i]);
        if (main && !shadow) {
          if (mg) {
            if (t < mg->t1[i]) {
              mg->t2[i] = mg->t1[i];
              mg->t1[i] = t;
            } else if (t < mg->t2[i]) {
              mg->t2[i] = t;
            }
          }
          if (t < c.out.t[i]) {
            c.out.t[i] = t;
            c.out.id[i] = tid;
            c.out.b1[i] = b1;
            c.out.b2[i] = b2;
          }
        }
        if (mg && m && (main || near_thr(t, r->tmin[i]))) mg->tmarg[i] = std::fmin(mg->tmarg[i], t);
      }
    }
}

inline bool rh_cull(const Node& n, const Tri& T, int mode, int node_kind, orc_stats* st) {
  st->cull_tests++;
  // 6DB primary rays use intersectsTri; every other cull is intersectsNode on the triangle's
  // bounds (the 5DB kernels, yetAnotherIntersection5DB.cl:78 included, use it in every mode)
  if (node_kind == 0 && mode == ORC_CLOSEST_PRIMARY) return cull_tri(n, T);
  return cull_box(node_kind, n, T.bmin, T.bmax);
}

// DFS of intersection6DB.cl:51-119 (and the IA / secondary / shadow variants) for one triangle.
// While flags are tracked, a cull that rejects on a marginal decision is descended anyway in
// shadow mode (no counts, no results; its leaf tests only record marginal candidates).
template <class F>
void rh_visit(const RH& h, int l, int k, int j, const Tri& T, int mode, int node_kind,
              bool use_valid, orc_stats* st, F&& at_leaf, bool shadow = false,
              bool chain_marg = false) {
  const Node& n = h.at(l, k, j);
  if (use_valid && !n.valid) return;
  orc_stats scratch{0, 0, 0, 0};
  g_marg = false;
  const bool pass = rh_cull(n, T, mode, node_kind, shadow ? &scratch : st);
  const bool m = g_track && g_marg;
  if (!pass) {
    if (!m) return;
    shadow = true;
  }
  chain_marg |= m;
  if (l == 0) {
    at_leaf(k, j, shadow, chain_marg);
    return;
  }
  for (int cj = 2 * j; cj <= 2 * j + 1; ++cj)
    for (int ck = 2 * k; ck <= 2 * k + 1; ++ck)
      if (ck < h.gw[l - 1] && cj < h.gh[l - 1])
        rh_visit(h, l - 1, ck, cj, T, mode, node_kind, use_valid, st, at_leaf, shadow, chain_marg);
}

void init_out(const orc_rays* r, int mode, Out& o) {
  size_t n = (size_t)r->W * r->H;
  for (size_t i = 0; i < n; ++i) {
    if (mode == ORC_OCCLUDED) {
      o.occ[i] = 0;
    } else {
      o.id[i] = -1;
      o.t[i] = kInf;
      o.b1[i] = 0.f;
      o.b2[i] = 0.f;
    }
  }
}

void trace_naive(const float* v9, uint32_t ntri, const orc_rays* r, int mode, Out& o,
                 orc_stats* st) {
  size_t n = (size_t)r->W * r->H;
  std::vector<Tri> tris(ntri);
  for (uint32_t i = 0; i < ntri; ++i) tris[i] = load_tri(v9, i);
  for (size_t i = 0; i < n; ++i) {
    // naive.cpp:244-245 (secondary rays with hit[k]==false are skipped), naive.cpp:364
    // (shadow rays are compacted to valid ones); primary rays are all traced.
    if (mode != ORC_CLOSEST_PRIMARY && !ray_valid(r, i)) continue;
    V3 org = ray_o(r, i), dir = ray_d(r, i);
    for (uint32_t k = 0; k < ntri; ++k) {
      float t, b1, b2;
      st->ray_tests++;
      if (mode == ORC_OCCLUDED) {
        // naive.cl:207-229
        if (!mt_naive_p(org, dir, tris[k], t, b1, b2)) continue;
        if (t < r->tmin[i] || t > r->tmax[i]) continue;
        o.occ[i] = 1;
        break;
      } else {
        // naive.cl:26-52: t >= mint, closer wins (ties -> smallest id)
        if (!mt_naive(org, dir, tris[k], t, b1, b2)) continue;
        if (t < r->tmin[i]) continue;
        if (o.id[i] != -1 && !(t < o.t[i])) continue;
        o.t[i] = t;
        o.id[i] = (int32_t)k;
        o.b1[i] = b1;
        o.b2[i] = b2;
      }
    }
  }
}

void trace_rh(const float* v9, uint32_t ntri, const orc_rays* r, const orc_params* p, int mode,
              Out& o, orc_stats* st, Margins* mg) {
  bool use_valid = mode != ORC_CLOSEST_PRIMARY;
  RH h = rh_build(r, p->cx, p->cy, p->height, use_valid, p->node);
  LeafCtx c{r, &h, mode, o, st, mg};
  int top = h.L - 1;
  for (uint32_t tid = 0; tid < ntri; ++tid) {
    Tri T = load_tri(v9, tid);
    for (int j = 0; j < h.gh[top]; ++j)
      for (int k = 0; k < h.gw[top]; ++k)
        rh_visit(h, top, k, j, T, mode, p->node, use_valid, st,
                 [&](int lk, int lj, bool shadow, bool cm) {
                   leaf_test(c, lk, lj, T, (int32_t)tid, shadow, cm);
                 });
  }
  if (mode != ORC_CLOSEST_PRIMARY || p->repair_run <= 0) return;
  // Repair pass (yetAnotherIntersection6DB.cl, intersectAllLeaves.cl:188-238), repairRun
  // times, triangles in increasing id order over the same candidate set.
  size_t n = (size_t)r->W * r->H;
  std::vector<int32_t> id0(o.id, o.id + n);
  orc_stats dummy{0, 0, 0, 0};
  const bool track = g_track;
  g_track = false;  // the flags describe the main pass
  for (int run = 0; run < p->repair_run; ++run)
    for (uint32_t tid = 0; tid < ntri; ++tid) {
      Tri T = load_tri(v9, tid);
      for (int j = 0; j < h.gh[top]; ++j)
        for (int k = 0; k < h.gw[top]; ++k)
          rh_visit(h, top, k, j, T, mode, p->node, false, &dummy, [&](int lk, int lj, bool shadow,
                                                                      bool) {
            if (shadow) return;
            int x0 = lk * h.cx, y0 = lj * h.cy;
            int x1 = std::min(r->W, x0 + h.cx), y1 = std::min(r->H, y0 + h.cy);
            for (int y = y0; y < y1; ++y)
              for (int x = x0; x < x1; ++x) {
                size_t i = (size_t)y * r->W + x;
                float t, b1, b2;
                if (!mt_repair(ray_o(r, i), ray_d(r, i), T, t, b1, b2)) continue;
                if (t < r->tmin[i]) continue;
                if (t > o.t[i] - kEps || o.id[i] == (int32_t)tid) continue;
                o.t[i] = t;
                o.id[i] = (int32_t)tid;
                o.b1[i] = b1;
                o.b2[i] = b2;
              }
          });
    }
  g_track = track;
  for (size_t i = 0; i < n; ++i)
    if (o.id[i] != id0[i]) st->repair_rays++;
}

// Lockstep RH x BVH product traversal of intersection6DB_BVH.cl:51-223: the RH node descends
// while it is interior, the BVH node while it is interior; leaf x leaf pairs test the leaf's
// primitives (only the first one for secondary rays, intersection6DB_BVH2.cl:88).
struct BvhCtx {
  const RH* h;
  const orc_gpu_node* nodes;
  const uint32_t* prim_order;
  const float* v9;
  int mode, node_kind;
  bool use_valid;
  LeafCtx* leaf;
  orc_stats* st;
};

// One step of the lockstep product traversal: the cull of (hierarchy node (l, k, j), BVH node
// ni).  Returns false when the pair stops the walk; `shadow` / `chain_marg` as in rh_visit.
inline bool bvh_step(BvhCtx& c, int l, int k, int j, uint32_t ni, bool& shadow, bool& chain_marg) {
  const orc_gpu_node& b = c.nodes[ni];
  if (bvh_dead(c.node_kind, b)) return false;
  const Node& n = c.h->at(l, k, j);
  if (c.use_valid && !n.valid) return false;
  V3 bmin = {b.ax, b.ay, b.az}, bmax = {b.bx, b.by, b.bz};
  if (!shadow) c.st->cull_tests++;
  g_marg = false;
  bool pass = cull_box(c.node_kind, n, bmin, bmax);
  const bool m = g_track && g_marg;
  if (!pass) {
    if (!m) return false;
    shadow = true;
  }
  chain_marg |= m;
  return true;
}

// Calls emit(k, j, bvh leaf, shadow, chain_marg) for every (hierarchy leaf, BVH leaf) pair the
// lockstep walk from (l, k, j, ni) reaches.
template <class F>
void bvh_walk(BvhCtx& c, int l, int k, int j, uint32_t ni, bool shadow, bool chain_marg, F&& emit) {
  if (!bvh_step(c, l, k, j, ni, shadow, chain_marg)) return;
  const orc_gpu_node& b = c.nodes[ni];
  bool rleaf = l == 0, bleaf = b.nPrimitives > 0;
  if (rleaf && bleaf) {
    emit(k, j, ni, shadow, chain_marg);
  } else if (rleaf) {
    bvh_walk(c, l, k, j, b.primOffset, shadow, chain_marg, emit);
    bvh_walk(c, l, k, j, b.primOffset + 1, shadow, chain_marg, emit);
  } else {
    for (int cj = 2 * j; cj <= 2 * j + 1; ++cj)
      for (int ck = 2 * k; ck <= 2 * k + 1; ++ck) {
        if (ck >= c.h->gw[l - 1] || cj >= c.h->gh[l - 1]) continue;
        if (bleaf) {
          bvh_walk(c, l - 1, ck, cj, ni, shadow, chain_marg, emit);
        } else {
          bvh_walk(c, l - 1, ck, cj, b.primOffset, shadow, chain_marg, emit);
          bvh_walk(c, l - 1, ck, cj, b.primOffset + 1, shadow, chain_marg, emit);
        }
      }
  }
}

// ---- pbrt BVH build (bvh.cpp:33-455) -----------------------------------------------------
struct BBoxP {
  V3 pMin{kInf, kInf, kInf}, pMax{-kInf, -kInf, -kInf};
  float& lo(int d) { return d == 0 ? pMin.x : d == 1 ? pMin.y : pMin.z; }
  float get_lo(int d) const { return d == 0 ? pMin.x : d == 1 ? pMin.y : pMin.z; }
  float get_hi(int d) const { return d == 0 ? pMax.x : d == 1 ? pMax.y : pMax.z; }
  float SurfaceArea() const {
    V3 d = sub(pMax, pMin);
    return 2.f * (d.x * d.y + d.x * d.z + d.y * d.z);
  }
  int MaximumExtent() const {
    V3 d = sub(pMax, pMin);
    if (d.x > d.y && d.x > d.z) return 0;
    if (d.y > d.z) return 1;
    return 2;
  }
};
// std::min / std::max semantics (pbrt uses them in Union, geometry.cpp)
inline float smin(float a, float b) { return (b < a) ? b : a; }
inline float smax(float a, float b) { return (a < b) ? b : a; }
inline BBoxP bunion(const BBoxP& b, V3 p) {
  BBoxP r;
  r.pMin = {smin(b.pMin.x, p.x), smin(b.pMin.y, p.y), smin(b.pMin.z, p.z)};
  r.pMax = {smax(b.pMax.x, p.x), smax(b.pMax.y, p.y), smax(b.pMax.z, p.z)};
  return r;
}
inline BBoxP bunion(const BBoxP& a, const BBoxP& b) {
  BBoxP r;
  r.pMin = {smin(a.pMin.x, b.pMin.x), smin(a.pMin.y, b.pMin.y), smin(a.pMin.z, b.pMin.z)};
  r.pMax = {smax(a.pMax.x, b.pMax.x), smax(a.pMax.y, b.pMax.y), smax(a.pMax.z, b.pMax.z)};
  return r;
}
inline float vget(V3 v, int d) { return d == 0 ? v.x : d == 1 ? v.y : v.z; }

struct PrimInfo {
  int primitiveNumber;
  V3 centroid;
  BBoxP bounds;
};
struct BuildNode {
  BBoxP bounds;
  int children[2] = {-1, -1};
  uint32_t splitAxis = 0, firstPrimOffset = 0, nPrimitives = 0;
};

struct BvhBuilder {
  uint32_t maxPrims;
  int split = 0;  // 0 sah, 1 middle, 2 equal (bvh.cpp:214-221)
  std::vector<BuildNode> nodes;
  std::vector<uint32_t> ordered;  // orderedPrims, with the repeats of the SAH fall-through

  int build(std::vector<PrimInfo>& bd, uint32_t start, uint32_t end) {
    int me = (int)nodes.size();
    nodes.emplace_back();
    BBoxP bbox;
    for (uint32_t i = start; i < end; ++i) bbox = bunion(bbox, bd[i].bounds);
    uint32_t nPrims = end - start;
    auto make_leaf = [&]() {
      uint32_t first = (uint32_t)ordered.size();
      for (uint32_t i = start; i < end; ++i) ordered.push_back((uint32_t)bd[i].primitiveNumber);
      nodes[me].firstPrimOffset = first;
      nodes[me].nPrimitives = nPrims;
      nodes[me].bounds = bbox;
    };
    if (nPrims == 1) {
      make_leaf();
      return me;
    }
    BBoxP cb;
    for (uint32_t i = start; i < end; ++i) cb = bunion(cb, bd[i].centroid);
    int dim = cb.MaximumExtent();
    uint32_t mid = (start + end) / 2;
    if (cb.get_hi(dim) == cb.get_lo(dim)) {
      make_leaf();
      return me;
    }
    auto equal_counts = [&]() {
      mid = (start + end) / 2;
      std::nth_element(&bd[start], &bd[mid], &bd[end - 1] + 1,
                       [dim](const PrimInfo& a, const PrimInfo& b) {
                         return vget(a.centroid, dim) < vget(b.centroid, dim);
                       });
    };
    if (split == 1) {
      // SPLIT_MIDDLE, bvh.cpp:333-340: CompareToMid is centroid < pmid.  An empty side (the
      // midpoint rounded onto an end of the range) makes the reference recurse without end;
      // equal counts stand in for that case.
      const float pmid = .5f * (cb.get_lo(dim) + cb.get_hi(dim));
      PrimInfo* pm = std::partition(&bd[start], &bd[end - 1] + 1, [&](const PrimInfo& p) {
        return vget(p.centroid, dim) < pmid;
      });
      mid = (uint32_t)(pm - &bd[0]);
      if (mid == start || mid == end) equal_counts();
    } else if (split == 2 || nPrims <= 4) {
      equal_counts();  // SPLIT_EQUAL_COUNTS (bvh.cpp:342-347), small SAH nodes (:351-356)
    } else {
      const int nB = 12;
      int cnt[nB] = {0};
      BBoxP bb[nB];
      float lo = cb.get_lo(dim), hi = cb.get_hi(dim);
      for (uint32_t i = start; i < end; ++i) {
        int b = (int)(nB * ((vget(bd[i].centroid, dim) - lo) / (hi - lo)));
        if (b == nB) b = nB - 1;
        cnt[b]++;
        bb[b] = bunion(bb[b], bd[i].bounds);
      }
      float cost[nB - 1];
      for (int i = 0; i < nB - 1; ++i) {
        BBoxP b0, b1;
        int c0 = 0, c1 = 0;
        for (int j = 0; j <= i; ++j) {
          b0 = bunion(b0, bb[j]);
          c0 += cnt[j];
        }
        for (int j = i + 1; j < nB; ++j) {
          b1 = bunion(b1, bb[j]);
          c1 += cnt[j];
        }
        cost[i] = .125f + (c0 * b0.SurfaceArea() + c1 * b1.SurfaceArea()) / bbox.SurfaceArea();
      }
      float minCost = cost[0];
      int best = 0;
      for (int i = 1; i < nB - 1; ++i)
        if (cost[i] < minCost) {
          minCost = cost[i];
          best = i;
        }
      if (nPrims > maxPrims || minCost < nPrims) {
        PrimInfo* pm = std::partition(&bd[start], &bd[end - 1] + 1, [&](const PrimInfo& p) {
          int b = (int)(nB * ((vget(p.centroid, dim) - lo) / (hi - lo)));
          if (b == nB) b = nB - 1;
          return b <= best;
        });
        mid = (uint32_t)(pm - &bd[0]);
      } else {
        // bvh.cpp:417-428: the primitives are recorded as a leaf, but without a return the
        // code goes on to InitInterior with mid = (start + end) / 2 and builds both halves,
        // which records them again (reachable only with maxPrims > 4)
        make_leaf();
        mid = (start + end) / 2;
      }
    }
    int c0 = build(bd, start, mid);
    int c1 = build(bd, mid, end);
    nodes[me].children[0] = c0;
    nodes[me].children[1] = c1;
    nodes[me].bounds = bunion(nodes[c0].bounds, nodes[c1].bounds);
    nodes[me].splitAxis = (uint32_t)dim;
    nodes[me].nPrimitives = 0;
    return me;
  }

  void run(const float* v9, uint32_t ntri) {
    std::vector<PrimInfo> bd(ntri);
    for (uint32_t i = 0; i < ntri; ++i) {
      const float* p = v9 + 9 * (size_t)i;
      V3 a{p[0], p[1], p[2]}, b{p[3], p[4], p[5]}, c{p[6], p[7], p[8]};
      BBoxP bb;
      bb.pMin = {smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)};
      bb.pMax = {smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)};
      bb = bunion(bb, c);
      bd[i].primitiveNumber = (int)i;
      bd[i].bounds = bb;
      bd[i].centroid = {.5f * bb.pMin.x + .5f * bb.pMax.x, .5f * bb.pMin.y + .5f * bb.pMax.y,
                        .5f * bb.pMin.z + .5f * bb.pMax.z};
    }
    nodes.reserve(2 * (size_t)ntri);
    if (ntri) build(bd, 0, ntri);
  }

  // GPUbufferWide (bvh.cpp:154-206)
  uint32_t gpu_wide(uint32_t height, orc_gpu_node* out, uint32_t* top) {
    struct Tmp {
      int ptr, parent;
      uint32_t level, childA;
    };
    size_t total = nodes.size();
    std::vector<Tmp> a(total);
    a[0] = {0, -1, 0, 0};
    uint32_t write = 1;
    int totalHeight = 0;
    for (size_t i = 0; i < total; ++i) {
      const BuildNode& bn = nodes[a[i].ptr];
      if (bn.nPrimitives > 0) continue;
      a[write] = {bn.children[0], -1, a[i].level + 1, 0};
      a[write + 1] = {bn.children[1], -1, a[i].level + 1, 0};
      a[i].childA = write;
      totalHeight = (int)a[i].level + 1;
      write += 2;
    }
    totalHeight = std::max(totalHeight - (int)height, 0);
    uint32_t nodeNum = 0, topLevel = 0;
    for (size_t i = 0; i < total; ++i) {
      const BuildNode& bn = nodes[a[i].ptr];
      if ((int)a[i].level < totalHeight && !bn.nPrimitives) continue;
      if ((int)a[i].level <= totalHeight) ++topLevel;
      if (a[i].parent != -1) out[a[i].parent].primOffset = nodeNum;
      out[nodeNum].ax = bn.bounds.pMin.x;
      out[nodeNum].ay = bn.bounds.pMin.y;
      out[nodeNum].az = bn.bounds.pMin.z;
      out[nodeNum].bx = bn.bounds.pMax.x;
      out[nodeNum].by = bn.bounds.pMax.y;
      out[nodeNum].bz = bn.bounds.pMax.z;
      out[nodeNum].nPrimitives = bn.nPrimitives;
      if (bn.nPrimitives == 0) a[a[i].childA].parent = (int)nodeNum;
      if (bn.nPrimitives) out[nodeNum].primOffset = bn.firstPrimOffset;
      ++nodeNum;
    }
    *top = topLevel;
    return nodeNum;
  }
};

// ---- pbrt CPU BVH (flattenBVHTree + Intersect/IntersectP, bvh.cpp:437-565) ---------------
struct LinNode {
  BBoxP bounds;
  uint32_t offset;  // leaf: primitivesOffset, interior: secondChildOffset
  uint32_t nPrimitives, axis;
};
struct CpuBvh {
  std::vector<LinNode> nodes;
  std::vector<uint32_t> order;  // BVH position -> input index
  std::vector<Tri> tris;        // BVH order
};

uint32_t flatten(const BvhBuilder& b, int ni, std::vector<LinNode>& out, uint32_t* offset) {
  const BuildNode& n = b.nodes[ni];
  uint32_t me = (*offset)++;
  out[me].bounds = n.bounds;
  if (n.nPrimitives > 0) {
    out[me].offset = n.firstPrimOffset;
    out[me].nPrimitives = n.nPrimitives;
    out[me].axis = 0;
  } else {
    out[me].axis = n.splitAxis;
    out[me].nPrimitives = 0;
    flatten(b, n.children[0], out, offset);
    out[me].offset = flatten(b, n.children[1], out, offset);
  }
  return me;
}

inline bool box_hit(const BBoxP& b, V3 o, V3 inv, const uint32_t neg[3], float mint, float maxt) {
  const V3* bounds[2] = {&b.pMin, &b.pMax};
  float tmin = (bounds[neg[0]]->x - o.x) * inv.x;
  float tmax = (bounds[1 - neg[0]]->x - o.x) * inv.x;
  float tymin = (bounds[neg[1]]->y - o.y) * inv.y;
  float tymax = (bounds[1 - neg[1]]->y - o.y) * inv.y;
  if ((tmin > tymax) || (tymin > tmax)) return false;
  if (tymin > tmin) tmin = tymin;
  if (tymax < tmax) tmax = tymax;
  float tzmin = (bounds[neg[2]]->z - o.z) * inv.z;
  float tzmax = (bounds[1 - neg[2]]->z - o.z) * inv.z;
  if ((tmin > tzmax) || (tzmin > tmax)) return false;
  if (tzmin > tmin) tmin = tzmin;
  if (tzmax < tmax) tmax = tzmax;
  return (tmin < maxt) && (tmax > mint);
}

// BVHAccel::Intersect (bvh.cpp:464-516); Triangle::Intersect shrinks maxt on each hit.
bool cpu_closest(const CpuBvh& b, V3 o, V3 d, float mint, float maxt, int32_t& id, float& tt,
                 float& ob1, float& ob2) {
  if (b.nodes.empty()) return false;
  bool hit = false;
  V3 inv = {1.f / d.x, 1.f / d.y, 1.f / d.z};
  uint32_t neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
  uint32_t todo[64], to = 0, ni = 0;
  while (true) {
    const LinNode& n = b.nodes[ni];
    if (box_hit(n.bounds, o, inv, neg, mint, maxt)) {
      if (n.nPrimitives > 0) {
        for (uint32_t i = 0; i < n.nPrimitives; ++i) {
          float t, b1, b2;
          const Tri& T = b.tris[n.offset + i];
          if (mt_pbrt(o, d, T, t, b1, b2) && !(t < mint || t > maxt)) {
            maxt = t;
            id = (int32_t)b.order[n.offset + i];
            tt = t;
            ob1 = b1;
            ob2 = b2;
            hit = true;
          }
        }
        if (to == 0) break;
        ni = todo[--to];
      } else {
        if (neg[n.axis]) {
          todo[to++] = ni + 1;
          ni = n.offset;
        } else {
          todo[to++] = n.offset;
          ni = ni + 1;
        }
      }
    } else {
      if (to == 0) break;
      ni = todo[--to];
    }
  }
  return hit;
}

bool cpu_occluded(const CpuBvh& b, V3 o, V3 d, float mint, float maxt) {
  if (b.nodes.empty()) return false;
  V3 inv = {1.f / d.x, 1.f / d.y, 1.f / d.z};
  uint32_t neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
  uint32_t todo[64], to = 0, ni = 0;
  while (true) {
    const LinNode& n = b.nodes[ni];
    if (box_hit(n.bounds, o, inv, neg, mint, maxt)) {
      if (n.nPrimitives > 0) {
        for (uint32_t i = 0; i < n.nPrimitives; ++i) {
          float t, b1, b2;
          if (mt_pbrt(o, d, b.tris[n.offset + i], t, b1, b2) && !(t < mint || t > maxt))
            return true;
        }
        if (to == 0) break;
        ni = todo[--to];
      } else {
        if (neg[n.axis]) {
          todo[to++] = ni + 1;
          ni = n.offset;
        } else {
          todo[to++] = n.offset;
          ni = ni + 1;
        }
      }
    } else {
      if (to == 0) break;
      ni = todo[--to];
    }
  }
  return false;
}

template <class F>
void parallel_for(size_t n, int nthreads, F&& fn) {
  if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  std::atomic<size_t> next{0};
  const size_t chunk = 256;
  auto work = [&]() {
    while (true) {
      size_t s = next.fetch_add(chunk);
      if (s >= n) break;
      size_t e = std::min(n, s + chunk);
      for (size_t i = s; i < e; ++i) fn(i);
    }
  };
  if (nthreads == 1) {
    work();
    return;
  }
  std::vector<std::thread> th;
  for (int i = 0; i < nthreads; ++i) th.emplace_back(work);
  for (auto& t : th) t.join();
}

// ---- per-ray path restatement ------------------------------------------------------------
// The library traces incoherent ray groups ray by ray: each ray walks a search BVH of the
// triangles (binned SAH, one triangle per leaf, boxes grown by 2^-11 * max(1, largest
// |coordinate|)), runs the strict leaf test on the triangles of the leaves it reaches and
// evaluates the reference's cull chain for the ones that pass.  This is a CPU restatement of
// that traversal (not of the reference): its results must equal the reference-semantics
// traces above, and its per-ray counts (interior nodes loaded, leaf tests, chain culls) are
// the work model of the per-ray kernel, so the tree is built by the same rules as the
// library's.
constexpr int kSDepth = 26;
constexpr int kSLeaf = 1;

struct SBoxC {
  float lo[3], hi[3];
};
struct SNodeC {
  SBoxC b[2];
  uint32_t ref[2];
};

struct SearchTree {
  std::vector<uint32_t> perm;  // sorted slot -> triangle id
  std::vector<Tri> tri;        // sorted order
  std::vector<SNodeC> nodes;
};

int s_levels(uint32_t n) {
  int l = 0;
  while (n > (uint32_t)kSLeaf) {
    n = (n + 1) / 2;
    ++l;
  }
  return l;
}

void s_tribox(const SearchTree& T, const float* v9, uint32_t a, uint32_t b, SBoxC& box) {
  for (int k = 0; k < 3; ++k) {
    box.lo[k] = kInf;
    box.hi[k] = -kInf;
  }
  for (uint32_t i = a; i < b; ++i) {
    const float* p = v9 + 9 * (size_t)T.perm[i];
    for (int v = 0; v < 3; ++v)
      for (int k = 0; k < 3; ++k) {
        box.lo[k] = std::min(box.lo[k], p[3 * v + k]);
        box.hi[k] = std::max(box.hi[k], p[3 * v + k]);
      }
  }
}

uint32_t s_build(SearchTree& T, const float* v9, const std::vector<uint32_t>& code, uint32_t a,
                 uint32_t b, int depth, float pad, SBoxC& box) {
  const uint32_t n = b - a;
  if (n <= (uint32_t)kSLeaf) {
    s_tribox(T, v9, a, b, box);
    return 0x80000000u | ((n - 1) << 27) | a;
  }
  const uint32_t node = (uint32_t)T.nodes.size();
  T.nodes.emplace_back();
  uint32_t m = a + n / 2;
  if (code[a] != code[b - 1]) {
    const int bit = 31 - __builtin_clz(code[a] ^ code[b - 1]);
    uint32_t lo = a, hi = b - 1;
    while (lo < hi) {
      const uint32_t md = lo + (hi - lo) / 2;
      if (code[md] & (1u << bit)) hi = md;
      else lo = md + 1;
    }
    if (depth + s_levels(lo - a) <= kSDepth && depth + s_levels(b - lo) <= kSDepth) m = lo;
  }
  SBoxC b0, b1;
  const uint32_t r0 = s_build(T, v9, code, a, m, depth + 1, pad, b0);
  const uint32_t r1 = s_build(T, v9, code, m, b, depth + 1, pad, b1);
  SNodeC& nd = T.nodes[node];
  for (int k = 0; k < 3; ++k) {
    nd.b[0].lo[k] = b0.lo[k] - pad;
    nd.b[0].hi[k] = b0.hi[k] + pad;
    nd.b[1].lo[k] = b1.lo[k] - pad;
    nd.b[1].hi[k] = b1.hi[k] + pad;
    box.lo[k] = std::min(b0.lo[k], b1.lo[k]);
    box.hi[k] = std::max(b0.hi[k], b1.hi[k]);
  }
  nd.ref[0] = r0;
  nd.ref[1] = r1;
  return node;
}

// Binned SAH build: 16 bins on each axis with a centroid extent, split cost nL*area(L) +
// nR*area(R) (the first minimum over the axes x, y, z in turn and the planes in order),
// stable partitions.  A range of at most kSLeaf triangles is a leaf unless its best split
// costs less than testing them all (node and triangle cost 1 per unit area; with kSLeaf = 1
// every leaf holds one triangle and the comparison never runs); a range whose
// centroids coincide on every axis, or whose SAH split would take the tree past kSDepth
// interior levels, splits at the median by (centroid, id) on the axis of largest centroid
// extent.
struct SPrim {
  float lo[3], hi[3], c[3];
};
constexpr float kSNodeCost = 1.f, kSTriCost = 1.f;

float s_area(const SBoxC& b) {
  const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
  return 2.f * ((dx * dy + dy * dz) + dz * dx);
}

void s_grow(SBoxC& b, const SPrim& p) {
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = std::min(b.lo[k], p.lo[k]);
    b.hi[k] = std::max(b.hi[k], p.hi[k]);
  }
}

void s_empty(SBoxC& b) {
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = kInf;
    b.hi[k] = -kInf;
  }
}

uint32_t s_build_sah(SearchTree& T, const std::vector<SPrim>& P, uint32_t a, uint32_t b, int depth,
                     float pad, SBoxC& box) {
  const uint32_t n = b - a;
  uint32_t* idx = T.perm.data();
  if (n == 1) {
    s_empty(box);
    for (uint32_t i = a; i < b; ++i) s_grow(box, P[idx[i]]);
    return 0x80000000u | ((n - 1) << 27) | a;
  }
  float clo[3] = {kInf, kInf, kInf}, chi[3] = {-kInf, -kInf, -kInf};
  for (uint32_t i = a; i < b; ++i)
    for (int k = 0; k < 3; ++k) {
      clo[k] = std::min(clo[k], P[idx[i]].c[k]);
      chi[k] = std::max(chi[k], P[idx[i]].c[k]);
    }
  int ax = 0;
  for (int k = 1; k < 3; ++k)
    if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
  constexpr int kB = 16;
  auto bin_of = [&](uint32_t t, int q) {
    const float scale = (float)kB / (chi[q] - clo[q]);
    const int k = (int)((P[t].c[q] - clo[q]) * scale);
    return std::min(std::max(k, 0), kB - 1);
  };
  int best_q = -1, best = -1;
  float best_cost = kInf;
  for (int q = 0; q < 3; ++q) {
    if (!(chi[q] - clo[q] > 0.f)) continue;
    SBoxC bb[kB];
    uint32_t cnt[kB] = {0};
    for (int k = 0; k < kB; ++k) s_empty(bb[k]);
    for (uint32_t i = a; i < b; ++i) {
      const int k = bin_of(idx[i], q);
      s_grow(bb[k], P[idx[i]]);
      ++cnt[k];
    }
    // costs of the kB - 1 planes: left side accumulated upward, right side downward
    float larea[kB], rarea[kB];
    uint32_t lcnt[kB], rcnt[kB];
    SBoxC acc;
    s_empty(acc);
    uint32_t c = 0;
    for (int k = 0; k < kB; ++k) {
      c += cnt[k];
      if (cnt[k])
        for (int z = 0; z < 3; ++z) {
          acc.lo[z] = std::min(acc.lo[z], bb[k].lo[z]);
          acc.hi[z] = std::max(acc.hi[z], bb[k].hi[z]);
        }
      larea[k] = c ? s_area(acc) : 0.f;
      lcnt[k] = c;
    }
    s_empty(acc);
    c = 0;
    for (int k = kB - 1; k >= 0; --k) {
      c += cnt[k];
      if (cnt[k])
        for (int z = 0; z < 3; ++z) {
          acc.lo[z] = std::min(acc.lo[z], bb[k].lo[z]);
          acc.hi[z] = std::max(acc.hi[z], bb[k].hi[z]);
        }
      rarea[k] = c ? s_area(acc) : 0.f;
      rcnt[k] = c;
    }
    for (int k = 0; k + 1 < kB; ++k) {
      if (lcnt[k] == 0 || rcnt[k + 1] == 0) continue;
      const float cost = (float)lcnt[k] * larea[k] + (float)rcnt[k + 1] * rarea[k + 1];
      if (cost < best_cost) {
        best_cost = cost;
        best = k;
        best_q = q;
      }
    }
  }
  uint32_t m = 0;
  bool median = best < 0;
  if (n <= (uint32_t)kSLeaf) {
    SBoxC all;
    s_empty(all);
    for (uint32_t i = a; i < b; ++i) s_grow(all, P[idx[i]]);
    const float A = s_area(all);
    if (best < 0 || !(kSNodeCost * A + kSTriCost * best_cost < kSTriCost * (float)n * A)) {
      box = all;
      return 0x80000000u | ((n - 1) << 27) | a;
    }
  }
  if (!median) {
    uint32_t* mid = std::stable_partition(idx + a, idx + b,
                                          [&](uint32_t t) { return bin_of(t, best_q) <= best; });
    m = (uint32_t)(mid - idx);
    if (depth + s_levels(m - a) > kSDepth || depth + s_levels(b - m) > kSDepth) median = true;
  }
  if (median && n <= (uint32_t)kSLeaf) {
    s_empty(box);
    for (uint32_t i = a; i < b; ++i) s_grow(box, P[idx[i]]);
    return 0x80000000u | ((n - 1) << 27) | a;
  }
  if (median) {
    std::stable_sort(idx + a, idx + b, [&](uint32_t s0, uint32_t s1) {
      return P[s0].c[ax] < P[s1].c[ax] || (P[s0].c[ax] == P[s1].c[ax] && s0 < s1);
    });
    m = a + n / 2;
  }
  const uint32_t node = (uint32_t)T.nodes.size();
  T.nodes.emplace_back();
  SBoxC b0, b1;
  const uint32_t r0 = s_build_sah(T, P, a, m, depth + 1, pad, b0);
  const uint32_t r1 = s_build_sah(T, P, m, b, depth + 1, pad, b1);
  SNodeC& nd = T.nodes[node];
  for (int k = 0; k < 3; ++k) {
    nd.b[0].lo[k] = b0.lo[k] - pad;
    nd.b[0].hi[k] = b0.hi[k] + pad;
    nd.b[1].lo[k] = b1.lo[k] - pad;
    nd.b[1].hi[k] = b1.hi[k] + pad;
    box.lo[k] = std::min(b0.lo[k], b1.lo[k]);
    box.hi[k] = std::max(b0.hi[k], b1.hi[k]);
  }
  nd.ref[0] = r0;
  nd.ref[1] = r1;
  return node;
}

SearchTree search_tree(const float* v9, uint32_t ntri) {
  SearchTree T;
  T.perm.resize(ntri);
  for (uint32_t i = 0; i < ntri; ++i) T.perm[i] = i;
  float big = 1.f;
  for (size_t i = 0; i < 9 * (size_t)ntri; ++i) big = std::max(big, std::fabs(v9[i]));
  const float pad = std::ldexp(big, -11);
  if (ntri == 0) return T;
  std::vector<SPrim> P(ntri);
  for (uint32_t i = 0; i < ntri; ++i) {
    const float* p = v9 + 9 * (size_t)i;
    for (int k = 0; k < 3; ++k) {
      P[i].lo[k] = std::min(std::min(p[k], p[3 + k]), p[6 + k]);
      P[i].hi[k] = std::max(std::max(p[k], p[3 + k]), p[6 + k]);
      P[i].c[k] = (P[i].lo[k] + P[i].hi[k]) * 0.5f;
    }
  }
  if (ntri <= (uint32_t)kSLeaf) {
    const uint32_t m = ntri == 1 ? 1 : ntri / 2;
    SNodeC nd;
    s_tribox(T, v9, 0, m, nd.b[0]);
    nd.ref[0] = 0x80000000u | ((m - 1) << 27);
    nd.b[1] = nd.b[0];
    nd.ref[1] = nd.ref[0];
    if (ntri > 1) {
      s_tribox(T, v9, m, ntri, nd.b[1]);
      nd.ref[1] = 0x80000000u | ((ntri - m - 1) << 27) | m;
    }
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 3; ++k) {
        nd.b[j].lo[k] -= pad;
        nd.b[j].hi[k] += pad;
      }
    T.nodes.push_back(nd);
  } else {
    SBoxC box;
    s_build_sah(T, P, 0, ntri, 1, pad, box);
  }
  T.tri.resize(ntri);
  for (uint32_t i = 0; i < ntri; ++i) T.tri[i] = load_tri(v9, T.perm[i]);
  return T;
}

constexpr float kSlackC = 0x1p-16f;
inline bool s_slab(V3 o, V3 inv, const SBoxC& b, float tlo, float thi, float& tn) {
  const float ax = (b.lo[0] - o.x) * inv.x, bx = (b.hi[0] - o.x) * inv.x;
  const float ay = (b.lo[1] - o.y) * inv.y, by = (b.hi[1] - o.y) * inv.y;
  const float az = (b.lo[2] - o.z) * inv.z, bz = (b.hi[2] - o.z) * inv.z;
  tn = std::fmax(std::fmax(std::fmin(ax, bx), std::fmin(ay, by)), std::fmin(az, bz));
  float tf = std::fmin(std::fmin(std::fmax(ax, bx), std::fmax(ay, by)), std::fmax(az, bz));
  tf = tf + std::fabs(tf) * kSlackC;
  return tn <= tf && tn <= thi && tf >= tlo;
}

struct SearchCtx {
  const SearchTree* T;
  const RH* h;
  int mode, node_kind, kind;
  bool repair;
  // ray-bvh: the reference BVH, its leaf of every triangle (bit 31: not first), parents, depths
  const orc_gpu_node* nodes;
  std::vector<uint32_t> tleaf, parent;
  std::vector<uint8_t> depth;
};

// The reference's chain for (the ray's leaf at image position (x, y), triangle or BVH leaf).
bool s_chain(const SearchCtx& c, int x, int y, uint32_t tri, uint32_t leaf, uint64_t& culls) {
  const RH& h = *c.h;
  const int top = h.L - 1;
  auto rnode = [&](int l) -> const Node& {
    return h.at(l, x / (h.cx << l), y / (h.cy << l));
  };
  if (c.kind == 1) {
    const Tri& T = c.T->tri[tri];
    for (int l = top; l >= 0; --l) {
      const Node& n = rnode(l);
      ++culls;
      if (!n.valid) return false;
      bool pass = c.node_kind == 0 && c.mode == ORC_CLOSEST_PRIMARY
                      ? cull_tri(n, T)
                      : cull_box(c.node_kind, n, T.bmin, T.bmax);
      if (!pass) return false;
    }
    return true;
  }
  const int d = c.depth[leaf];
  uint32_t node = leaf;
  int dn = d;
  for (int st = std::max(top, d); st >= 0; --st) {
    const int bd = std::min(st, d);
    while (dn > bd) {
      node = c.parent[node];
      --dn;
    }
    const Node& n = rnode(top - std::min(st, top));
    ++culls;
    if (!n.valid) return false;
    const orc_gpu_node& b = c.nodes[node];
    if (bvh_dead(c.node_kind, b)) return false;
    const V3 bmin = {b.ax, b.ay, b.az}, bmax = {b.bx, b.by, b.bz};
    if (!cull_box(c.node_kind, n, bmin, bmax)) return false;
  }
  return true;
}

struct SCount {
  uint64_t rays = 0, visits = 0, tests = 0, culls = 0, chains = 0;
};

void s_trace_ray(const SearchCtx& c, const orc_rays* r, size_t i, Out& o, SCount& cnt) {
  const int x = (int)(i % (size_t)r->W), y = (int)(i / (size_t)r->W);
  const V3 org = ray_o(r, i), d = ray_d(r, i);
  const float tmin = r->tmin[i], tmax = r->tmax[i];
  auto safe = [](float v) { return std::fabs(v) < 0x1p-80f ? (v < 0.f ? -0x1p-80f : 0x1p-80f) : v; };
  const V3 inv = {1.0f / safe(d.x), 1.0f / safe(d.y), 1.0f / safe(d.z)};
  const float tlo = tmin - std::fabs(tmin) * kSlackC;
  ++cnt.rays;
  float best_t = kInf;
  int32_t best_id = -1;
  float bb1 = 0.f, bb2 = 0.f;
  bool occluded = false;
  uint32_t stack[kSDepth + 2];
  int sp = 0;
  uint32_t cur = 0;
  const SearchTree& T = *c.T;
  for (;;) {
    if (!(cur & 0x80000000u)) {
      ++cnt.visits;
      const SNodeC& nd = T.nodes[cur];
      float thi = c.mode == ORC_OCCLUDED ? tmax : best_t;
      thi = thi + std::fabs(thi) * kSlackC;
      float tn0, tn1;
      const bool h0 = s_slab(org, inv, nd.b[0], tlo, thi, tn0);
      const bool h1 = s_slab(org, inv, nd.b[1], tlo, thi, tn1);
      if (h0 && h1) {
        const bool first1 = tn1 < tn0;
        stack[sp++] = first1 ? nd.ref[0] : nd.ref[1];
        cur = first1 ? nd.ref[1] : nd.ref[0];
        continue;
      }
      if (h0 || h1) {
        cur = h0 ? nd.ref[0] : nd.ref[1];
        continue;
      }
    } else {
      const uint32_t first = cur & ((1u << 27) - 1u), n = ((cur >> 27) & 15u) + 1u;
      for (uint32_t f = 0; f < n && !occluded; ++f) {
        const uint32_t slot = first + f;
        const uint32_t id = T.perm[slot];
        uint32_t leaf = 0;
        if (c.kind == 2) {
          leaf = c.tleaf[id];
          if (leaf == 0xffffffffu) continue;
          if (c.mode == ORC_CLOSEST_SECONDARY && c.node_kind != 2 && (leaf & 0x80000000u))
            continue;
          leaf &= 0x7fffffffu;
        }
        ++cnt.tests;
        float t, b1, b2;
        if (!mt_strict(org, d, T.tri[slot], t, b1, b2)) continue;
        if (c.mode == ORC_OCCLUDED) {
          if (!(t > tmin && t < tmax)) continue;
          ++cnt.chains;
          occluded = s_chain(c, x, y, slot, leaf, cnt.culls);
          continue;
        }
        if (!(t > tmin)) continue;
        if (!(t < best_t || (t == best_t && (int32_t)id < best_id))) continue;
        ++cnt.chains;
        if (!s_chain(c, x, y, slot, leaf, cnt.culls)) continue;
        best_t = t;
        best_id = (int32_t)id;
        bb1 = b1;
        bb2 = b2;
      }
      if (occluded) break;
    }
    if (sp == 0) break;
    cur = stack[--sp];
  }
  if (c.mode == ORC_OCCLUDED) {
    o.occ[i] = occluded ? 1 : 0;
  } else {
    o.id[i] = best_id;
    o.t[i] = best_t;
    o.b1[i] = bb1;
    o.b2[i] = bb2;
  }
}

}  // namespace

extern "C" {

int orc_mt_strict(const float o[3], const float d[3], const float v9[9], float out[3]) {
  Tri T = load_tri(v9, 0);
  return mt_strict({o[0], o[1], o[2]}, {d[0], d[1], d[2]}, T, out[0], out[1], out[2]);
}
int orc_mt_naive(const float o[3], const float d[3], const float v9[9], float out[3]) {
  Tri T = load_tri(v9, 0);
  return mt_naive({o[0], o[1], o[2]}, {d[0], d[1], d[2]}, T, out[0], out[1], out[2]);
}
int orc_mt_naive_p(const float o[3], const float d[3], const float v9[9], float out[3]) {
  Tri T = load_tri(v9, 0);
  return mt_naive_p({o[0], o[1], o[2]}, {d[0], d[1], d[2]}, T, out[0], out[1], out[2]);
}
int orc_mt_repair(const float o[3], const float d[3], const float v9[9], float out[3]) {
  Tri T = load_tri(v9, 0);
  return mt_repair({o[0], o[1], o[2]}, {d[0], d[1], d[2]}, T, out[0], out[1], out[2]);
}
int orc_mt_pbrt(const float o[3], const float d[3], const float v9[9], float out[3]) {
  Tri T = load_tri(v9, 0);
  return mt_pbrt({o[0], o[1], o[2]}, {d[0], d[1], d[2]}, T, out[0], out[1], out[2]);
}
int orc_cull_tri(const float node16[16], const float v9[9]) {
  return cull_tri(node_from16(node16), load_tri(v9, 0));
}
int orc_cull_node6(const float node16[16], const float bmin[3], const float bmax[3]) {
  return cull_node6(node_from16(node16), {bmin[0], bmin[1], bmin[2]}, {bmax[0], bmax[1], bmax[2]});
}
int orc_cull_5db(const float node16[16], const float bmin[3], const float bmax[3]) {
  return cull_5db(node_from16(node16), {bmin[0], bmin[1], bmin[2]}, {bmax[0], bmax[1], bmax[2]});
}
int orc_cull_ia(const float node16[16], const float bmin[3], const float bmax[3]) {
  return cull_ia(node_from16(node16), {bmin[0], bmin[1], bmin[2]}, {bmax[0], bmax[1], bmax[2]});
}

int orc_rh_levels(int32_t W, int32_t H, int32_t cx, int32_t cy, int32_t height, int32_t* dims) {
  if (W <= 0 || H <= 0 || cx <= 0 || cy <= 0 || height <= 0 || height > 60) return 0;
  int tmp[128];
  int L = rh_levels(W, H, cx, cy, height, tmp);
  for (int i = 0; i < 2 * L; ++i) dims[i] = tmp[i];
  return L;
}

int orc_rh_reference_well_formed(int32_t W, int32_t H, int32_t cx, int32_t cy, int32_t height) {
  if (W % cx || H % cy) return 0;
  int gw = W / cx, gh = H / cy;
  for (int i = 1; i < height; ++i) {
    if ((gw & 1) || (gh & 1)) return 0;
    gw /= 2;
    gh /= 2;
    if ((gw & 1) || (gh & 1)) break;
  }
  return 1;
}

int64_t orc_rh_build(const orc_rays* r, int32_t cx, int32_t cy, int32_t height, int32_t use_valid,
                     int32_t node, float* nodes) {
  RH h = rh_build(r, cx, cy, height, use_valid != 0, node);
  for (size_t i = 0; i < h.nodes.size(); ++i) node_to16(h.nodes[i], nodes + 16 * i);
  return (int64_t)h.nodes.size();
}

// Per-ray flags from the bookkeeping of a finished trace: bit 0 near tie (closest modes: the
// two nearest candidates within 1e-6 * t), bit 1 ulp-marginal (a candidate whose outcome rests
// on a marginal decision could change the result: closest modes, one no farther than the
// result; occlusion, one in the ray's interval while no candidate occludes it robustly).
void finish_flags(const orc_rays* r, int mode, const Out& o, const Margins& mg, uint8_t* flags) {
  const size_t n = (size_t)r->W * r->H;
  for (size_t i = 0; i < n; ++i) {
    uint8_t f = 0;
    if (mode == ORC_OCCLUDED) {
      if (mg.tmarg[i] == -kInf && !mg.robust[i]) f |= 2;
    } else {
      if (std::isfinite(mg.t1[i]) && std::isfinite(mg.t2[i]) &&
          mg.t2[i] - mg.t1[i] <= 1e-6f * mg.t1[i])
        f |= 1;
      const float tr = o.t[i];
      if (mg.tmarg[i] == -kInf || (mg.tmarg[i] < kInf && (!std::isfinite(tr) ||
                                                          mg.tmarg[i] <= tr * (1.f + 1e-6f))))
        f |= 2;
    }
    flags[i] = f;
  }
}

int orc_trace(const float* v9, uint32_t ntri, const orc_rays* r, const orc_params* p, int32_t mode,
              int32_t* id, float* t, float* b1, float* b2, uint8_t* occ, orc_stats* st,
              uint8_t* flags) {
  orc_stats local{0, 0, 0, 0};
  if (!st) st = &local;
  Out o{id, t, b1, b2, occ};
  init_out(r, mode, o);
  Margins mg;
  if (flags) mg.init((size_t)r->W * r->H);
  g_track = flags != nullptr;
  if (p->kind == 0) {
    trace_naive(v9, ntri, r, mode, o, st);
  } else if (p->kind == 1) {
    trace_rh(v9, ntri, r, p, mode, o, st, flags ? &mg : nullptr);
  } else {
    g_track = false;
    return 1;
  }
  g_track = false;
  if (flags) {
    if (p->kind == 0) std::memset(flags, 0, (size_t)r->W * r->H);
    else finish_flags(r, mode, o, mg, flags);
  }
  return 0;
}

int orc_trace_raybvh(const float* v9, uint32_t ntri, const orc_gpu_node* nodes, uint32_t nnodes,
                     const uint32_t* roots, uint32_t nroots, const uint32_t* prim_order,
                     const orc_rays* r, const orc_params* p, int32_t mode, int32_t* id, float* t,
                     float* b1, float* b2, uint8_t* occ, orc_stats* st, uint8_t* flags) {
  (void)ntri;
  (void)nnodes;
  orc_stats local{0, 0, 0, 0};
  if (!st) st = &local;
  Out o{id, t, b1, b2, occ};
  init_out(r, mode, o);
  Margins mg;
  if (flags) mg.init((size_t)r->W * r->H);
  g_track = flags != nullptr;
  bool use_valid = mode != ORC_CLOSEST_PRIMARY;
  RH h = rh_build(r, p->cx, p->cy, p->height, use_valid, p->node);
  LeafCtx lc{r, &h, mode, o, st, flags ? &mg : nullptr};
  BvhCtx c{&h, nodes, prim_order, v9, mode, p->node, use_valid, &lc, st};
  int top = h.L - 1;
  // The candidate set does not depend on visit order, but for the closest modes the per-ray
  // reduction must keep the smallest original id among equal t, which leaf_test's strict
  // compare does only for increasing ids: the reached (leaf, primitive) pairs are collected
  // and tested in increasing id order (occlusion tests them as they are reached).  Secondary
  // rays test only the first primitive of a BVH leaf (intersection6DB_BVH2.cl:88).
  struct Pair {
    int32_t tid;
    int k, j;
    bool shadow, cm;
  };
  std::vector<Pair> pairs;
  auto emit = [&](int k, int j, uint32_t ni, bool shadow, bool cm) {
    const orc_gpu_node& b = nodes[ni];
    // the 6DB kernel tests the first primitive only (intersection6DB_BVH2.cl:88), the 5DB
    // one every primitive (intersection5DB_BVH2.cl:84)
    uint32_t cnt = mode == ORC_CLOSEST_SECONDARY && p->node != 2 ? 1u : b.nPrimitives;
    for (uint32_t f = 0; f < cnt; ++f) {
      const uint32_t orig = prim_order[b.primOffset + f];
      if (mode == ORC_OCCLUDED) leaf_test(lc, k, j, load_tri(v9, orig), (int32_t)orig, shadow, cm);
      else pairs.push_back({(int32_t)orig, k, j, shadow, cm});
    }
  };
  for (uint32_t f = 0; f < nroots; ++f)
    for (int j = 0; j < h.gh[top]; ++j)
      for (int k = 0; k < h.gw[top]; ++k) bvh_walk(c, top, k, j, roots[f], false, false, emit);
  std::stable_sort(pairs.begin(), pairs.end(),
                   [](const Pair& a, const Pair& b) { return a.tid < b.tid; });
  for (const Pair& pr : pairs)
    leaf_test(lc, pr.k, pr.j, load_tri(v9, (uint32_t)pr.tid), pr.tid, pr.shadow, pr.cm);
  g_track = false;
  if (flags) finish_flags(r, mode, o, mg, flags);
  return 0;
}

int orc_search_trace(const float* v9, uint32_t ntri, const orc_gpu_node* nodes, uint32_t nnodes,
                     const uint32_t* roots, uint32_t nroots, const uint32_t* prim_order,
                     const orc_rays* r, const orc_params* p, int32_t mode, const uint8_t* ray_mask,
                     int32_t nthreads, int32_t* id, float* t, float* b1, float* b2, uint8_t* occ,
                     orc_search_stats* st) {
  if (p->kind != 1 && p->kind != 2) return 1;
  if (p->kind == 2 && (!nodes || !roots || !prim_order)) return 1;
  Out o{id, t, b1, b2, occ};
  init_out(r, mode, o);
  const SearchTree T = search_tree(v9, ntri);
  const bool use_valid = mode != ORC_CLOSEST_PRIMARY;
  const RH h = rh_build(r, p->cx, p->cy, p->height, use_valid, p->node);
  SearchCtx c;
  c.T = &T;
  c.h = &h;
  c.mode = mode;
  c.node_kind = p->node;
  c.kind = p->kind;
  c.repair = false;
  c.nodes = nodes;
  if (p->kind == 2) {
    c.tleaf.assign(ntri, 0xffffffffu);
    c.parent.assign(nnodes, 0xffffffffu);
    c.depth.assign(nnodes, 0);
    for (uint32_t f = 0; f < nroots; ++f) {
      std::vector<uint32_t> stk{roots[f]};
      c.parent[roots[f]] = roots[f];
      while (!stk.empty()) {
        const uint32_t n = stk.back();
        stk.pop_back();
        const orc_gpu_node& b = nodes[n];
        if (bvh_dead(p->node, b)) continue;
        if (b.nPrimitives) {
          for (uint32_t k = 0; k < b.nPrimitives; ++k)
            c.tleaf[prim_order[b.primOffset + k]] = n | (k > 0 ? 0x80000000u : 0u);
          continue;
        }
        for (uint32_t ch = b.primOffset; ch <= b.primOffset + 1; ++ch) {
          c.parent[ch] = n;
          c.depth[ch] = (uint8_t)(c.depth[n] + 1);
          stk.push_back(ch);
        }
      }
    }
  }
  const size_t n = (size_t)r->W * r->H;
  std::vector<SCount> per(n);
  if (ntri > 0)
    parallel_for(n, nthreads, [&](size_t i) {
      if (ray_mask && !ray_mask[i]) return;
      if (mode != ORC_CLOSEST_PRIMARY && !ray_valid(r, i)) return;
      s_trace_ray(c, r, i, o, per[i]);
    });
  if (st) {
    *st = orc_search_stats{0, 0, 0, 0, 0, 0};
    st->nodes = T.nodes.size();
    for (const SCount& q : per) {
      st->rays += q.rays;
      st->visits += q.visits;
      st->tests += q.tests;
      st->culls += q.culls;
      st->chains += q.chains;
    }
  }
  return 0;
}

void orc_shading(const float* v9, const float* uv6, uint32_t ntri, const orc_rays* r,
                 const int32_t* id, int32_t naive_rule, float* u, float* v, float* dpdu,
                 float* dpdv) {
  (void)ntri;
  size_t n = (size_t)r->W * r->H;
  static const float kDefUV[6] = {0.f, 0.f, 1.f, 0.f, 1.f, 1.f};
  for (size_t i = 0; i < n; ++i) {
    u[i] = v[i] = 0.f;
    for (int c = 0; c < 3; ++c) dpdu[3 * i + c] = dpdv[3 * i + c] = 0.f;
    if (id[i] < 0) continue;
    Tri T = load_tri(v9, (uint32_t)id[i]);
    const float* uv = uv6 ? uv6 + 6 * (size_t)id[i] : kDefUV;
    V3 d = ray_d(r, i), o = ray_o(r, i);
    // computeDpTuTv.cl:36-46
    V3 s1 = cross(d, T.e2);
    float div = dot(s1, T.e1);
    float inv = 1.0f / div;
    V3 dd = sub(o, T.v1);
    float b1 = dot(dd, s1) * inv;
    V3 s2 = cross(dd, T.e1);
    float b2 = dot(d, s2) * inv;
    // computeDpTuTv.cl:48-85
    float du1 = uv[0] - uv[4], du2 = uv[2] - uv[4];
    float dv1 = uv[1] - uv[5], dv2 = uv[3] - uv[5];
    V3 dp1 = sub(T.v1, T.v3), dp2 = sub(T.v2, T.v3);
    float det = du1 * dv2 - dv1 * du2;
    V3 pu, pv;
    // naive.cl:60-61 (inline copy in the naive kernels): fallback for |det| < EPS
    if (naive_rule ? (det > -kEps && det < kEps) : det == 0.f) {
      V3 tn = normalize(cross(T.e2, T.e1));
      if (std::fabs(tn.x) > std::fabs(tn.y)) {
        float il = 1.0f / std::sqrt(tn.x * tn.x + tn.z * tn.z);
        pu = {-tn.z * il, 0.f, tn.x * il};
      } else {
        float il = 1.0f / std::sqrt(tn.y * tn.y + tn.z * tn.z);
        pu = {0.f, tn.z * il, -tn.y * il};
      }
      pv = cross(tn, pu);
    } else {
      float id2 = 1.f / det;
      pu = {(dv2 * dp1.x - dv1 * dp2.x) * id2, (dv2 * dp1.y - dv1 * dp2.y) * id2,
            (dv2 * dp1.z - dv1 * dp2.z) * id2};
      pv = {(-du2 * dp1.x + du1 * dp2.x) * id2, (-du2 * dp1.y + du1 * dp2.y) * id2,
            (-du2 * dp1.z + du1 * dp2.z) * id2};
    }
    float b0 = 1 - b1 - b2;
    u[i] = (b0 * uv[0] + b1 * uv[2]) + b2 * uv[4];
    v[i] = (b0 * uv[1] + b1 * uv[3]) + b2 * uv[5];
    dpdu[3 * i] = pu.x; dpdu[3 * i + 1] = pu.y; dpdu[3 * i + 2] = pu.z;
    dpdv[3 * i] = pv.x; dpdv[3 * i + 1] = pv.y; dpdv[3 * i + 2] = pv.z;
  }
}

int64_t orc_bvh_order(const float* v9, uint32_t ntri, uint32_t max_prims, int32_t split,
                      uint32_t* order_out, uint64_t cap) {
  BvhBuilder b;
  b.maxPrims = std::min(255u, max_prims);
  b.split = split;
  b.run(v9, ntri);
  if (b.ordered.size() > cap) return -1;
  for (size_t i = 0; i < b.ordered.size(); ++i) order_out[i] = b.ordered[i];
  return (int64_t)b.ordered.size();
}

int64_t orc_bvh_build(const float* v9, uint32_t ntri, uint32_t max_prims, int32_t split,
                      int32_t gpu, uint32_t bvh_height, uint32_t* order_out,
                      orc_gpu_node* nodes_out, uint32_t* top_out) {
  if (max_prims > 4) return -1;  // the node layout of the fall-through is not restated here
  BvhBuilder b;
  b.maxPrims = std::min(255u, max_prims);
  b.split = split;
  b.run(v9, ntri);
  for (uint32_t i = 0; i < ntri; ++i) order_out[i] = b.ordered[i];
  if (!gpu) return (int64_t)b.nodes.size();
  if (b.nodes.empty()) {
    *top_out = 0;
    return 0;
  }
  return b.gpu_wide(bvh_height, nodes_out, top_out);
}

int64_t orc_raybvh_handoff(const float* v9, uint32_t ntri, uint32_t max_prims, int32_t split,
                           uint32_t bvh_height, orc_gpu_node* nodes_out, uint32_t* roots_out,
                           uint32_t* nroots_out, uint32_t* prim_order_out) {
  if (max_prims > 4) return -1;
  // global BVH order (rayBVH.cpp:144-146)
  BvhBuilder g;
  g.maxPrims = std::min(255u, max_prims);
  g.split = split;
  g.run(v9, ntri);
  std::vector<uint32_t> gorder = g.ordered;
  // 60k-triangle parts (rayBVH.cpp:288-297), per-part BVH (rayBVH.cpp:204-231)
  uint32_t parts = 1, partCount = ntri;
  const uint32_t kMax = 60000;
  if (ntri > kMax) {
    parts = (ntri + kMax - 1) / kMax;
    partCount = (ntri + parts - 1) / parts;
  }
  uint32_t node_base = 0, nroots = 0;
  for (uint32_t pi = 0; pi < parts; ++pi) {
    uint32_t lo = pi * partCount, hi = std::min(ntri, lo + partCount);
    std::vector<float> sub9((size_t)(hi - lo) * 9);
    for (uint32_t i = lo; i < hi; ++i)
      std::memcpy(&sub9[(size_t)(i - lo) * 9], v9 + 9 * (size_t)gorder[i], 9 * sizeof(float));
    BvhBuilder b;
    b.maxPrims = std::min(255u, max_prims);
    b.split = split;
    b.run(sub9.data(), hi - lo);
    std::vector<orc_gpu_node> tmp(2 * (size_t)(hi - lo) + 1);
    uint32_t top = 0;
    uint32_t nn = b.gpu_wide(bvh_height, tmp.data(), &top);
    for (uint32_t i = 0; i < nn; ++i) {
      orc_gpu_node g2 = tmp[i];
      if (g2.nPrimitives) g2.primOffset += lo;
      else g2.primOffset += node_base;
      nodes_out[node_base + i] = g2;
    }
    for (uint32_t i = 0; i < top; ++i) roots_out[nroots++] = node_base + i;
    for (uint32_t i = lo; i < hi; ++i) prim_order_out[i] = gorder[lo + b.ordered[i - lo]];
    node_base += nn;
  }
  *nroots_out = nroots;
  return node_base;
}

void* orc_cpu_bvh_create(const float* v9, uint32_t ntri, uint32_t max_prims) {
  BvhBuilder b;
  b.maxPrims = std::min(255u, max_prims);
  b.run(v9, ntri);
  CpuBvh* h = new CpuBvh();
  h->order = b.ordered;
  h->nodes.resize(b.nodes.size());
  uint32_t off = 0;
  if (!b.nodes.empty()) flatten(b, 0, h->nodes, &off);
  h->tris.resize(ntri);
  for (uint32_t i = 0; i < ntri; ++i) h->tris[i] = load_tri(v9, h->order[i]);
  return h;
}

void orc_cpu_bvh_destroy(void* h) { delete (CpuBvh*)h; }

void orc_cpu_bvh_closest(void* hp, const orc_rays* r, int32_t nthreads, int32_t* id, float* t,
                         float* b1, float* b2) {
  const CpuBvh& b = *(const CpuBvh*)hp;
  size_t n = (size_t)r->W * r->H;
  parallel_for(n, nthreads, [&](size_t i) {
    id[i] = -1;
    t[i] = kInf;
    b1[i] = b2[i] = 0.f;
    if (!ray_valid(r, i)) return;
    cpu_closest(b, ray_o(r, i), ray_d(r, i), r->tmin[i], r->tmax[i], id[i], t[i], b1[i], b2[i]);
  });
}

void orc_cpu_bvh_occluded(void* hp, const orc_rays* r, int32_t nthreads, uint8_t* occ) {
  const CpuBvh& b = *(const CpuBvh*)hp;
  size_t n = (size_t)r->W * r->H;
  parallel_for(n, nthreads, [&](size_t i) {
    occ[i] = 0;
    if (!ray_valid(r, i)) return;
    occ[i] = cpu_occluded(b, ray_o(r, i), ray_d(r, i), r->tmin[i], r->tmax[i]) ? 1 : 0;
  });
}

void orc_exact_closest(const float* v9, uint32_t ntri, const orc_rays* r, int32_t* id, float* t) {
  size_t n = (size_t)r->W * r->H;
  std::vector<Tri> tris(ntri);
  for (uint32_t i = 0; i < ntri; ++i) tris[i] = load_tri(v9, i);
  for (size_t i = 0; i < n; ++i) {
    id[i] = -1;
    t[i] = kInf;
    if (!ray_valid(r, i)) continue;
    float maxt = r->tmax[i];
    for (uint32_t k = 0; k < ntri; ++k) {
      float tt, b1, b2;
      if (!mt_pbrt(ray_o(r, i), ray_d(r, i), tris[k], tt, b1, b2)) continue;
      if (tt < r->tmin[i] || tt > maxt) continue;
      if (id[i] >= 0 && !(tt < t[i])) continue;
      maxt = tt;
      t[i] = tt;
      id[i] = (int32_t)k;
    }
  }
}

}  // extern "C"
