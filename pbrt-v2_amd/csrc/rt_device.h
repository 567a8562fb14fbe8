// rt_device.h -- FP32 device arithmetic of the ray-space accelerators (gfx950).
//
// Every function fixes its operation order so results are reproducible bit for bit:
// the library is compiled with -ffp-contract=off (no FMA contraction) and correctly rounded
// FP32 division and square root; min/max are written as selects.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtd {

// EPS of intersectAllLeaves.cl:8 and naive.cl:2 in FP32 (the kernels do not enable
// cl_khr_fp64, so their double literals are single precision).
constexpr float kEps = 0.002f;
constexpr float kNaiveLo = -1e-3f - kEps;            // naive.cl:38,43
constexpr float kNaiveHi = (1.0f + 1e-3f) + kEps;    // naive.cl:38,43
constexpr float kNaivePLo = -1e-3f;                  // naive.cl:219,224
constexpr float kNaivePHi = 1.0f + 1e-3f;            // naive.cl:219,224
constexpr float kRepHi1 = 1.0f + kEps;               // intersectAllLeaves.cl:217
constexpr float kRepHi2 = 1.0f - kEps;               // intersectAllLeaves.cl:222

struct V3 {
  float x, y, z;
};

__device__ __forceinline__ float fmn(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float fmx(float a, float b) { return b > a ? b : a; }
__device__ __forceinline__ V3 vmin(V3 a, V3 b) { return {fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)}; }
__device__ __forceinline__ V3 vmax(V3 a, V3 b) { return {fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// normalize(v) = v * (1/sqrt(v.v)); a zero vector is returned unchanged.
__device__ __forceinline__ V3 normalize(V3 v) {
  float l2 = dot(v, v);
  if (l2 == 0.f) return v;
  float inv = 1.0f / sqrtf(l2);
  return {v.x * inv, v.y * inv, v.z * inv};
}

// Triangle as the kernels use it: vertices plus the two edges from v1.
struct Tri {
  V3 v1, v2, v3, e1, e2;
};
__device__ __forceinline__ Tri make_tri(V3 a, V3 b, V3 c) {
  Tri t;
  t.v1 = a;
  t.v2 = b;
  t.v3 = c;
  t.e1 = sub(b, a);
  t.e2 = sub(c, a);
  return t;
}

// Strict Moller-Trumbore of intersectAllLeaves.cl:40-55: det != 0, b1 in (0,1), b2 > 0,
// b1 + b2 < 1.  Returns false when the barycentric test fails; t/b1/b2 are then undefined.
// Before the division, b1 = fl(dot(ov, s) * fl(1/det)) is rejected when it is certainly
// outside (0, 1): a zero or opposite-signed numerator gives b1 <= 0, and
// |numerator| >= fl(|det| * (1 + 2^-20)) gives b1 > 1 whatever the two roundings (|det| in
// the normal range, so both roundings are relative).  Only
// tests the exact sequence would reject are cut short, so the outcome is unchanged.
__device__ __forceinline__ bool mt_strict(V3 o, V3 d, const Tri& T, float& t, float& b1, float& b2) {
  V3 s = cross(d, T.e2);
  float det = dot(s, T.e1);
  if (det == 0.0f) return false;
  V3 ov = sub(o, T.v1);
  const float n1 = dot(ov, s);
  if (n1 == 0.0f || ((n1 < 0.0f) != (det < 0.0f))) return false;
  if (fabsf(det) >= 0x1p-100f && fabsf(n1) >= fabsf(det) * (1.0f + 0x1p-20f)) return false;
  float inv = 1.0f / det;
  b1 = n1 * inv;
  if (!(b1 > 0.f && b1 < 1.f)) return false;
  V3 s2 = cross(ov, T.e1);
  b2 = dot(d, s2) * inv;
  if (!(b2 > 0.f && (b1 + b2) < 1.f)) return false;
  t = dot(T.e2, s2) * inv;
  return true;
}

// Shrunk barycentric region of the repair pass (intersectAllLeaves.cl:216-222).  Only called
// on values that already passed mt_strict, whose region contains this one.
__device__ __forceinline__ bool repair_region(float b1, float b2) {
  return !(b1 < kEps || b1 > kRepHi1) && !(b2 < kEps || (b1 + b2) > kRepHi2);
}

// naive.cl:30-46 (TriRay closest).
__device__ __forceinline__ bool mt_naive(V3 o, V3 d, const Tri& T, float& t, float& b1, float& b2) {
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

// naive.cl:211-227 (IntersectionP).
__device__ __forceinline__ bool mt_naive_p(V3 o, V3 d, const Tri& T, float& t) {
  V3 s1 = cross(d, T.e2);
  float div = dot(s1, T.e1);
  if (div == 0.0f) return false;
  float inv = 1.0f / div;
  V3 dd = sub(o, T.v1);
  float b1 = dot(dd, s1) * inv;
  if (b1 < kNaivePLo || b1 > kNaivePHi) return false;
  V3 s2 = cross(dd, T.e1);
  float b2 = dot(d, s2) * inv;
  if (b2 < kNaivePLo || (b1 + b2) > kNaivePHi) return false;
  t = dot(T.e2, s2) * inv;
  return true;
}

// Barycentrics of a known winner (computeDpTuTv.cl:36-46, no range tests).
__device__ __forceinline__ void bary(V3 o, V3 d, const Tri& T, float& b1, float& b2) {
  V3 s1 = cross(d, T.e2);
  float div = dot(s1, T.e1);
  float inv = 1.0f / div;
  V3 dd = sub(o, T.v1);
  b1 = dot(dd, s1) * inv;
  V3 s2 = cross(dd, T.e1);
  b2 = dot(d, s2) * inv;
}

// Ray-hierarchy node: the component-wise box of directions and origins of its rays.
struct Node {
  V3 dmin, dmax, omin, omax;
};

// intersectsNode6DB.cl:53-63 / 88-98 with the uninitialised `ret` taken as false.
__device__ __forceinline__ bool box_logic(V3 tmin, V3 tmax, const Node& n) {
  V3 f0 = vmax(tmin, n.dmin);
  V3 f1 = vmin(tmax, n.dmax);
  bool ret = false;
  if (tmin.x < 0.f && tmax.x > 0.f) ret = (f0.x < f1.x + kEps);
  if (tmin.y < 0.f && tmax.y > 0.f) ret |= (f0.y < f1.y + kEps);
  if (tmin.z < 0.f && tmax.z > 0.f) ret |= (f0.z < f1.z + kEps);
  return ret || (f0.x < f1.x + kEps && f0.y < f1.y + kEps && f0.z < f1.z + kEps);
}

// Origin-box centre used by intersectsTri (intersectsNode6DB.cl:73).
__device__ __forceinline__ V3 tri_center(const Node& n) {
  return {(n.omin.x + n.omax.x) / 2, (n.omin.y + n.omax.y) / 2, (n.omin.z + n.omax.z) / 2};
}
// Direction box of a triangle seen from oc (intersectsNode6DB.cl:76-86).
__device__ __forceinline__ void tri_dirbox(V3 oc, const Tri& T, V3& tmin, V3& tmax) {
  V3 a = normalize(sub(T.v1, oc));
  tmin = a;
  tmax = a;
  V3 b = normalize(sub(T.v2, oc));
  tmin = vmin(tmin, b);
  tmax = vmax(tmax, b);
  V3 c = normalize(sub(T.v3, oc));
  tmin = vmin(tmin, c);
  tmax = vmax(tmax, c);
}
// intersectsTri, intersectsNode6DB.cl:66-99.
__device__ __forceinline__ bool cull_tri(const Node& n, const Tri& T) {
  V3 tmin, tmax;
  tri_dirbox(tri_center(n), T, tmin, tmax);
  return box_logic(tmin, tmax, n);
}

// intersectsNode, intersectsNode6DB.cl:6-64 (box Minkowski-grown by the origin half-extent).
__device__ __forceinline__ bool cull_node6(const Node& n, V3 bmin, V3 bmax) {
  V3 h = {(n.omax.x - n.omin.x) / 2, (n.omax.y - n.omin.y) / 2, (n.omax.z - n.omin.z) / 2};
  bmin = sub(bmin, h);
  bmax = add(bmax, h);
  V3 oc = add(h, n.omin);
  V3 r = normalize(sub(V3{bmin.x, bmin.y, bmin.z}, oc));
  V3 tmin = r, tmax = r;
#define RT_CORNER(X, Y, Z)              \
  r = normalize(sub(V3{X, Y, Z}, oc)); \
  tmin = vmin(tmin, r);                 \
  tmax = vmax(tmax, r);
  RT_CORNER(bmin.x, bmin.y, bmax.z)
  RT_CORNER(bmin.x, bmax.y, bmin.z)
  RT_CORNER(bmin.x, bmax.y, bmax.z)
  RT_CORNER(bmax.x, bmin.y, bmin.z)
  RT_CORNER(bmax.x, bmax.y, bmin.z)
  RT_CORNER(bmax.x, bmin.y, bmax.z)
  RT_CORNER(bmax.x, bmax.y, bmax.z)
#undef RT_CORNER
  return box_logic(tmin, tmax, n);
}

// intersectsNodeIA.cl:6-65 (interval slab test).
__device__ __forceinline__ void ia_axis(float lo, float hi, float dlo, float dhi, float& smin,
                                        float& smax) {
  float a = lo * dlo, b = hi * dlo, c = hi * dhi, d = lo * dhi;
  smin = fmn(fmn(fmn(a, b), c), d);
  smax = fmx(fmx(fmx(a, b), c), d);
}
__device__ __forceinline__ bool cull_ia(const Node& n, V3 bmin, V3 bmax) {
  V3 lo = sub(bmin, n.omax);
  V3 hi = sub(bmax, n.omin);
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
  return sx < sy + kEps;
}

// intersectsNode5DB.cl:1-83, node "box_uv", whose dmin/dmax hold the slope and dz intervals of
// its rays: dmin = (min d.y/d.x, min d.z, 0), dmax = (max d.y/d.x, max d.z, 0).  As C parses
// the kernel, line 79 tests range & 1 (0x1100 == 0x1100 is 1): some corner direction has y not
// below zero; on line 81 the first conjunct is a comparison (0 or 1) plus EPS, never zero, so
// the slope intervals never decide.  The node passes when a normalised corner direction of the
// origin-grown box has y >= 0, else when the corners' z range meets [dz min, dz max] within
// EPS.  Corners in the kernel's order.
__device__ __forceinline__ bool cull_5db(const Node& n, V3 bmin, V3 bmax) {
  V3 h = {(n.omax.x - n.omin.x) / 2, (n.omax.y - n.omin.y) / 2, (n.omax.z - n.omin.z) / 2};
  bmin = sub(bmin, h);
  bmax = add(bmax, h);
  V3 oc = add(h, n.omin);
  V3 r = normalize(sub(V3{bmin.x, bmin.y, bmin.z}, oc));
  bool up = !(r.y < 0.f);
  float zlo = r.z, zhi = r.z;
#define RT_UV_CORNER(X, Y, Z)           \
  r = normalize(sub(V3{X, Y, Z}, oc)); \
  up |= !(r.y < 0.f);                   \
  zlo = fmn(zlo, r.z);                  \
  zhi = fmx(zhi, r.z);
  RT_UV_CORNER(bmax.x, bmax.y, bmax.z)
  RT_UV_CORNER(bmin.x, bmax.y, bmin.z)
  RT_UV_CORNER(bmin.x, bmax.y, bmax.z)
  RT_UV_CORNER(bmin.x, bmin.y, bmax.z)
  RT_UV_CORNER(bmax.x, bmin.y, bmin.z)
  RT_UV_CORNER(bmax.x, bmax.y, bmin.z)
  RT_UV_CORNER(bmax.x, bmin.y, bmax.z)
#undef RT_UV_CORNER
  if (up) return true;
  return fmx(zlo, n.dmin.y) < fmn(zhi, n.dmax.y) + kEps;
}

// Exact conservative prefilters.  The normalised direction of every corner (or vertex) has
// its squared length between the sums of the per-axis minimum and maximum squares, and the
// exact test's scale factor (correctly rounded sqrt and division, both monotone) lies between
// the outward-widened hardware reciprocal square roots of those sums, so the normalised
// components lie between the bounds computed here from the same FP32 values.  box_logic is
// monotone under growing the direction box, so a reject here implies a reject of the exact
// test.  A zero-length corner (normalize returns it unchanged) is left to the exact test.
__device__ __forceinline__ void outer_axis(float a, float b, float inv_lo, float inv_hi, float& lo,
                                           float& hi) {
  const float la = a < 0.f ? a * inv_hi : a * inv_lo, ha = a < 0.f ? a * inv_lo : a * inv_hi;
  const float lb = b < 0.f ? b * inv_hi : b * inv_lo, hb = b < 0.f ? b * inv_lo : b * inv_hi;
  lo = fmn(la, lb);
  hi = fmx(ha, hb);
}
// The exact test scales by fl(1/fl(sqrt(l2))), within 2^-23 (relative) of 1/sqrt(l2) and
// monotone in l2.  The hardware reciprocal square root is within 1 ulp (2^-23) of
// 1/sqrt(l2); widened by 2^-16 outward it brackets the exact factor over [l2min, l2max]
// with a wide margin.  Tiny squared lengths (where the hardware may flush) pass.
__device__ __forceinline__ bool outer_box_logic(V3 lo, V3 hi, V3 mn, V3 mx, const Node& n) {
  const float l2min = (mn.x + mn.y) + mn.z;
  if (!(l2min >= 0x1p-100f)) return true;
  const float l2max = (mx.x + mx.y) + mx.z;
  const float inv_hi = __builtin_amdgcn_rsqf(l2min) * (1.0f + 0x1p-16f);
  const float inv_lo = __builtin_amdgcn_rsqf(l2max) * (1.0f - 0x1p-16f);
  V3 tmin, tmax;
  outer_axis(lo.x, hi.x, inv_lo, inv_hi, tmin.x, tmax.x);
  outer_axis(lo.y, hi.y, inv_lo, inv_hi, tmin.y, tmax.y);
  outer_axis(lo.z, hi.z, inv_lo, inv_hi, tmin.z, tmax.z);
  return box_logic(tmin, tmax, n);
}
// May return true when cull_node6 is false; never false when it is true.
__device__ __forceinline__ bool prefilter_node6(const Node& n, V3 bmin, V3 bmax) {
  V3 h = {(n.omax.x - n.omin.x) / 2, (n.omax.y - n.omin.y) / 2, (n.omax.z - n.omin.z) / 2};
  const V3 b0 = sub(bmin, h), b1 = add(bmax, h);
  const V3 oc = add(h, n.omin);
  const V3 lo = sub(b0, oc), hi = sub(b1, oc);
  const V3 ql = {lo.x * lo.x, lo.y * lo.y, lo.z * lo.z}, qh = {hi.x * hi.x, hi.y * hi.y, hi.z * hi.z};
  return outer_box_logic(lo, hi, vmin(ql, qh), vmax(ql, qh), n);
}
// May return true when cull_tri is false; never false when it is true.
__device__ __forceinline__ bool prefilter_tri(const Node& n, const Tri& T) {
  const V3 oc = tri_center(n);
  const V3 a = sub(T.v1, oc), b = sub(T.v2, oc), c = sub(T.v3, oc);
  const V3 qa = {a.x * a.x, a.y * a.y, a.z * a.z}, qb = {b.x * b.x, b.y * b.y, b.z * b.z},
           qc = {c.x * c.x, c.y * c.y, c.z * c.z};
  const V3 lo = vmin(vmin(a, b), c), hi = vmax(vmax(a, b), c);
  return outer_box_logic(lo, hi, vmin(vmin(qa, qb), qc), vmax(vmax(qa, qb), qc), n);
}

// Cluster tests: conservative for every triangle whose vertices lie in [cmin, cmax].  A
// vertex or grown-box corner component lies between the endpoint differences (FP32
// subtraction is monotone), its square at least that of the endpoint nearest to zero (0 when
// the interval straddles zero).
__device__ __forceinline__ void interval_squares(V3 lo, V3 hi, V3& mn, V3& mx) {
  const V3 ql = {lo.x * lo.x, lo.y * lo.y, lo.z * lo.z}, qh = {hi.x * hi.x, hi.y * hi.y, hi.z * hi.z};
  mn = vmin(ql, qh);
  mx = vmax(ql, qh);
  if (lo.x <= 0.f && hi.x >= 0.f) mn.x = 0.f;
  if (lo.y <= 0.f && hi.y >= 0.f) mn.y = 0.f;
  if (lo.z <= 0.f && hi.z >= 0.f) mn.z = 0.f;
}
__device__ __forceinline__ bool cluster_tri(const Node& n, V3 cmin, V3 cmax) {
  const V3 oc = tri_center(n);
  const V3 lo = sub(cmin, oc), hi = sub(cmax, oc);
  V3 mn, mx;
  interval_squares(lo, hi, mn, mx);
  return outer_box_logic(lo, hi, mn, mx, n);
}
__device__ __forceinline__ bool cluster_node6(const Node& n, V3 cmin, V3 cmax) {
  const V3 h = {(n.omax.x - n.omin.x) / 2, (n.omax.y - n.omin.y) / 2, (n.omax.z - n.omin.z) / 2};
  const V3 b0 = sub(cmin, h), b1 = add(cmax, h);
  const V3 oc = add(h, n.omin);
  const V3 lo = sub(b0, oc), hi = sub(b1, oc);
  V3 mn, mx;
  interval_squares(lo, hi, mn, mx);
  return outer_box_logic(lo, hi, mn, mx, n);
}

// Bound of cull_5db for every box inside [cmin, cmax] (one triangle's bounds, a cluster of
// triangles or of BVH boxes): never false when cull_5db is true for such a box.  A grown
// corner's components lie between the endpoint differences lo / hi (FP32 subtraction is
// monotone).  While hi.y < 0 every corner's y is negative; its normalised y stays negative
// unless the product underflows, which the test against 2^-100 excludes.  The normalised z then
// lies between outer_axis' bounds, and the z test is monotone in them.
__device__ __forceinline__ bool bound_5db(const Node& n, V3 cmin, V3 cmax) {
  const V3 h = {(n.omax.x - n.omin.x) / 2, (n.omax.y - n.omin.y) / 2, (n.omax.z - n.omin.z) / 2};
  const V3 b0 = sub(cmin, h), b1 = add(cmax, h);
  const V3 oc = add(h, n.omin);
  const V3 lo = sub(b0, oc), hi = sub(b1, oc);
  if (!(hi.y < 0.f)) return true;
  V3 mn, mx;
  interval_squares(lo, hi, mn, mx);
  const float l2min = (mn.x + mn.y) + mn.z;
  const float l2max = (mx.x + mx.y) + mx.z;
  if (!(l2min >= 0x1p-100f) || !(l2max <= 0x1p+100f)) return true;
  const float inv_hi = __builtin_amdgcn_rsqf(l2min) * (1.0f + 0x1p-16f);
  const float inv_lo = __builtin_amdgcn_rsqf(l2max) * (1.0f - 0x1p-16f);
  if (!(hi.y * inv_lo < -0x1p-100f)) return true;
  float zlo, zhi;
  outer_axis(lo.z, hi.z, inv_lo, inv_hi, zlo, zhi);
  return fmx(zlo, n.dmin.y) < fmn(zhi, n.dmax.y) + kEps;
}

// Order-preserving 32-bit key of a float (+0 and -0 share the key of +0), and the 64-bit
// nearest-hit key (t, id): the smallest key is the nearest hit, ties -> smallest id.
__device__ __forceinline__ uint32_t fkey(float t) {
  uint32_t b = __float_as_uint(t == 0.0f ? 0.0f : t);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fkey_inv(uint32_t k) {
  uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(b);
}
constexpr uint64_t kMissKey = ~0ull;
__device__ __forceinline__ uint64_t hit_key(float t, int32_t id) {
  return ((uint64_t)fkey(t) << 32) | (uint32_t)id;
}

}  // namespace rtd
