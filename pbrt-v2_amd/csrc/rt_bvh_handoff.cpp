// rt_bvh_handoff.cpp -- the CPU-built BVH that ray-bvh hands to the device.
//
// Host restatement of what RayBVH does before its first trace (rayBVH.cpp:144-146,204-276):
//   1. a BVHAccel over all triangles fixes the BVH primitive order;
//   2. that order is cut into parts of at most 60,000 triangles (MAX_VERTICES,
//      rayBVH.cpp:288-297);
//   3. every part gets its own BVHAccel (SAH, 12 buckets, maxBVHPrim primitives per leaf,
//      bvh.cpp:290-434) flattened breadth first by GPUbufferWide (bvh.cpp:154-206), which drops
//      all but the lowest BVHheight levels and keeps the remaining frontier as traversal roots.
// The parts are concatenated: node indices and primitive positions are rebased per part, and
// prim_order maps every primitive position back to the original triangle index (the remap of
// rayBVH.cpp:795-799).  Sorting uses std::partition / std::nth_element as bvh.cpp does, so
// the node layout follows the same libstdc++ algorithms.  The split methods are BVHAccel's
// (bvh.cpp:214-221,332-427): "sah" (12 buckets), "middle" (centroid midpoint) and "equal"
// (equal counts).
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#include "../../include/pbrt_amd_rt.h"

namespace {

struct P3 {
  float v[3];
};
struct Box {
  P3 lo{{std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
         std::numeric_limits<float>::infinity()}};
  P3 hi{{-std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
         -std::numeric_limits<float>::infinity()}};
  void grow(const P3& p) {
    for (int a = 0; a < 3; ++a) {
      lo.v[a] = std::min(lo.v[a], p.v[a]);
      hi.v[a] = std::max(hi.v[a], p.v[a]);
    }
  }
  void grow(const Box& b) {
    for (int a = 0; a < 3; ++a) {
      lo.v[a] = std::min(lo.v[a], b.lo.v[a]);
      hi.v[a] = std::max(hi.v[a], b.hi.v[a]);
    }
  }
  float area() const {
    float dx = hi.v[0] - lo.v[0], dy = hi.v[1] - lo.v[1], dz = hi.v[2] - lo.v[2];
    return 2.f * (dx * dy + dx * dz + dy * dz);
  }
  int widest() const {
    float dx = hi.v[0] - lo.v[0], dy = hi.v[1] - lo.v[1], dz = hi.v[2] - lo.v[2];
    if (dx > dy && dx > dz) return 0;
    return dy > dz ? 1 : 2;
  }
};

struct Item {
  int prim;
  P3 c;
  Box b;
};

struct BNode {
  Box box;
  int kid[2] = {-1, -1};
  uint32_t first = 0, count = 0;
};

class Builder {
 public:
  Builder(uint32_t max_prims, int split_method)
      : max_prims_(std::min(255u, max_prims)), method_(split_method) {}

  // Builds over `ntri` triangles (9 floats each); fills order (position -> input index).
  void build(const float* v9, uint32_t ntri) {
    std::vector<Item> items(ntri);
    for (uint32_t i = 0; i < ntri; ++i) {
      const float* p = v9 + 9 * (size_t)i;
      Box b;
      // BBox(p1, p2) then Union with p3 (Triangle::WorldBound)
      for (int a = 0; a < 3; ++a) {
        b.lo.v[a] = std::min(p[a], p[3 + a]);
        b.hi.v[a] = std::max(p[a], p[3 + a]);
      }
      b.grow(P3{{p[6], p[7], p[8]}});
      items[i].prim = (int)i;
      items[i].b = b;
      for (int a = 0; a < 3; ++a) items[i].c.v[a] = .5f * b.lo.v[a] + .5f * b.hi.v[a];
    }
    nodes_.clear();
    order_.clear();
    nodes_.reserve(2 * (size_t)ntri);
    if (ntri) split(items, 0, ntri);
  }

  const std::vector<uint32_t>& order() const { return order_; }

  // Breadth-first GPUNode layout keeping the lowest `height` levels (GPUbufferWide).
  uint32_t wide(uint32_t height, std::vector<rt_gpu_node>& out, uint32_t* top) const {
    out.clear();
    if (nodes_.empty()) {
      *top = 0;
      return 0;
    }
    struct Q {
      int node;
      uint32_t level;
      int parent;  // output index whose child A this entry is, or -1
      uint32_t kidA;
    };
    std::vector<Q> q;
    q.reserve(nodes_.size());
    q.push_back({0, 0, -1, 0});
    uint32_t deepest = 0;
    for (size_t i = 0; i < q.size(); ++i) {
      const BNode& n = nodes_[q[i].node];
      if (n.count) continue;
      q[i].kidA = (uint32_t)q.size();
      q.push_back({n.kid[0], q[i].level + 1, -1, 0});
      q.push_back({n.kid[1], q[i].level + 1, -1, 0});
      deepest = q[i].level + 1;
    }
    const int skip = std::max((int)deepest - (int)height, 0);
    uint32_t topn = 0;
    for (size_t i = 0; i < q.size(); ++i) {
      const BNode& n = nodes_[q[i].node];
      if ((int)q[i].level < skip && !n.count) continue;
      if ((int)q[i].level <= skip) ++topn;
      const uint32_t me = (uint32_t)out.size();
      if (q[i].parent >= 0) out[q[i].parent].prim_offset = me;
      rt_gpu_node g{};
      g.ax = n.box.lo.v[0];
      g.ay = n.box.lo.v[1];
      g.az = n.box.lo.v[2];
      g.bx = n.box.hi.v[0];
      g.by = n.box.hi.v[1];
      g.bz = n.box.hi.v[2];
      g.n_primitives = n.count;
      g.prim_offset = n.count ? n.first : 0;
      out.push_back(g);
      if (!n.count) q[q[i].kidA].parent = (int)me;
    }
    *top = topn;
    return (uint32_t)out.size();
  }

 private:
  static int bucket(float c, float lo, float hi) {
    int b = (int)(12 * ((c - lo) / (hi - lo)));
    return b == 12 ? 11 : b;
  }

  int split(std::vector<Item>& it, uint32_t s, uint32_t e) {
    const int me = (int)nodes_.size();
    nodes_.emplace_back();
    Box box;
    for (uint32_t i = s; i < e; ++i) box.grow(it[i].b);
    const uint32_t n = e - s;
    auto leaf = [&]() {
      nodes_[me].first = (uint32_t)order_.size();
      nodes_[me].count = n;
      nodes_[me].box = box;
      for (uint32_t i = s; i < e; ++i) order_.push_back((uint32_t)it[i].prim);
      return me;
    };
    if (n == 1) return leaf();
    Box cb;
    for (uint32_t i = s; i < e; ++i) cb.grow(it[i].c);
    const int ax = cb.widest();
    const float lo = cb.lo.v[ax], hi = cb.hi.v[ax];
    if (hi == lo) return leaf();
    uint32_t mid = (s + e) / 2;
    auto equal_counts = [&]() {
      mid = (s + e) / 2;
      std::nth_element(&it[s], &it[mid], &it[e - 1] + 1,
                       [ax](const Item& a, const Item& b) { return a.c.v[ax] < b.c.v[ax]; });
    };
    if (method_ == RT_SPLIT_MIDDLE) {
      // bvh.cpp:333-340.  When the midpoint rounds onto an end of the centroid range every
      // primitive lands on one side and the reference recurses without end; equal counts here.
      const float pmid = .5f * (lo + hi);
      Item* m = std::partition(&it[s], &it[e - 1] + 1,
                               [&](const Item& p) { return p.c.v[ax] < pmid; });
      mid = (uint32_t)(m - &it[0]);
      if (mid == s || mid == e) equal_counts();
    } else if (method_ == RT_SPLIT_EQUAL || n <= 4) {
      equal_counts();  // bvh.cpp:342-347, and SAH's small nodes (:351-356)
    } else {
      int cnt[12] = {0};
      Box bb[12];
      for (uint32_t i = s; i < e; ++i) {
        const int b = bucket(it[i].c.v[ax], lo, hi);
        ++cnt[b];
        bb[b].grow(it[i].b);
      }
      const float inv_area = box.area();
      float best = 0.f;
      int best_split = 0;
      for (int k = 0; k < 11; ++k) {
        Box left, right;
        int cl = 0, cr = 0;
        for (int j = 0; j <= k; ++j) {
          left.grow(bb[j]);
          cl += cnt[j];
        }
        for (int j = k + 1; j < 12; ++j) {
          right.grow(bb[j]);
          cr += cnt[j];
        }
        const float cost = .125f + (cl * left.area() + cr * right.area()) / inv_area;
        if (k == 0 || cost < best) {
          best = cost;
          best_split = k;
        }
      }
      if (n <= max_prims_ && !(best < (float)n)) {
        // bvh.cpp:417-428: the reference records these primitives as a leaf but has no return
        // there, so the node is rebuilt as an interior node over [s, (s+e)/2) and
        // [(s+e)/2, e), which appends the same primitives again.  Every later copy has a
        // larger index than the first one and never wins a tie against it: the first
        // occurrences, in this range's current order, are the order that counts.  Only
        // reachable with more than 4 primitives per leaf (the sort order, not the handoff).
        return leaf();
      }
      Item* m = std::partition(&it[s], &it[e - 1] + 1, [&](const Item& p) {
        return bucket(p.c.v[ax], lo, hi) <= best_split;
      });
      mid = (uint32_t)(m - &it[0]);
    }
    const int a = split(it, s, mid);
    const int b = split(it, mid, e);
    nodes_[me].kid[0] = a;
    nodes_[me].kid[1] = b;
    Box u = nodes_[a].box;
    u.grow(nodes_[b].box);
    nodes_[me].box = u;
    nodes_[me].count = 0;
    return me;
  }

  uint32_t max_prims_;
  int method_;
  std::vector<BNode> nodes_;
  std::vector<uint32_t> order_;
};

}  // namespace

extern "C" int rt_raybvh_build_handoff(const float* v9, uint32_t ntri, uint32_t max_prims,
                                       int32_t split_method, uint32_t bvh_height,
                                       rt_gpu_node** nodes_out, uint32_t* nnodes_out,
                                       uint32_t** roots_out, uint32_t* nroots_out,
                                       uint32_t** prim_order_out) {
  if (!nodes_out || !nnodes_out || !roots_out || !nroots_out || !prim_order_out)
    return RT_ERR_INVALID;
  if (ntri && !v9) return RT_ERR_INVALID;
  if (split_method < RT_SPLIT_SAH || split_method > RT_SPLIT_EQUAL) return RT_ERR_INVALID;
  if (max_prims == 0 || max_prims > 4) return RT_ERR_UNSUPPORTED;
  *nodes_out = nullptr;
  *roots_out = *prim_order_out = nullptr;
  *nnodes_out = *nroots_out = 0;
  // 1. global BVH order
  Builder global(max_prims, split_method);
  global.build(v9, ntri);
  const std::vector<uint32_t> gorder = global.order();
  // 2. parts of at most 60,000 triangles
  uint32_t parts = 1, per = ntri;
  if (ntri > 60000u) {
    parts = (ntri + 60000u - 1) / 60000u;
    per = (ntri + parts - 1) / parts;
  }
  std::vector<rt_gpu_node> all;
  std::vector<uint32_t> roots, order(ntri);
  std::vector<float> sub;
  std::vector<rt_gpu_node> tmp;
  for (uint32_t p = 0; p < parts; ++p) {
    const uint32_t lo = p * per, hi = std::min(ntri, lo + per);
    sub.resize((size_t)(hi - lo) * 9);
    for (uint32_t i = lo; i < hi; ++i)
      std::memcpy(&sub[(size_t)(i - lo) * 9], v9 + 9 * (size_t)gorder[i], 9 * sizeof(float));
    // 3. per-part BVH and breadth-first layout
    Builder b(max_prims, split_method);
    b.build(sub.data(), hi - lo);
    uint32_t top = 0;
    const uint32_t nn = b.wide(bvh_height, tmp, &top);
    const uint32_t base = (uint32_t)all.size();
    for (uint32_t i = 0; i < nn; ++i) {
      rt_gpu_node g = tmp[i];
      g.prim_offset += g.n_primitives ? lo : base;
      all.push_back(g);
    }
    for (uint32_t i = 0; i < top; ++i) roots.push_back(base + i);
    for (uint32_t i = lo; i < hi; ++i) order[i] = gorder[lo + b.order()[i - lo]];
  }
  rt_gpu_node* n = (rt_gpu_node*)std::malloc(sizeof(rt_gpu_node) * std::max<size_t>(all.size(), 1));
  uint32_t* r = (uint32_t*)std::malloc(sizeof(uint32_t) * std::max<size_t>(roots.size(), 1));
  uint32_t* o = (uint32_t*)std::malloc(sizeof(uint32_t) * std::max<size_t>(order.size(), 1));
  if (!n || !r || !o) {
    std::free(n);
    std::free(r);
    std::free(o);
    return RT_ERR_OOM;
  }
  if (!all.empty()) std::memcpy(n, all.data(), sizeof(rt_gpu_node) * all.size());
  if (!roots.empty()) std::memcpy(r, roots.data(), sizeof(uint32_t) * roots.size());
  if (!order.empty()) std::memcpy(o, order.data(), sizeof(uint32_t) * order.size());
  *nodes_out = n;
  *nnodes_out = (uint32_t)all.size();
  *roots_out = r;
  *nroots_out = (uint32_t)roots.size();
  *prim_order_out = o;
  return RT_OK;
}

extern "C" int rt_raybvh_free_handoff(rt_gpu_node* nodes, uint32_t* roots, uint32_t* prim_order) {
  std::free(nodes);
  std::free(roots);
  std::free(prim_order);
  return RT_OK;
}

extern "C" int rt_bvh_order(const float* v9, uint32_t ntri, uint32_t max_prims,
                            int32_t split_method, uint32_t* order) {
  if (ntri && (!v9 || !order)) return RT_ERR_INVALID;
  if (max_prims == 0) return RT_ERR_INVALID;
  if (split_method < RT_SPLIT_SAH || split_method > RT_SPLIT_EQUAL) return RT_ERR_INVALID;
  Builder b(max_prims, split_method);  // leaves of at most min(255, max_prims), as BVHAccel
  b.build(v9, ntri);
  const std::vector<uint32_t>& o = b.order();
  std::copy(o.begin(), o.end(), order);
  return RT_OK;
}
