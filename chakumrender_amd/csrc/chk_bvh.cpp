// Host BVH build: binned SAH binary tree -> 4-wide collapse -> 128-byte SoA
// nodes (host side) -> 64-byte quantized nodes (what the device traverses).
// See chk_bvh.h.
#include "chk_bvh.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace chk {
namespace {

struct Box {
  float lo[3], hi[3];
  void reset() {
    lo[0] = lo[1] = lo[2] = INFINITY;
    hi[0] = hi[1] = hi[2] = -INFINITY;
  }
  void grow(const Box &b) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], b.lo[k]);
      hi[k] = std::max(hi[k], b.hi[k]);
    }
  }
  void grow(const float *p) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], p[k]);
      hi[k] = std::max(hi[k], p[k]);
    }
  }
  double area() const {
    double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
    if (!(dx >= 0) || !(dy >= 0) || !(dz >= 0)) return 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
  }
};

struct Node2 {
  Box box;
  uint32_t left, right;  // internal: children; leaf: left = first, right = count | kLeafBit
  bool is_leaf() const { return (right & kLeafBit) != 0; }
};

struct Builder {
  const float *tris;
  std::vector<Box> pbox;
  std::vector<float> cent;  // 3 per prim
  std::vector<uint32_t> idx;
  std::vector<Node2> nodes;
  std::atomic<uint32_t> next{0};

  static constexpr int kBins = 32;
  static constexpr double kTrav = 1.0, kIsect = 1.0;

  uint32_t alloc2() { return next.fetch_add(2); }

  void build(uint32_t ni, uint32_t begin, uint32_t end, int depth) {
    Box b, cb;
    b.reset();
    cb.reset();
    for (uint32_t i = begin; i < end; ++i) {
      b.grow(pbox[idx[i]]);
      cb.grow(&cent[3 * (size_t)idx[i]]);
    }
    Node2 &nd = nodes[ni];
    nd.box = b;
    uint32_t n = end - begin;
    auto make_leaf = [&]() {
      nd.left = begin;
      nd.right = n | kLeafBit;
    };
    if (n == 1) {
      make_leaf();
      return;
    }
    // binned SAH over all three axes
    double best_cost = INFINITY;
    int best_axis = -1, best_bin = -1;
    double parent_area = b.area();
    for (int ax = 0; ax < 3; ++ax) {
      float ext = cb.hi[ax] - cb.lo[ax];
      if (!(ext > 0.0f)) continue;
      Box bb[kBins];
      uint32_t cnt[kBins] = {};
      for (int k = 0; k < kBins; ++k) bb[k].reset();
      float scale = kBins / ext;
      for (uint32_t i = begin; i < end; ++i) {
        uint32_t p = idx[i];
        int k = std::min(kBins - 1, (int)((cent[3 * (size_t)p + ax] - cb.lo[ax]) * scale));
        cnt[k]++;
        bb[k].grow(pbox[p]);
      }
      double ra[kBins];
      uint32_t rc[kBins];
      Box acc;
      acc.reset();
      uint32_t c = 0;
      for (int k = kBins - 1; k > 0; --k) {
        acc.grow(bb[k]);
        c += cnt[k];
        ra[k] = acc.area();
        rc[k] = c;
      }
      acc.reset();
      c = 0;
      for (int k = 0; k < kBins - 1; ++k) {
        acc.grow(bb[k]);
        c += cnt[k];
        if (c == 0 || rc[k + 1] == 0) continue;
        double cost = kTrav + kIsect * (acc.area() * c + ra[k + 1] * rc[k + 1]) /
                                  std::max(parent_area, 1e-30);
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = ax;
          best_bin = k;
        }
      }
    }
    double leaf_cost = kIsect * n;
    if (n <= kMaxLeaf && (best_axis < 0 || leaf_cost <= best_cost)) {
      make_leaf();
      return;
    }
    uint32_t mid;
    if (best_axis < 0) {
      // all centroids coincide: split the range in the middle
      mid = begin + n / 2;
    } else {
      float ext = cb.hi[best_axis] - cb.lo[best_axis];
      float scale = kBins / ext;
      float lo = cb.lo[best_axis];
      uint32_t *m = std::partition(&idx[begin], &idx[0] + end, [&](uint32_t p) {
        int k = std::min(kBins - 1, (int)((cent[3 * (size_t)p + best_axis] - lo) * scale));
        return k <= best_bin;
      });
      mid = (uint32_t)(m - &idx[0]);
      if (mid == begin || mid == end) mid = begin + n / 2;
    }
    uint32_t c = alloc2();
    nd.left = c;
    nd.right = c + 1;
    if (n > 4096) {
#pragma omp task firstprivate(c, begin, mid, depth)
      build(c, begin, mid, depth + 1);
#pragma omp task firstprivate(c, mid, end, depth)
      build(c + 1, mid, end, depth + 1);
#pragma omp taskwait
    } else {
      build(c, begin, mid, depth + 1);
      build(c + 1, mid, end, depth + 1);
    }
  }
};

struct Collapser {
  const std::vector<Node2> &n2;
  std::vector<Node4> &out;
  uint32_t max_depth = 0;

  // returns node4 index; *bound = stack bound of this subtree
  uint32_t collapse(uint32_t root2, uint32_t depth, uint32_t *bound) {
    uint32_t kids[4];
    int k = 0;
    if (n2[root2].is_leaf()) {
      kids[k++] = root2;
    } else {
      kids[k++] = n2[root2].left;
      kids[k++] = n2[root2].right;
      while (k < 4) {
        int best = -1;
        double ba = -1;
        for (int i = 0; i < k; ++i) {
          if (n2[kids[i]].is_leaf()) continue;
          double a = n2[kids[i]].box.area();
          if (a > ba) {
            ba = a;
            best = i;
          }
        }
        if (best < 0) break;
        uint32_t open = kids[best];
        kids[best] = n2[open].left;
        kids[k++] = n2[open].right;
      }
    }
    uint32_t me = (uint32_t)out.size();
    out.emplace_back();
    max_depth = std::max(max_depth, depth + 1);
    uint32_t child_code[4];
    uint32_t sub = 0;
    for (int i = 0; i < k; ++i) {
      const Node2 &c = n2[kids[i]];
      if (c.is_leaf()) {
        child_code[i] = leaf_make(c.left, c.right & ~kLeafBit);
      } else {
        uint32_t b = 0;
        child_code[i] = collapse(kids[i], depth + 1, &b);
        sub = std::max(sub, b);
      }
    }
    Node4 &nd = out[me];
    std::memset(&nd, 0, sizeof(nd));
    for (int i = 0; i < 4; ++i) {
      if (i < k) {
        const Box &b = n2[kids[i]].box;
        nd.lo_x[i] = b.lo[0];
        nd.hi_x[i] = b.hi[0];
        nd.lo_y[i] = b.lo[1];
        nd.hi_y[i] = b.hi[1];
        nd.lo_z[i] = b.lo[2];
        nd.hi_z[i] = b.hi[2];
        nd.child[i] = child_code[i];
      } else {
        // empty slot: lo = hi = +inf misses for every ray direction
        nd.lo_x[i] = nd.lo_y[i] = nd.lo_z[i] = INFINITY;
        nd.hi_x[i] = nd.hi_y[i] = nd.hi_z[i] = INFINITY;
        nd.child[i] = kEmptyChild;
      }
    }
    *bound = (uint32_t)(k - 1) + sub;
    return me;
  }
};

// One axis of a node: grid origin o (the lowest low plane), step s (a power
// of two with every high plane within 255 steps) and the children's indices,
// checked in quad precision so that o + q s never cuts into a box.
void quantize_axis(const Node4 &nd, const float *lo, const float *hi, float &o_out, float &s_out,
                   uint32_t &qlo, uint32_t &qhi) {
  float o = INFINITY, top = -INFINITY;
  for (int i = 0; i < 4; ++i) {
    if (nd.child[i] == kEmptyChild) continue;
    o = std::min(o, lo[i]);
    top = std::max(top, hi[i]);
  }
  qlo = qhi = 0;
  if (!(o <= top)) {  // no child: any grid (the slots are skipped by their id)
    o_out = 0.0f;
    s_out = 1.0f;
    return;
  }
  const __float128 O = o;
  const double extent = (double)((__float128)top - O);
  int k = -126;
  while (k < 127 && std::ldexp(255.0, k) < extent) ++k;
  // (finite boxes: extent < 2^129, so some k <= 121 always fits)
  for (; k <= 127; ++k) {
    const double s = std::ldexp(1.0, k);
    bool ok = true;
    uint32_t ql = 0, qh = 0;
    for (int i = 0; i < 4; ++i) {
      if (nd.child[i] == kEmptyChild) continue;
      long long a = (long long)std::floor((double)(((__float128)lo[i] - O) / s));
      long long b = (long long)std::ceil((double)(((__float128)hi[i] - O) / s));
      while (a > 0 && O + (__float128)a * s > (__float128)lo[i]) --a;
      while (O + (__float128)b * s < (__float128)hi[i]) ++b;
      if (a < 0 || b > 255) {
        ok = false;
        break;
      }
      ql |= (uint32_t)a << (8 * i);
      qh |= (uint32_t)b << (8 * i);
    }
    if (ok) {
      o_out = o;
      s_out = (float)s;
      qlo = ql;
      qhi = qh;
      return;
    }
  }
  o_out = o;  // not reached for finite boxes
  s_out = std::ldexp(1.0f, 127);
  qlo = 0;
  qhi = 0xFFFFFFFFu;
}

}  // namespace

bool build_bvh4(const float *tris, uint32_t n, Bvh4 &out, std::string *msg) {
  Builder B;
  B.tris = tris;
  std::vector<uint32_t> keep;
  keep.reserve(n);
  out.orig_to_leaf.assign(n, -1);
  for (uint32_t i = 0; i < n; ++i) {
    const float *p = tris + 9 * (size_t)i;
    bool finite = true;
    for (int k = 0; k < 9; ++k) finite = finite && std::isfinite(p[k]);
    if (!finite) continue;
    double e1[3], e2[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = (double)p[3 + k] - p[k];
      e2[k] = (double)p[6 + k] - p[k];
    }
    double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
           cz = e1[0] * e2[1] - e1[1] * e2[0];
    if (!(cx * cx + cy * cy + cz * cz > 0.0)) continue;  // zero area: never hit
    keep.push_back(i);
  }
  if (keep.size() >= (1u << 28)) {
    if (msg) *msg = "too many triangles for the 28-bit leaf index";
    return false;
  }
  uint32_t m = (uint32_t)keep.size();
  out.stats = BvhStats();
  out.stats.n_tris = m;
  out.nodes.clear();
  out.tris.clear();
  out.leaf_to_orig.clear();
  if (m == 0) {
    Node4 nd;
    std::memset(&nd, 0, sizeof(nd));
    for (int i = 0; i < 4; ++i) {
      nd.lo_x[i] = nd.lo_y[i] = nd.lo_z[i] = INFINITY;
      nd.hi_x[i] = nd.hi_y[i] = nd.hi_z[i] = INFINITY;
      nd.child[i] = kEmptyChild;
    }
    out.nodes.push_back(nd);
    Node4Q q;
    std::memset(&q, 0, sizeof(q));
    q.sx = q.sy = q.sz = 1.0f;
    for (int i = 0; i < 4; ++i) q.child[i] = kEmptyChild;
    out.qnodes.assign(1, q);
    out.stats.n_nodes = 1;
    out.stats.max_depth = 1;
    return true;
  }
  B.pbox.resize(m);
  B.cent.resize(3 * (size_t)m);
  B.idx.resize(m);
  for (uint32_t j = 0; j < m; ++j) {
    const float *p = tris + 9 * (size_t)keep[j];
    Box b;
    b.reset();
    b.grow(p);
    b.grow(p + 3);
    b.grow(p + 6);
    B.pbox[j] = b;
    for (int k = 0; k < 3; ++k) B.cent[3 * (size_t)j + k] = 0.5f * (b.lo[k] + b.hi[k]);
    B.idx[j] = j;
  }
  B.nodes.resize(2 * (size_t)m);
  B.next.store(1);
#pragma omp parallel
#pragma omp single
  B.build(0, 0, m, 0);
  B.nodes.resize(B.next.load());

  Collapser Cl{B.nodes, out.nodes};
  out.nodes.reserve(B.nodes.size() / 2 + 1);
  uint32_t bound = 0;
  Cl.collapse(0, 0, &bound);
  out.stats.n_nodes = (uint32_t)out.nodes.size();
  out.stats.max_depth = Cl.max_depth;
  out.stats.stack_bound = bound;
  uint32_t leaves = 0;
  for (const Node4 &nd : out.nodes)
    for (int i = 0; i < 4; ++i)
      if (nd.child[i] != kEmptyChild && (nd.child[i] & kLeafBit)) ++leaves;
  out.stats.n_leaves = leaves;

  out.tris.resize(m);
  out.leaf_to_orig.resize(m);
  for (uint32_t j = 0; j < m; ++j) {
    uint32_t orig = keep[B.idx[j]];
    const float *p = tris + 9 * (size_t)orig;
    TriRec &t = out.tris[j];
    t.a[0] = p[0]; t.a[1] = p[1]; t.a[2] = p[2];
    std::memcpy(&t.a[3], &orig, 4);
    t.b[0] = p[3]; t.b[1] = p[4]; t.b[2] = p[5]; t.b[3] = 0.0f;
    t.c[0] = p[6]; t.c[1] = p[7]; t.c[2] = p[8]; t.c[3] = 0.0f;
    out.leaf_to_orig[j] = orig;
    out.orig_to_leaf[orig] = (int32_t)j;
  }
  // SAH cost estimate of the 4-wide tree (relative units)
  double root_area = 0;
  {
    Box rb;
    rb.reset();
    for (uint32_t j = 0; j < m; ++j) rb.grow(B.pbox[j]);
    root_area = std::max(rb.area(), 1e-30);
  }
  double cost = 0;
  for (const Node4 &nd : out.nodes) {
    for (int i = 0; i < 4; ++i) {
      if (nd.child[i] == kEmptyChild) continue;
      Box b;
      b.lo[0] = nd.lo_x[i]; b.hi[0] = nd.hi_x[i];
      b.lo[1] = nd.lo_y[i]; b.hi[1] = nd.hi_y[i];
      b.lo[2] = nd.lo_z[i]; b.hi[2] = nd.hi_z[i];
      double a = b.area() / root_area;
      cost += (nd.child[i] & kLeafBit) ? a * leaf_count(nd.child[i]) : a;
    }
  }
  out.stats.sah_cost = cost + 1.0;
  out.qnodes.resize(out.nodes.size());
  for (size_t j = 0; j < out.nodes.size(); ++j) {
    const Node4 &nd = out.nodes[j];
    Node4Q &q = out.qnodes[j];
    quantize_axis(nd, nd.lo_x, nd.hi_x, q.ox, q.sx, q.qlx, q.qhx);
    quantize_axis(nd, nd.lo_y, nd.hi_y, q.oy, q.sy, q.qly, q.qhy);
    quantize_axis(nd, nd.lo_z, nd.hi_z, q.oz, q.sz, q.qlz, q.qhz);
    for (int i = 0; i < 4; ++i) q.child[i] = nd.child[i];
  }
  return true;
}

}  // namespace chk
