// The quantized BVH nodes and the traversal's ray / node test, on the CPU
// (built host-only from the kernels' own header chk_node_test.h and linked
// with the host BVH builder):
//  * every quantized child box contains its exact box (quad precision);
//  * every child box that an exact slab test (quad precision, on the exact
//    box) accepts for a ray, node_hits_q accepts too, with its key no
//    later than the exact entry distance up to FP32 rounding.
// Scenes: small random triangles in a unit cube, the same shifted and
// scaled by 1000, a flat scene (z = 0 everywhere), long slivers.  Rays:
// random origins and directions, axis-aligned and nearly axis-aligned ones,
// origins on box faces.  Prints one summary line.
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "chk_node_test.h"

using namespace chk;
typedef __float128 q128;

static bool exact_accept(const float lo[3], const float hi[3], const float o[3], const float d[3],
                         float tmax, q128 *tn_out) {
  q128 tn = 0, tf = tmax;
  for (int a = 0; a < 3; ++a) {
    // the traversal's direction: components below 1e-20 in magnitude are
    // taken as +-1e-20 (safe_rcp)
    const float da = std::fabs(d[a]) > 1e-20f ? d[a] : std::copysign(1e-20f, d[a]);
    q128 t0 = ((q128)lo[a] - (q128)o[a]) / (q128)da;
    q128 t1 = ((q128)hi[a] - (q128)o[a]) / (q128)da;
    if (t0 > t1) std::swap(t0, t1);
    if (t0 > tn) tn = t0;
    if (t1 < tf) tf = t1;
  }
  *tn_out = tn;
  return tn <= tf;
}

int main(int argc, char **argv) {
  const int n_rays = argc > 1 ? std::atoi(argv[1]) : 2000;
  long grid_accepts = 0, boxes = 0, contain_bad = 0, tests = 0, accepts = 0, missed = 0, key_late = 0, id_bad = 0;
  for (int sc = 0; sc < 4; ++sc) {
    std::mt19937 rng(1234 + sc);
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    const uint32_t n = sc == 3 ? 3000 : 20000;
    const float scale = sc == 1 ? 1000.0f : 1.0f, shift = sc == 1 ? 5000.0f : 0.0f;
    std::vector<float> tri(9 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
      const float c[3] = {U(rng), U(rng), sc == 2 ? 0.0f : U(rng)};
      for (int v = 0; v < 3; ++v)
        for (int a = 0; a < 3; ++a) {
          float e = (sc == 3 && a == 0) ? 0.5f * U(rng) : 0.02f * U(rng);
          if (sc == 2 && a == 2) e = 0.0f;
          tri[9 * i + 3 * v + a] = (c[a] + e) * scale + shift;
        }
    }
    Bvh4 b;
    std::string msg;
    if (!build_bvh4(tri.data(), n, b, &msg)) {
      std::printf("build failed: %s\n", msg.c_str());
      return 1;
    }
    for (size_t j = 0; j < b.nodes.size(); ++j) {
      const Node4 &e = b.nodes[j];
      const Node4Q &q = b.qnodes[j];
      const float *lo[3] = {e.lo_x, e.lo_y, e.lo_z}, *hi[3] = {e.hi_x, e.hi_y, e.hi_z};
      const float og[3] = {q.ox, q.oy, q.oz}, st[3] = {q.sx, q.sy, q.sz};
      const uint32_t ql[3] = {q.qlx, q.qly, q.qlz}, qh[3] = {q.qhx, q.qhy, q.qhz};
      for (int c = 0; c < 4; ++c) {
        if (q.child[c] != e.child[c]) ++id_bad;
        if (e.child[c] == kEmptyChild) continue;
        ++boxes;
        for (int a = 0; a < 3; ++a) {
          const q128 L = (q128)og[a] + (q128)((ql[a] >> (8 * c)) & 255u) * (q128)st[a];
          const q128 H = (q128)og[a] + (q128)((qh[a] >> (8 * c)) & 255u) * (q128)st[a];
          if (L > (q128)lo[a][c] || H < (q128)hi[a][c]) ++contain_bad;
        }
      }
    }
    // rays against every node (the test is per node, so no traversal needed)
    const float lo_s = -1.5f * scale + shift, span = 3.0f * scale;
    std::uniform_real_distribution<float> U01(0.0f, 1.0f);
    for (int r = 0; r < n_rays; ++r) {
      float o[3], d[3];
      for (int a = 0; a < 3; ++a) o[a] = lo_s + span * U01(rng);
      for (int a = 0; a < 3; ++a) d[a] = U(rng);
      const int kind = r % 5;
      if (kind == 1) {  // axis-aligned
        const int a = r % 3;
        d[0] = d[1] = d[2] = 0.0f;
        d[a] = (r & 8) ? 1.0f : -1.0f;
      } else if (kind == 2) {  // nearly axis-aligned
        d[r % 3] *= 1e-7f;
      }
      float nrm = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (!(nrm > 0.0f)) d[2] = nrm = 1.0f;
      for (int a = 0; a < 3; ++a) d[a] /= nrm;
      const float tmax = (kind == 4) ? 0.3f * scale : 1e30f;
      const f3 of = mk(o[0], o[1], o[2]);
      const f3 inv = mk(safe_rcp(d[0]), safe_rcp(d[1]), safe_rcp(d[2]));
      for (size_t j = 0; j < b.nodes.size(); j += (b.nodes.size() > 4000 ? 7 : 1)) {
        const Node4 &e = b.nodes[j];
        float origin[3] = {o[0], o[1], o[2]};
        if (kind == 3) {  // origin on a face of child 0's box
          origin[0] = e.lo_x[0];
        }
        const f3 oo = kind == 3 ? mk(origin[0], origin[1], origin[2]) : of;
        float key[4];
        uint32_t cid[4];
        node_hits_q(b.qnodes.data(), (uint32_t)j, oo, inv, tmax, key, cid);
        for (int c = 0; c < 4; ++c) {
          if (e.child[c] == kEmptyChild) {
            if (key[c] < kInf) ++missed;  // an empty slot must never be entered
            continue;
          }
          ++tests;
          const float blo[3] = {e.lo_x[c], e.lo_y[c], e.lo_z[c]};
          const float bhi[3] = {e.hi_x[c], e.hi_y[c], e.hi_z[c]};
          q128 tn;
          if (!exact_accept(blo, bhi, origin, d, tmax, &tn)) continue;
          ++accepts;
          if (!(key[c] < kInf)) {
            ++missed;
          } else if ((q128)key[c] > tn * (q128)(1.0 + 1e-6) + (q128)1e-30) {
            ++key_late;
          }
        }
      }
    }
    // rays aimed at the corners, edge midpoints and face centres of child
    // boxes (grazing and boundary cases), from random origins
    for (size_t j = 0; j < b.nodes.size(); j += (b.nodes.size() > 4000 ? 5 : 1)) {
      const Node4 &e = b.nodes[j];
      for (int c = 0; c < 4; ++c) {
        if (e.child[c] == kEmptyChild) continue;
        const float blo[3] = {e.lo_x[c], e.lo_y[c], e.lo_z[c]};
        const float bhi[3] = {e.hi_x[c], e.hi_y[c], e.hi_z[c]};
        for (int k = 0; k < 6; ++k) {
          float p[3], o[3], d[3];
          for (int a = 0; a < 3; ++a) {
            const int m = (k + a + (int)j) % 3;  // low, high or middle per axis
            p[a] = m == 0 ? blo[a] : m == 1 ? bhi[a] : 0.5f * (blo[a] + bhi[a]);
            o[a] = lo_s + span * U01(rng);
          }
          if (k == 5) o[(int)j % 3] = p[(int)j % 3];  // in the plane of a face
          for (int a = 0; a < 3; ++a) d[a] = p[a] - o[a];
          float nrm = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
          if (!(nrm > 0.0f)) continue;
          for (int a = 0; a < 3; ++a) d[a] /= nrm;
          const f3 inv = mk(safe_rcp(d[0]), safe_rcp(d[1]), safe_rcp(d[2]));
          float key[4];
          uint32_t cid[4];
          node_hits_q(b.qnodes.data(), (uint32_t)j, mk(o[0], o[1], o[2]), inv, 1e30f, key, cid);
          ++tests;
          q128 tn;
          if (!exact_accept(blo, bhi, o, d, 1e30f, &tn)) continue;
          ++accepts;
          if (!(key[c] < kInf)) {
            ++missed;
          } else if ((q128)key[c] > tn * (q128)(1.0 + 1e-6) + (q128)1e-30) {
            ++key_late;
          }
        }
      }
    }
  }
  // nodes whose child boxes lie exactly on their grid (o + q s are floats),
  // so a ray through a corner, edge or face of the box passes exactly
  // through a quantized plane: the rounding of the test alone decides
  {
    std::mt19937 rng(99);
    std::uniform_real_distribution<float> U01(0.0f, 1.0f);
    std::uniform_int_distribution<int> Q(0, 255), E(-12, 4);
    std::vector<Node4Q> qn(1);
    for (int it = 0; it < 20000; ++it) {
      Node4Q &q = qn[0];
      // origin a multiple of the step, below 2^20 steps: o + q s is a float
      const float st[3] = {std::ldexp(1.0f, E(rng)), std::ldexp(1.0f, E(rng)),
                           std::ldexp(1.0f, E(rng))};
      float og[3];
      for (int a = 0; a < 3; ++a) og[a] = std::floor((U01(rng) - 0.5f) * 2097152.0f) * st[a];
      q.ox = og[0], q.oy = og[1], q.oz = og[2], q.sx = st[0], q.sy = st[1], q.sz = st[2];
      uint32_t ql[3] = {0, 0, 0}, qh[3] = {0, 0, 0};
      float blo[4][3], bhi[4][3];
      for (int c = 0; c < 4; ++c) {
        q.child[c] = (uint32_t)c;
        for (int a = 0; a < 3; ++a) {
          int l = Q(rng), h = Q(rng);
          if (l > h) std::swap(l, h);
          ql[a] |= (uint32_t)l << (8 * c);
          qh[a] |= (uint32_t)h << (8 * c);
          blo[c][a] = og[a] + (float)l * st[a];  // exact (see above)
          bhi[c][a] = og[a] + (float)h * st[a];
        }
      }
      q.qlx = ql[0], q.qhx = qh[0], q.qly = ql[1], q.qhy = qh[1], q.qlz = ql[2], q.qhz = qh[2];
      const int c = it & 3;
      for (int k = 0; k < 8; ++k) {
        float p[3], o[3], d[3];
        for (int a = 0; a < 3; ++a) {
          const int m = (k >> a) & 1;
          p[a] = m ? bhi[c][a] : blo[c][a];
          const float span = 300.0f * st[a] + std::fabs(og[a]);
          o[a] = p[a] + span * (2.0f * U01(rng) - 1.0f) * (k == 7 ? 1e3f : 1.0f);
        }
        if (it % 5 == 0) o[it % 3] = p[it % 3];  // in the plane of a face
        if (k == 6)  // from the grid origin: no absolute rounding, only inv's
          for (int a = 0; a < 3; ++a) o[a] = og[a];
        for (int a = 0; a < 3; ++a) d[a] = p[a] - o[a];
        float nrm = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (!(nrm > 0.0f)) continue;
        for (int a = 0; a < 3; ++a) d[a] /= nrm;
        const f3 inv = mk(safe_rcp(d[0]), safe_rcp(d[1]), safe_rcp(d[2]));
        float key[4];
        uint32_t cid[4];
        node_hits_q(qn.data(), 0u, mk(o[0], o[1], o[2]), inv, 1e30f, key, cid);
        ++tests;
        q128 tn;
        if (!exact_accept(blo[c], bhi[c], o, d, 1e30f, &tn)) continue;
        ++accepts;
        ++grid_accepts;
        if (!(key[c] < kInf)) {
          ++missed;
        } else if ((q128)key[c] > tn * (q128)(1.0 + 1e-6) + (q128)1e-30) {
          ++key_late;
        }
      }
    }
  }
  std::printf("grid_accepts %ld ", grid_accepts);
  std::printf("boxes %ld contain_bad %ld id_bad %ld tests %ld accepts %ld missed %ld key_late %ld\n",
              boxes, contain_bad, id_bad, tests, accepts, missed, key_late);
  return 0;
}
