/*
 * ORACLE -- test infrastructure only (see cpu_ref.h).
 *
 * FP64 restatement of the reference's per-sample radiance path.  Each
 * function cites the reference file:line it follows.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   orc_tri_intersect, orc_closest_brute, orc_bvh_*  -- golden vectors from
 *       the compiled reference's BvhAccel::intersectT (ids mapped back to
 *       original indices), plus brute force == reference BVH on every ray and
 *       closed-form hits on axis-aligned triangles.
 *   orc_sphere_intersect, orc_spheres_brute -- golden vectors from the
 *       compiled reference's BvhAccel::intersectS over Sphere primitives;
 *       closed-form hits (axis rays through the centre: t = |c - o| - r).
 *   orc_occluded_brute -- golden vectors from the compiled reference's
 *       shadow() semantics (intersectT closest hit + 1e-5 margin).
 *   orc_brdf_eval -- golden vectors from the compiled MaterialType::brdf /
 *       pdfEval; reciprocity of the GGX/Beckmann/Phong lobes (ks=1,
 *       Fresnel is symmetric in h.wi/h.wo), Monte-Carlo normalisation of
 *       pdfEval, white-furnace value of the Lambert lobe.
 *   orc_brdf_sample -- histogram of sampled directions against pdfEval
 *       (chi-square), E[f cos / pdf] against the hemisphere quadrature of
 *       f cos.  (The reference's sampler draws from clock()-seeded tables,
 *       so no golden sample exists.)
 *   orc_M_GGX, orc_ltc_shade, orc_ltc_brdf_cos -- golden vectors from the
 *       compiled M_GGX and LightSource::getLightFromMeshSource_Analytic
 *       (ks in {0, 0.5, 1}); closed forms: the diffuse polygon integral of a
 *       full-hemisphere-covering polygon is 1 and of a coaxial disc-like
 *       square matches Lambert's formula; LTC with Minv = I reduces to the
 *       diffuse integral.
 *   orc_philox4x32_10 -- the Random123 known-answer vectors.
 *   orc_render -- per-pixel mean against an independent hemisphere
 *       estimator (pure BSDF sampling, no NEE) within 3 sigma; the
 *       white-furnace closed form of a closed Lambert box lit from inside.
 */
#include "cpu_ref.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double PI = 3.14159265359;  // mathPrimitives.h:4
const double INF = 1e20;          // mathPrimitives.h:5

struct V3 {
  double x, y, z;
};
inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
inline V3 ld(const double *p) { return V3{p[0], p[1], p[2]}; }
inline void st(double *p, V3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 mul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 vdiv(V3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Vec::operator% (mathPrimitives.cpp:16)
inline V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// Vec::norm (mathPrimitives.cpp:22): v * (1/sqrt(v.v))
inline V3 norm(V3 a) { return a * (1.0 / std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z)); }
inline double length(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// Vec::maxnorm (mathPrimitives.cpp:24-29)
inline V3 maxnorm(V3 a) {
  double m = a.x;
  if (m < a.y) m = a.y;
  if (m < a.z) m = a.z;
  return v3(a.x / m, a.y / m, a.z / m);
}

// ---------------------------------------------------------------- geometry
struct Tri {
  V3 A, B, C, n;  // n: unit (C-A)x(B-A), geometryPrimitives.h:75-79
  double area;
};

inline Tri make_tri(const double *p) {
  Tri t;
  t.A = ld(p);
  t.B = ld(p + 3);
  t.C = ld(p + 6);
  V3 nn = cross(t.C - t.A, t.B - t.A);
  t.area = 0.5 * length(nn);
  t.n = norm(nn);
  return t;
}

// Triangle::intersect (geometryPrimitives.cpp:29-71).  beta/gamma out for
// the grazing margin.
inline double tri_isect(const Tri &T, V3 o, V3 d, V3 *N, double *bary_margin) {
  double raydotn = dot(d, T.n);
  V3 Nf = T.n;
  if (raydotn > 0) {
    Nf = T.n * -1.0;
    raydotn = -raydotn;
  }
  if (N) *N = Nf;
  if (raydotn * raydotn <= 1e-10) return INF;  // eps, geometryPrimitives.h:62
  double t = dot(T.A - o, Nf) / raydotn;
  if (t < 0) return INF;
  V3 P = o + d * t;
  V3 v0 = P - T.A, v1 = T.B - T.A, v2 = T.C - T.A;
  double k0 = dot(v0, v1), k1 = dot(v0, v2);
  double b0 = dot(v1, v1), b1 = dot(v1, v2);
  double c0 = dot(v2, v1), c1 = dot(v2, v2);
  double gamma = (k0 * b1 - k1 * b0) / (c0 * b1 - c1 * b0);
  double beta = (k0 - gamma * c0) / b0;
  if (gamma + beta > 1) return INF;
  else if (0 > gamma || gamma > 1) return INF;
  else if (0 > beta || beta > 1) return INF;
  if (bary_margin) *bary_margin = std::min(std::min(gamma, beta), 1.0 - gamma - beta);
  return t;
}

// AABBox::intersect (geometryPrimitives.cpp:107-139).  The y and z axes
// use 1.f / d (a float literal promoted to double: same value).
inline double aabb_isect(V3 pmin, V3 pmax, V3 o, V3 d) {
  double t0 = -INF, t1 = INF, tmp;
  const double pm[3] = {pmin.x, pmin.y, pmin.z}, px[3] = {pmax.x, pmax.y, pmax.z};
  const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  for (int a = 0; a < 3; ++a) {
    double inv = 1. / dd[a];
    double tn = (pm[a] - oo[a]) * inv, tf = (px[a] - oo[a]) * inv;
    if (tn > tf) { tmp = tn; tn = tf; tf = tmp; }
    t0 = tn > t0 ? tn : t0;
    t1 = tf < t1 ? tf : t1;
    if (t0 > t1) return INF;
  }
  return t0;
}

// ---------------------------------------------------------------- reference BVH
// bvhAccel.h:6-47 and bvhAccel.cpp:13-249.
struct Box {
  V3 mn, mx;
};
inline Box box_union(const Box &a, const Box &b) {
  return Box{v3(std::min(a.mn.x, b.mn.x), std::min(a.mn.y, b.mn.y), std::min(a.mn.z, b.mn.z)),
             v3(std::max(a.mx.x, b.mx.x), std::max(a.mx.y, b.mx.y), std::max(a.mx.z, b.mx.z))};
}
inline Box box_union_p(const Box &a, V3 p) {
  // AABBox::uNion(b, p) keeps miN/maX macro semantics ((a) < (b) ? (a) : (b))
  auto mi = [](double a, double b) { return a < b ? a : b; };
  auto ma = [](double a, double b) { return a > b ? a : b; };
  return Box{v3(mi(a.mn.x, p.x), mi(a.mn.y, p.y), mi(a.mn.z, p.z)),
             v3(ma(a.mx.x, p.x), ma(a.mx.y, p.y), ma(a.mx.z, p.z))};
}
inline Box box_union_m(const Box &a, const Box &b) {
  auto mi = [](double a, double b) { return a < b ? a : b; };
  auto ma = [](double a, double b) { return a > b ? a : b; };
  return Box{v3(mi(a.mn.x, b.mn.x), mi(a.mn.y, b.mn.y), mi(a.mn.z, b.mn.z)),
             v3(ma(a.mx.x, b.mx.x), ma(a.mx.y, b.mx.y), ma(a.mx.z, b.mx.z))};
}
// AABBox::maximumExtent (geometryPrimitives.cpp:142-150)
inline int max_extent(const Box &b) {
  V3 d = b.mx - b.mn;
  if (d.x > d.y && d.x > d.z) return 0;
  else if (d.y > d.z) return 1;
  return 2;
}

struct PrimInfo {
  Box box;
  uint32_t idx;
  V3 centroid;
};

struct BuildNode {
  Box bounds;
  int child[2];
  uint32_t axis, first, count;
};

struct LinNode {
  Box bounds;
  uint32_t offset;  // leaf: primitivesOffset; interior: secondChildOffset
  uint32_t count;
  uint32_t axis;
};

struct RefBvh {
  std::vector<Tri> tris;      // reordered (bvhAccel.cpp:114-121)
  std::vector<uint32_t> ordered;
  std::vector<PrimInfo> build;
  std::vector<BuildNode> bnodes;
  std::vector<LinNode> nodes;
};

// BvhAccel::recursiveBuild (bvhAccel.cpp:24-93)
int rec_build(RefBvh &B, uint32_t start, uint32_t end) {
  int me = (int)B.bnodes.size();
  B.bnodes.push_back(BuildNode());
  Box bbox = B.build[start].box;
  for (uint32_t i = start + 1; i < end; ++i) bbox = box_union_m(bbox, B.build[i].box);
  uint32_t np = end - start;
  auto make_leaf = [&]() {
    uint32_t first = (uint32_t)B.ordered.size();
    for (uint32_t i = start; i < end; ++i) B.ordered.push_back(B.build[i].idx);
    BuildNode &nd = B.bnodes[me];
    nd.bounds = bbox;
    nd.first = first;
    nd.count = np;
    nd.child[0] = nd.child[1] = -1;
  };
  if (np == 1) {  // nPRIM_LEAF (bvhAccel.cpp:9)
    make_leaf();
    return me;
  }
  Box cb{B.build[start].centroid, B.build[start].centroid};
  for (uint32_t i = start + 1; i < end; ++i) cb = box_union_p(cb, B.build[i].centroid);
  int dim = max_extent(cb);
  if ((cb.mx.x == cb.mn.x && dim == 0) || (cb.mx.y == cb.mn.y && dim == 1) ||
      (cb.mx.z == cb.mn.z && dim == 2)) {
    make_leaf();
    return me;
  }
  V3 pmid = (cb.mn + cb.mx) * 0.5;
  double midc = dim == 0 ? pmid.x : dim == 1 ? pmid.y : pmid.z;
  PrimInfo *mp = std::partition(&B.build[start], &B.build[end - 1] + 1, [&](const PrimInfo &a) {
    return (dim == 0 && a.centroid.x < midc) || (dim == 1 && a.centroid.y < midc) ||
           (dim == 2 && a.centroid.z < midc);
  });
  uint32_t mid = (uint32_t)(mp - &B.build[0]);
  if (mid == end || mid == start) {
    make_leaf();
    return me;
  }
  int c0 = rec_build(B, start, mid);
  int c1 = rec_build(B, mid, end);
  BuildNode &nd = B.bnodes[me];
  nd.child[0] = c0;
  nd.child[1] = c1;
  nd.bounds = box_union_m(B.bnodes[c0].bounds, B.bnodes[c1].bounds);
  nd.axis = (uint32_t)dim;
  nd.count = 0;
  return me;
}

// BvhAccel::recursiveFlatten (bvhAccel.cpp:126-143)
uint32_t rec_flatten(RefBvh &B, int node, uint32_t *offset) {
  const BuildNode &bn = B.bnodes[node];
  uint32_t my = (*offset)++;
  LinNode &ln = B.nodes[my];
  ln.bounds = bn.bounds;
  if (bn.count > 0) {
    ln.offset = bn.first;
    ln.count = bn.count;
  } else {
    ln.axis = bn.axis;
    ln.count = 0;
    rec_flatten(B, bn.child[0], offset);
    uint32_t second = rec_flatten(B, bn.child[1], offset);
    B.nodes[my].offset = second;
  }
  return my;
}

RefBvh *bvh_build(const double *tris, uint32_t n) {
  RefBvh *B = new RefBvh();
  std::vector<Tri> orig(n);
  B->build.reserve(n);
  for (uint32_t i = 0; i < n; ++i) {
    orig[i] = make_tri(tris + 9 * (size_t)i);
    // Triangle ctor box = uNion(AABBox(A, B), C) (geometryPrimitives.h:70)
    Box b{v3(std::min(orig[i].A.x, orig[i].B.x), std::min(orig[i].A.y, orig[i].B.y),
             std::min(orig[i].A.z, orig[i].B.z)),
          v3(std::max(orig[i].A.x, orig[i].B.x), std::max(orig[i].A.y, orig[i].B.y),
             std::max(orig[i].A.z, orig[i].B.z))};
    b = box_union_p(b, orig[i].C);
    PrimInfo pi;
    pi.box = b;
    pi.idx = i;
    pi.centroid = b.mn * 0.5 + b.mx * 0.5;  // bvhAccel.h:11
    B->build.push_back(pi);
  }
  if (n == 0) return B;
  B->ordered.reserve(n);
  int root = rec_build(*B, 0, n);
  B->nodes.resize(B->bnodes.size());
  uint32_t off = 0;
  rec_flatten(*B, root, &off);
  B->tris.resize(n);
  for (uint32_t i = 0; i < n; ++i) B->tris[i] = orig[B->ordered[i]];
  return B;
}

// BvhAccel::intersectT (bvhAccel.cpp:200-249).  prune adds `tnear < t`.
bool bvh_closest(const RefBvh &B, V3 o, V3 d, int prune, double *t_out, V3 *N_out, int *rid,
                 uint32_t *nv, uint32_t *nt) {
  *t_out = INF;
  *rid = -1;
  if (B.nodes.empty()) return false;
  V3 inv = v3(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
  uint32_t neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
  uint32_t todo[128];
  uint32_t sp = 0, node = 0, visits = 0, tests = 0;
  double t = INF;
  bool hit = false;
  while (true) {
    const LinNode &ln = B.nodes[node];
    ++visits;
    double tb = aabb_isect(ln.bounds.mn, ln.bounds.mx, o, d);
    if (tb < INF && (!prune || tb < t)) {
      if (ln.count > 0) {
        for (uint32_t i = 0; i < ln.count; ++i) {
          V3 Nl;
          ++tests;
          double dd = tri_isect(B.tris[ln.offset + i], o, d, &Nl, nullptr);
          if (dd != 0 && dd < t) {  // `(d = ...) && d < t`, bvhAccel.cpp:221
            t = dd;
            *rid = (int)(ln.offset + i);
            *N_out = Nl;
            hit = true;
          }
        }
        if (sp == 0) break;
        node = todo[--sp];
      } else {
        if (sp >= 127) {
          fprintf(stderr, "orc_bvh_closest: stack overflow\n");
          abort();
        }
        if (neg[ln.axis]) {
          todo[sp++] = node + 1;
          node = ln.offset;
        } else {
          todo[sp++] = ln.offset;
          node = node + 1;
        }
      }
    } else {
      if (sp == 0) break;
      node = todo[--sp];
    }
  }
  *t_out = t;
  if (nv) *nv = visits;
  if (nt) *nt = tests;
  return hit;
}

// ---------------------------------------------------------------- materials
struct Mat {
  int brdf, light;
  double ks, alpha, pe, iior, eior, refl;
  V3 c, rad;
};
inline Mat to_mat(const chk_material &m) {
  Mat r;
  r.brdf = (int)m.brdf;
  r.light = (int)m.light;
  r.ks = m.ks;
  r.alpha = m.alpha;
  r.pe = m.phong_exp;
  r.iior = m.int_ior;
  r.eior = m.ext_ior;
  r.refl = m.reflectance;
  r.c = v3(m.color[0], m.color[1], m.color[2]);
  r.rad = v3(m.radiance[0], m.radiance[1], m.radiance[2]);
  return r;
}

// MaterialType::fresnel (materialTypes.cpp:156-178)
inline double fresnel(const Mat &m, double c) {
  double g = m.iior / m.eior;
  g *= g;
  g += c * c - 1.0;
  g = std::sqrt(g);
  double t1 = g - c, t2 = g + c;
  double temp = c * t2 - 1.0;
  double F = temp * temp;
  temp = c * t1 + 1.0;
  F /= (temp * temp);
  F += 1.0;
  temp = t1 / t2;
  temp *= temp;
  F *= 0.5 * temp;
  return F;
}
// MaterialType::ggxDist (materialTypes.cpp:180-191)
inline double ggx_D(const Mat &m, V3 h, V3 n) {
  double c2 = dot(h, n);
  c2 *= c2;
  double t2 = 1.0 / c2 - 1.0;
  double a2 = m.alpha * m.alpha;
  t2 += a2;
  t2 *= t2;
  return a2 / (PI * c2 * c2 * t2);
}
// MaterialType::ggxG1 (materialTypes.cpp:193-210)
inline double ggx_G1(const Mat &m, V3 v, V3 h, V3 n) {
  double c2 = dot(v, n);
  c2 *= c2;
  double t2 = 1.0 / c2 - 1.0;
  if (t2 <= 1e-15) return 1.0;
  else if (dot(h, v) <= 1e-15) return 0;
  t2 *= m.alpha * m.alpha;
  t2 += 1.0;
  t2 = 1.0 + std::sqrt(t2);
  return 2.0 / t2;
}
// MaterialType::beckmannDist (materialTypes.cpp:212-222)
inline double beck_D(const Mat &m, V3 h, V3 n) {
  double c2 = dot(h, n);
  c2 *= c2;
  double t2 = 1.0 / c2 - 1.0;
  double a2 = m.alpha * m.alpha;
  t2 /= a2;
  t2 *= -1.0;
  return std::exp(t2) / (PI * c2 * c2 * a2);
}
// rational G1 of beckmannG1 / phongG1 (materialTypes.cpp:224-258)
inline double rational_G1(double a) {
  if (a < 1.6) return (3.535 * a + 2.181 * a * a) / (1.0 + 2.276 * a + 2.577 * a * a);
  return 1;
}
inline double beck_G1(const Mat &m, V3 v, V3 h, V3 n) {
  double c2 = dot(v, n);
  c2 *= c2;
  double a = 1.0 / (m.alpha * std::sqrt(1.0 / c2 - 1.0));
  if (c2 >= 0.999999) return 1.0;
  else if (dot(h, v) <= 1e-15) return 0;
  return rational_G1(a);
}
inline double phong_G1(const Mat &m, V3 v, V3 h, V3 n) {
  double c2 = dot(v, n);
  c2 *= c2;
  double a = 0.5 * m.alpha + 1;
  a /= (1.0 / c2 - 1.0);
  a = std::sqrt(a);
  if (c2 >= 0.999999) return 1.0;
  else if (dot(h, v) <= 1e-15) return 0;
  return rational_G1(a);
}

// MaterialType::brdf (materialTypes.cpp:9-50) times reflectance
// (geometryPrimitives.cpp:7-9), brdfApprox not set.
double brdf(const Mat &m, V3 n, V3 wo_ref, V3 wo, V3 wi) {
  double r = 0.0;
  if (m.brdf == CHK_BRDF_CLASSIC_PHONG) {
    double cosine = dot(wo_ref, wi);
    r = (m.ks * (m.pe + 1) * std::pow(cosine > 0 ? cosine : 0, m.pe) / (2 * PI)) +
        (1 - m.ks) / PI;
  } else if (m.brdf == CHK_BRDF_GGX || m.brdf == CHK_BRDF_BECKMANN || m.brdf == CHK_BRDF_PHONG) {
    V3 h = norm(wo + wi);
    double F = fresnel(m, dot(h, wi));
    double D, G;
    if (m.brdf == CHK_BRDF_GGX) {
      D = ggx_D(m, h, n);
      G = ggx_G1(m, wo, h, n) * ggx_G1(m, wi, h, n);
    } else if (m.brdf == CHK_BRDF_BECKMANN) {
      D = beck_D(m, h, n);
      G = beck_G1(m, wo, h, n) * beck_G1(m, wi, h, n);
    } else {
      D = (m.alpha + 2) * std::pow(dot(h, n), m.alpha) / (2 * PI);
      G = phong_G1(m, wo, h, n) * phong_G1(m, wi, h, n);
    }
    r = (m.ks * F * D * G) / (4.0 * dot(wi, n) * dot(wo, n)) + (1 - m.ks) / PI;
  }
  return r * m.refl;
}

// MaterialType::pdfEval (materialTypes.cpp:89-116)
double pdf_eval(const Mat &m, V3 n, V3 wo_ref, V3 wo, V3 wi) {
  double diffuse = (1.0 - m.ks) * dot(wi, n) / PI;
  double spec = m.ks;
  if (m.brdf == CHK_BRDF_CLASSIC_PHONG) {
    double cosine = dot(wi, wo_ref);
    spec *= ((m.pe + 1) * std::pow(cosine > 0 ? cosine : 0, m.pe)) / (2 * PI);
  } else if (m.brdf == CHK_BRDF_GGX || m.brdf == CHK_BRDF_BECKMANN || m.brdf == CHK_BRDF_PHONG) {
    V3 h = norm(wo + wi);
    double D = 0;
    if (m.brdf == CHK_BRDF_GGX) D = ggx_D(m, h, n);
    else if (m.brdf == CHK_BRDF_BECKMANN) D = beck_D(m, h, n);
    else D = (m.alpha + 2) * std::pow(dot(h, n), m.alpha) / (2 * PI);
    spec *= D * dot(h, n) / (4.0 * dot(h, wo));
  }
  return spec + diffuse;
}

// Duff et al. 2017 orthonormal basis.
inline void onb(V3 n, V3 *t, V3 *b) {
  double sign = std::copysign(1.0, n.z);
  double a = -1.0 / (sign + n.z);
  double bb = n.x * n.y * a;
  *t = v3(1.0 + sign * n.x * n.x * a, sign * bb, -sign * n.x);
  *b = v3(bb, sign + n.y * n.y * a, -n.y);
}

// getBrdfDirectionSampleRandomized (materialTypes.cpp:136-153) with the
// samplers of domainSampler.cpp:154-314; u = (lobe, e1, e2).
V3 brdf_sample(const Mat &m, V3 n, V3 wo_ref, V3 wo, const double *u) {
  double e1 = u[1], e2 = u[2];
  double phi = 2 * PI * e2;
  if (u[0] < m.ks) {
    if (m.brdf == CHK_BRDF_CLASSIC_PHONG) {
      // getClassicPhongDirectionSamples (domainSampler.cpp:264-314)
      double ct = std::pow((1.0 - e1), 1.0 / (m.pe + 1));
      double th = std::acos(ct);
      double s = std::sin(th);
      V3 T, Bv;
      onb(wo_ref, &T, &Bv);
      return T * (s * std::cos(phi)) + Bv * (s * std::sin(phi)) + wo_ref * ct;
    }
    // getBrdfDirectionSamples (domainSampler.cpp:193-260)
    double ct = 0, s = 0;
    if (m.brdf == CHK_BRDF_GGX) {
      double tt = m.alpha * std::sqrt(e1) / std::sqrt(1 - e1 + 1e-10);
      ct = 1.0 / std::sqrt(1 + tt * tt);
      s = tt * ct;
    } else if (m.brdf == CHK_BRDF_BECKMANN) {
      double tt = std::sqrt(m.alpha * m.alpha * std::log(1 - e1 + 1e-20) * -1.0);
      ct = 1.0 / std::sqrt(1 + tt * tt);
      s = tt * ct;
    } else if (m.brdf == CHK_BRDF_PHONG) {
      ct = std::pow(e1, 1.0 / (m.alpha + 2.0));
      s = std::sqrt(1.0 - ct * ct);
    }
    V3 T, Bv;
    onb(n, &T, &Bv);
    V3 h = T * (s * std::cos(phi)) + Bv * (s * std::sin(phi)) + n * ct;
    return norm(h * 2.0 * dot(h, wo) - wo);
  }
  // getCosineDirectionSamples (domainSampler.cpp:154-189; table built at
  // :39-49 with theta = asin(sqrt(e1)))
  double th = std::asin(std::sqrt(e1));
  double s = std::sin(th), ct = std::cos(th);
  V3 T, Bv;
  onb(n, &T, &Bv);
  return T * (s * std::cos(phi)) + Bv * (s * std::sin(phi)) + n * ct;
}

// ---------------------------------------------------------------- LTC
std::vector<double> g_ltc;  // 4096 x 19

struct M3 {
  double a[9];
  double det;
};
inline void calc_det(M3 &m) {  // mat3::calcDet (mathPrimitives.h:165-169)
  const double *a = m.a;
  m.det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
          a[2] * (a[3] * a[7] - a[4] * a[6]);
}
// mat33 -> mat3 through the 9-argument column-major ctor (ltc.h:7-10,
// mathPrimitives.h:149-163).
inline M3 from_file(const double *f) {
  M3 m;
  m.a[0] = f[0]; m.a[3] = f[1]; m.a[6] = f[2];
  m.a[1] = f[3]; m.a[4] = f[4]; m.a[7] = f[5];
  m.a[2] = f[6]; m.a[5] = f[7]; m.a[8] = f[8];
  calc_det(m);
  return m;
}
inline V3 mv(const M3 &m, V3 v) {  // mat3::mul(Vec) (mathPrimitives.h:82-86)
  return v3(m.a[0] * v.x + m.a[1] * v.y + m.a[2] * v.z, m.a[3] * v.x + m.a[4] * v.y + m.a[5] * v.z,
            m.a[6] * v.x + m.a[7] * v.y + m.a[8] * v.z);
}
inline M3 mm(const M3 &p, const M3 &q) {  // mat3::mul(mat3) (mathPrimitives.h:49-63)
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.a[i * 3 + j] = p.a[i * 3 + 0] * q.a[0 * 3 + j] + p.a[i * 3 + 1] * q.a[1 * 3 + j] +
                       p.a[i * 3 + 2] * q.a[2 * 3 + j];
  r.det = p.det * q.det;
  return r;
}

// M_GGX (ltc.h:18-61).
void M_GGX(float theta, float alpha, M3 *M, M3 *Minv, double *amplitude) {
  const int size = 64;
  int t = std::max(0, std::min(size - 1, (int)std::floor(theta / (0.5 * PI) * size)));
  int a = std::max(0, std::min(size - 1, (int)std::floor(std::sqrt((double)alpha) * size)));
  int tp = std::min(size - 1, t + 1);
  const double *r0 = &g_ltc[(size_t)(a + t * size) * 19];
  *M = from_file(r0);
  *Minv = from_file(r0 + 9);
  *amplitude = r0[18];
  if (t == tp) return;
  float mult = (float)(theta / (0.5 * PI) * size - (float)t);
  const double *r1 = &g_ltc[(size_t)(a + tp * size) * 19];
  M3 Mn = from_file(r1), Mi = from_file(r1 + 9);
  double ampn = r1[18];
  for (int i = 0; i < 9; ++i) {
    M->a[i] = M->a[i] + (Mn.a[i] - M->a[i]) * mult;
    Minv->a[i] = Minv->a[i] + (Mi.a[i] - Minv->a[i]) * mult;
  }
  *amplitude = *amplitude + (ampn - *amplitude) * mult;
}

// MaterialType::getBrdfApproxParam (materialTypes.cpp:52-74), GGX, force.
void brdf_approx(V3 n, V3 wo, double alpha, M3 *M, M3 *Minv, double *amp) {
  M_GGX((float)std::acos(dot(n, wo)), (float)alpha, M, Minv, amp);
  V3 T1 = norm(wo - n * dot(n, wo));
  V3 T2 = norm(cross(n, T1));
  // mat3(T1, T2, n).transpose(): rows T1, T2, n
  M3 R;
  R.a[0] = T1.x; R.a[1] = T1.y; R.a[2] = T1.z;
  R.a[3] = T2.x; R.a[4] = T2.y; R.a[5] = T2.z;
  R.a[6] = n.x;  R.a[7] = n.y;  R.a[8] = n.z;
  calc_det(R);
  *Minv = mm(*Minv, R);
}

// LightSource::integrateEdge (lightSources_CV.cpp:8-25)
inline double edge_n(V3 v1, V3 v2, V3 n) {
  double th = std::acos(dot(v1, v2));
  double d = dot(cross(v1, v2), n);
  return d * ((th > 0.001) ? th / std::sin(th) : 1.0);
}
inline double edge_z(V3 v1, V3 v2) {
  double th = std::acos(dot(v1, v2));
  double d = cross(v1, v2).z;
  return d * ((th > 0.001) ? th / std::sin(th) : 1.0);
}

// LightSource::clipTriangle, general plane (lightSources_CV.cpp:28-126).
bool clip_general(V3 A, V3 B, V3 C, V3 n, V3 P, double nA, double nB, double nC, V3 *clip) {
  const double eps = 1e-15;
  bool iab = false, ibc = false, ica = false;
  V3 dAB = A - B, dBC = B - C, dCA = C - A;
  double nAB = dot(n, dAB), nBC = dot(n, dBC), nCA = dot(n, dCA);
  double pA = dot(n, P - A), pB = dot(n, P - B), pC = dot(n, P - C);
  int idx = 0;
  if (nAB * nAB > eps) {
    V3 I = dAB * (pA / nAB) + A;
    if (dot(I - A, B - I) > 0) {
      iab = true;
      if (nA < 0) { clip[idx++] = I; clip[idx++] = B; }
      else { clip[idx++] = A; clip[idx++] = I; }
    }
  } else if (pA * pA < eps) return false;
  if (nBC * nBC > eps) {
    V3 I = dBC * (pB / nBC) + B;
    if (dot(I - B, C - I) > 0) {
      ibc = true;
      if (nB < 0) { clip[idx++] = I; clip[idx++] = C; }
      else { clip[idx++] = B; clip[idx++] = I; }
    }
  } else if (pB * pB < eps) return false;
  if (nCA * nCA > eps) {
    V3 I = dCA * (pC / nCA) + C;
    // Vec k1 = iCA.x - C.x; Vec k2 = A.x - iCA.x (x components only, :101-102)
    if ((I.x - C.x) * (A.x - I.x) > 0) {
      ica = true;
      if (nC < 0) { clip[idx++] = I; clip[idx++] = A; }
      else { clip[idx++] = C; clip[idx++] = I; }
    }
  } else if (pC * pC < eps) return false;
  return ((int)iab + (int)ibc + (int)ica) > 1;
}

// LightSource::clipTriangle, plane z = 0 (lightSources_CV.cpp:129-219).
bool clip_z(V3 A, V3 B, V3 C, double nA, double nB, double nC, V3 *clip) {
  const double eps = 1e-15;
  bool iab = false, ibc = false, ica = false;
  V3 dAB = A - B, dBC = B - C, dCA = C - A;
  int idx = 0;
  if (dAB.z * dAB.z > eps) {
    V3 I = dAB * (-A.z / dAB.z) + A;
    if (dot(I - A, B - I) > 0) {
      iab = true;
      if (nA < 0) { clip[idx++] = I; clip[idx++] = B; }
      else { clip[idx++] = A; clip[idx++] = I; }
    }
  } else if (A.z * A.z < eps) return false;
  if (dBC.z * dBC.z > eps) {
    V3 I = dBC * (-B.z / dBC.z) + B;
    if (dot(I - B, C - I) > 0) {
      ibc = true;
      if (nB < 0) { clip[idx++] = I; clip[idx++] = C; }
      else { clip[idx++] = B; clip[idx++] = I; }
    }
  } else if (B.z * B.z < eps) return false;
  if (dCA.z * dCA.z > eps) {
    V3 I = dCA * (-C.z / dCA.z) + C;
    if ((I.x - C.x) * (A.x - I.x) > 0) {  // x-only test, :194-197
      ica = true;
      if (nC < 0) { clip[idx++] = I; clip[idx++] = A; }
      else { clip[idx++] = C; clip[idx++] = I; }
    }
  } else if (C.z * C.z < eps) return false;
  return ((int)iab + (int)ibc + (int)ica) > 1;
}

// getLightAnalyticDiffuse (lightSources_CV.cpp:221-269), one triangle,
// before the 1/(2 pi).
double diffuse_tri(V3 A, V3 B, V3 C, V3 n, V3 x) {
  V3 e1 = norm(A - x), e2 = norm(B - x), e3 = norm(C - x);
  double nA = dot(e1, n), nB = dot(e2, n), nC = dot(e3, n);
  double sum = 0;
  if (nA > 0 && nB > 0 && nC > 0) {
    sum = edge_n(e1, e2, n) + edge_n(e2, e3, n);
    sum += edge_n(e3, e1, n);
    // reference accumulates sum = a; sum += b; sum += c
  } else if (nA < 0 && nB < 0 && nC < 0) {
    return 0;
  } else {
    V3 clip[6] = {};
    if (!clip_general(A, B, C, n, x, nA, nB, nC, clip)) return 0;
    V3 c1 = norm(clip[0] - x), c2 = norm(clip[1] - x), c3 = norm(clip[2] - x),
       c4 = norm(clip[3] - x);
    sum = edge_n(c1, c2, n);
    sum += edge_n(c2, c3, n);
    sum += edge_n(c3, c4, n);
    sum += edge_n(c4, c1, n);
  }
  return sum > 0 ? sum : sum * -1;
}

// getLightAnalytic (lightSources_CV.cpp:271-321), one triangle, before the
// 1/(2 pi).
double ltc_tri(V3 A, V3 B, V3 C, V3 x, const M3 &Minv) {
  V3 e1 = mv(Minv, A - x), e2 = mv(Minv, B - x), e3 = mv(Minv, C - x);
  double sum = 0;
  if (e1.z > 0 && e2.z > 0 && e3.z > 0) {
    e1 = norm(e1); e2 = norm(e2); e3 = norm(e3);
    sum = edge_z(e1, e2);
    sum += edge_z(e2, e3);
    sum += edge_z(e3, e1);
  } else if (e1.z < 0 && e2.z < 0 && e3.z < 0) {
    return 0;
  } else {
    V3 clip[6] = {};
    if (!clip_z(e1, e2, e3, e1.z / length(e1), e2.z / length(e2), e3.z / length(e3), clip))
      return 0;
    V3 c1 = norm(clip[0]), c2 = norm(clip[1]), c3 = norm(clip[2]), c4 = norm(clip[3]);
    sum = edge_z(c1, c2);
    sum += edge_z(c2, c3);
    sum += edge_z(c3, c4);
    sum += edge_z(c4, c1);
  }
  return sum > 0 ? sum : -sum;
}

// ---------------------------------------------------------------- Philox
inline void philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
inline double u01(uint32_t x) { return (double)(x >> 8) * (1.0 / 16777216.0); }

// ---------------------------------------------------------------- direct lighting
// The estimators of shadeDI (shader.cpp:100-109) over mesh lights.
// Philox counter {point, estimate, slot, CHK_STREAM_DIRECT}; u = (x + 0.5) 2^-32.
inline double ud(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }

struct DLight {
  uint32_t first, count;  // run of scene triangles
  double area;            // MeshLight::area (sum of Triangle::area)
  uint32_t inv[300];      // MeshLight::invCDF
};

struct DSph {  // Sphere (geometryPrimitives.h:38-52)
  V3 c;
  double r;
  uint32_t mat;
};
struct DPt {  // PointSource (lightSources.h:22-31)
  V3 p, power;
};
// EnvSource (lightSources.h:33-109) with the tables of initSamples
// (lightSources.cpp:680-782): normalizationArea, the 2000-bin inverse CDF of
// the 3000-row marginal and, per row, of the 4000-column conditional.
constexpr uint32_t kEnvAreaRows = 2000, kEnvRows = 3000, kEnvCols = 4000, kEnvBins = 2000;
struct DEnv {
  bool on = false;
  std::vector<double> img;  // h x w x 3
  uint32_t w = 0, h = 0, axis = 0;
  V3 rad;                   // BaseSource radiance (col.maxnorm() * intensity)
  double area = 0;
  std::vector<uint32_t> inv_m, inv_c;
};

// The scene of prepareScene (prepareScene.cpp:54-240): the object lists and
// LightSource's mList, tList, sList and pList.
struct DScene {
  std::vector<double> td;
  std::vector<Tri> T;
  RefBvh *B = nullptr;
  std::vector<Mat> mats;
  std::vector<uint32_t> tri_mat;
  std::vector<DLight> L;    // mesh lights
  std::vector<DLight> TL;   // triangle lights, one triangle each
  std::vector<DSph> sph;    // spheres
  std::vector<uint32_t> ssrc;  // sphere sources -> their sphere
  std::vector<DPt> pts;     // point sources
  DEnv env;                 // environment light (on or not)
  uint32_t n_tris = 0;
  ~DScene() { delete B; }
};

// A primitive: a triangle index below n_tris, else n_tris + sphere index.
inline const Mat &prim_mat(const DScene &S, uint32_t prim) {
  return prim < S.n_tris ? S.mats[S.tri_mat[prim]] : S.mats[S.sph[prim - S.n_tris].mat];
}

// Sphere::intersect (geometryPrimitives.cpp:12-27): the first root beyond
// eps = 1e-4 (geometryPrimitives.h:39), else INF.
double sph_isect(const DSph &q, V3 o, V3 d) {
  const V3 op = q.c - o;
  const double b = dot(op, d);
  double det = b * b - dot(op, op) + q.r * q.r;
  if (det < 0) return INF;
  det = std::sqrt(det);
  double t;
  return (t = b - det) > 1e-4 ? t : ((t = b + det) > 1e-4 ? t : INF);
}
// BvhAccel::intersectS (bvhAccel.cpp:151-198) over the list: the nearest
// sphere under the strict d < t (a tie keeps the earlier sphere; the
// reference's sphere BVH order may keep another).  -1 on a miss.
int sph_closest(const DScene &S, V3 o, V3 d, double *t) {
  *t = INF;
  int id = -1;
  for (size_t k = 0; k < S.sph.size(); ++k) {
    const double dd = sph_isect(S.sph[k], o, d);
    if (dd && dd < *t) {
      *t = dd;
      id = (int)k;
    }
  }
  return id;
}

// MeshLight::initMeshLight (lightSources.cpp:784-807).
void init_mesh_light(const DScene &S, DLight &L) {
  L.area = 0;
  for (uint32_t i = 0; i < L.count; ++i) L.area += S.T[L.first + i].area;
  double cdf = 0;
  uint32_t last = 0;
  for (uint32_t i = 0; i < L.count; ++i) {
    cdf += S.T[L.first + i].area / L.area;
    uint32_t index = (uint32_t)(cdf * 300);
    if (index >= 300) index = 299;
    for (uint32_t j = last; j <= index; ++j) L.inv[j] = i;
    last = index + 1;
  }
  for (uint32_t j = last; j < 300; ++j) L.inv[j] = L.count - 1;
}

// MeshLight::getSamples, one point (lightSources.cpp:809-828): e picks the
// bin, e1 = sqrt(u), e2 = u.
V3 mesh_light_point(const DScene &S, const DLight &L, const uint32_t *r, uint32_t *tri) {
  double e = ud(r[0]), e1 = std::sqrt(ud(r[1])), e2 = ud(r[2]);
  uint32_t index = (uint32_t)(e * 300);
  index = index >= 300 ? 299 : index;
  *tri = L.first + L.inv[index];
  const Tri &t = S.T[*tri];
  return t.A * (1.0 - e1) + t.B * e1 * (1.0 - e2) + t.C * e2 * e1;
}

// MeshLight::intersect (lightSources.cpp:830-847): brute force, last
// triangle first, `(d = ...) && d < t`.
bool mesh_light_isect(const DScene &S, const DLight &L, V3 o, V3 d, uint32_t *tri, double *t,
                      V3 *N) {
  *t = INF;
  *tri = 0xFFFFFFFFu;
  for (int i = (int)L.count; i--;) {
    V3 Nl;
    double dd = tri_isect(S.T[L.first + i], o, d, &Nl, nullptr);
    if (dd != 0 && dd < *t) {
      *t = dd;
      *N = Nl;
      *tri = L.first + (uint32_t)i;
    }
  }
  return *t < INF;
}

// ---------------------------------------------------------------- environment light
inline double grey(V3 L) { return L.x * 0.3 + L.y * 0.59 + L.z * 0.11; }  // mathPrimitives.h:10
// (uint32_t) of a double as the reference's x86 build converts it: NaN -> 0
inline uint32_t to_u32(double v) { return v == v ? (uint32_t)v : 0u; }

// EnvSource::getThetaPhi + getRadiance (lightSources.h:57-101): the pixel of
// unit direction d.  Where a rounding excess makes the reference's acos or
// asin NaN, its row or column conversion gives 0 (to_u32); the row is
// clamped to the image (the reference reads one row past it at theta = pi).
uint32_t env_pixel(const DEnv &E, V3 d) {
  double x, y, z;
  if (E.axis == 0) {
    x = d.y; y = d.z; z = d.x;
  } else if (E.axis == 1) {
    x = d.x; y = d.z; z = d.y;
  } else {
    x = d.x; y = d.y; z = d.z;
  }
  const double theta = std::acos(z);
  const double st = std::sqrt(1 - z * z);
  double phi;
  if (st < 1e-12) {
    phi = 0.0;
  } else {
    x /= st;
    y /= st;
    if (y >= 0 && x >= 0)
      phi = std::asin(y);
    else if (y >= 0 && x <= 0)
      phi = PI - std::asin(y);
    else if (y <= 0 && x <= 0)
      phi = PI + std::asin(-y);
    else
      phi = 2 * PI - std::asin(-y);
  }
  uint32_t col = to_u32((phi / (2 * PI)) * E.w), row = to_u32((theta / PI) * E.h);
  col = col >= E.w ? E.w - 1 : col;
  row = row >= E.h ? E.h - 1 : row;
  return row * E.w + col;
}
inline V3 env_pixel_radiance(const DEnv &E, uint32_t p) {
  return v3(E.img[3 * (size_t)p] * E.rad.x, E.img[3 * (size_t)p + 1] * E.rad.y,
            E.img[3 * (size_t)p + 2] * E.rad.z);  // envMap[.].mult(radiance)
}
inline V3 env_radiance(const DEnv &E, V3 d) { return env_pixel_radiance(E, env_pixel(E, d)); }

// invertedCDF (lightSources.cpp:662-678): bin k of the running sum takes
// index i and fills the unset bins below it; bins past the total take the
// last index.
void inverted_cdf(const double *pdf, uint32_t length, double delta, uint32_t *inv, uint32_t nb) {
  double sum = 0;
  for (uint32_t j = 0; j < nb; ++j) inv[j] = 0;
  for (uint32_t i = 0; i < length; ++i) {
    sum += pdf[i] * delta;
    const uint32_t k = to_u32(sum * nb);
    inv[k >= nb ? nb - 1 : k] = i;
    for (uint32_t j = k - 1; j < nb && inv[j] == 0; j--) inv[j] = i;
  }
  for (uint32_t j = to_u32(sum * nb); j < nb; ++j) inv[j] = length - 1;
}

// Sum over phi of grey(getRadiance) on row i of a rows x 4000 grid
// (calcArea / calcMarginalDensity, lightSources.cpp:680-722).
double env_row_sum(const DEnv &E, uint32_t rows, uint32_t i, double *row_vals) {
  const double dth = PI / rows, dph = 2.0 * PI / kEnvCols;
  const double st = std::sin(dth * i), ct = std::cos(dth * i);
  double sum = 0;
  for (uint32_t j = 0; j < kEnvCols; ++j) {
    const double cp = std::cos(dph * j), sp = std::sin(dph * j);
    const double g = grey(env_radiance(E, v3(st * cp, st * sp, ct)));
    if (row_vals) row_vals[j] = g;
    sum += g;
  }
  return sum;
}

// EnvSource::initSamples (lightSources.cpp:740-782) up to the inverse CDFs.
void env_init(DEnv &E) {
  std::vector<double> rs(kEnvAreaRows);
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t i = 0; i < (int64_t)kEnvAreaRows; ++i) rs[i] = env_row_sum(E, kEnvAreaRows, (uint32_t)i, nullptr);
  const double dth_a = PI / kEnvAreaRows, dph = 2.0 * PI / kEnvCols;
  double area = 0;
  for (uint32_t i = 0; i < kEnvAreaRows; ++i) area += rs[i] * std::sin(dth_a * i);
  E.area = area * dth_a * dph;
  // marginal density (calcMarginalDensity, :702-722)
  const double dth = PI / kEnvRows;
  std::vector<double> pm(kEnvRows);
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t i = 0; i < (int64_t)kEnvRows; ++i) {
    const double st = std::sin(dth * i);
    pm[i] = (env_row_sum(E, kEnvRows, (uint32_t)i, nullptr) * st * dph) / E.area;
  }
  E.inv_m.assign(kEnvBins, 0);
  inverted_cdf(pm.data(), kEnvRows, PI / kEnvRows, E.inv_m.data(), kEnvBins);
  // conditional density per row (calcConditionalDensity, :724-738) and its
  // inverse CDF
  E.inv_c.assign((size_t)kEnvRows * kEnvBins, 0);
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t i = 0; i < (int64_t)kEnvRows; ++i) {
    std::vector<double> g(kEnvCols);
    env_row_sum(E, kEnvRows, (uint32_t)i, g.data());
    const double st = std::sin(dth * i);
    for (uint32_t j = 0; j < kEnvCols; ++j) g[j] = g[j] * st / (E.area * pm[i]);
    inverted_cdf(g.data(), kEnvCols, 2 * PI / kEnvCols, &E.inv_c[(size_t)i * kEnvBins], kEnvBins);
  }
}

// One importance sample from (e1, e2) (lightSources.cpp:765-779).
V3 env_sample(const DEnv &E, double e1, double e2) {
  const uint32_t ti = E.inv_m[(uint32_t)(e1 * kEnvBins)];
  const double theta = ti * (PI / kEnvRows);
  const uint32_t pi_ = E.inv_c[(size_t)ti * kEnvBins + (uint32_t)(e2 * kEnvBins)];
  const double phi = pi_ * (2.0 * PI) / kEnvCols;
  double x = std::sin(theta), y = x;
  x *= std::cos(phi);
  y *= std::sin(phi);
  return v3(x, y, std::cos(theta));
}

// shadow() (shader.cpp:76-97): the closest sphere, then the closest
// triangle, + 1e-5 < distance.
bool occluded(const DScene &S, V3 o, V3 d, double dist) {
  double t;
  if (sph_closest(S, o, d, &t) >= 0 && t + 1e-5 < dist) return true;
  V3 N;
  int rid;
  return bvh_closest(*S.B, o, d, 1, &t, &N, &rid, nullptr, nullptr) && t + 1e-5 < dist;
}

// The sampling cone of sphere source j from x (mis.cpp:103-115): axis
// (c - x).norm(), sin(theta_max) = r / |c - x|, cap area 2 pi (1 - cos
// theta_max) (domainSampler.cpp:325).  Inside the source theta_max = pi / 2
// (the reference's asin has no value there).
struct DCone {
  V3 w;
  double cos_max, cap;
};
DCone sph_cone(const DScene &S, uint32_t j, V3 x) {
  const DSph &q = S.sph[S.ssrc[j]];
  const V3 v = q.c - x;
  double sn = q.r / length(v);
  sn = sn > 1.0 ? 1.0 : sn;
  DCone C;
  C.w = norm(v);
  C.cap = 2.0 * PI * (1 - std::cos(std::asin(sn)));
  C.cos_max = std::sqrt(1 - sn * sn);
  return C;
}
// SphericalSampler::getSolidDirectionSamples, one direction
// (domainSampler.cpp:320-369), in the frame it builds from w.
V3 cone_dir(const DCone &C, double e1, double e2) {
  const V3 w = C.w;
  V3 X;
  if (w.x != 0)
    X = v3((-w.y - w.z) / w.x, 1.0, 1.0);
  else if (w.y != 0)
    X = v3(1.0, -w.z / w.y, 1.0);
  else
    X = v3(1.0, 1.0, 0.0);
  X = norm(X);
  const V3 Y = norm(cross(w, X));
  const double ct = 1.0 - (C.cap * e1) / (2.0 * PI);
  const double th = std::acos(ct);
  const double sth = std::sin(th), phi = 2 * PI * e2;
  const double a = sth * std::cos(phi), b = sth * std::sin(phi);
  return v3(a * X.x + b * Y.x + ct * w.x, a * X.y + b * Y.y + ct * w.y, a * X.z + b * Y.z + ct * w.z);
}

struct DPoint {
  V3 x, n, d;
  uint32_t prim;
};
struct DCtx {
  Mat m;
  V3 wo, wo_ref;
  bool apx;  // BRDFApprox::isSet (materialTypes.cpp:52-74: GGX with force)
  M3 M, Minv;
  double amp;
  V3 G;
};

// Per-point part of the estimators: wo_ref (lightSources.cpp:44-45),
// getBrdfApproxParam and the analytic term G (lightSources_CV.cpp:323-333).
void direct_setup(const DScene &S, const DPoint &P, bool need_G, DCtx *C) {
  C->m = prim_mat(S, P.prim);
  C->wo = P.d * -1.0;
  C->wo_ref = norm(P.n * 2.0 * dot(P.n, C->wo) + P.d);
  C->apx = C->m.brdf == CHK_BRDF_GGX;
  C->amp = 0;
  if (C->apx) brdf_approx(P.n, C->wo, C->m.alpha, &C->M, &C->Minv, &C->amp);
  C->G = v3(0, 0, 0);
  if (!need_G) return;
  V3 sl = v3(0, 0, 0), sd = v3(0, 0, 0);
  for (const DLight &L : S.L)
    for (uint32_t i = 0; i < L.count; ++i) {
      const Tri &t = S.T[L.first + i];
      V3 rad = S.mats[S.tri_mat[L.first + i]].rad;
      if (C->apx) sl = sl + rad * ltc_tri(t.A, t.B, t.C, P.x, C->Minv);
      sd = sd + rad * diffuse_tri(t.A, t.B, t.C, P.n, P.x);
    }
  // getLightAnalytic * amplitude * ks: an unset BRDFApprox has amplitude 0
  V3 lt = v3(sl.x / (2.0 * PI), sl.y / (2.0 * PI), sl.z / (2.0 * PI));
  V3 df = v3(sd.x / (2.0 * PI), sd.y / (2.0 * PI), sd.z / (2.0 * PI));
  C->G = C->apx ? lt * C->amp * C->m.ks : v3(0, 0, 0);
  C->G = C->G + df * (1.0 - C->m.ks);
}

// BasePrimitive::brdf with the LTC approximation when it is set
// (materialTypes.cpp:9-11,76-87; geometryPrimitives.cpp:7-9).
double brdf_apx(const DCtx &C, V3 n, V3 wi) {
  if (!C.apx) return brdf(C.m, n, C.wo_ref, C.wo, wi);
  V3 Lo = norm(mv(C.Minv, wi));
  V3 L_ = mv(C.M, Lo);
  double l = length(L_);
  double J = C.M.det / (l * l * l);
  double D = 1.0 / PI * std::max(0.0, Lo.z);
  double res = C.amp * D / J;
  return (C.m.ks * res / dot(wi, n) + (1 - C.m.ks) / PI) * C.m.refl;
}

inline void dphilox(uint32_t a, uint32_t b, uint32_t slot, uint32_t v, const uint32_t *key,
                    uint32_t *r) {
  uint32_t c[4] = {a, b, slot, CHK_STREAM_DIRECT + v};
  philox(c, key, r);
}

// Size of MIS::getSamples' sample set: 10 per sphere source, mesh light and
// triangle light, 60 for the BSDF (mis.h:8-12).
uint32_t mis_count(const DScene &S) {
  return (uint32_t)(S.ssrc.size() + S.L.size() + S.TL.size() + (S.env.on ? 1 : 0)) *
             CHK_MIS_LIGHT_SAMPLES +
         CHK_MIS_BSDF_SAMPLES;
}

// MIS::getSamples sample set (mis.cpp:103-191): sphere-source cones, mesh
// lights, triangle lights (getTriLightSurfaceSamples, domainSampler.cpp:
// 138-150, the distribution of a one-triangle mesh light), the BSDF.
void mis_samples(const DScene &S, const DCtx &C, const DPoint &P, uint32_t a, uint32_t b,
                 uint32_t v, const uint32_t *key, V3 *out) {
  uint32_t slot = 0, r[4], tri;
  for (uint32_t j = 0; j < S.ssrc.size(); ++j) {
    const DCone K = sph_cone(S, j, P.x);
    for (uint32_t q = 0; q < CHK_MIS_LIGHT_SAMPLES; ++q, ++slot) {
      dphilox(a, b, slot, v, key, r);
      out[slot] = cone_dir(K, ud(r[0]), ud(r[1]));
    }
  }
  for (const DLight &L : S.L)
    for (uint32_t q = 0; q < CHK_MIS_LIGHT_SAMPLES; ++q, ++slot) {
      dphilox(a, b, slot, v, key, r);
      out[slot] = norm(mesh_light_point(S, L, r, &tri) - P.x);
    }
  for (const DLight &L : S.TL)
    for (uint32_t q = 0; q < CHK_MIS_LIGHT_SAMPLES; ++q, ++slot) {
      dphilox(a, b, slot, v, key, r);
      out[slot] = norm(mesh_light_point(S, L, r, &tri) - P.x);
    }
  if (S.env.on)  // EnMpIS (mis.cpp:157-169)
    for (uint32_t q = 0; q < CHK_MIS_LIGHT_SAMPLES; ++q, ++slot) {
      dphilox(a, b, slot, v, key, r);
      out[slot] = env_sample(S.env, ud(r[0]), ud(r[1]));
    }
  // getBrdfDirectionSamples (materialTypes.cpp:118-134): floor(60 ks)
  // specular-lobe samples, then cosine samples
  const uint32_t nb = (uint32_t)(CHK_MIS_BSDF_SAMPLES * C.m.ks);
  for (uint32_t q = 0; q < CHK_MIS_BSDF_SAMPLES; ++q, ++slot) {
    dphilox(a, b, slot, v, key, r);
    double u[3] = {q < nb ? -1.0 : 2.0, ud(r[1]), ud(r[2])};
    out[slot] = brdf_sample(C.m, P.n, C.wo_ref, C.wo, u);
  }
}

// 1 / (sum_k N_k pdfVal_k(s) + 1e-20) (mis.cpp:34-95,209-220), the
// techniques in getSamples' order.
double mis_weight(const DScene &S, const DCtx &C, const DPoint &P, V3 s) {
  double denom = 0;
  for (uint32_t j = 0; j < S.ssrc.size(); ++j) {  // SoAnISSP (mis.cpp:35-39)
    const DCone K = sph_cone(S, j, P.x);
    denom += (double)CHK_MIS_LIGHT_SAMPLES * (dot(s, K.w) >= K.cos_max ? 1.0 / K.cap : 0.);
  }
  for (const DLight &L : S.L) {
    uint32_t tri;
    double d;
    V3 N = v3(0, 0, 0);
    double p = 0;
    if (mesh_light_isect(S, L, P.x, s, &tri, &d, &N)) {  // Ray r(x, sample)
      double c1 = dot(s, N);
      c1 = c1 < 0 ? -c1 : c1;
      p = ((1.0 / L.area) * d * d) / c1;
    }
    denom += (double)CHK_MIS_LIGHT_SAMPLES * p;
  }
  for (const DLight &L : S.TL) {  // SuArISTR (mis.cpp:68-78)
    V3 N = v3(0, 0, 0);
    double p = 0;
    const double d = tri_isect(S.T[L.first], P.x, s, &N, nullptr);
    if (d < INF) {
      double c1 = dot(s, N);
      c1 = c1 < 0 ? -c1 : c1;
      p = ((1.0 / L.area) * d * d) / c1;
    }
    denom += (double)CHK_MIS_LIGHT_SAMPLES * p;
  }
  if (S.env.on)  // EnMpIS (mis.cpp:84-87): grey(L) / normalizationArea
    denom += (double)CHK_MIS_LIGHT_SAMPLES * (grey(env_radiance(S.env, s)) * (1.0 / S.env.area));
  denom += (double)CHK_MIS_BSDF_SAMPLES * pdf_eval(C.m, P.n, C.wo_ref, C.wo, s);
  return 1.0 / (denom + 1e-20);
}

V3 direct_estimate(const DScene &S, const chk_direct_params &p, const DPoint &P, uint32_t a,
                   uint32_t b, uint32_t v = 0) {
  const uint32_t key[2] = {p.seed_lo, p.seed_hi};
  // the mesh-light estimators return 0 without a mesh light
  // (lightSources_CV.cpp:324,336,400), the MIS one without any source
  // (lightSources.cpp:33)
  if (p.estimator != CHK_DIRECT_MIS && S.L.empty()) return v3(0, 0, 0);
  if (p.estimator == CHK_DIRECT_MIS && S.L.empty() && S.TL.empty() && S.ssrc.empty() &&
      S.pts.empty() && !S.env.on)
    return v3(0, 0, 0);
  DCtx C;
  direct_setup(S, P, p.estimator != CHK_DIRECT_MIS, &C);
  if (p.estimator == CHK_DIRECT_ANALYTIC) return C.G;
  const V3 o = P.x + P.n * 1e-8;  // rr.o (lightSources.cpp:42, lightSources_CV.cpp:346,412)
  V3 sum = v3(0, 0, 0);
  if (p.estimator == CHK_DIRECT_ANALYTIC_CV) {
    const uint32_t n = std::min(p.cv_samples, 1000u);
    for (uint32_t j = 0; j < S.L.size(); ++j) {
      const DLight &L = S.L[j];
      V3 s = v3(0, 0, 0);
      const double pdf = L.area;  // getSamples returns 1/pdf
      for (uint32_t q = 0; q < n; ++q) {
        uint32_t r[4], tri;
        dphilox(a, b, j * n + q, v, key, r);
        V3 y = mesh_light_point(S, L, r, &tri);
        V3 w = norm(y - P.x);
        double d = length(y - P.x);
        double vis = occluded(S, o, w, d) ? 0.0 : 1.0;
        double cs = dot(w, P.n);
        double c1 = dot(w, S.T[tri].n);
        c1 = c1 < 0 ? -c1 : c1;
        if (cs < 0) continue;
        double fa = brdf(C.m, P.n, C.wo_ref, C.wo, w);
        double fx = brdf_apx(C, P.n, w);
        V3 rad = S.mats[S.tri_mat[tri]].rad;
        s = s + rad * ((vis * fa - fx) * cs * c1 / (d * d));
      }
      sum = sum + s * (pdf / n);
    }
    return sum + C.G;
  }
  if (p.estimator == CHK_DIRECT_MIS) {
    // point sources (lightSources.cpp:47-58)
    for (const DPt &Q : S.pts) {
      const V3 dv = Q.p - P.x;
      const V3 w = norm(dv);
      const double cs = dot(w, P.n);
      if (cs < 0) continue;
      if (occluded(S, o, w, length(dv))) continue;
      const double sq = dot(dv, dv);
      sum = sum + vdiv(Q.power, 4.0 * PI * sq) * (cs * brdf(C.m, P.n, C.wo_ref, C.wo, w));
    }
  }
  // MIS and MISCV share MIS::getSamples
  const uint32_t Sn = mis_count(S);
  std::vector<V3> smp(Sn);
  mis_samples(S, C, P, a, b, v, key, smp.data());
  for (uint32_t i = 0; i < Sn; ++i) {
    const V3 w = smp[i];
    const double wt = mis_weight(S, C, P, w);
    const double cs = dot(w, P.n);
    if (cs < 0) continue;
    if (p.estimator == CHK_DIRECT_MIS) {
      // lightSources.cpp:60-106: the first source hit and not shadowed, in
      // the order mesh lights, sphere sources, triangle lights
      const double thr = brdf(C.m, P.n, C.wo_ref, C.wo, w) * cs * wt;
      bool hit = false;
      for (const DLight &L : S.L) {
        uint32_t tri;
        double d;
        V3 N = v3(0, 0, 0);
        if (!mesh_light_isect(S, L, o, w, &tri, &d, &N)) continue;
        if (occluded(S, o, w, d)) continue;
        sum = sum + S.mats[S.tri_mat[tri]].rad * thr;
        hit = true;
        break;
      }
      for (uint32_t j = 0; j < S.ssrc.size() && !hit; ++j) {
        const DSph &q = S.sph[S.ssrc[j]];
        const double d = sph_isect(q, o, w);
        if (d >= INF) continue;
        if (occluded(S, o, w, d)) continue;
        sum = sum + S.mats[q.mat].rad * thr;
        hit = true;
      }
      for (uint32_t j = 0; j < S.TL.size() && !hit; ++j) {
        V3 N;
        const double d = tri_isect(S.T[S.TL[j].first], o, w, &N, nullptr);
        if (d >= INF) continue;
        if (occluded(S, o, w, d)) continue;
        sum = sum + S.mats[S.tri_mat[S.TL[j].first]].rad * thr;
        hit = true;
      }
      // lightSources.cpp:102-104: nothing hit at all
      if (!hit && S.env.on && !occluded(S, o, w, INF)) sum = sum + env_radiance(S.env, w) * thr;
    } else {
      // lightSources_CV.cpp:420-445: the first mesh light hit
      for (const DLight &L : S.L) {
        uint32_t tri;
        double d;
        V3 N = v3(0, 0, 0);
        if (!mesh_light_isect(S, L, o, w, &tri, &d, &N)) continue;
        double vis = 1.0;
        if ((p.flags & CHK_DIRECT_VISIBILITY) && occluded(S, o, w, d)) vis = 0.0;
        double fa = brdf(C.m, P.n, C.wo_ref, C.wo, w);
        double fx = brdf_apx(C, P.n, w);
        sum = sum + S.mats[S.tri_mat[tri]].rad * ((vis * fa - fx) * cs * wt);
        break;
      }
    }
  }
  return p.estimator == CHK_DIRECT_MISCV ? sum + C.G : sum;
}

// ---------------------------------------------------------------- reference shaders
// Philox counter of the shaders' own draws: {pixel, sample, vertex << 8 | k,
// CHK_STREAM_SHADE}; k = try index of a direction, 0xFF for Russian roulette.
inline void sphilox(uint32_t pix, uint32_t smp, uint32_t w, const uint32_t *key, uint32_t *r) {
  uint32_t c[4] = {pix, smp, w, CHK_STREAM_SHADE};
  philox(c, key, r);
}
// getCosineDirectionSamples distribution (domainSampler.cpp:154-189), 1/pdf pi
V3 cosine_dir(V3 n, double e1, double e2) {
  double st = std::sqrt(e1), ct = std::sqrt(1.0 - e1), phi = 2 * PI * e2;
  V3 T, Bv;
  onb(n, &T, &Bv);
  return T * (st * std::cos(phi)) + Bv * (st * std::sin(phi)) + n * ct;
}
// getHemiDirectionSamples distribution (uniform, domainSampler.cpp:88-99), 1/pdf 2 pi
V3 hemi_dir(V3 n, double e1, double e2) {
  double ct = e1, st = std::sqrt(std::max(0.0, 1.0 - e1 * e1)), phi = 2 * PI * e2;
  V3 T, Bv;
  onb(n, &T, &Bv);
  return T * (st * std::cos(phi)) + Bv * (st * std::sin(phi)) + n * ct;
}

struct Hit {
  bool hit;
  double t;
  V3 N;
  uint32_t prim;
};
// The shaders' closest hit (shader.cpp:128-136,344-351): the triangle hit
// (intersectTriangle) unless a sphere is strictly nearer (intersectSphere,
// whose normal is c - x turned against d, shader.cpp:41-47).
Hit closest(const DScene &S, V3 o, V3 d) {
  Hit h;
  int rid;
  h.hit = bvh_closest(*S.B, o, d, 1, &h.t, &h.N, &rid, nullptr, nullptr);
  h.prim = h.hit ? S.B->ordered[rid] : 0u;
  const double tT = h.hit ? h.t : INF;
  double tS;
  const int id = sph_closest(S, o, d, &tS);
  if (id >= 0 && tS < tT) {
    h.hit = true;
    h.t = tS;
    V3 N = norm(S.sph[id].c - (o + d * tS));
    if (dot(d, N) > 0) N = N * -1.0;
    h.N = N;
    h.prim = S.n_tris + (uint32_t)id;
  }
  return h;
}

// shadeImplicit (shader.cpp:276-337): up to 4 vertices, cosine sampling,
// an emitter ends the path with radiance.maxnorm() times the product.
V3 shade_implicit(const DScene &S, V3 o, V3 d, uint32_t pix, uint32_t smp, const uint32_t *key,
                  bool physical) {
  V3 product = v3(1, 1, 1), col = v3(0, 0, 0);
  for (uint32_t i = 0; i < 4; ++i) {
    Hit h = closest(S, o, d);
    if (!h.hit) {
      col = v3(0, 0, 0);
      break;
    }
    const Mat &m = prim_mat(S, h.prim);
    V3 x = o + d * h.t, n = h.N;
    if (m.light != CHK_LIGHT_NONE) {
      col = mul(physical ? m.rad : maxnorm(m.rad), product);
      break;
    }
    V3 wo = d * -1.0;
    V3 wo_ref = norm(n * 2.0 * dot(n, wo) + d);
    uint32_t r[4];
    sphilox(pix, smp, i << 8, key, r);
    V3 w = cosine_dir(n, ud(r[0]), ud(r[1]));
    double b = brdf(m, n, wo_ref, wo, w);
    product = product * (PI * 1.0 * b);  // pdfInv * cosine * brdf, cosine = 1
    product = mul(m.c, product);
    o = x + n * 1e-6;
    d = w;
  }
  return col;
}

// shadeExplicit (shader.cpp:111-207): shadeDI at every vertex, Russian
// roulette with probability 0.3, uniform-hemisphere continuation that draws
// again while it hits an emitter (at most CHK_SHADE_MAX_TRIES draws; the
// reference loops until it misses one), and DI_i + col * TP_i / 0.7 from
// the last vertex back.
V3 shade_explicit(const DScene &S, const chk_direct_params &dp, V3 o, V3 d, uint32_t pix,
                  uint32_t smp, const uint32_t *key, bool physical) {
  const double q = 0.3;
  Hit h = closest(S, o, d);
  if (!h.hit) return v3(0, 0, 0);
  const Mat *m = &prim_mat(S, h.prim);
  V3 x = o + d * h.t, n = h.N, rd = d;
  uint32_t prim = h.prim;
  if (m->light != CHK_LIGHT_NONE) return physical ? m->rad : maxnorm(m->rad);
  std::vector<V3> DI, TP;
  for (uint32_t i = 0; i < 100; ++i) {
    uint32_t r[4];
    sphilox(pix, smp, (i << 8) | 0xFFu, key, r);
    const double e = ud(r[0]);
    DPoint P{x, n, rd, prim};
    DI.push_back(mul(m->c, direct_estimate(S, dp, P, pix, smp, i)));  // shadeDI
    if (e < q) {
      TP.push_back(v3(0, 0, 0));
      break;
    }
    V3 wo = rd * -1.0;
    V3 wo_ref = norm(n * 2.0 * dot(n, wo) + rd);
    V3 so = x + n * 1e-6;
    V3 w = v3(0, 0, 1);
    Hit nh;
    int status = 2;
    for (uint32_t k = 0; k < CHK_SHADE_MAX_TRIES; ++k) {
      sphilox(pix, smp, (i << 8) | k, key, r);
      w = hemi_dir(n, ud(r[0]), ud(r[1]));
      nh = closest(S, so, w);
      if (!nh.hit) break;  // the ray escapes
      if (prim_mat(S, nh.prim).light != CHK_LIGHT_NONE) continue;
      status = 1;
      break;
    }
    if (status == 2) {
      TP.push_back(v3(0, 0, 0));
      break;
    }
    double cosine = dot(w, n);
    double b = brdf(*m, n, wo_ref, wo, w);
    double product = 2 * PI * cosine * b;  // pdf (= 1/pdf, 2 pi) * cosine * brdf
    TP.push_back(m->c * product);
    rd = w;
    x = so + w * nh.t;
    n = nh.N;
    prim = nh.prim;
    m = &prim_mat(S, prim);
  }
  V3 col = DI.back();
  for (int i = (int)DI.size() - 2; i >= 0; --i) col = DI[i] + vdiv(mul(col, TP[i]), 1 - q);
  return col;
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

double orc_tri_intersect(const double *A, const double *B, const double *C, const double *o,
                         const double *d, double *N) {
  double p[9] = {A[0], A[1], A[2], B[0], B[1], B[2], C[0], C[1], C[2]};
  Tri T = make_tri(p);
  V3 Nf;
  double t = tri_isect(T, ld(o), ld(d), &Nf, nullptr);
  if (N) st(N, Nf);
  return t;
}

double orc_aabb_intersect(const double *pmin, const double *pmax, const double *o,
                          const double *d) {
  return aabb_isect(ld(pmin), ld(pmax), ld(o), ld(d));
}

double orc_sphere_intersect(const double *sphere, const double *o, const double *d) {
  const double eps = 1e-4;  // geometryPrimitives.h:39
  const V3 op = v3(sphere[0] - o[0], sphere[1] - o[1], sphere[2] - o[2]);
  const double b = op.x * d[0] + op.y * d[1] + op.z * d[2];
  double det = b * b - (op.x * op.x + op.y * op.y + op.z * op.z) + sphere[3] * sphere[3];
  if (det < 0) return INF;
  det = std::sqrt(det);
  double t;
  return (t = b - det) > eps ? t : ((t = b + det) > eps ? t : INF);
}

void orc_spheres_brute(const double *spheres, uint32_t n, const double *rays, uint32_t m,
                       int32_t *id, double *t) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < (int64_t)m; ++r) {
    const double *q = rays + 6 * r;
    double best = INF;
    int32_t bid = -1;
    for (uint32_t i = 0; i < n; ++i) {
      const double d = orc_sphere_intersect(spheres + 4 * (size_t)i, q, q + 3);
      if (d && d < best) {  // bvhAccel.cpp:171
        best = d;
        bid = (int32_t)i;
      }
    }
    id[r] = bid;
    t[r] = best;
  }
}

void orc_closest_brute(const double *tris, uint32_t n, const double *rays, uint32_t m,
                       int32_t *id, double *t, double *N, double *t2, double *margin,
                       double *cos2) {
  std::vector<Tri> T(n);
  for (uint32_t i = 0; i < n; ++i) T[i] = make_tri(tris + 9 * (size_t)i);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t r = 0; r < (int64_t)m; ++r) {
    V3 o = ld(rays + 6 * r), d = ld(rays + 6 * r + 3);
    double best = INF, second = INF, bm = 0;
    int bi = -1;
    V3 bN = v3(0, 0, 0);
    // shader.cpp:61-67 loops i = n-1 .. 0 with strict d < t
    for (int64_t i = (int64_t)n - 1; i >= 0; --i) {
      V3 Nl;
      double mg = 0;
      double dd = tri_isect(T[i], o, d, &Nl, &mg);
      if (dd != 0 && dd < INF) {
        if (dd < best) {
          second = best;
          best = dd;
          bi = (int)i;
          bN = Nl;
          bm = mg;
        } else if (dd < second) {
          second = dd;
        }
      }
    }
    id[r] = bi;
    t[r] = best;
    if (N) st(N + 3 * r, bN);
    if (t2) t2[r] = second;
    if (margin) margin[r] = bi >= 0 ? bm : 0;
    if (cos2) {
      double c = bi >= 0 ? dot(d, T[bi].n) : 0;
      cos2[r] = c * c;
    }
  }
}

void orc_occluded_brute(const double *tris, uint32_t n, const double *rays, uint32_t m,
                        uint8_t *occluded) {
  std::vector<Tri> T(n);
  for (uint32_t i = 0; i < n; ++i) T[i] = make_tri(tris + 9 * (size_t)i);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t r = 0; r < (int64_t)m; ++r) {
    V3 o = ld(rays + 7 * r), d = ld(rays + 7 * r + 3);
    double dist = rays[7 * r + 6];
    uint8_t occ = 0;
    for (uint32_t i = 0; i < n && !occ; ++i) {
      double dd = tri_isect(T[i], o, d, nullptr, nullptr);
      // shader.cpp:88-90: closest hit t with t + eps < dist (eps 1e-5)
      if (dd != 0 && dd < INF && dd + 1e-5 < dist) occ = 1;
    }
    occluded[r] = occ;
  }
}

void *orc_bvh_build(const double *tris, uint32_t n) { return bvh_build(tris, n); }
void orc_bvh_free(void *b) { delete (RefBvh *)b; }
uint32_t orc_bvh_node_count(void *b) { return (uint32_t)((RefBvh *)b)->nodes.size(); }
void orc_bvh_ordered(void *b, uint32_t *ordered) {
  RefBvh *B = (RefBvh *)b;
  std::copy(B->ordered.begin(), B->ordered.end(), ordered);
}
void orc_bvh_closest(void *b, const double *rays, uint32_t m, int prune, int32_t *id,
                     int32_t *rid, double *t, double *N, uint32_t *nv, uint32_t *nt) {
  const RefBvh &B = *(RefBvh *)b;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t r = 0; r < (int64_t)m; ++r) {
    double tt;
    V3 Nn = v3(0, 0, 0);
    int ri;
    uint32_t a = 0, c = 0;
    bvh_closest(B, ld(rays + 6 * r), ld(rays + 6 * r + 3), prune, &tt, &Nn, &ri, &a, &c);
    if (id) id[r] = ri >= 0 ? (int32_t)B.ordered[ri] : -1;
    if (rid) rid[r] = ri;
    if (t) t[r] = tt;
    if (N) st(N + 3 * r, Nn);
    if (nv) nv[r] = a;
    if (nt) nt[r] = c;
  }
}

void orc_brdf_eval(const chk_material *mats, const uint32_t *mat_idx, uint32_t k,
                   const double *n, const double *wo_ref, const double *wo, const double *wi,
                   double *f, double *pdf) {
  for (uint32_t i = 0; i < k; ++i) {
    Mat m = to_mat(mats[mat_idx[i]]);
    V3 N = ld(n + 3 * i), R = ld(wo_ref + 3 * i), O = ld(wo + 3 * i), I = ld(wi + 3 * i);
    if (f) f[i] = brdf(m, N, R, O, I);
    if (pdf) pdf[i] = pdf_eval(m, N, R, O, I);
  }
}

void orc_brdf_sample(const chk_material *mats, const uint32_t *mat_idx, uint32_t k,
                     const double *n, const double *wo_ref, const double *wo, const double *u,
                     double *wi, double *pdf) {
  for (uint32_t i = 0; i < k; ++i) {
    Mat m = to_mat(mats[mat_idx[i]]);
    V3 N = ld(n + 3 * i), R = ld(wo_ref + 3 * i), O = ld(wo + 3 * i);
    V3 w = brdf_sample(m, N, R, O, u + 3 * i);
    st(wi + 3 * i, w);
    if (pdf) pdf[i] = pdf_eval(m, N, R, O, w);
  }
}

void orc_onb(const double *n, double *t, double *b) {
  V3 T, B;
  onb(ld(n), &T, &B);
  st(t, T);
  st(b, B);
}

void orc_ltc_set(const double *table) { g_ltc.assign(table, table + 4096 * 19); }

void orc_M_GGX(float theta, float alpha, double *M, double *Minv, double *amplitude,
               double *detM) {
  M3 a, b;
  double amp;
  M_GGX(theta, alpha, &a, &b, &amp);
  std::copy(a.a, a.a + 9, M);
  std::copy(b.a, b.a + 9, Minv);
  *amplitude = amp;
  *detM = a.det;
}

void orc_ltc_shade(uint32_t k, const double *x, const double *n, const double *v,
                   const double *alpha, const double *ks, const double *tris,
                   const double *radiance, uint32_t n_tris, double *rgb, double *ltc,
                   double *diffuse, double *amp) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)k; ++i) {
    V3 X = ld(x + 3 * i), N = ld(n + 3 * i), W = ld(v + 3 * i);
    M3 M, Minv;
    double a;
    brdf_approx(N, W, alpha[i], &M, &Minv, &a);
    V3 sl = v3(0, 0, 0), sd = v3(0, 0, 0);
    for (uint32_t j = 0; j < n_tris; ++j) {
      V3 A = ld(tris + 9 * j), B = ld(tris + 9 * j + 3), C = ld(tris + 9 * j + 6);
      V3 L = ld(radiance + 3 * j);
      sl = sl + L * ltc_tri(A, B, C, X, Minv);
      sd = sd + L * diffuse_tri(A, B, C, N, X);
    }
    // sumLight / (2.0 * PI) (lightSources_CV.cpp:268,320)
    V3 lt = v3(sl.x / (2.0 * PI), sl.y / (2.0 * PI), sl.z / (2.0 * PI));
    V3 df = v3(sd.x / (2.0 * PI), sd.y / (2.0 * PI), sd.z / (2.0 * PI));
    // getLightFromMeshSource_Analytic (lightSources_CV.cpp:329-330)
    V3 G = lt * a * ks[i];
    G = G + df * (1 - ks[i]);
    if (rgb) st(rgb + 3 * i, G);
    if (ltc) st(ltc + 3 * i, lt);
    if (diffuse) st(diffuse + 3 * i, df);
    if (amp) amp[i] = a;
  }
}

void orc_ltc_brdf_cos(uint32_t k, const double *n, const double *wo, const double *wi,
                      const double *alpha, double *val) {
  for (uint32_t i = 0; i < k; ++i) {
    V3 N = ld(n + 3 * i), O = ld(wo + 3 * i), I = ld(wi + 3 * i);
    M3 M, Minv;
    double a;
    brdf_approx(N, O, alpha[i], &M, &Minv, &a);
    // approxLTC_BRDF (materialTypes.cpp:76-87)
    V3 Lo = norm(mv(Minv, I));
    V3 L_ = mv(M, Lo);
    double l = length(L_);
    double J = M.det / (l * l * l);
    double D = 1.0 / PI * std::max(0.0, Lo.z);
    val[i] = a * D / J;
  }
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

void orc_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
  philox(ctr, key, out);
}

uint64_t orc_render(const float *tris_f, uint32_t n_tris, const uint32_t *tri_mat,
                    const chk_material *mats_in, uint32_t n_mats,
                    const chk_mesh_light *lights, uint32_t n_lights, const chk_camera *cam,
                    const chk_render_params *P, uint32_t row0, uint32_t row1,
                    uint32_t row_step, int n_threads, double *accum, double *sumsq,
                    uint64_t *n_rays) {
  std::vector<double> td((size_t)n_tris * 9);
  for (size_t i = 0; i < td.size(); ++i) td[i] = tris_f[i];
  RefBvh *B = bvh_build(td.data(), n_tris);
  std::vector<Tri> T(n_tris);
  for (uint32_t i = 0; i < n_tris; ++i) T[i] = make_tri(&td[9 * (size_t)i]);
  std::vector<Mat> mats(n_mats);
  for (uint32_t i = 0; i < n_mats; ++i) mats[i] = to_mat(mats_in[i]);
  // light bookkeeping: per-light total area and an exact area CDF
  std::vector<int> tri_light(n_tris, -1);
  std::vector<double> larea(n_lights), lcdf;
  std::vector<uint32_t> lcdf_off(n_lights + 1, 0);
  for (uint32_t j = 0; j < n_lights; ++j) {
    double a = 0;
    for (uint32_t i = 0; i < lights[j].n_tris; ++i) {
      uint32_t ti = lights[j].first_tri + i;
      tri_light[ti] = (int)j;
      a += T[ti].area;
    }
    larea[j] = a;
    lcdf_off[j] = (uint32_t)lcdf.size();
    double c = 0;
    for (uint32_t i = 0; i < lights[j].n_tris; ++i) {
      c += T[lights[j].first_tri + i].area / a;
      lcdf.push_back(c);
    }
  }
  lcdf_off[n_lights] = (uint32_t)lcdf.size();

  // camera basis (prepareScene.cpp:21-33)
  const uint32_t W = cam->width, H = cam->height;
  V3 cpos = v3(cam->pos[0], cam->pos[1], cam->pos[2]);
  V3 cdir = norm(v3(cam->dir[0], cam->dir[1], cam->dir[2]));
  V3 cx = norm(cross(cdir, v3(cam->up[0], cam->up[1], cam->up[2])));
  cx = cx * ((double)W / H) * std::tan(cam->fovx_deg * PI / 360.0);
  V3 cy = norm(cross(cx, cdir)) * std::tan(cam->fovy_deg * PI / 360.0);

  const uint32_t key[2] = {P->seed_lo, P->seed_hi};
  const uint32_t D = P->max_depth;
  const double invL = n_lights ? 1.0 / n_lights : 0.0;
  uint64_t paths = 0, rays = 0;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  if (row_step == 0) row_step = 1;
  const int64_t n_rows = row1 > row0 ? ((int64_t)row1 - row0 + row_step - 1) / row_step : 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : paths, rays)
  for (int64_t ri = 0; ri < n_rows; ++ri) {
    const int64_t row = (int64_t)row0 + ri * row_step;
    uint32_t y = H - 1 - (uint32_t)row;  // idx = (h - y - 1) * w + x (main.cpp:72)
    for (uint32_t x = 0; x < W; ++x) {
      uint32_t idx = (uint32_t)row * W + x;
      double acc[3] = {0, 0, 0}, acc2[3] = {0, 0, 0};
      for (uint32_t s = P->sample_offset; s < P->sample_offset + P->spp; ++s) {
        uint32_t r4[4];
        double jx = 0.34, jy = 0.20;  // main.cpp:23-24, k = 0
        if (P->jitter_mode == CHK_JITTER_BOX) {
          uint32_t c[4] = {idx, s, (0u << 2) | CHK_STREAM_CAM, 0u};
          philox(c, key, r4);
          jx = u01(r4[0]);
          jy = u01(r4[1]);
        }
        // main.cpp:77-79
        V3 dir = cx * ((double(x) + jx) / W - .5) + cy * ((double(y) + jy) / H - .5) + cdir;
        V3 o = cpos, d = norm(dir);
        V3 L = v3(0, 0, 0), thr = v3(1, 1, 1);
        double pdf_prev = 0;
        for (uint32_t depth = 0; depth <= D; ++depth) {
          double t;
          V3 Ng;
          int rid;
          ++rays;
          if (!bvh_closest(*B, o, d, 1, &t, &Ng, &rid, nullptr, nullptr)) break;
          uint32_t prim = B->ordered[rid];
          const Mat &m = mats[tri_mat[prim]];
          if (m.light != CHK_LIGHT_NONE) {
            if (depth == 0) {
              // shadeDirectOnly: ptr->m.getRadiance = radiance.maxnorm()
              // (shader.cpp:354-356, materialTypes.h:34-38)
              L = L + ((P->flags & CHK_RENDER_PHYSICAL_EMITTER) ? m.rad : maxnorm(m.rad));
            } else {
              int j = tri_light[prim];
              double w = 1.0;
              if (j >= 0) {
                double cl = std::fabs(dot(d, T[prim].n));
                double pl = invL * t * t / (larea[j] * cl);
                w = pdf_prev / (pdf_prev + pl);  // 1/(sum N_k p_k) * N_b p_b (mis.cpp:219)
              }
              L = L + mul(thr, m.rad) * w;
            }
            break;
          }
          if (depth == D) break;
          V3 X = o + d * t;
          V3 n = Ng;  // facing normal from Triangle::intersect
          V3 wo = d * -1.0;
          V3 wo_ref = norm(n * 2.0 * dot(n, wo) + d);  // lightSources.cpp:44-45
          // ---- next-event estimation (one light sample)
          if (n_lights) {
            uint32_t c[4] = {idx, s, (depth << 2) | CHK_STREAM_NEE, 0u};
            philox(c, key, r4);
            uint32_t j = std::min((uint32_t)(u01(r4[0]) * n_lights), n_lights - 1);
            double u3 = u01(r4[3]);
            uint32_t k0 = lcdf_off[j], k1 = lcdf_off[j + 1], kk = k0;
            while (kk + 1 < k1 && u3 >= lcdf[kk]) ++kk;
            uint32_t lt = lights[j].first_tri + (kk - k0);
            // MeshLight::getSamples point (lightSources.cpp:812-824)
            double e1 = std::sqrt(u01(r4[1])), e2 = u01(r4[2]);
            V3 yp = T[lt].A * (1.0 - e1) + T[lt].B * e1 * (1.0 - e2) + T[lt].C * e2 * e1;
            V3 dv = yp - X;
            double d2 = dot(dv, dv), dist = std::sqrt(d2);
            V3 wi = dv * (1.0 / dist);
            double cx_ = dot(wi, n);
            double cl = std::fabs(dot(wi, T[lt].n));
            if (cx_ > 0 && cl > 0) {
              double pl = invL * d2 / (larea[j] * cl);  // mis.cpp:55-66
              double f = brdf(m, n, wo_ref, wo, wi);
              double pb = pdf_eval(m, n, wo_ref, wo, wi);
              double wgt = 1.0 / (pl + pb);
              V3 Le = mats[tri_mat[lt]].rad;
              V3 contrib = mul(mul(thr, m.c), Le) * (f * cx_ * wgt);
              // shadow() (shader.cpp:76-97) from x + n*1e-8 (lightSources.cpp:36,42)
              V3 so = X + n * 1e-8;
              double ts;
              V3 Ns;
              int rs;
              ++rays;
              bool occ = bvh_closest(*B, so, wi, 1, &ts, &Ns, &rs, nullptr, nullptr) &&
                         (ts + 1e-5 < dist);
              if (!occ && std::isfinite(contrib.x + contrib.y + contrib.z)) L = L + contrib;
            }
          }
          // ---- BSDF sample continues the path
          {
            uint32_t c[4] = {idx, s, (depth << 2) | CHK_STREAM_BSDF, 0u};
            philox(c, key, r4);
            double u[3] = {u01(r4[0]), u01(r4[1]), u01(r4[2])};
            V3 wi = brdf_sample(m, n, wo_ref, wo, u);
            double cs = dot(wi, n);
            if (!(cs > 0)) break;
            double pb = pdf_eval(m, n, wo_ref, wo, wi);
            double f = brdf(m, n, wo_ref, wo, wi);
            if (!(pb > 0) || !std::isfinite(pb) || !std::isfinite(f)) break;
            thr = mul(thr, m.c) * (f * cs / pb);
            pdf_prev = pb;
            o = X + n * 1e-6;  // secondary-ray offset (shader.cpp:158,315)
            d = wi;
          }
        }
        acc[0] += L.x; acc[1] += L.y; acc[2] += L.z;
        acc2[0] += L.x * L.x; acc2[1] += L.y * L.y; acc2[2] += L.z * L.z;
        ++paths;
      }
      for (int c = 0; c < 3; ++c) {
        accum[3 * (size_t)idx + c] += acc[c];
        if (sumsq) sumsq[3 * (size_t)idx + c] += acc2[c];
      }
    }
  }
  delete B;
  if (n_rays) *n_rays = rays;
  return paths;
}

void *orc_direct_scene_desc(const chk_scene_desc *D) {
  const uint32_t n_tris = D->n_tris;
  DScene *S = new DScene();
  S->n_tris = n_tris;
  S->td.resize((size_t)n_tris * 9);
  for (size_t i = 0; i < S->td.size(); ++i) S->td[i] = D->tris[i];
  S->T.resize(n_tris);
  for (uint32_t i = 0; i < n_tris; ++i) S->T[i] = make_tri(&S->td[9 * (size_t)i]);
  S->B = bvh_build(S->td.data(), n_tris);
  S->mats.resize(D->n_mats);
  for (uint32_t i = 0; i < D->n_mats; ++i) S->mats[i] = to_mat(D->mats[i]);
  S->tri_mat.assign(D->tri_mat, D->tri_mat + n_tris);
  S->L.resize(D->n_mesh_lights);
  for (uint32_t j = 0; j < D->n_mesh_lights; ++j) {
    S->L[j].first = D->mesh_lights[j].first_tri;
    S->L[j].count = D->mesh_lights[j].n_tris;
    init_mesh_light(*S, S->L[j]);
  }
  S->TL.resize(D->n_tri_lights);
  for (uint32_t j = 0; j < D->n_tri_lights; ++j) {
    S->TL[j].first = D->tri_lights[j];
    S->TL[j].count = 1;
    init_mesh_light(*S, S->TL[j]);  // area = Triangle::area
  }
  for (uint32_t j = 0; j < D->n_spheres; ++j) {
    const chk_sphere &q = D->spheres[j];
    S->sph.push_back(DSph{v3(q.center[0], q.center[1], q.center[2]), q.radius, q.mat});
  }
  S->ssrc.assign(D->sphere_sources, D->sphere_sources + D->n_sphere_sources);
  for (uint32_t j = 0; j < D->n_point_sources; ++j) {
    const chk_point_source &q = D->point_sources[j];
    S->pts.push_back(DPt{v3(q.pos[0], q.pos[1], q.pos[2]), v3(q.power[0], q.power[1], q.power[2])});
  }
  if (D->env) {
    DEnv &E = S->env;
    E.on = true;
    E.w = D->env->width;
    E.h = D->env->height;
    E.axis = D->env->axis;
    E.img.assign(D->env->image, D->env->image + 3 * (size_t)E.w * E.h);
    E.rad = v3(D->env->radiance[0], D->env->radiance[1], D->env->radiance[2]);
    env_init(E);
  }
  return S;
}

void *orc_direct_scene(const float *tris, uint32_t n_tris, const uint32_t *tri_mat,
                       const chk_material *mats, uint32_t n_mats,
                       const chk_mesh_light *lights, uint32_t n_lights) {
  chk_scene_desc D;
  std::memset(&D, 0, sizeof(D));
  D.tris = tris;
  D.n_tris = n_tris;
  D.tri_mat = tri_mat;
  D.mats = mats;
  D.n_mats = n_mats;
  D.mesh_lights = lights;
  D.n_mesh_lights = n_lights;
  return orc_direct_scene_desc(&D);
}

uint32_t orc_mis_count(void *s) { return mis_count(*(const DScene *)s); }

double orc_env_tables(void *s, uint32_t *inv_marginal, uint32_t *inv_conditional) {
  const DEnv &E = ((const DScene *)s)->env;
  if (!E.on) return 0.0;
  if (inv_marginal) std::copy(E.inv_m.begin(), E.inv_m.end(), inv_marginal);
  if (inv_conditional) std::copy(E.inv_c.begin(), E.inv_c.end(), inv_conditional);
  return E.area;
}

void orc_env_radiance(void *s, uint32_t k, const double *dirs, double *out) {
  const DEnv &E = ((const DScene *)s)->env;
  for (uint32_t i = 0; i < k; ++i) st(out + 3 * i, env_radiance(E, ld(dirs + 3 * i)));
}

void orc_env_sample(void *s, uint32_t k, const double *u, double *out) {
  const DEnv &E = ((const DScene *)s)->env;
  for (uint32_t i = 0; i < k; ++i) st(out + 3 * i, env_sample(E, u[2 * i], u[2 * i + 1]));
}

void orc_direct_scene_free(void *s) { delete (DScene *)s; }

void orc_mis_weights(void *s, uint32_t k, const double *x, const double *n,
                     const double *wo_ref, const double *wo, const uint32_t *prim,
                     const double *samples, uint32_t S, double *weights) {
  const DScene &D = *(const DScene *)s;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)k; ++i) {
    DPoint P{ld(x + 3 * i), ld(n + 3 * i), v3(0, 0, 0), prim[i]};
    DCtx C;
    C.m = prim_mat(D, P.prim);
    C.wo_ref = ld(wo_ref + 3 * i);
    C.wo = ld(wo + 3 * i);
    for (uint32_t q = 0; q < S; ++q)
      weights[(size_t)i * S + q] = mis_weight(D, C, P, ld(samples + 3 * ((size_t)i * S + q)));
  }
}

void orc_mis_samples(void *s, const double *x, const double *n, const double *wo_ref,
                     const double *wo, uint32_t prim, uint32_t point, uint32_t estimate,
                     uint32_t seed_lo, uint32_t seed_hi, double *samples) {
  const DScene &D = *(const DScene *)s;
  DPoint P{ld(x), ld(n), v3(0, 0, 0), prim};
  DCtx C;
  C.m = prim_mat(D, prim);
  C.wo_ref = ld(wo_ref);
  C.wo = ld(wo);
  const uint32_t key[2] = {seed_lo, seed_hi};
  const uint32_t Sn = mis_count(D);
  std::vector<V3> out(Sn);
  mis_samples(D, C, P, point, estimate, 0, key, out.data());
  for (uint32_t q = 0; q < Sn; ++q) st(samples + 3 * q, out[q]);
}

void orc_direct_light(void *s, const chk_direct_params *p, uint32_t k, const double *x,
                      const double *n, const double *d, const uint32_t *prim, int n_threads,
                      double *rgb) {
  const DScene &D = *(const DScene *)s;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < (int64_t)k; ++i) {
    DPoint P{ld(x + 3 * i), ld(n + 3 * i), ld(d + 3 * i), prim[i]};
    st(rgb + 3 * i, direct_estimate(D, *p, P, (uint32_t)i, p->estimate));
  }
}

void orc_render_shade(void *s, const chk_camera *cam, const chk_render_params *P, uint32_t shader,
                      const chk_direct_params *dp, uint32_t row0, uint32_t row1,
                      uint32_t row_step, int n_threads, double *accum, double *sumsq) {
  const DScene &D = *(const DScene *)s;
  const uint32_t W = cam->width, H = cam->height;
  V3 cpos = v3(cam->pos[0], cam->pos[1], cam->pos[2]);
  V3 cdir = norm(v3(cam->dir[0], cam->dir[1], cam->dir[2]));
  V3 cx = norm(cross(cdir, v3(cam->up[0], cam->up[1], cam->up[2])));
  cx = cx * ((double)W / H) * std::tan(cam->fovx_deg * PI / 360.0);
  V3 cy = norm(cross(cx, cdir)) * std::tan(cam->fovy_deg * PI / 360.0);
  const uint32_t key[2] = {P->seed_lo, P->seed_hi};
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  if (row_step == 0) row_step = 1;
  const int64_t n_rows = row1 > row0 ? ((int64_t)row1 - row0 + row_step - 1) / row_step : 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t ri = 0; ri < n_rows; ++ri) {
    const int64_t row = (int64_t)row0 + ri * row_step;
    uint32_t y = H - 1 - (uint32_t)row;  // idx = (h - y - 1) * w + x (main.cpp:72)
    for (uint32_t x = 0; x < W; ++x) {
      uint32_t idx = (uint32_t)row * W + x;
      double acc[3] = {0, 0, 0}, acc2[3] = {0, 0, 0};
      for (uint32_t sm = P->sample_offset; sm < P->sample_offset + P->spp; ++sm) {
        double jx = 0.34, jy = 0.20;  // main.cpp:23-24, k = 0
        if (P->jitter_mode == CHK_JITTER_BOX) {
          uint32_t c[4] = {idx, sm, (0u << 2) | CHK_STREAM_CAM, 0u}, r4[4];
          philox(c, key, r4);
          jx = u01(r4[0]);
          jy = u01(r4[1]);
        }
        V3 dir = cx * ((double(x) + jx) / W - .5) + cy * ((double(y) + jy) / H - .5) + cdir;
        V3 d = norm(dir);
        V3 L = v3(0, 0, 0);
        const bool phys = (P->flags & CHK_RENDER_PHYSICAL_EMITTER) != 0;
        if (shader == CHK_SHADE_IMPLICIT) {
          L = shade_implicit(D, cpos, d, idx, sm, key, phys);
        } else if (shader == CHK_SHADE_EXPLICIT) {
          L = shade_explicit(D, *dp, cpos, d, idx, sm, key, phys);
        } else {
          // shadeDirectOnly (shader.cpp:339-363)
          const Hit h = closest(D, cpos, d);
          if (h.hit) {
            const Mat &m = prim_mat(D, h.prim);
            if (m.light != CHK_LIGHT_NONE) {
              L = phys ? m.rad : maxnorm(m.rad);
            } else {
              DPoint Q{cpos + d * h.t, h.N, d, h.prim};
              L = mul(m.c, direct_estimate(D, *dp, Q, idx, sm));  // shadeDI: c.mult(light)
            }
          }
        }
        acc[0] += L.x; acc[1] += L.y; acc[2] += L.z;
        acc2[0] += L.x * L.x; acc2[1] += L.y * L.y; acc2[2] += L.z * L.z;
      }
      for (int c = 0; c < 3; ++c) {
        accum[3 * (size_t)idx + c] += acc[c];
        if (sumsq) sumsq[3 * (size_t)idx + c] += acc2[c];
      }
    }
  }
}

}  // extern "C"
