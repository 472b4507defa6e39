// Direct lighting with the reference's shadeDI estimators (shader.cpp:100-109):
//   getLightFromMeshSource_Analytic       lightSources_CV.cpp:323-333
//   getLightFromMeshSource_AnalyticCV     lightSources_CV.cpp:335-397
//   getLightFromAllSources_MIS            lightSources.cpp:32-109
//   getLightFromMeshSource_Analytic_MISCV lightSources_CV.cpp:399-450
// with MIS::getSamples' sample set and weights (mis.cpp:18-223), and the
// pixel loop around shadeDirectOnly (shader.cpp:339-363).
//
// Everything that decides a value is FP64, like the reference: sample
// directions, MIS weights (brute force over every mesh light's triangles,
// as Pdf::pdfVal does), BRDFs, the LTC fit and the polygon integrals.  Only
// visibility comes from the FP32 BVH traversal (one any-hit ray per sample).
//
// Work is a wavefront over (shading point, sample slot) items in chunks:
//   k_dl_setup / k_dr_points   per point: material, wo_ref, LTC frame, G
//   k_dl_samples               per item: Philox sample, weight, light pick,
//                              the terms of its contribution, a shadow ray
//   any hit                    visibility of every item's shadow ray
//   k_dl_reduce                per point: the reference's sums in its order
// Items are slot-major (item = slot * P + point), so a wave works on one
// slot: one mesh light or the BSDF, wave-uniform branches and light loads.
#include <chrono>
#include <cmath>
#include <cstring>

#include "chk_env_dev.h"
#include "chk_internal.h"
#include "chk_light64.h"
#include "chk_ltc_dev.h"

namespace chk {
namespace {

constexpr int kDBlock = 128;
constexpr uint64_t kItemBudget = 1ull << 24;  // (point, slot) items per chunk
constexpr uint32_t kNone = 0xFFFFFFFFu;

enum PointFlags : uint32_t { kActive = 1, kApx = 2 };

// One shading point, FP64.
struct PointRec {
  d3 x, n, wo, wo_ref;
  d3 R0, R1, R2;   // rows of the rotated Minv (getBrdfApproxParam)
  double m[5];     // M a[0], a[2], a[4], a[6], a[8]
  double det, amp;
  d3 G;            // analytic term
  d3 pre;          // value of an inactive point (render: emitter or miss)
  uint32_t ca, cb; // Philox counter words 0 and 1
  uint32_t cd;     // path vertex (added to the counter's stream word)
  uint32_t path;   // the shader path this point belongs to
  uint32_t mat, leaf;
  uint32_t flags, nb;  // nb = floor(60 ks): specular-lobe BSDF slots
};

__device__ inline double ud(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }
__device__ inline d3 f4d(float4 v) { return dk(v.x, v.y, v.z); }

// MeshLight::intersect (lightSources.cpp:830-847): brute force from the last
// triangle, `(d = ...) && d < t`.  tri: index into the light-triangle table.
__device__ inline bool light_isect(const DeviceScene &S, uint32_t j, d3 o, d3 d, uint32_t *tri,
                                   double *t, d3 *N) {
  const LightRec L = S.lights[j];
  *t = kInf64;
  *tri = kNone;
  for (int i = (int)L.count; i--;) {
    d3 Nl;
    const double dd = tri_isect64p(S.light64[L.first + i], o, d, &Nl);
    if (dd != 0 && dd < *t) {
      *t = dd;
      *N = Nl;
      *tri = L.first + (uint32_t)i;
    }
  }
  return *t < kInf64;
}

// MeshLight::getSamples, one point (lightSources.cpp:809-828): the bin of
// e in the 300-bin inverse CDF picks the triangle; e1 = sqrt(u), e2 = u.
__device__ inline d3 light_point(const DeviceScene &S, uint32_t j, u4 r, uint32_t *tri) {
  const LightRec L = S.lights[j];
  const double e = ud(r.x), e1 = sqrt(ud(r.y)), e2 = ud(r.z);
  uint32_t bin = (uint32_t)(e * 300);
  bin = bin >= 300 ? 299 : bin;
  uint32_t k = L.first;
  while (k + 1 < L.first + L.count && (float)bin > S.light_tris[k].rad.w) ++k;
  *tri = k;
  const LightTri &lt = S.light_tris[k];
  return f4d(lt.a_cdf) * (1.0 - e1) + f4d(lt.b_area) * e1 * (1.0 - e2) +
         f4d(lt.c_leaf) * e2 * e1;
}

// ------------------------------------------------------------ spheres, sources
// Emitter code of a sample item: a light-triangle index (mesh or triangle
// light), or a sphere-source / point-source index or an environment pixel
// under its tag.  kOccS marks a sample whose shadow ray a sphere blocks.
constexpr uint32_t kTagSph = 0x10000000u, kTagPt = 0x20000000u, kTagEnv = 0x30000000u;
constexpr uint32_t kTagMask = 0x30000000u;
constexpr uint32_t kCodeMask = 0x0FFFFFFFu, kOccS = 0x80000000u;
// A closest hit on a sphere: leaf word kHitSph | sphere, its FP64 t in the
// hit's z (low word) and w (high word).
constexpr uint32_t kHitSph = 0x80000000u;

// Sphere::intersect (geometryPrimitives.cpp:12-27): the first root beyond
// eps = 1e-4 (geometryPrimitives.h:39), else INF.
__device__ inline double sph_isect(const SphereRec &q, d3 o, d3 d) {
  const d3 op = dk(q.c[0], q.c[1], q.c[2]) - o;
  const double b = dot(op, d);
  double det = b * b - dot(op, op) + q.r * q.r;
  if (det < 0) return kInf64;
  det = sqrt(det);
  double t = b - det;
  if (t > 1e-4) return t;
  t = b + det;
  return t > 1e-4 ? t : kInf64;
}
// BvhAccel::intersectS (bvhAccel.cpp:151-198) as a loop over the spheres:
// the nearest hit under the strict d < t (wave-uniform loads).
__device__ inline uint32_t sph_closest(const DeviceScene &S, d3 o, d3 d, double *t) {
  *t = kInf64;
  uint32_t id = kNone;
  for (uint32_t k = 0; k < S.n_spheres; ++k) {
    const double dd = sph_isect(S.spheres[k], o, d);
    if (dd < *t) {
      *t = dd;
      id = k;
    }
  }
  return id;
}
// The sphere half of shadow() (shader.cpp:83-85).
__device__ inline bool sph_occluded(const DeviceScene &S, d3 o, d3 d, double dist) {
  double t;
  return sph_closest(S, o, d, &t) != kNone && t + 1e-5 < dist;
}

// The sampling cone of sphere source j seen from x (mis.cpp:103-115): axis
// w = (c - x).norm(), sin(theta_max) = r / |c - x| and the cap area
// 2 pi (1 - cos theta_max) (domainSampler.cpp:325).  A point inside the
// source takes theta_max = pi / 2 (the reference's asin has no value there).
struct Cone {
  d3 w;
  double cos_max, cap;
};
__device__ inline Cone sph_cone(const DeviceScene &S, uint32_t j, d3 x) {
  const SphereRec &q = S.spheres[S.sph_src[j]];
  const d3 v = dk(q.c[0], q.c[1], q.c[2]) - x;
  double sn = q.r / sqrt(dot(v, v));
  sn = sn > 1.0 ? 1.0 : sn;
  Cone C;
  C.w = v * (1.0 / sqrt(dot(v, v)));
  C.cap = 2.0 * kPi64 * (1 - cos(asin(sn)));
  C.cos_max = sqrt(1 - sn * sn);
  return C;
}
// SphericalSampler::getSolidDirectionSamples, one direction
// (domainSampler.cpp:320-369): cos theta = 1 - A e1 / (2 pi), phi = 2 pi e2
// in the frame (newX, newY, w) the reference builds from w.
__device__ inline d3 cone_dir(const Cone &C, double e1, double e2) {
  const d3 w = C.w;
  d3 X;
  if (w.x != 0)
    X = dk((-w.y - w.z) / w.x, 1.0, 1.0);
  else if (w.y != 0)
    X = dk(1.0, -w.z / w.y, 1.0);
  else
    X = dk(1.0, 1.0, 0.0);
  X = X * (1.0 / sqrt(dot(X, X)));
  d3 Y = cross(w, X);
  Y = Y * (1.0 / sqrt(dot(Y, Y)));
  const double ct = 1.0 - (C.cap * e1) / (2.0 * kPi64);
  const double st = sin(acos(ct));
  double sp, cp;
  sincos(2.0 * kPi64 * e2, &sp, &cp);
  const double a = st * cp, b = st * sp;
  return dk(a * X.x + b * Y.x + ct * w.x, a * X.y + b * Y.y + ct * w.y,
            a * X.z + b * Y.z + ct * w.z);
}
// The radiance of an emitter code.
__device__ inline d3 code_radiance(const DeviceScene &S, uint32_t code) {
  const uint32_t j = code & kCodeMask;
  if ((code & kTagMask) == kTagEnv) return env_pixel_radiance(S.env, j);
  if ((code & kTagMask) == kTagSph) {
    const chk_material &m = S.mats[S.sphere_mat[S.sph_src[j]]];
    return dk(m.radiance[0], m.radiance[1], m.radiance[2]);
  }
  if ((code & kTagMask) == kTagPt) {
    const chk_point_source &L = S.points[j];
    return dk(L.power[0], L.power[1], L.power[2]);
  }
  return f4d(S.light_tris[j].rad);
}

// Closest-hit records (float4 t, leaf word, z, w) of the shading kernels.
__device__ inline bool hit_is_sph(float4 h) {
  const uint32_t l = __float_as_uint(h.y);
  return l != kNone && (l & kHitSph);
}
__device__ inline double hit_t(float4 h) {
  if (hit_is_sph(h)) return __hiloint2double(__float_as_int(h.w), __float_as_int(h.z));
  return (double)h.x;
}
__device__ inline bool missed(const DeviceScene &S, float4 h) {
  return !(h.x < kInf) || (!hit_is_sph(h) && __float_as_uint(h.y) >= S.n_leaf_tris);
}
__device__ inline uint32_t hit_mat(const DeviceScene &S, float4 h) {
  const uint32_t l = __float_as_uint(h.y);
  return hit_is_sph(h) ? S.sphere_mat[l & ~kHitSph] : S.leaf_info[l].x;
}
// the BVH leaf a ray leaving the hit skips (none for a sphere)
__device__ inline uint32_t hit_leaf(float4 h) {
  return hit_is_sph(h) ? kNone : __float_as_uint(h.y);
}
// The unit normal facing the ray: Triangle::intersect's (C-A)x(B-A) turned
// against d (geometryPrimitives.cpp:29-71), or the sphere's c - x turned
// against d (shader.cpp:41-47).
__device__ inline d3 hit_normal(const DeviceScene &S, float4 h, d3 x, d3 d) {
  d3 n;
  if (hit_is_sph(h)) {
    const SphereRec &q = S.spheres[__float_as_uint(h.y) & ~kHitSph];
    n = dk(q.c[0], q.c[1], q.c[2]) - x;
  } else {
    const float4 *tp = S.tris + 3 * (size_t)__float_as_uint(h.y);
    const d3 A = f4d(tp[0]), B = f4d(tp[1]), C = f4d(tp[2]);
    n = cross(C - A, B - A);
  }
  n = n * (1.0 / sqrt(dot(n, n)));
  if (dot(d, n) > 0) n = n * -1.0;
  return n;
}
// A caller's primitive: a triangle index below n_tris, else a sphere
// (n_tris + sphere index).  Returns false when it names neither.
__device__ inline bool prim_info(const DeviceScene &S, uint32_t o, uint32_t *mat, uint32_t *leaf) {
  if (o < S.n_tris) {
    const uint2 oi = S.orig_info[o];
    *leaf = oi.x;
    *mat = oi.y;
    return true;
  }
  if (o - S.n_tris < S.n_spheres) {
    *leaf = kNone;
    *mat = S.sphere_mat[o - S.n_tris];
    return true;
  }
  return false;
}

// BasePrimitive::brdf with the LTC approximation when it is set
// (materialTypes.cpp:9-11,76-87; geometryPrimitives.cpp:7-9).
__device__ inline double brdf_apx_f(const MatT<double> &m, const PointRec &P, d3 wi, double f);
__device__ inline double brdf_apx(const MatT<double> &m, const PointRec &P, d3 wi) {
  double f = 0.0, pdf;
  if (!(P.flags & kApx)) brdf_pdf(m, P.n, P.wo_ref, P.wo, wi, f, pdf);
  return brdf_apx_f(m, P, wi, f);
}
// The same with brdf(wi) given (returned when no approximation is set).
__device__ inline double brdf_apx_f(const MatT<double> &m, const PointRec &P, d3 wi, double f) {
  if (!(P.flags & kApx)) return f;
  const d3 Lo = normalize(dk(dot(P.R0, wi), dot(P.R1, wi), dot(P.R2, wi)));
  const d3 L_ = dk(P.m[0] * Lo.x + P.m[1] * Lo.z, P.m[2] * Lo.y, P.m[3] * Lo.x + P.m[4] * Lo.z);
  const double l = sqrt(dot(L_, L_));
  const double J = P.det / (l * l * l);
  const double D = 1.0 / kPi64 * fmax(0.0, Lo.z);
  const double res = P.amp * D / J;
  return (m.ks * res / dot(wi, P.n) + (1 - m.ks) / kPi64) * m.refl;
}

__device__ inline bool sampler_ok(uint32_t b) {
  return b == CHK_BRDF_CLASSIC_PHONG || b == CHK_BRDF_GGX || b == CHK_BRDF_BECKMANN ||
         b == CHK_BRDF_PHONG;
}

// Per-point part: wo_ref (lightSources.cpp:44-45), getBrdfApproxParam
// (materialTypes.cpp:52-74, GGX with force) and G = getLightAnalytic * amp
// * ks + getLightAnalyticDiffuse * (1 - ks) (lightSources_CV.cpp:323-333).
__device__ void setup_point(const DeviceScene &S, const double *__restrict__ ltc, d3 x, d3 n,
                            d3 d, uint32_t mat, uint32_t leaf, bool need_G, PointRec &P) {
  const MatT<double> m = load_mat<double>(S.mats + mat);
  P.x = x;
  P.n = n;
  P.wo = d * -1.0;
  P.wo_ref = normalize(n * 2.0 * dot(n, P.wo) + d);
  P.mat = mat;
  P.leaf = leaf;
  P.flags = kActive;
  P.nb = (uint32_t)((double)CHK_MIS_BSDF_SAMPLES * m.ks);
  P.amp = 0;
  P.G = dk(0, 0, 0);
  P.pre = dk(0, 0, 0);
  if (m.brdf == CHK_BRDF_GGX && ltc) {
    const LtcFit F = ltc_fit(ltc, (float)acos(dot(n, P.wo)), m.alpha);
    const d3 T1 = normalize(P.wo - n * dot(n, P.wo));
    const d3 T2 = normalize(cross(n, T1));
    P.R0 = T1 * F.mi[0] + n * F.mi[1];
    P.R1 = T2 * F.mi[2];
    P.R2 = T1 * F.mi[3] + n * F.mi[4];
#pragma unroll
    for (int c = 0; c < 5; ++c) P.m[c] = F.m[c];
    P.det = F.det;
    P.amp = F.amp;
    P.flags |= kApx;
  }
  if (!need_G) return;
  d3 sl = dk(0, 0, 0), sd = dk(0, 0, 0);
  const bool do_ltc = (P.flags & kApx) && m.ks != 0.0, do_dif = m.ks != 1.0;
  const d3 zero = dk(0, 0, 0), zax = dk(0, 0, 1);
  for (uint32_t j = 0; j < S.n_lights; ++j) {
    const LightRec L = S.lights[j];
    for (uint32_t i = 0; i < L.count; ++i) {
      const LightTri &lt = S.light_tris[L.first + i];
      const d3 A = f4d(lt.a_cdf), B = f4d(lt.b_area), C = f4d(lt.c_leaf), rad = f4d(lt.rad);
      if (do_ltc) {
        const d3 a = A - x, b = B - x, c = C - x;
        const d3 e1 = dk(dot(P.R0, a), dot(P.R1, a), dot(P.R2, a));
        const d3 e2 = dk(dot(P.R0, b), dot(P.R1, b), dot(P.R2, b));
        const d3 e3 = dk(dot(P.R0, c), dot(P.R1, c), dot(P.R2, c));
        sl = sl + rad * tri64<true>(e1, e2, e3, zax, zero);
      }
      if (do_dif) sd = sd + rad * tri64<false>(A, B, C, n, x);
    }
  }
  const double tp = 2.0 * kPi64;
  const d3 lt = dk(sl.x / tp, sl.y / tp, sl.z / tp), df = dk(sd.x / tp, sd.y / tp, sd.z / tp);
  P.G = do_ltc ? lt * P.amp * m.ks : dk(0, 0, 0);
  P.G = P.G + df * (1.0 - m.ks);
}

// Shading points of the entry point: (x, n, d, prim) -> PointRec.
__global__ __launch_bounds__(kDBlock) void k_dl_setup(
    DeviceScene S, const double *__restrict__ ltc, const double *__restrict__ X,
    const double *__restrict__ N, const double *__restrict__ D, const uint32_t *__restrict__ prim,
    uint32_t p0, uint32_t P, uint32_t estimate, bool need_G, PointRec *__restrict__ pts,
    uint32_t *__restrict__ err) {
  const uint32_t p = blockIdx.x * kDBlock + threadIdx.x;
  if (p >= P) return;
  const size_t g = (size_t)p0 + p;
  PointRec &R = pts[p];
  uint32_t mat = 0, leaf = kNone;
  bool ok = prim_info(S, prim[g], &mat, &leaf);
  if (ok) ok = sampler_ok(S.mats[mat].brdf);
  R.cd = 0;
  R.path = p;
  if (!ok) {
    atomicOr(err, 1u);
    R.flags = 0;
    R.pre = dk(0, 0, 0);
    R.ca = (uint32_t)g;
    R.cb = estimate;
    return;
  }
  setup_point(S, ltc, dk(X[3 * g], X[3 * g + 1], X[3 * g + 2]),
              dk(N[3 * g], N[3 * g + 1], N[3 * g + 2]), dk(D[3 * g], D[3 * g + 1], D[3 * g + 2]),
              mat, leaf, need_G, R);
  R.ca = (uint32_t)g;
  R.cb = estimate;
}

// A shadow ray from the point's offset origin to the FP64 point y, with the
// 1e-5 margin of shadow() (shader.cpp:82-90) and the point's own triangle
// skipped.
__device__ inline void shadow_ray(const PointRec &P, d3 y, float4 *__restrict__ ray, size_t i) {
  const f3 xf = mk((float)P.x.x, (float)P.x.y, (float)P.x.z);
  const f3 nf = mk((float)P.n.x, (float)P.n.y, (float)P.n.z);
  const f3 org = offset_ray(xf, nf);
  const d3 sd = y - dk(org.x, org.y, org.z);
  const double dist = sqrt(dot(sd, sd));
  const d3 dir = sd * (1.0 / dist);
  ray[2 * i] = make_float4(org.x, org.y, org.z, (float)(dist - 1e-5));
  ray[2 * i + 1] = make_float4((float)dir.x, (float)dir.y, (float)dir.z, __uint_as_float(P.leaf));
}
// A shadow ray along unit w for distance dist (shadow(rr, dist) with a far
// dist): same origin and margin as shadow_ray.
__device__ inline void shadow_dir(const PointRec &P, d3 w, double dist, float4 *__restrict__ ray,
                                  size_t i) {
  const f3 xf = mk((float)P.x.x, (float)P.x.y, (float)P.x.z);
  const f3 nf = mk((float)P.n.x, (float)P.n.y, (float)P.n.z);
  const f3 org = offset_ray(xf, nf);
  ray[2 * i] = make_float4(org.x, org.y, org.z, (float)(dist - 1e-5));
  ray[2 * i + 1] = make_float4((float)w.x, (float)w.y, (float)w.z, __uint_as_float(P.leaf));
}
__device__ inline void no_ray(float4 *__restrict__ ray, size_t i) {
  ray[2 * i] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
  ray[2 * i + 1] = make_float4(0.0f, 0.0f, 1.0f, __uint_as_float(kNone));
}

// 1 / (sum_k N_k pdfVal_k(s) + 1e-20) (mis.cpp:34-95,209-220), the
// techniques in MIS::getSamples' order: the sphere sources' cones, the mesh
// lights and the triangle lights (10 samples each), the environment (10),
// the BSDF (60).  EXT = false: a scene with mesh lights only (the other
// counts are 0), compiled without their code.
template <bool EXT>
__device__ inline double mis_weight_bp(const DeviceScene &S, d3 x, d3 s, double bsdf_pdf);
template <bool EXT>
__device__ inline double mis_weight(const DeviceScene &S, const MatT<double> &m, d3 x, d3 n,
                                    d3 wo_ref, d3 wo, d3 s) {
  double f, pdf;
  brdf_pdf(m, n, wo_ref, wo, s, f, pdf);
  return mis_weight_bp<EXT>(S, x, s, pdf);
}
// The same with the BSDF's pdfEval of s given (the callers that evaluate
// the BSDF at s anyway).
template <bool EXT>
__device__ inline double mis_weight_bp(const DeviceScene &S, d3 x, d3 s, double bsdf_pdf) {
  double denom = 0;
  if (EXT)
    for (uint32_t j = 0; j < S.n_sph_src; ++j) {  // SoAnISSP (mis.cpp:35-39)
      const Cone C = sph_cone(S, j, x);
      denom += (double)CHK_MIS_LIGHT_SAMPLES * (dot(s, C.w) >= C.cos_max ? 1.0 / C.cap : 0.);
    }
  const uint32_t n_area = S.n_lights + (EXT ? S.n_tri_lights : 0u);
  for (uint32_t j = 0; j < n_area; ++j) {  // SuArISME, SuArISTR
    uint32_t tri;
    double d;
    d3 N = dk(0, 0, 0);
    double p = 0;
    if (light_isect(S, j, x, s, &tri, &d, &N)) {  // Ray r(x, sample)
      double c1 = dot(s, N);
      c1 = c1 < 0 ? -c1 : c1;
      p = ((1.0 / S.light_area[j]) * d * d) / c1;
    }
    denom += (double)CHK_MIS_LIGHT_SAMPLES * p;
  }
  if (EXT && S.env.on)  // EnMpIS (mis.cpp:84-87): grey(L) / normalizationArea
    denom += (double)CHK_MIS_LIGHT_SAMPLES *
             (grey(env_pixel_radiance(S.env, env_pixel(S.env, s.x, s.y, s.z))) * (1.0 / S.env.area));
  denom += (double)CHK_MIS_BSDF_SAMPLES * bsdf_pdf;
  return 1.0 / (denom + 1e-20);
}

// Item terms: the contribution is rad(code) * (V ? a : b) with V the
// visibility of the item's shadow ray (1 when the estimator ignores it); a
// and b are the sample's term for V = 1 and V = 0, each formed in the
// reference's order.  A point-source item contributes rad / c * a.
struct Items {
  double *a, *b, *c;
  uint32_t *tri;
  float4 *ray;
};

// Sample slots of MIS::getSamples (mis.cpp:97-191): 10 per sphere source
// (cone), mesh light and triangle light (area), 10 for the environment light
// (importance), then 60 BSDF slots; the MIS estimator adds one slot per
// point source after them.
struct Slots {
  uint32_t sph, mesh, tri, env, mis;  // the first slot after each group; mis = total
};
__device__ inline Slots mis_slots(const DeviceScene &S) {
  Slots L;
  L.sph = S.n_sph_src * CHK_MIS_LIGHT_SAMPLES;
  L.mesh = L.sph + S.n_lights * CHK_MIS_LIGHT_SAMPLES;
  L.tri = L.mesh + S.n_tri_lights * CHK_MIS_LIGHT_SAMPLES;
  L.env = L.tri + (S.env.on ? CHK_MIS_LIGHT_SAMPLES : 0u);
  L.mis = L.env + CHK_MIS_BSDF_SAMPLES;
  return L;
}

// The direction of sample slot `slot` from its Philox draw r (mis.cpp:103-191).
template <bool EXT>
__device__ inline d3 slot_dir(const DeviceScene &S, const PointRec &R, const MatT<double> &m,
                              const Slots &L, uint32_t slot, u4 r) {
  if (EXT && slot < L.sph)
    return cone_dir(sph_cone(S, slot / CHK_MIS_LIGHT_SAMPLES, R.x), ud(r.x), ud(r.y));
  if (slot < L.tri) {
    uint32_t tri;
    return normalize(light_point(S, (slot - L.sph) / CHK_MIS_LIGHT_SAMPLES, r, &tri) - R.x);
  }
  if (EXT && slot < L.env) return env_sample(S.env, ud(r.x), ud(r.y));  // EnMpIS (mis.cpp:157-169)
  const uint32_t q = slot - L.env;
  return brdf_sample(m, R.n, R.wo_ref, R.wo, q < R.nb ? -1.0 : 2.0, ud(r.y), ud(r.z));
}

template <int EST, bool EXT>
__global__ __launch_bounds__(kDBlock) void k_dl_samples(DeviceScene S, const PointRec *__restrict__ pts,
                                                        uint32_t P, uint32_t n_slots, uint32_t cvn,
                                                        uint32_t flags, uint32_t key0,
                                                        uint32_t key1, Items I) {
  const size_t i = (size_t)blockIdx.x * kDBlock + threadIdx.x;
  if (i >= (size_t)P * n_slots) return;
  const uint32_t slot = (uint32_t)(i / P), p = (uint32_t)(i % P);
  const PointRec &R = pts[p];
  I.tri[i] = kNone;
  if (!(R.flags & kActive)) {
    no_ray(I.ray, i);
    return;
  }
  const MatT<double> m = load_mat<double>(S.mats + R.mat);
  const d3 o = R.x + R.n * 1e-8;  // rr.o (lightSources.cpp:42, lightSources_CV.cpp:346,412)
  if (EST == CHK_DIRECT_ANALYTIC_CV) {
    // lightSources_CV.cpp:357-385
    const u4 r = philox(u4{R.ca, R.cb, slot, CHK_STREAM_DIRECT + R.cd}, key0, key1);
    const uint32_t j = slot / cvn;
    uint32_t tri;
    const d3 y = light_point(S, j, r, &tri);
    const d3 w = normalize(y - R.x);
    const d3 dv = y - R.x;
    const double d = sqrt(dot(dv, dv));
    const double cs = dot(w, R.n);
    const LightTri64 &q = S.light64[tri];
    double c1 = dot(w, dk(q.n[0], q.n[1], q.n[2]));  // mesh[ids[i]].t.n
    c1 = c1 < 0 ? -c1 : c1;
    if (cs < 0) {
      no_ray(I.ray, i);
      return;
    }
    double fa, pdf;
    brdf_pdf(m, R.n, R.wo_ref, R.wo, w, fa, pdf);
    // (visibility * brdfActual - brdfApproximate) * cosine * cosine1 / (d * d)
    const double fx = brdf_apx_f(m, R, w, fa);
    I.a[i] = (fa - fx) * cs * c1 / (d * d);
    I.b[i] = (0.0 * fa - fx) * cs * c1 / (d * d);
    if (EXT && S.n_spheres && sph_occluded(S, o, w, d)) {
      I.tri[i] = tri | kOccS;
      no_ray(I.ray, i);
    } else {
      I.tri[i] = tri;
      shadow_ray(R, y, I.ray, i);
    }
    return;
  }
  const Slots L = mis_slots(S);
  if (EXT && EST == CHK_DIRECT_MIS && slot >= L.mis) {
    // point source j (lightSources.cpp:47-58): radiance / (4 pi |p - x|^2)
    // * cosine * brdf, after its shadow test
    const uint32_t j = slot - L.mis;
    const chk_point_source &Q = S.points[j];
    const d3 dv = dk(Q.pos[0], Q.pos[1], Q.pos[2]) - R.x;
    const d3 w = dv * (1.0 / sqrt(dot(dv, dv)));
    const double cs = dot(w, R.n);
    if (cs < 0) {
      no_ray(I.ray, i);
      return;
    }
    double fa, pdf;
    brdf_pdf(m, R.n, R.wo_ref, R.wo, w, fa, pdf);
    I.a[i] = cs * fa;
    I.b[i] = 0.0;
    I.c[i] = 4.0 * kPi64 * dot(dv, dv);
    if (S.n_spheres && sph_occluded(S, o, w, sqrt(dot(dv, dv)))) {
      I.tri[i] = kTagPt | j | kOccS;
      no_ray(I.ray, i);
    } else {
      I.tri[i] = kTagPt | j;
      shadow_ray(R, dk(Q.pos[0], Q.pos[1], Q.pos[2]), I.ray, i);
    }
    return;
  }
  const u4 r = philox(u4{R.ca, R.cb, slot, CHK_STREAM_DIRECT + R.cd}, key0, key1);
  const d3 w = slot_dir<EXT>(S, R, m, L, slot, r);
  // the weight is needed only for a sample above the horizon that reaches a
  // source (the reference computes every weight in getSamples and skips the
  // others, lightSources.cpp:62-66, lightSources_CV.cpp:416-420)
  const double cs = dot(w, R.n);
  if (cs < 0) {
    no_ray(I.ray, i);
    return;
  }
  uint32_t code = kNone;
  double pd = 0;
  if (EST == CHK_DIRECT_MIS) {
    // lightSources.cpp:60-106: the first source in the order mesh lights,
    // sphere sources, triangle lights that the ray hits and shadow() passes.
    // Every source's geometry is in the scene, so a source farther than the
    // nearest source hit + 1e-5 is shadowed by it: the candidate is the
    // first source within 1e-5 of the nearest hit.
    double dmin = kInf64;
    for (uint32_t j = 0; j < S.n_lights + (EXT ? S.n_tri_lights : 0u); ++j) {
      uint32_t tri;
      double d;
      d3 N;
      if (light_isect(S, j, o, w, &tri, &d, &N) && d < dmin) dmin = d;
    }
    if (EXT)
      for (uint32_t j = 0; j < S.n_sph_src; ++j) {
        const double d = sph_isect(S.spheres[S.sph_src[j]], o, w);
        if (d < dmin) dmin = d;
      }
    for (uint32_t j = 0; j < S.n_lights && code == kNone; ++j) {
      uint32_t tri;
      double d;
      d3 N;
      if (light_isect(S, j, o, w, &tri, &d, &N) && d <= dmin + 1e-5) {
        code = tri;
        pd = d;
      }
    }
    for (uint32_t j = 0; EXT && j < S.n_sph_src && code == kNone; ++j) {
      const double d = sph_isect(S.spheres[S.sph_src[j]], o, w);
      if (d < kInf64 && d <= dmin + 1e-5) {
        code = kTagSph | j;
        pd = d;
      }
    }
    for (uint32_t j = S.n_lights; EXT && j < S.n_lights + S.n_tri_lights && code == kNone; ++j) {
      uint32_t tri;
      double d;
      d3 N;
      if (light_isect(S, j, o, w, &tri, &d, &N) && d <= dmin + 1e-5) {
        code = tri;
        pd = d;
      }
    }
  } else {
    // lightSources_CV.cpp:420-445: the first mesh light in order hit
    for (uint32_t j = 0; j < S.n_lights && code == kNone; ++j) {
      uint32_t tri;
      double d;
      d3 N;
      if (light_isect(S, j, o, w, &tri, &d, &N)) {
        code = tri;
        pd = d;
      }
    }
  }
  bool env = false;
  if (EXT && EST == CHK_DIRECT_MIS && code == kNone && S.env.on) {
    // lightSources.cpp:102-104: no source hit; the environment counts when
    // nothing at all is hit (shadow(rr, INF))
    code = kTagEnv | env_pixel(S.env, w.x, w.y, w.z);
    pd = kInf64;
    env = true;
  }
  if (code == kNone) {
    no_ray(I.ray, i);
    return;
  }
  double fa, pdf;
  brdf_pdf(m, R.n, R.wo_ref, R.wo, w, fa, pdf);
  const double wt = mis_weight_bp<EXT>(S, R.x, w, pdf);
  if (EST == CHK_DIRECT_MIS) {
    I.a[i] = fa * cs * wt;  // throughput
    I.b[i] = 0.0;
  } else {
    // (visibility * brdfActual - brdfApproximate) * cosine * weights[i]
    const double fx = brdf_apx_f(m, R, w, fa);
    I.a[i] = (fa - fx) * cs * wt;
    I.b[i] = (0.0 * fa - fx) * cs * wt;
  }
  if (EST == CHK_DIRECT_MIS || (flags & CHK_DIRECT_VISIBILITY)) {
    if (EXT && S.n_spheres && sph_occluded(S, o, w, pd)) {
      code |= kOccS;
      no_ray(I.ray, i);
    } else if (env) {
      shadow_dir(R, w, pd, I.ray, i);
    } else {
      shadow_ray(R, o + w * pd, I.ray, i);
    }
  } else {
    no_ray(I.ray, i);
  }
  I.tri[i] = code;
}

// getLightFromMeshSource_Analytic_MISCV as shipped, without the shadow test
// it ignores (lightSources_CV.cpp:399-450), in one pass per point: the slots
// in order, each term formed and added as the reference forms and adds it,
// then G.  No item buffers and no shadow rays; every lane of a wave works on
// the same slot, so the light / BSDF branches stay wave-uniform.  out as
// k_dl_reduce.  Bounded to 3 waves per SIMD (168 VGPRs; 33 registers of
// loop-invariant point data spill to scratch): 19.2 ms for the 1080p
// shadeDirectOnly frame against 22.7 at 2 waves (205 VGPRs) and 19.0 at 4
// (284 bytes of spill).
template <bool EXT>
__global__ __launch_bounds__(kDBlock, 3) void k_dl_miscv_fused(DeviceScene S,
                                                            const PointRec *__restrict__ pts,
                                                            uint32_t P, uint32_t n_slots,
                                                            uint32_t key0, uint32_t key1,
                                                            bool colour, double *__restrict__ out) {
  const uint32_t p = blockIdx.x * kDBlock + threadIdx.x;
  if (p >= P) return;
  const PointRec &R = pts[p];
  d3 v = R.pre;
  if (R.flags & kActive) {
    const MatT<double> m = load_mat<double>(S.mats + R.mat);
    const d3 o = R.x + R.n * 1e-8;  // rr.o (lightSources_CV.cpp:412)
    const Slots L = mis_slots(S);
    d3 sum = dk(0, 0, 0);
    for (uint32_t slot = 0; slot < n_slots; ++slot) {
      const u4 r = philox(u4{R.ca, R.cb, slot, CHK_STREAM_DIRECT + R.cd}, key0, key1);
      const d3 w = slot_dir<EXT>(S, R, m, L, slot, r);
      const double cs = dot(w, R.n);
      if (cs < 0) continue;
      uint32_t code = kNone;
      for (uint32_t j = 0; j < S.n_lights && code == kNone; ++j) {
        uint32_t tri;
        double d;
        d3 N;
        if (light_isect(S, j, o, w, &tri, &d, &N)) code = tri;
      }
      if (code == kNone) continue;
      double fa, pdf;
      brdf_pdf(m, R.n, R.wo_ref, R.wo, w, fa, pdf);
      const double wt = mis_weight_bp<EXT>(S, R.x, w, pdf);
      const double fx = brdf_apx_f(m, R, w, fa);
      sum = sum + f4d(S.light_tris[code].rad) * ((1.0 * fa - fx) * cs * wt);
    }
    sum = sum + R.G;
    v = sum;
    if (colour) {
      const chk_material &mt = S.mats[R.mat];
      v = dk(mt.color[0] * sum.x, mt.color[1] * sum.y, mt.color[2] * sum.z);
    }
  }
  out[3 * (size_t)p] = v.x;
  out[3 * (size_t)p + 1] = v.y;
  out[3 * (size_t)p + 2] = v.z;
}

// getLightFromMeshSource_Analytic_MISCV as shipped, without the shadow test
// it ignores (lightSources_CV.cpp:399-450), in one pass per point and no
// item buffers or shadow rays.  One wave per shading point: lane l takes
// slots l, l + 64, l + 128, ..., forms each term as the reference forms it
// and adds its own terms in slot order; the 64 lane sums are then combined
// by a fixed butterfly (the same association in every lane) and G is added.
// The point's record and material are wave-uniform (scalar loads).  A chunk
// of P points is P waves here against P / 64 in k_dl_miscv_fused, which is
// what a small chunk needs to fill the GPU; for large chunks the lane
// kernel is faster (its slots, and so its branches, are wave-uniform; here
// a wave's lanes mix light and BSDF slots, and its third pass holds 12 of
// 140 slots).  The sum differs from the reference's sequential one by FP64
// rounding only, and is the same on every run.  out as k_dl_reduce.
constexpr int kMiscvWaves = 4;  // points (waves) per block
constexpr uint32_t kMiscvWavePoints = 600000;  // chunks below this size use it
template <bool EXT>
__global__ __launch_bounds__(64 * kMiscvWaves, 3) void k_dl_miscv_wave(
    DeviceScene S, const PointRec *__restrict__ pts, uint32_t P, uint32_t n_slots, uint32_t key0,
    uint32_t key1, bool colour, double *__restrict__ out) {
  const uint32_t p =
      __builtin_amdgcn_readfirstlane(blockIdx.x * kMiscvWaves + (threadIdx.x >> 6));
  const uint32_t lane = threadIdx.x & 63u;
  if (p >= P) return;  // wave-uniform
  const PointRec &R = pts[p];
  d3 v = R.pre;
  if (R.flags & kActive) {
    const MatT<double> m = load_mat<double>(S.mats + R.mat);
    const d3 o = R.x + R.n * 1e-8;  // rr.o (lightSources_CV.cpp:412)
    const Slots L = mis_slots(S);
    d3 sum = dk(0, 0, 0);
    for (uint32_t slot = lane; slot < n_slots; slot += 64) {
      const u4 r = philox(u4{R.ca, R.cb, slot, CHK_STREAM_DIRECT + R.cd}, key0, key1);
      const d3 w = slot_dir<EXT>(S, R, m, L, slot, r);
      const double cs = dot(w, R.n);
      if (cs < 0) continue;
      uint32_t code = kNone;
      for (uint32_t j = 0; j < S.n_lights && code == kNone; ++j) {
        uint32_t tri;
        double d;
        d3 N;
        if (light_isect(S, j, o, w, &tri, &d, &N)) code = tri;
      }
      if (code == kNone) continue;
      double fa, pdf;
      brdf_pdf(m, R.n, R.wo_ref, R.wo, w, fa, pdf);
      const double wt = mis_weight_bp<EXT>(S, R.x, w, pdf);
      const double fx = brdf_apx_f(m, R, w, fa);
      sum = sum + f4d(S.light_tris[code].rad) * ((1.0 * fa - fx) * cs * wt);
    }
#pragma unroll
    for (int k = 1; k < 64; k <<= 1) {
      sum.x += __shfl_xor(sum.x, k);
      sum.y += __shfl_xor(sum.y, k);
      sum.z += __shfl_xor(sum.z, k);
    }
    sum = sum + R.G;
    v = sum;
    if (colour) {
      const chk_material &mt = S.mats[R.mat];
      v = dk(mt.color[0] * sum.x, mt.color[1] * sum.y, mt.color[2] * sum.z);
    }
  }
  if (lane == 0) {
    out[3 * (size_t)p] = v.x;
    out[3 * (size_t)p + 1] = v.y;
    out[3 * (size_t)p + 2] = v.z;
  }
}

// Per point: the reference's sums in its order.  out: k x 3 doubles (the
// estimator's value) or, with colour, c * value for active points and the
// point's pre value for the others.
template <int EST>
__global__ __launch_bounds__(kDBlock) void k_dl_reduce(DeviceScene S, const PointRec *__restrict__ pts,
                                                       uint32_t P, uint32_t n_slots, uint32_t cvn,
                                                       bool use_vis, const Items I,
                                                       const uint8_t *__restrict__ occ,
                                                       bool colour, double *__restrict__ out) {
  const uint32_t p = blockIdx.x * kDBlock + threadIdx.x;
  if (p >= P) return;
  const PointRec &R = pts[p];
  d3 v = R.pre;
  if (R.flags & kActive) {
    d3 sum = dk(0, 0, 0);
    if (EST == CHK_DIRECT_ANALYTIC_CV) {
      for (uint32_t j = 0; j < S.n_lights; ++j) {
        d3 s = dk(0, 0, 0);
        for (uint32_t q = 0; q < cvn; ++q) {
          const size_t i = (size_t)(j * cvn + q) * P + p;
          const uint32_t tri = I.tri[i];
          if (tri == kNone) continue;
          const double V = (occ[i] || (tri & kOccS)) ? 0.0 : 1.0;
          s = s + f4d(S.light_tris[tri & ~kOccS].rad) * (V != 0.0 ? I.a[i] : I.b[i]);
        }
        sum = sum + s * (S.light_area[j] / cvn);
      }
      sum = sum + R.G;
    } else if (EST != CHK_DIRECT_ANALYTIC) {
      // MIS: the point sources first (lightSources.cpp:47-58), then the samples
      const uint32_t first = EST == CHK_DIRECT_MIS ? mis_slots(S).mis : 0u;
      for (uint32_t k = 0; k < n_slots; ++k) {
        const uint32_t slot = first + k < n_slots ? first + k : first + k - n_slots;
        const size_t i = (size_t)slot * P + p;
        const uint32_t code = I.tri[i];
        if (code == kNone) continue;
        const double V = ((use_vis && occ[i]) || (code & kOccS)) ? 0.0 : 1.0;
        const d3 rad = code_radiance(S, code & ~kOccS);
        if (EST == CHK_DIRECT_MIS) {
          if (V == 0.0) continue;
          if ((code & kTagMask) == kTagPt)
            sum = sum + dk(rad.x / I.c[i], rad.y / I.c[i], rad.z / I.c[i]) * I.a[i];
          else
            sum = sum + rad * I.a[i];
        } else {
          sum = sum + rad * (V != 0.0 ? I.a[i] : I.b[i]);
        }
      }
      if (EST == CHK_DIRECT_MISCV) sum = sum + R.G;
    } else {
      sum = R.G;
    }
    v = sum;
    if (colour) {
      const chk_material &mt = S.mats[R.mat];
      v = dk(mt.color[0] * sum.x, mt.color[1] * sum.y, mt.color[2] * sum.z);
    }
  }
  out[3 * (size_t)p] = v.x;
  out[3 * (size_t)p + 1] = v.y;
  out[3 * (size_t)p + 2] = v.z;
}

// MIS weights of caller samples.
__global__ __launch_bounds__(kDBlock) void k_mis_weights(
    DeviceScene S, const double *__restrict__ X, const double *__restrict__ N,
    const double *__restrict__ WR, const double *__restrict__ WO, const uint32_t *__restrict__ prim,
    uint32_t k, const double *__restrict__ smp, uint32_t n_s, double *__restrict__ wout,
    uint32_t *__restrict__ err) {
  const size_t i = (size_t)blockIdx.x * kDBlock + threadIdx.x;
  if (i >= (size_t)k * n_s) return;
  const size_t p = i / n_s;
  uint32_t mat, leaf;
  if (!prim_info(S, prim[p], &mat, &leaf)) {
    atomicOr(err, 1u);
    wout[i] = 0.0;
    return;
  }
  const MatT<double> m = load_mat<double>(S.mats + mat);
  const d3 x = dk(X[3 * p], X[3 * p + 1], X[3 * p + 2]);
  const d3 n = dk(N[3 * p], N[3 * p + 1], N[3 * p + 2]);
  const d3 wr = dk(WR[3 * p], WR[3 * p + 1], WR[3 * p + 2]);
  const d3 wo = dk(WO[3 * p], WO[3 * p + 1], WO[3 * p + 2]);
  const d3 s = dk(smp[3 * i], smp[3 * i + 1], smp[3 * i + 2]);
  wout[i] = mis_weight<true>(S, m, x, n, wr, wo, s);
}

// ------------------------------------------------------------ render pieces
struct CamD {
  double pos[3], dir[3], cx[3], cy[3];
  uint32_t W, H;
};

// Camera ray of item it (pixel-major, spp samples per pixel): main.cpp:72-79
// in FP64, FP32 copy for the traversal.
__global__ __launch_bounds__(kDBlock) void k_dr_camera(CamD cam, uint32_t pix0, uint32_t n_items,
                                                       uint32_t spp, uint32_t s0, uint32_t jitter,
                                                       uint32_t key0, uint32_t key1,
                                                       float4 *__restrict__ rays,
                                                       double *__restrict__ dirs) {
  const uint32_t it = blockIdx.x * kDBlock + threadIdx.x;
  if (it >= n_items) return;
  const uint32_t idx = pix0 + it / spp, s = s0 + it % spp;
  const uint32_t row = idx / cam.W, x = idx % cam.W;
  const uint32_t y = cam.H - 1u - row;
  double jx = 0.34, jy = 0.20;  // main.cpp:23-24, k = 0
  if (jitter == CHK_JITTER_BOX) {
    const u4 rr = philox(u4{idx, s, (0u << 2) | CHK_STREAM_CAM, 0u}, key0, key1);
    jx = u01(rr.x);
    jy = u01(rr.y);
  }
  const double fx = ((double)x + jx) / cam.W - .5, fy = ((double)y + jy) / cam.H - .5;
  d3 d = dk(cam.cx[0], cam.cx[1], cam.cx[2]) * fx + dk(cam.cy[0], cam.cy[1], cam.cy[2]) * fy +
         dk(cam.dir[0], cam.dir[1], cam.dir[2]);
  d = normalize(d);
  rays[2 * (size_t)it] = make_float4((float)cam.pos[0], (float)cam.pos[1], (float)cam.pos[2], kInf);
  rays[2 * (size_t)it + 1] = make_float4((float)d.x, (float)d.y, (float)d.z, __uint_as_float(kNone));
  dirs[3 * (size_t)it] = d.x;
  dirs[3 * (size_t)it + 1] = d.y;
  dirs[3 * (size_t)it + 2] = d.z;
}

// shadeDirectOnly (shader.cpp:339-363) up to the estimator: a miss gives 0,
// an emitter its radiance.maxnorm() (materialTypes.h:34-38; or its radiance),
// anything else becomes a shading point.
__global__ __launch_bounds__(kDBlock) void k_dr_points(DeviceScene S, const double *__restrict__ ltc,
                                                       CamD cam, uint32_t pix0, uint32_t n_items,
                                                       uint32_t spp, uint32_t s0, bool physical,
                                                       bool need_G, const float4 *__restrict__ hits,
                                                       const double *__restrict__ dirs,
                                                       PointRec *__restrict__ pts) {
  const uint32_t it = blockIdx.x * kDBlock + threadIdx.x;
  if (it >= n_items) return;
  PointRec &R = pts[it];
  R.ca = pix0 + it / spp;
  R.cb = s0 + it % spp;
  R.cd = 0;
  R.path = it;
  R.flags = 0;
  R.pre = dk(0, 0, 0);
  const float4 h = hits[it];
  if (missed(S, h)) return;
  const uint32_t mat = hit_mat(S, h);
  const chk_material &mt = S.mats[mat];
  if (mt.light != CHK_LIGHT_NONE) {
    d3 r = dk(mt.radiance[0], mt.radiance[1], mt.radiance[2]);
    if (!physical) {
      double mx = r.x;
      if (mx < r.y) mx = r.y;
      if (mx < r.z) mx = r.z;
      r = dk(r.x / mx, r.y / mx, r.z / mx);  // Vec::maxnorm (mathPrimitives.cpp:24-29)
    }
    R.pre = r;
    return;
  }
  if (!sampler_ok(mt.brdf)) return;  // rejected on the host before launch
  const d3 d = dk(dirs[3 * (size_t)it], dirs[3 * (size_t)it + 1], dirs[3 * (size_t)it + 2]);
  const d3 o = dk(cam.pos[0], cam.pos[1], cam.pos[2]);
  const d3 x = o + d * hit_t(h);
  setup_point(S, ltc, x, hit_normal(S, h, x, d), d, mat, hit_leaf(h), need_G, R);
}

// Per-pixel FP64 sums of the chunk's items in sample order (main.cpp:62,83).
__global__ __launch_bounds__(kDBlock) void k_dr_accum(const double *__restrict__ val, uint32_t pix0,
                                                      uint32_t n_pix, uint32_t spp,
                                                      double *__restrict__ accum,
                                                      double *__restrict__ sumsq) {
  const uint32_t q = blockIdx.x * kDBlock + threadIdx.x;
  if (q >= n_pix) return;
  double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
  for (uint32_t s = 0; s < spp; ++s) {
    const double *v = val + 3 * ((size_t)q * spp + s);
    for (int c = 0; c < 3; ++c) {
      a[c] += v[c];
      b[c] += v[c] * v[c];
    }
  }
  const size_t o = 3 * ((size_t)pix0 + q);
  for (int c = 0; c < 3; ++c) {
    accum[o + c] += a[c];
    if (sumsq) sumsq[o + c] += b[c];
  }
}

// ------------------------------------------------------------ shaders
// shadeExplicit and shadeImplicit (shader.cpp:111-207,276-337) as path
// wavefronts: one path per (pixel, sample) of the chunk.
enum PathState : uint32_t { kDead = 0, kReady = 1, kTracing = 2 };
struct PathSt {
  d3 T, col;     // throughput (prod of TP / (1 - q), or shadeImplicit's product), colour
  d3 o, d;       // FP64 ray of the pending trace
  float4 hit;    // the accepted hit of the next vertex (t, leaf)
  uint32_t state, tries, pk, pad;
};

constexpr double kRR = 0.3;  // terminationProbability (shader.cpp:112)

__device__ inline u4 shade_rng(uint32_t pix, uint32_t smp, uint32_t w, uint32_t k0, uint32_t k1) {
  return philox(u4{pix, smp, w, CHK_STREAM_SHADE}, k0, k1);
}
// getHemiDirectionSamples' distribution (uniform, domainSampler.cpp:88-99), 1/pdf = 2 pi
__device__ inline d3 hemi_dir(d3 n, double e1, double e2) {
  const double ct = e1, st = sqrt(fmax(0.0, 1.0 - e1 * e1));
  double sp, cp;
  sincos(2.0 * kPi64 * e2, &sp, &cp);
  d3 t, b;
  onb(n, t, b);
  return t * (st * cp) + b * (st * sp) + n * ct;
}
// getCosineDirectionSamples' distribution (domainSampler.cpp:154-189), 1/pdf = pi
__device__ inline d3 cosine_dir(d3 n, double e1, double e2) {
  const double st = sqrt(e1), ct = sqrt(1.0 - e1);
  double sp, cp;
  sincos(2.0 * kPi64 * e2, &sp, &cp);
  d3 t, b;
  onb(n, t, b);
  return t * (st * cp) + b * (st * sp) + n * ct;
}
__device__ inline void put_ray(float4 *__restrict__ rays, uint32_t p, d3 o, d3 d, uint32_t skip) {
  rays[2 * (size_t)p] = make_float4((float)o.x, (float)o.y, (float)o.z, kInf);
  rays[2 * (size_t)p + 1] = make_float4((float)d.x, (float)d.y, (float)d.z, __uint_as_float(skip));
}
__device__ inline void kill_ray(float4 *__restrict__ rays, uint32_t p) {
  rays[2 * (size_t)p].w = -1.0f;
}
__device__ inline bool is_emitter(const DeviceScene &S, float4 h, uint32_t *mat) {
  *mat = hit_mat(S, h);
  return S.mats[*mat].light != CHK_LIGHT_NONE;
}
// an emitter's value as the shaders return it (materialTypes.h:34-38)
__device__ inline d3 emitter_value(const chk_material &mt, bool physical) {
  d3 r = dk(mt.radiance[0], mt.radiance[1], mt.radiance[2]);
  if (physical) return r;
  double mx = r.x;
  if (mx < r.y) mx = r.y;
  if (mx < r.z) mx = r.z;
  return dk(r.x / mx, r.y / mx, r.z / mx);
}

// Paths start at the camera: T = 1, col = 0, tracing the camera ray.
__global__ __launch_bounds__(kDBlock) void k_sh_init(CamD cam, uint32_t P, const double *__restrict__ dirs,
                                                     PathSt *__restrict__ st) {
  const uint32_t p = blockIdx.x * kDBlock + threadIdx.x;
  if (p >= P) return;
  PathSt &s = st[p];
  s.T = dk(1, 1, 1);
  s.col = dk(0, 0, 0);
  s.o = dk(cam.pos[0], cam.pos[1], cam.pos[2]);
  s.d = dk(dirs[3 * (size_t)p], dirs[3 * (size_t)p + 1], dirs[3 * (size_t)p + 2]);
  s.state = kTracing;
  s.tries = 0;
  s.pk = 0;
}

// The sphere half of the shaders' closest hit (shader.cpp:128-136,344-351):
// after the triangle trace, the nearest sphere replaces the triangle hit when
// tS < tT (a tie keeps the triangle).  The FP64 ray is the path's (st) or the
// camera's (dirs).  Rays with a negative tmax are not traced.
__global__ __launch_bounds__(kDBlock) void k_sph_closest(DeviceScene S, uint32_t n, CamD cam,
                                                         const double *__restrict__ dirs,
                                                         const PathSt *__restrict__ st,
                                                         const float4 *__restrict__ rays,
                                                         float4 *__restrict__ hits) {
  const uint32_t p = blockIdx.x * kDBlock + threadIdx.x;
  if (p >= n || !(rays[2 * (size_t)p].w > 0.0f)) return;
  d3 o, d;
  if (st) {
    o = st[p].o;
    d = st[p].d;
  } else {
    o = dk(cam.pos[0], cam.pos[1], cam.pos[2]);
    d = dk(dirs[3 * (size_t)p], dirs[3 * (size_t)p + 1], dirs[3 * (size_t)p + 2]);
  }
  const float4 h = hits[p];
  const double tT = missed(S, h) ? kInf64 : (double)h.x;
  double tS;
  const uint32_t id = sph_closest(S, o, d, &tS);
  if (id != kNone && tS < tT)
    hits[p] = make_float4((float)tS, __uint_as_float(kHitSph | id),
                          __int_as_float(__double2loint(tS)), __int_as_float(__double2hiint(tS)));
}

// shadeExplicit's first hit (shader.cpp:128-140): a miss is 0, an emitter
// its value, anything else the first vertex.
__global__ __launch_bounds__(kDBlock) void k_ex_camera_hit(DeviceScene S, uint32_t P, bool physical,
                                                           const float4 *__restrict__ hits,
                                                           PathSt *__restrict__ st,
                                                           float4 *__restrict__ rays) {
  const uint32_t p = blockIdx.x * kDBlock + threadIdx.x;
  if (p >= P) return;
  PathSt &s = st[p];
  kill_ray(rays, p);  // traced again only once a continuation is drawn
  const float4 h = hits[p];
  uint32_t mat;
  if (missed(S, h)) {
    s.state = kDead;
  } else if (is_emitter(S, h, &mat)) {
    s.col = emitter_value(S.mats[mat], physical);
    s.state = kDead;
  } else {
    s.hit = h;
    s.state = kReady;
  }
}

// Every path at a new vertex becomes a shading point (appended, one atomic
// per wave; the order of points does not change any path's result).
__global__ __launch_bounds__(kDBlock) void k_ex_gather(DeviceScene S, const double *__restrict__ ltc,
                                                       uint32_t P, uint32_t pix0, uint32_t spp,
                                                       uint32_t s0, uint32_t vertex, bool need_G,
                                                       PathSt *__restrict__ st,
                                                       PointRec *__restrict__ pts,
                                                       uint32_t *__restrict__ count) {
  const uint32_t p = blockIdx.x * kDBlock + threadIdx.x;
  const bool act = p < P && st[p].state == kReady;
  const unsigned long long b = __ballot(act);
  if (b == 0) return;
  const uint32_t lane = __lane_id();
  const uint32_t first = (uint32_t)(__ffsll((long long)b) - 1);
  uint32_t base = 0;
  if (lane == first) base = atomicAdd(count, (uint32_t)__popcll(b));
  base = __shfl(base, (int)first);
  if (!act) return;
  const uint32_t k = base + (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
  PathSt &s = st[p];
  PointRec &R = pts[k];
  R.ca = pix0 + p / spp;
  R.cb = s0 + p % spp;
  R.cd = vertex;
  R.path = p;
  const d3 x = s.o + s.d * hit_t(s.hit);
  setup_point(S, ltc, x, hit_normal(S, s.hit, x, s.d), s.d, hit_mat(S, s.hit), hit_leaf(s.hit),
              need_G, R);
  s.pk = k;
}

// After shadeDI at this vertex (val[k] = c * estimate): add it, play
// Russian roulette, and draw the first continuation direction
// (shader.cpp:146-160).
__global__ __launch_bounds__(kDBlock) void k_ex_cont(uint32_t n, uint32_t vertex, uint32_t key0,
                                                     uint32_t key1, const PointRec *__restrict__ pts,
                                                     const double *__restrict__ val,
                                                     PathSt *__restrict__ st,
                                                     float4 *__restrict__ rays) {
  const uint32_t k = blockIdx.x * kDBlock + threadIdx.x;
  if (k >= n) return;
  const PointRec &R = pts[k];
  PathSt &s = st[R.path];
  const d3 di = dk(val[3 * (size_t)k], val[3 * (size_t)k + 1], val[3 * (size_t)k + 2]);
  s.col = s.col + s.T * di;
  const double e = ud(shade_rng(R.ca, R.cb, (vertex << 8) | 0xFFu, key0, key1).x);
  if (e < kRR || vertex + 1 >= 100) {
    s.state = kDead;
    kill_ray(rays, R.path);
    return;
  }
  const u4 r = shade_rng(R.ca, R.cb, vertex << 8, key0, key1);
  const d3 w = hemi_dir(R.n, ud(r.x), ud(r.y));
  s.o = R.x + R.n * 1e-6;  // secondaryRay.o (shader.cpp:154)
  s.d = w;
  s.tries = 0;
  s.state = kTracing;
  put_ray(rays, R.path, s.o, w, R.leaf);
}

// The continuation's trace came back: a miss ends the path (TP = 0), an
// emitter draws again, anything else is accepted with
// TP = c * (2 pi cos brdf) and T *= TP / (1 - q) (shader.cpp:162-201).
__global__ __launch_bounds__(kDBlock) void k_ex_retry(DeviceScene S, uint32_t P, uint32_t pix0,
                                                      uint32_t spp, uint32_t s0, uint32_t vertex,
                                                      uint32_t key0, uint32_t key1,
                                                      const float4 *__restrict__ hits,
                                                      const PointRec *__restrict__ pts,
                                                      PathSt *__restrict__ st,
                                                      float4 *__restrict__ rays,
                                                      uint32_t *__restrict__ count) {
  const uint32_t p = blockIdx.x * kDBlock + threadIdx.x;
  bool again = false;
  if (p < P && st[p].state == kTracing) {
    PathSt &s = st[p];
    const float4 h = hits[p];
    uint32_t mat;
    if (missed(S, h)) {
      s.state = kDead;
      kill_ray(rays, p);
    } else if (is_emitter(S, h, &mat)) {
      if (++s.tries >= CHK_SHADE_MAX_TRIES) {
        s.state = kDead;
        kill_ray(rays, p);
      } else {
        const PointRec &R = pts[s.pk];
        const u4 r = shade_rng(pix0 + p / spp, s0 + p % spp, (vertex << 8) | s.tries, key0, key1);
        s.d = hemi_dir(R.n, ud(r.x), ud(r.y));
        put_ray(rays, p, s.o, s.d, R.leaf);
        again = true;
      }
    } else {
      const PointRec &R = pts[s.pk];
      const MatT<double> m = load_mat<double>(S.mats + R.mat);
      const double cosine = dot(s.d, R.n);
      double f, pdf;
      brdf_pdf(m, R.n, R.wo_ref, R.wo, s.d, f, pdf);
      const double product = 2.0 * kPi64 * cosine * f;  // pdf * cosine * brdf, pdf = 2 pi
      const d3 TP = m.c * product;
      s.T = dk(s.T.x * TP.x / (1 - kRR), s.T.y * TP.y / (1 - kRR), s.T.z * TP.z / (1 - kRR));
      s.hit = h;
      s.state = kReady;
      kill_ray(rays, p);
    }
  }
  const unsigned long long b = __ballot(again);
  if (b && __lane_id() == (uint32_t)(__ffsll((long long)b) - 1))
    atomicAdd(count, (uint32_t)__popcll(b));
}

// One vertex of shadeImplicit (shader.cpp:285-333): a miss is 0, an
// emitter ends the path with radiance.maxnorm() times the product, anything
// else draws a cosine direction and multiplies the product by
// c * (pi * brdf).  After the fourth vertex the path is 0.
__global__ __launch_bounds__(kDBlock) void k_im_step(DeviceScene S, uint32_t P, uint32_t pix0,
                                                     uint32_t spp, uint32_t s0, uint32_t vertex,
                                                     bool physical, uint32_t key0, uint32_t key1,
                                                     const float4 *__restrict__ hits,
                                                     PathSt *__restrict__ st,
                                                     float4 *__restrict__ rays) {
  const uint32_t p = blockIdx.x * kDBlock + threadIdx.x;
  if (p >= P) return;
  PathSt &s = st[p];
  if (s.state != kTracing) return;
  const float4 h = hits[p];
  uint32_t mat;
  if (missed(S, h)) {
    s.col = dk(0, 0, 0);
    s.state = kDead;
    kill_ray(rays, p);
    return;
  }
  if (is_emitter(S, h, &mat)) {
    s.col = emitter_value(S.mats[mat], physical) * s.T;
    s.state = kDead;
    kill_ray(rays, p);
    return;
  }
  if (vertex >= 3) {
    s.state = kDead;
    kill_ray(rays, p);
    return;
  }
  const uint32_t leaf = hit_leaf(h);
  const d3 x = s.o + s.d * hit_t(h);
  const d3 n = hit_normal(S, h, x, s.d);
  const MatT<double> m = load_mat<double>(S.mats + mat);
  const d3 wo = s.d * -1.0;
  const d3 wo_ref = normalize(n * 2.0 * dot(n, wo) + s.d);
  const u4 r = shade_rng(pix0 + p / spp, s0 + p % spp, vertex << 8, key0, key1);
  const d3 w = cosine_dir(n, ud(r.x), ud(r.y));
  double f, pdf;
  brdf_pdf(m, n, wo_ref, wo, w, f, pdf);
  s.T = s.T * (kPi64 * 1.0 * f);  // product * (pdfInv * cosine * brdf), cosine = 1
  s.T = m.c * s.T;
  s.o = x + n * 1e-6;
  s.d = w;
  put_ray(rays, p, s.o, w, leaf);
}

// Per-pixel FP64 sums of the chunk's path colours in sample order.
__global__ __launch_bounds__(kDBlock) void k_sh_accum(const PathSt *__restrict__ st, uint32_t pix0,
                                                      uint32_t n_pix, uint32_t spp,
                                                      double *__restrict__ accum,
                                                      double *__restrict__ sumsq) {
  const uint32_t q = blockIdx.x * kDBlock + threadIdx.x;
  if (q >= n_pix) return;
  double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
  for (uint32_t s = 0; s < spp; ++s) {
    const d3 c = st[(size_t)q * spp + s].col;
    const double v[3] = {c.x, c.y, c.z};
    for (int j = 0; j < 3; ++j) {
      a[j] += v[j];
      b[j] += v[j] * v[j];
    }
  }
  const size_t o = 3 * ((size_t)pix0 + q);
  for (int j = 0; j < 3; ++j) {
    accum[o + j] += a[j];
    if (sumsq) sumsq[o + j] += b[j];
  }
}

// ------------------------------------------------------------ host side
inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// Slots per point.  The mesh-light estimators return 0 without a mesh light
// (lightSources_CV.cpp:324,336,400), getLightFromAllSources_MIS without any
// source (lightSources.cpp:33).
uint32_t n_slots_of(const chk_scene_s *sc, const chk_direct_params &p) {
  const DeviceScene &S = sc->d;
  const uint32_t n_mis =
      (S.n_sph_src + S.n_lights + S.n_tri_lights + (S.env.on ? 1u : 0u)) * CHK_MIS_LIGHT_SAMPLES +
      CHK_MIS_BSDF_SAMPLES;
  switch (p.estimator) {
    case CHK_DIRECT_ANALYTIC: return 0;
    case CHK_DIRECT_ANALYTIC_CV: return S.n_lights * std::min(p.cv_samples, 1000u);
    case CHK_DIRECT_MIS:
      return (S.n_lights || S.n_tri_lights || S.n_sph_src || S.n_points || S.env.on)
                 ? n_mis + S.n_points
                 : 0;
    default: return S.n_lights ? n_mis : 0;
  }
}

chk_status check_params(const chk_scene_s *sc, const chk_direct_params *p, const char *who) {
  if (!p || p->estimator > CHK_DIRECT_MISCV) {
    set_error(std::string(who) + ": null params or unknown estimator");
    return CHK_EINVAL;
  }
  if (p->estimator == CHK_DIRECT_ANALYTIC_CV && p->cv_samples == 0) {
    set_error(std::string(who) + ": ANALYTIC_CV needs cv_samples > 0");
    return CHK_EINVAL;
  }
  if (p->estimator != CHK_DIRECT_MIS && !sc->ctx->ltc_dev) {
    set_error(std::string(who) + ": the analytic term needs chk_set_ltc_table");
    return CHK_EINVAL;
  }
  return CHK_OK;
}

// Scratch of one chunk of P points with n_slots slots each (+ render buffers).
struct DWork {
  PointRec *pts;
  uint32_t *err;
  uint32_t *count;  // 64 bytes after err
  Items I;
  uint8_t *occ;
  double *val;     // P x 3
  float4 *crays;   // render: P camera rays
  float4 *chits;
  double *cdirs;   // P x 3
  PathSt *paths;   // render with a path shader: P paths
  uint32_t *ovf;   // the traversal's overflow list (TraceShape::ovf)
  uint32_t ovf_cap;
};
size_t dwork_bytes(uint32_t P, uint32_t n_slots, bool render) {
  const size_t I = (size_t)P * n_slots;
  size_t b = al256((size_t)P * sizeof(PointRec)) + 256;
  b += 3 * al256(I * 8) + al256(I * 4) + al256(I * 32) + al256(I);
  b += al256((size_t)P * 24);
  if (render)
    b += al256((size_t)P * 32) + al256((size_t)P * 16) + al256((size_t)P * 24) +
         al256((size_t)P * sizeof(PathSt));
  b += al256(ovf_words((uint32_t)std::max<size_t>(I, P)) * 4);
  return b;
}
chk_status dwork_get(chk_scene_s *sc, uint32_t P, uint32_t n_slots, bool render, DWork &W) {
  const size_t need = dwork_bytes(P, n_slots, render);
  if (sc->dwork_bytes < need) {
    if (sc->dwork) (void)hipFree(sc->dwork);
    sc->dwork = nullptr;
    sc->dwork_bytes = 0;
    if (hipMalloc(&sc->dwork, need) != hipSuccess) {
      sc->dwork = nullptr;
      set_error("direct lighting: cannot allocate " + std::to_string(need) + " bytes of scratch");
      return CHK_ENOMEM;
    }
    sc->dwork_bytes = need;
  }
  const size_t I = (size_t)P * n_slots;
  char *q = (char *)sc->dwork;
  W.pts = (PointRec *)q;
  q += al256((size_t)P * sizeof(PointRec));
  W.err = (uint32_t *)q;
  W.count = W.err + 16;
  q += 256;
  W.I.a = (double *)q;
  q += al256(I * 8);
  W.I.b = (double *)q;
  q += al256(I * 8);
  W.I.c = (double *)q;
  q += al256(I * 8);
  W.I.tri = (uint32_t *)q;
  q += al256(I * 4);
  W.I.ray = (float4 *)q;
  q += al256(I * 32);
  W.occ = (uint8_t *)q;
  q += al256(I);
  W.val = (double *)q;
  q += al256((size_t)P * 24);
  W.crays = W.chits = nullptr;
  W.cdirs = nullptr;
  W.paths = nullptr;
  if (render) {
    W.crays = (float4 *)q;
    q += al256((size_t)P * 32);
    W.chits = (float4 *)q;
    q += al256((size_t)P * 16);
    W.cdirs = (double *)q;
    q += al256((size_t)P * 24);
    W.paths = (PathSt *)q;
    q += al256((size_t)P * sizeof(PathSt));
  }
  W.ovf_cap = (uint32_t)std::max<size_t>(I, P);
  W.ovf = (uint32_t *)q;
  return CHK_OK;
}

uint32_t chunk_points(uint32_t n_slots, uint32_t total) {
  const uint64_t per = n_slots ? std::max<uint64_t>(256, kItemBudget / n_slots) : (1u << 20);
  return (uint32_t)std::min<uint64_t>(per, total);
}

template <int EST>
void launch_items(chk_scene_s *sc, const DWork &W, uint32_t P, uint32_t n_slots, uint32_t cvn,
                  const chk_direct_params &p, hipStream_t st) {
  const size_t I = (size_t)P * n_slots;
  if (!I) return;
  const dim3 g((unsigned)((I + kDBlock - 1) / kDBlock)), b(kDBlock);
  if (sc->extras)
    hipLaunchKernelGGL((k_dl_samples<EST, true>), g, b, 0, st, sc->d, W.pts, P, n_slots, cvn,
                       p.flags, p.seed_lo, p.seed_hi, W.I);
  else
    hipLaunchKernelGGL((k_dl_samples<EST, false>), g, b, 0, st, sc->d, W.pts, P, n_slots, cvn,
                       p.flags, p.seed_lo, p.seed_hi, W.I);
}

// MISCV as shipped runs fused (k_dl_miscv_wave): no per-item storage.
bool fused(const chk_direct_params &p) {
  return p.estimator == CHK_DIRECT_MISCV && !(p.flags & CHK_DIRECT_VISIBILITY);
}
// Slots that need per-item storage (0 for the fused estimator).
uint32_t item_slots(const chk_scene_s *sc, const chk_direct_params &p) {
  return fused(p) ? 0u : n_slots_of(sc, p);
}

bool uses_vis(const chk_direct_params &p) {
  return p.estimator == CHK_DIRECT_MIS || p.estimator == CHK_DIRECT_ANALYTIC_CV ||
         (p.estimator == CHK_DIRECT_MISCV && (p.flags & CHK_DIRECT_VISIBILITY));
}

// Samples, visibility and the reduction of one chunk whose points are set up.
chk_status run_items(chk_scene_s *sc, const DWork &W, uint32_t P, const chk_direct_params &p,
                     bool colour, uint64_t *shadow_rays, hipStream_t st) {
  const uint32_t n_slots = n_slots_of(sc, p);
  const uint32_t cvn = std::min(p.cv_samples, 1000u);
  if (fused(p)) {
    // Below 600,000 points a chunk cannot keep the point-per-lane kernel's
    // waves resident long enough (its last wave of points runs alone): one
    // wave per point instead (the path shaders' later vertices).  Measured
    // on the 1080p direct leg: shadeDirectOnly 19.2 ms (wave kernel
    // throughout: 29.4), shadeExplicit 38.9 ms (lane kernel throughout: 70.6).
    if (P < kMiscvWavePoints) {
      const dim3 g((P + kMiscvWaves - 1) / kMiscvWaves), b(64 * kMiscvWaves);
      if (sc->extras)
        hipLaunchKernelGGL((k_dl_miscv_wave<true>), g, b, 0, st, sc->d, W.pts, P, n_slots,
                           p.seed_lo, p.seed_hi, colour, W.val);
      else
        hipLaunchKernelGGL((k_dl_miscv_wave<false>), g, b, 0, st, sc->d, W.pts, P, n_slots,
                           p.seed_lo, p.seed_hi, colour, W.val);
    } else {
      const dim3 g((P + kDBlock - 1) / kDBlock), b(kDBlock);
      if (sc->extras)
        hipLaunchKernelGGL((k_dl_miscv_fused<true>), g, b, 0, st, sc->d, W.pts, P, n_slots,
                           p.seed_lo, p.seed_hi, colour, W.val);
      else
        hipLaunchKernelGGL((k_dl_miscv_fused<false>), g, b, 0, st, sc->d, W.pts, P, n_slots,
                           p.seed_lo, p.seed_hi, colour, W.val);
    }
    CHK_HIP(hipGetLastError());
    return CHK_OK;
  }
  switch (p.estimator) {
    case CHK_DIRECT_ANALYTIC_CV:
      launch_items<CHK_DIRECT_ANALYTIC_CV>(sc, W, P, n_slots, cvn, p, st);
      break;
    case CHK_DIRECT_MIS: launch_items<CHK_DIRECT_MIS>(sc, W, P, n_slots, cvn, p, st); break;
    case CHK_DIRECT_MISCV: launch_items<CHK_DIRECT_MISCV>(sc, W, P, n_slots, cvn, p, st); break;
    default: break;
  }
  CHK_HIP(hipGetLastError());
  const size_t I = (size_t)P * n_slots;
  const bool vis = uses_vis(p) && I > 0;
  if (vis) {
    const TraceShape shape{sc->stack_size, 128, W.ovf, W.ovf_cap, sc->shallow_stack};
    CHK_HIP(launch_any(sc, shape, W.I.ray, linear_range((uint32_t)I), W.occ, nullptr, nullptr,
                       true, false, nullptr, SlotScenes{}, st));
    if (shadow_rays) *shadow_rays += I;
  }
  const dim3 g((P + kDBlock - 1) / kDBlock), b(kDBlock);
  switch (p.estimator) {
    case CHK_DIRECT_ANALYTIC:
      hipLaunchKernelGGL(k_dl_reduce<CHK_DIRECT_ANALYTIC>, g, b, 0, st, sc->d, W.pts, P, n_slots,
                         cvn, vis, W.I, W.occ, colour, W.val);
      break;
    case CHK_DIRECT_ANALYTIC_CV:
      hipLaunchKernelGGL(k_dl_reduce<CHK_DIRECT_ANALYTIC_CV>, g, b, 0, st, sc->d, W.pts, P,
                         n_slots, cvn, vis, W.I, W.occ, colour, W.val);
      break;
    case CHK_DIRECT_MIS:
      hipLaunchKernelGGL(k_dl_reduce<CHK_DIRECT_MIS>, g, b, 0, st, sc->d, W.pts, P, n_slots, cvn,
                         vis, W.I, W.occ, colour, W.val);
      break;
    default:
      hipLaunchKernelGGL(k_dl_reduce<CHK_DIRECT_MISCV>, g, b, 0, st, sc->d, W.pts, P, n_slots,
                         cvn, vis, W.I, W.occ, colour, W.val);
      break;
  }
  CHK_HIP(hipGetLastError());
  return CHK_OK;
}

// Whether every non-emitter material has a BSDF sampler (found at scene
// creation).
bool materials_ok(const chk_scene_s *sc, std::string &why) {
  if (sc->bad_sampler_mat < 0) return true;
  why = "material " + std::to_string(sc->bad_sampler_mat) + " has no BSDF sampler (brdf " +
        std::to_string(sc->bad_sampler_brdf) + ")";
  return false;
}

}  // namespace
}  // namespace chk

using namespace chk;

extern "C" {

chk_status chk_mis_weights(chk_scene_t sc, const double *x, const double *n, const double *wo_ref,
                           const double *wo, const uint32_t *prim, uint32_t k,
                           const double *samples, uint32_t S, double *weights, void *stream) {
  CHK_TRY_BEGIN
  if (!sc) {
    set_error("chk_mis_weights: null scene");
    return CHK_EINVAL;
  }
  if (k == 0 || S == 0) return CHK_OK;
  if (!x || !n || !wo_ref || !wo || !prim || !samples || !weights) {
    set_error("chk_mis_weights: null argument");
    return CHK_EINVAL;
  }
  if ((uint64_t)k * S >= (1ull << 32)) {
    set_error("chk_mis_weights: k * S must stay below 2^32");
    return CHK_ECAPACITY;
  }
  CHK_HIP(hipSetDevice(sc->ctx->device));
  const hipStream_t st = (hipStream_t)stream;
  DWork W;
  const chk_status cs = dwork_get(sc, 1, 0, false, W);
  if (cs != CHK_OK) return cs;
  CHK_HIP(hipMemsetAsync(W.err, 0, 4, st));
  const size_t I = (size_t)k * S;
  hipLaunchKernelGGL(k_mis_weights, dim3((unsigned)((I + kDBlock - 1) / kDBlock)), dim3(kDBlock),
                     0, st, sc->d, x, n, wo_ref, wo, prim, k, samples, S, weights, W.err);
  CHK_HIP(hipGetLastError());
  uint32_t err = 0;
  CHK_HIP(hipMemcpyAsync(&err, W.err, 4, hipMemcpyDeviceToHost, st));
  CHK_HIP(hipStreamSynchronize(st));
  if (err) {
    set_error("chk_mis_weights: a prim index names no triangle or sphere");
    return CHK_EINVAL;
  }
  return CHK_OK;
  CHK_TRY_END
}

chk_status chk_direct_light(chk_scene_t sc, const double *x, const double *n, const double *d,
                            const uint32_t *prim, uint32_t k, const chk_direct_params *params,
                            double *rgb, void *stream) {
  CHK_TRY_BEGIN
  if (!sc) {
    set_error("chk_direct_light: null scene");
    return CHK_EINVAL;
  }
  chk_status cs = check_params(sc, params, "chk_direct_light");
  if (cs != CHK_OK) return cs;
  if (k == 0) return CHK_OK;
  if (!x || !n || !d || !prim || !rgb) {
    set_error("chk_direct_light: null argument");
    return CHK_EINVAL;
  }
  CHK_HIP(hipSetDevice(sc->ctx->device));
  const hipStream_t st = (hipStream_t)stream;
  const chk_direct_params p = *params;
  const uint32_t n_slots = item_slots(sc, p);
  const uint32_t Pc = chunk_points(n_slots, k);
  DWork W;
  cs = dwork_get(sc, Pc, n_slots, false, W);
  if (cs != CHK_OK) return cs;
  CHK_HIP(hipMemsetAsync(W.err, 0, 4, st));
  const bool need_G = p.estimator != CHK_DIRECT_MIS;
  for (uint32_t p0 = 0; p0 < k; p0 += Pc) {
    const uint32_t P = std::min(Pc, k - p0);
    hipLaunchKernelGGL(k_dl_setup, dim3((P + kDBlock - 1) / kDBlock), dim3(kDBlock), 0, st, sc->d,
                       (const double *)sc->ctx->ltc_dev, x, n, d, prim, p0, P, p.estimate, need_G,
                       W.pts, W.err);
    CHK_HIP(hipGetLastError());
    cs = run_items(sc, W, P, p, false, nullptr, st);
    if (cs != CHK_OK) return cs;
    CHK_HIP(hipMemcpyAsync(rgb + 3 * (size_t)p0, W.val, (size_t)P * 24, hipMemcpyDeviceToDevice,
                           st));
  }
  uint32_t err = 0;
  CHK_HIP(hipMemcpyAsync(&err, W.err, 4, hipMemcpyDeviceToHost, st));
  CHK_HIP(hipStreamSynchronize(st));
  if (err) {
    set_error("chk_direct_light: a prim index names no triangle or sphere, or names a material "
              "without a BSDF sampler (GGX, Beckmann, Phong, classic Phong)");
    return CHK_EINVAL;
  }
  return CHK_OK;
  CHK_TRY_END
}

chk_status chk_render_shade(chk_scene_t sc, const chk_camera *cam, const chk_render_params *P,
                            uint32_t shader, const chk_direct_params *direct, double *accum,
                            double *sumsq, chk_render_stats *stats, void *stream) {
  CHK_TRY_BEGIN
  if (!sc || !cam || !P || !accum) {
    set_error("chk_render_shade: null scene, camera, params or accum");
    return CHK_EINVAL;
  }
  if (shader > CHK_SHADE_IMPLICIT) {
    set_error("chk_render_shade: unknown shader");
    return CHK_EINVAL;
  }
  if (P->window[0] | P->window[1] | P->window[2] | P->window[3]) {
    set_error("chk_render_shade: params->window is for chk_render; the shaders render whole "
              "images");
    return CHK_EINVAL;
  }
  const bool implicit = shader == CHK_SHADE_IMPLICIT;
  chk_direct_params dp{};
  if (!implicit) {
    chk_status cs = check_params(sc, direct, "chk_render_shade");
    if (cs != CHK_OK) return cs;
    dp = *direct;
  }
  if (cam->width == 0 || cam->height == 0 || P->spp == 0) {
    set_error("chk_render_shade: zero width, height or spp");
    return CHK_EINVAL;
  }
  if ((uint64_t)cam->width * cam->height >= (1ull << 32)) {
    set_error("chk_render_shade: image too large");
    return CHK_ECAPACITY;
  }
  CHK_HIP(hipSetDevice(sc->ctx->device));
  std::string why;
  if (!materials_ok(sc, why)) {
    set_error("chk_render_shade: " + why);
    return CHK_EINVAL;
  }
  const hipStream_t st = (hipStream_t)stream;
  const auto t0 = std::chrono::steady_clock::now();
  // camera basis in FP64 (prepareScene.cpp:21-33)
  CamD C;
  {
    const double PI = 3.14159265359;
    double d[3] = {cam->dir[0], cam->dir[1], cam->dir[2]};
    const double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (double &v : d) v /= dl;
    const double up[3] = {cam->up[0], cam->up[1], cam->up[2]};
    double cx[3] = {d[1] * up[2] - d[2] * up[1], d[2] * up[0] - d[0] * up[2],
                    d[0] * up[1] - d[1] * up[0]};
    const double cl = std::sqrt(cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]);
    if (!(cl >= 1e-20)) {
      set_error("chk_render_shade: camera up and direction are parallel");
      return CHK_EINVAL;
    }
    const double sx = ((double)cam->width / cam->height) * std::tan(cam->fovx_deg * PI / 360.0) / cl;
    for (double &v : cx) v *= sx;
    double cy[3] = {cx[1] * d[2] - cx[2] * d[1], cx[2] * d[0] - cx[0] * d[2],
                    cx[0] * d[1] - cx[1] * d[0]};
    const double yl = std::sqrt(cy[0] * cy[0] + cy[1] * cy[1] + cy[2] * cy[2]);
    const double sy = std::tan(cam->fovy_deg * PI / 360.0) / yl;
    for (int c = 0; c < 3; ++c) {
      C.pos[c] = cam->pos[c];
      C.dir[c] = d[c];
      C.cx[c] = cx[c];
      C.cy[c] = cy[c] * sy;
    }
    C.W = cam->width;
    C.H = cam->height;
  }
  const uint32_t n_slots = implicit ? 0u : item_slots(sc, dp);
  const uint32_t n_pix = cam->width * cam->height, spp = P->spp;
  // whole pixels per chunk
  const uint32_t cap_items = chunk_points(n_slots, 0xFFFFFFFFu);
  if (spp > cap_items) {
    set_error("chk_render_shade: spp exceeds the items-per-chunk capacity (" +
              std::to_string(cap_items) + ")");
    return CHK_ECAPACITY;
  }
  const uint32_t pix_per = std::max(1u, cap_items / spp);
  const uint32_t Pc = (uint32_t)std::min<uint64_t>((uint64_t)pix_per * spp, (uint64_t)n_pix * spp);
  DWork W;
  chk_status cs = dwork_get(sc, Pc, n_slots, true, W);
  if (cs != CHK_OK) return cs;
  const bool need_G = !implicit && dp.estimator != CHK_DIRECT_MIS;
  const bool physical = (P->flags & CHK_RENDER_PHYSICAL_EMITTER) != 0;
  const TraceShape shape{sc->stack_size, 128, W.ovf, W.ovf_cap, sc->shallow_stack};
  const double *ltc = (const double *)sc->ctx->ltc_dev;
  uint64_t shadow = 0, extend = 0;
  uint32_t chunks = 0, launches = 0;
  // spheres after every triangle trace (camera rays: paths == nullptr)
  auto spheres = [&](uint32_t items, const PathSt *paths) -> hipError_t {
    if (sc->d.n_spheres == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sph_closest, dim3((items + kDBlock - 1) / kDBlock), dim3(kDBlock), 0, st,
                       sc->d, items, C, W.cdirs, paths, W.crays, W.chits);
    return hipGetLastError();
  };
  for (uint32_t pix0 = 0; pix0 < n_pix; pix0 += pix_per) {
    const uint32_t np = std::min(pix_per, n_pix - pix0);
    const uint32_t items = np * spp;
    const dim3 g((items + kDBlock - 1) / kDBlock), b(kDBlock);
    hipLaunchKernelGGL(k_dr_camera, g, b, 0, st, C, pix0, items, spp, P->sample_offset,
                       P->jitter_mode, P->seed_lo, P->seed_hi, W.crays, W.cdirs);
    CHK_HIP(hipGetLastError());
    CHK_HIP(launch_closest(sc, shape, W.crays, linear_range(items), W.chits, true, false, nullptr,
                           SlotScenes{}, st));
    CHK_HIP(spheres(items, nullptr));
    extend += items;
    ++launches;
    if (shader == CHK_SHADE_DIRECT_ONLY) {
      hipLaunchKernelGGL(k_dr_points, g, b, 0, st, sc->d, ltc, C, pix0, items, spp,
                         P->sample_offset, physical, need_G, W.chits, W.cdirs, W.pts);
      CHK_HIP(hipGetLastError());
      cs = run_items(sc, W, items, dp, true, &shadow, st);
      if (cs != CHK_OK) return cs;
      hipLaunchKernelGGL(k_dr_accum, dim3((np + kDBlock - 1) / kDBlock), dim3(kDBlock), 0, st,
                         W.val, pix0, np, spp, accum,
                         (P->flags & CHK_RENDER_SUMSQ) ? sumsq : nullptr);
      CHK_HIP(hipGetLastError());
      ++chunks;
      continue;
    }
    hipLaunchKernelGGL(k_sh_init, g, b, 0, st, C, items, W.cdirs, W.paths);
    CHK_HIP(hipGetLastError());
    if (implicit) {
      for (uint32_t v = 0; v < 4; ++v) {
        hipLaunchKernelGGL(k_im_step, g, b, 0, st, sc->d, items, pix0, spp, P->sample_offset, v,
                           physical, P->seed_lo, P->seed_hi, W.chits, W.paths, W.crays);
        CHK_HIP(hipGetLastError());
        if (v == 3) break;
        CHK_HIP(launch_closest(sc, shape, W.crays, linear_range(items), W.chits, true, false,
                               nullptr, SlotScenes{}, st));
        CHK_HIP(spheres(items, W.paths));
        ++launches;
        extend += items;  // upper bound: rays of paths still alive
      }
    } else {
      hipLaunchKernelGGL(k_ex_camera_hit, g, b, 0, st, sc->d, items, physical, W.chits, W.paths,
                         W.crays);
      CHK_HIP(hipGetLastError());
      for (uint32_t v = 0; v < 100; ++v) {
        // shadeDI at every live path's vertex
        CHK_HIP(launch_zero_u32(W.count, 1, st));
        hipLaunchKernelGGL(k_ex_gather, g, b, 0, st, sc->d, ltc, items, pix0, spp,
                           P->sample_offset, v, need_G, W.paths, W.pts, W.count);
        CHK_HIP(hipGetLastError());
        uint32_t n_act = 0;
        CHK_HIP(hipMemcpyAsync(&n_act, W.count, 4, hipMemcpyDeviceToHost, st));
        CHK_HIP(hipStreamSynchronize(st));
        if (n_act == 0) break;
        cs = run_items(sc, W, n_act, dp, true, &shadow, st);
        if (cs != CHK_OK) return cs;
        hipLaunchKernelGGL(k_ex_cont, dim3((n_act + kDBlock - 1) / kDBlock), dim3(kDBlock), 0, st,
                           n_act, v, P->seed_lo, P->seed_hi, W.pts, W.val, W.paths, W.crays);
        CHK_HIP(hipGetLastError());
        // continuation: draw again while the ray hits an emitter
        for (uint32_t t = 0; t < CHK_SHADE_MAX_TRIES; ++t) {
          CHK_HIP(launch_closest(sc, shape, W.crays, linear_range(items), W.chits, true, false,
                                 nullptr, SlotScenes{}, st));
          CHK_HIP(spheres(items, W.paths));
          ++launches;
          CHK_HIP(launch_zero_u32(W.count, 1, st));
          hipLaunchKernelGGL(k_ex_retry, g, b, 0, st, sc->d, items, pix0, spp, P->sample_offset, v,
                             P->seed_lo, P->seed_hi, W.chits, W.pts, W.paths, W.crays, W.count);
          CHK_HIP(hipGetLastError());
          uint32_t again = 0;
          CHK_HIP(hipMemcpyAsync(&again, W.count, 4, hipMemcpyDeviceToHost, st));
          CHK_HIP(hipStreamSynchronize(st));
          extend += n_act;  // upper bound: rays traced this round
          if (again == 0) break;
        }
      }
    }
    hipLaunchKernelGGL(k_sh_accum, dim3((np + kDBlock - 1) / kDBlock), dim3(kDBlock), 0, st,
                       W.paths, pix0, np, spp, accum,
                       (P->flags & CHK_RENDER_SUMSQ) ? sumsq : nullptr);
    CHK_HIP(hipGetLastError());
    ++chunks;
  }
  CHK_HIP(hipStreamSynchronize(st));
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->paths = (uint64_t)n_pix * spp;
    stats->rays_extend = extend;
    stats->rays_shadow = shadow;
    stats->chunks = chunks;
    stats->extend_launches = launches;
    stats->other_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return CHK_OK;
  CHK_TRY_END
}

chk_status chk_render_direct(chk_scene_t sc, const chk_camera *cam, const chk_render_params *P,
                             const chk_direct_params *direct, double *accum, double *sumsq,
                             chk_render_stats *stats, void *stream) {
  return chk_render_shade(sc, cam, P, CHK_SHADE_DIRECT_ONLY, direct, accum, sumsq, stats, stream);
}

}  // extern "C"
