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

#include "chk_internal.h"
#include "chk_ltc_dev.h"

namespace chk {
namespace {

constexpr double kInf64 = 1e20;  // mathPrimitives.h:5
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

// Triangle::intersect (geometryPrimitives.cpp:29-71) in FP64 with the
// triangle's constants precomputed (LightTri64): the reference's operations
// without recomputing the normal and the edge dot products per call.
__device__ inline double tri_isect64p(const LightTri64 &T, d3 o, d3 d, d3 *N) {
  const d3 n = dk(T.n[0], T.n[1], T.n[2]);
  double rn = dot(d, n);
  d3 Nf = n;
  if (rn > 0) {
    Nf = n * -1.0;
    rn = -rn;
  }
  *N = Nf;
  if (rn * rn <= 1e-10) return kInf64;
  const d3 A = dk(T.a[0], T.a[1], T.a[2]);
  const double t = dot(A - o, Nf) / rn;
  if (t < 0) return kInf64;
  const d3 v0 = (o + d * t) - A;
  const double k0 = dot(v0, dk(T.v1[0], T.v1[1], T.v1[2]));
  const double k1 = dot(v0, dk(T.v2[0], T.v2[1], T.v2[2]));
  const double gamma = (k0 * T.b1 - k1 * T.b0) / T.den;
  const double beta = (k0 - gamma * T.b1) / T.b0;  // c0 = v2.v1 = b1
  if (gamma + beta > 1) return kInf64;
  if (0 > gamma || gamma > 1) return kInf64;
  if (0 > beta || beta > 1) return kInf64;
  return t;
}

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

// BasePrimitive::brdf with the LTC approximation when it is set
// (materialTypes.cpp:9-11,76-87; geometryPrimitives.cpp:7-9).
__device__ inline double brdf_apx(const MatT<double> &m, const PointRec &P, d3 wi) {
  if (!(P.flags & kApx)) {
    double f, pdf;
    brdf_pdf(m, P.n, P.wo_ref, P.wo, wi, f, pdf);
    return f;
  }
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
  const uint32_t o = prim[g];
  uint32_t mat = 0, leaf = kNone;
  bool ok = o < S.n_tris;
  if (ok) {
    const uint2 oi = S.orig_info[o];
    leaf = oi.x;
    mat = oi.y;
    ok = sampler_ok(S.mats[mat].brdf);
  }
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
__device__ inline void no_ray(float4 *__restrict__ ray, size_t i) {
  ray[2 * i] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
  ray[2 * i + 1] = make_float4(0.0f, 0.0f, 1.0f, __uint_as_float(kNone));
}

// 1 / (sum_k N_k pdfVal_k(s) + 1e-20) (mis.cpp:34-95,209-220).
__device__ inline double mis_weight(const DeviceScene &S, const MatT<double> &m, d3 x, d3 n,
                                    d3 wo_ref, d3 wo, d3 s) {
  double denom = 0;
  for (uint32_t j = 0; j < S.n_lights; ++j) {
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
  double f, pdf;
  brdf_pdf(m, n, wo_ref, wo, s, f, pdf);
  denom += (double)CHK_MIS_BSDF_SAMPLES * pdf;
  return 1.0 / (denom + 1e-20);
}

// Item terms: the contribution is rad(tri) * ((V a - b) c) with V the
// visibility of the item's shadow ray (1 when the estimator ignores it).
struct Items {
  double *a, *b, *c;
  uint32_t *tri;
  float4 *ray;
};

template <int EST>
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
  const u4 r = philox(u4{R.ca, R.cb, slot, CHK_STREAM_DIRECT + R.cd}, key0, key1);
  if (EST == CHK_DIRECT_ANALYTIC_CV) {
    // lightSources_CV.cpp:357-385
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
    I.a[i] = fa;
    I.b[i] = brdf_apx(m, R, w);
    I.c[i] = cs * c1 / (d * d);
    I.tri[i] = tri;
    shadow_ray(R, y, I.ray, i);
    return;
  }
  // MIS::getSamples sample (mis.cpp:127-138,180-191)
  const uint32_t nls = S.n_lights * CHK_MIS_LIGHT_SAMPLES;
  d3 w;
  if (slot < nls) {
    uint32_t tri;
    w = normalize(light_point(S, slot / CHK_MIS_LIGHT_SAMPLES, r, &tri) - R.x);
  } else {
    const uint32_t q = slot - nls;
    w = brdf_sample(m, R.n, R.wo_ref, R.wo, q < R.nb ? -1.0 : 2.0, ud(r.y), ud(r.z));
  }
  const double wt = mis_weight(S, m, R.x, R.n, R.wo_ref, R.wo, w);
  const double cs = dot(w, R.n);
  if (cs < 0) {
    no_ray(I.ray, i);
    return;
  }
  const d3 o = R.x + R.n * 1e-8;  // rr.o (lightSources.cpp:42, lightSources_CV.cpp:412)
  uint32_t pick = kNone, ptri = kNone;
  double pd = 0;
  if (EST == CHK_DIRECT_MIS) {
    // lightSources.cpp:64-80: the first mesh light in order that the ray
    // hits and shadow() passes.  Every light triangle is in the scene, so a
    // light farther than the nearest light hit + 1e-5 is shadowed by it:
    // the candidate is the first light within 1e-5 of the nearest hit.
    double dmin = kInf64;
    for (uint32_t j = 0; j < S.n_lights; ++j) {
      uint32_t tri;
      double d;
      d3 N;
      if (light_isect(S, j, o, w, &tri, &d, &N) && d < dmin) dmin = d;
    }
    for (uint32_t j = 0; j < S.n_lights && pick == kNone; ++j) {
      uint32_t tri;
      double d;
      d3 N;
      if (light_isect(S, j, o, w, &tri, &d, &N) && d <= dmin + 1e-5) {
        pick = j;
        ptri = tri;
        pd = d;
      }
    }
  } else {
    // lightSources_CV.cpp:420-445: the first mesh light in order hit
    for (uint32_t j = 0; j < S.n_lights && pick == kNone; ++j) {
      uint32_t tri;
      double d;
      d3 N;
      if (light_isect(S, j, o, w, &tri, &d, &N)) {
        pick = j;
        ptri = tri;
        pd = d;
      }
    }
  }
  if (pick == kNone) {
    no_ray(I.ray, i);
    return;
  }
  double fa, pdf;
  brdf_pdf(m, R.n, R.wo_ref, R.wo, w, fa, pdf);
  if (EST == CHK_DIRECT_MIS) {
    I.a[i] = fa * cs * wt;  // throughput
    I.b[i] = 0.0;
    I.c[i] = 1.0;
  } else {
    I.a[i] = fa;
    I.b[i] = brdf_apx(m, R, w);
    I.c[i] = cs * wt;
  }
  I.tri[i] = ptri;
  if (EST == CHK_DIRECT_MIS || (flags & CHK_DIRECT_VISIBILITY))
    shadow_ray(R, o + w * pd, I.ray, i);
  else
    no_ray(I.ray, i);
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
          const double V = occ[i] ? 0.0 : 1.0;
          s = s + f4d(S.light_tris[tri].rad) * ((V * I.a[i] - I.b[i]) * I.c[i]);
        }
        sum = sum + s * (S.light_area[j] / cvn);
      }
      sum = sum + R.G;
    } else if (EST != CHK_DIRECT_ANALYTIC) {
      for (uint32_t slot = 0; slot < n_slots; ++slot) {
        const size_t i = (size_t)slot * P + p;
        const uint32_t tri = I.tri[i];
        if (tri == kNone) continue;
        const double V = (use_vis && occ[i]) ? 0.0 : 1.0;
        if (EST == CHK_DIRECT_MIS) {
          if (V == 0.0) continue;
          sum = sum + f4d(S.light_tris[tri].rad) * I.a[i];
        } else {
          sum = sum + f4d(S.light_tris[tri].rad) * ((V * I.a[i] - I.b[i]) * I.c[i]);
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
  const uint32_t o = prim[p];
  if (o >= S.n_tris) {
    atomicOr(err, 1u);
    wout[i] = 0.0;
    return;
  }
  const uint32_t mat = S.orig_info[o].y;
  const MatT<double> m = load_mat<double>(S.mats + mat);
  const d3 x = dk(X[3 * p], X[3 * p + 1], X[3 * p + 2]);
  const d3 n = dk(N[3 * p], N[3 * p + 1], N[3 * p + 2]);
  const d3 wr = dk(WR[3 * p], WR[3 * p + 1], WR[3 * p + 2]);
  const d3 wo = dk(WO[3 * p], WO[3 * p + 1], WO[3 * p + 2]);
  const d3 s = dk(smp[3 * i], smp[3 * i + 1], smp[3 * i + 2]);
  wout[i] = mis_weight(S, m, x, n, wr, wo, s);
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
  const uint32_t leaf = __float_as_uint(h.y);
  if (!(h.x < kInf) || leaf >= S.n_leaf_tris) return;
  const uint32_t mat = S.leaf_info[leaf].x;
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
  const float4 *tp = S.tris + 3 * (size_t)leaf;
  const d3 A = f4d(tp[0]), B = f4d(tp[1]), C = f4d(tp[2]);
  const d3 nn = cross(C - A, B - A);
  d3 n = nn * (1.0 / sqrt(dot(nn, nn)));
  if (dot(d, n) > 0) n = n * -1.0;  // Triangle::intersect's facing normal
  setup_point(S, ltc, o + d * (double)h.x, n, d, mat, leaf, need_G, R);
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
  const uint32_t leaf = __float_as_uint(h.y);
  *mat = S.leaf_info[leaf].x;
  return S.mats[*mat].light != CHK_LIGHT_NONE;
}
__device__ inline bool missed(const DeviceScene &S, float4 h) {
  return !(h.x < kInf) || __float_as_uint(h.y) >= S.n_leaf_tris;
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
// the facing unit normal of leaf triangle `leaf` for direction d
__device__ inline d3 facing_normal(const DeviceScene &S, uint32_t leaf, d3 d) {
  const float4 *tp = S.tris + 3 * (size_t)leaf;
  const d3 A = f4d(tp[0]), B = f4d(tp[1]), C = f4d(tp[2]);
  const d3 nn = cross(C - A, B - A);
  d3 n = nn * (1.0 / sqrt(dot(nn, nn)));
  if (dot(d, n) > 0) n = n * -1.0;
  return n;
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
  const uint32_t leaf = __float_as_uint(s.hit.y);
  PointRec &R = pts[k];
  R.ca = pix0 + p / spp;
  R.cb = s0 + p % spp;
  R.cd = vertex;
  R.path = p;
  setup_point(S, ltc, s.o + s.d * (double)s.hit.x, facing_normal(S, leaf, s.d), s.d,
              S.leaf_info[leaf].x, leaf, need_G, R);
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
  const uint32_t leaf = __float_as_uint(h.y);
  const d3 x = s.o + s.d * (double)h.x;
  const d3 n = facing_normal(S, leaf, s.d);
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

uint32_t n_slots_of(const chk_scene_s *sc, const chk_direct_params &p) {
  switch (p.estimator) {
    case CHK_DIRECT_ANALYTIC: return 0;
    case CHK_DIRECT_ANALYTIC_CV: return sc->d.n_lights * std::min(p.cv_samples, 1000u);
    default: return sc->d.n_lights * CHK_MIS_LIGHT_SAMPLES + CHK_MIS_BSDF_SAMPLES;
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
};
size_t dwork_bytes(uint32_t P, uint32_t n_slots, bool render) {
  const size_t I = (size_t)P * n_slots;
  size_t b = al256((size_t)P * sizeof(PointRec)) + 256;
  b += 3 * al256(I * 8) + al256(I * 4) + al256(I * 32) + al256(I);
  b += al256((size_t)P * 24);
  if (render)
    b += al256((size_t)P * 32) + al256((size_t)P * 16) + al256((size_t)P * 24) +
         al256((size_t)P * sizeof(PathSt));
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
  }
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
  if (I) {
    hipLaunchKernelGGL(k_dl_samples<EST>, dim3((unsigned)((I + kDBlock - 1) / kDBlock)),
                       dim3(kDBlock), 0, st, sc->d, W.pts, P, n_slots, cvn, p.flags, p.seed_lo,
                       p.seed_hi, W.I);
  }
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
    const TraceShape shape{sc->stack_size, 128};
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

// Whether every non-emitter material has a BSDF sampler.
bool materials_ok(const chk_scene_s *sc, std::string &why) {
  std::vector<chk_material> m(sc->n_mats);
  if (m.empty()) return true;
  if (hipMemcpy(m.data(), sc->d.mats, m.size() * sizeof(chk_material), hipMemcpyDeviceToHost) !=
      hipSuccess) {
    why = "cannot read the materials";
    return false;
  }
  for (size_t i = 0; i < m.size(); ++i) {
    const uint32_t b = m[i].brdf;
    const bool ok = b == CHK_BRDF_CLASSIC_PHONG || b == CHK_BRDF_GGX || b == CHK_BRDF_BECKMANN ||
                    b == CHK_BRDF_PHONG;
    if (m[i].light == CHK_LIGHT_NONE && !ok) {
      why = "material " + std::to_string(i) + " has no BSDF sampler (brdf " + std::to_string(b) + ")";
      return false;
    }
  }
  return true;
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
    set_error("chk_mis_weights: a prim index is not below the scene's triangle count");
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
  const uint32_t n_slots = n_slots_of(sc, p);
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
    set_error("chk_direct_light: a prim index is out of range or names a material without a "
              "BSDF sampler (GGX, Beckmann, Phong, classic Phong)");
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
  const uint32_t n_slots = implicit ? 0u : n_slots_of(sc, dp);
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
  const TraceShape shape{sc->stack_size, 128};
  const double *ltc = (const double *)sc->ctx->ltc_dev;
  uint64_t shadow = 0, extend = 0;
  uint32_t chunks = 0, launches = 0;
  for (uint32_t pix0 = 0; pix0 < n_pix; pix0 += pix_per) {
    const uint32_t np = std::min(pix_per, n_pix - pix0);
    const uint32_t items = np * spp;
    const dim3 g((items + kDBlock - 1) / kDBlock), b(kDBlock);
    hipLaunchKernelGGL(k_dr_camera, g, b, 0, st, C, pix0, items, spp, P->sample_offset,
                       P->jitter_mode, P->seed_lo, P->seed_hi, W.crays, W.cdirs);
    CHK_HIP(hipGetLastError());
    CHK_HIP(launch_closest(sc, shape, W.crays, linear_range(items), W.chits, true, false, nullptr,
                           SlotScenes{}, st));
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
