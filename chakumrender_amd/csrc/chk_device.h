// Device-side building blocks of the per-sample radiance path (gfx950).
// Material terms are templates: FP32 in the fused render kernels, FP64 in
// the per-sample entry points (the reference's own precision).
//
// Reference formulas restated here (FP64 in the reference):
//   Triangle::intersect          geometryPrimitives.cpp:29-71  (replaced by the
//                                watertight test, same accept rules)
//   MaterialType::brdf/pdfEval   materialTypes.cpp:9-50,89-116,156-258
//   BRDF samplers                domainSampler.cpp:154-314
//   MeshLight::getSamples point  lightSources.cpp:812-824
// FP32 rearrangements keep the reference's value but avoid cancellation:
// tan^2 is |h x n|^2 / (h.n)^2 instead of 1/cos^2 - 1, and the Phong power
// is exp(e * log1p(-(1 - cos))) with 1 - cos = |h x n|^2 / (1 + cos).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "chakum_types.h"

namespace chk {

constexpr float kPi = 3.14159265359f;  // mathPrimitives.h:4
constexpr float kInvPi = 1.0f / 3.14159265359f;
constexpr float kInf = 1e30f;

// 3-vectors in FP32 (the fused render kernels) and FP64 (the per-sample
// entry points, which take the reference's double vectors).
template <typename T>
struct v3 {
  T x, y, z;
};
using f3 = v3<float>;
using d3 = v3<double>;
template <typename T>
struct same {
  using type = T;
};
template <typename T>
__host__ __device__ inline v3<T> mk3(T x, T y, T z) {
  return v3<T>{x, y, z};
}
__host__ __device__ inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ inline d3 dk(double x, double y, double z) { return d3{x, y, z}; }
template <typename T>
__host__ __device__ inline v3<T> operator+(v3<T> a, v3<T> b) {
  return mk3(a.x + b.x, a.y + b.y, a.z + b.z);
}
template <typename T>
__host__ __device__ inline v3<T> operator-(v3<T> a, v3<T> b) {
  return mk3(a.x - b.x, a.y - b.y, a.z - b.z);
}
template <typename T>
__host__ __device__ inline v3<T> operator*(v3<T> a, typename same<T>::type s) {
  return mk3(a.x * s, a.y * s, a.z * s);
}
template <typename T>
__host__ __device__ inline v3<T> operator*(v3<T> a, v3<T> b) {
  return mk3(a.x * b.x, a.y * b.y, a.z * b.z);
}
template <typename T>
__host__ __device__ inline T dot(v3<T> a, v3<T> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename T>
__host__ __device__ inline v3<T> cross(v3<T> a, v3<T> b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// scalar math by precision (FP32: hardware approximations where the
// render's error budget allows them)
__device__ inline float rsq(float a) { return rsqrtf(a); }
__device__ inline double rsq(double a) { return 1.0 / sqrt(a); }
__device__ inline float sqr_t(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ inline double sqr_t(double a) { return sqrt(a); }
__device__ inline float ex(float a) { return __expf(a); }
__device__ inline double ex(double a) { return exp(a); }
__device__ inline float lg1p(float a) { return log1pf(a); }
__device__ inline double lg1p(double a) { return log1p(a); }
__device__ inline float pw(float a, float b) { return __powf(a, b); }
__device__ inline double pw(double a, double b) { return pow(a, b); }
__device__ inline float fm(float a, float b, float c) { return fmaf(a, b, c); }
__device__ inline double fm(double a, double b, double c) { return fma(a, b, c); }
__device__ inline float fab(float a) { return fabsf(a); }
__device__ inline double fab(double a) { return fabs(a); }
__device__ inline float fmx(float a, float b) { return fmaxf(a, b); }
__device__ inline double fmx(double a, double b) { return fmax(a, b); }
__device__ inline float csign(float a, float b) { return copysignf(a, b); }
__device__ inline double csign(double a, double b) { return copysign(a, b); }
__device__ inline void sncs(float a, float *s, float *c) { __sincosf(a, s, c); }
__device__ inline void sncs(double a, double *s, double *c) { sincos(a, s, c); }
// a / b: FP32 through the hardware reciprocal (v_rcp_f32, 1 ulp) instead of
// the correctly rounded division sequence; FP64 exact
__device__ inline float dv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ inline double dv(double a, double b) { return a / b; }

template <typename T>
__device__ inline v3<T> normalize(v3<T> a) {
  return a * rsq(dot(a, a));
}
__device__ inline f3 xyz(float4 v) { return mk(v.x, v.y, v.z); }
__device__ inline float comp(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
struct u4 {
  uint32_t x, y, z, w;
};
__device__ inline u4 philox(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}
__device__ inline float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// ------------------------------------------------------------------ ONB
// Duff et al., "Building an Orthonormal Basis, Revisited", JCGT 6(1) 2017.
template <typename T>
__device__ inline void onb(v3<T> n, v3<T> &t, v3<T> &b) {
  const T sign = csign(T(1), n.z);
  const T a = dv(T(-1), sign + n.z);
  const T bb = n.x * n.y * a;
  t = mk3(T(1) + sign * n.x * n.x * a, sign * bb, -sign * n.x);
  b = mk3(bb, sign + n.y * n.y * a, -n.y);
}

// ------------------------------------------------------------------ materials
template <typename T>
struct MatT {
  uint32_t brdf, light;
  T ks, alpha, pe, iior, eior, refl;
  v3<T> c, rad;
};
using Mat = MatT<float>;
template <typename T = float>
__device__ inline MatT<T> load_mat(const chk_material *__restrict__ m) {
  MatT<T> r;
  r.brdf = m->brdf;
  r.light = m->light;
  r.ks = m->ks;
  r.alpha = m->alpha;
  r.pe = m->phong_exp;
  r.iior = m->int_ior;
  r.eior = m->ext_ior;
  r.refl = m->reflectance;
  r.c = mk3<T>(m->color[0], m->color[1], m->color[2]);
  r.rad = mk3<T>(m->radiance[0], m->radiance[1], m->radiance[2]);
  return r;
}

// pi and 1/pi with the reference's literal (mathPrimitives.h:4)
template <typename T>
__device__ inline T pi_v() {
  return T(3.14159265359);
}
template <typename T>
__device__ inline T inv_pi_v() {
  return T(1) / T(3.14159265359);
}

// fresnel (materialTypes.cpp:156-178)
template <typename T>
__device__ inline T fresnel(const MatT<T> &m, T c) {
  T g = dv(m.iior, m.eior);
  g = sqr_t(fm(g, g, fm(c, c, T(-1))));
  const T t1 = g - c, t2 = g + c;
  const T a = fm(c, t2, T(-1)), b = fm(c, t1, T(1));
  const T q = dv(t1, t2);
  return T(0.5) * q * q * (T(1) + dv(a * a, b * b));
}

// Microfacet terms from cos = h.n and sin^2 = |h x n|^2.
// ggxDist (materialTypes.cpp:180-191): a2 / (pi cos^4 (a2 + tan2)^2)
template <typename T>
__device__ inline T ggx_D(T a2, T c, T s2) {
  const T den = fm(a2, c * c, s2);  // cos^2 (a2 + tan2)
  return dv(a2, pi_v<T>() * den * den);
}
// beckmannDist (materialTypes.cpp:212-222): exp(-tan2/a2) / (pi cos^4 a2)
template <typename T>
__device__ inline T beck_D(T a2, T c, T s2) {
  const T c2 = c * c;
  return dv(ex(dv(-s2, c2 * a2)), pi_v<T>() * c2 * c2 * a2);
}
// Phong D (materialTypes.cpp:39,107): (e+2)/(2 pi) cos^e
template <typename T>
__device__ inline T phong_D(T e, T c, T s2) {
  // cos <= 0 (a direction below the surface, which pdfEval still weighs):
  // pow's own value, finite for an integral exponent and NaN otherwise
  if (!(c > T(0))) return (e + T(2)) * (T(0.5) * inv_pi_v<T>()) * pw(c, e);
  const T omc = s2 / (T(1) + c);  // 1 - cos, accurate near 1
  return (e + T(2)) * (T(0.5) * inv_pi_v<T>()) * ex(e * lg1p(-omc));
}
// ggxG1 (materialTypes.cpp:193-210); v.n = cv, tan^2 = s2v / cv^2
template <typename T>
__device__ inline T ggx_G1(T a2, T cv, T s2v, T hv) {
  const T c2 = cv * cv;
  if (s2v <= T(1e-15) * c2) return T(1);
  if (hv <= T(1e-15)) return T(0);
  return dv(T(2), T(1) + sqr_t(T(1) + dv(a2 * s2v, c2)));
}
// rational G1 of beckmannG1/phongG1 (materialTypes.cpp:224-258)
template <typename T>
__device__ inline T rational_G1(T a) {
  if (a < T(1.6))
    return dv(T(3.535) * a + T(2.181) * a * a, T(1) + T(2.276) * a + T(2.577) * a * a);
  return T(1);
}
template <typename T>
__device__ inline T beck_G1(T alpha, T cv, T s2v, T hv) {
  const T c2 = cv * cv;
  if (c2 >= T(0.999999)) return T(1);
  if (hv <= T(1e-15)) return T(0);
  // a = 1 / (alpha tan) = |cv| / (alpha sqrt(s2v))
  return rational_G1(dv(fab(cv), alpha * sqr_t(s2v)));
}
template <typename T>
__device__ inline T phong_G1(T e, T cv, T s2v, T hv) {
  const T c2 = cv * cv;
  if (c2 >= T(0.999999)) return T(1);
  if (hv <= T(1e-15)) return T(0);
  // a = sqrt((0.5 e + 1) / tan^2)
  return rational_G1(sqr_t(dv((T(0.5) * e + T(1)) * c2, s2v)));
}

// |a x b|^2 for unit vectors: sin^2 of the angle between them.
template <typename T>
__device__ inline T sin2(v3<T> a, v3<T> b) {
  const v3<T> c = cross(a, b);
  return dot(c, c);
}

// BasePrimitive::brdf = MaterialType::brdf * reflectance (scalar) and
// MaterialType::pdfEval, evaluated together (they share h and D).
template <typename T>
__device__ inline void brdf_pdf(const MatT<T> &m, v3<T> n, v3<T> wo_ref, v3<T> wo, v3<T> wi,
                                T &f, T &pdf) {
  const T ip = inv_pi_v<T>();
  const T cwi = dot(wi, n);
  const T diffuse_f = (T(1) - m.ks) * ip;
  const T diffuse_p = (T(1) - m.ks) * cwi * ip;
  if (m.brdf == CHK_BRDF_CLASSIC_PHONG) {
    const T cs = fmx(dot(wo_ref, wi), T(0));
    T lobe = (m.pe + T(1)) * pw(cs, m.pe) * (T(0.5) * ip);
    if (m.pe == T(0)) lobe = T(0.5) * ip;
    f = (m.ks * lobe + diffuse_f) * m.refl;
    pdf = m.ks * lobe + diffuse_p;
    return;
  }
  if (m.brdf == CHK_BRDF_GGX || m.brdf == CHK_BRDF_BECKMANN || m.brdf == CHK_BRDF_PHONG) {
    const v3<T> h = normalize(wo + wi);
    const T ch = dot(h, n), s2h = sin2(h, n);
    const T cwo = dot(wo, n);
    const T hwi = dot(h, wi), hwo = dot(h, wo);
    const T F = fresnel(m, hwi);
    T D, G;
    const T s2o = sin2(wo, n), s2i = sin2(wi, n);
    if (m.brdf == CHK_BRDF_GGX) {
      const T a2 = m.alpha * m.alpha;
      D = ggx_D(a2, ch, s2h);
      G = ggx_G1(a2, cwo, s2o, hwo) * ggx_G1(a2, cwi, s2i, hwi);
    } else if (m.brdf == CHK_BRDF_BECKMANN) {
      const T a2 = m.alpha * m.alpha;
      D = beck_D(a2, ch, s2h);
      G = beck_G1(m.alpha, cwo, s2o, hwo) * beck_G1(m.alpha, cwi, s2i, hwi);
    } else {
      D = phong_D(m.alpha, ch, s2h);
      G = phong_G1(m.alpha, cwo, s2o, hwo) * phong_G1(m.alpha, cwi, s2i, hwi);
    }
    f = (dv(m.ks * F * D * G, T(4) * cwi * cwo) + diffuse_f) * m.refl;
    pdf = dv(m.ks * D * ch, T(4) * hwo) + diffuse_p;
    return;
  }
  f = T(0);
  pdf = m.ks + diffuse_p;  // pdfEval's fall-through (materialTypes.cpp:111-114)
}

// getBrdfDirectionSampleRandomized (materialTypes.cpp:136-153).
template <typename T>
__device__ inline v3<T> brdf_sample(const MatT<T> &m, v3<T> n, v3<T> wo_ref, v3<T> wo, T ulobe,
                                    T e1, T e2) {
  T sphi, cphi;
  sncs(T(2) * pi_v<T>() * e2, &sphi, &cphi);
  v3<T> t, b;
  if (ulobe < m.ks) {
    if (m.brdf == CHK_BRDF_CLASSIC_PHONG) {
      // getClassicPhongDirectionSamples (domainSampler.cpp:264-314)
      const T ct = pw(T(1) - e1, dv(T(1), m.pe + T(1)));
      const T st = sqr_t(fmx(T(0), T(1) - ct * ct));
      onb(wo_ref, t, b);
      return t * (st * cphi) + b * (st * sphi) + wo_ref * ct;
    }
    T ct = T(1), st = T(0);
    if (m.brdf == CHK_BRDF_GGX) {
      // tan = alpha sqrt(e1 / (1 - e1)) (domainSampler.cpp:225)
      const T t2 = dv(m.alpha * m.alpha * e1, T(1) - e1);
      ct = rsq(T(1) + t2);
      st = sqr_t(t2) * ct;
    } else if (m.brdf == CHK_BRDF_BECKMANN) {
      // tan^2 = -alpha^2 ln(1 - e1) (domainSampler.cpp:230)
      const T t2 = -m.alpha * m.alpha * lg1p(-e1);
      ct = rsq(T(1) + t2);
      st = sqr_t(t2) * ct;
    } else if (m.brdf == CHK_BRDF_PHONG) {
      ct = pw(e1, dv(T(1), m.alpha + T(2)));  // domainSampler.cpp:235
      st = sqr_t(fmx(T(0), T(1) - ct * ct));
    }
    onb(n, t, b);
    const v3<T> h = t * (st * cphi) + b * (st * sphi) + n * ct;
    return normalize(h * (T(2) * dot(h, wo)) - wo);
  }
  // cosine lobe: sin(theta) = sqrt(e1) (domainSampler.cpp:39-49,154-189)
  const T st = sqr_t(e1), ct = sqr_t(T(1) - e1);
  onb(n, t, b);
  return t * (st * cphi) + b * (st * sphi) + n * ct;
}

// ------------------------------------------------------------------ geometry
// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013),
// two-sided, accepting t in (0, tmax) like Triangle::intersect + the
// strict `d < t` of intersectT (bvhAccel.cpp:221), and rejecting the
// near-parallel hits the reference rejects: (d.n)^2 <= 1e-10 |d|^2 for the
// unit normal n (geometryPrimitives.cpp:42).
struct RayPre {
  int kx, ky, kz;
  float Sx, Sy, Sz;
};
__device__ inline RayPre ray_pre(f3 d) {
  RayPre p;
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  p.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  p.kx = p.kz == 2 ? 0 : p.kz + 1;
  p.ky = p.kx == 2 ? 0 : p.kx + 1;
  float dz = comp(d, p.kz);
  if (dz < 0.0f) {
    int tmp = p.kx;
    p.kx = p.ky;
    p.ky = tmp;
  }
  p.Sx = comp(d, p.kx) / dz;
  p.Sy = comp(d, p.ky) / dz;
  p.Sz = 1.0f / dz;
  return p;
}

// Returns true and updates t/u/v on a hit closer than t.
__device__ inline bool tri_hit(const RayPre &p, f3 o, f3 d, f3 v0, f3 v1, f3 v2, float &t,
                               float &bu, float &bv) {
  f3 A = v0 - o, B = v1 - o, C = v2 - o;
  float Akz = comp(A, p.kz), Bkz = comp(B, p.kz), Ckz = comp(C, p.kz);
  float Ax = fmaf(-p.Sx, Akz, comp(A, p.kx)), Ay = fmaf(-p.Sy, Akz, comp(A, p.ky));
  float Bx = fmaf(-p.Sx, Bkz, comp(B, p.kx)), By = fmaf(-p.Sy, Bkz, comp(B, p.ky));
  float Cx = fmaf(-p.Sx, Ckz, comp(C, p.kx)), Cy = fmaf(-p.Sy, Ckz, comp(C, p.ky));
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  float det = U + V + W;
  if (det == 0.0f) return false;
  float T = U * (p.Sz * Akz) + V * (p.Sz * Bkz) + W * (p.Sz * Ckz);
  // 0 < T/det < t, sign-aware (two-sided)
  if (det > 0.0f) {
    if (!(T > 0.0f) || !(T < t * det)) return false;
  } else {
    if (!(T < 0.0f) || !(T > t * det)) return false;
  }
  // (d.n)^2 <= 1e-10 rejection with the unit normal (reference eps)
  f3 N = cross(v2 - v0, v1 - v0);
  float dn = dot(d, N);
  if (dn * dn <= 1e-10f * dot(N, N) * dot(d, d)) return false;
  float inv = 1.0f / det;
  t = T * inv;
  bu = V * inv;  // weight of v1 (beta)
  bv = W * inv;  // weight of v2 (gamma)
  return true;
}

// Offset a ray origin off a surface along n (Waechter & Binder, "A Fast and
// Robust Method for Avoiding Self-Intersection", Ray Tracing Gems 2019).
__device__ inline f3 offset_ray(f3 p, f3 n) {
  const float origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f, int_scale = 256.0f;
  int ox = (int)(int_scale * n.x), oy = (int)(int_scale * n.y), oz = (int)(int_scale * n.z);
  f3 pi = mk(__int_as_float(__float_as_int(p.x) + (p.x < 0 ? -ox : ox)),
             __int_as_float(__float_as_int(p.y) + (p.y < 0 ? -oy : oy)),
             __int_as_float(__float_as_int(p.z) + (p.z < 0 ? -oz : oz)));
  return mk(fabsf(p.x) < origin ? fmaf(float_scale, n.x, p.x) : pi.x,
            fabsf(p.y) < origin ? fmaf(float_scale, n.y, p.y) : pi.y,
            fabsf(p.z) < origin ? fmaf(float_scale, n.z, p.z) : pi.z);
}

}  // namespace chk
