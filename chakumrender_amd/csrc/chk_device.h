// Device-side building blocks of the per-sample radiance path (gfx950, FP32).
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

struct f3 {
  float x, y, z;
};
__host__ __device__ inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ inline f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__host__ __device__ inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline f3 cross(f3 a, f3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline f3 normalize(f3 a) { return a * rsqrtf(dot(a, a)); }
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
__device__ inline void onb(f3 n, f3 &t, f3 &b) {
  float sign = copysignf(1.0f, n.z);
  float a = -1.0f / (sign + n.z);
  float bb = n.x * n.y * a;
  t = mk(1.0f + sign * n.x * n.x * a, sign * bb, -sign * n.x);
  b = mk(bb, sign + n.y * n.y * a, -n.y);
}

// ------------------------------------------------------------------ materials
struct Mat {
  uint32_t brdf, light;
  float ks, alpha, pe, iior, eior, refl;
  f3 c, rad;
};
__device__ inline Mat load_mat(const chk_material *__restrict__ m) {
  Mat r;
  r.brdf = m->brdf;
  r.light = m->light;
  r.ks = m->ks;
  r.alpha = m->alpha;
  r.pe = m->phong_exp;
  r.iior = m->int_ior;
  r.eior = m->ext_ior;
  r.refl = m->reflectance;
  r.c = mk(m->color[0], m->color[1], m->color[2]);
  r.rad = mk(m->radiance[0], m->radiance[1], m->radiance[2]);
  return r;
}

// fresnel (materialTypes.cpp:156-178)
__device__ inline float fresnel(const Mat &m, float c) {
  float g = m.iior / m.eior;
  g = sqrtf(fmaf(g, g, fmaf(c, c, -1.0f)));
  float t1 = g - c, t2 = g + c;
  float a = fmaf(c, t2, -1.0f), b = fmaf(c, t1, 1.0f);
  float q = t1 / t2;
  return 0.5f * q * q * (1.0f + (a * a) / (b * b));
}

// Microfacet terms from cos = h.n and sin^2 = |h x n|^2.
// ggxDist (materialTypes.cpp:180-191): a2 / (pi cos^4 (a2 + tan2)^2)
__device__ inline float ggx_D(float a2, float c, float s2) {
  float c2 = c * c;
  float den = fmaf(a2, c2, s2);  // cos^2 (a2 + tan2)
  return a2 / (kPi * den * den);
}
// beckmannDist (materialTypes.cpp:212-222): exp(-tan2/a2) / (pi cos^4 a2)
__device__ inline float beck_D(float a2, float c, float s2) {
  float c2 = c * c;
  return __expf(-s2 / (c2 * a2)) / (kPi * c2 * c2 * a2);
}
// Phong D (materialTypes.cpp:39,107): (e+2)/(2 pi) cos^e
__device__ inline float phong_D(float e, float c, float s2) {
  float omc = s2 / (1.0f + c);  // 1 - cos, accurate near 1
  return (e + 2.0f) * (0.5f * kInvPi) * __expf(e * log1pf(-omc));
}
// ggxG1 (materialTypes.cpp:193-210); v.n = cv, tan^2 = s2v / cv^2
__device__ inline float ggx_G1(float a2, float cv, float s2v, float hv) {
  float c2 = cv * cv;
  if (s2v <= 1e-15f * c2) return 1.0f;
  if (hv <= 1e-15f) return 0.0f;
  return 2.0f / (1.0f + sqrtf(1.0f + a2 * s2v / c2));
}
// rational G1 of beckmannG1/phongG1 (materialTypes.cpp:224-258)
__device__ inline float rational_G1(float a) {
  if (a < 1.6f) return (3.535f * a + 2.181f * a * a) / (1.0f + 2.276f * a + 2.577f * a * a);
  return 1.0f;
}
__device__ inline float beck_G1(float alpha, float cv, float s2v, float hv) {
  float c2 = cv * cv;
  if (c2 >= 0.999999f) return 1.0f;
  if (hv <= 1e-15f) return 0.0f;
  // a = 1 / (alpha tan) = |cv| / (alpha sqrt(s2v))
  return rational_G1(fabsf(cv) / (alpha * sqrtf(s2v)));
}
__device__ inline float phong_G1(float e, float cv, float s2v, float hv) {
  float c2 = cv * cv;
  if (c2 >= 0.999999f) return 1.0f;
  if (hv <= 1e-15f) return 0.0f;
  // a = sqrt((0.5 e + 1) / tan^2)
  return rational_G1(sqrtf((0.5f * e + 1.0f) * c2 / s2v));
}

// |a x b|^2 for unit vectors: sin^2 of the angle between them.
__device__ inline float sin2(f3 a, f3 b) {
  f3 c = cross(a, b);
  return dot(c, c);
}

// BasePrimitive::brdf = MaterialType::brdf * reflectance (scalar) and
// MaterialType::pdfEval, evaluated together (they share h and D).
__device__ inline void brdf_pdf(const Mat &m, f3 n, f3 wo_ref, f3 wo, f3 wi, float &f, float &pdf) {
  float cwi = dot(wi, n);
  float diffuse_f = (1.0f - m.ks) * kInvPi;
  float diffuse_p = (1.0f - m.ks) * cwi * kInvPi;
  if (m.brdf == CHK_BRDF_CLASSIC_PHONG) {
    float cs = fmaxf(dot(wo_ref, wi), 0.0f);
    float lobe = (m.pe + 1.0f) * __powf(cs, m.pe) * (0.5f * kInvPi);
    if (m.pe == 0.0f) lobe = 0.5f * kInvPi;
    f = (m.ks * lobe + diffuse_f) * m.refl;
    pdf = m.ks * lobe + diffuse_p;
    return;
  }
  if (m.brdf == CHK_BRDF_GGX || m.brdf == CHK_BRDF_BECKMANN || m.brdf == CHK_BRDF_PHONG) {
    f3 h = normalize(wo + wi);
    float ch = dot(h, n), s2h = sin2(h, n);
    float cwo = dot(wo, n);
    float hwi = dot(h, wi), hwo = dot(h, wo);
    float F = fresnel(m, hwi);
    float D, G;
    float s2o = sin2(wo, n), s2i = sin2(wi, n);
    if (m.brdf == CHK_BRDF_GGX) {
      float a2 = m.alpha * m.alpha;
      D = ggx_D(a2, ch, s2h);
      G = ggx_G1(a2, cwo, s2o, hwo) * ggx_G1(a2, cwi, s2i, hwi);
    } else if (m.brdf == CHK_BRDF_BECKMANN) {
      float a2 = m.alpha * m.alpha;
      D = beck_D(a2, ch, s2h);
      G = beck_G1(m.alpha, cwo, s2o, hwo) * beck_G1(m.alpha, cwi, s2i, hwi);
    } else {
      D = phong_D(m.alpha, ch, s2h);
      G = phong_G1(m.alpha, cwo, s2o, hwo) * phong_G1(m.alpha, cwi, s2i, hwi);
    }
    f = ((m.ks * F * D * G) / (4.0f * cwi * cwo) + diffuse_f) * m.refl;
    pdf = m.ks * D * ch / (4.0f * hwo) + diffuse_p;
    return;
  }
  f = 0.0f;
  pdf = m.ks + diffuse_p;  // pdfEval's fall-through (materialTypes.cpp:111-114)
}

// getBrdfDirectionSampleRandomized (materialTypes.cpp:136-153).
__device__ inline f3 brdf_sample(const Mat &m, f3 n, f3 wo_ref, f3 wo, float ulobe, float e1,
                                 float e2) {
  float sphi, cphi;
  __sincosf(2.0f * kPi * e2, &sphi, &cphi);
  f3 t, b;
  if (ulobe < m.ks) {
    if (m.brdf == CHK_BRDF_CLASSIC_PHONG) {
      // getClassicPhongDirectionSamples (domainSampler.cpp:264-314)
      float ct = __powf(1.0f - e1, 1.0f / (m.pe + 1.0f));
      float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
      onb(wo_ref, t, b);
      return t * (st * cphi) + b * (st * sphi) + wo_ref * ct;
    }
    float ct = 1.0f, st = 0.0f;
    if (m.brdf == CHK_BRDF_GGX) {
      // tan = alpha sqrt(e1 / (1 - e1)) (domainSampler.cpp:225)
      float t2 = m.alpha * m.alpha * e1 / (1.0f - e1);
      ct = rsqrtf(1.0f + t2);
      st = sqrtf(t2) * ct;
    } else if (m.brdf == CHK_BRDF_BECKMANN) {
      // tan^2 = -alpha^2 ln(1 - e1) (domainSampler.cpp:230)
      float t2 = -m.alpha * m.alpha * log1pf(-e1);
      ct = rsqrtf(1.0f + t2);
      st = sqrtf(t2) * ct;
    } else if (m.brdf == CHK_BRDF_PHONG) {
      ct = __powf(e1, 1.0f / (m.alpha + 2.0f));  // domainSampler.cpp:235
      st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    }
    onb(n, t, b);
    f3 h = t * (st * cphi) + b * (st * sphi) + n * ct;
    return normalize(h * (2.0f * dot(h, wo)) - wo);
  }
  // cosine lobe: sin(theta) = sqrt(e1) (domainSampler.cpp:39-49,154-189)
  float st = sqrtf(e1), ct = sqrtf(1.0f - e1);
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
