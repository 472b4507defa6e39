// Analytic polygon-light shading: LightSource::getLightFromMeshSource_Analytic
// (lightSources_CV.cpp:323-333) for many shading points.
//
// Per point: M_GGX (ltc.h:18-61) reads two rows of the 64 x 64 GGX fit
// (only Minv entries a[0], a[2], a[4], a[6], a[8] are non-zero, so a row is
// 5 floats + amplitude), lerps them in theta, rotates into the frame of
// getBrdfApproxParam (materialTypes.cpp:52-74) and integrates every light
// triangle in the cosine domain (getLightAnalytic, lightSources_CV.cpp:271-321)
// and, for ks < 1, in world space (getLightAnalyticDiffuse, :221-269).
// The reference's clipping rules are kept, including the x-only "between"
// test on the CA edge (lightSources_CV.cpp:101-102,194-195).  The edge
// integral uses theta = atan2(|v1 x v2|, v1.v2), equal to the reference's
// acos(v1.v2) but well conditioned in FP32.
#include "chk_device.h"
#include "chk_internal.h"

namespace chk {
namespace {

constexpr int kLtcBlock = 256;
// A point whose edge terms cancel by more than the context's kappa (sum of
// |terms| over |sum|, radiance- and part-weighted) is recomputed by the FP64
// pass: each FP32 edge term carries a few ulps, so kappa bounds the relative
// error of the point at about kappa * 3e-7.

// integrateEdge (lightSources_CV.cpp:8-25): (v1 x v2).n * theta/sin(theta)
// for v_i = p_i/|p_i|, from the unnormalised p1, p2 and c = p1 x p2 with
// cn = c.n.  Since (v1 x v2).n = cn/(|p1||p2|) and sin(theta) = |c|/(|p1||p2|),
// the term is cn * theta/|c|, theta = atan2(|c|, p1.p2): no normalisation,
// and the caller forms c from a light edge (p1 x (p2 - p1)) without the
// cancellation of crossing two nearly parallel unit vectors.
__device__ inline float edge_u(f3 p1, f3 p2, f3 c, float cn) {
  const float s = sqrtf(dot(c, c));
  const float th = atan2f(s, dot(p1, p2));
  return th > 0.001f ? cn * (th / s) : cn * rsqrtf(dot(p1, p1) * dot(p2, p2));
}

// clipTriangle against z = 0 (lightSources_CV.cpp:129-219)
__device__ inline bool clip_z(f3 A, f3 B, f3 C, float nA, float nB, float nC, f3 *clip) {
  const float eps = 1e-15f;
  int cnt = 0, idx = 0;
  const f3 dAB = A - B, dBC = B - C, dCA = C - A;
  if (dAB.z * dAB.z > eps) {
    const f3 I = dAB * (-A.z / dAB.z) + A;
    if (dot(I - A, B - I) > 0.0f) {
      ++cnt;
      if (nA < 0.0f) { clip[idx++] = I; clip[idx++] = B; }
      else { clip[idx++] = A; clip[idx++] = I; }
    }
  } else if (A.z * A.z < eps) {
    return false;
  }
  if (dBC.z * dBC.z > eps) {
    const f3 I = dBC * (-B.z / dBC.z) + B;
    if (dot(I - B, C - I) > 0.0f) {
      ++cnt;
      if (idx < 4) {
        if (nB < 0.0f) { clip[idx++] = I; clip[idx++] = C; }
        else { clip[idx++] = B; clip[idx++] = I; }
      }
    }
  } else if (B.z * B.z < eps) {
    return false;
  }
  if (dCA.z * dCA.z > eps) {
    const f3 I = dCA * (-C.z / dCA.z) + C;
    if ((I.x - C.x) * (A.x - I.x) > 0.0f) {
      ++cnt;
      if (idx < 4) {
        if (nC < 0.0f) { clip[idx++] = I; clip[idx++] = A; }
        else { clip[idx++] = C; clip[idx++] = I; }
      }
    }
  } else if (C.z * C.z < eps) {
    return false;
  }
  return cnt > 1;
}

// clipTriangle against the plane through P with normal n
// (lightSources_CV.cpp:28-126)
__device__ inline bool clip_plane(f3 A, f3 B, f3 C, f3 n, f3 P, float nA, float nB, float nC,
                                  f3 *clip) {
  const float eps = 1e-15f;
  int cnt = 0, idx = 0;
  const f3 dAB = A - B, dBC = B - C, dCA = C - A;
  const float nAB = dot(n, dAB), nBC = dot(n, dBC), nCA = dot(n, dCA);
  const float pA = dot(n, P - A), pB = dot(n, P - B), pC = dot(n, P - C);
  if (nAB * nAB > eps) {
    const f3 I = dAB * (pA / nAB) + A;
    if (dot(I - A, B - I) > 0.0f) {
      ++cnt;
      if (nA < 0.0f) { clip[idx++] = I; clip[idx++] = B; }
      else { clip[idx++] = A; clip[idx++] = I; }
    }
  } else if (pA * pA < eps) {
    return false;
  }
  if (nBC * nBC > eps) {
    const f3 I = dBC * (pB / nBC) + B;
    if (dot(I - B, C - I) > 0.0f) {
      ++cnt;
      if (idx < 4) {
        if (nB < 0.0f) { clip[idx++] = I; clip[idx++] = C; }
        else { clip[idx++] = B; clip[idx++] = I; }
      }
    }
  } else if (pB * pB < eps) {
    return false;
  }
  if (nCA * nCA > eps) {
    const f3 I = dCA * (pC / nCA) + C;
    if ((I.x - C.x) * (A.x - I.x) > 0.0f) {
      ++cnt;
      if (idx < 4) {
        if (nC < 0.0f) { clip[idx++] = I; clip[idx++] = A; }
        else { clip[idx++] = C; clip[idx++] = I; }
      }
    }
  } else if (pC * pC < eps) {
    return false;
  }
  return cnt > 1;
}

__device__ inline f3 row_mul(f3 r0, f3 r1, f3 r2, f3 v) { return mk(dot(r0, v), dot(r1, v), dot(r2, v)); }

__global__ __launch_bounds__(kLtcBlock) void k_ltc(const float4 *__restrict__ tab,
                                                   const float *__restrict__ X,
                                                   const float *__restrict__ N,
                                                   const float *__restrict__ V,
                                                   const float *__restrict__ alpha,
                                                   const float *__restrict__ ks, uint32_t k,
                                                   const float *__restrict__ lt,
                                                   const float *__restrict__ lrad, uint32_t nlt,
                                                   float *__restrict__ rgb,
                                                   uint32_t *__restrict__ fb_list,
                                                   uint32_t *__restrict__ fb_count, float kappa) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t lane = __lane_id();
  for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < k; base += stride) {
    const uint32_t i = base + lane;
    bool flag = false;
    if (i < k) {
    const size_t o = 3 * (size_t)i;
    const f3 x = mk(X[o], X[o + 1], X[o + 2]);
    const f3 n = mk(N[o], N[o + 1], N[o + 2]);
    const f3 wo = mk(V[o], V[o + 1], V[o + 2]);
    const float a = alpha[i];
    const float kss = ks[i];
    // ---- M_GGX(acos(n.wo), alpha) (ltc.h:18-61), bins in double like the reference
    // float(acos(double(n.v))) as M_GGX receives it (ltc.h:18-19); float
    // products are exact in double, so the dot matches the reference's
    const float theta =
        (float)acos((double)n.x * wo.x + (double)n.y * wo.y + (double)n.z * wo.z);
    const double tq = (double)theta / (0.5 * 3.14159265359) * 64.0;
    const int tb = max(0, min(63, (int)floor(tq)));
    const int ab = max(0, min(63, (int)floor(sqrt((double)a) * 64.0)));
    const int tp = min(63, tb + 1);
    const float4 r0a = tab[2 * (ab + tb * 64)], r0b = tab[2 * (ab + tb * 64) + 1];
    float m0 = r0a.x, m2 = r0a.y, m4 = r0a.z, m6 = r0a.w, m8 = r0b.x, amp = r0b.y;
    if (tb != tp) {
      const float mult = (float)(tq - (double)(float)tb);
      const float4 r1a = tab[2 * (ab + tp * 64)], r1b = tab[2 * (ab + tp * 64) + 1];
      m0 = fmaf(r1a.x - m0, mult, m0);
      m2 = fmaf(r1a.y - m2, mult, m2);
      m4 = fmaf(r1a.z - m4, mult, m4);
      m6 = fmaf(r1a.w - m6, mult, m6);
      m8 = fmaf(r1b.x - m8, mult, m8);
      amp = fmaf(r1b.y - amp, mult, amp);
    }
    // ---- frame of getBrdfApproxParam (materialTypes.cpp:59-68)
    const f3 T1 = normalize(wo - n * dot(n, wo));
    const f3 T2 = normalize(cross(n, T1));
    // Minv * R, R rows (T1, T2, n); Minv = [[m0,0,m2],[0,m4,0],[m6,0,m8]]
    const f3 R0 = T1 * m0 + n * m2, R1 = T2 * m4, R2 = T1 * m6 + n * m8;
    float sl0 = 0.0f, sl1 = 0.0f, sl2 = 0.0f, sd0 = 0.0f, sd1 = 0.0f, sd2 = 0.0f;
    // conditioning: sum of |edge terms| against |edge sum| per part
    float al = 0.0f, bl = 0.0f, ad = 0.0f, bd = 0.0f;
    const bool do_ltc = kss != 0.0f, do_dif = kss != 1.0f;
    for (uint32_t j = 0; j < nlt; ++j) {
      const float *p = lt + 9 * (size_t)j;
      const f3 A = mk(p[0], p[1], p[2]), B = mk(p[3], p[4], p[5]), C = mk(p[6], p[7], p[8]);
      const f3 rad = mk(lrad[3 * j], lrad[3 * j + 1], lrad[3 * j + 2]);
      const float rmax = fmaxf(fmaxf(rad.x, rad.y), rad.z);
      if (do_ltc) {
        const f3 e1 = row_mul(R0, R1, R2, A - x), e2 = row_mul(R0, R1, R2, B - x),
                 e3 = row_mul(R0, R1, R2, C - x);
        float sum = 0.0f, mag = 0.0f;
        bool use = true;
        if (e1.z > 0.0f && e2.z > 0.0f && e3.z > 0.0f) {
          const f3 c1 = cross(e1, row_mul(R0, R1, R2, B - A)),
                   c2 = cross(e2, row_mul(R0, R1, R2, C - B)),
                   c3 = cross(e3, row_mul(R0, R1, R2, A - C));
          const float t1 = edge_u(e1, e2, c1, c1.z), t2 = edge_u(e2, e3, c2, c2.z),
                      t3 = edge_u(e3, e1, c3, c3.z);
          sum = t1 + t2 + t3;
          mag = fabsf(t1) + fabsf(t2) + fabsf(t3);
        } else if (e1.z < 0.0f && e2.z < 0.0f && e3.z < 0.0f) {
          use = false;
        } else {
          f3 cl[4] = {mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)};
          if (clip_z(e1, e2, e3, e1.z * rsqrtf(dot(e1, e1)), e2.z * rsqrtf(dot(e2, e2)),
                     e3.z * rsqrtf(dot(e3, e3)), cl)) {
            const f3 c1 = cross(cl[0], cl[1]), c2 = cross(cl[1], cl[2]), c3 = cross(cl[2], cl[3]),
                     c4 = cross(cl[3], cl[0]);
            const float t1 = edge_u(cl[0], cl[1], c1, c1.z), t2 = edge_u(cl[1], cl[2], c2, c2.z),
                        t3 = edge_u(cl[2], cl[3], c3, c3.z), t4 = edge_u(cl[3], cl[0], c4, c4.z);
            sum = t1 + t2 + t3 + t4;
            mag = fabsf(t1) + fabsf(t2) + fabsf(t3) + fabsf(t4);
          } else {
            use = false;
          }
        }
        if (use) {
          sum = fabsf(sum);
          sl0 = fmaf(rad.x, sum, sl0);
          sl1 = fmaf(rad.y, sum, sl1);
          sl2 = fmaf(rad.z, sum, sl2);
          al = fmaf(rmax, mag, al);
          bl = fmaf(rmax, sum, bl);
        }
      }
      if (do_dif) {
        const f3 p1 = A - x, p2 = B - x, p3 = C - x;
        // signs of n.(V - x) / |V - x| as the reference tests them
        const float nA = dot(p1, n), nB = dot(p2, n), nC = dot(p3, n);
        float sum = 0.0f, mag = 0.0f;
        bool use = true;
        if (nA > 0.0f && nB > 0.0f && nC > 0.0f) {
          const f3 c1 = cross(p1, B - A), c2 = cross(p2, C - B), c3 = cross(p3, A - C);
          const float t1 = edge_u(p1, p2, c1, dot(c1, n)), t2 = edge_u(p2, p3, c2, dot(c2, n)),
                      t3 = edge_u(p3, p1, c3, dot(c3, n));
          sum = t1 + t2 + t3;
          mag = fabsf(t1) + fabsf(t2) + fabsf(t3);
        } else if (nA < 0.0f && nB < 0.0f && nC < 0.0f) {
          use = false;
        } else {
          f3 cl[4] = {mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)};
          if (clip_plane(A, B, C, n, x, nA, nB, nC, cl)) {
            const f3 q1 = cl[0] - x, q2 = cl[1] - x, q3 = cl[2] - x, q4 = cl[3] - x;
            const f3 c1 = cross(q1, q2), c2 = cross(q2, q3), c3 = cross(q3, q4),
                     c4 = cross(q4, q1);
            const float t1 = edge_u(q1, q2, c1, dot(c1, n)), t2 = edge_u(q2, q3, c2, dot(c2, n)),
                        t3 = edge_u(q3, q4, c3, dot(c3, n)), t4 = edge_u(q4, q1, c4, dot(c4, n));
            sum = t1 + t2 + t3 + t4;
            mag = fabsf(t1) + fabsf(t2) + fabsf(t3) + fabsf(t4);
          } else {
            use = false;
          }
        }
        if (use) {
          sum = fabsf(sum);
          sd0 = fmaf(rad.x, sum, sd0);
          sd1 = fmaf(rad.y, sum, sd1);
          sd2 = fmaf(rad.z, sum, sd2);
          ad = fmaf(rmax, mag, ad);
          bd = fmaf(rmax, sum, bd);
        }
      }
    }
    // Sum / (2 pi), then Analytic = LTC * amp * ks + diffuse * (1 - ks)
    const float inv2pi = 1.0f / (2.0f * kPi);
    const float wl = inv2pi * amp * kss, wd = inv2pi * (1.0f - kss);
    rgb[o] = (do_ltc ? sl0 * wl : 0.0f) + (do_dif ? sd0 * wd : 0.0f);
    rgb[o + 1] = (do_ltc ? sl1 * wl : 0.0f) + (do_dif ? sd1 * wd : 0.0f);
    rgb[o + 2] = (do_ltc ? sl2 * wl : 0.0f) + (do_dif ? sd2 * wd : 0.0f);
    // an ill-conditioned point (edge terms cancel) is recomputed in FP64
    const float num = (do_ltc ? wl * al : 0.0f) + (do_dif ? wd * ad : 0.0f);
    const float den = (do_ltc ? wl * bl : 0.0f) + (do_dif ? wd * bd : 0.0f);
    flag = fb_list != nullptr && num > kappa * den;
    }
    // compacted append of flagged points (wave ballot, one atomic per wave)
    const unsigned long long mask = __ballot(flag);
    if (mask) {
      const uint32_t cnt = (uint32_t)__popcll(mask);
      const int leader = __ffsll((long long)mask) - 1;
      uint32_t b = 0;
      if ((int)lane == leader) b = atomicAdd(fb_count, cnt);
      b = __shfl(b, leader);
      const unsigned long long below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
      if (flag) fb_list[b + (uint32_t)__popcll(below)] = i;
    }
  }
}

// ---------------------------------------------------------------- FP64 pass
// The same computation as getLightFromMeshSource_Analytic in FP64, from the
// FP64 table, for the points the FP32 pass flagged.
__device__ inline double ddot(d3 a, d3 b) { return dot(a, b); }
__device__ inline d3 dcross(d3 a, d3 b) { return cross(a, b); }
__device__ inline d3 dnorm(d3 a) { return normalize(a); }
__device__ inline double dedge(d3 v1, d3 v2, d3 n) {  // lightSources_CV.cpp:8-15
  const double th = acos(ddot(v1, v2));
  return ddot(dcross(v1, v2), n) * ((th > 0.001) ? th / sin(th) : 1.0);
}

__device__ bool dclip(d3 A, d3 B, d3 C, d3 n, d3 P, double nA, double nB, double nC, d3 *clip) {
  const double eps = 1e-15;
  int cnt = 0, idx = 0;
  const d3 dAB = A - B, dBC = B - C, dCA = C - A;
  const double nAB = ddot(n, dAB), nBC = ddot(n, dBC), nCA = ddot(n, dCA);
  const double pA = ddot(n, P - A), pB = ddot(n, P - B), pC = ddot(n, P - C);
  if (nAB * nAB > eps) {
    const d3 I = dAB * (pA / nAB) + A;
    if (ddot(I - A, B - I) > 0) {
      ++cnt;
      if (nA < 0) { clip[idx++] = I; clip[idx++] = B; }
      else { clip[idx++] = A; clip[idx++] = I; }
    }
  } else if (pA * pA < eps) {
    return false;
  }
  if (nBC * nBC > eps) {
    const d3 I = dBC * (pB / nBC) + B;
    if (ddot(I - B, C - I) > 0) {
      ++cnt;
      if (idx < 4) {
        if (nB < 0) { clip[idx++] = I; clip[idx++] = C; }
        else { clip[idx++] = B; clip[idx++] = I; }
      }
    }
  } else if (pB * pB < eps) {
    return false;
  }
  if (nCA * nCA > eps) {
    const d3 I = dCA * (pC / nCA) + C;
    if ((I.x - C.x) * (A.x - I.x) > 0) {
      ++cnt;
      if (idx < 4) {
        if (nC < 0) { clip[idx++] = I; clip[idx++] = A; }
        else { clip[idx++] = C; clip[idx++] = I; }
      }
    }
  } else if (pC * pC < eps) {
    return false;
  }
  return cnt > 1;
}

// One triangle of either part: the LTC part is the plane z = 0 through the
// origin of the transformed frame (n = (0,0,1), P = 0), the diffuse part the
// plane through x with normal n; both use the same clip rules.
__device__ double dtri(d3 A, d3 B, d3 C, d3 n, d3 P) {
  d3 e1 = dnorm(A - P), e2 = dnorm(B - P), e3 = dnorm(C - P);
  const double nA = ddot(e1, n), nB = ddot(e2, n), nC = ddot(e3, n);
  double sum;
  if (nA > 0 && nB > 0 && nC > 0) {
    sum = dedge(e1, e2, n);
    sum += dedge(e2, e3, n);
    sum += dedge(e3, e1, n);
  } else if (nA < 0 && nB < 0 && nC < 0) {
    return 0.0;
  } else {
    d3 cl[4] = {dk(0, 0, 0), dk(0, 0, 0), dk(0, 0, 0), dk(0, 0, 0)};
    if (!dclip(A, B, C, n, P, nA, nB, nC, cl)) return 0.0;
    const d3 c1 = dnorm(cl[0] - P), c2 = dnorm(cl[1] - P), c3 = dnorm(cl[2] - P),
             c4 = dnorm(cl[3] - P);
    sum = dedge(c1, c2, n);
    sum += dedge(c2, c3, n);
    sum += dedge(c3, c4, n);
    sum += dedge(c4, c1, n);
  }
  return fabs(sum);
}

__global__ __launch_bounds__(128) void k_ltc64(const double *__restrict__ tab64,
                                              const float *__restrict__ X,
                                              const float *__restrict__ N,
                                              const float *__restrict__ V,
                                              const float *__restrict__ alpha,
                                              const float *__restrict__ ks,
                                              const float *__restrict__ lt,
                                              const float *__restrict__ lrad, uint32_t nlt,
                                              float *__restrict__ rgb,
                                              const uint32_t *__restrict__ fb_list,
                                              const uint32_t *__restrict__ fb_count) {
  const uint32_t cnt = *fb_count;
  const double PI = 3.14159265359;
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += gridDim.x * blockDim.x) {
    const uint32_t i = fb_list[q];
    const size_t o = 3 * (size_t)i;
    const d3 x = dk(X[o], X[o + 1], X[o + 2]), n = dk(N[o], N[o + 1], N[o + 2]),
             wo = dk(V[o], V[o + 1], V[o + 2]);
    const double kss = ks[i];
    // M_GGX((float)acos(n.wo), (float)alpha) (ltc.h:18-61, materialTypes.cpp:54)
    const float theta = (float)acos(ddot(n, wo));
    const double tq = (double)theta / (0.5 * PI) * 64.0;
    const int tb = max(0, min(63, (int)floor(tq)));
    const int ab = max(0, min(63, (int)floor(sqrt((double)alpha[i]) * 64.0)));
    const int tp = min(63, tb + 1);
    const double *r0 = tab64 + 6 * (ab + tb * 64);
    double m[6] = {r0[0], r0[1], r0[2], r0[3], r0[4], r0[5]};
    if (tb != tp) {
      const float mult = (float)(tq - (double)(float)tb);
      const double *r1 = tab64 + 6 * (ab + tp * 64);
      for (int c = 0; c < 6; ++c) m[c] = m[c] + (r1[c] - m[c]) * (double)mult;
    }
    const d3 T1 = dnorm(wo - n * ddot(n, wo));
    const d3 T2 = dnorm(dcross(n, T1));
    const d3 R0 = T1 * m[0] + n * m[1], R1 = T2 * m[2], R2 = T1 * m[3] + n * m[4];
    double sl[3] = {0, 0, 0}, sd[3] = {0, 0, 0};
    const d3 zero = dk(0, 0, 0), zax = dk(0, 0, 1);
    for (uint32_t j = 0; j < nlt; ++j) {
      const float *p = lt + 9 * (size_t)j;
      const d3 A = dk(p[0], p[1], p[2]), B = dk(p[3], p[4], p[5]), C = dk(p[6], p[7], p[8]);
      const double rad[3] = {lrad[3 * j], lrad[3 * j + 1], lrad[3 * j + 2]};
      if (kss != 0.0) {
        const d3 a = A - x, b = B - x, c = C - x;
        const d3 e1 = dk(ddot(R0, a), ddot(R1, a), ddot(R2, a));
        const d3 e2 = dk(ddot(R0, b), ddot(R1, b), ddot(R2, b));
        const d3 e3 = dk(ddot(R0, c), ddot(R1, c), ddot(R2, c));
        // clipTriangle(e1, e2, e3, e.z/|e|, ...) against z = 0 (lightSources_CV.cpp:301)
        const double s = dtri(e1, e2, e3, zax, zero);
        for (int ch = 0; ch < 3; ++ch) sl[ch] += rad[ch] * s;
      }
      if (kss != 1.0) {
        const double s = dtri(A, B, C, n, x);
        for (int ch = 0; ch < 3; ++ch) sd[ch] += rad[ch] * s;
      }
    }
    for (int ch = 0; ch < 3; ++ch) {
      double G = sl[ch] / (2.0 * PI) * m[5] * kss;
      G = G + sd[ch] / (2.0 * PI) * (1 - kss);
      rgb[o + ch] = (float)G;
    }
  }
}

}  // namespace
}  // namespace chk

using namespace chk;

extern "C" {

chk_status chk_set_ltc_table(chk_ctx_t ctx, const double *table, size_t n_doubles) {
  if (!ctx || !table || n_doubles != 4096u * 19u) {
    set_error("chk_set_ltc_table: expected 4096 x 19 doubles");
    return CHK_EINVAL;
  }
  CHK_HIP(hipSetDevice(ctx->device));
  // packed row: Minv a[0], a[2], a[4], a[6], a[8] (= file entries 0, 6, 4,
  // 2, 8 through mat3's column-major 9-argument ctor), amplitude, 0, 0
  std::vector<float> packed(4096 * 8, 0.0f);
  for (int r = 0; r < 4096; ++r) {
    const double *f = table + 19 * (size_t)r + 9;
    float *q = &packed[8 * (size_t)r];
    q[0] = (float)f[0];
    q[1] = (float)f[6];
    q[2] = (float)f[4];
    q[3] = (float)f[2];
    q[4] = (float)f[8];
    q[5] = (float)table[19 * (size_t)r + 18];
    for (int z : {1, 3, 5, 7})
      if (f[z] != 0.0) {
        set_error("chk_set_ltc_table: Minv entry that should be zero is not");
        return CHK_EINVAL;
      }
  }
  if (!ctx->ltc_tab) {
    if (hipMalloc(&ctx->ltc_tab, packed.size() * 4) != hipSuccess) {
      ctx->ltc_tab = nullptr;
      set_error("chk_set_ltc_table: cannot allocate the table");
      return CHK_ENOMEM;
    }
  }
  CHK_HIP(hipMemcpy(ctx->ltc_tab, packed.data(), packed.size() * 4, hipMemcpyHostToDevice));
  // FP64 copy for the FP64 pass: same five Minv entries and the amplitude
  std::vector<double> p64(4096 * 6);
  for (int r = 0; r < 4096; ++r) {
    const double *f = table + 19 * (size_t)r + 9;
    double *q = &p64[6 * (size_t)r];
    q[0] = f[0];
    q[1] = f[6];
    q[2] = f[4];
    q[3] = f[2];
    q[4] = f[8];
    q[5] = table[19 * (size_t)r + 18];
  }
  if (!ctx->ltc_dev) {
    if (hipMalloc(&ctx->ltc_dev, p64.size() * 8) != hipSuccess) {
      ctx->ltc_dev = nullptr;
      set_error("chk_set_ltc_table: cannot allocate the FP64 table");
      return CHK_ENOMEM;
    }
  }
  CHK_HIP(hipMemcpy(ctx->ltc_dev, p64.data(), p64.size() * 8, hipMemcpyHostToDevice));
  ctx->ltc_host.assign(table, table + n_doubles);
  return CHK_OK;
}

chk_status chk_ltc_eval(chk_ctx_t ctx, const float *x, const float *n, const float *v,
                        const float *alpha, const float *ks, uint32_t k, const float *light_tris,
                        const float *radiance, uint32_t n_light_tris, float *rgb, void *stream) {
  if (!ctx) {
    set_error("chk_ltc_eval: null context");
    return CHK_EINVAL;
  }
  if (!ctx->ltc_tab) {
    set_error("chk_ltc_eval: call chk_set_ltc_table first");
    return CHK_EINVAL;
  }
  if (k == 0) return CHK_OK;
  if (!x || !n || !v || !alpha || !ks || !rgb || (n_light_tris && (!light_tris || !radiance))) {
    set_error("chk_ltc_eval: null argument");
    return CHK_EINVAL;
  }
  CHK_HIP(hipSetDevice(ctx->device));
  const size_t need = 256 + 4 * (size_t)k;
  if (ctx->ltc_scratch_bytes < need) {
    if (ctx->ltc_scratch) (void)hipFree(ctx->ltc_scratch);
    ctx->ltc_scratch = nullptr;
    ctx->ltc_scratch_bytes = 0;
    if (hipMalloc(&ctx->ltc_scratch, need) != hipSuccess) {
      set_error("chk_ltc_eval: cannot allocate the FP64-pass list");
      return CHK_ENOMEM;
    }
    ctx->ltc_scratch_bytes = need;
  }
  uint32_t *fb_count = (uint32_t *)ctx->ltc_scratch;
  uint32_t *fb_list = (uint32_t *)((char *)ctx->ltc_scratch + 256);
  hipStream_t st = (hipStream_t)stream;
  CHK_HIP(launch_zero_u32(fb_count, 1, st));
  const uint32_t blocks = std::min<uint32_t>((k + kLtcBlock - 1) / kLtcBlock, (uint32_t)ctx->n_cu * 16);
  hipLaunchKernelGGL(k_ltc, dim3(blocks), dim3(kLtcBlock), 0, st, (const float4 *)ctx->ltc_tab, x, n,
                     v, alpha, ks, k, light_tris, radiance, n_light_tris, rgb, fb_list, fb_count,
                     ctx->ltc_kappa);
  CHK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_ltc64, dim3(ctx->n_cu * 2), dim3(128), 0, st, (const double *)ctx->ltc_dev, x,
                     n, v, alpha, ks, light_tris, radiance, n_light_tris, rgb, fb_list, fb_count);
  CHK_HIP(hipGetLastError());
  return CHK_OK;
}

chk_status chk_ltc_set_fp64_threshold(chk_ctx_t ctx, float kappa) {
  if (!ctx || !(kappa >= 0.0f)) {
    set_error("chk_ltc_set_fp64_threshold: null context or kappa not >= 0");
    return CHK_EINVAL;
  }
  ctx->ltc_kappa = kappa;
  return CHK_OK;
}

chk_status chk_ltc_fp64_count(chk_ctx_t ctx, uint32_t *count, void *stream) {
  if (!ctx || !count) {
    set_error("chk_ltc_fp64_count: null argument");
    return CHK_EINVAL;
  }
  *count = 0;
  if (!ctx->ltc_scratch) return CHK_OK;
  CHK_HIP(hipSetDevice(ctx->device));
  CHK_HIP(hipMemcpyAsync(count, ctx->ltc_scratch, 4, hipMemcpyDeviceToHost, (hipStream_t)stream));
  CHK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return CHK_OK;
}

}  // extern "C"
