// Analytic polygon-light shading: LightSource::getLightFromMeshSource_Analytic
// (lightSources_CV.cpp:323-333) for many shading points, in FP64.
//
// Per point: M_GGX (ltc.h:18-61) reads two rows of the 64 x 64 GGX fit
// (only Minv entries a[0], a[2], a[4], a[6], a[8] are non-zero, so a row is
// 5 doubles + amplitude), lerps them in theta, rotates into the frame of
// getBrdfApproxParam (materialTypes.cpp:52-74) and integrates every light
// triangle in the cosine domain (getLightAnalytic, lightSources_CV.cpp:271-321)
// and, for ks < 1, in world space (getLightAnalyticDiffuse, :221-269).
// The reference's clipping rules are kept, including the x-only "between"
// test on the CA edge (lightSources_CV.cpp:101-102,194-195).
//
// Why FP64: the CA-edge rule and the horizon tests make the reference's
// result discontinuous in the vertex coordinates, and the LTC transform of a
// small-roughness lobe amplifies rounding; an FP32 evaluation of these
// formulas lands up to 1e-2 away from the reference on random inputs, so
// the kernel computes in the reference's precision (FP32 in, FP32 out).
//
// One shading point per lane; the light triangles are wave-uniform (scalar
// loads); the two table rows come from L2.
#include "chk_device.h"
#include "chk_internal.h"

namespace chk {
namespace {

constexpr int kLtcBlock = 256;
constexpr double kPi64 = 3.14159265359;  // mathPrimitives.h:4

// integrateEdge (lightSources_CV.cpp:8-25): (v1 x v2).n * theta/sin(theta)
// with v_i = p_i/|p_i|, from the unnormalised p1, p2, c = p1 x p2 and
// cn = c.n: (v1 x v2).n = cn/(|p1||p2|) and sin(theta) = |c|/(|p1||p2|), so
// the term is cn * theta/|c| with theta = atan2(|c|, p1.p2) (the
// reference's acos(v1.v2), without its loss of precision near 0).
__device__ inline double edge64(d3 p1, d3 p2, d3 c, double cn) {
  const double s = sqrt(dot(c, c));
  const double th = atan2(s, dot(p1, p2));
  return th > 0.001 ? cn * (th / s) : cn / sqrt(dot(p1, p1) * dot(p2, p2));
}

// n.v for the clip plane: the LTC part clips against z = 0 in its own
// frame (ZPLANE), the diffuse part against the plane through x with normal n.
template <bool ZPLANE>
__device__ inline double ndot(d3 n, d3 v) {
  return ZPLANE ? v.z : dot(n, v);
}

// clipTriangle (lightSources_CV.cpp:28-126 and 129-219): the 4-point
// polygon of the part of ABC above the plane; false if it has none.
// nA, nB, nC: signs of the vertices (only their sign is read).
template <bool ZPLANE>
__device__ bool clip64(d3 A, d3 B, d3 C, d3 n, d3 P, double nA, double nB, double nC, d3 *clip) {
  const double eps = 1e-15;
  int cnt = 0, idx = 0;
  const d3 dAB = A - B, dBC = B - C, dCA = C - A;
  const double nAB = ndot<ZPLANE>(n, dAB), nBC = ndot<ZPLANE>(n, dBC),
               nCA = ndot<ZPLANE>(n, dCA);
  const double pA = ndot<ZPLANE>(n, P - A), pB = ndot<ZPLANE>(n, P - B),
               pC = ndot<ZPLANE>(n, P - C);
  if (nAB * nAB > eps) {
    const d3 I = dAB * (pA / nAB) + A;
    if (dot(I - A, B - I) > 0) {
      ++cnt;
      if (nA < 0) { clip[idx++] = I; clip[idx++] = B; }
      else { clip[idx++] = A; clip[idx++] = I; }
    }
  } else if (pA * pA < eps) {
    return false;
  }
  if (nBC * nBC > eps) {
    const d3 I = dBC * (pB / nBC) + B;
    if (dot(I - B, C - I) > 0) {
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
    // the reference tests only the x components on this edge
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

// |sum of edge terms| of one triangle (lightSources_CV.cpp:241-262 and
// 293-314).  ZPLANE: LTC frame (P = 0, n = z).
template <bool ZPLANE>
__device__ double tri64(d3 A, d3 B, d3 C, d3 n, d3 P) {
  const d3 p1 = A - P, p2 = B - P, p3 = C - P;
  // signs of n.(V - P)/|V - P|, as the reference tests them
  const double nA = ndot<ZPLANE>(n, p1), nB = ndot<ZPLANE>(n, p2), nC = ndot<ZPLANE>(n, p3);
  double sum;
  if (nA > 0 && nB > 0 && nC > 0) {
    const d3 c1 = cross(p1, p2), c2 = cross(p2, p3), c3 = cross(p3, p1);
    sum = edge64(p1, p2, c1, ndot<ZPLANE>(n, c1)) + edge64(p2, p3, c2, ndot<ZPLANE>(n, c2)) +
          edge64(p3, p1, c3, ndot<ZPLANE>(n, c3));
  } else if (nA < 0 && nB < 0 && nC < 0) {
    return 0.0;
  } else {
    d3 cl[4] = {dk(0, 0, 0), dk(0, 0, 0), dk(0, 0, 0), dk(0, 0, 0)};
    if (!clip64<ZPLANE>(A, B, C, n, P, nA, nB, nC, cl)) return 0.0;
    const d3 q1 = cl[0] - P, q2 = cl[1] - P, q3 = cl[2] - P, q4 = cl[3] - P;
    const d3 c1 = cross(q1, q2), c2 = cross(q2, q3), c3 = cross(q3, q4), c4 = cross(q4, q1);
    sum = edge64(q1, q2, c1, ndot<ZPLANE>(n, c1)) + edge64(q2, q3, c2, ndot<ZPLANE>(n, c2)) +
          edge64(q3, q4, c3, ndot<ZPLANE>(n, c3)) + edge64(q4, q1, c4, ndot<ZPLANE>(n, c4));
  }
  return fabs(sum);
}

__global__ __launch_bounds__(kLtcBlock) void k_ltc(const double *__restrict__ tab,
                                                   const float *__restrict__ X,
                                                   const float *__restrict__ N,
                                                   const float *__restrict__ V,
                                                   const float *__restrict__ alpha,
                                                   const float *__restrict__ ks, uint32_t k,
                                                   const float *__restrict__ lt,
                                                   const float *__restrict__ lrad, uint32_t nlt,
                                                   float *__restrict__ rgb) {
  const uint32_t i = blockIdx.x * kLtcBlock + threadIdx.x;
  if (i >= k) return;
  const size_t o = 3 * (size_t)i;
  const d3 x = dk(X[o], X[o + 1], X[o + 2]), n = dk(N[o], N[o + 1], N[o + 2]),
           wo = dk(V[o], V[o + 1], V[o + 2]);
  const double kss = ks[i];
  // M_GGX((float)acos(n.wo), (float)alpha) (ltc.h:18-61, materialTypes.cpp:54):
  // bins from the float theta and sqrt(alpha) in double, lerp with a float
  // weight
  const float theta = (float)acos(dot(n, wo));
  const double tq = (double)theta / (0.5 * kPi64) * 64.0;
  const int tb = max(0, min(63, (int)floor(tq)));
  const int ab = max(0, min(63, (int)floor(sqrt((double)alpha[i]) * 64.0)));
  const int tp = min(63, tb + 1);
  const double *r0 = tab + 6 * (ab + tb * 64);
  double m[6] = {r0[0], r0[1], r0[2], r0[3], r0[4], r0[5]};
  if (tb != tp) {
    const float mult = (float)(tq - (double)(float)tb);
    const double *r1 = tab + 6 * (ab + tp * 64);
#pragma unroll
    for (int c = 0; c < 6; ++c) m[c] = m[c] + (r1[c] - m[c]) * (double)mult;
  }
  // frame of getBrdfApproxParam (materialTypes.cpp:59-68); rows of Minv * R
  const d3 T1 = normalize(wo - n * dot(n, wo));
  const d3 T2 = normalize(cross(n, T1));
  const d3 R0 = T1 * m[0] + n * m[1], R1 = T2 * m[2], R2 = T1 * m[3] + n * m[4];
  double sl0 = 0, sl1 = 0, sl2 = 0, sd0 = 0, sd1 = 0, sd2 = 0;
  const d3 zero = dk(0, 0, 0), zax = dk(0, 0, 1);
  const bool do_ltc = kss != 0.0, do_dif = kss != 1.0;
  for (uint32_t j = 0; j < nlt; ++j) {
    const float *p = lt + 9 * (size_t)j;
    const d3 A = dk(p[0], p[1], p[2]), B = dk(p[3], p[4], p[5]), C = dk(p[6], p[7], p[8]);
    const double r0c = lrad[3 * j], r1c = lrad[3 * j + 1], r2c = lrad[3 * j + 2];
    if (do_ltc) {
      const d3 a = A - x, b = B - x, c = C - x;
      const d3 e1 = dk(dot(R0, a), dot(R1, a), dot(R2, a));
      const d3 e2 = dk(dot(R0, b), dot(R1, b), dot(R2, b));
      const d3 e3 = dk(dot(R0, c), dot(R1, c), dot(R2, c));
      const double s = tri64<true>(e1, e2, e3, zax, zero);
      sl0 += r0c * s;
      sl1 += r1c * s;
      sl2 += r2c * s;
    }
    if (do_dif) {
      const double s = tri64<false>(A, B, C, n, x);
      sd0 += r0c * s;
      sd1 += r1c * s;
      sd2 += r2c * s;
    }
  }
  // Sum / (2 pi), then Analytic = LTC * amp * ks + diffuse * (1 - ks)
  const double wl = m[5] * kss / (2.0 * kPi64), wd = (1.0 - kss) / (2.0 * kPi64);
  rgb[o] = (float)(sl0 * wl + sd0 * wd);
  rgb[o + 1] = (float)(sl1 * wl + sd1 * wd);
  rgb[o + 2] = (float)(sl2 * wl + sd2 * wd);
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
  // 2, 8 through mat3's column-major 9-argument ctor), amplitude
  std::vector<double> p64(4096 * 6);
  for (int r = 0; r < 4096; ++r) {
    const double *f = table + 19 * (size_t)r + 9;
    for (int z : {1, 3, 5, 7})
      if (f[z] != 0.0) {
        set_error("chk_set_ltc_table: Minv entry that should be zero is not");
        return CHK_EINVAL;
      }
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
      set_error("chk_set_ltc_table: cannot allocate the table");
      return CHK_ENOMEM;
    }
  }
  CHK_HIP(hipMemcpy(ctx->ltc_dev, p64.data(), p64.size() * 8, hipMemcpyHostToDevice));
  return CHK_OK;
}

chk_status chk_ltc_eval(chk_ctx_t ctx, const float *x, const float *n, const float *v,
                        const float *alpha, const float *ks, uint32_t k, const float *light_tris,
                        const float *radiance, uint32_t n_light_tris, float *rgb, void *stream) {
  if (!ctx) {
    set_error("chk_ltc_eval: null context");
    return CHK_EINVAL;
  }
  if (!ctx->ltc_dev) {
    set_error("chk_ltc_eval: call chk_set_ltc_table first");
    return CHK_EINVAL;
  }
  if (k == 0) return CHK_OK;
  if (!x || !n || !v || !alpha || !ks || !rgb || (n_light_tris && (!light_tris || !radiance))) {
    set_error("chk_ltc_eval: null argument");
    return CHK_EINVAL;
  }
  CHK_HIP(hipSetDevice(ctx->device));
  hipLaunchKernelGGL(k_ltc, dim3((k + kLtcBlock - 1) / kLtcBlock), dim3(kLtcBlock), 0,
                     (hipStream_t)stream, (const double *)ctx->ltc_dev, x, n, v, alpha, ks, k,
                     light_tris, radiance, n_light_tris, rgb);
  CHK_HIP(hipGetLastError());
  return CHK_OK;
}

}  // extern "C"
