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
#include "chk_internal.h"
#include "chk_ltc_dev.h"

namespace chk {
namespace {

constexpr int kLtcBlock = 256;

// The LTC frame of triangle j seen from x: its vertices relative to x,
// multiplied by the rows of Minv * R.
__device__ inline void ltc_vertices(const float *__restrict__ lt, uint32_t j, d3 x, d3 R0, d3 R1,
                                    d3 R2, d3 &e1, d3 &e2, d3 &e3) {
  const float *p = lt + 9 * (size_t)j;
  const d3 a = dk(p[0], p[1], p[2]) - x, b = dk(p[3], p[4], p[5]) - x,
           c = dk(p[6], p[7], p[8]) - x;
  e1 = dk(dot(R0, a), dot(R1, a), dot(R2, a));
  e2 = dk(dot(R0, b), dot(R1, b), dot(R2, b));
  e3 = dk(dot(R0, c), dot(R1, c), dot(R2, c));
}

// Light triangles are taken 32 at a time.  First pass: every lane
// integrates the triangles entirely above its horizon (three edges, no
// clipping; the value of tri64<true>) and marks those that straddle it.
// Second pass: each lane walks its own marked triangles (the lanes of a wave
// in step, each on its next one), clipping and integrating them (tri64's
// clip branch); a lane with none left repeats a harmless evaluation.  The
// clip code thus runs only as often as the lane with the most straddling
// triangles needs it.  Triangles below the horizon contribute 0 in both.
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
  // M_GGX((float)acos(n.wo), (float)alpha) (ltc.h:18-61, materialTypes.cpp:54)
  const LtcFit F = ltc_fit(tab, (float)acos(dot(n, wo)), (double)alpha[i]);
  const double *m = F.mi;
  // frame of getBrdfApproxParam (materialTypes.cpp:59-68); rows of Minv * R
  const d3 T1 = normalize(wo - n * dot(n, wo));
  const d3 T2 = normalize(cross(n, T1));
  const d3 R0 = T1 * m[0] + n * m[1], R1 = T2 * m[2], R2 = T1 * m[3] + n * m[4];
  double sl0 = 0, sl1 = 0, sl2 = 0, sd0 = 0, sd1 = 0, sd2 = 0;
  const d3 zero = dk(0, 0, 0), zax = dk(0, 0, 1);
  const bool do_ltc = kss != 0.0, do_dif = kss != 1.0;
  for (uint32_t j0 = 0; j0 < nlt; j0 += 32) {
    const uint32_t jn = min(32u, nlt - j0);
    uint32_t straddle = 0;
    for (uint32_t jj = 0; jj < jn; ++jj) {
      const uint32_t j = j0 + jj;
      const double r0c = lrad[3 * j], r1c = lrad[3 * j + 1], r2c = lrad[3 * j + 2];
      if (do_ltc) {
        d3 e1, e2, e3;
        ltc_vertices(lt, j, x, R0, R1, R2, e1, e2, e3);
        const bool above = e1.z > 0 && e2.z > 0 && e3.z > 0;
        const bool below = e1.z < 0 && e2.z < 0 && e3.z < 0;
        straddle |= (!above && !below) ? 1u << jj : 0u;
        const d3 c1 = cross(e1, e2), c2 = cross(e2, e3), c3 = cross(e3, e1);
        const double sum = edge64(e1, e2, c1, c1.z) + edge64(e2, e3, c2, c2.z) +
                           edge64(e3, e1, c3, c3.z);
        const double sv = above ? fabs(sum) : 0.0;
        sl0 += r0c * sv;
        sl1 += r1c * sv;
        sl2 += r2c * sv;
      }
      if (do_dif) {
        const float *p = lt + 9 * (size_t)j;
        const d3 A = dk(p[0], p[1], p[2]), B = dk(p[3], p[4], p[5]), C = dk(p[6], p[7], p[8]);
        const double sv = tri64<false>(A, B, C, n, x);
        sd0 += r0c * sv;
        sd1 += r1c * sv;
        sd2 += r2c * sv;
      }
    }
    while (__any(straddle != 0)) {  // wave-uniform loop
      const bool mine = straddle != 0;
      const uint32_t jj = mine ? (uint32_t)__builtin_ctz(straddle) : 0u;
      straddle &= straddle - 1u;
      const uint32_t j = j0 + jj;
      d3 e1, e2, e3;
      ltc_vertices(lt, j, x, R0, R1, R2, e1, e2, e3);
      Clip4 cl;
      cl.q0 = cl.q1 = cl.q2 = cl.q3 = zero;
      cl.idx = 0;
      const bool ok = clip64<true>(e1, e2, e3, zax, zero, e1.z, e2.z, e3.z, cl);
      const d3 c1 = cross(cl.q0, cl.q1), c2 = cross(cl.q1, cl.q2), c3 = cross(cl.q2, cl.q3),
               c4 = cross(cl.q3, cl.q0);
      const double sum = edge64(cl.q0, cl.q1, c1, c1.z) + edge64(cl.q1, cl.q2, c2, c2.z) +
                         edge64(cl.q2, cl.q3, c3, c3.z) + edge64(cl.q3, cl.q0, c4, c4.z);
      const double sv = (mine && ok) ? fabs(sum) : 0.0;
      sl0 += lrad[3 * j] * sv;
      sl1 += lrad[3 * j + 1] * sv;
      sl2 += lrad[3 * j + 2] * sv;
    }
  }
  // Sum / (2 pi), then Analytic = LTC * amp * ks + diffuse * (1 - ks)
  const double wl = F.amp * kss / (2.0 * kPi64), wd = (1.0 - kss) / (2.0 * kPi64);
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
  // packed record (kLtcRec doubles): Minv a[0], a[2], a[4], a[6], a[8]
  // (= file entries 0, 6, 4, 2, 8 through mat3's column-major 9-argument
  // ctor), amplitude, the same five entries of M, and det M as
  // mat3::calcDet computes it (mathPrimitives.h:165-169)
  std::vector<double> p64(4096 * (size_t)kLtcRec);
  for (int r = 0; r < 4096; ++r) {
    const double *fm = table + 19 * (size_t)r, *fi = fm + 9;
    for (int z : {1, 3, 5, 7})
      if (fi[z] != 0.0 || fm[z] != 0.0) {
        set_error("chk_set_ltc_table: an M or Minv entry that should be zero is not");
        return CHK_EINVAL;
      }
    double *q = &p64[kLtcRec * (size_t)r];
    const int pick[5] = {0, 6, 4, 2, 8};
    for (int c = 0; c < 5; ++c) {
      q[c] = fi[pick[c]];
      q[6 + c] = fm[pick[c]];
    }
    q[5] = fm[18];
    // a[] of M in mat3 order (column-major ctor), then calcDet
    const double a[9] = {fm[0], fm[3], fm[6], fm[1], fm[4], fm[7], fm[2], fm[5], fm[8]};
    q[11] = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
            a[2] * (a[3] * a[7] - a[4] * a[6]);
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
