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
// One shading point per lane; the light triangles are staged in LDS in
// FP64; the two table rows come from L2.
#include "chk_internal.h"
#include "chk_ltc_dev.h"

namespace chk {
namespace {

constexpr int kLtcBlock = 256;
constexpr int kLtcWaves = kLtcBlock / 64;
constexpr int kChunk = 32;  // light triangles per chunk

// Light triangles of the current chunk in FP64 (vertices, radiance), and per
// wave: the 64 points' LTC frames, the (triangle, point) pairs of the chunk
// whose triangle lies above the point's horizon, and the results of one
// window of 64 pairs.
struct LtcLds {
  double tri[kChunk][9];
  double rad[kChunk][3];
  struct Wave {
    double frame[12][64];  // x, then the rows of Minv * R, component-major
    uint16_t pair[kChunk * 64];  // triangle << 6 | lane, triangle-major
    unsigned long long mask[kChunk];
    uint32_t base[kChunk];
    double res[2][64];
  } w[kLtcWaves];
  uint32_t quads;  // the chunk's triangles pair up into quads (ltc_quad_pairs)
};

__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// The LTC frame of triangle t seen from x: its vertices relative to x,
// multiplied by the rows of Minv * R.
__device__ inline void ltc_vertices(const double *p, d3 x, d3 R0, d3 R1, d3 R2, d3 &e1, d3 &e2,
                                    d3 &e3) {
  const d3 a = dk(p[0], p[1], p[2]) - x, b = dk(p[3], p[4], p[5]) - x,
           c = dk(p[6], p[7], p[8]) - x;
  e1 = dk(dot(R0, a), dot(R1, a), dot(R2, a));
  e2 = dk(dot(R0, b), dot(R1, b), dot(R2, b));
  e3 = dk(dot(R0, c), dot(R1, c), dot(R2, c));
}

// |sum of the three edge terms| of a triangle above the horizon
// (lightSources_CV.cpp:293-314 without clipping).
__device__ inline double ltc_unclipped(d3 e1, d3 e2, d3 e3) {
  const d3 c1 = cross(e1, e2), c2 = cross(e2, e3), c3 = cross(e3, e1);
  return fabs(edge64(e1, e2, c1, c1.z) + edge64(e2, e3, c2, c2.z) + edge64(e3, e1, c3, c3.z));
}

// The same for a triangle straddling the horizon: clipTriangle
// (lightSources_CV.cpp:129-219) to z >= 0, then the 4-point polygon.
__device__ inline double ltc_clipped(d3 e1, d3 e2, d3 e3) {
  const d3 zero = dk(0, 0, 0), zax = dk(0, 0, 1);
  Clip4 cl;
  cl.q0 = cl.q1 = cl.q2 = cl.q3 = zero;
  cl.idx = 0;
  const bool ok = clip64<true>(e1, e2, e3, zax, zero, e1.z, e2.z, e3.z, cl);
  const d3 c1 = cross(cl.q0, cl.q1), c2 = cross(cl.q1, cl.q2), c3 = cross(cl.q2, cl.q3),
           c4 = cross(cl.q3, cl.q0);
  const double sum = edge64(cl.q0, cl.q1, c1, c1.z) + edge64(cl.q1, cl.q2, c2, c2.z) +
                     edge64(cl.q2, cl.q3, c3, c3.z) + edge64(cl.q3, cl.q0, c4, c4.z);
  return ok ? fabs(sum) : 0.0;
}

// The wave's (triangle, point) pairs of one kind (bits: this lane's
// triangles of the chunk), listed triangle-major and evaluated 64 at a time,
// each lane on the next pair whatever its point (the point's frame comes
// from LDS); each lane then adds its own pairs' results in triangle order.
template <bool CLIP>
__device__ inline void ltc_pairs(const LtcLds &S, LtcLds::Wave &Wl, uint32_t bits, uint32_t jn,
                                 uint32_t lane, double &s0, double &s1, double &s2) {
  const unsigned long long below_me = (1ull << lane) - 1ull;
  uint32_t np = 0;
  for (uint32_t t = 0; t < jn; ++t) {
    const bool mine = (bits >> t) & 1u;
    const unsigned long long mt = __ballot(mine);
    if (mine) Wl.pair[np + __popcll(mt & below_me)] = (uint16_t)((t << 6) | lane);
    if (lane == 0) {
      Wl.mask[t] = mt;
      Wl.base[t] = np;
    }
    np += (uint32_t)__popcll(mt);
  }
  wave_lds_sync();
  for (uint32_t w0 = 0; w0 < np; w0 += 64) {
    {
      const uint32_t e = min(w0 + lane, np - 1);
      const uint32_t pr = Wl.pair[e], t = pr >> 6, P = pr & 63u;
      const d3 xp = dk(Wl.frame[0][P], Wl.frame[1][P], Wl.frame[2][P]);
      const d3 Q0 = dk(Wl.frame[3][P], Wl.frame[4][P], Wl.frame[5][P]);
      const d3 Q1 = dk(Wl.frame[6][P], Wl.frame[7][P], Wl.frame[8][P]);
      const d3 Q2 = dk(Wl.frame[9][P], Wl.frame[10][P], Wl.frame[11][P]);
      d3 e1, e2, e3;
      ltc_vertices(S.tri[t], xp, Q0, Q1, Q2, e1, e2, e3);
      Wl.res[0][lane] = CLIP ? ltc_clipped(e1, e2, e3) : ltc_unclipped(e1, e2, e3);
    }
    wave_lds_sync();
    // this lane's pairs in the window, in triangle order
    while (bits != 0u) {
      const uint32_t t = (uint32_t)__builtin_ctz(bits);
      const uint32_t idx = Wl.base[t] + (uint32_t)__popcll(Wl.mask[t] & below_me);
      if (idx >= w0 + 64) break;
      const double sv = Wl.res[0][idx - w0];
      s0 += S.rad[t][0] * sv;
      s1 += S.rad[t][1] * sv;
      s2 += S.rad[t][2] * sv;
      bits &= bits - 1u;
    }
    wave_lds_sync();  // the window's results are read before the next is written
  }
}

// The same for a chunk whose triangles pair up into quads (triangles 2q,
// 2q + 1 = (a, b, c), (a, c, d), as MESHSOURCE quads are listed): one unit
// per quad with either triangle above the horizon (which: bit 0 for the
// first, bit 1 for the second).  A unit transforms the four corners once and
// evaluates five edges, the diagonal's term serving both triangles (c -> a
// for the first, its negation a -> c for the second); each triangle's sum
// keeps its own absolute value, as in the reference.
__device__ inline void ltc_quad_pairs(const LtcLds &S, LtcLds::Wave &Wl, uint32_t above,
                                      uint32_t nq, uint32_t lane, double &s0, double &s1,
                                      double &s2) {
  const unsigned long long below_me = (1ull << lane) - 1ull;
  uint32_t np = 0;
  for (uint32_t q = 0; q < nq; ++q) {
    const uint32_t which = (above >> (2 * q)) & 3u;
    const unsigned long long mt = __ballot(which != 0u);
    if (which) Wl.pair[np + __popcll(mt & below_me)] = (uint16_t)((q << 8) | (which << 6) | lane);
    if (lane == 0) {
      Wl.mask[q] = mt;
      Wl.base[q] = np;
    }
    np += (uint32_t)__popcll(mt);
  }
  wave_lds_sync();
  uint32_t units = 0;  // this lane's quads with a triangle above
  for (uint32_t q = 0; q < nq; ++q) units |= ((above >> (2 * q)) & 3u) ? 1u << q : 0u;
  for (uint32_t w0 = 0; w0 < np; w0 += 64) {
    {
      const uint32_t e = min(w0 + lane, np - 1);
      const uint32_t pr = Wl.pair[e], q = pr >> 8, P = pr & 63u;
      const d3 xp = dk(Wl.frame[0][P], Wl.frame[1][P], Wl.frame[2][P]);
      const d3 Q0 = dk(Wl.frame[3][P], Wl.frame[4][P], Wl.frame[5][P]);
      const d3 Q1 = dk(Wl.frame[6][P], Wl.frame[7][P], Wl.frame[8][P]);
      const d3 Q2 = dk(Wl.frame[9][P], Wl.frame[10][P], Wl.frame[11][P]);
      d3 ea, eb, ec, ed;
      ltc_vertices(S.tri[2 * q], xp, Q0, Q1, Q2, ea, eb, ec);
      {
        const double *p2 = S.tri[2 * q + 1];
        const d3 dd = dk(p2[6], p2[7], p2[8]) - xp;
        ed = dk(dot(Q0, dd), dot(Q1, dd), dot(Q2, dd));
      }
      const d3 cab = cross(ea, eb), cbc = cross(eb, ec), cca = cross(ec, ea), ccd = cross(ec, ed),
               cda = cross(ed, ea);
      const double Eca = edge64(ec, ea, cca, cca.z);
      Wl.res[0][lane] = fabs(edge64(ea, eb, cab, cab.z) + edge64(eb, ec, cbc, cbc.z) + Eca);
      Wl.res[1][lane] = fabs(-Eca + edge64(ec, ed, ccd, ccd.z) + edge64(ed, ea, cda, cda.z));
    }
    wave_lds_sync();
    // this lane's units in the window, in quad (so triangle) order
    while (units != 0u) {
      const uint32_t q = (uint32_t)__builtin_ctz(units);
      const uint32_t idx = Wl.base[q] + (uint32_t)__popcll(Wl.mask[q] & below_me);
      if (idx >= w0 + 64) break;
      const uint32_t which = (above >> (2 * q)) & 3u;
      if (which & 1u) {
        const double sv = Wl.res[0][idx - w0];
        s0 += S.rad[2 * q][0] * sv;
        s1 += S.rad[2 * q][1] * sv;
        s2 += S.rad[2 * q][2] * sv;
      }
      if (which & 2u) {
        const double sv = Wl.res[1][idx - w0];
        s0 += S.rad[2 * q + 1][0] * sv;
        s1 += S.rad[2 * q + 1][1] * sv;
        s2 += S.rad[2 * q + 1][2] * sv;
      }
      units &= units - 1u;
    }
    wave_lds_sync();  // the window's results are read before the next is written
  }
}

// Light triangles are taken kChunk at a time.  For the LTC part each lane
// classifies the chunk's triangles against its horizon (z of the three
// transformed vertices): above, straddling or below.  The wave then
// evaluates its pairs with a triangle above (three edges), then those with
// a triangle straddling (clipped, four edges), each as a compact list
// (ltc_pairs), so no lane works on a triangle below its horizon and no lane
// waits for the one with the most straddling triangles.  A lane's sums take
// the terms of the triangles above in triangle order, then those of the
// straddling ones, as a loop over all triangles adding 0 for those below
// would.  The diffuse part (ks < 1) evaluates every triangle in each lane
// (tri64<false>).
__global__ __launch_bounds__(kLtcBlock) void k_ltc(const double *__restrict__ tab,
                                                   const float *__restrict__ X,
                                                   const float *__restrict__ N,
                                                   const float *__restrict__ V,
                                                   const float *__restrict__ alpha,
                                                   const float *__restrict__ ks, uint32_t k,
                                                   const float *__restrict__ lt,
                                                   const float *__restrict__ lrad, uint32_t nlt,
                                                   float *__restrict__ rgb) {
  __shared__ LtcLds S;
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  LtcLds::Wave &Wl = S.w[wv];
  const uint32_t i = blockIdx.x * kLtcBlock + threadIdx.x;
  const bool valid = i < k;
  const size_t o = 3 * (size_t)(valid ? i : 0);
  const d3 x = dk(X[o], X[o + 1], X[o + 2]), n = dk(N[o], N[o + 1], N[o + 2]),
           wo = dk(V[o], V[o + 1], V[o + 2]);
  const double kss = ks[valid ? i : 0];
  // M_GGX((float)acos(n.wo), (float)alpha) (ltc.h:18-61, materialTypes.cpp:54)
  const LtcFit F = ltc_fit(tab, (float)acos(dot(n, wo)), (double)alpha[valid ? i : 0]);
  const double *m = F.mi;
  // frame of getBrdfApproxParam (materialTypes.cpp:59-68); rows of Minv * R
  const d3 T1 = normalize(wo - n * dot(n, wo));
  const d3 T2 = normalize(cross(n, T1));
  const d3 R0 = T1 * m[0] + n * m[1], R1 = T2 * m[2], R2 = T1 * m[3] + n * m[4];
  {
    const double fr[12] = {x.x, x.y, x.z, R0.x, R0.y, R0.z, R1.x, R1.y, R1.z, R2.x, R2.y, R2.z};
#pragma unroll
    for (int c = 0; c < 12; ++c) Wl.frame[c][lane] = fr[c];
  }
  double sl0 = 0, sl1 = 0, sl2 = 0, sd0 = 0, sd1 = 0, sd2 = 0;
  const bool do_ltc = valid && kss != 0.0, do_dif = valid && kss != 1.0;
  for (uint32_t j0 = 0; j0 < nlt; j0 += kChunk) {
    const uint32_t jn = min((uint32_t)kChunk, nlt - j0);
    __syncthreads();  // the previous chunk's triangles are no longer read
    for (uint32_t q = threadIdx.x; q < jn * 12; q += kLtcBlock) {
      const uint32_t t = q / 12, c = q % 12;
      if (c < 9)
        S.tri[t][c] = lt[9 * (size_t)(j0 + t) + c];
      else
        S.rad[t][c - 9] = lrad[3 * (size_t)(j0 + t) + c - 9];
    }
    if (threadIdx.x == 0) S.quads = 1u;
    __syncthreads();
    // quads: an even count, each odd triangle (a', c', d) repeating corners
    // a and c of the triangle (a, b, c) before it
    if (threadIdx.x < jn / 2 || (threadIdx.x == 0 && (jn & 1u))) {
      const uint32_t q = threadIdx.x;
      const double *u = S.tri[2 * q], *v = S.tri[(2 * q + 1) % kChunk];
      const bool ok = (jn & 1u) == 0u && v[0] == u[0] && v[1] == u[1] && v[2] == u[2] &&
                      v[3] == u[6] && v[4] == u[7] && v[5] == u[8];
      if (!ok) S.quads = 0u;
    }
    __syncthreads();
    const bool quads = S.quads != 0u;
    uint32_t above = 0, straddle = 0;
    if (do_ltc && quads) {  // the shared corners a, c once per quad
      for (uint32_t q = 0; q < jn / 2; ++q) {
        const double *p = S.tri[2 * q], *p2 = S.tri[2 * q + 1];
        const double za = dot(R2, dk(p[0], p[1], p[2]) - x), zb = dot(R2, dk(p[3], p[4], p[5]) - x),
                     zc = dot(R2, dk(p[6], p[7], p[8]) - x),
                     zd = dot(R2, dk(p2[6], p2[7], p2[8]) - x);
        const bool a1 = za > 0 && zb > 0 && zc > 0, b1 = za < 0 && zb < 0 && zc < 0;
        const bool a2 = za > 0 && zc > 0 && zd > 0, b2 = za < 0 && zc < 0 && zd < 0;
        above |= (a1 ? 1u << (2 * q) : 0u) | (a2 ? 2u << (2 * q) : 0u);
        straddle |= (!a1 && !b1 ? 1u << (2 * q) : 0u) | (!a2 && !b2 ? 2u << (2 * q) : 0u);
      }
    } else if (do_ltc) {
      for (uint32_t t = 0; t < jn; ++t) {
        const double *p = S.tri[t];
        const double z1 = dot(R2, dk(p[0], p[1], p[2]) - x), z2 = dot(R2, dk(p[3], p[4], p[5]) - x),
                     z3 = dot(R2, dk(p[6], p[7], p[8]) - x);
        const bool a = z1 > 0 && z2 > 0 && z3 > 0, b = z1 < 0 && z2 < 0 && z3 < 0;
        above |= a ? 1u << t : 0u;
        straddle |= (!a && !b) ? 1u << t : 0u;
      }
    }
    if (quads)  // block-uniform
      ltc_quad_pairs(S, Wl, above, jn / 2, lane, sl0, sl1, sl2);
    else
      ltc_pairs<false>(S, Wl, above, jn, lane, sl0, sl1, sl2);
    ltc_pairs<true>(S, Wl, straddle, jn, lane, sl0, sl1, sl2);
    if (__any(do_dif)) {  // wave-uniform (tri64 votes across the wave)
      for (uint32_t t = 0; t < jn; ++t) {
        const double *p = S.tri[t];
        const d3 A = dk(p[0], p[1], p[2]), B = dk(p[3], p[4], p[5]), C = dk(p[6], p[7], p[8]);
        const double sv = tri64<false>(A, B, C, n, x);
        if (do_dif) {
          sd0 += S.rad[t][0] * sv;
          sd1 += S.rad[t][1] * sv;
          sd2 += S.rad[t][2] * sv;
        }
      }
    }
  }
  if (!valid) return;
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
