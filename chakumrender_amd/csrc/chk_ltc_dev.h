// FP64 device pieces of the reference's polygon-light integrals and of
// its LTC fit, shared by the analytic-shading kernel (chk_ltc.hip) and the
// direct-lighting estimators (chk_direct.hip).
//
//   integrateEdge            lightSources_CV.cpp:8-25
//   clipTriangle (2 forms)   lightSources_CV.cpp:28-126, 129-219
//   per-triangle sums        lightSources_CV.cpp:241-262, 293-314
//   M_GGX                    ltc.h:18-61
#pragma once

#include "chk_device.h"

namespace chk {

constexpr double kPi64 = 3.14159265359;  // mathPrimitives.h:4

// Packed LTC record (chk_set_ltc_table): Minv a[0], a[2], a[4], a[6], a[8],
// amplitude, M a[0], a[2], a[4], a[6], a[8], det M.  The other four entries
// of both matrices are zero in the fit.
constexpr int kLtcRec = 12;

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

// Points of the clipped polygon, written in pairs at positions 0-1 then 2-3
// (a third pair is dropped); named registers rather than an array indexed
// at run time, which would live in scratch memory.
__device__ inline d3 sel3(bool c, d3 a, d3 b) {
  return dk(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z);
}
struct Clip4 {
  d3 q0, q1, q2, q3;
  int idx;
  __device__ void put(d3 X, d3 Y) {
    const bool a = idx == 0, b = idx == 2;
    q0 = sel3(a, X, q0);
    q1 = sel3(a, Y, q1);
    q2 = sel3(b, X, q2);
    q3 = sel3(b, Y, q3);
    idx += 2;
  }
};

// clipTriangle (lightSources_CV.cpp:28-126 and 129-219): the 4-point
// polygon of the part of ABC above the plane; false if it has none.
// nA, nB, nC: signs of the vertices (only their sign is read).
template <bool ZPLANE>
__device__ bool clip64(d3 A, d3 B, d3 C, d3 n, d3 P, double nA, double nB, double nC,
                       Clip4 &clip) {
  const double eps = 1e-15;
  int cnt = 0;
  const d3 dAB = A - B, dBC = B - C, dCA = C - A;
  const double nAB = ndot<ZPLANE>(n, dAB), nBC = ndot<ZPLANE>(n, dBC),
               nCA = ndot<ZPLANE>(n, dCA);
  const double pA = ndot<ZPLANE>(n, P - A), pB = ndot<ZPLANE>(n, P - B),
               pC = ndot<ZPLANE>(n, P - C);
  if (nAB * nAB > eps) {
    const d3 I = dAB * (pA / nAB) + A;
    if (dot(I - A, B - I) > 0) {
      ++cnt;
      if (nA < 0) clip.put(I, B);
      else clip.put(A, I);
    }
  } else if (pA * pA < eps) {
    return false;
  }
  if (nBC * nBC > eps) {
    const d3 I = dBC * (pB / nBC) + B;
    if (dot(I - B, C - I) > 0) {
      ++cnt;
      if (nB < 0) clip.put(I, C);
      else clip.put(B, I);
    }
  } else if (pB * pB < eps) {
    return false;
  }
  if (nCA * nCA > eps) {
    const d3 I = dCA * (pC / nCA) + C;
    // the reference tests only the x components on this edge
    if ((I.x - C.x) * (A.x - I.x) > 0) {
      ++cnt;
      if (nC < 0) clip.put(I, A);
      else clip.put(C, I);
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
    Clip4 cl;
    cl.q0 = cl.q1 = cl.q2 = cl.q3 = dk(0, 0, 0);
    cl.idx = 0;
    if (!clip64<ZPLANE>(A, B, C, n, P, nA, nB, nC, cl)) return 0.0;
    const d3 q1 = cl.q0 - P, q2 = cl.q1 - P, q3 = cl.q2 - P, q4 = cl.q3 - P;
    const d3 c1 = cross(q1, q2), c2 = cross(q2, q3), c3 = cross(q3, q4), c4 = cross(q4, q1);
    sum = edge64(q1, q2, c1, ndot<ZPLANE>(n, c1)) + edge64(q2, q3, c2, ndot<ZPLANE>(n, c2)) +
          edge64(q3, q4, c3, ndot<ZPLANE>(n, c3)) + edge64(q4, q1, c4, ndot<ZPLANE>(n, c4));
  }
  return fabs(sum);
}


// M_GGX(theta, alpha) (ltc.h:18-61): bins from the float theta and
// sqrt(alpha) in double, lerp in theta with a float weight; det M stays the
// floor bin's (the lerp does not recompute mat3::det).
struct LtcFit {
  double mi[5];  // Minv a[0], a[2], a[4], a[6], a[8]
  double amp;
  double m[5];   // M a[0], a[2], a[4], a[6], a[8]
  double det;
};
__device__ inline LtcFit ltc_fit(const double *__restrict__ tab, float theta, double alpha) {
  const double tq = (double)theta / (0.5 * kPi64) * 64.0;
  const int tb = max(0, min(63, (int)floor(tq)));
  const int ab = max(0, min(63, (int)floor(sqrt(alpha) * 64.0)));
  const int tp = min(63, tb + 1);
  const double *r0 = tab + kLtcRec * (ab + tb * 64);
  double v[11];
#pragma unroll
  for (int c = 0; c < 11; ++c) v[c] = r0[c];
  if (tb != tp) {
    const float mult = (float)(tq - (double)(float)tb);
    const double *r1 = tab + kLtcRec * (ab + tp * 64);
#pragma unroll
    for (int c = 0; c < 11; ++c) v[c] = v[c] + (r1[c] - v[c]) * (double)mult;
  }
  LtcFit f;
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    f.mi[c] = v[c];
    f.m[c] = v[6 + c];
  }
  f.amp = v[5];
  f.det = r0[11];
  return f;
}

}  // namespace chk
