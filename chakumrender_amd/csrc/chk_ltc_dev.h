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

// 1/x (normal x of either sign) and 1/sqrt(x) (positive normal x): the hardware approximation
// (v_rcp_f64 / v_rsq_f64, relative error below 2^-24 on gfx950) and one
// Newton step (relative error below 5e-15, measured over 2^20 arguments
// spanning 2^-30..2^31), cheaper than a divide or sqrt.
__device__ inline double rcp64(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(fma(-x, r, 1.0), r, r);
}
__device__ inline double rsq64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  return fma(0.5 * y, fma(-x * y, y, 1.0), y);
}

// atan2(y, x) for y >= 0 (an angle in [0, pi]): t = min/max of (y, |x|) in
// [0, 1], atan(t) = t P(t^2) with P the degree-14 Chebyshev fit of
// atan(sqrt(w))/sqrt(w) on [0, 1] (relative error 2.9e-13 in FP64;
// gen_atan_poly.py 14), then the octant.  Branch free.  The edge terms of a
// triangle seen nearly edge-on cancel by up to ~5 orders of magnitude, so
// the fit keeps 13 digits (degree 12, at 1.1e-11, moved such points' FP32
// results by 4e-7 relative).
__device__ inline double atan2_upper(double y, double x) {
  const double ax = fabs(x);
  const double mx = fmax(y, ax), mn = fmin(y, ax);
  const double t = mn * rcp64(fmax(mx, 1e-300));  // 0 for x = y = 0
  const double w = t * t;
  double p = 0.0001704604046618824;
  p = fma(p, w, -0.0014666686334116588);
  p = fma(p, w, 0.005936376596452122);
  p = fma(p, w, -0.01523249190907656);
  p = fma(p, w, 0.028410612112699588);
  p = fma(p, w, -0.042324616793255394);
  p = fma(p, w, 0.054582394542429476);
  p = fma(p, w, -0.06536228811849373);
  p = fma(p, w, 0.07663177357754078);
  p = fma(p, w, -0.09086360660727331);
  p = fma(p, w, 0.11110640790649381);
  p = fma(p, w, -0.14285684661254053);
  p = fma(p, w, 0.1999999901342103);
  p = fma(p, w, -0.3333333332025471);
  p = fma(p, w, 0.99999999999971);
  double a = t * p;
  a = y > ax ? 1.5707963267948966 - a : a;
  return x < 0.0 ? 3.141592653589793 - a : a;
}

// integrateEdge (lightSources_CV.cpp:8-25): (v1 x v2).n * theta/sin(theta)
// with v_i = p_i/|p_i|, from the unnormalised p1, p2, c = p1 x p2 and
// cn = c.n: (v1 x v2).n = cn/(|p1||p2|) and sin(theta) = |c|/(|p1||p2|), so
// the term is cn * theta/|c| with theta = atan2(|c|, p1.p2) (the
// reference's acos(v1.v2), without its loss of precision near 0).  For
// theta <= 0.001 the reference uses the limit cn/(|p1||p2|), which is
// cn (theta/|c|) (sin(theta)/theta), here with the series 1 - theta^2/6
// + theta^4/120 (exact to 1e-22 there).  A zero edge (c = 0, so cn = 0)
// gives 0: |c|^2 is floored at 1e-300, so every factor stays finite.
__device__ inline double edge64(d3 p1, d3 p2, d3 c, double cn) {
  const double cc = dot(c, c);
  const double rs = rsq64(fmax(cc, 1e-300));
  const double th = atan2_upper(cc * rs, dot(p1, p2));
  const double t2 = th * th;
  const double sr = fma(t2, fma(t2, 1.0 / 120.0, -1.0 / 6.0), 1.0);
  return cn * ((th * rs) * (th > 0.001 ? 1.0 : sr));
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
  __device__ void put(bool on, d3 X, d3 Y) {
    const bool a = on && idx == 0, b = on && idx == 2;
    q0 = sel3(a, X, q0);
    q1 = sel3(a, Y, q1);
    q2 = sel3(b, X, q2);
    q3 = sel3(b, Y, q3);
    idx += on ? 2 : 0;
  }
};

// clipTriangle (lightSources_CV.cpp:28-126 and 129-219): the 4-point
// polygon of the part of ABC above the plane; false if it has none.
// nA, nB, nC: signs of the vertices (only their sign is read).  Evaluated
// without branches (the reference's early returns become a flag), so the
// lanes of a wave that clip and those that do not stay converged.
template <bool ZPLANE>
__device__ bool clip64(d3 A, d3 B, d3 C, d3 n, d3 P, double nA, double nB, double nC,
                       Clip4 &clip) {
  const double eps = 1e-15;
  int cnt = 0;
  bool dead = false;
  const d3 dAB = A - B, dBC = B - C, dCA = C - A;
  const double nAB = ndot<ZPLANE>(n, dAB), nBC = ndot<ZPLANE>(n, dBC),
               nCA = ndot<ZPLANE>(n, dCA);
  const double pA = ndot<ZPLANE>(n, P - A), pB = ndot<ZPLANE>(n, P - B),
               pC = ndot<ZPLANE>(n, P - C);
  {
    const bool cuts = nAB * nAB > eps;
    const d3 I = dAB * (pA * rcp64(cuts ? nAB : 1.0)) + A;
    const bool hit = cuts && dot(I - A, B - I) > 0;
    dead = dead || (!cuts && pA * pA < eps);
    cnt += hit ? 1 : 0;
    clip.put(hit, sel3(nA < 0, I, A), sel3(nA < 0, B, I));
  }
  {
    const bool cuts = nBC * nBC > eps;
    const d3 I = dBC * (pB * rcp64(cuts ? nBC : 1.0)) + B;
    const bool hit = cuts && dot(I - B, C - I) > 0;
    dead = dead || (!cuts && pB * pB < eps);
    cnt += hit ? 1 : 0;
    clip.put(hit, sel3(nB < 0, I, B), sel3(nB < 0, C, I));
  }
  {
    const bool cuts = nCA * nCA > eps;
    const d3 I = dCA * (pC * rcp64(cuts ? nCA : 1.0)) + C;
    // the reference tests only the x components on this edge
    const bool hit = cuts && (I.x - C.x) * (A.x - I.x) > 0;
    dead = dead || (!cuts && pC * pC < eps);
    cnt += hit ? 1 : 0;
    clip.put(hit, sel3(nC < 0, I, C), sel3(nC < 0, A, I));
  }
  return cnt > 1 && !dead;
}

// |sum of edge terms| of one triangle (lightSources_CV.cpp:241-262 and
// 293-314).  ZPLANE: LTC frame (P = 0, n = z).  Every lane evaluates one
// 4-point polygon: the clipped one, or ABC with C repeated (a zero edge, so
// the sum is the reference's three-edge sum in its order); the clip itself
// runs only when some lane of the wave straddles the plane.
template <bool ZPLANE>
__device__ double tri64(d3 A, d3 B, d3 C, d3 n, d3 P) {
  const d3 p1 = A - P, p2 = B - P, p3 = C - P;
  // signs of n.(V - P)/|V - P|, as the reference tests them
  const double nA = ndot<ZPLANE>(n, p1), nB = ndot<ZPLANE>(n, p2), nC = ndot<ZPLANE>(n, p3);
  const bool above = nA > 0 && nB > 0 && nC > 0;
  const bool below = nA < 0 && nB < 0 && nC < 0;
  Clip4 cl;
  cl.q0 = cl.q1 = cl.q2 = cl.q3 = dk(0, 0, 0);
  cl.idx = 0;
  bool clipped = false;
  if (__any(!above && !below))  // wave-uniform
    clipped = clip64<ZPLANE>(A, B, C, n, P, nA, nB, nC, cl);
  const d3 q1 = sel3(above, p1, cl.q0 - P), q2 = sel3(above, p2, cl.q1 - P),
           q3 = sel3(above, p3, cl.q2 - P), q4 = sel3(above, p3, cl.q3 - P);
  const d3 c1 = cross(q1, q2), c2 = cross(q2, q3), c3 = cross(q3, q4), c4 = cross(q4, q1);
  const double sum = edge64(q1, q2, c1, ndot<ZPLANE>(n, c1)) + edge64(q2, q3, c2, ndot<ZPLANE>(n, c2)) +
                     edge64(q3, q4, c3, ndot<ZPLANE>(n, c3)) + edge64(q4, q1, c4, ndot<ZPLANE>(n, c4));
  return (above || (!below && clipped)) ? fabs(sum) : 0.0;
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
