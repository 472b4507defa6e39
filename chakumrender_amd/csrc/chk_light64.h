// FP64 light triangles: the per-triangle constants of Triangle::intersect
// (geometryPrimitives.cpp:29-71) computed once at scene creation, and the
// test itself.  Host and device code (the scene builder fills the records;
// a CPU test compares the test with the oracle's restatement).
#pragma once

#include <cmath>

#include "chk_device.h"
#include "chk_internal.h"

namespace chk {

constexpr double kInf64 = 1e20;  // mathPrimitives.h:5

// The constants of a light triangle with FP32 corners p[9] (A, B, C): A,
// v1 = B - A, v2 = C - A, the unit normal (C-A)x(B-A) normalised as
// Vec::norm (v * (1 / sqrt(v.v))), b0 = v1.v1, b1 = v1.v2 (= c0), c1 =
// v2.v2, den = c0 b1 - c1 b0, each as the reference forms it.
__host__ inline LightTri64 make_light64(const float *p) {
  LightTri64 Q;
  double v1[3], v2[3];
  for (int c = 0; c < 3; ++c) {
    Q.a[c] = p[c];
    v1[c] = Q.v1[c] = (double)p[3 + c] - p[c];
    v2[c] = Q.v2[c] = (double)p[6 + c] - p[c];
  }
  const double nx = v2[1] * v1[2] - v2[2] * v1[1], ny = v2[2] * v1[0] - v2[0] * v1[2],
               nz = v2[0] * v1[1] - v2[1] * v1[0];
  const double inv = 1.0 / std::sqrt(nx * nx + ny * ny + nz * nz);
  Q.n[0] = nx * inv;
  Q.n[1] = ny * inv;
  Q.n[2] = nz * inv;
  Q.b0 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
  Q.b1 = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
  Q.c1 = v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2];
  Q.den = Q.b1 * Q.b1 - Q.c1 * Q.b0;
  return Q;
}

// Triangle::intersect (geometryPrimitives.cpp:29-71) in FP64 with the
// triangle's constants precomputed (LightTri64): the reference's operations
// without recomputing the normal and the edge dot products per call.
// A miss is decided without the divisions whose quotient only the miss
// tests read, where the sign or size of the numerator already decides them
// (the results are those of the reference's divide-then-compare):
//  * t = tn / rn with rn < 0: t < 0 exactly when tn > 0 (|rn| <= 1 for a
//    unit direction and unit normal, so a positive tn never underflows to
//    -0 in the quotient);
//  * gamma = gn / den with den < 0 (b1^2 - c1 b0, Cauchy-Schwarz): gamma < 0
//    exactly when gn > 0, and gamma > 1 when gn < den (1 + 2^-40) (the
//    quotient then exceeds 1 + 2^-41, which rounds above 1);
//  * beta = bn / b0 with b0 > 0: beta < 0 exactly when bn < 0, beta > 1 when
//    bn > b0 (1 + 2^-40).
// Numerators within 1e-280 of zero take the divisions (no underflow
// question), as do degenerate triangles (den >= 0).
__host__ __device__ inline double tri_isect64p(const LightTri64 &T, d3 o, d3 d, d3 *N) {
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
  const double tn = dot(A - o, Nf);
  if (tn > 1e-280) return kInf64;  // t < 0
  const double t = tn / rn;
  if (t < 0) return kInf64;
  const d3 v0 = (o + d * t) - A;
  const double k0 = dot(v0, dk(T.v1[0], T.v1[1], T.v1[2]));
  const double k1 = dot(v0, dk(T.v2[0], T.v2[1], T.v2[2]));
  const double gn = k0 * T.b1 - k1 * T.b0;
  const bool fast = T.den < 0;
  constexpr double kOver = 1.0 + 0x1p-40;
  if (fast && (gn > 1e-280 || gn < T.den * kOver)) return kInf64;  // gamma < 0 or > 1
  const double gamma = gn / T.den;
  const double bn = k0 - gamma * T.b1;  // c0 = v2.v1 = b1
  if (fast && (bn < -1e-280 || bn > T.b0 * kOver)) return kInf64;  // beta < 0 or > 1
  const double beta = bn / T.b0;
  if (gamma + beta > 1) return kInf64;
  if (0 > gamma || gamma > 1) return kInf64;
  if (0 > beta || beta > 1) return kInf64;
  return t;
}

}  // namespace chk
