// CPU check of chk::tri_isect64p (chk_light64.h), the FP64 light-triangle test
// of the direct-lighting kernels, whose misses skip divisions: on random and
// edge-grazing rays its result (t, normal) must equal
//  * the divide-then-compare form of Triangle::intersect
//    (geometryPrimitives.cpp:29-71) on the same precomputed constants, and
//  * with WITH_ORACLE, the oracle's restatement orc_tri_intersect, which
//    recomputes the constants per call as the reference does,
// bit for bit.  Built host-only by tests/test_light64_misses.py.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "chk_light64.h"

#ifdef WITH_ORACLE
extern "C" double orc_tri_intersect(const double *A, const double *B, const double *C,
                                    const double *o, const double *d, double *N);
#endif

using namespace chk;

// Triangle::intersect with the quotients formed before any miss test.
static double divide_then_compare(const LightTri64 &T, d3 o, d3 d, d3 *N) {
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
  const double beta = (k0 - gamma * T.b1) / T.b0;
  if (gamma + beta > 1) return kInf64;
  if (0 > gamma || gamma > 1) return kInf64;
  if (0 > beta || beta > 1) return kInf64;
  return t;
}

static bool bits_eq(double a, double b) { return std::memcmp(&a, &b, 8) == 0; }

int main(int argc, char **argv) {
  const long n_tris = argc > 1 ? std::atol(argv[1]) : 4000;
  const int per_tri = 250;
  std::mt19937_64 g(0x5EED);
  std::uniform_real_distribution<double> U(-1, 1);
  long cases = 0, hits = 0, bad = 0, bad_oracle = 0, tests_oracle = 0;
  for (long k = 0; k < n_tris; ++k) {
    float p[9];
    const double sc = std::pow(10.0, 2 * U(g));
    for (int c = 0; c < 9; ++c) p[c] = (float)(U(g) * sc + (c % 3 == 1 ? 3.0 : 0.0));
    const LightTri64 Q = make_light64(p);
    double A[3], B[3], C[3];
    for (int c = 0; c < 3; ++c) A[c] = p[c], B[c] = p[3 + c], C[c] = p[6 + c];
    for (int r = 0; r < per_tri; ++r) {
      const d3 o = dk(3 * sc * U(g), 3 * sc * U(g) + 3, 3 * sc * U(g));
      double u = std::fabs(U(g)), v = std::fabs(U(g));
      const int kind = r % 5;
      if (kind == 0) {
        if (u + v > 1) u = 1 - u, v = 1 - v;  // inside
      } else if (kind == 1) {
        v = 1 - u;  // on the BC edge
      } else if (kind == 2) {
        v = 0;  // on the AB edge
      } else if (kind == 3) {
        u = 0;  // on the AC edge
      } else {
        u = 2 * U(g), v = 2 * U(g);  // anywhere in the plane
      }
      if (kind >= 1 && kind <= 3) {  // a few ulps to either side of the edge
        for (int s = 0; s < 1 + r % 4; ++s) u = std::nextafter(u, (r & 1) ? 2.0 : -2.0);
      }
      d3 tgt;
      tgt.x = A[0] + (B[0] - A[0]) * v + (C[0] - A[0]) * u;
      tgt.y = A[1] + (B[1] - A[1]) * v + (C[1] - A[1]) * u;
      tgt.z = A[2] + (B[2] - A[2]) * v + (C[2] - A[2]) * u;
      d3 d = tgt - o;
      d = d * (1.0 / std::sqrt(dot(d, d)));
      if (r % 7 == 0) d = d * -1.0;  // away from the plane
      d3 N1, N2;
      const double t1 = tri_isect64p(Q, o, d, &N1);
      const double t2 = divide_then_compare(Q, o, d, &N2);
      ++cases;
      if (t1 < kInf64) ++hits;
      if (!bits_eq(t1, t2) || (t1 < kInf64 && (!bits_eq(N1.x, N2.x) || !bits_eq(N1.y, N2.y) ||
                                            !bits_eq(N1.z, N2.z)))) {
        if (bad < 5) std::printf("mismatch: %.17g vs %.17g\n", t1, t2);
        ++bad;
      }
#ifdef WITH_ORACLE
      double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z}, N3[3];
      const double t3 = orc_tri_intersect(A, B, C, oo, dd, N3);
      ++tests_oracle;
      if (!bits_eq(t1, t3) ||
          (t1 < kInf64 && (!bits_eq(N1.x, N3[0]) || !bits_eq(N1.y, N3[1]) || !bits_eq(N1.z, N3[2])))) {
        if (bad_oracle < 5) std::printf("oracle mismatch: %.17g vs %.17g\n", t1, t3);
        ++bad_oracle;
      }
#endif
    }
  }
  std::printf("cases %ld hits %ld mismatches %ld oracle_cases %ld oracle_mismatches %ld\n", cases,
              hits, bad, tests_oracle, bad_oracle);
  return (bad || bad_oracle) ? 1 : 0;
}
