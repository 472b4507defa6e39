/*
 * ORACLE -- test infrastructure only.
 *
 * extern "C" entry points over the reference's OWN compiled code (built by
 * oracle/Makefile from the sources under the reference checkout into
 * oracle/_ref/libchkref.so).  tests/golden/make_golden.py calls them to write
 * the golden fixtures; tests compare the FP64 restatement (cpu_ref.cpp)
 * against them when the library is present.  Nothing here is shipped.
 *
 * Reference entry points used:
 *   BvhAccel::initAccel / intersectT       bvhAccel.cpp:13-22,200-249
 *   BasePrimitive::brdf, MaterialType::pdfEval
 *                                          geometryPrimitives.cpp:7-9,
 *                                          materialTypes.cpp:9-50,89-116
 *   M_GGX                                  ltc.h:18-61 (compiled into materialTypes.o)
 *   LightSource::getLightFromMeshSource_Analytic
 *                                          lightSources_CV.cpp:323-333
 */
#include <stdint.h>
#include <cstring>
#include <vector>

#include "mathPrimitives.h"
#include "geometryPrimitives.h"
#include "materialTypes.h"
#include "bvhAccel.h"
#include "lightSources.h"

// Defined in materialTypes.o (ltc.h is a definition-bearing header).
void M_GGX(const float theta, const float alpha, mat3 &M, mat3 &Minv, double &amplitude);

static MaterialType make_material(int brdf, double ks, double alpha, double pe, double iior,
                                  double eior) {
  if (brdf == CLASSIC_PHONG) return MaterialType(pe, ks);
  return MaterialType(alpha, iior, eior, ks, (BRDFType)brdf);
}

extern "C" {

/* Closest hit through BvhAccel::intersectT.  The original index of each
 * triangle travels in its `reflectance` field (unused by intersect), which
 * undoes the in-place reorder of bvhAccel.cpp:114-121. */
void ref_bvh_closest(const double *tris, uint32_t n, const double *rays, uint32_t m,
                     int32_t *id, double *t, double *N) {
  Triangle *list = new Triangle[n];
  for (uint32_t i = 0; i < n; ++i) {
    const double *p = tris + 9 * (size_t)i;
    list[i] = Triangle(Vec(p[0], p[1], p[2]), Vec(p[3], p[4], p[5]), Vec(p[6], p[7], p[8]),
                       Vec(1, 1, 1), (double)i, MaterialType());
  }
  BvhAccel *acc = new BvhAccel((uint8_t *)list, n, sizeof(Triangle));
  acc->initAccel();
  for (uint32_t r = 0; r < m; ++r) {
    const double *q = rays + 6 * (size_t)r;
    Ray ray(Vec(q[0], q[1], q[2]), Vec(q[3], q[4], q[5]));
    double tt;
    Vec NN;
    int rid;
    bool hit = acc->intersectT(ray, tt, NN, rid);
    id[r] = hit ? (int32_t)list[rid].reflectance : -1;
    t[r] = tt;
    if (N) {
      N[3 * r] = hit ? NN.x : 0;
      N[3 * r + 1] = hit ? NN.y : 0;
      N[3 * r + 2] = hit ? NN.z : 0;
    }
  }
  delete acc;  // leaks its nodes, as the reference does (no destructor)
  delete[] list;
}

/* BvhAccel::intersectS over Sphere primitives (bvhAccel.cpp:151-198,
 * geometryPrimitives.cpp:12-27): closest sphere id (ORIGINAL index, carried
 * in the reflectance field like ref_bvh_closest) and t per ray.
 * spheres: n x 4 (center, radius); rays: m x 6 (o, d). */
void ref_bvh_spheres(const double *spheres, uint32_t n, const double *rays, uint32_t m,
                     int32_t *id, double *t) {
  Sphere *list = new Sphere[n];
  for (uint32_t i = 0; i < n; ++i) {
    const double *p = spheres + 4 * (size_t)i;
    list[i] = Sphere(p[3], Vec(p[0], p[1], p[2]), Vec(1, 1, 1), (double)i, MaterialType());
  }
  BvhAccel *acc = new BvhAccel((uint8_t *)list, n, sizeof(Sphere));
  acc->initAccel();
  for (uint32_t r = 0; r < m; ++r) {
    const double *q = rays + 6 * (size_t)r;
    Ray ray(Vec(q[0], q[1], q[2]), Vec(q[3], q[4], q[5]));
    double tt;
    int rid;
    bool hit = acc->intersectS(ray, tt, rid);
    id[r] = hit ? (int32_t)list[rid].reflectance : -1;
    t[r] = tt;
  }
  delete acc;
  delete[] list;
}

/* shadow() of shader.cpp:88-90 for the triangle BVH: occluded iff the
 * closest hit has t + 1e-5 < dist.  rays: m x 7 (o, d, dist). */
void ref_bvh_occluded(const double *tris, uint32_t n, const double *rays, uint32_t m,
                      uint8_t *occ) {
  Triangle *list = new Triangle[n];
  for (uint32_t i = 0; i < n; ++i) {
    const double *p = tris + 9 * (size_t)i;
    list[i] = Triangle(Vec(p[0], p[1], p[2]), Vec(p[3], p[4], p[5]), Vec(p[6], p[7], p[8]));
  }
  BvhAccel *acc = new BvhAccel((uint8_t *)list, n, sizeof(Triangle));
  acc->initAccel();
  for (uint32_t r = 0; r < m; ++r) {
    const double *q = rays + 7 * (size_t)r;
    Ray ray(Vec(q[0], q[1], q[2]), Vec(q[3], q[4], q[5]));
    double tt;
    Vec NN;
    int rid;
    occ[r] = (acc->intersectT(ray, tt, NN, rid) && tt + 1e-5 < q[6]) ? 1 : 0;
  }
  delete acc;
  delete[] list;
}

/* BasePrimitive::brdf and MaterialType::pdfEval for k configurations. */
void ref_brdf(int brdf, double ks, double alpha, double pe, double iior, double eior,
              double reflectance, uint32_t k, const double *n, const double *wo_ref,
              const double *wo, const double *wi, double *f, double *pdf) {
  MaterialType mt = make_material(brdf, ks, alpha, pe, iior, eior);
  Triangle tri(Vec(0, 0, 0), Vec(1, 0, 0), Vec(0, 1, 0), Vec(1, 1, 1), reflectance, mt);
  for (uint32_t i = 0; i < k; ++i) {
    Vec N(n[3 * i], n[3 * i + 1], n[3 * i + 2]);
    Vec R(wo_ref[3 * i], wo_ref[3 * i + 1], wo_ref[3 * i + 2]);
    Vec O(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
    Vec I(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
    f[i] = tri.brdf(N, R, O, I);
    pdf[i] = mt.pdfEval(N, R, O, I);
  }
}

/* M_GGX: M and Minv as mat3::a (row-major), amplitude, M.det. */
void ref_M_GGX(float theta, float alpha, double *M, double *Minv, double *amp, double *det) {
  mat3 a, b;
  double am;
  M_GGX(theta, alpha, a, b, am);
  std::memcpy(M, a.a, sizeof(a.a));
  std::memcpy(Minv, b.a, sizeof(b.a));
  *amp = am;
  *det = a.det;
}

/* LightSource::getLightFromMeshSource_Analytic for k shading points of one
 * GGX material each (alpha[i], ks[i], IOR 1.5/1) lit by ONE mesh light made
 * of n_tris triangles with per-triangle radiance. */
void ref_ltc_shade(uint32_t k, const double *x, const double *n, const double *v,
                   const double *alpha, const double *ks, const double *tris,
                   const double *radiance, uint32_t n_tris, double *rgb) {
  LightSource ls;
  MeshLight ml;
  for (uint32_t j = 0; j < n_tris; ++j) {
    const double *p = tris + 9 * (size_t)j;
    TriLight tl(Vec(p[0], p[1], p[2]), Vec(p[3], p[4], p[5]), Vec(p[6], p[7], p[8]),
                Vec(1, 1, 1), 1.0);
    tl.radiance = Vec(radiance[3 * j], radiance[3 * j + 1], radiance[3 * j + 2]);
    ml.add(tl);
  }
  ls.addMSource(ml);
  for (uint32_t i = 0; i < k; ++i) {
    MaterialType mt(alpha[i], 1.5, 1.0, ks[i], GGX);
    Triangle prim(Vec(0, 0, 0), Vec(1, 0, 0), Vec(0, 1, 0), Vec(1, 1, 1), 1.0, mt);
    Vec X(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    Vec N(n[3 * i], n[3 * i + 1], n[3 * i + 2]);
    Vec V(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    Ray r(X + V, V * -1.0);
    Vec G = ls.getLightFromMeshSource_Analytic(r, N, X, &prim);
    rgb[3 * i] = G.x;
    rgb[3 * i + 1] = G.y;
    rgb[3 * i + 2] = G.z;
  }
}

}  // extern "C"
