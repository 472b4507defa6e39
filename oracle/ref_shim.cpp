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
 *   LightSource::getLightFromMeshSource_AnalyticCV
 *                                          lightSources_CV.cpp:335-397
 *   LightSource::getLightFromAllSources_MIS
 *                                          lightSources.cpp:32-109
 *   MIS::getSamples                        mis.cpp:97-223
 *   shadow()                               shader.cpp:76-97
 *   loadAccels, loadLightSampler           objects.cpp:38-56, mis.cpp:225-234
 *   initSceneParser, prepareScreenCamera, prepareObjects, prepareLightSources
 *                                          prepareScene.cpp:1-281 (with sceneParser.cpp
 *                                          and the vendored tinyxml2)
 */
#include <stdint.h>
#include <cstring>
#include <vector>

#include "mathPrimitives.h"
#include "geometryPrimitives.h"
#include "materialTypes.h"
#include "bvhAccel.h"
#include "lightSources.h"
#include "mis.h"
#include "objects.h"
#include "prepareScene.h"
#include "ppm.h"
#include <unistd.h>

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

/* The reference's scene state for its direct-lighting estimators, set up
 * the way its front end does (prepareScene.cpp:54-241 then loadAccels() and
 * loadLightSampler()): scene triangles in vTriangleList, each mesh light as
 * a MeshLight of TriLights (LightSource::addMSource also appends them to
 * vTriangleList), triangle lights by addTSource, sphere objects in
 * vSphereList, sphere sources by addSSource (which appends the source's
 * sphere), point sources by addPSource, and the dummy triangle.
 *   tris    n_tris x 9 doubles; tri_mat: material per triangle
 *   md      n_mats x 14: brdf, ks, alpha, phong_exp, int_ior, ext_ior,
 *           reflectance, colour[3], radiance[3], light
 *   lights  n_lights x 2: first triangle, triangle count (emitters)
 *   tl      n_tl triangle indices (emitters)
 *   sph     n_sph x 4: centre, radius; sph_mat: material per sphere
 *   ssrc    n_ss sphere indices (emitters)
 *   pts     n_pts x 6: position, power (the PointSource radiance field)
 * The primitive of a shading point is an input triangle index, or n_tris +
 * sphere index. */
static std::vector<Triangle> g_prims;  // the BasePrimitive of every input triangle
static std::vector<Sphere> g_sph;      // ... and of every input sphere
static uint32_t g_ntris = 0;

static BasePrimitive *prim_ptr(uint32_t prim) {
  return prim < g_ntris ? (BasePrimitive *)&g_prims[prim] : (BasePrimitive *)&g_sph[prim - g_ntris];
}

void ref_direct_setup_ex(const double *tris, uint32_t n_tris, const uint32_t *tri_mat,
                         const double *md, uint32_t n_mats, const uint32_t *lights,
                         uint32_t n_lights, const uint32_t *tl, uint32_t n_tl, const double *sph,
                         const uint32_t *sph_mat, uint32_t n_sph, const uint32_t *ssrc,
                         uint32_t n_ss, const double *pts, uint32_t n_pts) {
  (void)n_mats;
  vTriangleList.clear();
  vSphereList.clear();
  g_ntris = n_tris;
  std::vector<int> is_light(n_tris, 0);
  for (uint32_t j = 0; j < n_lights; ++j)
    for (uint32_t i = 0; i < lights[2 * j + 1]; ++i) is_light[lights[2 * j] + i] = 1;
  for (uint32_t j = 0; j < n_tl; ++j) is_light[tl[j]] = 1;
  g_prims.assign(n_tris, Triangle());
  for (uint32_t o = 0; o < n_tris; ++o) {
    const double *p = tris + 9 * (size_t)o, *m = md + 14 * (size_t)tri_mat[o];
    MaterialType mt = make_material((int)m[0], m[1], m[2], m[3], m[4], m[5]);
    g_prims[o] = Triangle(Vec(p[0], p[1], p[2]), Vec(p[3], p[4], p[5]), Vec(p[6], p[7], p[8]),
                          Vec(m[7], m[8], m[9]), m[6], mt);
    if (!is_light[o]) vTriangleList.push_back(g_prims[o]);
  }
  std::vector<int> is_src(n_sph, 0);
  for (uint32_t j = 0; j < n_ss; ++j) is_src[ssrc[j]] = 1;
  g_sph.assign(n_sph, Sphere());
  for (uint32_t j = 0; j < n_sph; ++j) {
    const double *q = sph + 4 * (size_t)j, *m = md + 14 * (size_t)sph_mat[j];
    MaterialType mt = make_material((int)m[0], m[1], m[2], m[3], m[4], m[5]);
    g_sph[j] = Sphere(q[3], Vec(q[0], q[1], q[2]), Vec(m[7], m[8], m[9]), m[6], mt);
    if (!is_src[j]) vSphereList.push_back(g_sph[j]);
  }
  lSource = new LightSource();
  for (uint32_t j = 0; j < n_lights; ++j) {
    MeshLight ml;
    for (uint32_t i = 0; i < lights[2 * j + 1]; ++i) {
      const uint32_t o = lights[2 * j] + i;
      const double *p = tris + 9 * (size_t)o, *m = md + 14 * (size_t)tri_mat[o];
      TriLight t(Vec(p[0], p[1], p[2]), Vec(p[3], p[4], p[5]), Vec(p[6], p[7], p[8]),
                 Vec(1, 1, 1), 1.0);
      t.radiance = Vec(m[10], m[11], m[12]);
      ml.add(t);
    }
    ml.initMeshLight();
    lSource->addMSource(ml);
  }
  for (uint32_t j = 0; j < n_tl; ++j) {
    const uint32_t o = tl[j];
    const double *p = tris + 9 * (size_t)o, *m = md + 14 * (size_t)tri_mat[o];
    TriLight t(Vec(p[0], p[1], p[2]), Vec(p[3], p[4], p[5]), Vec(p[6], p[7], p[8]), Vec(1, 1, 1),
               1.0);
    t.radiance = Vec(m[10], m[11], m[12]);
    t.t.m = MaterialType(PLANAR, t.radiance);
    lSource->addTSource(t);
  }
  for (uint32_t j = 0; j < n_ss; ++j) {
    const double *q = sph + 4 * (size_t)ssrc[j], *m = md + 14 * (size_t)sph_mat[ssrc[j]];
    SphereSource src(q[3], Vec(q[0], q[1], q[2]), Vec(1, 1, 1), 1.0);
    src.radiance = Vec(m[10], m[11], m[12]);
    src.p = Sphere(q[3], Vec(q[0], q[1], q[2]), Vec(0, 0, 0), 0, MaterialType(VOLUME, src.radiance));
    lSource->addSSource(src);
  }
  for (uint32_t j = 0; j < n_pts; ++j) {
    const double *q = pts + 6 * (size_t)j;
    PointSource ps(Vec(q[0], q[1], q[2]), Vec(1, 1, 1), 1.0);
    ps.radiance = Vec(q[3], q[4], q[5]);
    lSource->addPSource(ps);
  }
  // prepareScene.cpp:241
  vTriangleList.push_back(Triangle(1e6, 1e6, 1e6, Vec(0.0, 0, 0.9), 0.9, MaterialType(1.0, 0.5)));
  loadAccels();
  loadLightSampler();
}

/* ref_direct_setup_ex with mesh lights only and the dummy sphere of
 * prepareScene.cpp:238. */
void ref_direct_setup(const double *tris, uint32_t n_tris, const uint32_t *tri_mat,
                      const double *md, uint32_t n_mats, const uint32_t *lights,
                      uint32_t n_lights) {
  std::vector<double> m(md, md + 14 * (size_t)n_mats);
  // MaterialType(200.2, 1): classic Phong, exponent 200.2, ks 1, colour .999
  const double dm[14] = {1, 1, 0, 200.2, 0, 0, 1.0, .999, .999, .999, 0, 0, 0, 0};
  m.insert(m.end(), dm, dm + 14);
  const double dummy[4] = {27, 1e5, 47, 1};
  const uint32_t dmat = n_mats;
  ref_direct_setup_ex(tris, n_tris, tri_mat, m.data(), n_mats + 1, lights, n_lights, NULL, 0,
                      dummy, &dmat, 1, NULL, 0, NULL, 0);
}

/* An EnvSource (lightSources.h:33-109) added to the scene of the last
 * ref_direct_setup(_ex) call, the way prepareLightSources would
 * (prepareScene.cpp:150-154, commented out there): EnvSource(image, w, h,
 * axis, col, intensity) runs initSamples (lightSources.cpp:662-782), then
 * loadLightSampler() hands it to MIS.  image: h x w x 3 doubles. */
static EnvSource *g_env = NULL;
void ref_env_setup(const double *image, uint32_t w, uint32_t h, uint32_t axis, const double *col,
                   double intensity) {
  std::vector<Vec> img((size_t)w * h);
  for (size_t i = 0; i < img.size(); ++i) img[i] = Vec(image[3 * i], image[3 * i + 1], image[3 * i + 2]);
  g_env = new EnvSource(&img[0], w, h, axis, Vec(col[0], col[1], col[2]), intensity);
  lSource->addESource(*g_env);
  loadLightSampler();
}
double ref_env_area() { return lSource->getEnvSourcePtr()->normalizationArea; }
/* EnvSource::getRadiance for k unit directions (k x 3 in, k x 3 out). */
void ref_env_radiance(uint32_t k, const double *dirs, double *out) {
  EnvSource *e = lSource->getEnvSourcePtr();
  for (uint32_t i = 0; i < k; ++i) {
    Vec L = e->getRadiance(Vec(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
    out[3 * i] = L.x;
    out[3 * i + 1] = L.y;
    out[3 * i + 2] = L.z;
  }
}
/* The first n of initSamples' importance samples (n x 3). */
void ref_env_samples(uint32_t n, double *out) {
  EnvSource *e = lSource->getEnvSourcePtr();
  for (uint32_t i = 0; i < n && i < EnvSource::nSamples * EnvSource::multiplier; ++i) {
    out[3 * i] = e->impSamples[i].x;
    out[3 * i + 1] = e->impSamples[i].y;
    out[3 * i + 2] = e->impSamples[i].z;
  }
}

/* readImage / writeImage (ppm.cpp:36-152) in directory `dir`, where they
 * open their fixed file names (env1.ppm, env1_cpy.ppm).  ref_read_ppm
 * returns readImage's status, the size in wh and up to cap pixels. */
int ref_read_ppm(const char *dir, double *img, uint32_t cap, uint32_t *wh) {
  char cwd[4096];
  if (!getcwd(cwd, sizeof cwd) || chdir(dir) != 0) return -1;
  Vec *image = NULL;
  uint32_t w = 0, h = 0;
  const int ok = readImage(image, w, h);
  if (chdir(cwd) != 0) return -1;
  if (!ok) return 0;
  wh[0] = w;
  wh[1] = h;
  for (uint32_t i = 0; i < w * h && i < cap; ++i) {
    img[3 * i] = image[i].x;
    img[3 * i + 1] = image[i].y;
    img[3 * i + 2] = image[i].z;
  }
  delete[] image;
  return 1;
}
int ref_write_ppm(const char *dir, const double *img, uint32_t w, uint32_t h) {
  char cwd[4096];
  if (!getcwd(cwd, sizeof cwd) || chdir(dir) != 0) return -1;
  std::vector<Vec> image((size_t)w * h);
  for (size_t i = 0; i < image.size(); ++i) image[i] = Vec(img[3 * i], img[3 * i + 1], img[3 * i + 2]);
  const int ok = writeImage(&image[0], w, h);
  if (chdir(cwd) != 0) return -1;
  return ok;
}

/* MIS::getSamples for k points (primitive = the point's input triangle):
 * samples (k x 200 x 3) and weights (k x 200); returns the sample count. */
uint32_t ref_mis_get_samples(uint32_t k, const double *x, const double *n, const double *wo_ref,
                             const double *wo, const uint32_t *prim, double *samples,
                             double *weights) {
  uint32_t S = 0;
  for (uint32_t i = 0; i < k; ++i) {
    Vec smp[200];
    double w[200] = {0};
    S = lightSampler->getSamples(Vec(x[3 * i], x[3 * i + 1], x[3 * i + 2]),
                                 Vec(n[3 * i], n[3 * i + 1], n[3 * i + 2]),
                                 Vec(wo_ref[3 * i], wo_ref[3 * i + 1], wo_ref[3 * i + 2]),
                                 Vec(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), prim_ptr(prim[i]),
                                 smp, w);
    for (uint32_t q = 0; q < S && q < 200; ++q) {
      double *o = samples + 3 * ((size_t)i * 200 + q);
      o[0] = smp[q].x;
      o[1] = smp[q].y;
      o[2] = smp[q].z;
      weights[(size_t)i * 200 + q] = w[q];
    }
  }
  return S;
}

/* reps calls of one estimator per point (0: getLightFromMeshSource_Analytic,
 * 1: _AnalyticCV with cv samples, 2: getLightFromAllSources_MIS); mean and
 * the variance of the mean from `batches` batch means (the estimators seed
 * their generators from clock(), so successive calls can share streams). */
void ref_direct_stats(int est, uint32_t cv, uint32_t reps, uint32_t batches, uint32_t k,
                      const double *x, const double *n, const double *d, const uint32_t *prim,
                      double *mean, double *var) {
  if (batches == 0 || batches > reps) batches = reps;
  const uint32_t per = reps / batches;
  for (uint32_t i = 0; i < k; ++i) {
    Vec X(x[3 * i], x[3 * i + 1], x[3 * i + 2]), N(n[3 * i], n[3 * i + 1], n[3 * i + 2]),
        D(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
    Ray r(X - D, D);
    std::vector<Vec> bm(batches);
    for (uint32_t b = 0; b < batches; ++b) {
      Vec s;
      for (uint32_t q = 0; q < per; ++q) {
        BasePrimitive *bp = prim_ptr(prim[i]);
        Vec v = est == 0   ? lSource->getLightFromMeshSource_Analytic(r, N, X, bp)
                : est == 1 ? lSource->getLightFromMeshSource_AnalyticCV(r, N, X, bp, cv)
                           : lSource->getLightFromAllSources_MIS(r, N, X, bp);
        s = s + v;
      }
      bm[b] = s * (1.0 / per);
    }
    double m[3] = {0, 0, 0}, v2[3] = {0, 0, 0};
    for (uint32_t b = 0; b < batches; ++b) {
      m[0] += bm[b].x;
      m[1] += bm[b].y;
      m[2] += bm[b].z;
    }
    for (int c = 0; c < 3; ++c) m[c] /= batches;
    for (uint32_t b = 0; b < batches; ++b) {
      const double dv[3] = {bm[b].x - m[0], bm[b].y - m[1], bm[b].z - m[2]};
      for (int c = 0; c < 3; ++c) v2[c] += dv[c] * dv[c];
    }
    for (int c = 0; c < 3; ++c) {
      mean[3 * i + c] = m[c];
      var[3 * i + c] = batches > 1 ? v2[c] / (batches - 1) / batches : 0.0;
    }
  }
}

/* The reference's scene front end on one XML file, in main.cpp's order
 * (main.cpp:33-45): initSceneParser, prepareScreenCamera, prepareObjects,
 * prepareLightSources.  Outputs (capacity `cap` triangles / spheres):
 *   counts  [n_tris, n_mesh_lights, n_tri_lights, n_spheres, n_sphere_sources,
 *            n_point_sources, width, height]
 *   cam     camPos, camDir, cx, cy (12 doubles)
 *   tris    vTriangleList in order: A, B, C (9 doubles each)
 *   tmat    per triangle 15 doubles: brdf, light, specularCoef, alpha,
 *           phongExp, intIOR, extIOR, reflectance, c[3], radiance[3]
 *   mesh    triangles per mesh light, then per triangle light (1 each)
 *   sph     vSphereList: centre, radius, then the 15 material doubles
 *   ssrc    sphere sources: centre, radius, radiance[3] (7 doubles each)
 * Returns 0 on a parse error. */
static void put_mat(double *o, const MaterialType &m, const Vec &c, double refl) {
  o[0] = m.b;
  o[1] = m.l;
  o[2] = m.specularCoef;
  o[3] = m.alpha;
  o[4] = m.phongExp;
  o[5] = m.intIOR;
  o[6] = m.extIOR;
  o[7] = refl;
  o[8] = c.x;
  o[9] = c.y;
  o[10] = c.z;
  o[11] = m.radiance.x;
  o[12] = m.radiance.y;
  o[13] = m.radiance.z;
  o[14] = 0;
}

int ref_parse_scene(const char *path, uint32_t *counts, double *cam, double *tris, double *tmat,
                    uint32_t *mesh, double *sph, double *ssrc, uint32_t cap) {
  vTriangleList.clear();
  vSphereList.clear();
  if (!initSceneParser(path)) return 0;
  uint32_t w, h;
  Vec camPos, camDir, cx, cy;
  if (!prepareScreenCamera(w, h, camPos, camDir, cx, cy, false)) return 0;
  if (!prepareObjects(false)) return 0;
  if (!prepareLightSources(false)) return 0;
  const Vec cv[4] = {camPos, camDir, cx, cy};
  for (int k = 0; k < 4; ++k) {
    cam[3 * k] = cv[k].x;
    cam[3 * k + 1] = cv[k].y;
    cam[3 * k + 2] = cv[k].z;
  }
  uint32_t nm, nt, nss;
  MeshLight *ml = lSource->getMeshLightPtr(nm);
  TriLight *tl = lSource->getTriLightPtr(nt);
  SphereSource *ss = lSource->getSphereSourcePtr(nss);
  const uint32_t nT = (uint32_t)vTriangleList.size(), nS = (uint32_t)vSphereList.size();
  if (nT > cap || nS > cap || nm + nt > cap || nss > cap) return 0;
  for (uint32_t i = 0; i < nT; ++i) {
    const Triangle &t = vTriangleList[i];
    const Vec v[3] = {t.A, t.B, t.C};
    for (int k = 0; k < 3; ++k) {
      tris[9 * i + 3 * k] = v[k].x;
      tris[9 * i + 3 * k + 1] = v[k].y;
      tris[9 * i + 3 * k + 2] = v[k].z;
    }
    put_mat(tmat + 15 * (size_t)i, t.m, t.c, t.reflectance);
  }
  for (uint32_t j = 0; j < nm; ++j) mesh[j] = (uint32_t)ml[j].mesh.size();
  for (uint32_t j = 0; j < nt; ++j) mesh[nm + j] = 1;
  for (uint32_t i = 0; i < nS; ++i) {
    const Sphere &q = vSphereList[i];
    double *o = sph + 19 * (size_t)i;
    o[0] = q.p.x;
    o[1] = q.p.y;
    o[2] = q.p.z;
    o[3] = q.r;
    put_mat(o + 4, q.m, q.c, q.reflectance);
  }
  for (uint32_t i = 0; i < nss; ++i) {
    double *o = ssrc + 7 * (size_t)i;
    o[0] = ss[i].p.p.x;
    o[1] = ss[i].p.p.y;
    o[2] = ss[i].p.p.z;
    o[3] = ss[i].p.r;
    o[4] = ss[i].radiance.x;
    o[5] = ss[i].radiance.y;
    o[6] = ss[i].radiance.z;
  }
  counts[0] = nT;
  counts[1] = nm;
  counts[2] = nt;
  counts[3] = nS;
  counts[4] = nss;
  counts[5] = sceneParser->findPointSources(false);
  counts[6] = w;
  counts[7] = h;
  delete[] ml;
  delete[] tl;
  delete[] ss;
  return 1;
}

}  // extern "C"
