/*
 * ORACLE -- test infrastructure only.
 *
 * FP64 CPU restatement of the reference's per-sample radiance path
 * (Sayantan2048/ChakumRender).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, and only as the checker or the timed
 * CPU baseline; the product library never links or calls it.
 *
 * Every function cites the reference file:line it restates.  What pins each
 * function (golden vectors from the compiled reference, closed forms,
 * invariants) is listed in the header comment of oracle/cpu_ref.cpp.
 */
#ifndef CHK_ORACLE_CPU_REF_H
#define CHK_ORACLE_CPU_REF_H

#include <stdint.h>
#include "chakum_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- geometry ---------------------------------------------------------- */
/* Triangle::intersect (geometryPrimitives.cpp:29-71) with the constructor's
 * normal (C-A)x(B-A) normalised (geometryPrimitives.h:75-79); returns t or
 * 1e20 and the ray-facing unit normal N. */
double orc_tri_intersect(const double *A, const double *B, const double *C,
                         const double *o, const double *d, double *N);
/* AABBox::intersect (geometryPrimitives.cpp:107-139): t0 or 1e20. */
double orc_aabb_intersect(const double *pmin, const double *pmax,
                          const double *o, const double *d);

/* Brute-force closest hit over all triangles with the reference test and the
 * reference's `(d = t_i) && d < t` update (shader.cpp:55-70 semantics).
 * tris: n x 9 doubles (A,B,C).  rays: m x 6 (o, d).  Outputs per ray: id
 * (-1 on miss), t, N[3] (may be NULL), t2 = nearest t of any OTHER triangle
 * (1e20 if none), margin = min(beta, gamma, 1-beta-gamma) of the winning hit
 * (grazing measure), cos2 = (d.n)^2 of the winning hit. */
void orc_closest_brute(const double *tris, uint32_t n, const double *rays, uint32_t m,
                       int32_t *id, double *t, double *N, double *t2, double *margin,
                       double *cos2);

/* Sphere::intersect (geometryPrimitives.cpp:12-27): first root beyond
 * eps = 1e-4 (geometryPrimitives.h:39), else 1e20.  Brute-force closest over
 * n spheres (n x 4: center, radius) with the strict d < t of
 * BvhAccel::intersectS (bvhAccel.cpp:171): id -1 when nothing is hit. */
double orc_sphere_intersect(const double *sphere, const double *o, const double *d);
void orc_spheres_brute(const double *spheres, uint32_t n, const double *rays, uint32_t m,
                       int32_t *id, double *t);

/* Shadow query of shader.cpp:76-97 for the triangle list alone: occluded iff
 * some triangle hit has 0 < t and t + 1e-5 < dist.  rays: m x 7 (o, d, dist). */
void orc_occluded_brute(const double *tris, uint32_t n, const double *rays, uint32_t m,
                        uint8_t *occluded);

/* Reference BVH (bvhAccel.cpp:13-249): midpoint split, one primitive per
 * leaf, DFS flatten, near-child-first stack traversal.  prune=0 follows the
 * reference exactly (no t-pruning, bvhAccel.cpp:217); prune=1 skips boxes
 * whose entry t exceeds the best hit, which changes only the work. */
void *orc_bvh_build(const double *tris, uint32_t n);
void orc_bvh_free(void *bvh);
uint32_t orc_bvh_node_count(void *bvh);
/* ordered[i] = original index of reordered primitive i (bvhAccel.cpp:114-121) */
void orc_bvh_ordered(void *bvh, uint32_t *ordered);
/* Closest hit; id is the ORIGINAL triangle index (or -1), rid the reordered
 * index intersectT returns (or -1).  Counters may be NULL. */
void orc_bvh_closest(void *bvh, const double *rays, uint32_t m, int prune,
                     int32_t *id, int32_t *rid, double *t, double *N,
                     uint32_t *node_visits, uint32_t *tri_tests);

/* ---- materials --------------------------------------------------------- */
/* BasePrimitive::brdf = MaterialType::brdf * reflectance
 * (geometryPrimitives.cpp:7-9, materialTypes.cpp:9-50) and
 * MaterialType::pdfEval (materialTypes.cpp:89-116), scalar, for k
 * configurations; mat_idx selects from mats.  Vectors are k x 3. */
void orc_brdf_eval(const chk_material *mats, const uint32_t *mat_idx, uint32_t k,
                   const double *n, const double *wo_ref, const double *wo, const double *wi,
                   double *f, double *pdf);
/* One BSDF sample from three uniforms (lobe, u1, u2):
 * getBrdfDirectionSampleRandomized (materialTypes.cpp:136-153) with the
 * half-vector samplers of domainSampler.cpp:193-260, the classic-Phong lobe
 * of :264-314 and cosine sampling of :154-189, in the tangent frame of
 * orc_onb.  u: k x 3. */
void orc_brdf_sample(const chk_material *mats, const uint32_t *mat_idx, uint32_t k,
                     const double *n, const double *wo_ref, const double *wo,
                     const double *u, double *wi, double *pdf);
/* Orthonormal basis (t, b) around unit n: Duff et al., "Building an
 * Orthonormal Basis, Revisited", JCGT 6(1) 2017. */
void orc_onb(const double *n, double *t, double *b);

/* ---- LTC --------------------------------------------------------------- */
/* Installs the 4096 x 19 table (M[9], Minv[9] in ltc.inc file order,
 * amplitude) extracted from headers/ltc.inc. */
void orc_ltc_set(const double *table);
/* M_GGX (headers/ltc.h:18-61): M, Minv as row-major 3x3 (mat3::a order),
 * amplitude, and the determinant the reference carries in M.det
 * (floor-bin value, mathPrimitives.h:149-169). */
void orc_M_GGX(float theta, float alpha, double *M, double *Minv, double *amplitude, double *detM);
/* LightSource::getLightFromMeshSource_Analytic (lightSources_CV.cpp:323-333)
 * for k shading points of a GGX material.  Per point: x[3], n[3], v[3]
 * (direction TO the viewer, i.e. -r.d), alpha, ks.  Lights: n_tris
 * triangles (9 doubles each) with per-triangle radiance[3], visited in
 * array order like the reference's mesh-light loops
 * (lightSources_CV.cpp:225-266,277-318).  Outputs (k x 3, may be NULL):
 * rgb (total), ltc (getLightAnalytic, before amplitude*ks), diffuse
 * (getLightAnalyticDiffuse, before 1-ks), and amp (k, the LTC amplitude). */
void orc_ltc_shade(uint32_t k, const double *x, const double *n, const double *v,
                   const double *alpha, const double *ks,
                   const double *tris, const double *radiance, uint32_t n_tris,
                   double *rgb, double *ltc, double *diffuse, double *amp);
/* BRDF-times-cosine approximation of approxLTC_BRDF
 * (materialTypes.cpp:52-87) after getBrdfApproxParam(force=true). */
void orc_ltc_brdf_cos(uint32_t k, const double *n, const double *wo, const double *wi,
                      const double *alpha, double *val);

/* Threads of the OpenMP loops that take no thread count (n > 0). */
void orc_set_threads(int n);

/* ---- RNG ----------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void orc_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out);

/* ---- integrator ---------------------------------------------------------- */
/* FP64 path tracer restating the configs' estimator from reference parts:
 * camera rays (main.cpp:72-79, prepareScene.cpp:24-33), closest hit
 * (bvhAccel.cpp:200-249), BRDF eval/sample (materialTypes.cpp:9-153),
 * mesh-light area sampling (lightSources.cpp:809-828), balance-heuristic
 * MIS weight 1/(sum N_k p_k) (mis.cpp:209-220) with one light and one BSDF
 * sample per vertex, shadow test (shader.cpp:76-97), FP64 per-pixel
 * accumulation (main.cpp:62,83).  Randomness is Philox with the counter
 * layout of chakum_types.h.  accum: w*h*3 doubles (sums over samples,
 * added to existing contents); sumsq: optional w*h*3 doubles.
 * Only image rows row0, row0 + row_step, ... < row1 are rendered (bounded
 * CPU samples).
 * Returns the number of paths traced; *n_rays receives rays traced. */
uint64_t orc_render(const float *tris, uint32_t n_tris, const uint32_t *tri_mat,
                    const chk_material *mats, uint32_t n_mats,
                    const chk_mesh_light *lights, uint32_t n_lights,
                    const chk_camera *cam, const chk_render_params *params,
                    uint32_t row0, uint32_t row1, uint32_t row_step, int n_threads,
                    double *accum, double *sumsq, uint64_t *n_rays);

/* ---- direct lighting (shadeDI's estimators) ------------------------------ */
/* A scene for the estimators: triangles (FP32 in, FP64 math), materials,
 * mesh lights with MeshLight::initMeshLight's 300-bin inverse CDF
 * (lightSources.cpp:784-807) and the reference-semantics BVH for shadow().
 * Needs orc_ltc_set before the estimators that use the LTC fit. */
void *orc_direct_scene(const float *tris, uint32_t n_tris, const uint32_t *tri_mat,
                       const chk_material *mats, uint32_t n_mats,
                       const chk_mesh_light *lights, uint32_t n_lights);
/* The same with every kind of primitive and source of chk_scene_desc:
 * spheres (Sphere::intersect, geometryPrimitives.cpp:12-27, in shadow() and
 * the shaders' closest hit), triangle lights (TriLight), sphere sources
 * (SphereSource: cone sampling, mis.cpp:103-115, domainSampler.cpp:320-369)
 * and point sources (PointSource, lightSources.cpp:47-58).  A primitive
 * index at or above n_tris names sphere index - n_tris. */
void *orc_direct_scene_desc(const chk_scene_desc *desc);
void orc_direct_scene_free(void *scene);
/* Size of MIS::getSamples' sample set for this scene (10 per sphere source,
 * mesh light and triangle light, 60 BSDF). */
uint32_t orc_mis_count(void *scene);
/* The scene's environment light (EnvSource, lightSources.h:33-109,
 * lightSources.cpp:662-782): normalizationArea (returned; 0 without one),
 * the 2000-entry inverse CDF of the marginal and the 3000 x 2000 inverse CDFs
 * of the conditionals (either may be NULL); getRadiance of k unit
 * directions; the importance sample of k pairs (e1, e2). */
double orc_env_tables(void *scene, uint32_t *inv_marginal, uint32_t *inv_conditional);
void orc_env_radiance(void *scene, uint32_t k, const double *dirs, double *out);
void orc_env_sample(void *scene, uint32_t k, const double *u, double *out);
/* MIS weights (mis.cpp:18-95,209-220) of S given samples per point
 * (samples k x S x 3, weights k x S). */
void orc_mis_weights(void *scene, uint32_t k, const double *x, const double *n,
                     const double *wo_ref, const double *wo, const uint32_t *prim,
                     const double *samples, uint32_t S, double *weights);
/* The MIS::getSamples sample set of one point (mis.cpp:97-191) drawn with
 * the Philox counter of chakum_types.h: orc_mis_count unit directions. */
void orc_mis_samples(void *scene, const double *x, const double *n, const double *wo_ref,
                     const double *wo, uint32_t prim, uint32_t point, uint32_t estimate,
                     uint32_t seed_lo, uint32_t seed_hi, double *samples);
/* One estimate per point of getLightFromMeshSource_Analytic / _AnalyticCV /
 * getLightFromAllSources_MIS / _Analytic_MISCV (lightSources_CV.cpp:323-450,
 * lightSources.cpp:32-109) with shadow() of shader.cpp:76-97 over the
 * reference BVH.  x, n, d (k x 3), prim (k), rgb (k x 3). */
void orc_direct_light(void *scene, const chk_direct_params *p, uint32_t k, const double *x,
                      const double *n, const double *d, const uint32_t *prim, int n_threads,
                      double *rgb);
/* The pixel loop of main.cpp:62-94 around one of the reference's shaders
 * (chk_shader): shadeDirectOnly (shader.cpp:339-363), shadeExplicit
 * (:111-207, with dp's estimator in shadeDI) or shadeImplicit (:276-337).
 * Rows row0, row0 + row_step, ... < row1; accum/sumsq as orc_render. */
void orc_render_shade(void *scene, const chk_camera *cam, const chk_render_params *P,
                      uint32_t shader, const chk_direct_params *dp, uint32_t row0, uint32_t row1,
                      uint32_t row_step, int n_threads, double *accum, double *sumsq);

#ifdef __cplusplus
}
#endif
#endif
