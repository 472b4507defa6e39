/*
 * Plain-C scene records shared by the HIP library (libchakum.so) and its
 * callers.  Every record is a fixed-layout POD: no torch types, no C++.
 *
 * Enumerations keep the reference's numeric values so scene data written
 * for the reference maps one to one:
 *   BRDFType  headers/materialTypes.h:7   {BRDF_NONE, CLASSIC_PHONG, GGX, BECKMANN, PHONG, GGXAPPROX}
 *   LightType headers/materialTypes.h:6   {NONE, POINT, VOLUME, PLANAR, MESH}
 */
#ifndef CHAKUM_TYPES_H
#define CHAKUM_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum chk_brdf_type {
  CHK_BRDF_NONE = 0,
  CHK_BRDF_CLASSIC_PHONG = 1,
  CHK_BRDF_GGX = 2,
  CHK_BRDF_BECKMANN = 3,
  CHK_BRDF_PHONG = 4,
  CHK_BRDF_GGXAPPROX = 5
};

enum chk_light_type {
  CHK_LIGHT_NONE = 0,
  CHK_LIGHT_POINT = 1,
  CHK_LIGHT_VOLUME = 2,
  CHK_LIGHT_PLANAR = 3,
  CHK_LIGHT_MESH = 4
};

/* One material.  Mirrors MaterialType (headers/materialTypes.h:20-120) plus
 * the per-primitive colour `c` and `reflectance` of BasePrimitive
 * (headers/geometryPrimitives.h:24-36), which the reference stores beside
 * the material in every triangle.  56 bytes. */
typedef struct chk_material {
  uint32_t brdf;        /* chk_brdf_type */
  uint32_t light;       /* chk_light_type; != NONE marks an emitter */
  float ks;             /* specularCoef in [0,1] (materialTypes.h:88-89 clamps it) */
  float alpha;          /* GGX/Beckmann roughness, or PHONG exponent */
  float phong_exp;      /* CLASSIC_PHONG exponent */
  float int_ior;        /* interior IOR */
  float ext_ior;        /* exterior IOR */
  float reflectance;    /* BasePrimitive::reflectance */
  float color[3];       /* BasePrimitive::c */
  float radiance[3];    /* emitter radiance = col.maxnorm()*intensity (lightSources.h:12-16) */
} chk_material;

/* A mesh light: a contiguous run of emitter triangles in the scene's
 * triangle array (MeshLight, headers/lightSources.h:141-153; the reference
 * appends light triangles to the scene list in LightSource::addMSource,
 * lightSources.h:197-203). */
typedef struct chk_mesh_light {
  uint32_t first_tri;
  uint32_t n_tris;
} chk_mesh_light;

/* A sphere primitive: Sphere (headers/geometryPrimitives.h:38-52) with its
 * material (colour, reflectance and MaterialType in a chk_material).  FP64
 * like the reference's Sphere.  40 bytes. */
typedef struct chk_sphere {
  double center[3];
  double radius;
  uint32_t mat;
  uint32_t pad;
} chk_sphere;

/* A point light: PointSource (headers/lightSources.h:22-31).  power is its
 * BaseSource radiance field, col.maxnorm() * power (lightSources.h:12-16);
 * the radiance at x is power / (4 pi |p - x|^2).  48 bytes. */
typedef struct chk_point_source {
  double pos[3];
  double power[3];
} chk_point_source;

/* An environment light: EnvSource (headers/lightSources.h:33-109).
 *   image     height x width x 3 doubles; row 0 is theta = 0 about the map's
 *             polar axis (the first row readImage loads, ppm.cpp:36-115)
 *   axis      the world axis that is the map's polar axis: 0 x, 1 y, 2 z
 *             (EnvSource::getRadiance, lightSources.h:80-94)
 *   radiance  the BaseSource radiance getRadiance multiplies each pixel by,
 *             col.maxnorm() * intensity
 * The library copies the image and builds initSamples' tables
 * (lightSources.cpp:662-782) on the GPU.  40 bytes. */
typedef struct chk_env_source {
  const double *image;
  uint32_t width, height, axis, pad;
  double radiance[3];
} chk_env_source;

/* Everything a scene holds, in the order LightSource and the object lists
 * keep it (lightSources.h:158-203, objects.h:7-17).  Light geometry is part
 * of the primitive lists, as LightSource::addMSource / addTSource / addSSource
 * append it there:
 *   mesh_lights     MeshLight: runs of emitter triangles
 *   tri_lights      TriLight: one emitter triangle each (triangle indices)
 *   sphere_sources  SphereSource: emitter spheres (sphere indices); the
 *                   sphere is both the visible sphere and the source
 *   point_sources   PointSource: not geometry
 *   env             EnvSource or NULL (LightSource::addESource keeps one) */
typedef struct chk_scene_desc {
  const float *tris;            /* n_tris x 9 (A, B, C) */
  const uint32_t *tri_mat;      /* n_tris */
  const chk_material *mats;     /* n_mats */
  const chk_mesh_light *mesh_lights;
  const uint32_t *tri_lights;
  const chk_sphere *spheres;
  const uint32_t *sphere_sources;
  const chk_point_source *point_sources;
  const chk_env_source *env;
  uint32_t n_tris, n_mats, n_mesh_lights, n_tri_lights;
  uint32_t n_spheres, n_sphere_sources, n_point_sources, reserved;
} chk_scene_desc;

/* Pinhole camera with the reference's conventions (prepareScene.cpp:12-36,
 * main.cpp:72-79): cx = |dir x up| * (w/h) * tan(fovx/2),
 * cy = |cx x dir| * tan(fovy/2); fov in degrees. */
typedef struct chk_camera {
  float pos[3];
  float dir[3];
  float up[3];
  float fovx_deg;
  float fovy_deg;
  uint32_t width;
  uint32_t height;
} chk_camera;

enum chk_mis_mode {
  CHK_MIS_BALANCE = 0,      /* NEE + BSDF sampling, balance heuristic (mis.cpp:209-220 with N=1 each) */
  CHK_MIS_LEARNED = 1       /* learned light selection + mixing weight, defensive mixture */
};

enum chk_jitter_mode {
  CHK_JITTER_BOX = 0,       /* uniform sub-pixel jitter per sample */
  CHK_JITTER_REFERENCE = 1  /* fixed (0.34, 0.20) offset of main.cpp:23-24,77 */
};

typedef struct chk_render_params {
  uint32_t spp;             /* samples per pixel */
  uint32_t sample_offset;   /* first sample index (progressive / resumable renders) */
  uint32_t max_depth;       /* path vertices with next-event estimation; the BSDF ray
                               leaving the last one is traced for emission only */
  uint32_t mis_mode;        /* chk_mis_mode */
  uint32_t jitter_mode;     /* chk_jitter_mode */
  uint32_t seed_lo;         /* Philox key */
  uint32_t seed_hi;
  uint32_t flags;           /* CHK_RENDER_* bits */
  uint32_t window[4];       /* x0, y0, w, h: render only this rectangle of the image (row 0
                               at the top; each pixel exactly as in a whole-image render);
                               all zero = the whole image.  chk_render and its count, learned,
                               no-allocation and batched forms */
} chk_render_params;

/* A directly seen emitter returns its radiance (physical) instead of
 * radiance.maxnorm(), the reference's camera-ray value (materialTypes.h:34-38,
 * shader.cpp:354-356). */
#define CHK_RENDER_PHYSICAL_EMITTER 1u
/* Also accumulate the per-pixel sum of squared per-sample values
 * (the sumsq buffer receives W*H*3 doubles). */
#define CHK_RENDER_SUMSQ            2u

/* Direct-lighting estimators of the reference's shadeDI (shader.cpp:100-109),
 * one per line of its selection list. */
enum chk_direct_estimator {
  CHK_DIRECT_ANALYTIC = 0,     /* getLightFromMeshSource_Analytic (lightSources_CV.cpp:323-333) */
  CHK_DIRECT_ANALYTIC_CV = 1,  /* getLightFromMeshSource_AnalyticCV (lightSources_CV.cpp:335-397) */
  CHK_DIRECT_MIS = 2,          /* getLightFromAllSources_MIS (lightSources.cpp:32-109) */
  CHK_DIRECT_MISCV = 3         /* getLightFromMeshSource_Analytic_MISCV (lightSources_CV.cpp:399-450),
                                  the estimator shadeDI ships with */
};

/* MISCV: apply the shadow test that the reference computes and then
 * ignores (its visibility is 1 either way, lightSources_CV.cpp:433-438). */
#define CHK_DIRECT_VISIBILITY 1u

typedef struct chk_direct_params {
  uint32_t estimator;    /* chk_direct_estimator */
  uint32_t cv_samples;   /* ANALYTIC_CV: samples per mesh light (its nSamples; capped at
                            1000 like MAX_MESH_SAMPLES, lightSources.cpp:24) */
  uint32_t flags;        /* CHK_DIRECT_* bits */
  uint32_t estimate;     /* estimate index (second Philox counter word; point entry only) */
  uint32_t seed_lo;      /* Philox key */
  uint32_t seed_hi;
} chk_direct_params;

/* Philox counter of the direct-lighting estimators:
 *   ctr = { point index (or pixel), estimate index (or sample), slot,
 *           CHK_STREAM_DIRECT + path vertex (0 except in shadeExplicit) }
 * with one call per sample slot; a 32-bit output x maps to (x + 0.5) * 2^-32.
 * The slots follow MIS::getSamples (mis.cpp:97-191): 10 per sphere source
 * (its cone), then 10 per mesh light and 10 per triangle light (by area), then
 * 60 for the BSDF (the first floor(60 ks) from the specular lobe); the MIS
 * estimator adds one slot per point source after them (no draw).  With mesh
 * lights only, slots 10 j .. 10 j + 9 are mesh light j.  ANALYTIC_CV uses
 * slots j n .. j n + n - 1 for mesh light j. */
#define CHK_STREAM_DIRECT 0x44490000u

/* The reference's shaders (shader.cpp:362-378 selects one at compile time). */
enum chk_shader {
  CHK_SHADE_DIRECT_ONLY = 0,  /* shadeDirectOnly (shader.cpp:339-363), the shipped one */
  CHK_SHADE_EXPLICIT = 1,     /* shadeExplicit (shader.cpp:111-207) */
  CHK_SHADE_IMPLICIT = 2      /* shadeImplicit (shader.cpp:276-337) */
};
/* Philox counter of the shaders' own draws:
 *   ctr = { pixel, sample, (vertex << 8) | k, CHK_STREAM_SHADE }
 * k = the try index of a continuation direction, 0xFF for Russian roulette. */
#define CHK_STREAM_SHADE 0x53480000u
/* shadeExplicit draws a new direction while the last one hit an emitter
 * (shader.cpp:159-183, unbounded there); here at most this many draws, after
 * which the path ends as if the ray had escaped. */
#define CHK_SHADE_MAX_TRIES 64u
#define CHK_MIS_LIGHT_SAMPLES 10u  /* nSuArIS, mis.h:9 */
#define CHK_MIS_BSDF_SAMPLES 60u   /* nBrIS, mis.h:11 */

/* Philox counter layout shared by every implementation of the path:
 *   ctr = { pixel index, sample index, (depth << 2) | stream, request id }
 *   key = { seed_lo, seed_hi }
 * stream 0: camera jitter (depth 0 only) / next-event estimation,
 * stream 1: BSDF sampling.  A 32-bit output x maps to (x >> 8) * 2^-24. */
#define CHK_STREAM_NEE  0u
#define CHK_STREAM_BSDF 1u
#define CHK_STREAM_CAM  2u

#ifdef __cplusplus
}
#endif
#endif
