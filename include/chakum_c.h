/*
 * libchakum -- C ABI of the MI355X per-sample radiance core.
 *
 * Plain pointers and sizes only.  "dev" pointers are device (HBM) pointers
 * the caller allocated (hipMalloc, or any allocator such as torch's);
 * "host" pointers are ordinary host memory.  `stream` is a hipStream_t passed
 * as void* (NULL = the default stream).  No call throws, exits or prints:
 * every entry point returns a chk_status and leaves a message for
 * chk_last_error() on failure.
 *
 * Each entry point names the reference interface it replaces
 * (Sayantan2048/ChakumRender, paths relative to the reference root).
 */
#ifndef CHAKUM_C_H
#define CHAKUM_C_H

#include <stddef.h>
#include <stdint.h>

#include "chakum_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum chk_status {
  CHK_OK = 0,
  CHK_EINVAL = 1,     /* bad argument (null pointer, size 0 where not allowed, bad enum) */
  CHK_ENOMEM = 2,     /* device or host allocation failed */
  CHK_EDEVICE = 3,    /* HIP runtime error */
  CHK_ECAPACITY = 4,  /* a fixed capacity would be exceeded (the reference overflows silently,
                         e.g. bvhAccel.cpp:208 todo[64], lightSources.cpp:38 samples[200]) */
  CHK_EINTERNAL = 5
} chk_status;

typedef struct chk_ctx_s *chk_ctx_t;
typedef struct chk_scene_s *chk_scene_t;

/* Per-ray record for trace calls: 32 bytes, two float4.
 *   org_tmax = (o.x, o.y, o.z, tmax)   hits count only for 0 < t < tmax
 *   dir_skip = (d.x, d.y, d.z, skip)   skip: ORIGINAL triangle index never
 *                                      reported (raw int32 bits, -1 = none)
 * The reference's closest-hit query (BvhAccel::intersectT) is tmax = 1e30,
 * skip = -1; its shadow query shadow(ray, dist) (shader.cpp:76-97) is
 * tmax = dist - 1e-5. */
typedef struct chk_ray {
  float org_tmax[4];
  float dir_skip[4];
} chk_ray;

/* Closest-hit record: 16 bytes.  tri = ORIGINAL triangle index (the
 * reference returns an index into its reordered array, bvhAccel.cpp:223),
 * -1 on a miss (t = 1e30).  u, v = barycentric weights of vertices B and C
 * (beta, gamma of geometryPrimitives.cpp:63-64). */
typedef struct chk_hit {
  float t;
  int32_t tri;
  float u;
  float v;
} chk_hit;

typedef struct chk_scene_info {
  uint32_t n_tris;          /* triangles passed in */
  uint32_t n_tris_bvh;      /* triangles in the BVH (zero-area ones are dropped) */
  uint32_t n_nodes;         /* 128-byte 4-wide nodes */
  uint32_t n_leaves;
  uint32_t max_depth;
  uint32_t stack_bound;     /* worst-case traversal stack entries */
  uint32_t stack_size;      /* LDS stack entries per lane the kernels use */
  uint32_t n_lights;
  uint64_t device_bytes;    /* scene bytes resident in HBM */
  double sah_cost;
  double build_ms;
} chk_scene_info;

typedef struct chk_render_stats {
  uint64_t paths;           /* camera paths traced */
  uint64_t rays_extend;     /* closest-hit rays */
  uint64_t rays_shadow;     /* any-hit rays */
  uint64_t node_visits;     /* filled by chk_render_count only */
  uint64_t tri_tests;       /* filled by chk_render_count only */
  uint64_t shadow_node_visits;
  uint64_t shadow_tri_tests;
  uint32_t chunks;
  uint32_t extend_launches;
  double extend_ms;         /* sum of closest-hit kernel times (CHK_RENDER_PROFILE) */
  double shadow_ms;
  double shade_ms;
  double other_ms;
} chk_render_stats;

/* Time every kernel launch with HIP events into chk_render_stats. */
#define CHK_RENDER_PROFILE 0x100u

const char *chk_status_string(chk_status s);
/* Detail message of the last failing call on this thread. */
const char *chk_last_error(void);
/* Library version string. */
const char *chk_version(void);

/* Device memory and streams through the library's own HIP runtime, so a
 * caller needs no second runtime in its process (no reference counterpart:
 * the reference is host-only).  kind: 0 host->device, 1 device->host,
 * 2 device->device.  chk_memzero clears with a fill kernel (safe to record
 * in a HIP graph).  A NULL stream is the default stream. */
chk_status chk_device_alloc(int device, size_t bytes, void **dev);
chk_status chk_device_free(void *dev);
chk_status chk_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream);
chk_status chk_memzero(void *dev, size_t bytes, void *stream);
chk_status chk_stream_create(int device, void **stream);
chk_status chk_stream_destroy(void *stream);
chk_status chk_stream_sync(void *stream);

/* Context: one per device.  Replaces the reference's process globals
 * (objects.h:7-17, bvhAccel.h:78, lightSources.h:251, mis.h:39). */
chk_status chk_create(int device, chk_ctx_t *out);
chk_status chk_destroy(chk_ctx_t ctx);
/* Installs the 64 x 64 GGX LTC fit (4096 records x 19 doubles: M[9],
 * Minv[9] in the order of headers/ltc.inc, amplitude), the data M_GGX reads
 * (headers/ltc.h:18-61, headers/ltc.inc). host pointer. */
chk_status chk_set_ltc_table(chk_ctx_t ctx, const double *table_host, size_t n_doubles);

/* Scene + accelerator.  Replaces loadAccels() (objects.cpp:38-56) and
 * BvhAccel(list, n, size) + initAccel() (bvhAccel.h:71-72,
 * bvhAccel.cpp:13-22): copies the caller's arrays (never reorders them),
 * builds the BVH on the host and uploads everything.  All pointers host.
 *   tris     n_tris x 9 floats (A, B, C), the Triangle vertices
 *   tri_mat  n_tris material indices
 *   mats     n_mats materials (MaterialType + BasePrimitive colour,
 *            reflectance)
 *   lights   n_lights mesh lights, each a run of emitter triangles
 *            (MeshLight, lightSources.h:141-153) */
chk_status chk_scene_create(chk_ctx_t ctx, const float *tris, uint32_t n_tris,
                            const uint32_t *tri_mat, const chk_material *mats, uint32_t n_mats,
                            const chk_mesh_light *lights, uint32_t n_lights, chk_scene_t *out);
chk_status chk_scene_destroy(chk_scene_t scene);
chk_status chk_scene_get_info(chk_scene_t scene, chk_scene_info *info);

/* Closest hit for n rays.  Replaces BvhAccel::intersectT
 * (bvhAccel.cpp:200-249) batched.  rays, hits: dev. */
chk_status chk_trace_closest(chk_scene_t scene, const chk_ray *rays, uint32_t n, chk_hit *hits,
                             void *stream);
/* Occlusion for n rays: occluded[i] = 1 iff some triangle is hit with
 * 0 < t < tmax.  Replaces shadow() (shader.cpp:76-97) batched.  dev. */
chk_status chk_trace_any(chk_scene_t scene, const chk_ray *rays, uint32_t n, uint8_t *occluded,
                         void *stream);

/* BRDF value and sampling pdf for k configurations.  Replaces
 * BasePrimitive::brdf (geometryPrimitives.cpp:7-9 -> MaterialType::brdf,
 * materialTypes.cpp:9-50; headers/materialTypes.h:100) and
 * MaterialType::pdfEval (materialTypes.cpp:89-116; materialTypes.h:103).
 * Like the reference's Vec arguments, the directions are FP64 and the
 * evaluation is FP64.  mats (n_mats), mat_idx (k), n, wo_ref, wo, wi
 * (k x 3), f, pdf (k): all dev.  flags: 0, or CHK_BRDF_FP32 to evaluate
 * with the FP32 arithmetic of the fused render kernels on the FP32
 * rounding of the inputs. */
#define CHK_BRDF_FP32 1u
chk_status chk_brdf_eval(const chk_material *mats, uint32_t n_mats, const uint32_t *mat_idx,
                         const double *n, const double *wo_ref, const double *wo,
                         const double *wi, uint32_t k, uint32_t flags, double *f, double *pdf,
                         void *stream);
/* One BSDF sample per configuration from three uniforms u (k x 3: lobe,
 * e1, e2).  Replaces MaterialType::getBrdfDirectionSampleRandomized
 * (materialTypes.cpp:136-153; materialTypes.h:102) with explicit random
 * numbers instead of a clock()-seeded generator.  wi (k x 3), pdf (k): dev.
 * FP64 unless flags has CHK_BRDF_FP32. */
chk_status chk_brdf_sample(const chk_material *mats, uint32_t n_mats, const uint32_t *mat_idx,
                           const double *n, const double *wo_ref, const double *wo,
                           const double *u, uint32_t k, uint32_t flags, double *wi, double *pdf,
                           void *stream);

/* Analytic polygon-light shading of k points of a GGX material (alpha[i],
 * ks[i]) by n_light_tris emitter triangles with per-triangle radiance.
 * Replaces LightSource::getLightFromMeshSource_Analytic
 * (lightSources_CV.cpp:323-333: LTC integral getLightAnalytic :271-321 with
 * M_GGX ltc.h:18-61, plus the diffuse integral :221-269).  x, n, v (k x 3;
 * v = direction to the viewer), alpha, ks (k), light_tris (n x 9),
 * radiance (n x 3), rgb (k x 3): dev.  Needs chk_set_ltc_table. */
chk_status chk_ltc_eval(chk_ctx_t ctx, const float *x, const float *n, const float *v,
                        const float *alpha, const float *ks, uint32_t k,
                        const float *light_tris, const float *radiance, uint32_t n_light_tris,
                        float *rgb, void *stream);
/* Render: the pixel loop of main.cpp:62-94 around shade() (shader.cpp:362),
 * with the configs' estimator: path tracing to params->max_depth vertices,
 * next-event estimation from the mesh lights and BSDF sampling combined by
 * the balance heuristic of mis.cpp:209-220.  accum (W*H*3 doubles, dev)
 * receives per-pixel sums over the params->spp samples (added to its
 * contents, FP64, row 0 = top of the image like main.cpp:72); sumsq (dev,
 * may be NULL) the sums of squares when params->flags has CHK_RENDER_SUMSQ.
 * stats (host) may be NULL. */
chk_status chk_render(chk_scene_t scene, const chk_camera *camera, const chk_render_params *params,
                      double *accum, double *sumsq, chk_render_stats *stats, void *stream);
/* chk_render with counting traversal kernels: fills node_visits /
 * tri_tests (slower; for the work model, not for timing). */
chk_status chk_render_count(chk_scene_t scene, const chk_camera *camera,
                            const chk_render_params *params, double *accum,
                            chk_render_stats *stats, void *stream);

/* ---- learned light-selection proposal -----------------------------------
 * No reference counterpart (the reference gives every mesh light the same
 * sample count, mis.h:9, and the balance heuristic, mis.cpp:209-220).  A
 * 64-64-64-64-33 ReLU MLP (BF16 weights, FP32 accumulation on MFMA) maps 7
 * inputs in [0, 1] -- position in the scene box, normal and view direction as
 * (theta/pi, phi/2pi) -- through a 64-wide frequency encoding to 32
 * light-selection logits and one mixing logit.  The pmf is the defensive
 * mixture q' = 0.9 softmax + 0.1 uniform and the mixing weight
 * c = 0.1 + 0.8 sigmoid, so the estimator stays unbiased for any weights.
 * Parameters: 14625 floats, W1[64][64] b1[64] W2 b2 W3 b3 W4[64][33] b4[33]
 * (W[in][out]).  A new proposal has zero output weights: uniform pmf and
 * c = 1/2, i.e. plain balance-heuristic MIS. */
typedef struct chk_proposal_s *chk_proposal_t;
chk_status chk_proposal_create(chk_ctx_t ctx, uint32_t n_lights, uint64_t seed,
                               chk_proposal_t *out);
chk_status chk_proposal_destroy(chk_proposal_t p);
/* host float[14625]; set resets the Adam state. */
chk_status chk_proposal_set_params(chk_proposal_t p, const float *host, size_t n);
chk_status chk_proposal_get_params(chk_proposal_t p, float *host, size_t n);
/* inputs k x 7, probs k x 32 (q'), mix k (c), logits k x 33 (may be NULL): dev. */
chk_status chk_proposal_eval(chk_proposal_t p, const float *inputs, uint32_t k, float *probs,
                             float *mix, float *logits, void *stream);
/* Parameter gradient (dev float[14625], overwritten) of a loss whose
 * gradient with respect to the 33 logits is dlogits (k x 33, dev). */
chk_status chk_proposal_grad(chk_proposal_t p, const float *inputs, const float *dlogits,
                             uint32_t k, float *grad, void *stream);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) with learning rate lr. */
chk_status chk_proposal_adam(chk_proposal_t p, const float *grad, float lr, void *stream);

#ifdef __cplusplus
}
#endif
#endif
