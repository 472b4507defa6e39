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
  uint32_t n_nodes;         /* 4-wide nodes (64 bytes each on the device) */
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
  double proposal_ms;       /* proposal inputs, inference and training */
  uint32_t train_steps;     /* Adam steps taken by chk_render_learned */
  uint32_t stack_retries;   /* rays traced again with the tree's full traversal stack after
                               overflowing the shorter one (chk_scene_set_trace_stack) */
} chk_render_stats;

/* Time every kernel launch with HIP events into chk_render_stats. */
#define CHK_RENDER_PROFILE 0x100u
/* Run a render's chunks one after another on the caller's stream.  By default
 * a plain render (or a learned one that does not train) of several chunks
 * alternates two queue sets and streams, so that consecutive chunks overlap;
 * this flag gives each kernel launch the GPU to itself, for timing. */
#define CHK_RENDER_ONE_STREAM 0x200u

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
/* The whole scene of the reference's prepareScene (prepareScene.cpp:54-240:
 * prepareLights + prepareObjects): triangles and spheres with their
 * materials, mesh, triangle, sphere and point light sources
 * (chk_scene_desc).  chk_scene_create is this call with triangles and mesh
 * lights only.  Spheres and the triangle, sphere and point sources are seen by
 * the direct-lighting entry points and chk_render_shade; chk_render,
 * chk_render_count, chk_render_batch and chk_render_learned (the mesh-light
 * path tracer) reject a scene that has them.  At most 4096 spheres and 1024
 * sources of each of those kinds (CHK_ECAPACITY). */
chk_status chk_scene_create_desc(chk_ctx_t ctx, const chk_scene_desc *desc, chk_scene_t *out);
chk_status chk_scene_destroy(chk_scene_t scene);
/* Traversal stack per lane (no reference counterpart: the reference's
 * traversal keeps a 64-entry stack and never checks it, bvhAccel.cpp:208).
 * A tree whose worst case needs more than `entries` (8, 16 or 32, default
 * 16) is traced by the renders with `entries` LDS entries per lane, at
 * higher occupancy, and a ray that overflows them is traced again with the
 * full stack; results are the same either way.
 * chk_render_stats.stack_retries counts those rays.  8 is for testing the
 * second trace: rays of ordinary scenes overflow it. */
chk_status chk_scene_set_trace_stack(chk_scene_t scene, uint32_t entries);
chk_status chk_scene_get_info(chk_scene_t scene, chk_scene_info *info);
/* Accelerator of chk_trace_closest and chk_trace_any: CHK_ACCEL_BVH (the
 * default; BvhAccel, bvhAccel.h:71-74) or CHK_ACCEL_DUMMY, the reference's
 * DummyAccel (dummyAccel.h:11-20, dummyAccel.cpp:10-59): the triangles in
 * their original order in floor(sqrt(n)) runs with one box each, every
 * triangle of every box the ray crosses tested in order, the first of
 * equally near ones kept.  Hits and ids are those of the BVH (the reference's
 * DummyAccel reports the run index as the id, dummyAccel.cpp:49; here it is
 * the triangle's, and the run is id / floor(sqrt(n)), at most the last run).
 * The renders always traverse the BVH. */
#define CHK_ACCEL_BVH 0u
#define CHK_ACCEL_DUMMY 1u
chk_status chk_scene_set_accel(chk_scene_t scene, uint32_t kind);

/* Closest hit for n rays.  Replaces BvhAccel::intersectT
 * (bvhAccel.cpp:200-249) batched.  rays, hits: dev. */
chk_status chk_trace_closest(chk_scene_t scene, const chk_ray *rays, uint32_t n, chk_hit *hits,
                             void *stream);
/* Occlusion for n rays: occluded[i] = 1 iff some triangle is hit with
 * 0 < t < tmax.  Replaces shadow() (shader.cpp:76-97) batched.  dev. */
chk_status chk_trace_any(chk_scene_t scene, const chk_ray *rays, uint32_t n, uint8_t *occluded,
                         void *stream);

/* Closest sphere for n rays: Sphere::intersect (geometryPrimitives.cpp:12-27:
 * nearest root beyond eps = 1e-4) in FP64 over the n_spheres spheres
 * (n_spheres x 4 floats: centre, radius), keeping a hit only when it is
 * strictly nearer than the ray's tmax and every earlier hit.  Replaces
 * BvhAccel::intersectS (bvhAccel.cpp:151-198; pass tmax = INFINITY for its
 * semantics).  hits[i].tri = original sphere index or 0xFFFFFFFF; u = v = 0.
 * spheres, rays, hits: dev. */
chk_status chk_trace_spheres(const float *spheres, uint32_t n_spheres, const chk_ray *rays,
                             uint32_t n, chk_hit *hits, void *stream);

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

/* chk_render for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference, which renders on the host.  Same
 * arguments, results and errors as chk_render -- every image equals
 * chk_render's bit for bit -- but the path queues, counters and overflow
 * list live in a fixed array of CHK_NOALLOC_SCRATCH_BYTES of device memory
 * that the library holds on each device it runs on until the process ends,
 * so the call makes no hipMalloc/hipFree.  Paths are rendered in chunks of
 * at most 356,352 (the most that array holds), so spp above that is
 * CHK_ECAPACITY.  Not available with CHK_RENDER_PROFILE (CHK_EINVAL).  All
 * calls on one device share its array: calls that overlap on one device
 * (other streams, other threads) race and must be serialised by the
 * caller.  Each call clears the counters it uses; nothing else carries
 * over between calls. */
#define CHK_NOALLOC_SCRATCH_BYTES (64u << 20)
chk_status chk_render_noalloc(chk_scene_t scene, const chk_camera *camera,
                              const chk_render_params *params, double *accum, double *sumsq,
                              chk_render_stats *stats, void *stream);

/* ---- direct lighting: the reference's shadeDI estimators ----------------
 * Shading points are the reference's (r, n, x, primitive) arguments:
 *   x     hit point (k x 3 doubles)
 *   n     unit normal facing the viewer, as intersectT returns it (k x 3)
 *   d     unit direction of the ray that hit x, r.d (k x 3)
 *   prim  the primitive x lies on, whose material is the BasePrimitive's:
 *         the original triangle index, or n_tris + sphere index (k)
 * All FP64 in and out, like the reference's Vec arguments.  The light
 * sources are the scene's (chk_scene_desc): the mesh-light estimators see the
 * mesh lights, getLightFromAllSources_MIS every kind.  Materials must be GGX,
 * Beckmann, Phong or classic Phong (the reference has no sampler for the
 * others). */

/* MIS weights of given samples.  Replaces the weight half of
 * MIS::getSamples (mis.cpp:18-95,209-220): weight = 1 / (sum over the
 * sampling techniques of N_k p_k(sample) + 1e-20) with N = 10 per sphere
 * source (uniform cone, 1 / cap area inside it), per mesh light and per
 * triangle light (area sampling, solid-angle pdf d^2 / (A |cos|) from
 * Ray(x, sample)), and N = 60 for the BSDF (MaterialType::pdfEval).  wo_ref and wo as the
 * reference passes them (lightSources.cpp:44-45,60).  samples: k x S x 3
 * unit directions (point-major); weights: k x S.  All dev. */
chk_status chk_mis_weights(chk_scene_t scene, const double *x, const double *n,
                           const double *wo_ref, const double *wo, const uint32_t *prim,
                           uint32_t k, const double *samples, uint32_t S, double *weights,
                           void *stream);
/* One estimate of direct lighting per shading point, the value the
 * reference's estimator returns (before shadeDI multiplies the colour):
 * chk_direct_params selects getLightFromMeshSource_Analytic, ..._AnalyticCV,
 * getLightFromAllSources_MIS or ..._Analytic_MISCV.  Samples come from
 * Philox (counter in chakum_types.h) in the reference's allocation and
 * distributions; shadow rays are traced on the scene.  ANALYTIC, ANALYTIC_CV
 * and MISCV need chk_set_ltc_table.  x, n, d, prim, rgb (k x 3): dev. */
chk_status chk_direct_light(chk_scene_t scene, const double *x, const double *n,
                            const double *d, const uint32_t *prim, uint32_t k,
                            const chk_direct_params *params, double *rgb, void *stream);
/* Render with the reference's shipped shade() = shadeDirectOnly
 * (shader.cpp:339-363): camera ray, closest hit, an emitter's
 * radiance.maxnorm() (or its radiance with CHK_RENDER_PHYSICAL_EMITTER),
 * else colour times the selected direct-lighting estimator.  params: spp,
 * sample_offset, jitter_mode, seed and flags as chk_render (max_depth and
 * mis_mode unused); direct->estimate unused (the sample index takes its
 * place).  accum, sumsq: as chk_render. */
chk_status chk_render_direct(chk_scene_t scene, const chk_camera *camera,
                             const chk_render_params *params, const chk_direct_params *direct,
                             double *accum, double *sumsq, chk_render_stats *stats,
                             void *stream);
/* Render with one of the reference's shaders (chk_shader), the compile-time
 * choice of shade() (shader.cpp:362-378):
 *   CHK_SHADE_DIRECT_ONLY  as chk_render_direct
 *   CHK_SHADE_EXPLICIT     shadeExplicit (shader.cpp:111-207): colour times
 *                          the selected direct-lighting estimator at every
 *                          vertex, Russian roulette with probability 0.3,
 *                          uniform-hemisphere continuation that draws again
 *                          while it hits an emitter (CHK_SHADE_MAX_TRIES)
 *   CHK_SHADE_IMPLICIT     shadeImplicit (shader.cpp:276-337): up to 4
 *                          cosine-sampled vertices, an emitter's
 *                          radiance.maxnorm() ends the path (direct unused)
 * params and accumulators as chk_render_direct. */
chk_status chk_render_shade(chk_scene_t scene, const chk_camera *camera,
                            const chk_render_params *params, uint32_t shader,
                            const chk_direct_params *direct, double *accum, double *sumsq,
                            chk_render_stats *stats, void *stream);

/* ---- batched requests ----------------------------------------------------
 * No single reference counterpart: the reference renders one image per
 * process (main.cpp:62-94).  Many render requests -- each with its own
 * scene, camera, parameters (CHK_MIS_BALANCE) and accumulators -- run in one
 * wavefront: requests are ordered by cost (W * H * spp, shortest first);
 * those below a quarter of tuning->chunk_paths paths are packed into shared
 * chunks that mix requests and scenes, larger ones get chunks of their own
 * (single-scene kernels).  Chunks are launched in that order, alternating two
 * queue sets and streams, so consecutive chunks overlap.  Each request's
 * image equals its chk_render image exactly (same keys, same per-path
 * arithmetic, same per-pixel summation order). */
typedef struct chk_request {
  chk_scene_t scene;
  chk_camera camera;
  chk_render_params params;
  double *accum;            /* W*H*3, dev: += per-pixel sums */
  double *sumsq;            /* dev or NULL; used with CHK_RENDER_SUMSQ */
} chk_request;

typedef struct chk_batch_tuning {
  uint32_t chunk_paths;     /* paths in flight per chunk, 2^16..2^26 (0: 2^24) */
  uint32_t trace_block;     /* traversal threads per block: 64, 128, 256 (0: 128) */
  uint32_t shade_block;     /* shading threads per block: 128, 256, 512 (0: 256) */
  uint32_t use_graph;       /* capture the whole batch (both streams) into one
                               hipGraph: a call is one hipGraphLaunch; the
                               executable graph is reused by the next identical
                               call */
} chk_batch_tuning;

typedef struct chk_batch_stats {
  uint64_t paths;
  uint64_t rays_extend;
  uint64_t rays_shadow;
  uint32_t chunks;
  uint32_t segments;        /* (request, pixel range) pieces */
  double batch_ms;          /* device time from the first launch to the last */
  double capture_ms;        /* host time spent capturing graphs (0 if reused) */
  uint32_t graph_cached;    /* 1: the call replayed graphs of the previous call */
  uint32_t reserved;
} chk_batch_stats;

/* Render n requests.  tuning NULL: defaults.  latency_ms (host, n, may be
 * NULL): per request, device time from the batch start to the end of the
 * chunk that finished it (device clock stamps written after each chunk, in
 * both modes).  Synchronises the stream. */
chk_status chk_render_batch(chk_ctx_t ctx, const chk_request *reqs, uint32_t n,
                            const chk_batch_tuning *tuning, double *latency_ms,
                            chk_batch_stats *stats, void *stream);
/* Autotuner: times every chunk_paths (2^22 .. 2^25) x trace_block (64, 128)
 * candidate on the batch (scratch accumulators; the requests' own images are
 * not touched), then eager launches against graph replay for the fastest;
 * the graph is chosen unless it is more than 1 % slower.  best_ms (may be
 * NULL): its batch time. */
chk_status chk_batch_autotune(chk_ctx_t ctx, const chk_request *reqs, uint32_t n,
                              chk_batch_tuning *best, double *best_ms, void *stream);

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
/* inputs k x 7, probs k x 32 (q', 16-byte aligned), mix k (c), logits k x 33
 * (may be NULL): dev. */
chk_status chk_proposal_eval(chk_proposal_t p, const float *inputs, uint32_t k, float *probs,
                             float *mix, float *logits, void *stream);
/* Parameter gradient (dev float[14625], overwritten) of a loss whose
 * gradient with respect to the 33 logits is dlogits (k x 33, dev). */
chk_status chk_proposal_grad(chk_proposal_t p, const float *inputs, const float *dlogits,
                             uint32_t k, float *grad, void *stream);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) with learning rate lr. */
chk_status chk_proposal_adam(chk_proposal_t p, const float *grad, float lr, void *stream);

/* One training record of chk_render_learned: a vertex's light sample and
 * BSDF sample as the learned render writes them (40 bytes).
 *   light     light j drawn from q' (0xFFFFFFFF: no light sample)
 *   gl        f cos * albedo * radiance of the light sample, before visibility
 *   pl, pbl   light-sampling (q'_j times the area pdf) and BSDF pdfs of it
 *   fb        f cos * albedo of the BSDF sample
 *   le        emitted radiance the BSDF sample's ray found (0: none)
 *   plb, pbb  light-sampling and BSDF pdfs of the BSDF sample
 *   sh_slot   index of the light sample's occlusion flag (0xFFFFFFFF: none)
 *   light_b   light the BSDF sample hit (0xFFFFFFFF: none) */
typedef struct chk_train_record {
  uint32_t light;
  float gl, pl, pbl, fb, le, plb, pbb;
  uint32_t sh_slot, light_b;
} chk_train_record;

/* The loss gradient chk_render_learned's online training feeds to
 * chk_proposal_grad, for k records at the proposal's current parameters
 * (no reference counterpart).  Per record, with q'_c and c the current
 * network's outputs for inputs[i] and q'_s = probs[i] the pmf the samples
 * were drawn with:
 *   light head: a = gl V / pl (V = 1 - occluded[sh_slot]) times
 *     grad_z(-log q'_c,j) -- the KL divergence from the light-contribution
 *     distribution to q'_c -- or, with CHK_TRAIN_LIGHTS_CHI2, a^2 q'_s,j /
 *     q'_c,j times it (the second moment of the light estimate);
 *   mixing head: d F^2 / dz_32 of the vertex's two-sample estimate
 *     F = c g_l / (c pl + (1-c) pbl) + (1-c) g_b / (c plb + (1-c) pbb);
 *   the light head's rows divided by the batch's sum of their weights (a,
 *   or a^2 q'_s,j / q'_c,j), the mixing head's by the batch's sum of
 *   |d F^2 / dz_32|.
 * rec, inputs (k x 7), probs (k x 32), occluded (indexed by sh_slot),
 * dlogits (k x 33, out): dev.  k <= 65536; flags: CHK_TRAIN_*. */
chk_status chk_proposal_loss_grad(chk_proposal_t p, const chk_train_record *rec,
                                  const float *inputs, const float *probs,
                                  const uint8_t *occluded, uint32_t k, uint32_t flags,
                                  float *dlogits, void *stream);

/* Online training inside chk_render_learned. */
typedef struct chk_train_params {
  uint32_t train_spp;        /* calls with params->sample_offset < train_spp train */
  uint32_t batch;            /* records per Adam step, 16..65536 */
  uint32_t steps_per_bounce; /* Adam steps after each bounce while training */
  float lr;                  /* Adam learning rate */
  uint32_t flags;            /* CHK_TRAIN_* */
} chk_train_params;
#define CHK_TRAIN_FREEZE_MIX 1u    /* no gradient into the mixing logit */
#define CHK_TRAIN_FREEZE_LIGHTS 2u /* no gradient into the light logits */
#define CHK_TRAIN_LIGHTS_CHI2 4u   /* light head: second moment of the NEE
                                      estimate instead of the KL divergence to
                                      the light-contribution distribution */

/* chk_render with the learned proposal: at every vertex the network (for
 * the scene's n_lights lights) gives q' and c; the light of next-event
 * estimation is drawn from q', and the light and BSDF samples are combined
 * with weights c p_l / (c p_l + (1-c) p_b) and its complement, which keeps
 * the estimator unbiased for any network output.  params->mis_mode must be
 * CHK_MIS_LEARNED.  With train != NULL and params->sample_offset <
 * train->train_spp the network is trained online on the render's own
 * samples: after each bounce, train->steps_per_bounce Adam steps on the
 * previous bounce's records (light choice, its shadowed contribution, and
 * the emission found by the BSDF sample), minimising the KL divergence from
 * the light-contribution distribution to q' and the second moment of the
 * vertex's direct-light estimate over c.  An untrained proposal (zero output layer, q' = 1/L, c = 1/2)
 * is the estimator of chk_render; the images differ only where the float
 * CDF of q' puts a light-selection boundary one ulp away from floor(u L). */
chk_status chk_render_learned(chk_scene_t scene, const chk_camera *camera,
                              const chk_render_params *params, chk_proposal_t proposal,
                              const chk_train_params *train, double *accum, double *sumsq,
                              chk_render_stats *stats, void *stream);

#ifdef __cplusplus
}
#endif
#endif
