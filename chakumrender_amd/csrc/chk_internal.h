// Internal structures shared by the host API and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "chakum_c.h"
#include "chk_bvh.h"

namespace chk {

void set_error(const std::string &msg);

#define CHK_HIP(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::chk::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));         \
      return CHK_EDEVICE;                                                          \
    }                                                                              \
  } while (0)

// Entry points catch every C++ exception: none crosses the C ABI.
#define CHK_TRY_BEGIN try {
#define CHK_TRY_END                                   \
  }                                                   \
  catch (const std::bad_alloc &) {                    \
    set_error("host allocation failed");              \
    return CHK_ENOMEM;                                \
  }                                                   \
  catch (const std::exception &ex) {                  \
    set_error(std::string("internal error: ") + ex.what()); \
    return CHK_EINTERNAL;                             \
  }                                                   \
  catch (...) {                                       \
    set_error("internal error");                      \
    return CHK_EINTERNAL;                             \
  }

// Device-resident light data for next-event estimation.
struct LightTri {
  float4 a_cdf;    // A, cumulative area fraction within its light
  float4 b_area;   // B, triangle area
  float4 c_leaf;   // C, leaf index (raw bits)
  float4 rad;      // radiance (emitter material); w: last bin of the
                   // reference's 300-bin inverse CDF that maps to this
                   // triangle (MeshLight::initMeshLight, lightSources.cpp:784-807)
};
// FP64 constants of one light triangle for Triangle::intersect
// (geometryPrimitives.cpp:29-71), from its FP32 vertices: A, v1 = B - A,
// v2 = C - A, the unit normal (C-A)x(B-A), b0 = v1.v1, b1 = v1.v2 (= c0),
// c1 = v2.v2 and den = c0 b1 - c1 b0.
struct LightTri64 {
  double a[3], v1[3], v2[3], n[3];
  double b0, b1, c1, den;
};
struct LightRec {
  uint32_t first;  // into the LightTri table
  uint32_t count;
  float area;      // total area of the mesh light
  float pad;
};
// An environment light (EnvSource, lightSources.h:33-109) with the tables
// of its initSamples (lightSources.cpp:740-782).
struct EnvDev {
  const double *img = nullptr;      // h x w x 3
  const uint32_t *inv_m = nullptr;  // inverse CDF of the marginal, kEnvBins entries
  const uint32_t *inv_c = nullptr;  // inverse CDFs of the conditionals, kEnvRows x kEnvBins
  double rad[3] = {0, 0, 0};        // BaseSource radiance (col.maxnorm() * intensity)
  double area = 0;                  // normalizationArea
  uint32_t w = 0, h = 0, axis = 0, on = 0;
};
// A sphere primitive in FP64 (Sphere, geometryPrimitives.h:38-52).
struct SphereRec {
  double c[3];
  double r;
};

struct DeviceScene {
  const Node4Q *nodes = nullptr;
  const float4 *tris = nullptr;        // 3 float4 per leaf triangle
  const uint2 *leaf_info = nullptr;    // (material, light or 0xFFFFFFFF)
  const chk_material *mats = nullptr;
  const LightTri *light_tris = nullptr;
  const LightRec *lights = nullptr;
  const LightTri64 *light64 = nullptr;  // parallel to light_tris
  const uint2 *orig_info = nullptr;    // per original triangle: (leaf or 0xFFFFFFFF, material)
  const double *light_area = nullptr;  // per mesh light: MeshLight::area in FP64
  uint32_t n_lights = 0;               // mesh lights: LightRec 0 .. n_lights - 1
  // Sources beyond the mesh lights and the sphere list, used by the direct-
  // lighting estimators and the reference shaders.  Triangle lights are
  // LightRec n_lights .. n_lights + n_tri_lights - 1 (one triangle each).
  const SphereRec *spheres = nullptr;
  const uint32_t *sphere_mat = nullptr;
  const uint32_t *sph_src = nullptr;   // SphereSource j -> its sphere
  const chk_point_source *points = nullptr;
  uint32_t n_tri_lights = 0, n_spheres = 0, n_sph_src = 0, n_points = 0;
  EnvDev env;
  uint32_t n_leaf_tris = 0;
  uint32_t n_tris = 0;                 // original triangle count
  // box of the BVH triangles (degenerate ones excluded): the learned
  // proposal's position input is (x - box_lo) * box_inv in [0, 1]
  float box_lo[3] = {0, 0, 0}, box_inv[3] = {1, 1, 1};
};

}  // namespace chk

struct chk_ctx_s {
  int device = 0;
  int n_cu = 256;
  double *ltc_dev = nullptr;  // packed LTC table (kLtcRec doubles per record, chk_ltc_dev.h)
  void *batch_state = nullptr;  // queues, tables and cached graphs of chk_render_batch
};

struct chk_scene_s {
  chk_ctx_s *ctx = nullptr;
  chk::DeviceScene d;
  void *blob = nullptr;        // one device allocation holding all scene arrays
  size_t blob_bytes = 0;
  chk_scene_info info{};
  int stack_size = 32;
  // render scratch (grown on demand)
  void *work = nullptr;
  size_t work_bytes = 0;
  void *lwork = nullptr;  // learned-proposal buffers of the render
  size_t lwork_bytes = 0;
  // a second path-queue set and stream: chk_render runs every other chunk
  // there, so one chunk's last, sparse bounces overlap the next one's first
  void *work2 = nullptr, *lwork2 = nullptr;
  size_t work2_bytes = 0, lwork2_bytes = 0;
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  void *dwork = nullptr;  // direct-lighting estimators' chunk buffers
  size_t dwork_bytes = 0;
  uint32_t n_mats = 0;
  // spheres, triangle, sphere or point sources: only the direct-lighting
  // entry points and chk_render_shade handle them
  bool extras = false;
  void *env_mem = nullptr;  // environment image and tables
  // the first non-emitter material without a BSDF sampler (-1: none)
  int shallow_stack = 16;  // chk_scene_set_trace_stack (8, 16 or 32)
  // process-unique creation number: a batch's cached graphs are keyed on it,
  // so a new scene at a freed scene's address never replays them
  uint64_t gen = 0;
  // accelerator of the trace entry points (chk_scene_set_accel); the run
  // boxes of CHK_ACCEL_DUMMY, built on first use
  uint32_t accel = 0;
  float4 *dummy_box = nullptr;
  uint32_t dummy_runs = 0;
  int bad_sampler_mat = -1;
  uint32_t bad_sampler_brdf = 0;
};

namespace chk {

// Work layout of a traversal launch: kSegs segments of segcap items each
// (item (r, k) at index r * segcap + k).  Segment r holds count[r *
// kSegStride] items when count is set, else the items of [r * segcap,
// (r + 1) * segcap) below n (a plain array of n items).  segcap is a
// multiple of 512 (the largest block size of the consumers).
constexpr uint32_t kSegs = 8;
constexpr uint32_t kSegStride = 64;  // counters 256 bytes apart
struct SegRange {
  const uint32_t *count;
  uint32_t n, segcap;
  __host__ __device__ uint32_t count_of(uint32_t r) const {
    if (count) return count[r * kSegStride];
    const uint32_t lo = r * segcap;
    return n > lo ? (n - lo < segcap ? n - lo : segcap) : 0u;
  }
};
inline uint32_t seg_capacity(uint32_t n) {
  const uint32_t per = (uint32_t)(((uint64_t)n + kSegs - 1) / kSegs);
  return (per + 511u) & ~511u;  // a multiple of every block size used (<= 512)
}
constexpr uint32_t kMaxTraceBatch = 1u << 31;  // rays per trace call
// Spheres are tested by brute force in FP64 (the reference's sphere BVH
// gives the same closest hit), so their number is bounded; every triangle,
// sphere or point source costs sample slots per shading point.
constexpr uint32_t kMaxSpheres = 4096;
constexpr uint32_t kMaxSources = 1024;
// An environment image's pixel index must fit an emitter code (28 bits).
constexpr uint64_t kMaxEnvPixels = 1ull << 28;
// Builds the environment light's device image and tables into sc (scene
// creation; synchronises).  Returns false with a message on failure.
bool build_env(chk_scene_s *sc, const chk_env_source &E, std::string &err);
inline SegRange linear_range(uint32_t n) { return SegRange{nullptr, n, seg_capacity(n)}; }

// Kernel launchers (chk_trace.hip).
// Launch shape: LDS stack entries per lane (32, 64 or 96, at least the
// scene's stack bound) and threads per block (64, 128 or 256).
struct TraceShape {
  int stack;
  int block;
  // Optional overflow list (ovf[0] = count, ovf[64 ...] = ray indices, room
  // for ovf_cap rays, at least the launch's ray slots).  With it, a tree
  // whose worst-case stack exceeds 32 entries is traced with a 32-entry stack
  // and the rays that overflow it are traced again with the full stack.
  uint32_t *ovf = nullptr;
  uint32_t ovf_cap = 0;
  int shallow = 16;  // the shorter stack: 16 or 32 (8 only for single-scene 128-thread launches)
  uint32_t stream = 0;  // >0: ray-replacement traversal with this many blocks per segment
};
// Blocks per segment of the ray-replacement traversal in the renders: the
// closest-hit rays after the camera rays and the shadow rays.  Camera rays
// keep the one-ray-per-lane kernel: a wave's 64 samples of one pixel walk
// the same nodes, and replacing finished rays would break that up.
constexpr uint32_t kStreamBlocksClosest = 1024, kStreamBlocksAny = 768;
// Words of an overflow list for n ray slots.
inline size_t ovf_words(uint32_t n) { return 64 + (size_t)n; }
// Per-ray scenes of a batched launch: ray i belongs to scene[slot_seg[i]]
// (scene == nullptr: every ray uses the launch's scene).
struct SlotScenes {
  const DeviceScene *scene;
  const uint32_t *slot_seg;
};
// Closest hit over rays (chk_ray layout) laid out by W.  leaf_ids:
// report/skip leaf indices instead of original ones.  counters (dev, 4 x
// u64, may be null): [node visits, tri tests] accumulated when `count`,
// [2] = lanes whose stack overflowed, [3] += rays traced.  s may be null
// when M.scene is set.
hipError_t launch_closest(const chk_scene_s *s, const TraceShape &T, const float4 *rays,
                          const SegRange &W, float4 *hits, bool leaf_ids, bool count,
                          unsigned long long *counters, const SlotScenes &M, hipStream_t st);
// Any hit.  occluded (may be null) receives 0/1; with path_L, unoccluded
// rays add shadow_L[i].xyz to path_L[bits(shadow_L[i].w)].
hipError_t launch_any(const chk_scene_s *s, const TraceShape &T, const float4 *rays,
                      const SegRange &W, uint8_t *occluded, const float4 *shadow_L,
                      float4 *path_L, bool leaf_ids, bool count, unsigned long long *counters,
                      const SlotScenes &M, hipStream_t st);
hipError_t launch_zero_u32(uint32_t *p, uint32_t n, hipStream_t st);
// DummyAccel (chk_scene_set_accel): the run boxes of the scene's original
// triangle order, and closest / any hit over them.
hipError_t launch_dummy_boxes(const chk_scene_s *s, uint32_t n_runs, uint32_t run_len,
                              float4 *box, hipStream_t st);
hipError_t launch_dummy_trace(const chk_scene_s *s, bool any, const float4 *box, uint32_t n_runs,
                              const float4 *rays, uint32_t n, float4 *hits, uint8_t *occluded,
                              hipStream_t st);
// Closest sphere per ray (spheres: ns x 4 floats, centre and radius).
hipError_t launch_spheres(const float *spheres, uint32_t ns, const float4 *rays, uint32_t n,
                          float4 *hits, hipStream_t st);

// Learned proposal (chk_mlp.hip).  Inference over the items of R: inputs
// (7 per item), probs (32 per item), mix (1 per item), logits (33, may be null).
struct TrainRec;
hipError_t launch_mlp_eval(const chk_proposal_s *p, const float *inputs, const SegRange &R,
                           float *probs, float *mix, float *logits, hipStream_t st);
// One Adam step on `batch` records of a finished bounce (see chk_mlp.hip).
chk_status proposal_train_step(chk_proposal_s *p, const SegRange &R, const TrainRec *rec,
                               const float *inputs, const float *probs, const float *mix,
                               const uint8_t *s_occ, uint32_t batch, uint32_t stride,
                               uint32_t offset, float lr, uint32_t flags, hipStream_t st);
uint32_t proposal_n_lights(const chk_proposal_s *p);
const chk_ctx_s *proposal_ctx(const chk_proposal_s *p);

int pick_stack_size(uint32_t bound);

// Caller-provided render workspace: the path queues, counters and overflow
// list of render_impl carved from [ptr, ptr + bytes) instead of the scene's
// growable buffer; the chunk size is the largest that fits.
struct Workspace {
  void *ptr;
  size_t bytes;
};
// The render path (chk_render.hip).  ws == nullptr: the scene's own buffers,
// grown with hipMalloc on demand.
chk_status render_impl(chk_scene_s *sc, const chk_camera *cam, const chk_render_params *P,
                       double *accum, double *sumsq, chk_render_stats *stats, bool count,
                       chk_proposal_s *prop, const chk_train_params *train, hipStream_t st,
                       const Workspace *ws = nullptr);
// Looks up this device's copy of the fixed render scratch array (called by
// chk_create with the device current, once per device; later calls return
// the first result).
hipError_t noalloc_scratch_init(int device);

// Batched requests (chk_render.hip).
chk_status batch_impl(chk_ctx_s *ctx, const chk_request *R, uint32_t n,
                      const chk_batch_tuning *tuning, double *latency_ms, chk_batch_stats *stats,
                      hipStream_t st);
chk_status batch_autotune(chk_ctx_s *ctx, const chk_request *R, uint32_t n,
                          chk_batch_tuning *best, double *best_ms, hipStream_t st);
void batch_state_free(void *state);

}  // namespace chk
