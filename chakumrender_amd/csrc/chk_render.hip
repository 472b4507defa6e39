// Wavefront integrator for the configs' estimator (path tracing with
// next-event estimation and balance-heuristic MIS), FP32 shading and FP64
// per-pixel accumulation.
//
// The reference renders one camera ray per pixel through a recursive
// shade() (main.cpp:66-85, shader.cpp:362).  Here a render is split into
// chunks of whole pixels x spp paths; per chunk:
//   k_reset, k_camera                              camera rays  (main.cpp:72-79)
//   repeat depth = 0..max_depth:
//     k_reset_bounce
//     closest hit over the live-ray queue          (bvhAccel.cpp:200-249)
//     k_shade: emission + MIS weight, one light sample, one BSDF sample;
//              survivors and shadow rays are appended to compacted queues
//              by a block-wide ballot prefix + one atomic per block
//   Every queue is split into kSegs segments with their own counters; the
//   blocks of class r = blockIdx % kSegs read and append segment r only, so
//   a path stays in one segment (one XCD under round-robin dispatch) for its
//   whole life and appends never contend across XCDs.
//     any hit over the shadow queue                (shader.cpp:76-97)
//   k_accumulate: per-pixel FP64 sums, G lanes per pixel (main.cpp:62,83)
// Every path owns one FP32 radiance slot; queue order never changes a
// path's result (its randomness is keyed by pixel, sample and depth).
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>

#include "chk_device.h"
#include "chk_internal.h"
#include "chk_mlp.h"

namespace chk {
namespace {

constexpr uint32_t kChunkPaths = 1u << 24;  // paths in flight per chunk of a training pass
// and of the other renders: 2^25 measured 1.4 % faster than 2^24 on the
// 1080p headline with two streams (2^26: 7 % slower)
constexpr uint32_t kChunkPathsPlain = 1u << 25;
constexpr int kShadeBlock = 256;

// control words in the work buffer: per-segment queue counters, each
// segment's counter kSegStride words (256 bytes) after the previous one
enum Ctrl : uint32_t {
  kQ0 = 0,                        // live-ray counts, queue 0
  kQ1 = kSegs * kSegStride,       // live-ray counts, queue 1
  kQS = 2 * kSegs * kSegStride,   // shadow-ray counts
  kQP = 3 * kSegs * kSegStride,   // previous bounce's live-ray counts (training)
  kCtrlWords = 4 * kSegs * kSegStride
};
// 64-bit statistics words: closest-hit group, then any-hit group, in the
// order of the traversal kernels' counters
enum Stat : uint32_t {
  kNodeVisits = 0,
  kTriTests = 1,
  kOverflow = 2,
  kRaysExt = 3,
  kShNodeVisits = 4,
  kShTriTests = 5,
  kShOverflow = 6,
  kRaysShadow = 7,
  kStatWords = 8
};

struct CamArgs {
  float pos[3], dir[3], cx[3], cy[3];
  uint32_t W, H;
};

// Arguments of one request segment: the request's scene, camera and
// parameters, and the pixels [pix0, pix0 + n_pix) whose paths occupy chunk
// path ids [p0, p0 + n_pix * spp).
struct RenderArgs {
  DeviceScene S;
  CamArgs cam;
  uint32_t spp, sample_offset, max_depth, jitter_mode, flags, key0, key1, req;
  uint32_t pix0, n_pix, p0;  // pixels: indices into the window, row-major
  uint32_t win_x0, win_y0, win_w, win_h;
  float inv_nl;
};

// Image pixel (row-major index into the W x H image, row 0 at the top) of
// window pixel li: keys and accumulators use it, so a window renders its
// pixels exactly as the whole image does.
__host__ __device__ inline uint32_t image_pixel(const RenderArgs &A, uint32_t li) {
  if (A.win_w == A.cam.W) return A.win_y0 * A.cam.W + li;  // full rows
  return (A.win_y0 + li / A.win_w) * A.cam.W + A.win_x0 + li % A.win_w;
}

struct Buffers {
  uint32_t *ctrl;
  unsigned long long *stats;
  float4 *q_ray[2];
  float4 *q_thr[2];
  uint32_t *q_pid[2];
  float4 *hits;
  float4 *s_ray;
  float4 *s_L;
  float4 *path_L;
  uint32_t *ovf;  // the traversal's overflow list (TraceShape::ovf), qcap rays
};

// qcap: queue slots (kSegs * segcap); cap: paths per chunk
size_t buffers_bytes(uint32_t qcap, uint32_t cap) {
  size_t b = (size_t)kCtrlWords * 4 + 256 + 256;
  b += 2 * (size_t)qcap * (32 + 16 + 4);
  b += (size_t)qcap * (16 + 32 + 16);
  b += (size_t)cap * 16;
  b += ovf_words(qcap) * 4;
  return b + 16 * 256;
}

// The largest chunk (a multiple of 4096 paths, at most kChunkPaths) whose
// buffers fit in `bytes`; 0 when not even 4096 paths fit.
uint32_t workspace_chunk_paths(size_t bytes) {
  uint32_t lo = 0, hi = kChunkPaths / 4096u;
  while (lo < hi) {  // largest k with buffers_bytes(qcap(4096 k), 4096 k) <= bytes
    const uint32_t mid = (lo + hi + 1) / 2, cap = mid * 4096u;
    if (buffers_bytes(kSegs * seg_capacity(cap), cap) <= bytes)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo * 4096u;
}

Buffers carve(void *base, uint32_t qcap, uint32_t cap) {
  Buffers B;
  char *p = (char *)base;
  auto take = [&](size_t bytes) {
    char *r = p;
    p += (bytes + 255) & ~(size_t)255;
    return (void *)r;
  };
  B.ctrl = (uint32_t *)take(kCtrlWords * 4);
  B.stats = (unsigned long long *)take(kStatWords * 8);
  for (int k = 0; k < 2; ++k) {
    B.q_ray[k] = (float4 *)take((size_t)qcap * 32);
    B.q_thr[k] = (float4 *)take((size_t)qcap * 16);
    B.q_pid[k] = (uint32_t *)take((size_t)qcap * 4);
  }
  B.hits = (float4 *)take((size_t)qcap * 16);
  B.s_ray = (float4 *)take((size_t)qcap * 32);
  B.s_L = (float4 *)take((size_t)qcap * 16);
  B.path_L = (float4 *)take((size_t)cap * 16);
  B.ovf = (uint32_t *)take(ovf_words(qcap) * 4);
  return B;
}

// Learned-proposal buffers, double-buffered by bounce parity.
struct LBuffers {
  uint32_t *prev[2];
  float *inputs[2], *probs[2], *mix[2];
  TrainRec *rec[2];
  uint8_t *s_occ;
};

size_t lbuffers_bytes(uint32_t qcap) {
  return 2 * (size_t)qcap * (4 + 28 + 128 + 4 + sizeof(TrainRec)) + (size_t)qcap + 16 * 256;
}

LBuffers lcarve(void *base, uint32_t qcap) {
  LBuffers L;
  char *p = (char *)base;
  auto take = [&](size_t bytes) {
    char *r = p;
    p += (bytes + 255) & ~(size_t)255;
    return (void *)r;
  };
  for (int k = 0; k < 2; ++k) {
    L.prev[k] = (uint32_t *)take((size_t)qcap * 4);
    L.inputs[k] = (float *)take((size_t)qcap * 28);
    L.probs[k] = (float *)take((size_t)qcap * 128);
    L.mix[k] = (float *)take((size_t)qcap * 4);
    L.rec[k] = (TrainRec *)take((size_t)qcap * sizeof(TrainRec));
  }
  L.s_occ = (uint8_t *)take(qcap);
  return L;
}

// Block-wide compacting append into two queues: every thread of the block
// calls it once; threads with pa (pb) get consecutive slots after the
// block's base in queue a (b), ordered by key ka (kb) in 0..kBins-1 and, for
// equal keys, by thread.  The keys are the rays' direction octants: the
// waves of the next traversal then hold rays of one octant from a few
// neighbouring paths, which walk the same nodes in the same order.  One
// atomic per queue per block; the order is a function of the block's
// contents only.
constexpr uint32_t kBins = 8;
__device__ inline uint32_t octant(float4 dir) {
  return (dir.x < 0.0f ? 1u : 0u) | (dir.y < 0.0f ? 2u : 0u) | (dir.z < 0.0f ? 4u : 0u);
}
template <int NW>
__device__ inline void block_append2(bool pa, bool pb, uint32_t ka, uint32_t kb, uint32_t *ca,
                                     uint32_t *cb, uint32_t &sa, uint32_t &sb) {
  __shared__ uint32_t cnt[2][kBins][NW];
  __shared__ uint32_t base[2];
  const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  unsigned long long mine_a = 0, mine_b = 0;
#pragma unroll
  for (uint32_t k = 0; k < kBins; ++k) {
    const unsigned long long ma = __ballot(pa && ka == k), mb = __ballot(pb && kb == k);
    mine_a = ka == k ? ma : mine_a;
    mine_b = kb == k ? mb : mine_b;
    if (lane == 0) {
      cnt[0][k][w] = (uint32_t)__popcll(ma);
      cnt[1][k][w] = (uint32_t)__popcll(mb);
    }
  }
  __syncthreads();
  // exclusive scan over (bin, wave) in that order, then the block's base
  if (threadIdx.x < 2) {
    uint32_t tot = 0;
    for (uint32_t k = 0; k < kBins; ++k)
      for (int v = 0; v < NW; ++v) {
        const uint32_t c = cnt[threadIdx.x][k][v];
        cnt[threadIdx.x][k][v] = tot;
        tot += c;
      }
    base[threadIdx.x] = tot ? atomicAdd(threadIdx.x ? cb : ca, tot) : 0u;
  }
  __syncthreads();
  sa = base[0] + cnt[0][pa ? ka : 0][w] + (uint32_t)__popcll(mine_a & below);
  sb = base[1] + cnt[1][pb ? kb : 0][w] + (uint32_t)__popcll(mine_b & below);
}

__global__ void k_reset(uint32_t *ctrl, unsigned long long *stats, int clear_stats) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kCtrlWords) ctrl[i] = 0u;
  if (clear_stats && i < kStatWords) stats[i] = 0ull;
}

// zero the output queue's and the shadow queue's segment counters; the
// output queue held the previous bounce's live rays, whose counts are kept
// in kQP for training on that bounce's records
__global__ void k_reset_bounce(uint32_t *ctrl, uint32_t out_q) {
  const uint32_t t = threadIdx.x;
  if (t < kSegs) {
    ctrl[kQP + t * kSegStride] = ctrl[out_q + t * kSegStride];
    ctrl[out_q + t * kSegStride] = 0u;
  } else if (t < 2 * kSegs) {
    ctrl[kQS + (t - kSegs) * kSegStride] = 0u;
  }
}

// Camera rays: main.cpp:72-79 with prepareScene.cpp:24-33's cx, cy.
// Live-ray counts of the camera queue: the chunk's `total` paths, path p
// in queue segment p / per, per = ceil(total / kSegs).
__global__ void k_queue_counts(uint32_t *__restrict__ ctrl, uint32_t total) {
  const uint32_t r = threadIdx.x;
  const uint32_t per = (total + kSegs - 1) / kSegs;
  if (r < kSegs) {
    const uint32_t lo = r * per;
    ctrl[kQ0 + r * kSegStride] = total > lo ? min(per, total - lo) : 0u;
  }
}

// Camera rays of one request segment (paths p0 .. p0 + n_pix * spp of a
// chunk of `total` paths).  q_seg (batched launches): the segment index of
// every slot.
__global__ void k_camera(RenderArgs A, uint32_t segcap, uint32_t total,
                         float4 *__restrict__ q_ray, float4 *__restrict__ q_thr,
                         uint32_t *__restrict__ q_pid, float4 *__restrict__ path_L,
                         uint32_t *__restrict__ q_seg, uint32_t seg) {
  const uint32_t n = A.n_pix * A.spp;
  const uint32_t per = (total + kSegs - 1) / kSegs;
  const uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x;
  if (lp >= n) return;
  const uint32_t p = A.p0 + lp;
  const uint32_t r = p / per;
  const size_t slot = (size_t)r * segcap + (p - r * per);
  if (q_seg) q_seg[slot] = seg;
  const uint32_t idx = image_pixel(A, A.pix0 + lp / A.spp);
  const uint32_t s = A.sample_offset + lp % A.spp;
  const uint32_t row = idx / A.cam.W, x = idx % A.cam.W;
  const uint32_t y = A.cam.H - 1u - row;
  float jx = 0.34f, jy = 0.20f;  // main.cpp:23-24, k = 0
  if (A.jitter_mode == CHK_JITTER_BOX) {
    u4 rr = philox(u4{idx, s, (0u << 2) | CHK_STREAM_CAM, A.req}, A.key0, A.key1);
    jx = u01(rr.x);
    jy = u01(rr.y);
  }
  const float fx = ((float)x + jx) / (float)A.cam.W - 0.5f;
  const float fy = ((float)y + jy) / (float)A.cam.H - 0.5f;
  f3 d = mk(A.cam.cx[0], A.cam.cx[1], A.cam.cx[2]) * fx +
         mk(A.cam.cy[0], A.cam.cy[1], A.cam.cy[2]) * fy +
         mk(A.cam.dir[0], A.cam.dir[1], A.cam.dir[2]);
  d = normalize(d);
  q_ray[2 * slot] = make_float4(A.cam.pos[0], A.cam.pos[1], A.cam.pos[2], kInf);
  q_ray[2 * slot + 1] = make_float4(d.x, d.y, d.z, __uint_as_float(0xFFFFFFFFu));
  q_thr[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
  q_pid[slot] = p;
  path_L[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__device__ inline f3 maxnorm(f3 a) {  // Vec::maxnorm (mathPrimitives.cpp:24-29)
  float m = a.x;
  if (m < a.y) m = a.y;
  if (m < a.z) m = a.z;
  return mk(a.x / m, a.y / m, a.z / m);
}

__device__ inline bool finite3(f3 a) {
  return isfinite(a.x) && isfinite(a.y) && isfinite(a.z);
}

__device__ inline float lum(f3 a) { return (a.x + a.y + a.z) * (1.0f / 3.0f); }

// Buffers of the learned proposal for one bounce (indexed like the in-queue).
struct LearnedIO {
  const float *probs, *mix;             // q' (32 per item) and c of this bounce
  const float *probs_prev, *mix_prev;   // the same for the previous bounce
  const uint32_t *in_prev;              // item's previous vertex (previous bounce index)
  uint32_t *out_prev;                   // written for appended continuations
  TrainRec *rec, *rec_prev;             // records of this / the previous bounce
};

// Per-slot request of a batched launch: tab[seg] are the arguments of the
// request segment a path belongs to.
struct BatchIO {
  const RenderArgs *tab;
  const uint32_t *in_seg;
  uint32_t *out_seg, *s_seg;
};

// Shade the vertex of queue slot i (off < total): emission with its MIS
// weight, one light sample, one BSDF sample.  Outputs the continuation ray,
// the shadow ray and, for LEARNED, the vertex's training record.
template <bool LEARNED>
__device__ __forceinline__ void shade_vertex(const RenderArgs &A, uint32_t depth, size_t i,
                                             const float4 *__restrict__ in_ray,
                                             const float4 *__restrict__ in_thr,
                                             const uint32_t *__restrict__ in_pid,
                                             const float4 *__restrict__ hits,
                                             float4 *__restrict__ path_L, const LearnedIO &P,
                                             bool &emit_ray, bool &emit_sh, float4 &nr0,
                                             float4 &nr1, float4 &nthr, float4 &sr0, float4 &sr1,
                                             float4 &sl, uint32_t &pid, TrainRec &rec) {
  const DeviceScene &S = A.S;
  const float4 h = hits[i];
  const uint32_t leaf = __float_as_uint(h.y);
  pid = in_pid[i];
  if (leaf != 0xFFFFFFFFu) {
    const float4 r1 = in_ray[2 * i + 1];
    const float4 thr4 = in_thr[i];
    const f3 d = xyz(r1);
    const f3 thr = xyz(thr4);
    const uint2 info = S.leaf_info[leaf];
    const Mat m = load_mat(S.mats + info.x);
    const float4 *tp = S.tris + 3 * (size_t)leaf;
    const f3 A0 = xyz(tp[0]), B0 = xyz(tp[1]), C0 = xyz(tp[2]);
    // reference normal (C-A)x(B-A) (geometryPrimitives.h:75-79)
    const f3 Ng = normalize(cross(C0 - A0, B0 - A0));
    if (m.light != CHK_LIGHT_NONE) {
      f3 Le;
      if (depth == 0) {
        // directly seen emitter: getRadiance = radiance.maxnorm()
        // (materialTypes.h:34-38, shader.cpp:354-356)
        Le = (A.flags & CHK_RENDER_PHYSICAL_EMITTER) ? m.rad : maxnorm(m.rad);
      } else {
        float w = 1.0f;
        if (info.y != 0xFFFFFFFFu) {
          const float cl = fabsf(dot(d, Ng));
          const float pb = thr4.w;
          const float pa = dv(h.x * h.x, S.lights[info.y].area * cl);  // mis.cpp:55-66
          if (LEARNED) {
            const uint32_t prev = P.in_prev[i];
            const float c = P.mix_prev[prev];
            const float pl = P.probs_prev[32 * (size_t)prev + info.y] * pa;
            w = dv((1.0f - c) * pb, c * pl + (1.0f - c) * pb);
            if (P.rec_prev) {
              P.rec_prev[prev].le = lum(m.rad);
              P.rec_prev[prev].plb = pl;
              P.rec_prev[prev].light_b = info.y;
            }
          } else {
            const float pl = A.inv_nl * pa;
            w = dv(pb, pb + pl);  // balance heuristic, N_light = N_bsdf = 1 (mis.cpp:209-220)
          }
        }
        Le = thr * m.rad * w;
      }
      if (finite3(Le)) {
        float4 L = path_L[pid];
        L.x += Le.x;
        L.y += Le.y;
        L.z += Le.z;
        path_L[pid] = L;
      }
    } else if (depth < A.max_depth) {
      const float bw = 1.0f - h.z - h.w;
      const f3 X = A0 * bw + B0 * h.z + C0 * h.w;
      const f3 n = dot(d, Ng) > 0.0f ? Ng * -1.0f : Ng;
      const f3 wo = d * -1.0f;
      const f3 wo_ref = normalize(n * (2.0f * dot(n, wo)) + d);  // lightSources.cpp:44-45
      const f3 org = offset_ray(X, n);
      const uint32_t pix = image_pixel(A, A.pix0 + (pid - A.p0) / A.spp);
      const uint32_t smp = A.sample_offset + (pid - A.p0) % A.spp;
      const float c = LEARNED ? P.mix[i] : 0.5f;
      // ---- next-event estimation: one point on one mesh light
      if (S.n_lights) {
        const u4 rr = philox(u4{pix, smp, (depth << 2) | CHK_STREAM_NEE, A.req}, A.key0, A.key1);
        uint32_t j;
        float qj;
        if (LEARNED) {
          // inverse CDF of q' over the scene's lights: the row (128 bytes)
          // is loaded at once, the walk runs in registers
          const float4 *q4 = reinterpret_cast<const float4 *>(P.probs + 32 * i);
          float4 qv[8];
#pragma unroll
          for (int t = 0; t < 8; ++t) qv[t] = q4[t];
          const float u = u01(rr.x);
          float acc = 0.0f;
          j = S.n_lights - 1u;
          qj = 0.0f;
          bool found = false;
#pragma unroll
          for (uint32_t k = 0; k < 32; ++k) {
            const float4 v = qv[k >> 2];
            const float qk = (k & 3) == 0 ? v.x : (k & 3) == 1 ? v.y : (k & 3) == 2 ? v.z : v.w;
            if (!found && k + 1 < S.n_lights) {
              acc += qk;
              if (u < acc) {
                j = k;
                qj = qk;
                found = true;
              }
            }
            if (k == S.n_lights - 1u && !found) qj = qk;
          }
        } else {
          j = min((uint32_t)(u01(rr.x) * (float)S.n_lights), S.n_lights - 1u);
          qj = A.inv_nl;
        }
        const LightRec L = S.lights[j];
        const float u3 = u01(rr.w);
        uint32_t k = L.first;
        while (k + 1 < L.first + L.count && u3 >= S.light_tris[k].a_cdf.w) ++k;
        const LightTri &lt = S.light_tris[k];
        const f3 la = xyz(lt.a_cdf), lb = xyz(lt.b_area), lc = xyz(lt.c_leaf);
        // MeshLight::getSamples point (lightSources.cpp:812-824)
        const float e1 = sqrtf(u01(rr.y)), e2 = u01(rr.z);
        const f3 yp = la * (1.0f - e1) + lb * (e1 * (1.0f - e2)) + lc * (e2 * e1);
        const f3 dy = yp - X;
        const float d2 = dot(dy, dy);
        const float dist = sqrtf(d2);
        const f3 wi = dy * dv(1.0f, dist);
        const float cx = dot(wi, n);
        const f3 nl = cross(lc - la, lb - la);
        const float cl = fabsf(dot(wi, nl)) * rsqrtf(dot(nl, nl));
        if (cx > 0.0f && cl > 0.0f) {
          float f, pb;
          brdf_pdf(m, n, wo_ref, wo, wi, f, pb);
          const float pl = dv(qj * d2, L.area * cl);  // mis.cpp:55-66
          const f3 lrad = xyz(lt.rad);
          // c-weighted balance heuristic; c = 1/2 gives f cos / (p_l + p_b)
          const f3 cc =
              thr * m.c * lrad * dv(f * cx * (2.0f * c), 2.0f * c * pl + 2.0f * (1.0f - c) * pb);
          if (LEARNED) {
            rec.light = j;
            rec.gl = f * cx * lum(m.c * lrad);
            rec.pl = pl;
            rec.pbl = pb;
          }
          if (finite3(cc) && (cc.x > 0.0f || cc.y > 0.0f || cc.z > 0.0f)) {
            // aim from the offset origin at the sampled point, so the ray
            // reaches the light exactly at t = sdist (shader.cpp:82-90)
            const f3 sd = yp - org;
            const float sdist = sqrtf(dot(sd, sd));
            const f3 sdir = sd * dv(1.0f, sdist);
            emit_sh = true;
            sr0 = make_float4(org.x, org.y, org.z, sdist - 1e-5f);
            sr1 = make_float4(sdir.x, sdir.y, sdir.z, __uint_as_float(leaf));
            sl = make_float4(cc.x, cc.y, cc.z, __uint_as_float(pid));
          }
        } else if (LEARNED) {
          rec.light = j;  // a zero-contribution sample still trains q'
          rec.pl = qj;
        }
      }
      // ---- BSDF sample continues the path
      {
        const u4 rr = philox(u4{pix, smp, (depth << 2) | CHK_STREAM_BSDF, A.req}, A.key0, A.key1);
        const f3 wi = brdf_sample(m, n, wo_ref, wo, u01(rr.x), u01(rr.y), u01(rr.z));
        const float cs = dot(wi, n);
        if (cs > 0.0f) {
          float f, pb;
          brdf_pdf(m, n, wo_ref, wo, wi, f, pb);
          if (pb > 0.0f && isfinite(pb) && isfinite(f)) {
            // the weight 2(1-c) times the 1/2 of the emission weight at
            // the next vertex is 1 for the plain path
            const f3 t2 = thr * m.c * dv(f * cs, pb);
            if (LEARNED) {
              rec.fb = f * cs * lum(m.c);
              rec.pbb = pb;
            }
            if (finite3(t2) && (t2.x > 0.0f || t2.y > 0.0f || t2.z > 0.0f)) {
              emit_ray = true;
              nr0 = make_float4(org.x, org.y, org.z, kInf);
              nr1 = make_float4(wi.x, wi.y, wi.z, __uint_as_float(leaf));
              nthr = make_float4(t2.x, t2.y, t2.z, pb);
            }
          }
        }
      }
    }
  }
}

// One vertex of every live path.  Block b shades items of segment
// r = b % kSegs and appends to segment r of the output queues.
// LEARNED: the light is drawn from the proposal's q' instead of uniformly,
// and light and BSDF samples are combined with the c-weighted balance
// heuristic w_l = c p_l / (c p_l + (1-c) p_b), w_b = 1 - w_l (any c in
// (0,1) keeps the two-sample estimator unbiased; c = 1/2 and q' = 1/L give
// the plain path exactly).
// MULTI: every slot names its request's arguments (Q.tab[Q.in_seg[i]]), so
// one launch shades the paths of many requests and scenes.
template <bool LEARNED, bool MULTI, int BS>
__global__ __launch_bounds__(BS) void k_shade(
    RenderArgs A0, uint32_t depth, uint32_t segcap, const float4 *__restrict__ in_ray,
    const float4 *__restrict__ in_thr, const uint32_t *__restrict__ in_pid,
    const float4 *__restrict__ hits, float4 *__restrict__ out_ray, float4 *__restrict__ out_thr,
    uint32_t *__restrict__ out_pid, float4 *__restrict__ s_ray, float4 *__restrict__ s_L,
    float4 *__restrict__ path_L, uint32_t *__restrict__ ctrl, uint32_t in_q, uint32_t out_q,
    LearnedIO P, BatchIO Q) {
  const uint32_t r = blockIdx.x % kSegs, qb = blockIdx.x / kSegs;
  const uint32_t total = ctrl[in_q + r * kSegStride];
  if (qb * BS >= total) return;  // whole block idle (uniform)
  const uint32_t off = qb * BS + threadIdx.x;
  const size_t i = (size_t)r * segcap + off;
  bool emit_ray = false, emit_sh = false;
  float4 nr0, nr1, nthr, sr0, sr1, sl;
  uint32_t pid = 0, seg = 0;
  TrainRec rec;
  rec.light = 0xFFFFFFFFu;
  rec.gl = rec.pl = rec.pbl = rec.fb = rec.le = rec.plb = rec.pbb = 0.0f;
  rec.sh_slot = 0xFFFFFFFFu;
  rec.light_b = 0xFFFFFFFFu;
  if (off < total) {
    if (MULTI) {
      seg = Q.in_seg[i];
      shade_vertex<LEARNED>(Q.tab[seg], depth, i, in_ray, in_thr, in_pid, hits, path_L, P,
                            emit_ray, emit_sh, nr0, nr1, nthr, sr0, sr1, sl, pid, rec);
    } else {
      shade_vertex<LEARNED>(A0, depth, i, in_ray, in_thr, in_pid, hits, path_L, P, emit_ray,
                            emit_sh, nr0, nr1, nthr, sr0, sr1, sl, pid, rec);
    }
  }
  uint32_t slot, sslot;
  block_append2<BS / 64>(emit_ray, emit_sh, octant(nr1), octant(sr1), ctrl + out_q + r * kSegStride,
                         ctrl + kQS + r * kSegStride, slot, sslot);
  const size_t seg0 = (size_t)r * segcap;
  if (emit_ray) {
    const size_t o = seg0 + slot;
    out_ray[2 * o] = nr0;
    out_ray[2 * o + 1] = nr1;
    out_thr[o] = nthr;
    out_pid[o] = pid;
    if (LEARNED) P.out_prev[o] = (uint32_t)i;
    if (MULTI) Q.out_seg[o] = seg;
  }
  if (emit_sh) {
    const size_t o = seg0 + sslot;
    s_ray[2 * o] = sr0;
    s_ray[2 * o + 1] = sr1;
    s_L[o] = sl;
    if (LEARNED) rec.sh_slot = (uint32_t)o;
    if (MULTI) Q.s_seg[o] = seg;
  }
  if (LEARNED && off < total && P.rec) P.rec[i] = rec;
}

// Proposal inputs of every live vertex: position in the scene box, shading
// normal and view direction as (theta/pi, phi/2pi), each in [0, 1].
__device__ inline void dir_angles(f3 v, float &a, float &b) {
  a = acosf(fminf(fmaxf(v.z, -1.0f), 1.0f)) * kInvPi;
  const float ph = atan2f(v.y, v.x) * (0.5f * kInvPi);
  b = ph < 0.0f ? ph + 1.0f : ph;
}

__global__ __launch_bounds__(256) void k_prop_in(RenderArgs A, uint32_t segcap,
                                                 const float4 *__restrict__ in_ray,
                                                 const float4 *__restrict__ hits,
                                                 const uint32_t *__restrict__ ctrl, uint32_t in_q,
                                                 float *__restrict__ inputs) {
  const DeviceScene &S = A.S;
  const uint32_t r = blockIdx.x % kSegs, qb = blockIdx.x / kSegs;
  const uint32_t off = qb * 256 + threadIdx.x;
  if (off >= ctrl[in_q + r * kSegStride]) return;
  const size_t i = (size_t)r * segcap + off;
  float s[7] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  const float4 h = hits[i];
  const uint32_t leaf = __float_as_uint(h.y);
  if (leaf != 0xFFFFFFFFu) {
    const f3 d = xyz(in_ray[2 * i + 1]);
    const float4 *tp = S.tris + 3 * (size_t)leaf;
    const f3 A0 = xyz(tp[0]), B0 = xyz(tp[1]), C0 = xyz(tp[2]);
    const f3 Ng = normalize(cross(C0 - A0, B0 - A0));
    const f3 n = dot(d, Ng) > 0.0f ? Ng * -1.0f : Ng;
    const f3 X = A0 * (1.0f - h.z - h.w) + B0 * h.z + C0 * h.w;
    s[0] = fminf(fmaxf((X.x - S.box_lo[0]) * S.box_inv[0], 0.0f), 1.0f);
    s[1] = fminf(fmaxf((X.y - S.box_lo[1]) * S.box_inv[1], 0.0f), 1.0f);
    s[2] = fminf(fmaxf((X.z - S.box_lo[2]) * S.box_inv[2], 0.0f), 1.0f);
    dir_angles(n, s[3], s[4]);
    dir_angles(d * -1.0f, s[5], s[6]);
  }
  float *o = inputs + 7 * i;
#pragma unroll
  for (int k = 0; k < 7; ++k) o[k] = s[k];
}

// Per-pixel FP64 accumulation (main.cpp:62,83).  G lanes share a pixel
// (acc_lanes: up to 8 below 4096 spp, where a chunk holds many pixels; 64
// from 4096 spp, where it holds a few hundred): lane j adds samples j,
// j + G, ... in order, then the G partial sums are combined by a fixed
// butterfly, so the result depends only on spp, and G lanes read a pixel's
// consecutive path radiances together.
__host__ __device__ inline uint32_t acc_lanes(uint32_t spp) {
  const uint32_t cap = spp >= 4096 ? 64u : 8u;
  uint32_t g = 1;
  while (g < cap && 2 * g <= spp) g *= 2;
  return g;
}
__global__ __launch_bounds__(256) void k_accumulate(RenderArgs A, const float4 *__restrict__ path_L,
                                                    double *__restrict__ accum,
                                                    double *__restrict__ sumsq) {
  const uint32_t G = acc_lanes(A.spp);
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t q = t / G, j = t % G;
  const bool live = q < A.n_pix;
  double s0 = 0, s1 = 0, s2 = 0, q0 = 0, q1 = 0, q2 = 0;
  if (live) {
    const float4 *p = path_L + A.p0 + (size_t)q * A.spp;
#pragma unroll 4
    for (uint32_t s = j; s < A.spp; s += G) {
      const float4 L = p[s];
      const double a = L.x, b = L.y, c = L.z;
      s0 += a;
      s1 += b;
      s2 += c;
      q0 += a * a;
      q1 += b * b;
      q2 += c * c;
    }
  }
  // every lane of the wave takes part (G divides 64, so groups never
  // straddle waves)
  for (uint32_t o = G / 2; o > 0; o /= 2) {
    s0 += __shfl_xor(s0, o);
    s1 += __shfl_xor(s1, o);
    s2 += __shfl_xor(s2, o);
    q0 += __shfl_xor(q0, o);
    q1 += __shfl_xor(q1, o);
    q2 += __shfl_xor(q2, o);
  }
  if (!live || j != 0) return;
  const size_t o = 3 * (size_t)image_pixel(A, A.pix0 + q);
  accum[o] += s0;
  accum[o + 1] += s1;
  accum[o + 2] += s2;
  if (sumsq) {
    sumsq[o] += q0;
    sumsq[o + 1] += q1;
    sumsq[o + 2] += q2;
  }
}

// Grid of k_accumulate for n_pix pixels of spp samples.
inline uint32_t accumulate_blocks(uint32_t n_pix, uint32_t spp) {
  return (uint32_t)(((uint64_t)n_pix * acc_lanes(spp) + 255) / 256);
}

struct Profiler {
  bool on = false;
  hipStream_t st = nullptr;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> evs;
  hipEvent_t a = nullptr;
  int kind = 0;
  void begin(int k) {
    if (!on) return;
    kind = k;
    (void)hipEventCreate(&a);
    (void)hipEventRecord(a, st);
  }
  void end() {
    if (!on) return;
    hipEvent_t b;
    (void)hipEventCreate(&b);
    (void)hipEventRecord(b, st);
    evs.push_back({kind, {a, b}});
  }
  void collect(chk_render_stats *s) {
    if (!on) return;
    for (auto &e : evs) {
      (void)hipEventSynchronize(e.second.second);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e.second.first, e.second.second);
      if (s) {
        if (e.first == 0) {
          s->extend_ms += ms;
        } else if (e.first == 1) {
          s->shade_ms += ms;
        } else if (e.first == 2) {
          s->shadow_ms += ms;
        } else if (e.first == 4) {
          s->proposal_ms += ms;
        } else {
          s->other_ms += ms;
        }
      }
      (void)hipEventDestroy(e.second.first);
      (void)hipEventDestroy(e.second.second);
    }
    evs.clear();
  }
};

// Request arguments from the ABI records: the camera basis of
// prepareScene.cpp:21-33 in double, the parameters, the scene.
chk_status make_args(const chk_scene_s *sc, const chk_camera *cam, const chk_render_params *P,
                     RenderArgs &A, const char *who) {
  std::memset(&A, 0, sizeof(A));
  A.S = sc->d;
  {
    // camera basis in double (prepareScene.cpp:21-33)
    const double PI = 3.14159265359;
    double d[3] = {cam->dir[0], cam->dir[1], cam->dir[2]};
    double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (double &v : d) v /= dl;
    double up[3] = {cam->up[0], cam->up[1], cam->up[2]};
    double cx[3] = {d[1] * up[2] - d[2] * up[1], d[2] * up[0] - d[0] * up[2],
                    d[0] * up[1] - d[1] * up[0]};
    double cl = std::sqrt(cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]);
    if (!(cl >= 1e-20)) {
      set_error(std::string(who) + ": camera up and direction are parallel");
      return CHK_EINVAL;
    }
    double sx = ((double)cam->width / cam->height) * std::tan(cam->fovx_deg * PI / 360.0) / cl;
    for (double &v : cx) v *= sx;
    double cy[3] = {cx[1] * d[2] - cx[2] * d[1], cx[2] * d[0] - cx[0] * d[2],
                    cx[0] * d[1] - cx[1] * d[0]};
    double yl = std::sqrt(cy[0] * cy[0] + cy[1] * cy[1] + cy[2] * cy[2]);
    double sy = std::tan(cam->fovy_deg * PI / 360.0) / yl;
    for (int k = 0; k < 3; ++k) {
      A.cam.pos[k] = cam->pos[k];
      A.cam.dir[k] = (float)d[k];
      A.cam.cx[k] = (float)cx[k];
      A.cam.cy[k] = (float)(cy[k] * sy);
    }
    A.cam.W = cam->width;
    A.cam.H = cam->height;
  }
  A.spp = P->spp;
  A.sample_offset = P->sample_offset;
  A.max_depth = P->max_depth;
  A.jitter_mode = P->jitter_mode;
  A.flags = P->flags;
  A.key0 = P->seed_lo;
  A.key1 = P->seed_hi;
  A.req = 0;
  A.inv_nl = sc->d.n_lights ? 1.0f / (float)sc->d.n_lights : 0.0f;
  const uint32_t *w = P->window;
  if ((w[0] | w[1] | w[2] | w[3]) == 0) {
    A.win_x0 = A.win_y0 = 0;
    A.win_w = cam->width;
    A.win_h = cam->height;
  } else {
    if (w[2] == 0 || w[3] == 0 || w[0] >= cam->width || w[1] >= cam->height ||
        w[2] > cam->width - w[0] || w[3] > cam->height - w[1]) {
      set_error(std::string(who) + ": window (x0, y0, w, h) is empty or leaves the image");
      return CHK_EINVAL;
    }
    A.win_x0 = w[0];
    A.win_y0 = w[1];
    A.win_w = w[2];
    A.win_h = w[3];
  }
  return CHK_OK;
}

}  // namespace

chk_status render_impl(chk_scene_s *sc, const chk_camera *cam, const chk_render_params *P,
                       double *accum, double *sumsq, chk_render_stats *stats, bool count,
                       chk_proposal_s *prop, const chk_train_params *train, hipStream_t st,
                       const Workspace *ws) {
  if (!cam || !P || !accum) {
    set_error("chk_render: null camera, params or accum");
    return CHK_EINVAL;
  }
  const bool learned = prop != nullptr;
  if (cam->width == 0 || cam->height == 0 || P->spp == 0) {
    set_error("chk_render: zero width, height or spp");
    return CHK_EINVAL;
  }
  if (P->mis_mode != (learned ? CHK_MIS_LEARNED : CHK_MIS_BALANCE)) {
    set_error(learned ? "chk_render_learned: params->mis_mode must be CHK_MIS_LEARNED"
                      : "chk_render: mis_mode other than CHK_MIS_BALANCE needs chk_render_learned");
    return CHK_EINVAL;
  }
  bool training = false;
  if (learned) {
    if (proposal_ctx(prop) != sc->ctx) {
      set_error("chk_render_learned: proposal and scene belong to different contexts");
      return CHK_EINVAL;
    }
    if (proposal_n_lights(prop) != sc->d.n_lights) {
      set_error("chk_render_learned: proposal has " + std::to_string(proposal_n_lights(prop)) +
                " lights, the scene " + std::to_string(sc->d.n_lights));
      return CHK_EINVAL;
    }
    if (train && P->sample_offset < train->train_spp) {
      if (train->batch < 16 || train->batch > (1u << 16) || train->steps_per_bounce == 0 ||
          !(train->lr > 0.0f) ||
          (train->flags & ~(CHK_TRAIN_FREEZE_MIX | CHK_TRAIN_FREEZE_LIGHTS | CHK_TRAIN_LIGHTS_CHI2))) {
        set_error("chk_render_learned: train batch must be 16..65536, steps >= 1, lr > 0, "
                  "known flags");
        return CHK_EINVAL;
      }
      training = true;
    }
  }
  // paths per chunk: fixed (a training pass keeps the smaller chunks, so its
  // Adam steps per pass do not depend on this choice), or the most a
  // caller's workspace holds
  const uint32_t chunk_paths =
      ws ? workspace_chunk_paths(ws->bytes) : (training ? kChunkPaths : kChunkPathsPlain);
  if (P->spp > chunk_paths) {
    set_error("chk_render: spp " + std::to_string(P->spp) + " exceeds the " +
              std::to_string(chunk_paths) + " paths per chunk" +
              (ws ? " that the fixed scratch holds" : ""));
    return CHK_ECAPACITY;
  }
  if ((uint64_t)cam->width * cam->height >= (1ull << 32)) {
    set_error("chk_render: image too large");
    return CHK_ECAPACITY;
  }
  RenderArgs A;
  {
    const chk_status as = make_args(sc, cam, P, A, "chk_render");
    if (as != CHK_OK) return as;
  }
  // pixels rendered: the window's, in row-major window order
  const uint64_t n_pix_total = (uint64_t)A.win_w * A.win_h;
  const uint32_t pix_per_chunk = std::max<uint32_t>(1u, chunk_paths / P->spp);
  const uint32_t cap = (uint32_t)std::min<uint64_t>((uint64_t)pix_per_chunk * P->spp,
                                                   n_pix_total * P->spp);
  const uint32_t segcap = seg_capacity(cap);
  const uint32_t qcap = kSegs * segcap;
  const size_t need = buffers_bytes(qcap, cap);
  void *work = sc->work;
  if (ws) {
    if (need > ws->bytes) {  // workspace_chunk_paths guarantees it fits
      set_error("chk_render: internal workspace sizing error");
      return CHK_EINTERNAL;
    }
    work = ws->ptr;
  } else if (sc->work_bytes < need) {
    if (sc->work) (void)hipFree(sc->work);
    sc->work = nullptr;
    sc->work_bytes = 0;
    if (hipMalloc(&sc->work, need) != hipSuccess) {
      set_error("chk_render: cannot allocate the path queues");
      return CHK_ENOMEM;
    }
    sc->work_bytes = need;
    work = sc->work;
  }
  Buffers B = carve(work, qcap, cap);
  CHK_HIP(launch_zero_u32(B.ovf, 2, st));  // overflow count, retries
  // Plain renders of several chunks alternate two queue sets and streams:
  // chunk c + 1 starts while chunk c finishes its last bounces, which hold
  // few rays and leave most of the GPU idle.  Chunks hold disjoint pixels,
  // so their accumulations never meet; the statistics are shared atomics.
  const uint64_t n_chunks_total = (n_pix_total + pix_per_chunk - 1) / pix_per_chunk;
  // (learned renders only when they do not train: training steps update the
  // network between the bounces of one chunk)
  const bool dual = (!learned || !training) && !ws && !count && n_chunks_total > 1 &&
                    !(P->flags & CHK_RENDER_ONE_STREAM);
  Buffers B2 = B;
  if (dual) {
    if (sc->work2_bytes < need) {
      if (sc->work2) (void)hipFree(sc->work2);
      sc->work2 = nullptr;
      sc->work2_bytes = 0;
      if (hipMalloc(&sc->work2, need) != hipSuccess) {
        set_error("chk_render: cannot allocate the second path-queue set");
        return CHK_ENOMEM;
      }
      sc->work2_bytes = need;
    }
    if (!sc->st2) CHK_HIP(hipStreamCreateWithFlags(&sc->st2, hipStreamNonBlocking));
    if (!sc->ev_fork) CHK_HIP(hipEventCreateWithFlags(&sc->ev_fork, hipEventDisableTiming));
    if (!sc->ev_join) CHK_HIP(hipEventCreateWithFlags(&sc->ev_join, hipEventDisableTiming));
    B2 = carve(sc->work2, qcap, cap);
    B2.stats = B.stats;
    CHK_HIP(launch_zero_u32(B2.ovf, 2, st));
  }
  LBuffers LB{};
  if (learned) {
    const size_t lneed = lbuffers_bytes(qcap);
    if (sc->lwork_bytes < lneed) {
      if (sc->lwork) (void)hipFree(sc->lwork);
      sc->lwork = nullptr;
      sc->lwork_bytes = 0;
      if (hipMalloc(&sc->lwork, lneed) != hipSuccess) {
        set_error("chk_render_learned: cannot allocate the proposal buffers");
        return CHK_ENOMEM;
      }
      sc->lwork_bytes = lneed;
    }
    LB = lcarve(sc->lwork, qcap);
  }
  LBuffers LB2 = LB;
  if (learned && dual) {
    const size_t lneed = lbuffers_bytes(qcap);
    if (sc->lwork2_bytes < lneed) {
      if (sc->lwork2) (void)hipFree(sc->lwork2);
      sc->lwork2 = nullptr;
      sc->lwork2_bytes = 0;
      if (hipMalloc(&sc->lwork2, lneed) != hipSuccess) {
        set_error("chk_render_learned: cannot allocate the second proposal-buffer set");
        return CHK_ENOMEM;
      }
      sc->lwork2_bytes = lneed;
    }
    LB2 = lcarve(sc->lwork2, qcap);
  }

  Profiler prof;
  prof.on = (P->flags & CHK_RENDER_PROFILE) != 0;
  prof.st = st;
  const uint32_t sgrid = kSegs * (segcap / kShadeBlock);
  uint32_t chunks = 0, ext_launches = 0, train_steps = 0;
  auto qbase = [](uint32_t q) { return q == 0 ? (uint32_t)kQ0 : (uint32_t)kQ1; };

  hipLaunchKernelGGL(k_reset, dim3((kCtrlWords + 255) / 256), dim3(256), 0, st, B.ctrl, B.stats,
                     1);
  if (dual) {
    CHK_HIP(hipEventRecord(sc->ev_fork, st));
    CHK_HIP(hipStreamWaitEvent(sc->st2, sc->ev_fork, 0));
  }
  const hipStream_t st_main = st;
  for (uint64_t pix0 = 0; pix0 < n_pix_total; pix0 += pix_per_chunk) {
    const bool second = dual && (chunks & 1u);
    const hipStream_t st_c = second ? sc->st2 : st_main;
    Buffers &Bc = second ? B2 : B;
    const LBuffers &LBc = second ? LB2 : LB;
    prof.st = st_c;
    ++chunks;
    A.pix0 = (uint32_t)pix0;
    A.n_pix = (uint32_t)std::min<uint64_t>(pix_per_chunk, n_pix_total - pix0);
    const uint32_t n_paths = A.n_pix * A.spp;
    prof.begin(3);
    hipLaunchKernelGGL(k_reset, dim3((kCtrlWords + 255) / 256), dim3(256), 0, st_c, Bc.ctrl, Bc.stats,
                       0);
    hipLaunchKernelGGL(k_queue_counts, dim3(1), dim3(64), 0, st_c, Bc.ctrl, n_paths);
    hipLaunchKernelGGL(k_camera, dim3((n_paths + 255) / 256), dim3(256), 0, st_c, A, segcap,
                       n_paths, Bc.q_ray[0], Bc.q_thr[0], Bc.q_pid[0], Bc.path_L, nullptr, 0u);
    prof.end();
    for (uint32_t depth = 0; depth <= P->max_depth; ++depth) {
      const uint32_t in = depth & 1u, out = in ^ 1u;
      hipLaunchKernelGGL(k_reset_bounce, dim3(1), dim3(64), 0, st_c, Bc.ctrl, qbase(out));
      prof.begin(0);
      hipError_t e = launch_closest(sc,
                                    TraceShape{sc->stack_size, 128, Bc.ovf, qcap, sc->shallow_stack,
                                               depth ? kStreamBlocksClosest : 0u},
                                    Bc.q_ray[in],
                                    SegRange{Bc.ctrl + qbase(in), 0u, segcap}, Bc.hits, true, count,
                                    Bc.stats + kNodeVisits, SlotScenes{}, st_c);
      prof.end();
      ++ext_launches;
      if (e != hipSuccess) {
        set_error(std::string("chk_render: closest-hit launch: ") + hipGetErrorString(e));
        return CHK_EDEVICE;
      }
      const SegRange inR{Bc.ctrl + qbase(in), 0u, segcap};
      LearnedIO io{};
      if (learned) {
        prof.begin(4);
        hipLaunchKernelGGL(k_prop_in, dim3(kSegs * (segcap / 256)), dim3(256), 0, st_c, A, segcap,
                           Bc.q_ray[in], Bc.hits, Bc.ctrl, qbase(in), LBc.inputs[in]);
        e = launch_mlp_eval(prop, LBc.inputs[in], inR, LBc.probs[in], LBc.mix[in], nullptr, st_c);
        prof.end();
        if (e != hipSuccess) {
          set_error(std::string("chk_render_learned: proposal launch: ") + hipGetErrorString(e));
          return CHK_EDEVICE;
        }
        io.probs = LBc.probs[in];
        io.mix = LBc.mix[in];
        io.probs_prev = LBc.probs[out];
        io.mix_prev = LBc.mix[out];
        io.in_prev = LBc.prev[in];
        io.out_prev = LBc.prev[out];
        // records only feed the training steps of a training pass
        io.rec = training ? LBc.rec[in] : nullptr;
        io.rec_prev = training ? LBc.rec[out] : nullptr;
      }
      prof.begin(1);
      if (learned)
        hipLaunchKernelGGL((k_shade<true, false, kShadeBlock>), dim3(sgrid), dim3(kShadeBlock), 0,
                           st_c, A, depth, segcap, Bc.q_ray[in], Bc.q_thr[in], Bc.q_pid[in], Bc.hits,
                           Bc.q_ray[out], Bc.q_thr[out], Bc.q_pid[out], Bc.s_ray, Bc.s_L, Bc.path_L,
                           Bc.ctrl, qbase(in), qbase(out), io, BatchIO{});
      else
        hipLaunchKernelGGL((k_shade<false, false, kShadeBlock>), dim3(sgrid), dim3(kShadeBlock), 0,
                           st_c, A, depth, segcap, Bc.q_ray[in], Bc.q_thr[in], Bc.q_pid[in], Bc.hits,
                           Bc.q_ray[out], Bc.q_thr[out], Bc.q_pid[out], Bc.s_ray, Bc.s_L, Bc.path_L,
                           Bc.ctrl, qbase(in), qbase(out), io, BatchIO{});
      prof.end();
      // the previous bounce's records are complete now (its shadow rays
      // were traced, this bounce found the emitters its BSDF rays hit)
      if (training && depth > 0) {
        prof.begin(4);
        const uint32_t stride = std::max(1u, cap / train->batch);
        for (uint32_t t = 0; t < train->steps_per_bounce; ++t) {
          chk_status ts = proposal_train_step(
              prop, SegRange{Bc.ctrl + kQP, 0u, segcap}, LBc.rec[out], LBc.inputs[out],
              LBc.probs[out], LBc.mix[out], LBc.s_occ, train->batch, stride,
              (uint32_t)(((uint64_t)t * stride) / train->steps_per_bounce), train->lr,
              train->flags, st_c);
          if (ts != CHK_OK) return ts;
          ++train_steps;
        }
        prof.end();
      }
      if (depth < P->max_depth) {
        prof.begin(2);
        e = launch_any(sc,
                       TraceShape{sc->stack_size, 128, Bc.ovf, qcap, sc->shallow_stack,
                                  kStreamBlocksAny},
                       Bc.s_ray,
                       SegRange{Bc.ctrl + kQS, 0u, segcap}, training ? LBc.s_occ : nullptr, Bc.s_L,
                       Bc.path_L, true, count, Bc.stats + kShNodeVisits, SlotScenes{}, st_c);
        prof.end();
        if (e != hipSuccess) {
          set_error(std::string("chk_render: any-hit launch: ") + hipGetErrorString(e));
          return CHK_EDEVICE;
        }
      }
    }
    prof.begin(3);
    hipLaunchKernelGGL(k_accumulate, dim3(accumulate_blocks(A.n_pix, A.spp)), dim3(256), 0, st_c, A,
                       Bc.path_L,
                       accum, (P->flags & CHK_RENDER_SUMSQ) ? sumsq : nullptr);
    prof.end();
  }
  CHK_HIP(hipGetLastError());
  if (dual) {
    CHK_HIP(hipEventRecord(sc->ev_join, sc->st2));
    CHK_HIP(hipStreamWaitEvent(st, sc->ev_join, 0));
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    unsigned long long h[kStatWords];
    CHK_HIP(hipMemcpyAsync(h, B.stats, sizeof(h), hipMemcpyDeviceToHost, st));
    uint32_t retries = 0, retries2 = 0;
    CHK_HIP(hipMemcpyAsync(&retries, B.ovf + 1, 4, hipMemcpyDeviceToHost, st));
    if (dual) CHK_HIP(hipMemcpyAsync(&retries2, B2.ovf + 1, 4, hipMemcpyDeviceToHost, st));
    CHK_HIP(hipStreamSynchronize(st));
    stats->stack_retries = retries + retries2;
    stats->paths = n_pix_total * P->spp;
    stats->rays_extend = h[kRaysExt];
    stats->rays_shadow = h[kRaysShadow];
    stats->node_visits = h[kNodeVisits];
    stats->tri_tests = h[kTriTests];
    stats->shadow_node_visits = h[kShNodeVisits];
    stats->shadow_tri_tests = h[kShTriTests];
    stats->chunks = chunks;
    stats->extend_launches = ext_launches;
    stats->train_steps = train_steps;
    prof.collect(stats);
    if (h[kOverflow] || h[kShOverflow]) {
      set_error("chk_render: traversal stack overflow");
      return CHK_EINTERNAL;
    }
  } else {
    prof.collect(nullptr);
  }
  return CHK_OK;
}

// ====================================================================== batching
// Many requests (scenes, cameras, parameters) in one wavefront: requests are
// ordered by cost (W * H * spp, shortest first).  Requests below a quarter
// of chunk_paths are packed into shared chunks of at most chunk_paths paths,
// which hold segments (pixel ranges) of several requests: every queue slot
// carries its segment index, so the traversal kernels find the ray's scene
// and the shading kernel the request's arguments.  A larger request gets
// chunks of its own and runs the single-scene kernels, as chk_render would.
// Per chunk the sequence is the single-request one.  With use_graph the
// whole batch -- every chunk on its alternating stream, the fork and join
// between the two streams, the statistics reset and the per-chunk time
// stamps -- is captured into one hipGraph, and a call is one hipGraphLaunch;
// the executable graph is kept until a call with other requests captures
// again or the buffers it names are reallocated.  Eager calls in between
// leave it valid: every call re-uploads its own tables before it launches
// anything.  Per-request latencies come from device time stamps written
// after each chunk (the same kernels in both modes).

namespace {

struct BatchSeg {
  uint32_t req, pix0, n_pix, p0;
};
struct BatchChunk {
  uint32_t first_seg, n_seg, total;
};

struct BatchState {
  void *work = nullptr, *work2 = nullptr;  // two path-queue sets (chunks alternate)
  size_t work_bytes = 0, work2_bytes = 0;
  hipStream_t st2 = nullptr;  // the stream of every other chunk
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  void *tab = nullptr;
  size_t tab_bytes = 0;
  std::vector<unsigned char> key;  // request records + tuning of the cached graph
  hipGraphExec_t exec = nullptr;   // the whole batch, captured
  hipStream_t cap = nullptr, cap2 = nullptr;  // the capture's two streams
  hipEvent_t cev_fork = nullptr, cev_join = nullptr;
  hipEvent_t ev_start = nullptr, ev_end = nullptr;
  void drop_graphs() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key.clear();
  }
  ~BatchState() {
    drop_graphs();
    if (cap) (void)hipStreamDestroy(cap);
    if (cap2) (void)hipStreamDestroy(cap2);
    if (cev_fork) (void)hipEventDestroy(cev_fork);
    if (cev_join) (void)hipEventDestroy(cev_join);
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_end) (void)hipEventDestroy(ev_end);
    if (st2) (void)hipStreamDestroy(st2);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (work) (void)hipFree(work);
    if (work2) (void)hipFree(work2);
    if (tab) (void)hipFree(tab);
  }
};

chk_status ensure(void *&ptr, size_t &have, size_t need, const char *what, bool &moved) {
  if (have >= need) return CHK_OK;
  moved = true;
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  have = 0;
  if (hipMalloc(&ptr, need) != hipSuccess) {
    set_error(std::string("chk_render_batch: cannot allocate ") + what);
    return CHK_ENOMEM;
  }
  have = need;
  return CHK_OK;
}

// The shading launch of a batch chunk: MULTI (every slot names its request)
// or, for a chunk that holds one request, the single-request kernel.
template <int BS>
void launch_shade_batch(uint32_t grid, hipStream_t st, uint32_t depth, uint32_t segcap,
                        const Buffers &B, uint32_t in, uint32_t out, uint32_t qin, uint32_t qout,
                        const BatchIO &Q, const RenderArgs *single) {
  if (single)
    hipLaunchKernelGGL((k_shade<false, false, BS>), dim3(grid), dim3(BS), 0, st, *single, depth,
                       segcap, B.q_ray[in], B.q_thr[in], B.q_pid[in], B.hits, B.q_ray[out],
                       B.q_thr[out], B.q_pid[out], B.s_ray, B.s_L, B.path_L, B.ctrl, qin, qout,
                       LearnedIO{}, BatchIO{});
  else
    hipLaunchKernelGGL((k_shade<false, true, BS>), dim3(grid), dim3(BS), 0, st, RenderArgs{}, depth,
                       segcap, B.q_ray[in], B.q_thr[in], B.q_pid[in], B.hits, B.q_ray[out],
                       B.q_thr[out], B.q_pid[out], B.s_ray, B.s_L, B.path_L, B.ctrl, qin, qout,
                       LearnedIO{}, Q);
}

// The device's constant-rate clock after a chunk's last kernel (per-request
// latency inside one graph launch, where no host event can be recorded
// between chunks).
__global__ void k_stamp(unsigned long long *p) {
  if (threadIdx.x == 0) *p = wall_clock64();
}

}  // namespace

void batch_state_free(void *p) { delete static_cast<BatchState *>(p); }

chk_status batch_impl(chk_ctx_s *ctx, const chk_request *R, uint32_t n,
                      const chk_batch_tuning *tuning, double *latency_ms, chk_batch_stats *stats,
                      hipStream_t st) {
  chk_batch_tuning T{1u << 24, 128, 256, 0};
  if (tuning) {
    T = *tuning;
    if (T.chunk_paths == 0) T.chunk_paths = 1u << 24;
    if (T.trace_block == 0) T.trace_block = 128;
    if (T.shade_block == 0) T.shade_block = 256;
  }
  if (T.chunk_paths < (1u << 16) || T.chunk_paths > (1u << 26) ||
      (T.trace_block != 64 && T.trace_block != 128 && T.trace_block != 256) ||
      (T.shade_block != 128 && T.shade_block != 256 && T.shade_block != 512)) {
    set_error("chk_render_batch: tuning wants chunk_paths 2^16..2^26, trace_block 64/128/256, "
              "shade_block 128/256/512");
    return CHK_EINVAL;
  }
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (n == 0) return CHK_OK;
  // ---- requests
  std::vector<RenderArgs> args(n);
  uint32_t max_depth = 0;
  int stack = 32;
  for (uint32_t i = 0; i < n; ++i) {
    const chk_request &q = R[i];
    const std::string who = "chk_render_batch: request " + std::to_string(i);
    if (!q.scene || q.scene->ctx != ctx || !q.accum) {
      set_error(who + ": null scene or accum, or scene of another context");
      return CHK_EINVAL;
    }
    if (q.camera.width == 0 || q.camera.height == 0 || q.params.spp == 0) {
      set_error(who + ": zero width, height or spp");
      return CHK_EINVAL;
    }
    if ((uint64_t)q.camera.width * q.camera.height >= (1ull << 32)) {
      set_error(who + ": image too large");
      return CHK_ECAPACITY;
    }
    if (q.params.mis_mode != CHK_MIS_BALANCE) {
      set_error(who + ": batches render with CHK_MIS_BALANCE");
      return CHK_EINVAL;
    }
    if (q.params.spp > T.chunk_paths) {
      set_error(who + ": spp exceeds chunk_paths");
      return CHK_ECAPACITY;
    }
    const chk_status as = make_args(q.scene, &q.camera, &q.params, args[i], who.c_str());
    if (as != CHK_OK) return as;
    max_depth = std::max(max_depth, q.params.max_depth);
    stack = std::max(stack, q.scene->stack_size);
  }
  // ---- shortest request first, then pack chunks
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return (uint64_t)args[a].win_w * args[a].win_h * args[a].spp <
           (uint64_t)args[b].win_w * args[b].win_h * args[b].spp;
  });
  std::vector<BatchSeg> segs;
  std::vector<BatchChunk> chunks;
  std::vector<uint32_t> last_chunk(n, 0);
  uint64_t total_paths = 0;
  {
    BatchChunk cur{0, 0, 0};
    const uint64_t own = T.chunk_paths / 4;
    for (uint32_t i : order) {
      const uint32_t npix = args[i].win_w * args[i].win_h, spp = args[i].spp;
      const bool alone = (uint64_t)npix * spp >= own;
      if (alone && cur.n_seg) {  // close the shared chunk
        chunks.push_back(cur);
        cur = BatchChunk{(uint32_t)segs.size(), 0, 0};
      }
      uint32_t pix = 0;
      while (pix < npix) {
        const uint32_t room = (T.chunk_paths - cur.total) / spp;
        if (room == 0) {
          chunks.push_back(cur);
          cur = BatchChunk{(uint32_t)segs.size(), 0, 0};
          continue;
        }
        const uint32_t take = std::min(room, npix - pix);
        segs.push_back(BatchSeg{i, pix, take, cur.total});
        ++cur.n_seg;
        cur.total += take * spp;
        pix += take;
        total_paths += (uint64_t)take * spp;
        last_chunk[i] = (uint32_t)chunks.size();
      }
      if (alone) {  // nothing shares the request's last chunk
        chunks.push_back(cur);
        cur = BatchChunk{(uint32_t)segs.size(), 0, 0};
      }
    }
    if (cur.n_seg) chunks.push_back(cur);
  }
  // ---- state, buffers, tables
  BatchState *bs = static_cast<BatchState *>(ctx->batch_state);
  if (!bs) {
    bs = new BatchState();
    ctx->batch_state = bs;
  }
  const uint32_t cap = T.chunk_paths;
  const uint32_t segcap = seg_capacity(cap);
  const uint32_t qcap = kSegs * segcap;
  const size_t need = buffers_bytes(qcap, cap) + 3 * (size_t)qcap * 4 + 4 * 256;
  // graph key: the request records, the tuning, and per request its scene's
  // creation number and stack (a graph bakes in the traversal kernel chosen
  // for the stack; a scene handle's address alone can be reused)
  std::vector<unsigned char> key(sizeof(chk_request) * n + sizeof(T) + 12 * (size_t)n);
  std::memcpy(key.data(), R, sizeof(chk_request) * n);
  std::memcpy(key.data() + sizeof(chk_request) * n, &T, sizeof(T));
  for (uint32_t i = 0; i < n; ++i) {
    unsigned char *k = key.data() + sizeof(chk_request) * n + sizeof(T) + 12 * (size_t)i;
    const uint64_t g = R[i].scene->gen;
    const uint32_t sz = (uint32_t)R[i].scene->stack_size;
    std::memcpy(k, &g, 8);
    std::memcpy(k + 8, &sz, 4);
  }
  // captured graphs hold the buffer addresses: any reallocation drops them
  bool moved = false;
  chk_status cs = ensure(bs->work, bs->work_bytes, need, "the path queues", moved);
  if (cs != CHK_OK) return cs;
  // chunks alternate two queue sets and streams (as in chk_render)
  const bool dual = chunks.size() > 1;
  if (dual) {
    cs = ensure(bs->work2, bs->work2_bytes, need, "the second path-queue set", moved);
    if (cs != CHK_OK) return cs;
    if (!bs->st2) CHK_HIP(hipStreamCreateWithFlags(&bs->st2, hipStreamNonBlocking));
    if (!bs->ev_fork) CHK_HIP(hipEventCreateWithFlags(&bs->ev_fork, hipEventDisableTiming));
    if (!bs->ev_join) CHK_HIP(hipEventCreateWithFlags(&bs->ev_join, hipEventDisableTiming));
  }
  const size_t tab_need =
      segs.size() * (sizeof(RenderArgs) + sizeof(DeviceScene)) + (chunks.size() + 2) * 8 + 768;
  cs = ensure(bs->tab, bs->tab_bytes, tab_need, "the request tables", moved);
  if (cs != CHK_OK) return cs;
  if (moved) bs->drop_graphs();
  const bool cached = T.use_graph && bs->key == key && bs->exec != nullptr;
  if (T.use_graph && !cached) bs->drop_graphs();
  // queue set k: path queues, counters and the slots' segment indices
  struct QSet {
    Buffers B;
    uint32_t *q_seg[2];
    uint32_t *s_seg;
  } qs[2];
  for (int k = 0; k < (dual ? 2 : 1); ++k) {
    void *w = k ? bs->work2 : bs->work;
    qs[k].B = carve(w, qcap, cap);
    uint32_t *segbuf = (uint32_t *)((char *)w + ((buffers_bytes(qcap, cap) + 255) & ~(size_t)255));
    qs[k].q_seg[0] = segbuf;
    qs[k].q_seg[1] = segbuf + qcap;
    qs[k].s_seg = segbuf + 2 * (size_t)qcap;
  }
  if (dual) qs[1].B.stats = qs[0].B.stats;  // one set of statistics counters
  const Buffers &B0 = qs[0].B;
  std::vector<RenderArgs> seg_args(segs.size());
  std::vector<DeviceScene> seg_scene(segs.size());
  for (size_t k = 0; k < segs.size(); ++k) {
    RenderArgs a = args[segs[k].req];
    a.pix0 = segs[k].pix0;
    a.n_pix = segs[k].n_pix;
    a.p0 = segs[k].p0;
    seg_args[k] = a;
    seg_scene[k] = a.S;
  }
  RenderArgs *d_args = (RenderArgs *)bs->tab;
  DeviceScene *d_scene =
      (DeviceScene *)((char *)bs->tab + ((segs.size() * sizeof(RenderArgs) + 255) & ~(size_t)255));
  // time stamps: [0] at the start, [c + 1] after chunk c
  unsigned long long *d_stamp =
      (unsigned long long *)((char *)d_scene +
                             ((segs.size() * sizeof(DeviceScene) + 255) & ~(size_t)255));
  CHK_HIP(hipStreamSynchronize(st));
  CHK_HIP(hipMemcpy(d_args, seg_args.data(), segs.size() * sizeof(RenderArgs),
                    hipMemcpyHostToDevice));
  CHK_HIP(hipMemcpy(d_scene, seg_scene.data(), segs.size() * sizeof(DeviceScene),
                    hipMemcpyHostToDevice));
  auto qbase = [](uint32_t q) { return q == 0 ? (uint32_t)kQ0 : (uint32_t)kQ1; };
  // ---- the launches of one chunk (queue set c % 2 when dual)
  auto emit_chunk = [&](uint32_t c, hipStream_t s) -> hipError_t {
    const QSet &X = qs[dual ? (c & 1u) : 0u];
    const Buffers &B = X.B;
    uint32_t *const *q_seg = X.q_seg;
    uint32_t *s_seg = X.s_seg;
    const TraceShape shape{stack, (int)T.trace_block, B.ovf, qcap, 16};
    const BatchChunk &C = chunks[c];
    // a chunk of one request (a large request fills chunks alone) runs the
    // single-scene kernels: no per-slot scene and argument lookups
    const RenderArgs *single = C.n_seg == 1 ? &seg_args[C.first_seg] : nullptr;
    const chk_scene_s *ssc = single ? R[segs[C.first_seg].req].scene : nullptr;
    const TraceShape sshape{ssc ? ssc->stack_size : stack, (int)T.trace_block, B.ovf, qcap, 16};
    hipLaunchKernelGGL(k_reset, dim3((kCtrlWords + 255) / 256), dim3(256), 0, s, B.ctrl, B.stats,
                       0);
    hipLaunchKernelGGL(k_queue_counts, dim3(1), dim3(64), 0, s, B.ctrl, C.total);
    for (uint32_t k = 0; k < C.n_seg; ++k) {
      const RenderArgs &a = seg_args[C.first_seg + k];
      const uint32_t np = a.n_pix * a.spp;
      hipLaunchKernelGGL(k_camera, dim3((np + 255) / 256), dim3(256), 0, s, a, segcap, C.total,
                         B.q_ray[0], B.q_thr[0], B.q_pid[0], B.path_L, single ? nullptr : q_seg[0],
                         k);
    }
    for (uint32_t depth = 0; depth <= max_depth; ++depth) {
      const uint32_t in = depth & 1u, out = in ^ 1u;
      hipLaunchKernelGGL(k_reset_bounce, dim3(1), dim3(64), 0, s, B.ctrl, qbase(out));
      // ray replacement after the camera rays and for shadow rays (as chk_render)
      TraceShape cshape = single ? sshape : shape, ashape = single ? sshape : shape;
      cshape.stream = depth ? kStreamBlocksClosest : 0u;
      ashape.stream = kStreamBlocksAny;
      hipError_t e =
          single ? launch_closest(ssc, cshape, B.q_ray[in], SegRange{B.ctrl + qbase(in), 0u, segcap},
                                  B.hits, true, false, B.stats + kNodeVisits, SlotScenes{}, s)
                 : launch_closest(nullptr, cshape, B.q_ray[in],
                                  SegRange{B.ctrl + qbase(in), 0u, segcap}, B.hits, true, false,
                                  B.stats + kNodeVisits,
                                  SlotScenes{d_scene + C.first_seg, q_seg[in]}, s);
      if (e != hipSuccess) return e;
      const BatchIO Q{d_args + C.first_seg, q_seg[in], q_seg[out], s_seg};
      const uint32_t sgrid = kSegs * (segcap / T.shade_block);
      if (T.shade_block == 128)
        launch_shade_batch<128>(sgrid, s, depth, segcap, B, in, out, qbase(in), qbase(out), Q,
                                single);
      else if (T.shade_block == 512)
        launch_shade_batch<512>(sgrid, s, depth, segcap, B, in, out, qbase(in), qbase(out), Q,
                                single);
      else
        launch_shade_batch<256>(sgrid, s, depth, segcap, B, in, out, qbase(in), qbase(out), Q,
                                single);
      if (depth < max_depth) {
        e = single ? launch_any(ssc, ashape, B.s_ray, SegRange{B.ctrl + kQS, 0u, segcap}, nullptr,
                                B.s_L, B.path_L, true, false, B.stats + kShNodeVisits,
                                SlotScenes{}, s)
                   : launch_any(nullptr, ashape, B.s_ray, SegRange{B.ctrl + kQS, 0u, segcap},
                                nullptr, B.s_L, B.path_L, true, false, B.stats + kShNodeVisits,
                                SlotScenes{d_scene + C.first_seg, s_seg}, s);
        if (e != hipSuccess) return e;
      }
    }
    for (uint32_t k = 0; k < C.n_seg; ++k) {
      const RenderArgs &a = seg_args[C.first_seg + k];
      const chk_request &q = R[segs[C.first_seg + k].req];
      hipLaunchKernelGGL(k_accumulate, dim3(accumulate_blocks(a.n_pix, a.spp)), dim3(256), 0, s, a,
                         B.path_L,
                         q.accum, (q.params.flags & CHK_RENDER_SUMSQ) ? q.sumsq : nullptr);
    }
    return hipGetLastError();
  };
  // ---- the whole batch on streams a (chunks 0, 2, ...) and b (1, 3, ...):
  // statistics reset (outside the chunks, which clear only their queues),
  // fork, chunks with a time stamp after each, join
  auto emit_all = [&](hipStream_t a, hipStream_t b, hipEvent_t fork,
                      hipEvent_t join) -> hipError_t {
    hipLaunchKernelGGL(k_reset, dim3((kCtrlWords + 255) / 256), dim3(256), 0, a, B0.ctrl,
                       B0.stats, 1);
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, a, d_stamp);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && dual) {
      e = hipEventRecord(fork, a);
      if (e == hipSuccess) e = hipStreamWaitEvent(b, fork, 0);
    }
    for (uint32_t c = 0; c < chunks.size() && e == hipSuccess; ++c) {
      const hipStream_t s = dual && (c & 1u) ? b : a;
      e = emit_chunk(c, s);
      if (e == hipSuccess) {
        hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, d_stamp + c + 1);
        e = hipGetLastError();
      }
    }
    if (e == hipSuccess && dual) {
      e = hipEventRecord(join, b);
      if (e == hipSuccess) e = hipStreamWaitEvent(a, join, 0);
    }
    return e;
  };
  if (!bs->ev_start) CHK_HIP(hipEventCreate(&bs->ev_start));
  if (!bs->ev_end) CHK_HIP(hipEventCreate(&bs->ev_end));
  // ---- capture (graph mode, not cached): one graph for the whole batch
  double capture_ms = 0.0;
  if (T.use_graph && !cached) {
    if (!bs->cap) CHK_HIP(hipStreamCreateWithFlags(&bs->cap, hipStreamNonBlocking));
    if (!bs->cap2) CHK_HIP(hipStreamCreateWithFlags(&bs->cap2, hipStreamNonBlocking));
    if (!bs->cev_fork) CHK_HIP(hipEventCreateWithFlags(&bs->cev_fork, hipEventDisableTiming));
    if (!bs->cev_join) CHK_HIP(hipEventCreateWithFlags(&bs->cev_join, hipEventDisableTiming));
    const auto t0 = std::chrono::steady_clock::now();
    CHK_HIP(hipStreamBeginCapture(bs->cap, hipStreamCaptureModeThreadLocal));
    const hipError_t e = emit_all(bs->cap, bs->cap2, bs->cev_fork, bs->cev_join);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(bs->cap, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      bs->drop_graphs();
      set_error(std::string("chk_render_batch: graph capture: ") +
                hipGetErrorString(e != hipSuccess ? e : e2));
      return CHK_EDEVICE;
    }
    hipGraphExec_t x = nullptr;
    const hipError_t e3 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e3 != hipSuccess) {
      bs->drop_graphs();
      set_error(std::string("chk_render_batch: graph instantiate: ") + hipGetErrorString(e3));
      return CHK_EDEVICE;
    }
    bs->exec = x;
    bs->key = key;
    capture_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  // ---- run: one graph launch, or the same launches eagerly
  hipError_t err = hipEventRecord(bs->ev_start, st);
  if (err == hipSuccess)
    err = T.use_graph ? hipGraphLaunch(bs->exec, st)
                      : emit_all(st, bs->st2, bs->ev_fork, bs->ev_join);
  if (err == hipSuccess) err = hipEventRecord(bs->ev_end, st);
  if (err == hipSuccess) err = hipEventSynchronize(bs->ev_end);
  if (err != hipSuccess) {
    set_error(std::string("chk_render_batch: ") + hipGetErrorString(err));
    return CHK_EDEVICE;
  }
  if (latency_ms) {
    std::vector<unsigned long long> stamp(chunks.size() + 1);
    err = hipMemcpy(stamp.data(), d_stamp, stamp.size() * 8, hipMemcpyDeviceToHost);
    int khz = 0;
    if (err == hipSuccess) err = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate,
                                                       ctx->device);
    if (err != hipSuccess || khz <= 0) {
      set_error(std::string("chk_render_batch: time stamps: ") +
                (err != hipSuccess ? hipGetErrorString(err) : "no wall clock rate"));
      return CHK_EDEVICE;
    }
    for (uint32_t i = 0; i < n; ++i)
      latency_ms[i] = (double)(stamp[last_chunk[i] + 1] - stamp[0]) / (double)khz;
  }
  unsigned long long h[kStatWords];
  err = hipMemcpy(h, B0.stats, sizeof(h), hipMemcpyDeviceToHost);
  if (err != hipSuccess) {
    set_error(std::string("chk_render_batch: statistics readback: ") + hipGetErrorString(err));
    return CHK_EDEVICE;
  }
  if (h[kOverflow] || h[kShOverflow]) {
    set_error("chk_render_batch: traversal stack overflow");
    return CHK_EINTERNAL;
  }
  if (stats) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, bs->ev_start, bs->ev_end);
    stats->paths = total_paths;
    stats->rays_extend = h[kRaysExt];
    stats->rays_shadow = h[kRaysShadow];
    stats->chunks = (uint32_t)chunks.size();
    stats->segments = (uint32_t)segs.size();
    stats->batch_ms = ms;
    stats->capture_ms = capture_ms;
    stats->graph_cached = cached ? 1u : 0u;
  }
  return CHK_OK;
}

// Autotuner: time every (chunk_paths, trace_block) candidate on the batch
// itself (the chunk size that suits a request mix depends on the mix at
// its real spp), rendering into scratch accumulators, then eager against
// graph replay for the fastest.  The first call of each chunk size grows
// the buffers; each candidate is timed on its second call.  The graph is
// kept unless it is more than 1 % slower than eager launches: a call is
// then one hipGraphLaunch, and the host thread issues nothing else.
chk_status batch_autotune(chk_ctx_s *ctx, const chk_request *R, uint32_t n,
                          chk_batch_tuning *best, double *best_ms, hipStream_t st) {
  if (n == 0) {
    *best = chk_batch_tuning{1u << 24, 128, 256, 0};
    return CHK_OK;
  }
  size_t maxpix = 0;
  for (uint32_t i = 0; i < n; ++i)
    maxpix = std::max<size_t>(maxpix, (size_t)R[i].camera.width * R[i].camera.height);
  double *scratch = nullptr;
  if (hipMalloc(&scratch, maxpix * 3 * sizeof(double) * 2) != hipSuccess) {
    set_error("chk_batch_autotune: cannot allocate scratch accumulators");
    return CHK_ENOMEM;
  }
  std::vector<chk_request> Q(R, R + n);
  for (auto &q : Q) {
    // every request adds into the same scratch: the values are discarded
    q.accum = scratch;
    q.sumsq = scratch + maxpix * 3;
  }
  const uint32_t chunk_c[] = {1u << 22, 1u << 23, 1u << 24, 1u << 25};
  const uint32_t tb_c[] = {64, 128};
  double bms = 1e30;
  chk_batch_tuning bt{1u << 24, 128, 256, 0};
  chk_status s = CHK_OK;
  for (uint32_t c : chunk_c)
    for (uint32_t tb : tb_c) {
      chk_batch_tuning t{c, tb, 256, 0};
      chk_batch_stats bst;
      for (int rep = 0; rep < (tb == tb_c[0] ? 2 : 1) && s == CHK_OK; ++rep)
        s = batch_impl(ctx, Q.data(), n, &t, nullptr, &bst, st);
      if (s != CHK_OK) break;
      if (bst.batch_ms < bms) {
        bms = bst.batch_ms;
        bt = t;
      }
    }
  if (s == CHK_OK) {
    chk_batch_tuning g = bt;
    g.use_graph = 1;
    chk_batch_stats bst;
    for (int rep = 0; rep < 3 && s == CHK_OK; ++rep) s = batch_impl(ctx, Q.data(), n, &g, nullptr, &bst, st);
    if (s == CHK_OK && bst.batch_ms <= 1.01 * bms) {
      bms = bst.batch_ms;
      bt = g;
    }
    BatchState *bs = static_cast<BatchState *>(ctx->batch_state);
    if (bs) bs->drop_graphs();  // they render into the scratch accumulators
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(scratch);
  if (s != CHK_OK) return s;
  *best = bt;
  if (best_ms) *best_ms = bms;
  return CHK_OK;
}

}  // namespace chk

// ====================================================================== no-allocation render
// chk_render for callers that cannot allocate device memory (code under HIP
// graph capture, serving loops, harnesses that forbid hipMalloc): the path
// queues live in one fixed array of device memory that the library holds
// for the life of the process, one copy per device.  The host-side name
// `scratch` is a stand-in; every use goes through the device address that
// hipGetSymbolAddress returned for the current device.

static __device__ __align__(256) unsigned char scratch[CHK_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
bool g_scratch_done[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
void *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

namespace chk {

hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_done[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
    g_scratch_done[device] = true;
  }
  return g_scratch_err[device];
}

}  // namespace chk

extern "C" chk_status chk_render_noalloc(chk_scene_t sc, const chk_camera *camera,
                                         const chk_render_params *params, double *accum,
                                         double *sumsq, chk_render_stats *stats, void *stream) {
  using namespace chk;
  CHK_TRY_BEGIN
  if (!sc) {
    set_error("chk_render_noalloc: null scene");
    return CHK_EINVAL;
  }
  if (sc->extras) {
    set_error("chk_render_noalloc: the scene has spheres or triangle, sphere, point or "
              "environment sources; only the direct-lighting entry points and "
              "chk_render_shade render those");
    return CHK_EINVAL;
  }
  if (params && (params->flags & CHK_RENDER_PROFILE)) {
    set_error("chk_render_noalloc: CHK_RENDER_PROFILE creates events; not available here");
    return CHK_EINVAL;
  }
  const int dev = sc->ctx->device;
  CHK_HIP(hipSetDevice(dev));
  void *ptr = nullptr;
  {
    const hipError_t e = noalloc_scratch_init(dev);
    if (e != hipSuccess) {
      set_error(std::string("chk_render_noalloc: no scratch array on this device: ") +
                hipGetErrorString(e));
      return CHK_EDEVICE;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    ptr = g_scratch_ptr[dev];
  }
  const Workspace ws{ptr, (size_t)CHK_NOALLOC_SCRATCH_BYTES};
  return render_impl(sc, camera, params, accum, sumsq, stats, false, nullptr, nullptr,
                     (hipStream_t)stream, &ws);
  CHK_TRY_END
}
