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
//   k_accumulate: per-pixel FP64 sums in sample order (main.cpp:62,83)
// Every path owns one FP32 radiance slot; queue order never changes a
// path's result (its randomness is keyed by pixel, sample and depth).
#include <chrono>
#include <cmath>
#include <cstring>

#include "chk_device.h"
#include "chk_internal.h"

namespace chk {
namespace {

constexpr uint32_t kChunkPaths = 1u << 22;  // paths in flight per chunk
constexpr int kShadeBlock = 256;

// control words in the work buffer: per-segment queue counters, each
// segment's counter kSegStride words (256 bytes) after the previous one
enum Ctrl : uint32_t {
  kQ0 = 0,                        // live-ray counts, queue 0
  kQ1 = kSegs * kSegStride,       // live-ray counts, queue 1
  kQS = 2 * kSegs * kSegStride,   // shadow-ray counts
  kCtrlWords = 3 * kSegs * kSegStride
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

struct RenderArgs {
  DeviceScene S;
  CamArgs cam;
  uint32_t spp, sample_offset, max_depth, jitter_mode, flags, key0, key1, req;
  uint32_t pix0, n_pix;
  float inv_nl;
};

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
};

// qcap: queue slots (kSegs * segcap); cap: paths per chunk
size_t buffers_bytes(uint32_t qcap, uint32_t cap) {
  size_t b = (size_t)kCtrlWords * 4 + 256 + 256;
  b += 2 * (size_t)qcap * (32 + 16 + 4);
  b += (size_t)qcap * (16 + 32 + 16);
  b += (size_t)cap * 16;
  return b + 16 * 256;
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
  return B;
}

__device__ inline uint32_t lane_id() { return __lane_id(); }

// Block-wide compacting append into two queues: every thread of the block
// calls it once; threads with pa (pb) get consecutive slots after the
// block's base in queue a (b).  One atomic per queue per block.
template <int NW>
__device__ inline void block_append2(bool pa, bool pb, uint32_t *ca, uint32_t *cb, uint32_t &sa,
                                     uint32_t &sb) {
  __shared__ uint32_t cnt[2][NW];
  __shared__ uint32_t base[2];
  const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
  const unsigned long long ma = __ballot(pa), mb = __ballot(pb);
  if (lane == 0) {
    cnt[0][w] = (uint32_t)__popcll(ma);
    cnt[1][w] = (uint32_t)__popcll(mb);
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    uint32_t tot = 0;
    for (int k = 0; k < NW; ++k) tot += cnt[threadIdx.x][k];
    base[threadIdx.x] = tot ? atomicAdd(threadIdx.x ? cb : ca, tot) : 0u;
  }
  __syncthreads();
  uint32_t oa = base[0], ob = base[1];
  for (uint32_t k = 0; k < w; ++k) {
    oa += cnt[0][k];
    ob += cnt[1][k];
  }
  const unsigned long long below = (1ull << lane) - 1ull;
  sa = oa + (uint32_t)__popcll(ma & below);
  sb = ob + (uint32_t)__popcll(mb & below);
}

__global__ void k_reset(uint32_t *ctrl, unsigned long long *stats, int clear_stats) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kCtrlWords) ctrl[i] = 0u;
  if (clear_stats && i < kStatWords) stats[i] = 0ull;
}

// zero the output queue's and the shadow queue's segment counters
__global__ void k_reset_bounce(uint32_t *ctrl, uint32_t out_q) {
  const uint32_t t = threadIdx.x;
  if (t < kSegs) ctrl[out_q + t * kSegStride] = 0u;
  else if (t < 2 * kSegs) ctrl[kQS + (t - kSegs) * kSegStride] = 0u;
}

// Camera rays: main.cpp:72-79 with prepareScene.cpp:24-33's cx, cy.
// Path p of the chunk goes to segment p / per, per = ceil(n / kSegs).
__global__ void k_camera(RenderArgs A, uint32_t segcap, float4 *__restrict__ q_ray,
                         float4 *__restrict__ q_thr, uint32_t *__restrict__ q_pid,
                         float4 *__restrict__ path_L, uint32_t *__restrict__ ctrl) {
  const uint32_t n = A.n_pix * A.spp;
  const uint32_t per = (n + kSegs - 1) / kSegs;
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kSegs) {
    const uint32_t lo = p * per;
    ctrl[kQ0 + p * kSegStride] = n > lo ? min(per, n - lo) : 0u;
  }
  if (p >= n) return;
  const uint32_t r = p / per;
  const size_t slot = (size_t)r * segcap + (p - r * per);
  const uint32_t idx = A.pix0 + p / A.spp;
  const uint32_t s = A.sample_offset + p % A.spp;
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

// One vertex of every live path.  Block b shades items of segment
// r = b % kSegs and appends to segment r of the output queues.
__global__ __launch_bounds__(kShadeBlock) void k_shade(
    RenderArgs A, uint32_t depth, uint32_t segcap, const float4 *__restrict__ in_ray,
    const float4 *__restrict__ in_thr, const uint32_t *__restrict__ in_pid,
    const float4 *__restrict__ hits, float4 *__restrict__ out_ray, float4 *__restrict__ out_thr,
    uint32_t *__restrict__ out_pid, float4 *__restrict__ s_ray, float4 *__restrict__ s_L,
    float4 *__restrict__ path_L, uint32_t *__restrict__ ctrl, uint32_t in_q, uint32_t out_q) {
  const DeviceScene &S = A.S;
  const uint32_t r = blockIdx.x % kSegs, qb = blockIdx.x / kSegs;
  const uint32_t total = ctrl[in_q + r * kSegStride];
  if (qb * kShadeBlock >= total) return;  // whole block idle (uniform)
  const uint32_t off = qb * kShadeBlock + threadIdx.x;
  const size_t i = (size_t)r * segcap + off;
  {
    bool emit_ray = false, emit_sh = false;
    float4 nr0, nr1, nthr, sr0, sr1, sl;
    uint32_t pid = 0;
    if (off < total) {
      const float4 h = hits[i];
      const uint32_t leaf = __float_as_uint(h.y);
      pid = in_pid[i];
      if (leaf != 0xFFFFFFFFu) {
        const float4 r1 = in_ray[2 * (size_t)i + 1];
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
              const float pl = A.inv_nl * h.x * h.x / (S.lights[info.y].area * cl);
              const float pb = thr4.w;
              w = pb / (pb + pl);  // balance heuristic, N_light = N_bsdf = 1 (mis.cpp:209-220)
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
          const uint32_t pix = A.pix0 + pid / A.spp;
          const uint32_t smp = A.sample_offset + pid % A.spp;
          // ---- next-event estimation: one point on one mesh light
          if (S.n_lights) {
            const u4 r = philox(u4{pix, smp, (depth << 2) | CHK_STREAM_NEE, A.req}, A.key0, A.key1);
            const uint32_t j = min((uint32_t)(u01(r.x) * (float)S.n_lights), S.n_lights - 1u);
            const LightRec L = S.lights[j];
            const float u3 = u01(r.w);
            uint32_t k = L.first;
            while (k + 1 < L.first + L.count && u3 >= S.light_tris[k].a_cdf.w) ++k;
            const LightTri &lt = S.light_tris[k];
            const f3 la = xyz(lt.a_cdf), lb = xyz(lt.b_area), lc = xyz(lt.c_leaf);
            // MeshLight::getSamples point (lightSources.cpp:812-824)
            const float e1 = sqrtf(u01(r.y)), e2 = u01(r.z);
            const f3 yp = la * (1.0f - e1) + lb * (e1 * (1.0f - e2)) + lc * (e2 * e1);
            const f3 dv = yp - X;
            const float d2 = dot(dv, dv);
            const float dist = sqrtf(d2);
            const f3 wi = dv * (1.0f / dist);
            const float cx = dot(wi, n);
            const f3 nl = cross(lc - la, lb - la);
            const float cl = fabsf(dot(wi, nl)) * rsqrtf(dot(nl, nl));
            if (cx > 0.0f && cl > 0.0f) {
              float f, pb;
              brdf_pdf(m, n, wo_ref, wo, wi, f, pb);
              const float pl = A.inv_nl * d2 / (L.area * cl);  // mis.cpp:55-66
              const f3 c = thr * m.c * xyz(lt.rad) * (f * cx / (pl + pb));
              if (finite3(c) && (c.x > 0.0f || c.y > 0.0f || c.z > 0.0f)) {
                // aim from the offset origin at the sampled point, so the ray
                // reaches the light exactly at t = sdist (shader.cpp:82-90)
                const f3 sd = yp - org;
                const float sdist = sqrtf(dot(sd, sd));
                const f3 sdir = sd * (1.0f / sdist);
                emit_sh = true;
                sr0 = make_float4(org.x, org.y, org.z, sdist - 1e-5f);
                sr1 = make_float4(sdir.x, sdir.y, sdir.z, __uint_as_float(leaf));
                sl = make_float4(c.x, c.y, c.z, __uint_as_float(pid));
              }
            }
          }
          // ---- BSDF sample continues the path
          {
            const u4 r = philox(u4{pix, smp, (depth << 2) | CHK_STREAM_BSDF, A.req}, A.key0, A.key1);
            const f3 wi = brdf_sample(m, n, wo_ref, wo, u01(r.x), u01(r.y), u01(r.z));
            const float cs = dot(wi, n);
            if (cs > 0.0f) {
              float f, pb;
              brdf_pdf(m, n, wo_ref, wo, wi, f, pb);
              if (pb > 0.0f && isfinite(pb) && isfinite(f)) {
                const f3 t2 = thr * m.c * (f * cs / pb);
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
    uint32_t slot, sslot;
    block_append2<kShadeBlock / 64>(emit_ray, emit_sh, ctrl + out_q + r * kSegStride,
                                    ctrl + kQS + r * kSegStride, slot, sslot);
    const size_t seg0 = (size_t)r * segcap;
    if (emit_ray) {
      const size_t o = seg0 + slot;
      out_ray[2 * o] = nr0;
      out_ray[2 * o + 1] = nr1;
      out_thr[o] = nthr;
      out_pid[o] = pid;
    }
    if (emit_sh) {
      const size_t o = seg0 + sslot;
      s_ray[2 * o] = sr0;
      s_ray[2 * o + 1] = sr1;
      s_L[o] = sl;
    }
  }
}

// Per-pixel FP64 accumulation in sample order (main.cpp:62,83).
__global__ void k_accumulate(RenderArgs A, const float4 *__restrict__ path_L,
                             double *__restrict__ accum, double *__restrict__ sumsq) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= A.n_pix) return;
  double s0 = 0, s1 = 0, s2 = 0, q0 = 0, q1 = 0, q2 = 0;
  const float4 *p = path_L + (size_t)q * A.spp;
  for (uint32_t s = 0; s < A.spp; ++s) {
    const float4 L = p[s];
    const double a = L.x, b = L.y, c = L.z;
    s0 += a;
    s1 += b;
    s2 += c;
    q0 += a * a;
    q1 += b * b;
    q2 += c * c;
  }
  const size_t o = 3 * ((size_t)A.pix0 + q);
  accum[o] += s0;
  accum[o + 1] += s1;
  accum[o + 2] += s2;
  if (sumsq) {
    sumsq[o] += q0;
    sumsq[o + 1] += q1;
    sumsq[o + 2] += q2;
  }
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

}  // namespace

chk_status render_impl(chk_scene_s *sc, const chk_camera *cam, const chk_render_params *P,
                       double *accum, double *sumsq, chk_render_stats *stats, bool count,
                       hipStream_t st) {
  if (!cam || !P || !accum) {
    set_error("chk_render: null camera, params or accum");
    return CHK_EINVAL;
  }
  if (cam->width == 0 || cam->height == 0 || P->spp == 0) {
    set_error("chk_render: zero width, height or spp");
    return CHK_EINVAL;
  }
  if (P->spp > kChunkPaths) {
    set_error("chk_render: spp exceeds the paths-per-chunk capacity");
    return CHK_ECAPACITY;
  }
  if (P->mis_mode != CHK_MIS_BALANCE) {
    set_error("chk_render: mis_mode other than CHK_MIS_BALANCE is not available in this entry point");
    return CHK_EINVAL;
  }
  const uint64_t n_pix_total = (uint64_t)cam->width * cam->height;
  if (n_pix_total >= (1ull << 32)) {
    set_error("chk_render: image too large");
    return CHK_ECAPACITY;
  }
  const uint32_t pix_per_chunk = std::max<uint32_t>(1u, kChunkPaths / P->spp);
  const uint32_t cap = (uint32_t)std::min<uint64_t>((uint64_t)pix_per_chunk * P->spp,
                                                   n_pix_total * P->spp);
  const uint32_t segcap = seg_capacity(cap);
  const uint32_t qcap = kSegs * segcap;
  const size_t need = buffers_bytes(qcap, cap);
  if (sc->work_bytes < need) {
    if (sc->work) (void)hipFree(sc->work);
    sc->work = nullptr;
    sc->work_bytes = 0;
    if (hipMalloc(&sc->work, need) != hipSuccess) {
      set_error("chk_render: cannot allocate the path queues");
      return CHK_ENOMEM;
    }
    sc->work_bytes = need;
  }
  Buffers B = carve(sc->work, qcap, cap);

  RenderArgs A;
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
      set_error("chk_render: camera up and direction are parallel");
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

  Profiler prof;
  prof.on = (P->flags & CHK_RENDER_PROFILE) != 0;
  prof.st = st;
  const uint32_t sgrid = kSegs * (segcap / kShadeBlock);
  uint32_t chunks = 0, ext_launches = 0;
  auto qbase = [](uint32_t q) { return q == 0 ? (uint32_t)kQ0 : (uint32_t)kQ1; };

  hipLaunchKernelGGL(k_reset, dim3((kCtrlWords + 255) / 256), dim3(256), 0, st, B.ctrl, B.stats,
                     1);
  for (uint64_t pix0 = 0; pix0 < n_pix_total; pix0 += pix_per_chunk) {
    ++chunks;
    A.pix0 = (uint32_t)pix0;
    A.n_pix = (uint32_t)std::min<uint64_t>(pix_per_chunk, n_pix_total - pix0);
    const uint32_t n_paths = A.n_pix * A.spp;
    prof.begin(3);
    hipLaunchKernelGGL(k_reset, dim3((kCtrlWords + 255) / 256), dim3(256), 0, st, B.ctrl, B.stats,
                       0);
    hipLaunchKernelGGL(k_camera, dim3((n_paths + 255) / 256), dim3(256), 0, st, A, segcap,
                       B.q_ray[0], B.q_thr[0], B.q_pid[0], B.path_L, B.ctrl);
    prof.end();
    for (uint32_t depth = 0; depth <= P->max_depth; ++depth) {
      const uint32_t in = depth & 1u, out = in ^ 1u;
      hipLaunchKernelGGL(k_reset_bounce, dim3(1), dim3(64), 0, st, B.ctrl, qbase(out));
      prof.begin(0);
      hipError_t e = launch_closest(sc, B.q_ray[in], SegRange{B.ctrl + qbase(in), 0u, segcap},
                                    B.hits, true, count, B.stats + kNodeVisits, st);
      prof.end();
      ++ext_launches;
      if (e != hipSuccess) {
        set_error(std::string("chk_render: closest-hit launch: ") + hipGetErrorString(e));
        return CHK_EDEVICE;
      }
      prof.begin(1);
      hipLaunchKernelGGL(k_shade, dim3(sgrid), dim3(kShadeBlock), 0, st, A, depth, segcap,
                         B.q_ray[in], B.q_thr[in], B.q_pid[in], B.hits, B.q_ray[out],
                         B.q_thr[out], B.q_pid[out], B.s_ray, B.s_L, B.path_L, B.ctrl, qbase(in),
                         qbase(out));
      prof.end();
      if (depth < P->max_depth) {
        prof.begin(2);
        e = launch_any(sc, B.s_ray, SegRange{B.ctrl + kQS, 0u, segcap}, nullptr, B.s_L, B.path_L,
                       true, count, B.stats + kShNodeVisits, st);
        prof.end();
        if (e != hipSuccess) {
          set_error(std::string("chk_render: any-hit launch: ") + hipGetErrorString(e));
          return CHK_EDEVICE;
        }
      }
    }
    prof.begin(3);
    hipLaunchKernelGGL(k_accumulate, dim3((A.n_pix + 255) / 256), dim3(256), 0, st, A, B.path_L,
                       accum, (P->flags & CHK_RENDER_SUMSQ) ? sumsq : nullptr);
    prof.end();
  }
  CHK_HIP(hipGetLastError());
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    unsigned long long h[kStatWords];
    CHK_HIP(hipMemcpyAsync(h, B.stats, sizeof(h), hipMemcpyDeviceToHost, st));
    CHK_HIP(hipStreamSynchronize(st));
    stats->paths = n_pix_total * P->spp;
    stats->rays_extend = h[kRaysExt];
    stats->rays_shadow = h[kRaysShadow];
    stats->node_visits = h[kNodeVisits];
    stats->tri_tests = h[kTriTests];
    stats->shadow_node_visits = h[kShNodeVisits];
    stats->shadow_tri_tests = h[kShTriTests];
    stats->chunks = chunks;
    stats->extend_launches = ext_launches;
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

}  // namespace chk
