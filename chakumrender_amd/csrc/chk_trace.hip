// Closest-hit and any-hit BVH traversal for gfx950.
//
// Restates BvhAccel::intersectT (bvhAccel.cpp:200-249) and shadow()
// (shader.cpp:76-97) over the 4-wide tree of chk_bvh.cpp (64-byte nodes with
// 8-bit quantized child boxes):
//  * one ray per lane; the rays of a launch are split into kSegs segments
//    and block b works on segment b % kSegs, so with round-robin dispatch
//    each XCD (and its L2) serves one segment; no global atomics;
//  * one ray per lane, the current node in VGPRs, pending nodes on a
//    per-lane stack in LDS laid out [slot][lane] (bank-conflict free);
//  * child boxes are tested conservatively, after Ize's robust slab test
//    ("Robust BVH Ray Traversal", JCGT 2(2) 2013; see node_hits_q),
//    children ordered by entry distance;
//  * triangles use the watertight test (chk_device.h) with the reference's
//    accept rules.
#include "chk_device.h"
#include "chk_internal.h"
#include "chk_node_test.h"

namespace chk {
namespace {

// Traversal with postponed leaves (Aila and Laine, "Understanding the
// Efficiency of Ray Traversal on GPUs", HPG 2009): a lane that reaches a leaf
// keeps it aside and goes on down inner nodes while other lanes of its wave
// still descend; the wave leaves the inner-node loop once nearly every lane
// holds a leaf (or has no node left) and then tests its leaves together.  Inner
// nodes and triangle tests thus each run with more of the wave's lanes
// active.  The leaves a ray tests, and so its closest hit, are those of the
// one-node-per-step order; only the order among them can change (the first
// of two exactly equally near triangles found wins, as in
// BvhAccel::intersectT's strict test).
template <bool ANY, bool COUNT, int STACK, int BS>
__device__ inline void traverse(const DeviceScene &S, uint32_t *__restrict__ stk, f3 o, f3 d,
                                float tmax, uint32_t skip, bool leaf_ids, float &t_out,
                                uint32_t &id_out, float &u_out, float &v_out, bool &hit_any,
                                uint32_t &leaf_out, uint32_t &nv, uint32_t &nt,
                                uint32_t &overflow) {
  const RayPre pre = ray_pre(d);
  const f3 inv = mk(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
  float t = tmax;
  uint32_t best = 0xFFFFFFFFu, best_leaf = 0u;
  float bu = 0.0f, bv = 0.0f;
  uint32_t entry = 0;  // root; valid while live
  bool live = true;
  uint32_t pend = 0;   // postponed leaf (leaves carry kLeafBit, so 0 is none)
  int sp = 0;
  hit_any = false;
  auto pop = [&]() {
    if (sp == 0) {
      live = false;
    } else {
      --sp;
      entry = stk[sp * BS];
    }
  };
  while (true) {
    // ---- inner nodes
    while (true) {
      // leave once fewer than 4 lanes still search without a leaf (waiting
      // for the last one measured 3.6 % slower; 2 to 6 lanes within 1 %), if
      // some lane holds one
      const bool searching = live && !(entry & kLeafBit);
      if (!__any(searching) ||
          (__popcll(__ballot(searching && pend == 0)) < 4 && __any(pend != 0)))
        break;  // wave-uniform
      if (!searching) continue;
      if (COUNT) ++nv;
      float key[4];
      uint32_t cid[4];
      const int nh = node_hits_q(S.nodes, entry, o, inv, t, key, cid);
      // sort (key, cid) ascending: 5 compare-exchanges
#define CHK_CX(a, b)                                      \
  {                                                       \
    bool sw = key[b] < key[a];                            \
    float tk = sw ? key[b] : key[a];                      \
    key[b] = sw ? key[a] : key[b];                        \
    key[a] = tk;                                          \
    uint32_t tc = sw ? cid[b] : cid[a];                   \
    cid[b] = sw ? cid[a] : cid[b];                        \
    cid[a] = tc;                                          \
  }
      CHK_CX(0, 1) CHK_CX(2, 3) CHK_CX(0, 2) CHK_CX(1, 3) CHK_CX(1, 2)
#undef CHK_CX
      if (nh > 0) {
        entry = cid[0];
        // push the other nh - 1 far to near, so the nearer ones pop first:
        // cid[1 + k] to slot sp + nh - 2 - k, without a loop (entries past
        // the stack are dropped and the ray marked, as one push at a time)
        const int np = nh - 1;
        if (sp + np > STACK) overflow = 1;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int pos = sp + np - 1 - k;
          if (k < np && pos < STACK) stk[pos * BS] = cid[1 + k];
        }
        sp = min(sp + np, STACK);
      } else {
        pop();
      }
      if (live && (entry & kLeafBit) && pend == 0) {  // postpone the first leaf
        pend = entry;
        pop();
      }
    }
    // ---- leaves: the postponed one, then any that the lane holds next
    while (true) {
      uint32_t lf = pend;
      pend = 0;
      if (lf == 0 && live && (entry & kLeafBit)) {
        lf = entry;
        pop();
      }
      if (!__any(lf != 0)) break;  // wave-uniform
      if (lf == 0) continue;
      const uint32_t first = leaf_first(lf), cnt = leaf_count(lf);
      for (uint32_t k = 0; k < cnt; ++k) {
        const uint32_t li = first + k;
        const float4 *tp = S.tris + 3 * (size_t)li;
        const float4 a = tp[0], b = tp[1], c = tp[2];
        const uint32_t tid = leaf_ids ? li : __float_as_uint(a.w);
        if (COUNT) ++nt;
        if (tid == skip) continue;
        float tt = t, uu, vv;
        if (tri_hit(pre, o, d, xyz(a), xyz(b), xyz(c), tt, uu, vv)) {
          t = tt;
          best = tid;
          best_leaf = li;
          bu = uu;
          bv = vv;
          hit_any = true;
          if (ANY) break;
        }
      }
      if (ANY && hit_any) live = false;  // occluded: nothing more to test
    }
    if (!__any(live)) break;  // wave-uniform
    if (!live) continue;
  }
  t_out = best == 0xFFFFFFFFu ? 1e30f : t;
  id_out = best;
  leaf_out = best_leaf;
  u_out = bu;
  v_out = bv;
}

// Distance to the winning triangle's plane in FP64 (the reference's
// t = ((A - o).N) / (d.N), geometryPrimitives.cpp:44): the watertight test
// decides hit / miss, this gives t to FP64 accuracy for the caller.
__device__ inline float refine_t(const float4 *tp, f3 o, f3 d) {
  const double ax = tp[0].x, ay = tp[0].y, az = tp[0].z;
  const double e1x = (double)tp[1].x - ax, e1y = (double)tp[1].y - ay, e1z = (double)tp[1].z - az;
  const double e2x = (double)tp[2].x - ax, e2y = (double)tp[2].y - ay, e2z = (double)tp[2].z - az;
  const double nx = e2y * e1z - e2z * e1y, ny = e2z * e1x - e2x * e1z, nz = e2x * e1y - e2y * e1x;
  const double num = (ax - o.x) * nx + (ay - o.y) * ny + (az - o.z) * nz;
  const double den = (double)d.x * nx + (double)d.y * ny + (double)d.z * nz;
  return (float)(num / den);
}

// One ray through the launch's scene or, in a batched launch, its own.
template <bool ANY, bool COUNT, int STACK, int BS, bool MULTI>
__device__ inline void trace_ray(const DeviceScene &S0, const SlotScenes &M, uint32_t *stk, size_t i,
                                 float4 ot, float4 ds, int leaf_ids, float &t, uint32_t &id,
                                 float &u, float &v, bool &any, uint32_t &leaf, uint32_t &nv,
                                 uint32_t &nt, uint32_t &overflow) {
  if (MULTI) {
    const DeviceScene *ps = M.scene + M.slot_seg[i];
    DeviceScene S;
    S.nodes = ps->nodes;
    S.tris = ps->tris;
    traverse<ANY, COUNT, STACK, BS>(S, stk, xyz(ot), xyz(ds), ot.w, __float_as_uint(ds.w),
                                    leaf_ids != 0, t, id, u, v, any, leaf, nv, nt, overflow);
  } else {
    traverse<ANY, COUNT, STACK, BS>(S0, stk, xyz(ot), xyz(ds), ot.w, __float_as_uint(ds.w),
                                    leaf_ids != 0, t, id, u, v, any, leaf, nv, nt, overflow);
  }
}

// A traced ray's outputs: the closest hit (t in FP64 accuracy for the
// caller-facing query), or the occlusion flag and, for an unoccluded shadow
// ray, its contribution added to its path.
template <bool ANY, bool MULTI>
__device__ inline void finish_ray(const DeviceScene &S, size_t i, float4 ot, float4 ds, float t,
                                  uint32_t id, float u, float v, bool any, uint32_t leaf,
                                  float4 *__restrict__ hits, uint8_t *__restrict__ occluded,
                                  const float4 *__restrict__ shadow_L,
                                  float4 *__restrict__ path_L, int leaf_ids) {
  if (!MULTI && !ANY && !leaf_ids && id != 0xFFFFFFFFu)
    t = refine_t(S.tris + 3 * (size_t)leaf, xyz(ot), xyz(ds));
  if (ANY) {
    if (occluded) occluded[i] = any ? 1 : 0;
    if (path_L && !any) {
      const float4 c = shadow_L[i];
      const uint32_t pid = __float_as_uint(c.w);
      float4 L = path_L[pid];
      L.x += c.x;
      L.y += c.y;
      L.z += c.z;
      path_L[pid] = L;
    }
  } else {
    hits[i] = make_float4(t, __uint_as_float(id), u, v);
  }
}

// One ray per thread: block b traces rays q * BS ... of segment
// r = b % kSegs, q = b / kSegs, then q + bps, q + 2 bps, ... (bps blocks
// per segment); blocks past a segment's live count exit at once.  (A
// persistent grid whose waves fetched batches from a per-segment cursor
// measured no faster once chunks hold 2^24 paths: the idle blocks of a
// launch cost less than the fetch loop.  Capping bps at 2^19 threads per
// segment against covering the whole capacity: within 1 %.)
template <bool ANY, bool COUNT, int STACK, int BS, bool MULTI>
__global__ __launch_bounds__(BS) void k_trace(DeviceScene S0, const float4 *__restrict__ rays,
                                              SegRange W, float4 *__restrict__ hits,
                                              uint8_t *__restrict__ occluded,
                                              const float4 *__restrict__ shadow_L,
                                              float4 *__restrict__ path_L, int leaf_ids,
                                              unsigned long long *__restrict__ counters,
                                              SlotScenes M, uint32_t *__restrict__ ovf,
                                              uint32_t ovf_cap) {
  __shared__ uint32_t stack_mem[STACK * BS];
  uint32_t *stk = stack_mem + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u;
  // block -> (segment, block within the segment); consecutive blocks go to
  // consecutive XCDs, so each segment's rays stay on one XCD's L2
  const uint32_t r = blockIdx.x % kSegs, q = blockIdx.x / kSegs;
  if (counters && blockIdx.x == 0 && threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (uint32_t k = 0; k < kSegs; ++k) tot += W.count_of(k);
    atomicAdd(&counters[3], tot);  // rays traced by this launch
  }
  const uint32_t cnt = W.count_of(r);
  if (q * BS >= cnt) return;  // whole block idle (uniform)
  // a grid of bps blocks per segment walks the segment's live rays
  uint32_t nv = 0, nt = 0, lost = 0;  // lost: overflowing rays the list had no room for
  const uint32_t nq = gridDim.x / kSegs;  // blocks per segment
  for (uint32_t qq = q; qq * BS < cnt; qq += nq) {  // block-uniform
    const uint32_t off = qq * BS + threadIdx.x;
    if (off >= cnt) continue;
    const size_t i = (size_t)r * W.segcap + off;
    const float4 ot = rays[2 * i], ds = rays[2 * i + 1];
    float t, u, v;
    uint32_t id, leaf, overflow = 0;
    bool any;
    trace_ray<ANY, COUNT, STACK, BS, MULTI>(S0, M, stk, i, ot, ds, leaf_ids, t, id, u, v, any, leaf,
                                            nv, nt, overflow);
    bool later = false;
    if (overflow && ovf) {
      // the stack was too small for this ray: listed, traced again with the
      // full stack by k_trace_retry
      const uint32_t k = atomicAdd(ovf, 1u);
      if (k < ovf_cap) {
        ovf[64 + k] = (uint32_t)i;
        overflow = 0;
        later = true;
      }
    }
    lost += overflow;
    if (!later)
      finish_ray<ANY, MULTI>(S0, i, ot, ds, t, id, u, v, any, leaf, hits, occluded, shadow_L,
                             path_L, leaf_ids);
  }
  // wave-uniform condition: every lane of the wave takes part in the shuffles
  if (COUNT || __any(lost != 0)) {
    unsigned long long a = nv, b = nt, c = lost;
    for (int o = 32; o > 0; o >>= 1) {
      a += __shfl_xor(a, o);
      b += __shfl_xor(b, o);
      c += __shfl_xor(c, o);
    }
    if (lane == 0 && counters) {
      if (COUNT) {
        atomicAdd(&counters[0], a);
        atomicAdd(&counters[1], b);
      }
      if (c) atomicAdd(&counters[2], c);
    }
  }
}

// Traversal with ray replacement (Aila and Laine's persistent threads with
// replacement of terminated rays, HPG 2009): block q of segment r owns the
// segment's rays [lo, hi) (an equal share); each lane starts on one, and
// after every leaf round the lanes whose ray has finished write its result
// and take the next ray of the block's share from a cursor in LDS, so a
// wave keeps its lanes busy instead of waiting for its slowest ray.  Each
// ray is traversed, and its result written, exactly as in k_trace.
// Measured on the 1080p headline render (closest hit after the camera rays
// at 1,024 blocks per segment, shadow rays at 768): 2,085 -> 2,293 Mpaths/s;
// for the coherent camera rays it is slower than k_trace (their waves hold
// 64 samples of one pixel, which replacement scatters).
// (The any-hit instantiation is bounded to at least 7 waves per SIMD: 71
// VGPRs; unbounded the compiler takes 75, 6 waves per SIMD, and its launches
// run about 5 % slower.)
template <bool ANY, int STACK, int BS, bool MULTI>
__global__ __launch_bounds__(BS, ANY ? 7 : 1) void k_trace_stream(
    DeviceScene S, const float4 *__restrict__ rays, SegRange W, float4 *__restrict__ hits,
    uint8_t *__restrict__ occluded, const float4 *__restrict__ shadow_L,
    float4 *__restrict__ path_L, int leaf_ids, unsigned long long *__restrict__ counters,
    SlotScenes M, uint32_t *__restrict__ ovf, uint32_t ovf_cap) {
  __shared__ uint32_t stack_mem[STACK * BS];
  __shared__ uint32_t cursor;
  uint32_t *stk = stack_mem + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t r = blockIdx.x % kSegs, q = blockIdx.x / kSegs, nb = gridDim.x / kSegs;
  if (counters && blockIdx.x == 0 && threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (uint32_t k = 0; k < kSegs; ++k) tot += W.count_of(k);
    atomicAdd(&counters[3], tot);  // rays traced by this launch
  }
  const uint32_t cnt = W.count_of(r);
  const uint32_t lo = (uint32_t)((uint64_t)cnt * q / nb);
  const uint32_t hi = (uint32_t)((uint64_t)cnt * (q + 1) / nb);
  if (lo >= hi) return;  // block-uniform
  if (threadIdx.x == 0) cursor = lo + BS;
  __syncthreads();
  const size_t seg0 = (size_t)r * W.segcap;
  // this lane's ray and its traversal state
  uint32_t idx = lo + threadIdx.x;
  bool has = idx < hi;
  f3 o = mk(0, 0, 0), d = mk(0, 0, 1), inv = mk(0, 0, 1);
  RayPre pre = ray_pre(d);
  float t = 0.0f, bu = 0.0f, bv = 0.0f;
  uint32_t skip = 0, best = 0xFFFFFFFFu, best_leaf = 0, entry = 0, pend = 0, lost = 0, spid = 0;
  f3 sc = mk(0, 0, 0);
  int sp = 0;
  bool live = false, hit_any = false, ovfl = false;
  // the ray's scene (a batched launch: the scene of the ray's request)
  const Node4Q *nodes = S.nodes;
  const float4 *tris = S.tris;
  auto start = [&]() {
    const size_t i = seg0 + idx;
    if (MULTI) {
      const DeviceScene *ps = M.scene + M.slot_seg[i];
      nodes = ps->nodes;
      tris = ps->tris;
    }
    const float4 ot = rays[2 * i], ds = rays[2 * i + 1];
    o = xyz(ot);
    d = xyz(ds);
    t = ot.w;
    skip = __float_as_uint(ds.w);
    // an unoccluded shadow ray adds its contribution to its path: the
    // contribution and the path id are read now, so that the add waits for
    // one load (the path's radiance) instead of two in a row
    if (ANY && path_L) {
      const float4 c = shadow_L[i];
      sc = xyz(c);
      spid = __float_as_uint(c.w);
    }
    pre = ray_pre(d);
    inv = mk(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
    best = 0xFFFFFFFFu;
    best_leaf = 0;
    bu = bv = 0.0f;
    entry = 0;
    pend = 0;
    sp = 0;
    live = true;
    hit_any = false;
    ovfl = false;
  };
  if (has) start();
  auto pop = [&]() {
    if (sp == 0) {
      live = false;
    } else {
      --sp;
      entry = stk[sp * BS];
    }
  };
  bool drained = false;  // the block's share is handed out (wave-uniform)
  while (true) {
    // ---- inner nodes (as traverse)
    while (true) {
      const bool searching = live && !(entry & kLeafBit);
      if (!__any(searching) ||
          (__popcll(__ballot(searching && pend == 0)) < 4 && __any(pend != 0)))
        break;  // wave-uniform
      if (!searching) continue;
      float key[4];
      uint32_t cid[4];
      const int nh = node_hits_q(nodes, entry, o, inv, t, key, cid);
#define CHK_CX(a, b)                                      \
  {                                                       \
    bool sw = key[b] < key[a];                            \
    float tk = sw ? key[b] : key[a];                      \
    key[b] = sw ? key[a] : key[b];                        \
    key[a] = tk;                                          \
    uint32_t tc = sw ? cid[b] : cid[a];                   \
    cid[b] = sw ? cid[a] : cid[b];                        \
    cid[a] = tc;                                          \
  }
      CHK_CX(0, 1) CHK_CX(2, 3) CHK_CX(0, 2) CHK_CX(1, 3) CHK_CX(1, 2)
#undef CHK_CX
      if (nh > 0) {
        entry = cid[0];
        const int np = nh - 1;  // pushed far to near, as in traverse
        if (sp + np > STACK) ovfl = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int pos = sp + np - 1 - k;
          if (k < np && pos < STACK) stk[pos * BS] = cid[1 + k];
        }
        sp = min(sp + np, STACK);
      } else {
        pop();
      }
      if (live && (entry & kLeafBit) && pend == 0) {
        pend = entry;
        pop();
      }
    }
    // ---- leaves (as traverse)
    while (true) {
      uint32_t lf = pend;
      pend = 0;
      if (lf == 0 && live && (entry & kLeafBit)) {
        lf = entry;
        pop();
      }
      if (!__any(lf != 0)) break;  // wave-uniform
      if (lf == 0) continue;
      const uint32_t first = leaf_first(lf), lcnt = leaf_count(lf);
      for (uint32_t k = 0; k < lcnt; ++k) {
        const uint32_t li = first + k;
        const float4 *tp = tris + 3 * (size_t)li;
        const float4 a = tp[0], b = tp[1], c = tp[2];
        const uint32_t tid = leaf_ids ? li : __float_as_uint(a.w);
        if (tid == skip) continue;
        float tt = t, uu, vv;
        if (tri_hit(pre, o, d, xyz(a), xyz(b), xyz(c), tt, uu, vv)) {
          t = tt;
          best = tid;
          best_leaf = li;
          bu = uu;
          bv = vv;
          hit_any = true;
          if (ANY) break;
        }
      }
      if (ANY && hit_any) live = false;
    }
    // ---- finished rays: results (or the overflow list), then the next ray
    if (has && !live) {
      const size_t i = seg0 + idx;
      bool later = false;
      if (ovfl && ovf) {
        const uint32_t k = atomicAdd(ovf, 1u);
        if (k < ovf_cap) {
          ovf[64 + k] = (uint32_t)i;
          later = true;
        }
      }
      if (ovfl && !later) ++lost;
      if (!later && ANY) {  // as finish_ray, with the shadow record already at hand
        if (occluded) occluded[i] = hit_any ? 1 : 0;
        if (path_L && !hit_any) {
          float4 L = path_L[spid];
          L.x += sc.x;
          L.y += sc.y;
          L.z += sc.z;
          path_L[spid] = L;
        }
      } else if (!later) {
        const float tr = best == 0xFFFFFFFFu ? 1e30f : t;
        finish_ray<ANY, MULTI>(S, i, make_float4(o.x, o.y, o.z, 0.0f),
                               make_float4(d.x, d.y, d.z, 0.0f), tr, best, bu, bv, hit_any,
                               best_leaf, hits, occluded, shadow_L, path_L, leaf_ids);
      }
      has = false;
    }
    const unsigned long long need = __ballot(!has);
    if (need != 0 && !drained) {  // wave-uniform
      const uint32_t n = (uint32_t)__popcll(need);
      const uint32_t leader = (uint32_t)__ffsll((long long)need) - 1;
      uint32_t b0 = 0;
      if (lane == leader) b0 = atomicAdd(&cursor, n);
      b0 = __shfl(b0, leader);
      if (b0 + n >= hi) drained = true;
      const uint32_t my = b0 + (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
      if (!has && my < hi) {
        idx = my;
        has = true;
        start();
      }
    }
    if (!__any(has)) break;  // wave-uniform
  }
  if (__any(lost != 0)) {
    unsigned long long c = lost;
    for (int k = 32; k > 0; k >>= 1) c += __shfl_xor(c, k);
    if (lane == 0 && counters && c) atomicAdd(&counters[2], c);
  }
}

// The rays k_trace listed as overflowing its stack, traced with a stack of
// STACK entries (the tree's worst case); results written as k_trace writes
// them.  A grid-stride loop over the list, whose length only the device
// knows.
template <bool ANY, int STACK, int BS, bool MULTI>
__global__ __launch_bounds__(BS) void k_trace_retry(DeviceScene S0, const float4 *__restrict__ rays,
                                                    float4 *__restrict__ hits,
                                                    uint8_t *__restrict__ occluded,
                                                    const float4 *__restrict__ shadow_L,
                                                    float4 *__restrict__ path_L, int leaf_ids,
                                                    unsigned long long *__restrict__ counters,
                                                    SlotScenes M, const uint32_t *__restrict__ ovf,
                                                    uint32_t ovf_cap) {
  __shared__ uint32_t stack_mem[STACK * BS];
  uint32_t *stk = stack_mem + threadIdx.x;
  const uint32_t n = ovf[0] < ovf_cap ? ovf[0] : ovf_cap;
  if (blockIdx.x == 0 && threadIdx.x == 0 && n) atomicAdd(const_cast<uint32_t *>(ovf) + 1, n);
  for (uint32_t k = blockIdx.x * BS + threadIdx.x; k < n; k += gridDim.x * BS) {
    const size_t i = ovf[64 + k];
    const float4 ot = rays[2 * i], ds = rays[2 * i + 1];
    float t, u, v;
    uint32_t id, leaf, nv = 0, nt = 0, overflow = 0;
    bool any;
    trace_ray<ANY, false, STACK, BS, MULTI>(S0, M, stk, i, ot, ds, leaf_ids, t, id, u, v, any,
                                            leaf, nv, nt, overflow);
    finish_ray<ANY, MULTI>(S0, i, ot, ds, t, id, u, v, any, leaf, hits, occluded, shadow_L, path_L,
                           leaf_ids);
    if (overflow && counters) atomicAdd(&counters[2], 1ull);
  }
}

// Closest sphere per ray: Sphere::intersect (geometryPrimitives.cpp:12-27,
// eps 1e-4) over every sphere in FP64, with the strict d < t of
// BvhAccel::intersectS (bvhAccel.cpp:171) and the ray's tmax as the initial
// t.  The sphere list is wave-uniform (scalar loads); spheres in the
// configs are a handful, so a brute-force loop beats a tree.
__global__ __launch_bounds__(256) void k_spheres(const float4 *__restrict__ sph, uint32_t ns,
                                                 const float4 *__restrict__ rays, uint32_t n,
                                                 float4 *__restrict__ hits) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 ot = rays[2 * (size_t)i], ds = rays[2 * (size_t)i + 1];
  const double ox = ot.x, oy = ot.y, oz = ot.z, dx = ds.x, dy = ds.y, dz = ds.z;
  const double eps = 1e-4;  // geometryPrimitives.h:39
  double best = ot.w;
  uint32_t id = 0xFFFFFFFFu;
  for (uint32_t k = 0; k < ns; ++k) {
    const float4 c = sph[k];
    const double px = c.x - ox, py = c.y - oy, pz = c.z - oz;
    const double b = px * dx + py * dy + pz * dz;
    double det = b * b - (px * px + py * py + pz * pz) + (double)c.w * c.w;
    if (det < 0.0) continue;
    det = sqrt(det);
    double t = b - det;
    if (!(t > eps)) {
      t = b + det;
      if (!(t > eps)) continue;
    }
    if (t < best) {
      best = t;
      id = k;
    }
  }
  hits[i] = make_float4(id == 0xFFFFFFFFu ? 1e30f : (float)best, __uint_as_float(id), 0.0f, 0.0f);
}

// ---------------------------------------------------------------- DummyAccel
// The reference's other accelerator (DummyAccel, dummyAccel.cpp:10-59): the
// triangles in their original order split into floor(sqrt(n)) runs of
// floor(sqrt(n)) (the last run takes the rest), one box per run; a ray tests
// the triangles of every run whose box it crosses, in order, keeping a hit
// only when strictly nearer (so the first of equally near triangles wins).
// Run boxes: (lo, first) (hi, end) as two float4.
__global__ __launch_bounds__(256) void k_dummy_boxes(DeviceScene S, uint32_t n_runs,
                                                     uint32_t run_len, float4 *__restrict__ box) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_runs) return;
  const uint32_t first = g * run_len, end = g + 1 == n_runs ? S.n_tris : first + run_len;
  float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
  for (uint32_t t = first; t < end; ++t) {
    const uint32_t leaf = S.orig_info[t].x;
    if (leaf == 0xFFFFFFFFu) continue;  // zero-area triangle: never hit
    const float4 *tp = S.tris + 3 * (size_t)leaf;
    for (int v = 0; v < 3; ++v) {
      const float4 q = tp[v];
      lo[0] = fminf(lo[0], q.x), hi[0] = fmaxf(hi[0], q.x);
      lo[1] = fminf(lo[1], q.y), hi[1] = fmaxf(hi[1], q.y);
      lo[2] = fminf(lo[2], q.z), hi[2] = fmaxf(hi[2], q.z);
    }
  }
  box[2 * g] = make_float4(lo[0], lo[1], lo[2], __uint_as_float(first));
  box[2 * g + 1] = make_float4(hi[0], hi[1], hi[2], __uint_as_float(end));
}

// One ray per lane; the runs and their triangles are wave-uniform.
template <bool ANY>
__global__ __launch_bounds__(256) void k_dummy_trace(DeviceScene S, const float4 *__restrict__ box,
                                                     uint32_t n_runs,
                                                     const float4 *__restrict__ rays, uint32_t n,
                                                     float4 *__restrict__ hits,
                                                     uint8_t *__restrict__ occluded) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  const float4 ot = live ? rays[2 * (size_t)i] : make_float4(0, 0, 0, 0);
  const float4 ds = live ? rays[2 * (size_t)i + 1] : make_float4(0, 0, 1, 0);
  const f3 o = xyz(ot), d = xyz(ds);
  const uint32_t skip = __float_as_uint(ds.w);
  const RayPre pre = ray_pre(d);
  const f3 inv = mk(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
  const float kGrow = 1.0f + 2.0f * 3.5762786865234375e-07f;  // 1 + 2*gamma_3
  float t = live ? ot.w : 0.0f, bu = 0.0f, bv = 0.0f;
  uint32_t best = 0xFFFFFFFFu, best_leaf = 0u;
  bool done = !live;
  for (uint32_t g = 0; g < n_runs; ++g) {
    const float4 lo = box[2 * g], hi = box[2 * g + 1];
    const float tx0 = (lo.x - o.x) * inv.x, tx1 = (hi.x - o.x) * inv.x;
    const float ty0 = (lo.y - o.y) * inv.y, ty1 = (hi.y - o.y) * inv.y;
    const float tz0 = (lo.z - o.z) * inv.z, tz1 = (hi.z - o.z) * inv.z;
    const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), t));
    const bool in_box = !done && tn <= tf * kGrow;
    if (!__any(in_box)) continue;  // wave-uniform
    const uint32_t first = __float_as_uint(lo.w), end = __float_as_uint(hi.w);
    for (uint32_t tid = first; tid < end; ++tid) {
      const uint32_t leaf = S.orig_info[tid].x;
      if (leaf == 0xFFFFFFFFu) continue;  // uniform
      const float4 *tp = S.tris + 3 * (size_t)leaf;
      float tt = t, uu, vv;
      if (in_box && tid != skip && tri_hit(pre, o, d, xyz(tp[0]), xyz(tp[1]), xyz(tp[2]), tt, uu, vv)) {
        t = tt;
        best = tid;
        best_leaf = leaf;
        bu = uu;
        bv = vv;
        if (ANY) done = true;
      }
    }
    if (ANY && !__any(!done)) break;  // every lane occluded (or idle)
  }
  if (!live) return;
  if (ANY) {
    occluded[i] = best != 0xFFFFFFFFu ? 1 : 0;
  } else {
    const float tr = best == 0xFFFFFFFFu ? 1e30f : refine_t(S.tris + 3 * (size_t)best_leaf, o, d);
    hits[i] = make_float4(tr, __uint_as_float(best), bu, bv);
  }
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

template <bool ANY, bool COUNT, int STACK, int BS, bool MULTI>
hipError_t launch_t(const chk_scene_s *s, const TraceShape &T, const float4 *rays,
                    const SegRange &W, float4 *hits, uint8_t *occ, const float4 *sL, float4 *pL,
                    bool leaf_ids, unsigned long long *counters, const SlotScenes &M,
                    hipStream_t st, uint32_t *ovf, uint32_t ovf_cap) {
  // at most 2^19 threads per segment (4,096 blocks of 128; about 16 blocks
  // per CU of the segment's XCD, 8 times over): a sparse bounce dispatches
  // no more blocks than that, a full one loops
  const uint32_t bps = std::min(W.segcap / BS, (1u << 19) / BS);
  const uint32_t grid = kSegs * bps;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((k_trace<ANY, COUNT, STACK, BS, MULTI>), dim3(grid), dim3(BS), 0, st,
                     s ? s->d : DeviceScene{}, rays, W, hits, occ, sL, pL, leaf_ids ? 1 : 0,
                     counters, M, ovf, ovf_cap);
  return hipGetLastError();
}

// The traversal at a given stack depth (8 only for single-scene 128-thread
// launches) and block size.
template <bool ANY, bool COUNT, bool MULTI>
hipError_t launch_s(const chk_scene_s *s, int stack, const TraceShape &T, const float4 *rays,
                    const SegRange &W, float4 *hits, uint8_t *occ, const float4 *sL, float4 *pL,
                    bool leaf_ids, unsigned long long *counters, const SlotScenes &M,
                    hipStream_t st, uint32_t *ovf = nullptr, uint32_t ovf_cap = 0) {
#define CHK_LAUNCH(STK, BSZ)                                                                 \
  return launch_t<ANY, COUNT, STK, BSZ, MULTI>(s, T, rays, W, hits, occ, sL, pL, leaf_ids,    \
                                               counters, M, st, ovf, ovf_cap)
#define CHK_BLOCKS(STK)            \
  switch (T.block) {               \
    case 64: CHK_LAUNCH(STK, 64);  \
    case 256: CHK_LAUNCH(STK, 256); \
    default: CHK_LAUNCH(STK, 128); \
  }
  if (!MULTI && !COUNT && stack == 8 && T.block == 128) {
    CHK_LAUNCH(8, 128);
  }
  switch (stack) {
    case 16: CHK_BLOCKS(16);
    case 32: CHK_BLOCKS(32);
    case 64: CHK_BLOCKS(64);
    default: CHK_BLOCKS(96);
  }
#undef CHK_BLOCKS
#undef CHK_LAUNCH
}

// The listed rays again with the tree's full stack (fixed grid: four
// blocks per CU, 8 waves per CU within the LDS of the 64-entry stack; an
// empty list costs one short launch).
template <bool ANY, bool MULTI>
hipError_t launch_retry(const chk_scene_s *s, int stack, const float4 *rays, float4 *hits,
                        uint8_t *occ, const float4 *sL, float4 *pL, bool leaf_ids,
                        unsigned long long *counters, const SlotScenes &M, hipStream_t st,
                        const uint32_t *ovf, uint32_t ovf_cap) {
  const DeviceScene S = s ? s->d : DeviceScene{};
  const dim3 g(1024), b(128);
  if (stack <= 32)
    hipLaunchKernelGGL((k_trace_retry<ANY, 32, 128, MULTI>), g, b, 0, st, S, rays, hits, occ, sL,
                       pL, leaf_ids ? 1 : 0, counters, M, ovf, ovf_cap);
  else if (stack <= 64)
    hipLaunchKernelGGL((k_trace_retry<ANY, 64, 128, MULTI>), g, b, 0, st, S, rays, hits, occ, sL,
                       pL, leaf_ids ? 1 : 0, counters, M, ovf, ovf_cap);
  else
    hipLaunchKernelGGL((k_trace_retry<ANY, 96, 128, MULTI>), g, b, 0, st, S, rays, hits, occ, sL,
                       pL, leaf_ids ? 1 : 0, counters, M, ovf, ovf_cap);
  return hipGetLastError();
}

template <bool ANY>
hipError_t launch_kind(const chk_scene_s *s, const TraceShape &T, const float4 *rays,
                       const SegRange &W, float4 *hits, uint8_t *occ, const float4 *sL,
                       float4 *pL, bool leaf_ids, bool count, unsigned long long *counters,
                       const SlotScenes &M, hipStream_t st) {
  const bool multi = M.scene != nullptr;
  if (count)  // the counted render (work model): single scene, 128-thread blocks, full stack
    return launch_s<ANY, true, false>(s, T.stack, TraceShape{T.stack, 128}, rays, W, hits, occ,
                                      sL, pL, leaf_ids, counters, M, st);
  int shallow = T.shallow;
  if (shallow == 8 && (multi || T.block != 128)) shallow = 16;
  if (T.stream > 0 && T.ovf && T.stack > 16 && shallow == 16 && (T.block == 64 || T.block == 128)) {
    // ray replacement over the 16-entry stack, then the overflow list again
    // with the full stack (as below); T.stream counts 128-thread blocks
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, T.ovf, 1u);
    const DeviceScene S = s ? s->d : DeviceScene{};
    const uint32_t bps = std::min<uint32_t>(W.segcap / T.block, T.stream * (128 / T.block));
    const dim3 g(kSegs * bps), b(T.block);
    const int li = leaf_ids ? 1 : 0;
    if (T.block == 64) {
      if (multi)
        hipLaunchKernelGGL((k_trace_stream<ANY, 16, 64, true>), g, b, 0, st, S, rays, W, hits, occ,
                           sL, pL, li, counters, M, T.ovf, T.ovf_cap);
      else
        hipLaunchKernelGGL((k_trace_stream<ANY, 16, 64, false>), g, b, 0, st, S, rays, W, hits, occ,
                           sL, pL, li, counters, M, T.ovf, T.ovf_cap);
    } else {
      if (multi)
        hipLaunchKernelGGL((k_trace_stream<ANY, 16, 128, true>), g, b, 0, st, S, rays, W, hits,
                           occ, sL, pL, li, counters, M, T.ovf, T.ovf_cap);
      else
        hipLaunchKernelGGL((k_trace_stream<ANY, 16, 128, false>), g, b, 0, st, S, rays, W, hits,
                           occ, sL, pL, li, counters, M, T.ovf, T.ovf_cap);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return multi ? launch_retry<ANY, true>(s, T.stack, rays, hits, occ, sL, pL, leaf_ids, counters,
                                           M, st, T.ovf, T.ovf_cap)
                 : launch_retry<ANY, false>(s, T.stack, rays, hits, occ, sL, pL, leaf_ids,
                                            counters, M, st, T.ovf, T.ovf_cap);
  }
  if (T.ovf && T.stack > shallow) {
    // a stack shorter than the tree's worst case for every ray (more waves
    // per CU), the rare overflowing ray again with the full stack
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, T.ovf, 1u);
    hipError_t e = multi ? launch_s<ANY, false, true>(s, shallow, T, rays, W, hits, occ, sL, pL,
                                                      leaf_ids, counters, M, st, T.ovf, T.ovf_cap)
                         : launch_s<ANY, false, false>(s, shallow, T, rays, W, hits, occ, sL, pL,
                                                       leaf_ids, counters, M, st, T.ovf,
                                                       T.ovf_cap);
    if (e != hipSuccess) return e;
    return multi ? launch_retry<ANY, true>(s, T.stack, rays, hits, occ, sL, pL, leaf_ids,
                                           counters, M, st, T.ovf, T.ovf_cap)
                 : launch_retry<ANY, false>(s, T.stack, rays, hits, occ, sL, pL, leaf_ids,
                                            counters, M, st, T.ovf, T.ovf_cap);
  }
  return multi ? launch_s<ANY, false, true>(s, T.stack, T, rays, W, hits, occ, sL, pL, leaf_ids,
                                            counters, M, st)
               : launch_s<ANY, false, false>(s, T.stack, T, rays, W, hits, occ, sL, pL,
                                             leaf_ids, counters, M, st);
}

}  // namespace

int pick_stack_size(uint32_t bound) {
  if (bound <= 32) return 32;
  if (bound <= 64) return 64;
  if (bound <= 96) return 96;
  return -1;
}

hipError_t launch_closest(const chk_scene_s *s, const TraceShape &T, const float4 *rays,
                          const SegRange &W, float4 *hits, bool leaf_ids, bool count,
                          unsigned long long *counters, const SlotScenes &M, hipStream_t st) {
  return launch_kind<false>(s, T, rays, W, hits, nullptr, nullptr, nullptr, leaf_ids, count,
                            counters, M, st);
}

hipError_t launch_any(const chk_scene_s *s, const TraceShape &T, const float4 *rays,
                      const SegRange &W, uint8_t *occluded, const float4 *shadow_L,
                      float4 *path_L, bool leaf_ids, bool count, unsigned long long *counters,
                      const SlotScenes &M, hipStream_t st) {
  return launch_kind<true>(s, T, rays, W, nullptr, occluded, shadow_L, path_L, leaf_ids, count,
                           counters, M, st);
}

hipError_t launch_spheres(const float *spheres, uint32_t ns, const float4 *rays, uint32_t n,
                          float4 *hits, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_spheres, dim3((n + 255) / 256), dim3(256), 0, st,
                     reinterpret_cast<const float4 *>(spheres), ns, rays, n, hits);
  return hipGetLastError();
}

hipError_t launch_dummy_boxes(const chk_scene_s *s, uint32_t n_runs, uint32_t run_len,
                              float4 *box, hipStream_t st) {
  if (n_runs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dummy_boxes, dim3((n_runs + 255) / 256), dim3(256), 0, st, s->d, n_runs,
                     run_len, box);
  return hipGetLastError();
}

hipError_t launch_dummy_trace(const chk_scene_s *s, bool any, const float4 *box, uint32_t n_runs,
                              const float4 *rays, uint32_t n, float4 *hits, uint8_t *occluded,
                              hipStream_t st) {
  if (n == 0) return hipSuccess;
  const dim3 g((n + 255) / 256), b(256);
  if (any)
    hipLaunchKernelGGL(k_dummy_trace<true>, g, b, 0, st, s->d, box, n_runs, rays, n, nullptr,
                       occluded);
  else
    hipLaunchKernelGGL(k_dummy_trace<false>, g, b, 0, st, s->d, box, n_runs, rays, n, hits,
                       nullptr);
  return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t *p, uint32_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

}  // namespace chk
