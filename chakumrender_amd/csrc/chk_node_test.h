// Ray / node test of the traversal kernels over the quantized 4-wide nodes
// (chk_bvh.h Node4Q).  Host and device: tests/native/bvh_quant_check.cpp
// runs it on the CPU against an exact slab test.
#pragma once

#include "chk_bvh.h"
#include "chk_device.h"

namespace chk {

__host__ __device__ inline float safe_rcp(float x) {
  return 1.0f / (fabsf(x) > 1e-20f ? x : copysignf(1e-20f, x));
}

// The children of quantized inner node `entry` whose boxes the ray enters
// before t: entry distances in key (kInf for the others), ids in cid, and
// their number.  Axis a's plane o_a + q s_a is crossed at
// t = b_a + q c_a with b_a = (o_a - ro_a) inv_a and c_a = s_a inv_a (exact:
// s_a is a power of two), one fma per plane.  The near plane of axis a is
// the low one when inv_a > 0 and the high one otherwise, so the near and far
// grid indices are picked once per node and a slab needs no min / max.  The
// rounding of b_a (at most 2^-23 |b_a|) is absolute in t, so b_a is moved by
// 2^-20 of its magnitude, down for the near distances and up for the far
// ones.  The rounding left is relative and the far distance is scaled by
// kGrow for it as in Ize's test.
// A box that the exact slab test accepts is therefore always accepted.
// Empty slots are skipped by their id.
__host__ __device__ inline int node_hits_q(const Node4Q *nodes, uint32_t entry, f3 o, f3 inv,
                                           float t, float key[4], uint32_t cid[4]) {
  const float kGrow = 1.0f + 2.0f * 3.5762786865234375e-07f;  // 1 + 2*gamma_3
  const uint4 *np = reinterpret_cast<const uint4 *>(nodes + entry);
  const uint4 w0 = np[0], w1 = np[1], w2 = np[2], ch = np[3];
  cid[0] = ch.x, cid[1] = ch.y, cid[2] = ch.z, cid[3] = ch.w;
  const bool sx = inv.x < 0.0f, sy = inv.y < 0.0f, sz = inv.z < 0.0f;
  const uint32_t qnx = sx ? w1.y : w1.x, qfx = sx ? w1.x : w1.y;
  const uint32_t qny = sy ? w1.w : w1.z, qfy = sy ? w1.z : w1.w;
  const uint32_t qnz = sz ? w2.y : w2.x, qfz = sz ? w2.x : w2.y;
  const float mx = (__builtin_bit_cast(float, w0.x) - o.x) * inv.x;
  const float my = (__builtin_bit_cast(float, w0.y) - o.y) * inv.y;
  const float mz = (__builtin_bit_cast(float, w0.z) - o.z) * inv.z;
  // b_a -+ 2^-20 |b_a|: the near distances lowered, the far ones raised
  const float bnx = fmaf(-fabsf(mx), 0x1p-20f, mx), bfx = fmaf(fabsf(mx), 0x1p-20f, mx);
  const float bny = fmaf(-fabsf(my), 0x1p-20f, my), bfy = fmaf(fabsf(my), 0x1p-20f, my);
  const float bnz = fmaf(-fabsf(mz), 0x1p-20f, mz), bfz = fmaf(fabsf(mz), 0x1p-20f, mz);
  const float cx = __builtin_bit_cast(float, w0.w) * inv.x;
  const float cy = __builtin_bit_cast(float, w2.z) * inv.y;
  const float cz = __builtin_bit_cast(float, w2.w) * inv.z;
  int nh = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int sh = 8 * c;
    const float tnx = fmaf((float)((qnx >> sh) & 255u), cx, bnx);
    const float tfx = fmaf((float)((qfx >> sh) & 255u), cx, bfx);
    const float tny = fmaf((float)((qny >> sh) & 255u), cy, bny);
    const float tfy = fmaf((float)((qfy >> sh) & 255u), cy, bfy);
    const float tnz = fmaf((float)((qnz >> sh) & 255u), cz, bnz);
    const float tfz = fmaf((float)((qfz >> sh) & 255u), cz, bfz);
    const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
    const float tf = fminf(fminf(tfx, tfy), fminf(tfz, t));
    const bool h = cid[c] != kEmptyChild && tn <= tf * kGrow;
    key[c] = h ? tn : kInf;
    nh += h ? 1 : 0;
  }
  return nh;
}

}  // namespace chk
