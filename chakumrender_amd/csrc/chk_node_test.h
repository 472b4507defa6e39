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
// s_a is a power of two), one fma per plane.  The rounding of b_a (at most
// 2^-23 |b_a|) is absolute in t, so b_a is formed twice, from o_a - ro_a
// moved down and up by 2^-20 of its magnitude: the low planes take the
// lowered one and the high planes the raised one, which for either sign of
// inv_a only widens the slab.  The rounding left is relative and the far
// distance is scaled by kGrow for it as in Ize's test.  A box that the exact
// slab test accepts is therefore always accepted.  Empty slots are skipped
// by their id.
__host__ __device__ inline int node_hits_q(const Node4Q *nodes, uint32_t entry, f3 o, f3 inv, float t,
                                  float key[4], uint32_t cid[4]) {
  const float kGrow = 1.0f + 2.0f * 3.5762786865234375e-07f;  // 1 + 2*gamma_3
  const uint4 *np = reinterpret_cast<const uint4 *>(nodes + entry);
  const uint4 w0 = np[0], w1 = np[1], w2 = np[2], ch = np[3];
  cid[0] = ch.x, cid[1] = ch.y, cid[2] = ch.z, cid[3] = ch.w;
  const float dx = __builtin_bit_cast(float, w0.x) - o.x;
  const float dy = __builtin_bit_cast(float, w0.y) - o.y;
  const float dz = __builtin_bit_cast(float, w0.z) - o.z;
  const float blx = fmaf(-fabsf(dx), 0x1p-20f, dx) * inv.x;
  const float bhx = fmaf(fabsf(dx), 0x1p-20f, dx) * inv.x;
  const float bly = fmaf(-fabsf(dy), 0x1p-20f, dy) * inv.y;
  const float bhy = fmaf(fabsf(dy), 0x1p-20f, dy) * inv.y;
  const float blz = fmaf(-fabsf(dz), 0x1p-20f, dz) * inv.z;
  const float bhz = fmaf(fabsf(dz), 0x1p-20f, dz) * inv.z;
  const float cx = __builtin_bit_cast(float, w0.w) * inv.x;
  const float cy = __builtin_bit_cast(float, w2.z) * inv.y;
  const float cz = __builtin_bit_cast(float, w2.w) * inv.z;
  int nh = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int sh = 8 * c;
    const float tx0 = fmaf((float)((w1.x >> sh) & 255u), cx, blx);
    const float tx1 = fmaf((float)((w1.y >> sh) & 255u), cx, bhx);
    const float ty0 = fmaf((float)((w1.z >> sh) & 255u), cy, bly);
    const float ty1 = fmaf((float)((w1.w >> sh) & 255u), cy, bhy);
    const float tz0 = fmaf((float)((w2.x >> sh) & 255u), cz, blz);
    const float tz1 = fmaf((float)((w2.y >> sh) & 255u), cz, bhz);
    const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), t));
    const bool h = cid[c] != kEmptyChild && tn <= tf * kGrow;
    key[c] = h ? tn : kInf;
    nh += h ? 1 : 0;
  }
  return nh;
}

}  // namespace chk
