// Learned light-selection proposal: a 64-64-64-64-33 ReLU MLP with BF16
// weights and FP32 accumulation on gfx950 MFMA (v_mfma_f32_16x16x16_bf16).
//
// No reference counterpart: the reference samples every mesh light with the
// same count (mis.h:9, mis.cpp:127-138) and weights samples with the balance
// heuristic (mis.cpp:209-220).  The network replaces the light-selection pmf
// and the light/BSDF mixing constants of that heuristic; a defensive mixture
// keeps every light's pmf >= eps/L and the mixing weight in [eps, 1-eps], so
// the estimator stays unbiased for any weights.
//
// Layouts (the whole computation keeps points in the MFMA N dimension):
//   Y^T (out x pts) = W^T (out x in) . X^T (in x pts)
// The 16x16 D tile of one layer (lane l: rows 4(l/16)+v, column l%16) is
// exactly the B operand of the next layer (lane l: k rows 4(l/16)+j,
// column l%16), so activations never leave registers between layers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace chk {

constexpr int kMlpIn = 64;       // encoded input width
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 33;      // 32 light logits + 1 mixing logit
constexpr int kMlpOutPad = 48;   // three 16-row tiles
constexpr int kMaxProposalLights = 32;
// flat parameter vector: W1[64x64] b1[64] W2 b2 W3 b3 W4[64x33] b4[33]
// (W[in][out], in-major)
constexpr int kMlpParams = 3 * (64 * 64 + 64) + 64 * 33 + 33;  // 14625
constexpr float kPropEpsQ = 0.1f;   // defensive share of the uniform light pmf
constexpr float kPropEpsC = 0.1f;   // mixing weight kept in [eps, 1 - eps]

inline int mlp_param_offset_w(int layer) { return layer * (64 * 64 + 64); }
inline int mlp_param_offset_b(int layer) { return layer * (64 * 64 + 64) + 64 * (layer < 3 ? 64 : 33); }

// Fragment arrays (bf16 bits as uint16), per layer l = 0..3 with
// OT = 4 (l < 3) or 3 (l = 3) output tiles and 4 input tiles:
//   F[l]: forward A fragments, index ((ot * 4 + kt) * 64 + lane) * 4 + j
//         = W[in = kt*16 + 4*(lane/16) + j][out = ot*16 + lane%16]
//   G[l]: backward A fragments, index ((it * OT + ot) * 64 + lane) * 4 + j
//         = W[in = it*16 + lane%16][out = ot*16 + 4*(lane/16) + j]
// Out-of-range outputs (>= 33) are zero.
//   I[l]: inference A fragments for v_mfma_f32_16x16x32_bf16, index
//         ((ot * 2 + kt) * 64 + lane) * 8 + j
//         = W[in = perm_l(kt*32 + 8*(lane/16) + j)][out = ot*16 + lane%16]
//         with the input orders of the inference kernel: perm_enc for l = 0
//         (lane group g computes octave g of every input), perm_hid for
//         l > 0 (the 16 activations a lane holds after a layer, in order).
constexpr int kFragF0 = 0;
constexpr int kFragLayerF = 16 * 64 * 4;  // 4096 bf16 per 64x64 layer
constexpr int kFragLastF = 12 * 64 * 4;   // 3072 for the output layer
constexpr int kFragTotal = 3 * kFragLayerF + kFragLastF;  // per F or G set

struct MlpDev {
  const uint16_t *F;  // kFragTotal bf16
  const uint16_t *G;  // kFragTotal bf16
  const uint16_t *I;  // kFragTotal bf16
  const float *bias;  // b1 b2 b3 (64 each) b4 (33 -> 48 padded with 0): 240 floats
  uint32_t n_lights;
};

// Input orders of the inference fragments.  K position kp of a 32-wide K
// tile pair: lane group g = (kp % 32) / 8 holds lane-local value
// u = (kp / 32) * 8 + kp % 8 (0..15).
//   perm_enc: u < 14: sin (u even) or cos (u odd) of 2^g pi s_(u/2), i.e.
//   encoding feature 9 (u/2) + 1 + g or 9 (u/2) + 5 + g; u = 14, 15: feature
//   9 e (the input s_e) for e = 2g + u - 14 < 7, else 63 (the constant 1).
//   perm_hid: after a layer a lane holds rows 16 ot + 4 g + v of the four
//   output tiles; u = 4 ot + v.
__host__ __device__ constexpr int perm_enc(int kp) {
  const int g = (kp % 32) / 8, u = (kp / 32) * 8 + kp % 8;
  return u < 14 ? 9 * (u / 2) + ((u & 1) ? 5 : 1) + g
                : (2 * g + u - 14 < 7 ? 9 * (2 * g + u - 14) : 63);
}
__host__ __device__ constexpr int perm_hid(int kp) {
  const int g = (kp % 32) / 8, u = (kp / 32) * 8 + kp % 8;
  return 16 * (u / 4) + 4 * g + u % 4;
}

// Per-vertex training record of the learned render (40 bytes), written by
// the shading kernel at the vertex, completed by the shadow pass (occlusion
// of the light sample) and by the next bounce (emission found by the BSDF
// sample).  Scalars are luminances (mean of RGB) of the unweighted integrand
// g = f * L * cos in solid-angle measure.
struct TrainRec {
  uint32_t light;    // light j drawn from q' (0xFFFFFFFF: no NEE sample)
  float gl;          // g of the light sample, before visibility
  float pl, pbl;     // light-sampling and BSDF pdfs of the light sample
  float fb;          // f * cos * albedo of the BSDF sample
  float le;          // emitted radiance its ray found (0: none)
  float plb, pbb;    // light-sampling and BSDF pdfs of the BSDF sample
  uint32_t sh_slot;  // shadow-queue slot of the light sample (0xFFFFFFFF: none)
  uint32_t light_b;  // light the BSDF sample hit (0xFFFFFFFF: none)
};
static_assert(sizeof(TrainRec) == 40, "TrainRec is 40 bytes");

}  // namespace chk
