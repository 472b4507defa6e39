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
constexpr int kFragF0 = 0;
constexpr int kFragLayerF = 16 * 64 * 4;  // 4096 bf16 per 64x64 layer
constexpr int kFragLastF = 12 * 64 * 4;   // 3072 for the output layer
constexpr int kFragTotal = 3 * kFragLayerF + kFragLastF;  // per F or G set

struct MlpDev {
  const uint16_t *F;  // kFragTotal bf16
  const uint16_t *G;  // kFragTotal bf16
  const float *bias;  // b1 b2 b3 (64 each) b4 (33 -> 48 padded with 0): 240 floats
  uint32_t n_lights;
};

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
