// Learned proposal network on MFMA: inference, backward deltas, weight
// gradients (split-K MFMA GEMM), deterministic reduction and Adam.  See
// chk_mlp.h for the layouts.
#include <cmath>
#include <cstring>
#include <vector>

#include "chk_device.h"
#include "chk_internal.h"
#include "chk_mlp.h"

namespace chk {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
// Round to nearest even, as f2bf does for every finite value; gfx950 converts
// two values per v_cvt_pk_bf16_f32.
__device__ inline short bf16_bits(float a) { return __builtin_bit_cast(short, (__bf16)a); }
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float fl2 __attribute__((ext_vector_type(2)));
typedef uint32_t u2 __attribute__((ext_vector_type(2)));
__device__ inline uint32_t pack2(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((fl2{a, b}), bf2));
}
__device__ inline s4 pack4(float a, float b, float c, float d) {
  return __builtin_bit_cast(s4, (u2{pack2(a, b), pack2(c, d)}));
}
__device__ inline float bf2f(short b) { return __uint_as_float(((uint32_t)(uint16_t)b) << 16); }

__device__ inline f4 mfma(s4 a, s4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Frequency encoding of the 7 inputs: for scalar i, features 9i + r are
// s, sin(2^k pi s) (r = 1+k), cos(2^k pi s) (r = 5+k), k = 0..3; feature 63 = 1.
__device__ inline float encode(int f, const float *s) {
  if (f >= 63) return 1.0f;
  const int i = f / 9, r = f - 9 * i;
  const float x = s[i];
  if (r == 0) return x;
  const int k = r <= 4 ? r - 1 : r - 5;
  float sn, cs;
  sincospif((float)(1 << k) * x, &sn, &cs);
  return r <= 4 ? sn : cs;
}

__device__ inline void encode_b(const float *s, int lane, s4 xin[4]) {
  const int g = lane >> 4;
#pragma unroll
  for (int kt = 0; kt < 4; ++kt) {
    const int f0 = kt * 16 + 4 * g;
    xin[kt] = pack4(encode(f0, s), encode(f0 + 1, s), encode(f0 + 2, s), encode(f0 + 3, s));
  }
}

__device__ inline const s4 *frag_f(const MlpDev &W, int l) {
  return reinterpret_cast<const s4 *>(W.F + (l < 3 ? l * kFragLayerF : 3 * kFragLayerF));
}
__device__ inline const s4 *frag_g(const MlpDev &W, int l) {
  return reinterpret_cast<const s4 *>(W.G + (l < 3 ? l * kFragLayerF : 3 * kFragLayerF));
}

// One layer over NG groups of 16 points: the accumulators start at the bias,
// each A fragment (from F, global or LDS) feeds the NG groups' MFMAs.  OT
// output tiles; hidden layers then apply ReLU and round to bf16 B operands.
template <int NG, int OT>
__device__ inline void layer_acc(const s4 *F, const float *b, const s4 (&in)[NG][4], int lane,
                                 f4 (&acc)[NG][OT]) {
  const int g = lane >> 4;
#pragma unroll
  for (int ot = 0; ot < OT; ++ot) {
    const f4 bias = *reinterpret_cast<const f4 *>(b + ot * 16 + 4 * g);
    const s4 a0 = F[ot * 4 * 64 + lane];
#pragma unroll
    for (int n = 0; n < NG; ++n) acc[n][ot] = mfma(a0, in[n][0], bias);
#pragma unroll
    for (int kt = 1; kt < 4; ++kt) {
      const s4 a = F[(ot * 4 + kt) * 64 + lane];
#pragma unroll
      for (int n = 0; n < NG; ++n) acc[n][ot] = mfma(a, in[n][kt], acc[n][ot]);
    }
  }
}
// ReLU after rounding: bf16 bit patterns order as signed 16-bit integers
// (negative values and -0 below +0), so max(bits, 0) is bf16(max(x, 0)) up to
// the sign of zero, two lanes per v_pk_max_i16.
__device__ inline s4 relu_bf16(s4 v) { return __builtin_elementwise_max(v, s4{0, 0, 0, 0}); }
template <int NG>
__device__ inline void hidden_n(const s4 *F, const float *b, const s4 (&in)[NG][4], int lane,
                                s4 (&out)[NG][4]) {
  f4 acc[NG][4];
  layer_acc<NG, 4>(F, b, in, lane, acc);
#pragma unroll
  for (int n = 0; n < NG; ++n)
#pragma unroll
    for (int ot = 0; ot < 4; ++ot)
      out[n][ot] = relu_bf16(pack4(acc[n][ot][0], acc[n][ot][1], acc[n][ot][2], acc[n][ot][3]));
}

__device__ inline void hidden(const MlpDev &W, int l, const s4 in[4], int lane, s4 out[4]) {
  const s4 (&i1)[1][4] = *reinterpret_cast<const s4 (*)[1][4]>(in);
  s4 (&o1)[1][4] = *reinterpret_cast<s4 (*)[1][4]>(out);
  hidden_n<1>(frag_f(W, l), W.bias + 64 * l, i1, lane, o1);
}

// Softmax over the first n_lights logits of each point (column), defensive
// mixture, and the mixing weight.  Lanes l, l^16, l^32, l^48 share a point.
// FULL: n_lights is 32 (no logit is masked).
template <bool FULL>
__device__ inline void proposal_head(const f4 out[3], int lane, uint32_t L, float q[8], float &mix,
                                     float &sum_e, float &mx) {
  const int g = lane >> 4;
  float e[8];
  mx = -INFINITY;
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int o = t * 16 + 4 * g + v;
      e[t * 4 + v] = FULL || (uint32_t)o < L ? out[t][v] : -INFINITY;
      mx = fmaxf(mx, e[t * 4 + v]);
    }
  mx = fmaxf(mx, __shfl_xor(mx, 16));
  mx = fmaxf(mx, __shfl_xor(mx, 32));
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    e[k] = __expf(e[k] - mx);  // 0 for a masked logit (mx is finite)
    s += e[k];
  }
  s += __shfl_xor(s, 16);
  s += __shfl_xor(s, 32);
  sum_e = s;
  // v_rcp_f32 (1 ulp) for the two reciprocals and the logistic below
  const float invL = FULL ? 1.0f / 32.0f : __builtin_amdgcn_rcpf((float)L);
  const float inv_s = __builtin_amdgcn_rcpf(s);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int o = (k >> 2) * 16 + 4 * g + (k & 3);
    q[k] = FULL || (uint32_t)o < L ? (1.0f - kPropEpsQ) * (e[k] * inv_s) + kPropEpsQ * invL : 0.0f;
  }
  // mixing logit: output 32 = tile 2, row 0 -> lanes of group 0, v = 0
  const float zc = __shfl(out[2][0], lane & 15);
  mix = kPropEpsC + (1.0f - 2.0f * kPropEpsC) * __builtin_amdgcn_rcpf(1.0f + __expf(-zc));
}

// ------------------------------------------------------------------ kernels
// Inference over the items of R on v_mfma_f32_16x16x32_bf16.  A block
// stages the inference fragments and the biases in LDS once; each wave then
// takes NG groups of 16 points at a time from its segment
// (r = blockIdx % kSegs).  A point's four lanes (lane % 16, groups
// g = lane / 16) each hold 16 of its values in B-operand order:
//  * encoding: lane group g evaluates octave g, sin and cos of
//    2^g pi s_i for the 7 inputs (v_sin_f32 / v_cos_f32 take revolutions:
//    2^(g-1) s_i), plus two of the 8 linear and constant features, and
//    packs them to bf16 in registers (input order perm_enc);
//  * after each layer a lane's D rows are the next layer's B operand as
//    they stand (input order perm_hid), so activations never leave
//    registers;
//  * every A fragment read from LDS feeds the NG groups' MFMAs.
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef short sh8 __attribute__((ext_vector_type(8)));
constexpr int kEvalBlock = 512;
constexpr int kEvalWaves = kEvalBlock / 64;
constexpr int kEvalGroups = 2;  // groups of 16 points per wave step

__device__ inline f4 mfma32(bf8 a, bf8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
typedef uint32_t u4v __attribute__((ext_vector_type(4)));
__device__ inline bf8 pack8(f4 a, f4 b) {
  return __builtin_bit_cast(bf8, (u4v{pack2(a[0], a[1]), pack2(a[2], a[3]), pack2(b[0], b[1]),
                                      pack2(b[2], b[3])}));
}
__device__ inline bf8 relu8(bf8 v) {
  return __builtin_bit_cast(bf8, __builtin_elementwise_max(__builtin_bit_cast(sh8, v), sh8{}));
}

template <int NG, int OT>
__device__ inline void layer32(const bf8 *F, const float *b, const bf8 (&in)[NG][2], int lane,
                               f4 (&acc)[NG][OT]) {
  const int g = lane >> 4;
#pragma unroll
  for (int ot = 0; ot < OT; ++ot) {
    const f4 bias = *reinterpret_cast<const f4 *>(b + ot * 16 + 4 * g);
    const bf8 a0 = F[(ot * 2) * 64 + lane], a1 = F[(ot * 2 + 1) * 64 + lane];
#pragma unroll
    for (int n = 0; n < NG; ++n) acc[n][ot] = mfma32(a1, in[n][1], mfma32(a0, in[n][0], bias));
  }
}
template <int NG>
__device__ inline void hidden32(const bf8 *F, const float *b, const bf8 (&in)[NG][2], int lane,
                                bf8 (&out)[NG][2]) {
  f4 acc[NG][4];
  layer32<NG, 4>(F, b, in, lane, acc);
#pragma unroll
  for (int n = 0; n < NG; ++n)
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) out[n][kt] = relu8(pack8(acc[n][2 * kt], acc[n][2 * kt + 1]));
}

constexpr bool is_permutation(bool enc) {
  bool seen[64] = {};
  for (int kp = 0; kp < 64; ++kp) {
    const int f = enc ? perm_enc(kp) : perm_hid(kp);
    if (f < 0 || f > 63 || seen[f]) return false;
    seen[f] = true;
  }
  return true;
}
static_assert(is_permutation(true) && is_permutation(false), "inference input orders");

// Orders the wave's LDS writes before its reads of other lanes' entries (and
// the reads before the next writes).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Rows of q' of one group of 16 points, staged so that each store
// instruction writes 1 KB of whole rows (pitch 36 floats: 16-byte aligned,
// fewer bank conflicts than 32).
constexpr int kStagePitch = 36;
struct EvalLds {
  bf8 I[kFragTotal / 8];
  __attribute__((aligned(16))) float bias[240];
  __attribute__((aligned(16))) float stage[kEvalWaves][16 * kStagePitch];
};

// Inputs of NG groups of 16 points from base.  A point past the end reads
// the last point's inputs, so it computes that point's results bit for bit
// and may store them again (the same values to the same addresses).
// buf[n][7], buf[n][8]: inputs 2g and 2g + 1 of lane group g (input 6
// twice for g = 3), loaded as such rather than selected from buf[n][0..6]
// with a lane-dependent index, which would put the array in scratch memory.
template <int NG>
__device__ inline void eval_load(float (&buf)[NG][9], const float *__restrict__ inputs,
                                 size_t seg0, uint32_t base, uint32_t k, int lane) {
  const int g = lane >> 4;
  const int e0 = 2 * g, e1 = min(2 * g + 1, 6);
#pragma unroll
  for (int n = 0; n < NG; ++n) {
    const float *in = inputs + 7 * (seg0 + min(base + 16 * n + (lane & 15), k - 1));
#pragma unroll
    for (int i = 0; i < 7; ++i) buf[n][i] = in[i];
    buf[n][7] = in[e0];
    buf[n][8] = in[e1];
  }
}

// Lane group g's 16 B-operand values of its point: sin and cos of
// 2^g pi s_i (perm_enc order), inputs 2g and 2g + 1 (the constant 1 for
// g = 3).
template <int NG>
__device__ inline void eval_encode(const float (&s)[NG][9], int lane, bf8 (&x)[NG][2]) {
  const int g = lane >> 4;
  const float rev = 0.5f * (float)(1 << g);  // revolutions per unit input at octave g
#pragma unroll
  for (int n = 0; n < NG; ++n) {
    float v[16];
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      const float a = s[n][i] * rev;
      v[2 * i] = __builtin_amdgcn_sinf(a);
      v[2 * i + 1] = __builtin_amdgcn_cosf(a);
    }
    v[14] = s[n][7];
    v[15] = g == 3 ? 1.0f : s[n][8];
    x[n][0] = pack8(f4{v[0], v[1], v[2], v[3]}, f4{v[4], v[5], v[6], v[7]});
    x[n][1] = pack8(f4{v[8], v[9], v[10], v[11]}, f4{v[12], v[13], v[14], v[15]});
  }
}

// The four layers of NG groups.
template <int NG>
__device__ inline void eval_layers(const EvalLds &S, const bf8 (&x)[NG][2], int lane,
                                   f4 (&out)[NG][3]) {
  // an offset the compiler cannot see through, so that the fragment reads
  // stay in the loop instead of being hoisted into 120 registers
  uint32_t z = 0;
  asm volatile("" : "+s"(z));
  const bf8 *I = S.I + z;
  const float *bias = S.bias + z;
  bf8 h1[NG][2], h2[NG][2];
  hidden32<NG>(I, bias, x, lane, h1);
  hidden32<NG>(I + kFragLayerF / 8, bias + 64, h1, lane, h2);
  hidden32<NG>(I + 2 * kFragLayerF / 8, bias + 128, h2, lane, h1);
  layer32<NG, 3>(I + 3 * kFragLayerF / 8, bias + 192, h1, lane, out);
}

// Each wave runs wave steps of NG groups of 16 points, g0 += stride.  The
// next step's inputs are loaded before this step's results are stored, and
// no branch skips a store (points past the end store the last point's
// results again): vector memory operations complete in order, so the wait
// for those inputs is then a wait for the loads alone, not for the stores.
template <int NG, bool FULL>
__global__ __launch_bounds__(kEvalBlock) void k_mlp_eval(MlpDev W, const float *__restrict__ inputs,
                                                         SegRange R, float *__restrict__ probs,
                                                         float *__restrict__ mix_out,
                                                         float *__restrict__ logits) {
  __shared__ EvalLds S;
  const bf8 *gI = reinterpret_cast<const bf8 *>(W.I);
  for (uint32_t i = threadIdx.x; i < kFragTotal / 8; i += kEvalBlock) S.I[i] = gI[i];
  for (uint32_t i = threadIdx.x; i < 240; i += kEvalBlock) S.bias[i] = W.bias[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, g = lane >> 4;
  const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t r = blockIdx.x % kSegs;
  const uint32_t k = R.count_of(r);
  const size_t seg0 = (size_t)r * R.segcap;
  const uint32_t wave = (blockIdx.x / kSegs) * kEvalWaves + wv;
  const uint32_t nw = (gridDim.x / kSegs) * kEvalWaves;
  const uint32_t stride = nw * 16 * NG;
  uint32_t g0 = wave * 16 * NG;
  if (g0 >= k) return;
  float s[NG][9];
  bf8 x[NG][2];
  eval_load<NG>(s, inputs, seg0, g0, k, lane);
  eval_encode<NG>(s, lane, x);
  for (;;) {
    f4 out[NG][3];
    eval_layers<NG>(S, x, lane, out);
    float q[NG][8], mix[NG];
#pragma unroll
    for (int n = 0; n < NG; ++n) {
      float se, mx;
      proposal_head<FULL>(out[n], lane, W.n_lights, q[n], mix[n], se, mx);
    }
    const uint32_t g1 = g0 + stride;
    eval_load<NG>(s, inputs, seg0, g1, k, lane);
    asm volatile("" ::: "memory");  // the loads above stay before the stores
    float *st = S.stage[wv];
#pragma unroll
    for (int n = 0; n < NG; ++n) {
      const size_t p = seg0 + min(g0 + 16 * n + (lane & 15), k - 1);
      // lane (point p, group g) holds q'_(4g..4g+3) and q'_(16+4g..16+4g+3)
      // of its row; after the exchange, lane l stores 16 bytes of row l / 8
      // (then of row 8 + l / 8) at float 4 (l % 8): 1 KB of rows per store
      float *w = st + kStagePitch * (lane & 15) + 4 * g;
      *reinterpret_cast<f4 *>(w) = f4{q[n][0], q[n][1], q[n][2], q[n][3]};
      *reinterpret_cast<f4 *>(w + 16) = f4{q[n][4], q[n][5], q[n][6], q[n][7]};
      wave_lds_sync();
      const f4 r0 = *reinterpret_cast<const f4 *>(st + kStagePitch * (lane >> 3) + 4 * (lane & 7));
      const f4 r1 =
          *reinterpret_cast<const f4 *>(st + kStagePitch * (8 + (lane >> 3)) + 4 * (lane & 7));
      wave_lds_sync();
      // a row past the end holds the last point's values: stored there again
      const size_t d0 = seg0 + min(g0 + 16 * n + (lane >> 3), k - 1);
      const size_t d1 = seg0 + min(g0 + 16 * n + 8 + (lane >> 3), k - 1);
      // streaming stores: the rows are read by a later kernel, not again here
      __builtin_nontemporal_store(r0, reinterpret_cast<f4 *>(probs + 32 * d0 + 4 * (lane & 7)));
      __builtin_nontemporal_store(r1, reinterpret_cast<f4 *>(probs + 32 * d1 + 4 * (lane & 7)));
      mix_out[p] = mix[n];  // the four lanes of a point hold the same value
      if (logits) {
#pragma unroll
        for (int t = 0; t < 3; ++t)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int o = t * 16 + 4 * g + v;
            if (o < kMlpOut) logits[33 * p + o] = out[n][t][v];
          }
      }
    }
    if (g1 >= k) break;
    g0 = g1;
    eval_encode<NG>(s, lane, x);
  }
}

// Forward + backward deltas for a batch; stores the layer inputs X_l and
// output deltas D_l (bf16, feature-major [feature][kpad]) for the weight
// gradient GEMM.  dlogits: k x 33 (upstream gradient of the loss).
__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(MlpDev W, const float *__restrict__ inputs,
                                                     const float *__restrict__ dlogits, uint32_t k,
                                                     uint32_t kpad, uint16_t *__restrict__ X,
                                                     uint16_t *__restrict__ D) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint32_t nw = gridDim.x * (blockDim.x >> 6);
  uint16_t *X0 = X, *X1 = X + 64 * (size_t)kpad, *X2 = X + 128 * (size_t)kpad,
           *X3 = X + 192 * (size_t)kpad;
  uint16_t *D1 = D, *D2 = D + 64 * (size_t)kpad, *D3 = D + 128 * (size_t)kpad,
           *D4 = D + 192 * (size_t)kpad;
  for (uint32_t g0 = wave * 16; g0 < kpad; g0 += nw * 16) {
    const uint32_t p = g0 + (lane & 15);
    const bool valid = p < k;
    float s[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) s[i] = valid ? inputs[7 * (size_t)p + i] : 0.0f;
    s4 x[4], h1[4], h2[4], h3[4];
    encode_b(s, lane, x);
    hidden(W, 0, x, lane, h1);
    hidden(W, 1, h1, lane, h2);
    hidden(W, 2, h2, lane, h3);
    // store layer inputs (bf16) feature-major
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const size_t o = (size_t)(t * 16 + 4 * g + v) * kpad + p;
        X0[o] = (uint16_t)x[t][v];
        X1[o] = (uint16_t)h1[t][v];
        X2[o] = (uint16_t)h2[t][v];
        X3[o] = (uint16_t)h3[t][v];
      }
    // upstream deltas of the 33 outputs (masked lights and padding: 0)
    s4 d4[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      float dv[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int o = t * 16 + 4 * g + v;
        const bool live = valid && o < kMlpOut && (o == 32 || (uint32_t)o < W.n_lights);
        dv[v] = live ? dlogits[33 * (size_t)p + o] : 0.0f;
      }
      d4[t] = pack4(dv[0], dv[1], dv[2], dv[3]);
#pragma unroll
      for (int v = 0; v < 4; ++v) D4[(size_t)(t * 16 + 4 * g + v) * kpad + p] = (uint16_t)d4[t][v];
    }
    // dH3 = W4 . d4 (K = 48 outputs), ReLU mask of h3
    s4 d3[4], d2[4], d1[4];
    {
      const s4 *G = frag_g(W, 3);
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ot = 0; ot < 3; ++ot) acc = mfma(G[(it * 3 + ot) * 64 + lane], d4[ot], acc);
        float m[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) m[v] = bf2f(h3[it][v]) > 0.0f ? acc[v] : 0.0f;
        d3[it] = pack4(m[0], m[1], m[2], m[3]);
      }
    }
    // dH2 = W3 . d3, dH1 = W2 . d2
    {
      const s4 *G = frag_g(W, 2);
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) acc = mfma(G[(it * 4 + ot) * 64 + lane], d3[ot], acc);
        float m[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) m[v] = bf2f(h2[it][v]) > 0.0f ? acc[v] : 0.0f;
        d2[it] = pack4(m[0], m[1], m[2], m[3]);
      }
    }
    {
      const s4 *G = frag_g(W, 1);
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) acc = mfma(G[(it * 4 + ot) * 64 + lane], d2[ot], acc);
        float m[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) m[v] = bf2f(h1[it][v]) > 0.0f ? acc[v] : 0.0f;
        d1[it] = pack4(m[0], m[1], m[2], m[3]);
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const size_t o = (size_t)(t * 16 + 4 * g + v) * kpad + p;
        D1[o] = (uint16_t)d1[t][v];
        D2[o] = (uint16_t)d2[t][v];
        D3[o] = (uint16_t)d3[t][v];
      }
  }
}

// Weight gradient tiles: dW_l^T (out x in) = D_l (out x pts) . X_l^T (pts x in),
// plus a fifth input tile of ones whose column 0 is the bias gradient.
// One wave per (tile, K chunk of 256 points); partial sums in D layout.
constexpr int kChunk = 256;
constexpr int kTiles = 3 * 4 * 5 + 3 * 5;  // 75

__device__ inline void tile_decode(int tile, int &l, int &mt, int &nt) {
  if (tile < 60) {
    l = tile / 20;
    mt = (tile % 20) / 5;
    nt = tile % 5;
  } else {
    l = 3;
    mt = (tile - 60) / 5;
    nt = (tile - 60) % 5;
  }
}

__global__ __launch_bounds__(64) void k_mlp_wgrad(const uint16_t *__restrict__ X,
                                                  const uint16_t *__restrict__ D, uint32_t kpad,
                                                  uint32_t nchunk, float *__restrict__ partial) {
  const int lane = threadIdx.x;
  const int tile = blockIdx.x / nchunk;
  const uint32_t chunk = blockIdx.x % nchunk;
  int l, mt, nt;
  tile_decode(tile, l, mt, nt);
  const uint16_t *Dl = D + (size_t)(64 * l) * kpad;
  const uint16_t *Xl = X + (size_t)(64 * l) * kpad;
  const int m = mt * 16 + (lane & 15), n = nt * 16 + (lane & 15);
  f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  const s4 ones = (lane & 15) == 0 ? s4{(short)0x3F80, (short)0x3F80, (short)0x3F80, (short)0x3F80}
                                   : s4{0, 0, 0, 0};
  const uint32_t p0 = chunk * kChunk, p1 = min(p0 + kChunk, kpad);
  for (uint32_t pb = p0; pb < p1; pb += 16) {
    const uint32_t pp = pb + 4 * (lane >> 4);
    const s4 a = *reinterpret_cast<const s4 *>(Dl + (size_t)m * kpad + pp);
    const s4 b = nt < 4 ? *reinterpret_cast<const s4 *>(Xl + (size_t)n * kpad + pp) : ones;
    acc = mfma(a, b, acc);
  }
  float *o = partial + ((size_t)tile * nchunk + chunk) * 256 + lane * 4;
  o[0] = acc[0];
  o[1] = acc[1];
  o[2] = acc[2];
  o[3] = acc[3];
}

// Sum partials over chunks (fixed order) into the flat gradient vector.
__global__ void k_mlp_reduce(const float *__restrict__ partial, uint32_t nchunk,
                             float *__restrict__ grad) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // tile * 256 + element
  if (idx >= kTiles * 256) return;
  const int tile = idx >> 8, e = idx & 255;
  const int lane = e >> 2, v = e & 3;
  int l, mt, nt;
  tile_decode(tile, l, mt, nt);
  const int out = mt * 16 + 4 * (lane >> 4) + v;
  const int in = nt * 16 + (lane & 15);
  const int O = l < 3 ? 64 : 33;
  if (out >= O) return;
  if (nt == 4 && in != 64) return;
  float s = 0.0f;
  for (uint32_t c = 0; c < nchunk; ++c) s += partial[((size_t)tile * nchunk + c) * 256 + e];
  const int wo = l * (64 * 64 + 64);
  if (nt < 4) grad[wo + in * O + out] = s;
  else grad[wo + 64 * O + out] = s;
}

__global__ void k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, int n, float lr_t, float b1, float b2,
                       float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

// Rebuild the bf16 fragments and the padded bias vector from the FP32 params.
__global__ void k_mlp_pack(const float *__restrict__ P, uint16_t *__restrict__ F,
                           uint16_t *__restrict__ G, uint16_t *__restrict__ I,
                           float *__restrict__ bias) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < 240) {
    const int l = idx < 192 ? idx / 64 : 3, o = idx < 192 ? idx % 64 : idx - 192;
    const int O = l < 3 ? 64 : 33;
    bias[idx] = o < O ? P[l * (64 * 64 + 64) + 64 * O + o] : 0.0f;
  }
  if (idx >= kFragTotal) return;
  int l, r;
  if (idx < 3 * kFragLayerF) {
    l = idx / kFragLayerF;
    r = idx % kFragLayerF;
  } else {
    l = 3;
    r = idx - 3 * kFragLayerF;
  }
  const int OT = l < 3 ? 4 : 3, O = l < 3 ? 64 : 33;
  const int j = r & 3, lane = (r >> 2) & 63, t = r >> 8;
  const float *W = P + l * (64 * 64 + 64);
  {  // F: t = ot * 4 + kt
    const int ot = t / 4, kt = t % 4;
    const int in = kt * 16 + 4 * (lane >> 4) + j, out = ot * 16 + (lane & 15);
    F[idx] = out < O ? f2bf(W[in * O + out]) : 0;
  }
  {  // G: t = it * OT + ot
    const int it = t / OT, ot = t % OT;
    const int in = it * 16 + (lane & 15), out = ot * 16 + 4 * (lane >> 4) + j;
    G[idx] = out < O ? f2bf(W[in * O + out]) : 0;
  }
  {  // I: ((ot * 2 + kt) * 64 + lane) * 8 + j
    const int j8 = r & 7, lane8 = (r >> 3) & 63, t8 = r >> 9;
    const int ot = t8 / 2, kt = t8 % 2;
    const int kp = kt * 32 + 8 * (lane8 >> 4) + j8;
    const int in = l == 0 ? perm_enc(kp) : perm_hid(kp), out = ot * 16 + (lane8 & 15);
    I[idx] = out < O ? f2bf(W[in * O + out]) : 0;
  }
}

// ------------------------------------------------------------ online training
// One training row per selected record of a finished bounce.  The records
// were drawn with the proposal of their bounce (q'_s, c_s); the gradients are
// taken at the current parameters (q'_c, c_c from a fresh forward pass over
// the batch), with q'_s only in the importance weights, so every Adam step
// sees where the network already is.
//  * light selection, default: KL(p* || q'_c), p*_j proportional to light j's
//    contribution: the sample j ~ q'_s with a = g_l V / p_l gives
//    a grad(-log q'_c,j);  CHK_TRAIN_LIGHTS_CHI2: the second moment
//    sum_j E[v_j^2] / q'_c,j of the NEE estimate, a^2 (q'_s,j / q'_c,j)
//    grad(-log q'_c,j).  grad_z log q'_i = (1 - eps) q_i (e_i - q) / q'_i.
//  * mixing weight c: F(c) = c g_l / p_c(w_l) + (1 - c) g_b / p_c(w_b),
//    p_c = c p_l + (1 - c) p_b, is unbiased for any c (the directions do not
//    depend on c), so 2 F dF/dc at c = c_c on the record's own pair is an
//    unbiased estimate of dE[F^2]/dc.
// Each head's rows are divided by the head's batch-mean magnitude
// (k_train_norm).
struct TrainArgs {
  SegRange R;               // records of the bounce (counts per segment)
  const TrainRec *rec;
  const float *inputs, *probs, *mix;
  const uint8_t *s_occ;     // occlusion per shadow slot (same segment)
  uint32_t n_lights, batch, stride, offset, flags;
};

__device__ inline bool pick_record(const SegRange &R, uint32_t v, size_t &idx) {
  for (uint32_t r = 0; r < kSegs; ++r) {
    const uint32_t c = R.count_of(r);
    if (v < c) {
      idx = (size_t)r * R.segcap + v;
      return true;
    }
    v -= c;
  }
  return false;
}

// Rows of the batch: record index (0xFFFFFFFF: none) and network input.
__global__ __launch_bounds__(256) void k_train_gather(TrainArgs T, float *__restrict__ bin,
                                                      uint32_t *__restrict__ ridx) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= T.batch) return;
  size_t idx;
  const bool ok = pick_record(T.R, b * T.stride + T.offset, idx) && T.rec[idx].light < T.n_lights;
  ridx[b] = ok ? (uint32_t)idx : 0xFFFFFFFFu;
  for (int i = 0; i < 7; ++i) bin[7 * (size_t)b + i] = ok ? T.inputs[7 * idx + i] : 0.0f;
}

__global__ __launch_bounds__(256) void k_train_rows(TrainArgs T, const uint32_t *__restrict__ ridx,
                                                    const float *__restrict__ cur_probs,
                                                    const float *__restrict__ cur_mix,
                                                    float *__restrict__ dl,
                                                    float *__restrict__ wa,
                                                    float *__restrict__ wh) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= T.batch) return;
  float *row = dl + 33 * (size_t)b;
  for (int k = 0; k < 33; ++k) row[k] = 0.0f;
  wa[b] = 0.0f;
  wh[b] = 0.0f;
  const uint32_t idx = ridx[b];
  if (idx == 0xFFFFFFFFu) return;
  const TrainRec R = T.rec[idx];
  const float *qs = T.probs + 32 * (size_t)idx;  // proposal the samples were drawn with
  const float *qc = cur_probs + 32 * (size_t)b;  // current proposal
  const double c = cur_mix[b];
  const double eq = kPropEpsQ, ec = kPropEpsC, invL = 1.0 / T.n_lights;
  const double vis = (R.sh_slot != 0xFFFFFFFFu && !T.s_occ[R.sh_slot]) ? 1.0 : 0.0;
  const double gl = (double)R.gl * vis;
  const double pl = R.pl, pbl = R.pbl, gb = (double)R.fb * R.le, plb = R.plb, pbb = R.pbb;
  // light part: w grad(-log q'_c,j), added as -w (1 - eps) (q_j/q'_j) (e_j - q) at q = q_c
  if (!(T.flags & CHK_TRAIN_FREEZE_LIGHTS)) {
    const double a = pl > 0.0 ? gl / pl : 0.0;
    const uint32_t j = R.light;
    const double qpc = qc[j];
    const double w = (T.flags & CHK_TRAIN_LIGHTS_CHI2) ? a * a * (double)qs[j] / qpc : a;
    if (w > 0.0 && isfinite(w)) {
      const double qj = (qpc - eq * invL) / (1.0 - eq);
      const double s = -w * (1.0 - eq) * qj / qpc;
      for (uint32_t k = 0; k < T.n_lights; ++k) {
        const double qk = ((double)qc[k] - eq * invL) / (1.0 - eq);
        row[k] = (float)(s * ((k == j ? 1.0 : 0.0) - qk));
      }
      wa[b] = (float)w;
    }
  }
  // mixing part: 2 F dF/dc at the current c
  if (!(T.flags & CHK_TRAIN_FREEZE_MIX)) {
    const double pcl = c * pl + (1.0 - c) * pbl, pcb = c * plb + (1.0 - c) * pbb;
    const double al = (gl > 0.0 && pcl > 0.0) ? gl / pcl : 0.0;
    const double ab = (gb > 0.0 && pcb > 0.0) ? gb / pcb : 0.0;
    const double F = c * al + (1.0 - c) * ab;
    double dF = al - ab;
    if (al > 0.0) dF -= c * al * (pl - pbl) / pcl;
    if (ab > 0.0) dF -= (1.0 - c) * ab * (plb - pbb) / pcb;
    const double H = 2.0 * F * dF;
    if (isfinite(H) && H != 0.0) {
      const double sg = (c - ec) / (1.0 - 2.0 * ec);  // sigmoid of the mixing logit
      const double gz = H * (1.0 - 2.0 * ec) * sg * (1.0 - sg);
      row[32] = (float)gz;
      wh[b] = (float)fabs(gz);
    }
  }
}

// Batch means of the two parts' magnitudes, in a fixed summation order.
__global__ __launch_bounds__(1024) void k_train_norm(const float *__restrict__ wa,
                                                     const float *__restrict__ wh, uint32_t batch,
                                                     float *__restrict__ scale) {
  __shared__ double sa[1024], sh[1024], na[1024], nh[1024];
  double a = 0, h = 0, ca = 0, ch = 0;
  for (uint32_t b = threadIdx.x; b < batch; b += 1024) {
    a += wa[b];
    h += wh[b];
    ca += wa[b] > 0.0f ? 1.0 : 0.0;
    ch += wh[b] > 0.0f ? 1.0 : 0.0;
  }
  sa[threadIdx.x] = a;
  sh[threadIdx.x] = h;
  na[threadIdx.x] = ca;
  nh[threadIdx.x] = ch;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sa[threadIdx.x] += sa[threadIdx.x + w];
      sh[threadIdx.x] += sh[threadIdx.x + w];
      na[threadIdx.x] += na[threadIdx.x + w];
      nh[threadIdx.x] += nh[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // rows without a sample count as zero gradient in the mean over the batch
    scale[0] = sa[0] > 0.0 ? (float)((double)batch / sa[0]) : 0.0f;
    scale[1] = sh[0] > 0.0 ? (float)((double)batch / sh[0]) : 0.0f;
  }
}

// Normalise each head by its batch mean magnitude.  No per-row clipping:
// the largest rows are the samples of rarely chosen lights (weight
// (g / (q'_j p_a))^2), and clipping them would bias the light head towards
// the lights it already prefers.
__global__ __launch_bounds__(256) void k_train_scale(float *__restrict__ dl, uint32_t batch,
                                                     uint32_t n_lights,
                                                     const float *__restrict__ scale) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  float *row = dl + 33 * (size_t)b;
  const float sa = scale[0] / (float)batch, sh = scale[1] / (float)batch;
  for (uint32_t k = 0; k < n_lights; ++k) row[k] *= sa;
  row[32] *= sh;
}

}  // namespace chk

// ====================================================================== API
struct chk_proposal_s {
  chk_ctx_s *ctx = nullptr;
  uint32_t n_lights = 0;
  float *params = nullptr;  // FP32 master
  float *adam_m = nullptr, *adam_v = nullptr;
  uint16_t *F = nullptr, *G = nullptr, *I = nullptr;
  float *bias = nullptr;
  uint64_t step = 0;
  // training scratch
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  // online-training batch (rows, deltas, magnitudes, scales, gradient)
  void *train = nullptr;
  size_t train_bytes = 0;
};

namespace chk {

MlpDev proposal_dev(const chk_proposal_s *p) {
  MlpDev W;
  W.F = p->F;
  W.G = p->G;
  W.I = p->I;
  W.bias = p->bias;
  W.n_lights = p->n_lights;
  return W;
}

hipError_t launch_mlp_eval(const chk_proposal_s *p, const float *inputs, const SegRange &R,
                           float *probs, float *mix, float *logits, hipStream_t st) {
  // probs rows are 128 bytes, written as two 16-byte stores per lane
  if (((uintptr_t)probs & 15) != 0) return hipErrorInvalidValue;
  // per segment: wave steps of 16 kEvalGroups points; at most two blocks
  // (16 waves) per CU over the whole grid
  const uint32_t steps = (R.segcap + 16 * kEvalGroups - 1) / (16 * kEvalGroups);
  const uint32_t per_seg = std::max<uint32_t>(
      1u, std::min<uint32_t>((steps + kEvalWaves - 1) / kEvalWaves,
                             std::max<uint32_t>(1u, (uint32_t)p->ctx->n_cu * 2 / kSegs)));
  if (p->n_lights == (uint32_t)kMaxProposalLights)
    hipLaunchKernelGGL((k_mlp_eval<kEvalGroups, true>), dim3(kSegs * per_seg), dim3(kEvalBlock), 0,
                       st, proposal_dev(p), inputs, R, probs, mix, logits);
  else
    hipLaunchKernelGGL((k_mlp_eval<kEvalGroups, false>), dim3(kSegs * per_seg), dim3(kEvalBlock),
                       0, st, proposal_dev(p), inputs, R, probs, mix, logits);
  return hipGetLastError();
}

static hipError_t pack(chk_proposal_s *p, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_pack, dim3((kFragTotal + 255) / 256), dim3(256), 0, st, p->params, p->F,
                     p->G, p->I, p->bias);
  return hipGetLastError();
}

// grad (dev, kMlpParams floats) of the loss whose output gradients are dlogits.
chk_status proposal_grad(chk_proposal_s *p, const float *inputs, const float *dlogits, uint32_t k,
                         float *grad, hipStream_t st) {
  const uint32_t kpad = std::max<uint32_t>(16, (k + 15) / 16 * 16);
  const uint32_t nchunk = (kpad + kChunk - 1) / kChunk;
  const size_t bx = (size_t)256 * kpad * 2, bd = (size_t)240 * kpad * 2;
  const size_t bp = (size_t)kTiles * nchunk * 256 * 4;
  const size_t need = bx + bd + bp + 1024;
  if (p->scratch_bytes < need) {
    if (p->scratch) (void)hipFree(p->scratch);
    p->scratch = nullptr;
    p->scratch_bytes = 0;
    if (hipMalloc(&p->scratch, need) != hipSuccess) {
      set_error("proposal: cannot allocate training scratch");
      return CHK_ENOMEM;
    }
    p->scratch_bytes = need;
  }
  uint16_t *X = (uint16_t *)p->scratch;
  uint16_t *D = (uint16_t *)((char *)p->scratch + bx);
  float *partial = (float *)((char *)p->scratch + ((bx + bd + 255) & ~(size_t)255));
  const uint32_t groups = kpad / 16;
  const uint32_t blocks = std::min<uint32_t>((groups + 3) / 4, (uint32_t)p->ctx->n_cu * 8);
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(blocks), dim3(256), 0, st, proposal_dev(p), inputs, dlogits,
                     k, kpad, X, D);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(kTiles * nchunk), dim3(64), 0, st, X, D, kpad, nchunk,
                     partial);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kTiles * 256 + 255) / 256), dim3(256), 0, st, partial,
                     nchunk, grad);
  CHK_HIP(hipGetLastError());
  return CHK_OK;
}

chk_status proposal_adam(chk_proposal_s *p, const float *grad, float lr, hipStream_t st) {
  p->step += 1;
  const double b1 = 0.9, b2 = 0.999;
  const double t = (double)p->step;
  const float lr_t = (float)(lr * std::sqrt(1.0 - std::pow(b2, t)) / (1.0 - std::pow(b1, t)));
  hipLaunchKernelGGL(k_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, p->params, p->adam_m,
                     p->adam_v, grad, kMlpParams, lr_t, (float)b1, (float)b2, 1e-8f);
  CHK_HIP(hipGetLastError());
  CHK_HIP(pack(p, st));
  return CHK_OK;
}

uint32_t proposal_n_lights(const chk_proposal_s *p) { return p->n_lights; }
const chk_ctx_s *proposal_ctx(const chk_proposal_s *p) { return p->ctx; }

// The loss gradient with respect to the 33 logits of `batch` records of a
// finished bounce (rows b * stride + offset of the records in segment order;
// stride >= 1 spreads the batch over the whole image) at the current
// parameters: *bin_out the batch's network inputs, *dl_out its dlogits
// (batch x 33), both in the proposal's training buffer.
chk_status proposal_train_rows(chk_proposal_s *p, const SegRange &R, const TrainRec *rec,
                               const float *inputs, const float *probs, const float *mix,
                               const uint8_t *s_occ, uint32_t batch, uint32_t stride,
                               uint32_t offset, uint32_t flags, hipStream_t st, float **bin_out,
                               float **dl_out, float **grad_out) {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b_in = al((size_t)batch * 7 * 4), b_dl = al((size_t)batch * 33 * 4);
  const size_t b_w = al((size_t)batch * 4), b_pr = al((size_t)batch * 32 * 4);
  const size_t b_g = al((size_t)kMlpParams * 4);
  const size_t need = b_in + b_dl + 4 * b_w + b_pr + 256 + b_g;
  if (p->train_bytes < need) {
    if (p->train) (void)hipFree(p->train);
    p->train = nullptr;
    p->train_bytes = 0;
    if (hipMalloc(&p->train, need) != hipSuccess) {
      set_error("proposal: cannot allocate the training batch");
      return CHK_ENOMEM;
    }
    p->train_bytes = need;
  }
  char *q = (char *)p->train;
  float *bin = (float *)q;
  q += b_in;
  float *dl = (float *)q;
  q += b_dl;
  float *wa = (float *)q;
  q += b_w;
  float *wh = (float *)q;
  q += b_w;
  uint32_t *ridx = (uint32_t *)q;
  q += b_w;
  float *cmix = (float *)q;
  q += b_w;
  float *cprobs = (float *)q;
  q += b_pr;
  float *scale = (float *)q;
  q += 256;
  float *grad = (float *)q;
  TrainArgs T;
  T.R = R;
  T.rec = rec;
  T.inputs = inputs;
  T.probs = probs;
  T.mix = mix;
  T.s_occ = s_occ;
  T.n_lights = p->n_lights;
  T.batch = batch;
  T.stride = std::max(1u, stride);
  T.offset = offset;
  T.flags = flags;
  const dim3 g((batch + 255) / 256);
  hipLaunchKernelGGL(k_train_gather, g, dim3(256), 0, st, T, bin, ridx);
  CHK_HIP(launch_mlp_eval(p, bin, linear_range(batch), cprobs, cmix, nullptr, st));
  hipLaunchKernelGGL(k_train_rows, g, dim3(256), 0, st, T, ridx, cprobs, cmix, dl, wa, wh);
  hipLaunchKernelGGL(k_train_norm, dim3(1), dim3(1024), 0, st, wa, wh, batch, scale);
  hipLaunchKernelGGL(k_train_scale, g, dim3(256), 0, st, dl, batch, p->n_lights, scale);
  CHK_HIP(hipGetLastError());
  *bin_out = bin;
  *dl_out = dl;
  *grad_out = grad;
  return CHK_OK;
}

// One online training step on `batch` records of a finished bounce.
chk_status proposal_train_step(chk_proposal_s *p, const SegRange &R, const TrainRec *rec,
                               const float *inputs, const float *probs, const float *mix,
                               const uint8_t *s_occ, uint32_t batch, uint32_t stride,
                               uint32_t offset, float lr, uint32_t flags, hipStream_t st) {
  float *bin = nullptr, *dl = nullptr, *grad = nullptr;
  chk_status s = proposal_train_rows(p, R, rec, inputs, probs, mix, s_occ, batch, stride, offset,
                                     flags, st, &bin, &dl, &grad);
  if (s != CHK_OK) return s;
  s = proposal_grad(p, bin, dl, batch, grad, st);
  if (s != CHK_OK) return s;
  return proposal_adam(p, grad, lr, st);
}

}  // namespace chk

using namespace chk;

extern "C" {

chk_status chk_proposal_create(chk_ctx_t ctx, uint32_t n_lights, uint64_t seed,
                               chk_proposal_t *out) {
  if (!ctx || !out) {
    set_error("chk_proposal_create: null argument");
    return CHK_EINVAL;
  }
  *out = nullptr;
  if (n_lights == 0 || n_lights > (uint32_t)kMaxProposalLights) {
    set_error("chk_proposal_create: the proposal supports 1..32 lights");
    return CHK_ECAPACITY;
  }
  CHK_HIP(hipSetDevice(ctx->device));
  chk_proposal_s *p = new chk_proposal_s();
  p->ctx = ctx;
  p->n_lights = n_lights;
  const size_t pb = (size_t)kMlpParams * 4;
  bool ok = hipMalloc(&p->params, pb) == hipSuccess && hipMalloc(&p->adam_m, pb) == hipSuccess &&
            hipMalloc(&p->adam_v, pb) == hipSuccess &&
            hipMalloc(&p->F, kFragTotal * 2) == hipSuccess &&
            hipMalloc(&p->G, kFragTotal * 2) == hipSuccess &&
            hipMalloc(&p->I, kFragTotal * 2) == hipSuccess &&
            hipMalloc(&p->bias, 256 * 4) == hipSuccess;
  if (!ok) {
    chk_proposal_destroy(p);
    set_error("chk_proposal_create: cannot allocate the network");
    return CHK_ENOMEM;
  }
  // He-uniform hidden layers from a seeded splitmix64; zero output layer so
  // an untrained network is the uniform pmf with mixing weight 1/2 (the
  // balance heuristic with one light and one BSDF sample).
  std::vector<float> h(kMlpParams, 0.0f);
  uint64_t st = seed ^ 0x9E3779B97F4A7C15ull;
  auto next = [&]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  const double lim = std::sqrt(6.0 / 64.0);
  for (int l = 0; l < 3; ++l)
    for (int i = 0; i < 64 * 64; ++i) h[l * (64 * 64 + 64) + i] = (float)((2.0 * next() - 1.0) * lim);
  chk_status s = chk_proposal_set_params(p, h.data(), h.size());
  if (s != CHK_OK) {
    chk_proposal_destroy(p);
    return s;
  }
  *out = p;
  return CHK_OK;
}

chk_status chk_proposal_destroy(chk_proposal_t p) {
  if (!p) return CHK_OK;
  (void)hipSetDevice(p->ctx->device);
  for (void *q : {(void *)p->params, (void *)p->adam_m, (void *)p->adam_v, (void *)p->F,
                  (void *)p->G, (void *)p->I, (void *)p->bias, p->scratch, p->train})
    if (q) (void)hipFree(q);
  delete p;
  return CHK_OK;
}

chk_status chk_proposal_set_params(chk_proposal_t p, const float *host, size_t n) {
  if (!p || !host || n != (size_t)kMlpParams) {
    set_error("chk_proposal_set_params: expected 14625 floats");
    return CHK_EINVAL;
  }
  CHK_HIP(hipSetDevice(p->ctx->device));
  CHK_HIP(hipMemcpy(p->params, host, n * 4, hipMemcpyHostToDevice));
  CHK_HIP(hipMemset(p->adam_m, 0, n * 4));
  CHK_HIP(hipMemset(p->adam_v, 0, n * 4));
  p->step = 0;
  CHK_HIP(pack(p, nullptr));
  CHK_HIP(hipDeviceSynchronize());
  return CHK_OK;
}

chk_status chk_proposal_get_params(chk_proposal_t p, float *host, size_t n) {
  if (!p || !host || n != (size_t)kMlpParams) {
    set_error("chk_proposal_get_params: expected 14625 floats");
    return CHK_EINVAL;
  }
  CHK_HIP(hipSetDevice(p->ctx->device));
  CHK_HIP(hipDeviceSynchronize());
  CHK_HIP(hipMemcpy(host, p->params, n * 4, hipMemcpyDeviceToHost));
  return CHK_OK;
}

chk_status chk_proposal_eval(chk_proposal_t p, const float *inputs, uint32_t k, float *probs,
                             float *mix, float *logits, void *stream) {
  if (!p || (k && (!inputs || !probs || !mix))) {
    set_error("chk_proposal_eval: null argument");
    return CHK_EINVAL;
  }
  if (k == 0) return CHK_OK;
  if (((uintptr_t)probs & 15) != 0) {
    set_error("chk_proposal_eval: probs must be 16-byte aligned");
    return CHK_EINVAL;
  }
  CHK_HIP(hipSetDevice(p->ctx->device));
  CHK_HIP(launch_mlp_eval(p, inputs, linear_range(k), probs, mix, logits, (hipStream_t)stream));
  return CHK_OK;
}

chk_status chk_proposal_grad(chk_proposal_t p, const float *inputs, const float *dlogits,
                             uint32_t k, float *grad, void *stream) {
  if (!p || !grad || (k && (!inputs || !dlogits))) {
    set_error("chk_proposal_grad: null argument");
    return CHK_EINVAL;
  }
  CHK_HIP(hipSetDevice(p->ctx->device));
  return proposal_grad(p, inputs, dlogits, k, grad, (hipStream_t)stream);
}

chk_status chk_proposal_loss_grad(chk_proposal_t p, const chk_train_record *rec,
                                  const float *inputs, const float *probs,
                                  const uint8_t *occluded, uint32_t k, uint32_t flags,
                                  float *dlogits, void *stream) {
  CHK_TRY_BEGIN
  static_assert(sizeof(chk_train_record) == sizeof(TrainRec), "record layouts");
  if (!p || (k && (!rec || !inputs || !probs || !occluded || !dlogits))) {
    set_error("chk_proposal_loss_grad: null argument");
    return CHK_EINVAL;
  }
  if (k > (1u << 16) ||
      (flags & ~(CHK_TRAIN_FREEZE_MIX | CHK_TRAIN_FREEZE_LIGHTS | CHK_TRAIN_LIGHTS_CHI2))) {
    set_error("chk_proposal_loss_grad: at most 65536 records and known CHK_TRAIN_* flags");
    return CHK_EINVAL;
  }
  if (k == 0) return CHK_OK;
  CHK_HIP(hipSetDevice(p->ctx->device));
  const hipStream_t st = (hipStream_t)stream;
  float *bin = nullptr, *dl = nullptr, *grad = nullptr;
  const chk_status s = proposal_train_rows(p, linear_range(k), (const TrainRec *)rec, inputs,
                                           probs, nullptr, occluded, k, 1, 0, flags, st, &bin,
                                           &dl, &grad);
  if (s != CHK_OK) return s;
  CHK_HIP(hipMemcpyAsync(dlogits, dl, (size_t)k * 33 * sizeof(float), hipMemcpyDeviceToDevice,
                         st));
  return CHK_OK;
  CHK_TRY_END
}

chk_status chk_proposal_adam(chk_proposal_t p, const float *grad, float lr, void *stream) {
  if (!p || !grad || !(lr > 0.0f)) {
    set_error("chk_proposal_adam: null argument or lr <= 0");
    return CHK_EINVAL;
  }
  CHK_HIP(hipSetDevice(p->ctx->device));
  return proposal_adam(p, grad, lr, (hipStream_t)stream);
}

}  // extern "C"
