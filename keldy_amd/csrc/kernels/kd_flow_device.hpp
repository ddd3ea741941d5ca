// kd_flow_device.hpp -- the learned coupling-flow warper: parameter layout and
// helpers shared by the flow kernels (kd_flow.hip) and the C-ABI.
//
// Map T: l in (0,1)^n -> x in (0,1)^n.  y = logit(l); L affine coupling layers;
// x = sigmoid(y).  Layer k conditions on the dims i with (i + k) even (set A)
// and transforms the others (set B):
//     a = y_A;  h1 = relu(W1 a + b1);  h2 = relu(W2 h1 + b2);  o = W3 h2 + b3
//     s = S tanh(o[0:nB]),  t = o[nB:2nB],   y_B <- y_B * exp(s) + t
// Hidden width KD_FLOW_HIDDEN.  With one conditioning input c (nc = 1: the
// time-conditioned flow, c a scaled external time) every layer's input layer
// is h1 = relu(W1 a + Wc c + b1): one map T_c per value of c, each a
// bijection with its own exact Jacobian.
//
// Tails.  Near a face of the cube an affine map in logit space is a power law:
// with a = exp(sum of the s a coordinate receives), x ~ 1 - (1-l)^a, so the
// density q(x) ~ (1-x)^(1/a - 1) and the estimator's variance integral
// int g^2 / q diverges wherever g does not vanish on the face once a < 1/2.
// The scale bound S = 0.9 ln2 / ceil(L/2) keeps a > 2^-0.9 (each coordinate is
// transformed by ceil(L/2) layers at most): the variance stays finite.  The W2 and W3 products run on bf16 MFMA with
// FP32 accumulation; coordinates and the Jacobian
//     log dx/dl = -sum log(l (1-l)) + sum_k sum s + sum log(x (1-x))
// are FP64.  The map is a bijection with an exact triangular Jacobian whatever
// precision the network runs in (s and t depend on y_A only), so the warped
// estimator stays unbiased.
#pragma once

#include <stdint.h>

// shared by hipcc (kernels, C-ABI) and the host compiler (host warper class)
#if defined(__HIPCC__)
#define KD_HD __host__ __device__
#else
#define KD_HD
#endif

#define KD_FLOW_HIDDEN 64
#define KD_FLOW_MAX_LAYERS 16

namespace kd {

// Parameter layout of layer k (floats), nA = |A_k|, nB = |B_k|:
//   W1 [H][nA], b1 [H], W2 [H][H], b2 [H], W3 [2 nB][H], b3 [2 nB], and with nc = 1 Wc [H]
KD_HD inline int flow_nA(int n, int k) { return (n + 1 - (k & 1)) / 2; }
KD_HD inline int flow_nB(int n, int k) { return n - flow_nA(n, k); }
KD_HD inline int flow_layer_size(int n, int k, int nc = 0) {
  const int H = KD_FLOW_HIDDEN, nA = flow_nA(n, k), nB = flow_nB(n, k);
  return H * nA + H + H * H + H + 2 * nB * H + 2 * nB + nc * H;
}
KD_HD inline int flow_layer_offset(int n, int k, int nc = 0) {
  int off = 0;
  for (int q = 0; q < k; q++) off += flow_layer_size(n, q, nc);
  return off;
}
KD_HD inline int flow_n_params(int n, int layers, int nc = 0) { return flow_layer_offset(n, layers, nc); }
// bound S of the per-layer log-scale s = S tanh(o) (see "Tails" above)
KD_HD inline float flow_s_scale(int layers) { return 0.62383246f / (float)((layers + 1) / 2); } // 0.9 ln 2

// bf16 operand copies, per layer: W2 [H][H] (out, in), W2^T [H][H] (in, out),
// W3 [16][H] (rows >= 2 nB zero), W3^T [H][16].
constexpr int kFlowPackW2 = 0;
constexpr int kFlowPackW2t = KD_FLOW_HIDDEN * KD_FLOW_HIDDEN;
constexpr int kFlowPackW3 = 2 * KD_FLOW_HIDDEN * KD_FLOW_HIDDEN;
constexpr int kFlowPackW3t = kFlowPackW3 + 16 * KD_FLOW_HIDDEN;
constexpr int kFlowPackPerLayer = kFlowPackW3t + 16 * KD_FLOW_HIDDEN;

struct FlowParams {
  const float *w;     // FP32 parameters (layout above)
  const uint16_t *pk; // bf16 operand copies, kFlowPackPerLayer per layer
  int layers;
  float s_scale;      // flow_s_scale(layers)
  int nc;             // conditioning inputs: 0 or 1
  float cond;         // the conditioning value of the map (nc = 1)
};

// float -> bfloat16 bits, round to nearest even (finite inputs)
KD_HD inline uint16_t f2bf(float f) {
  union {
    float f;
    uint32_t u;
  } v{f};
  uint32_t u = v.u;
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
KD_HD inline float bf2f(uint16_t b) {
  union {
    uint32_t u;
    float f;
  } v{(uint32_t)b << 16};
  return v.f;
}

} // namespace kd
