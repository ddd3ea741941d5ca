// kd_flow.hip -- the learned coupling-flow warper on gfx950: the map l -> x
// used in front of the warper chain, and its training step (inverse pass +
// backward + Adam), all hand-written.  The network is defined in
// kd_flow_device.hpp.
//
// One 64-lane wave per workgroup owns 64 points (lane = point).  Per coupling
// layer: the input layer (K = nA <= 6) runs on the VALU in FP32; the hidden
// 64x64 product and the output product run on bf16 MFMA
// (v_mfma_f32_16x16x16_bf16, FP32 accumulate) from LDS tiles:
//   A operand: lane l holds A[l % 16][4 (l / 16) + e], e = 0..3
//   B operand: lane l holds B[4 (l / 16) + e][l % 16]
//   D        : lane l holds D[4 (l / 16) + e][l % 16]
// The backward pass uses the D layout of one product directly as the A
// operand of the next (the weight gradients reduce over points, K = points).
// Training needs no stored activations: each coupling layer is inverted
// exactly, so the backward pass recomputes the inputs of layer k from the
// outputs of the inverse pass (the network sees bit-identical inputs).
#include "kd_device.hpp"
#include "kd_flow_device.hpp"
#include "kd_kernels.hpp"

namespace kd {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int FH = KD_FLOW_HIDDEN;
constexpr int FSH = FH + 8; // bf16 row stride of the [64][64] LDS tiles
constexpr int FSO = 20;     // bf16 row stride of the [64][16] LDS tiles
constexpr int FSF = 17;     // float row stride of the output tile
constexpr double kFlowClamp = 1.0 / 1099511627776.0; // 2^-40: keeps logit(l) finite

struct __align__(16) FlowLdsFwd {
  uint16_t H1[64 * FSH];
  uint16_t H2[64 * FSH];
  float O[64 * FSF];
};
struct __align__(16) FlowLdsTrain {
  FlowLdsFwd f;
  uint16_t dZ[64 * FSH];
  uint16_t dO[64 * FSO];
  uint16_t Ab[64 * FSO];
  float Cv[64]; // the points' conditioning values (nc = 1)
};

__device__ __forceinline__ v4f mfma16(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4s pack4(float a, float b, float c, float d) {
  v4s r;
  r[0] = (short)f2bf(a);
  r[1] = (short)f2bf(b);
  r[2] = (short)f2bf(c);
  r[3] = (short)f2bf(d);
  return r;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// log(sigmoid(y) (1 - sigmoid(y)))
__device__ __forceinline__ double log_sig_prime(double y) {
  const double ay = fabs(y);
  return -ay - 2.0 * log1p(exp(-ay));
}

// Network of coupling layer k for the wave's 64 points: a[nA] (and the
// lane's conditioning value cv when fp.nc = 1) -> s[nB], t[nB].
// Leaves h1 (H1) and h2 (H2) of every point in LDS as bf16.
template <int N>
__device__ void flow_net(const FlowParams &fp, int k, const float (&a)[N], float cv, float (&s)[N], float (&t)[N],
                         FlowLdsFwd &L, int lane) {
  const int nA = flow_nA(N, k), nB = flow_nB(N, k);
  const float *W1 = fp.w + flow_layer_offset(N, k, fp.nc);
  const float *b1 = W1 + FH * nA;
  const float *b2 = b1 + FH + FH * FH;
  const float *b3 = b2 + FH + 2 * nB * FH;
  const float *Wc = b3 + 2 * nB; // read only when fp.nc = 1
  const uint16_t *pk = fp.pk + (size_t)k * kFlowPackPerLayer;
  const int r = lane & 15, g = lane >> 4;

  // input layer on the VALU: h1 = relu(W1 a + b1)
#pragma unroll 2
  for (int j0 = 0; j0 < FH; j0 += 4) {
    uint32_t hb[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int j = j0 + e;
      float z = b1[j];
      if (fp.nc) z = fmaf(Wc[j], cv, z);
      for (int i = 0; i < nA; i++) z = fmaf(W1[j * nA + i], a[i], z);
      hb[e] = f2bf(fmaxf(z, 0.f));
    }
    *reinterpret_cast<uint2 *>(&L.H1[lane * FSH + j0]) = make_uint2(hb[0] | (hb[1] << 16), hb[2] | (hb[3] << 16));
  }
  __syncthreads();

  // hidden layer on MFMA: z2[p][j'] = sum_j h1[p][j] W2[j'][j]
  v4f acc[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc[mt][nt] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 4; ks++) {
    v4s A[4], B[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
      A[mt] = *reinterpret_cast<const v4s *>(&L.H1[(16 * mt + r) * FSH + 16 * ks + 4 * g]);
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
      B[nt] = *reinterpret_cast<const v4s *>(&pk[kFlowPackW2 + (16 * nt + r) * FH + 16 * ks + 4 * g]);
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
      for (int nt = 0; nt < 4; nt++) acc[mt][nt] = mfma16(A[mt], B[nt], acc[mt][nt]);
  }
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    const int j = 16 * nt + r;
    const float bj = b2[j];
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
      for (int e = 0; e < 4; e++) L.H2[(16 * mt + 4 * g + e) * FSH + j] = f2bf(fmaxf(acc[mt][nt][e] + bj, 0.f));
  }
  __syncthreads();

  // output layer on MFMA: o[p][q] = sum_j h2[p][j] W3[q][j]  (q < 2 nB <= 16)
  v4f o[4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++) o[mt] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 4; ks++) {
    const v4s B = *reinterpret_cast<const v4s *>(&pk[kFlowPackW3 + r * FH + 16 * ks + 4 * g]);
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const v4s A = *reinterpret_cast<const v4s *>(&L.H2[(16 * mt + r) * FSH + 16 * ks + 4 * g]);
      o[mt] = mfma16(A, B, o[mt]);
    }
  }
  if (r < 2 * nB) {
    const float bq = b3[r];
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
      for (int e = 0; e < 4; e++) L.O[(16 * mt + 4 * g + e) * FSF + r] = o[mt][e] + bq;
  }
  __syncthreads();
  for (int q = 0; q < nB; q++) {
    s[q] = fp.s_scale * tanhf(L.O[lane * FSF + q]);
    t[q] = L.O[lane * FSF + nB + q];
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// The map l -> x (forward direction; the warper's map_reverse) for n points.
// Points come from the Sobol generator (k0+1+i) or from in_l.
template <int N> __global__ __launch_bounds__(64) void flow_map_kernel(FlowParams fp, FlowMapIO io) {
  __shared__ FlowLdsFwd L;
  __shared__ uint32_t sC[N * 31];
  __shared__ uint32_t sS[N];
  const int lane = threadIdx.x;
  if (!io.in_l) {
    for (int i = lane; i < N * 31; i += 64) sC[i] = io.Csr[i];
    if (lane < N) sS[lane] = io.shfr[lane];
  }
  __syncthreads();
  const int64_t nb = (io.n_points + 63) / 64;
  for (int64_t b = blockIdx.x; b < nb; b += gridDim.x) {
    const int64_t pi = b * 64 + lane;
    const bool valid = pi < io.n_points;
    double y[N];
    double ldj = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) {
      double l = 0.5;
      if (valid) {
        l = io.in_l ? io.in_l[pi * N + i] : sobol_coord(sC, sS, i, io.k0 + 1u + (uint32_t)pi);
        if (io.out_l) io.out_l[pi * N + i] = l;
      }
      l = fmin(fmax(l, kFlowClamp), 1.0 - kFlowClamp);
      y[i] = log(l) - log1p(-l);
      ldj -= log(l) + log1p(-l);
    }
    for (int k = 0; k < fp.layers; k++) {
      float a[N], s[N], t[N];
      int na = 0;
#pragma unroll
      for (int i = 0; i < N; i++)
        if (((i + k) & 1) == 0) a[na++] = (float)y[i];
      flow_net<N>(fp, k, a, fp.cond, s, t, L, lane);
      int nb2 = 0;
#pragma unroll
      for (int i = 0; i < N; i++)
        if (((i + k) & 1) == 1) {
          y[i] = y[i] * exp((double)s[nb2]) + (double)t[nb2];
          ldj += (double)s[nb2];
          nb2++;
        }
    }
    if (valid) {
#pragma unroll
      for (int i = 0; i < N; i++) {
        io.out_x[pi * N + i] = 1.0 / (1.0 + exp(-y[i]));
        ldj += log_sig_prime(y[i]);
      }
      io.out_jac[pi] = exp(ldj);
    }
  }
}

// ---------------------------------------------------------------------------
// Backward of coupling layer k (see the header for the loss).  In: per lane
// gB = dL/dv^k_B, vB = v^k_B, s, the per-point loss weight c and a = v^k_A;
// LDS holds h1, h2 of the recomputed network.  Out: da = dL/da and this wave's
// gradient of the layer's parameters in gout.
template <int N>
__device__ void flow_backward(const FlowParams &fp, int k, const float (&a)[N], const float (&s)[N],
                              const double (&vB)[N], const float (&gB)[N], float c, float (&da)[N],
                              FlowLdsTrain &L, float *__restrict__ gout, int lane) {
  const int nA = flow_nA(N, k), nB = flow_nB(N, k);
  const int off = flow_layer_offset(N, k, fp.nc);
  const float *W1 = fp.w + off;
  const int oW1 = off, ob1 = oW1 + FH * nA, oW2 = ob1 + FH, ob2 = oW2 + FH * FH, oW3 = ob2 + FH,
            ob3 = oW3 + 2 * nB * FH, oWc = ob3 + 2 * nB;
  const uint16_t *pk = fp.pk + (size_t)k * kFlowPackPerLayer;
  const int r = lane & 15, g = lane >> 4;

  // dL/do for the s outputs (through S tanh: ds/do = S - s^2 / S) and the t outputs
  float dov[2 * N];
  for (int q = 0; q < nB; q++) {
    const float ds = -gB[q] * (float)vB[q] + c;
    dov[q] = ds * (fp.s_scale - s[q] * s[q] / fp.s_scale);
    dov[nB + q] = -gB[q] * expf(-s[q]);
  }
#pragma unroll
  for (int q = 0; q < 16; q++) {
    L.dO[lane * FSO + q] = q < 2 * nB ? f2bf(dov[q < 2 * N ? q : 0]) : (uint16_t)0;
    L.Ab[lane * FSO + q] = q < nA ? f2bf(a[q < N ? q : 0]) : (uint16_t)0;
  }
  for (int q = 0; q < 2 * nB; q++) {
    const float v = wave_sum(dov[q]);
    if (lane == 0) gout[ob3 + q] = v;
  }
  __syncthreads();

  // dh2 = dO W3 (K = 16), dz2 = dh2 * [h2 > 0]
  v4f acc[4][4];
  {
    v4s A[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) A[mt] = *reinterpret_cast<const v4s *>(&L.dO[(16 * mt + r) * FSO + 4 * g]);
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      const v4s B = *reinterpret_cast<const v4s *>(&pk[kFlowPackW3t + (16 * nt + r) * 16 + 4 * g]);
#pragma unroll
      for (int mt = 0; mt < 4; mt++) acc[mt][nt] = mfma16(A[mt], B, v4f{0.f, 0.f, 0.f, 0.f});
    }
  }
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    float sb = 0.f;
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int p = 16 * mt + 4 * g + e, j = 16 * nt + r;
        const float v = bf2f(L.f.H2[p * FSH + j]) > 0.f ? acc[mt][nt][e] : 0.f;
        acc[mt][nt][e] = v;
        L.dZ[p * FSH + j] = f2bf(v);
        sb += v;
      }
    sb += __shfl_xor(sb, 16, 64);
    sb += __shfl_xor(sb, 32, 64);
    if (g == 0) gout[ob2 + 16 * nt + r] = sb;
  }
  __syncthreads();

  // dW2[j'][j] = sum_p dz2[p][j'] h1[p][j]  (A from the D layout of dz2)
#pragma unroll
  for (int ta = 0; ta < 4; ta++) {
    v4f dw[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) dw[nt] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const v4s A = pack4(acc[mt][ta][0], acc[mt][ta][1], acc[mt][ta][2], acc[mt][ta][3]);
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        v4s B;
#pragma unroll
        for (int e = 0; e < 4; e++) B[e] = (short)L.f.H1[(16 * mt + 4 * g + e) * FSH + 16 * nt + r];
        dw[nt] = mfma16(A, B, dw[nt]);
      }
    }
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int e = 0; e < 4; e++) gout[oW2 + (16 * ta + 4 * g + e) * FH + 16 * nt + r] = dw[nt][e];
  }

  // dh1 = dz2 W2, dz1 = dh1 * [h1 > 0]
  v4f acc1[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc1[mt][nt] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 4; ks++) {
    v4s A[4], B[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
      A[mt] = *reinterpret_cast<const v4s *>(&L.dZ[(16 * mt + r) * FSH + 16 * ks + 4 * g]);
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
      B[nt] = *reinterpret_cast<const v4s *>(&pk[kFlowPackW2t + (16 * nt + r) * FH + 16 * ks + 4 * g]);
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
      for (int nt = 0; nt < 4; nt++) acc1[mt][nt] = mfma16(A[mt], B[nt], acc1[mt][nt]);
  }
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    float sb = 0.f, sc = 0.f;
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int p = 16 * mt + 4 * g + e, j = 16 * nt + r;
        const float v = bf2f(L.f.H1[p * FSH + j]) > 0.f ? acc1[mt][nt][e] : 0.f;
        acc1[mt][nt][e] = v;
        sb += v;
        if (fp.nc) sc = fmaf(v, L.Cv[p], sc); // dWc[j] = sum_p dz1[p][j] c_p
      }
    sb += __shfl_xor(sb, 16, 64);
    sb += __shfl_xor(sb, 32, 64);
    if (g == 0) gout[ob1 + 16 * nt + r] = sb;
    if (fp.nc) {
      sc += __shfl_xor(sc, 16, 64);
      sc += __shfl_xor(sc, 32, 64);
      if (g == 0) gout[oWc + 16 * nt + r] = sc;
    }
  }

  // dW1[j][i] = sum_p dz1[p][j] a[p][i]
#pragma unroll
  for (int ta = 0; ta < 4; ta++) {
    v4f dw = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const v4s A = pack4(acc1[mt][ta][0], acc1[mt][ta][1], acc1[mt][ta][2], acc1[mt][ta][3]);
      v4s B;
#pragma unroll
      for (int e = 0; e < 4; e++) B[e] = (short)L.Ab[(16 * mt + 4 * g + e) * FSO + r];
      dw = mfma16(A, B, dw);
    }
    if (r < nA) {
#pragma unroll
      for (int e = 0; e < 4; e++) gout[oW1 + (16 * ta + 4 * g + e) * nA + r] = dw[e];
    }
  }

  // dW3[q][j] = sum_p dO[p][q] h2[p][j]
  {
    v4f dw[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) dw[nt] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
      v4s A;
#pragma unroll
      for (int e = 0; e < 4; e++) A[e] = (short)L.dO[(16 * ks + 4 * g + e) * FSO + r];
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        v4s B;
#pragma unroll
        for (int e = 0; e < 4; e++) B[e] = (short)L.f.H2[(16 * ks + 4 * g + e) * FSH + 16 * nt + r];
        dw[nt] = mfma16(A, B, dw[nt]);
      }
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int q = 4 * g + e;
      if (q < 2 * nB) {
#pragma unroll
        for (int nt = 0; nt < 4; nt++) gout[oW3 + q * FH + 16 * nt + r] = dw[nt][e];
      }
    }
  }
  __syncthreads();

  // dz1 to LDS (dz2 is no longer read), then da = W1^T dz1 per point
#pragma unroll
  for (int nt = 0; nt < 4; nt++)
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
      for (int e = 0; e < 4; e++) L.dZ[(16 * mt + 4 * g + e) * FSH + 16 * nt + r] = f2bf(acc1[mt][nt][e]);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < N; i++) da[i] = 0.f;
  for (int j = 0; j < FH; j++) {
    const float d = bf2f(L.dZ[lane * FSH + j]);
    for (int i = 0; i < nA; i++) da[i] = fmaf(d, W1[j * nA + i], da[i]);
  }
  __syncthreads();
}

// One training step's gradient for 64 points per workgroup: inverse pass
// x -> l accumulating -log q, then the backward pass layer 0 .. L-1.
// Loss = sum_p c_p (-log q(x_p)), c_p = w_p / (w_mean * batch).
template <int N> __global__ __launch_bounds__(64) void flow_train_kernel(FlowParams fp, FlowTrainIO io) {
  __shared__ FlowLdsTrain L;
  const int lane = threadIdx.x;
  const int64_t pi = io.offset + (int64_t)blockIdx.x * 64 + lane;
  const bool valid = pi < io.n_points;
  float *gout = io.part + (size_t)blockIdx.x * io.n_params;
  double v[N];
  double nlq = 0.0;
  const double c = valid ? io.w[pi] * io.inv_norm : 0.0;
  // the point's conditioning value (nc = 1), also kept in LDS for the Wc gradient
  const float cv = io.c && valid ? (float)io.c[pi] : fp.cond;
  L.Cv[lane] = cv;
#pragma unroll
  for (int i = 0; i < N; i++) {
    double x = valid ? io.x[pi * N + i] : 0.5;
    x = fmin(fmax(x, kFlowClamp), 1.0 - kFlowClamp);
    v[i] = log(x) - log1p(-x);
    nlq += log(x) + log1p(-x);
  }
  for (int k = fp.layers - 1; k >= 0; k--) {
    float a[N], s[N], t[N];
    int na = 0;
#pragma unroll
    for (int i = 0; i < N; i++)
      if (((i + k) & 1) == 0) a[na++] = (float)v[i];
    flow_net<N>(fp, k, a, cv, s, t, L.f, lane);
    int nb = 0;
#pragma unroll
    for (int i = 0; i < N; i++)
      if (((i + k) & 1) == 1) {
        v[i] = (v[i] - (double)t[nb]) * exp(-(double)s[nb]);
        nlq += (double)s[nb];
        nb++;
      }
  }
  float gv[N];
#pragma unroll
  for (int i = 0; i < N; i++) {
    nlq -= log_sig_prime(v[i]);
    gv[i] = (float)(-(1.0 - 2.0 / (1.0 + exp(-v[i]))) * c);
  }
  const double lsum = wave_sum_d(c * nlq);
  if (lane == 0) io.loss_part[blockIdx.x] = lsum;

  for (int k = 0; k < fp.layers; k++) {
    float a[N], s[N], t[N], gB[N], da[N];
    double vB[N];
    int na = 0, nb = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
      if (((i + k) & 1) == 0) {
        a[na++] = (float)v[i];
      } else {
        vB[nb] = v[i];
        gB[nb] = gv[i];
        nb++;
      }
    }
    flow_net<N>(fp, k, a, cv, s, t, L.f, lane);
    flow_backward<N>(fp, k, a, s, vB, gB, (float)c, da, L, gout, lane);
    na = 0;
    nb = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
      if (((i + k) & 1) == 0) {
        gv[i] += da[na++];
      } else {
        gv[i] = gB[nb] * expf(-s[nb]);
        v[i] = vB[nb] * exp((double)s[nb]) + (double)t[nb];
        nb++;
      }
    }
  }
}

// grad[i] = sum over blocks of part[blk][i] (fixed order); loss likewise.
// grad[i] = sum over blocks of part[blk][i]: 64 parameters per workgroup, its
// 4 waves sum the blocks b = w (mod 4) (coalesced rows), then a fixed-order
// combine in LDS -- enough workgroups to cover the chip, deterministic.
__global__ __launch_bounds__(256) void flow_grad_reduce_kernel(const float *__restrict__ part, int nblk,
                                                               int n_params, float *__restrict__ grad,
                                                               const double *__restrict__ loss_part,
                                                               double *__restrict__ loss) {
  __shared__ float sh[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  float s = 0.f;
  if (i < n_params)
    for (int b = w; b < nblk; b += 4) s += part[(size_t)b * n_params + i];
  sh[w][lane] = s;
  __syncthreads();
  if (w == 0 && i < n_params) grad[i] = (sh[0][lane] + sh[1][lane]) + (sh[2][lane] + sh[3][lane]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double l = 0.0;
    for (int b = 0; b < nblk; b++) l += loss_part[b];
    *loss = l;
  }
}

// Adam (bias-corrected), step t >= 1.
__global__ __launch_bounds__(256) void flow_adam_kernel(float *__restrict__ w, float *__restrict__ m,
                                                        float *__restrict__ v, const float *__restrict__ g,
                                                        int n, float lr, float b1, float b2, float eps, float c1,
                                                        float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// bf16 operand copies of W2, W2^T, W3 (padded to 16 rows), W3^T.
__global__ __launch_bounds__(256) void flow_pack_kernel(const float *__restrict__ w, uint16_t *__restrict__ pk,
                                                        int n, int layers, int nc) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= layers * kFlowPackPerLayer) return;
  const int k = idx / kFlowPackPerLayer, e = idx - k * kFlowPackPerLayer;
  const int nA = flow_nA(n, k), nB = flow_nB(n, k);
  const float *W2 = w + flow_layer_offset(n, k, nc) + FH * nA + FH;
  const float *W3 = W2 + FH * FH + FH;
  float val;
  if (e < kFlowPackW2t) {
    val = W2[e];
  } else if (e < kFlowPackW3) {
    const int q = e - kFlowPackW2t, j = q / FH, jp = q % FH; // W2^T[j][j'] = W2[j'][j]
    val = W2[jp * FH + j];
  } else if (e < kFlowPackW3t) {
    const int q = e - kFlowPackW3, o = q / FH, j = q % FH;
    val = o < 2 * nB ? W3[o * FH + j] : 0.f;
  } else {
    const int q = e - kFlowPackW3t, j = q / 16, o = q % 16; // W3^T[j][o]
    val = o < 2 * nB ? W3[o * FH + j] : 0.f;
  }
  pk[idx] = f2bf(val);
}

// w[i] = |value_i|^power and per-block sums of w.
__global__ __launch_bounds__(256) void flow_abs_kernel(const double *__restrict__ vals, int64_t n, double power,
                                                       double *__restrict__ w, double *__restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double m = hypot(vals[2 * i], vals[2 * i + 1]);
    const double a = power == 1.0 ? m : pow(m, power);
    w[i] = a;
    s += a;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// ---------------------------------------------------------------------------
#define KD_FLOW_SWITCH(n, CALL)                                                                           \
  switch (n) {                                                                                            \
  case 1: return CALL(1);                                                                                 \
  case 2: return CALL(2);                                                                                 \
  case 3: return CALL(3);                                                                                 \
  case 4: return CALL(4);                                                                                 \
  case 5: return CALL(5);                                                                                 \
  case 6: return CALL(6);                                                                                 \
  case 7: return CALL(7);                                                                                 \
  case 8: return CALL(8);                                                                                 \
  case 9: return CALL(9);                                                                                 \
  case 10: return CALL(10);                                                                               \
  case 11: return CALL(11);                                                                               \
  case 12: return CALL(12);                                                                               \
  default: return hipErrorInvalidValue;                                                                   \
  }

template <int N> static hipError_t launch_flow_map_n(const FlowParams &fp, const FlowMapIO &io, int grid, hipStream_t st) {
  hipLaunchKernelGGL(flow_map_kernel<N>, dim3(grid), dim3(64), 0, st, fp, io);
  return hipGetLastError();
}
template <int N>
static hipError_t launch_flow_train_n(const FlowParams &fp, const FlowTrainIO &io, int grid, hipStream_t st) {
  hipLaunchKernelGGL(flow_train_kernel<N>, dim3(grid), dim3(64), 0, st, fp, io);
  return hipGetLastError();
}

hipError_t launch_flow_map(int order, const FlowParams &fp, const FlowMapIO &io, int grid, hipStream_t st) {
#define KD_FM(n) launch_flow_map_n<n>(fp, io, grid, st)
  KD_FLOW_SWITCH(order, KD_FM)
#undef KD_FM
}

hipError_t launch_flow_train(int order, const FlowParams &fp, const FlowTrainIO &io, int grid, hipStream_t st) {
#define KD_FT(n) launch_flow_train_n<n>(fp, io, grid, st)
  KD_FLOW_SWITCH(order, KD_FT)
#undef KD_FT
}

hipError_t launch_flow_grad_reduce(const float *part, int nblk, int n_params, float *grad, const double *loss_part,
                                   double *loss, hipStream_t st) {
  hipLaunchKernelGGL(flow_grad_reduce_kernel, dim3((n_params + 63) / 64), dim3(256), 0, st, part, nblk, n_params,
                     grad, loss_part, loss);
  return hipGetLastError();
}

hipError_t launch_flow_adam(float *w, float *m, float *v, const float *g, int n, float lr, float b1, float b2,
                            float eps, int step, hipStream_t st) {
  const float c1 = 1.f / (1.f - powf(b1, (float)step));
  const float c2 = 1.f / (1.f - powf(b2, (float)step));
  hipLaunchKernelGGL(flow_adam_kernel, dim3((n + 255) / 256), dim3(256), 0, st, w, m, v, g, n, lr, b1, b2, eps, c1,
                     c2);
  return hipGetLastError();
}

hipError_t launch_flow_pack(const float *w, uint16_t *pk, int order, int layers, int nc, hipStream_t st) {
  const int tot = layers * kFlowPackPerLayer;
  hipLaunchKernelGGL(flow_pack_kernel, dim3((tot + 255) / 256), dim3(256), 0, st, w, pk, order, layers, nc);
  return hipGetLastError();
}

hipError_t launch_flow_abs(const double *vals, int64_t n, double power, double *w, double *part, int grid,
                           hipStream_t st) {
  hipLaunchKernelGGL(flow_abs_kernel, dim3(grid), dim3(256), 0, st, vals, n, power, w, part);
  return hipGetLastError();
}

} // namespace kd
