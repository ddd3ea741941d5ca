// kd_split.hip -- the Keldysh sum split between a Schur-complement tree over
// the leading vertices, shared by the configurations, and one bordered LU per
// configuration over the rest (orders 6..10).
//
// Eliminating the vertices in the order 0, 1, ..., n-1 without row exchanges
// (as the per-configuration kernel does), the Schur complement after the
// first k vertices depends only on their Keldysh indices sigma_0..sigma_{k-1}
// and is a matrix over the remaining doubled index set:
//     level-k table  T_k[(i,s)][(j,t)],  index 2(i-k)+s for i >= k, border
//                    2(n-k), size D_k = 2(n-k)+1 (T_0 = the Wick matrix)
//     child with sigma_k = s:  T_{k+1}[a][b] = T_k[a+2][b+2]
//                              - T_k[a+2][s] T_k[s][b+2] / T_k[s][s]
// Each of the 2^K tables of level K serves the 2^(n-K) configurations below
// it: configuration sigma takes rows and columns 2(i-K)+sigma_i of its node's
// table and the border, a bordered LU of size n-K+1 finishes it, det A is the
// product of the K path pivots and its own n-K pivots, and the bordered Schur
// complement is the corner left at the end.
//
// One point per workgroup of TPB threads (64 up to order 7, 128 above): the
// workgroup builds the Wick matrix in LDS, computes levels 1..K breadth first
// (one (child, row, column chunk) per thread; the levels alternate between two
// LDS regions), then each thread runs the bordered LUs of its 2^n / TPB
// configurations from its node's table in registers -- at orders 7 and 8 two
// of them, paired on the last vertex's Keldysh index.  Order 8 (K = 3): 2,094
// complex multiply-adds in the tree and 128 x 80 in the paired LUs, 12.3k in
// all against 52.2k for a full bordered LU per configuration.
//
// Pivot check as in the per-configuration kernel: every pivot on a
// configuration's path (the tree's and its own) against PivotTol times
// max(|its original diagonal entry|, the vertex block's largest |Re|+|Im|); a
// flagged configuration is recomputed by partial-pivoting FP64 LU from the
// Wick matrix.
#include "kd_device.hpp"
#include "kd_kernels.hpp"
#include "kd_lu.hpp"
#include "kd_pivoted.hpp"
#include "kd_split_lu.hpp"

namespace kd {

// paired LUs where they fit (SplitShape::PAIR)
#ifndef KD_SPLIT_PAIR
#define KD_SPLIT_PAIR 1
#endif
// threads per workgroup (one point per workgroup) at orders >= 8
#ifndef KD_SPLIT_THREADS
#define KD_SPLIT_THREADS 128
#endif

template <int N> struct SplitShape {
  static constexpr int K = N <= 7 ? 2 : 3;                // tree levels
  static constexpr int M = N - K;                         // vertices per configuration LU
  // threads per workgroup: 2 configurations per thread from order 7 on (paired)
  static constexpr int TPB = N <= 7 ? 64 : (1 << N) < KD_SPLIT_THREADS ? (1 << N) : KD_SPLIT_THREADS;
  static constexpr int NW = TPB / 64;                     // waves per workgroup
  static constexpr int CPT = (1 << N) / TPB;              // configurations per thread
  // paired LUs (split_config_pair) at 2 or more configurations per thread; with
  // more than 5 vertices per LU the pair's registers spill
  static constexpr bool PAIR = KD_SPLIT_PAIR && CPT >= 2 && M >= 2 && M <= 5;
  static constexpr int S0 = wick_stride(N);               // Wick matrix row stride
  static constexpr int Dk(int k) { return 2 * (N - k) + 1; }
  static constexpr int Sk(int k) { return k == 0 ? S0 : wick_stride(N - k); }
  static constexpr int lvl(int k) { return (1 << k) * Dk(k) * Sk(k); } // entries of level k
  // level k (1..K) lives in region k % 2
  static constexpr int region(int r) {
    int s = 1;
    for (int k = 1; k <= K; k++)
      if (k % 2 == r && lvl(k) > s) s = lvl(k);
    return s;
  }
};

// Level k -> k+1: NCH = 2 NPAR children of NPAR = 2^k parent tables (Dp rows at
// stride Sp).  Child m = pm + NPAR s eliminates vertex k with Keldysh index s
// from parent pm.  Work item = (child, row, column chunk): each child row is
// cut into CH chunks so that the level spreads over the workgroup and every
// thread does a few multiply-adds.  The item holding a child's row 0, chunk 0
// also writes its path determinant and pivot flag (heap order: level k node m
// at NPAR + m).  thr[2k+s] is the pivot threshold of that choice.
template <typename T, typename PT, int Dp, int Sp, int Sc, int NPAR, int TPB>
__device__ __forceinline__ void split_level(const PT *__restrict__ par, cx<T> *__restrict__ ch,
                                            double2 *__restrict__ dpath, int *__restrict__ fpath,
                                            const T *__restrict__ thr, int k, int t) {
  constexpr int Dc = Dp - 2, NCH = 2 * NPAR;
  constexpr int CH0 = TPB / (NCH * Dc);
  constexpr int CH = CH0 < 1 ? 1 : CH0 > Dc ? Dc : CH0; // column chunks per row
  constexpr int CW = (Dc + CH - 1) / CH;                // columns per chunk
  for (int r = t; r < NCH * Dc * CH; r += TPB) {
    const int q = r / CH, c0 = (r - q * CH) * CW;
    const int m = q / Dc, a = q - m * Dc;
    const int pm = m & (NPAR - 1), s = m / NPAR;
    const PT *P = par + pm * (Dp * Sp);
    const PT *Pa = P + (a + 2) * Sp + 2 + c0;
    const PT *Ps = P + s * Sp + 2 + c0;
    // every operand of the item loaded before any result is stored: the
    // compiler cannot move these loads above the stores into `ch` by itself,
    // and a load-wait-store round trip per column would serialise the item
    cx<T> pa[CW], ps[CW];
#pragma unroll
    for (int b = 0; b < CW; b++)
      if (Dc % CW == 0 || c0 + b < Dc) {
        pa[b] = ldx<T>(Pa[b]);
        ps[b] = ldx<T>(Ps[b]);
      }
    const cx<T> piv = ldx<T>(P[s * Sp + s]);
    const cx<T> l = cmul<T>(ldx<T>(P[(a + 2) * Sp + s]), cinv<T>(piv));
    cx<T> *C = ch + m * (Dc * Sc) + a * Sc + c0;
#pragma unroll
    for (int b = 0; b < CW; b++)
      if (Dc % CW == 0 || c0 + b < Dc) C[b] = cfms<T>(pa[b], l, ps[b]);
    if (a == 0 && c0 == 0) {
      const double2 d = dpath[NPAR + pm];
      dpath[2 * NPAR + m] = make_double2(d.x * (double)piv.x - d.y * (double)piv.y,
                                         d.x * (double)piv.y + d.y * (double)piv.x);
      fpath[2 * NPAR + m] = fpath[NPAR + pm] | (int)!(cabs1<T>(piv) >= thr[2 * k + s]);
    }
  }
  __syncthreads();
}

#ifndef KD_SPLIT_WAVES_PER_EU
#define KD_SPLIT_WAVES_PER_EU 2
#endif
template <int N, typename T>
__global__ __launch_bounds__(SplitShape<N>::TPB) __attribute__((amdgpu_waves_per_eu(KD_SPLIT_WAVES_PER_EU))) void
integrand_split_kernel(KdParams p, PrepBuf buf) {
  using S = SplitShape<N>;
  constexpr int K = S::K, M = S::M, TPB = S::TPB, NW = S::NW, CPT = S::CPT;
  constexpr int D0 = 2 * N + 1, S0 = S::S0;
  __shared__ double2 W[D0 * S0];
  __shared__ cx<T> R0[S::region(0)], R1[S::region(1)];
  __shared__ double2 dpath[2 << K];
  __shared__ int fpath[2 << K];
  __shared__ T thr[2 * N];
  __shared__ double Us[N];
  __shared__ double2 red[NW];
  __shared__ double redf[NW];
  __shared__ unsigned long long Ssc; // max |Re|+|Im| of the vertex block (bits of a double >= 0)
  extern __shared__ double2 sbins[];

  const int t = threadIdx.x;
  if (p.nbins > 0) {
    for (int i = t; i < p.nbins; i += TPB) sbins[i] = make_double2(0.0, 0.0);
    __syncthreads();
  }
  const int64_t nbatch = (p.n_points + 63) / 64;
  double acc_re = 0.0, acc_im = 0.0, acc_dom = 0.0, acc_fb = 0.0;
  // One compacted point per item, numbered slot-major (see integrand_kernel):
  // item = grp * nbatch + b; a workgroup takes every gridDim.x-th item.  Each
  // wave scans a window of the workgroup's next 64 items at once -- lane l
  // loads the batch count of item l of the window -- and a ballot leaves the
  // live ones (slot below its batch's count), which are then processed in
  // order: a sparse domain costs one vector load per 64 items, not a dependent
  // load per item.  Both waves compute the same window.  The next live item's
  // vertex times and Jacobian are loaded while the current one is computed.
  const int lane = t & 63;
  const int64_t G = gridDim.x;
  int wgrp = nbatch > 0 ? (int)(blockIdx.x / nbatch) : 64; // no points: no items (zero partials)
  int64_t wb = blockIdx.x - (int64_t)wgrp * nbatch;        // the window's first item
  uint64_t mask = 0;                                        // live items of the window not yet done
  int gl = 0, bl = 0;                                       // this lane's item of the window
  double u_n = 0.0, jac_n = 0.0;
  for (;;) {
    if (!mask) {
      if (wgrp >= 64) break;
      const int64_t bw = wb + (int64_t)lane * G;
      gl = wgrp + (int)(bw / nbatch);
      bl = (int)(bw % nbatch);
      mask = __ballot(gl < 64 && gl < buf.cnt[bl]);
      const int64_t adv = wb + 64 * G;
      wgrp += (int)(adv / nbatch);
      wb = adv % nbatch;
      if (!mask) continue;
      const int j1 = __builtin_ctzll(mask);
      const int64_t s1 = (int64_t)__shfl(bl, j1, 64) * 64 + __shfl(gl, j1, 64);
      u_n = buf.u[s1 * N + (t < N ? t : N - 1)];
      jac_n = buf.jac[s1];
    }
    const int j = __builtin_ctzll(mask);
    mask &= mask - 1;
    const int g = __shfl(gl, j, 64);
    const int64_t b = __shfl(bl, j, 64);
    const int64_t slot = b * 64 + g;
    const double uc = u_n, jj = jac_n;
    if (mask) {
      const int j2 = __builtin_ctzll(mask);
      const int64_t s2 = (int64_t)__shfl(bl, j2, 64) * 64 + __shfl(gl, j2, 64);
      // every thread loads, so that no thread keeps an old value: a partial
      // load makes the compiler copy the register and wait for the load at once
      u_n = buf.u[s2 * N + (t < N ? t : N - 1)];
      jac_n = buf.jac[s2];
    }
    // Opaque per-point copy of the thread id: keeps the thread-only index math
    // of the matrix build, the tree levels and the configurations' LDS
    // addresses from being hoisted out of the point loop into registers.
    int tq = t;
    __asm__ volatile("" : "+v"(tq));
    if (tq < N) Us[tq] = uc;
    if (t == 0) {
      Ssc = 0ull;
      dpath[1] = make_double2(1.0, 0.0);
      fpath[1] = 0;
    }
    __syncthreads();
    // Wick matrix, interleaved by vertex (wick_direct.cpp:92-101 picks);
    // wmax: this thread's largest |Re|+|Im| of the vertex block
    double wmax = 0.0;
    for (int q = tq; q < wick_items<N>(); q += TPB) wmax = fmax(wmax, wick_item<N>(p, Us, W, S0, q));
    // block maximum: across the wave by shuffles, then one LDS atomic per wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) wmax = fmax(wmax, __shfl_xor(wmax, off, 64));
    if ((tq & 63) == 0) atomicMax(&Ssc, (unsigned long long)__double_as_longlong(wmax));
    __syncthreads();
    if (tq < 2 * N) {
      const double2 wd = W[tq * S0 + tq];
      const T sc = (T)__longlong_as_double((long long)Ssc);
      thr[tq] = PivotTol<T>::v * fmax(fmax((T)(fabs(wd.x) + fabs(wd.y)), sc), PivotTol<T>::floor);
    }
    __syncthreads();

    // levels 1..K, breadth first
    static_for<0, K>([&](auto KI) {
      constexpr int k = KI;
      cx<T> *dst = (k + 1) % 2 ? R1 : R0;
      if constexpr (k == 0)
        split_level<T, double2, D0, S0, S::Sk(1), 1, TPB>(W, dst, dpath, fpath, thr, 0, tq);
      else
        split_level<T, cx<T>, S::Dk(k), S::Sk(k), S::Sk(k + 1), (1 << k), TPB>(k % 2 ? R1 : R0, dst, dpath, fpath,
                                                                             thr, k, tq);
    });
    const cx<T> *LK = K % 2 ? R1 : R0;

    // The configurations: slot m of this thread.  Single: c = m TPB + t, local
    // bits c mod 2^M, node c >> M.  Paired (2 or more per thread): pair
    // c = (m / 2) TPB + t, local bits 0..M-2 from c mod 2^(M-1), node
    // c >> (M-1), last vertex's index m mod 2.
    constexpr bool PAIR = S::PAIR;
    auto cfg_of = [&](int m) {
      if constexpr (PAIR) {
        const int c = (m >> 1) * TPB + tq;
        const int loc = (c & ((1 << (M - 1)) - 1)) | ((m & 1) << (M - 1));
        return (c >> (M - 1)) | (loc << K);
      } else {
        const int c = m * TPB + tq;
        return (c >> M) | ((c & ((1 << M) - 1)) << K);
      }
    };
    double sre = 0.0, sim = 0.0;
    uint32_t badmask = 0;
    if constexpr (PAIR) {
#pragma unroll 1
      for (int mm = 0; mm < CPT / 2; mm++) {
        const int c = mm * TPB + tq;
        const int loc = c & ((1 << (M - 1)) - 1), node = c >> (M - 1);
        const int cfg0 = node | (loc << K);
        double2 v[2];
        bool bad[2];
        split_config_pair<M, T, cx<T>, false>(LK + node * (S::Dk(K) * S::Sk(K)), loc, thr + 2 * K,
                                              dpath[(1 << K) + node], v, bad);
        const bool fp = fpath[(1 << K) + node] != 0;
        bad[0] = bad[0] || fp;
        bad[1] = bad[1] || fp;
        const double sg = (__popc(cfg0) & 1) ? -1.0 : 1.0; // the index-1 configuration has the other sign
        sre += bad[0] ? 0.0 : sg * v[0].x;
        sim += bad[0] ? 0.0 : sg * v[0].y;
        sre += bad[1] ? 0.0 : -sg * v[1].x;
        sim += bad[1] ? 0.0 : -sg * v[1].y;
        badmask |= ((uint32_t)bad[0] << (2 * mm)) | ((uint32_t)bad[1] << (2 * mm + 1));
      }
    } else {
#pragma unroll 1
      for (int mm = 0; mm < CPT; mm++) {
        const int c = mm * TPB + tq;
        const int loc = c & ((1 << M) - 1), node = c >> M;
        const int cfg = node | (loc << K);
        bool bad;
        const double2 v = split_config<M, T>(LK + node * (S::Dk(K) * S::Sk(K)), loc, thr + 2 * K,
                                             dpath[(1 << K) + node], bad);
        bad = bad || fpath[(1 << K) + node] != 0;
        const double sg = (__popc(cfg) & 1) ? -1.0 : 1.0;
        sre += bad ? 0.0 : sg * v.x;
        sim += bad ? 0.0 : sg * v.y;
        badmask |= (uint32_t)bad << mm;
      }
    }
    // flagged configurations: partial-pivoting FP64 LU, outside the hot loop
    int nfb = 0;
    while (badmask) {
      const int m = __builtin_ctz(badmask);
      badmask &= badmask - 1;
      const int cfg = cfg_of(m);
      const double2 v = config_value_pivoted(W, cfg, N, S0);
      const double sg = (__popc(cfg) & 1) ? -1.0 : 1.0;
      sre += sg * v.x;
      sim += sg * v.y;
      nfb++;
    }
    acc_fb += (double)nfb;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      sre += __shfl_xor(sre, off, 64);
      sim += __shfl_xor(sim, off, 64);
    }
    if ((t & 63) == 0) red[t >> 6] = make_double2(sre, sim);
    __syncthreads();
    if (t == 0) {
      sre = red[0].x;
      sim = red[0].y;
#pragma unroll
      for (int w = 1; w < NW; w++) {
        sre += red[w].x;
        sim += red[w].y;
      }
      if (sqrt(sre * sre + sim * sim) < p.cutoff) {
        sre = 0.0;
        sim = 0.0;
      }
      sre *= jj;
      sim *= jj;
      if (p.mode == MODE_RUN) {
        acc_re += sre;
        acc_im += sim;
        acc_dom += 1.0;
        if (p.nbins > 0) {
          double tmin = Us[0];
#pragma unroll
          for (int i = 1; i < N; i++) tmin = fmin(tmin, Us[i]);
          const int bi = bin_index(p, tmin);
          if (bi >= 0) {
            atomicAdd(&sbins[bi].x, sre);
            atomicAdd(&sbins[bi].y, sim);
          }
        }
      } else {
        const int64_t idx = b * 64 + buf.idx[slot];
        p.out_vals[2 * idx] = sre;
        p.out_vals[2 * idx + 1] = sim;
        if (p.out_dom) p.out_dom[idx] = 1;
        if (p.out_jac) p.out_jac[idx] = jj;
      }
    }
    __syncthreads(); // W, the levels and red are reused by the next point
  }

  if (p.mode == MODE_RUN) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc_fb += __shfl_xor(acc_fb, off, 64);
    if ((t & 63) == 0) redf[t >> 6] = acc_fb;
    __syncthreads();
    if (t == 0) {
      double fb = redf[0];
#pragma unroll
      for (int w = 1; w < NW; w++) fb += redf[w];
      double *o = p.partials + 4 * (int64_t)blockIdx.x;
      o[0] = acc_re;
      o[1] = acc_im;
      o[2] = acc_dom;
      o[3] = fb;
    }
    if (p.nbins > 0) {
      __syncthreads();
      for (int i = t; i < p.nbins; i += TPB) p.bin_part[(int64_t)blockIdx.x * p.nbins + i] = sbins[i];
    }
  }
}

template <int N, typename T>
static hipError_t launch_split_n(const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st) {
  const size_t shm = p.nbins > 0 ? sizeof(double2) * (size_t)p.nbins : 0;
  hipLaunchKernelGGL((integrand_split_kernel<N, T>), dim3(grid), dim3(SplitShape<N>::TPB), shm, st, p, buf);
  return hipGetLastError();
}

int split_threads(int order) {
  switch (order) {
  case 6:
    return SplitShape<6>::TPB;
  case 7:
    return SplitShape<7>::TPB;
  case 8:
    return SplitShape<8>::TPB;
  case 9:
    return SplitShape<9>::TPB;
  case 10:
    return SplitShape<10>::TPB;
  default:
    return 0;
  }
}

bool split_paired(int order) {
  switch (order) {
  case 6:
    return SplitShape<6>::PAIR;
  case 7:
    return SplitShape<7>::PAIR;
  case 8:
    return SplitShape<8>::PAIR;
  case 9:
    return SplitShape<9>::PAIR;
  case 10:
    return SplitShape<10>::PAIR;
  default:
    return false;
  }
}

hipError_t launch_split(int order, int precision, const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st) {
  const bool f32 = lu_is_fp32(order, precision);
#define KD_SPLIT_CASE(n)                                                                                          \
  case n:                                                                                                         \
    return f32 ? launch_split_n<n, float>(p, buf, grid, st) : launch_split_n<n, double>(p, buf, grid, st);
  switch (order) {
    KD_SPLIT_CASE(6)
    KD_SPLIT_CASE(7)
    KD_SPLIT_CASE(8)
    KD_SPLIT_CASE(9)
    KD_SPLIT_CASE(10)
  default:
    return hipErrorInvalidValue;
  }
#undef KD_SPLIT_CASE
}

} // namespace kd
