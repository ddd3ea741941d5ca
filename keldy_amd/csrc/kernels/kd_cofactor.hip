// kd_cofactor.hip -- the kernel method on the device (gfx950).
//
// integrand_g_kernel (impurity_oneband/wick_kernel.cpp:42-138) and
// integrand_g_kernel_single_omega (wick_kernel_single_omega.cpp:41-140): for
// each Keldysh configuration sigma of the n vertices, the first row of the
// comatrix of S1 = [[g(X, X), r], [c, V']] (the external point X first, then
// the vertices; det_expansion.hpp:44-74), times the configuration's parity and
// det S2, one entry per vertex.
//
// Algebra used instead of the reference's LU + triangular solve of S1:
// - the vertex cofactors of S1's first row do not involve that row: with
//   c_i = g(x_i, X), C_{0,i+1} = -det(V'_{i<-c}) = -(adj(V') c)_i (Cramer);
// - V' is the vertex block V with its last diagonal entry set to zero
//   (wick_kernel.cpp:119), and S2 = V (g0 is spin-diagonal with the same
//   table for both spins, and S2 uses the same picks);
// - one unpivoted LU of V carried over the extra column c serves both: the
//   first n-1 elimination steps are common, only the last pivot differs,
//   p' = p - delta (delta = the zeroed entry).  With D = prod_{i<n-1} U_ii,
//   det S2 = D p and adj(V') c = D y', where y' solves U' y' = p' L^-1 c by
//   back substitution scaled by p' -- no division by p', so V' may be
//   singular (it is, at order 1).
// Entry i = parity * det S2 * C_{0,i+1} = -parity * D^2 * p * y'_i.
//
// A configuration whose first n-1 pivots fail the pivot check is recomputed
// from n+1 partial-pivoting FP64 determinants (det V and the n Cramer
// determinants) outside the hot loop.
#include "kd_device.hpp"
#include "kd_kernels.hpp"
#include "kd_lu.hpp"

namespace kd {

// emit(integral_constant<i>, entry) receives the entries as the back
// substitution finalises them (vertex N-1 first); a flagged configuration
// emits zeros (its entries come from the pivoted fallback).
template <int N, typename F>
__device__ __forceinline__ void cofactor_config(const double2 *__restrict__ Wq, int cfg, double sc, bool &bad,
                                                F &&emit) {
  constexpr int DS = wick_stride(N);
  int pr[N], pc[N];
  static_for<0, N>([&](auto I) {
    constexpr int i = I;
    const int pi = 2 * i + ((cfg >> i) & 1);
    pr[i] = pi * DS;
    pc[i] = pi;
  });
  double2 Uh[N][N + 1]; // normalised rows 0..N-2 of U, and L^-1 c in column N
  double2 dprod = cmk<double>(1.0, 0.0);
  double2 pfull = cmk<double>(0.0, 0.0), zlast = cmk<double>(0.0, 0.0), delta = cmk<double>(0.0, 0.0);
  bool flag = false;
  static_for<0, N>([&](auto I) {
    constexpr int i = I;
    double2 row[N + 1];
    static_for<0, N + 1>([&](auto J) {
      constexpr int j = J;
      if constexpr (j < N)
        row[j] = Wq[pr[i] + pc[j]];
      else
        row[j] = Wq[pr[i] + 2 * N];
    });
    const double diag0 = fmax(cabs1<double>(row[i]), sc);
    if constexpr (i == N - 1) delta = row[i];
    static_for<0, i>([&](auto K) {
      constexpr int k = K;
      const double2 l = row[k];
      static_for<k + 1, N + 1>([&](auto J) {
        constexpr int j = J;
        row[j] = cfms<double>(row[j], l, Uh[k][j]);
      });
    });
    if constexpr (i < N - 1) {
      const double2 piv = row[i];
      const double pm = cabs1<double>(piv);
      flag = flag || !(pm >= PivotTol<double>::v * diag0) || pm == 0.0;
      dprod = cmul<double>(dprod, piv);
      const double2 inv = cinv<double>(piv);
      static_for<i + 1, N + 1>([&](auto J) {
        constexpr int j = J;
        Uh[i][j] = cmul<double>(row[j], inv);
      });
    } else {
      pfull = row[N - 1];
      zlast = row[N];
    }
    __builtin_amdgcn_sched_barrier(0);
  });
  const double2 pp = cmk<double>(pfull.x - delta.x, pfull.y - delta.y);
  const double sg = flag ? 0.0 : (__popc(cfg) & 1) ? 1.0 : -1.0; // minus the parity (0: flagged)
  double2 coef = cmul<double>(cmul<double>(dprod, dprod), pfull);
  coef.x *= sg;
  coef.y *= sg;
  double2 yh[N];
  yh[N - 1] = zlast;
  emit(std::integral_constant<int, N - 1>{}, cmul<double>(coef, zlast));
  static_for<0, N - 1>([&](auto I) {
    constexpr int i = N - 2 - (int)I;
    double2 s = cmul<double>(pp, Uh[i][N]);
    static_for<i + 1, N>([&](auto J) {
      constexpr int j = J;
      s = cfms<double>(s, Uh[i][j], yh[j]);
    });
    yh[i] = s;
    emit(std::integral_constant<int, i>{}, cmul<double>(coef, s));
  });
  bad = flag;
}

// Determinant by LU with partial pivoting (max |Re|+|Im|, first on ties); a destroyed.
__device__ __forceinline__ double2 det_pivoted(double2 *a, int m) {
  double2 det = make_double2(1.0, 0.0);
  for (int k = 0; k < m; k++) {
    int p = k;
    double best = cabs1<double>(a[k * m + k]);
    for (int r = k + 1; r < m; r++) {
      const double v = cabs1<double>(a[r * m + k]);
      if (v > best) {
        best = v;
        p = r;
      }
    }
    if (best == 0.0) return make_double2(0.0, 0.0);
    if (p != k) {
      for (int j = 0; j < m; j++) {
        const double2 t = a[k * m + j];
        a[k * m + j] = a[p * m + j];
        a[p * m + j] = t;
      }
      det.x = -det.x;
      det.y = -det.y;
    }
    const double2 piv = a[k * m + k];
    det = cmul<double>(det, piv);
    const double2 inv = cinv<double>(piv);
    for (int r = k + 1; r < m; r++) {
      const double2 l = cmul<double>(a[r * m + k], inv);
      for (int j = k + 1; j < m; j++) a[r * m + j] = cfms<double>(a[r * m + j], l, a[k * m + j]);
    }
  }
  return det;
}

// Fallback for a flagged configuration: entry i = parity * det V * (-det V'_{i<-c}).
__device__ __noinline__ void cofactor_config_pivoted(const double2 *__restrict__ Wq, int cfg, int N, int DS,
                                                    double2 *v) {
  int pk[KD_MAX_ORDER];
  for (int i = 0; i < N; i++) pk[i] = 2 * i + ((cfg >> i) & 1);
  double2 a[KD_MAX_ORDER * KD_MAX_ORDER];
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) a[i * N + j] = Wq[pk[i] * DS + pk[j]];
  const double2 d2 = det_pivoted(a, N);
  const double sg = (__popc(cfg) & 1) ? 1.0 : -1.0;
  for (int c = 0; c < N; c++) {
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++) a[i * N + j] = Wq[pk[i] * DS + (j == c ? 2 * N : pk[j])];
    if (c != N - 1) a[(N - 1) * N + N - 1] = make_double2(0.0, 0.0);
    double2 r = cmul<double>(d2, det_pivoted(a, N));
    r.x *= sg;
    r.y *= sg;
    v[c] = r;
  }
}

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 shfl_xor2(double2 a, int off) {
  return make_double2(__shfl_xor(a.x, off, 64), __shfl_xor(a.y, off, 64));
}

// One 64-thread workgroup (one wave) per work item, persistent over the
// compacted points of the prep stage (same item numbering as the Keldysh-sum
// kernel).  G = min(2^N, 64) lanes per point, 2^N / G configurations per lane
// (configuration = lane + G m): Keldysh indices of vertices 0..log2(G)-1 are
// fixed per lane, those of the higher vertices follow m.
// EV: per-configuration entries to p.out_vals (kernel-method evaluation);
// otherwise per-vertex sums only (runs, single omega, samples).
template <int N, bool EV> __global__ __launch_bounds__(64) void cofactor_kernel(KdParams p, PrepBuf buf) {
  constexpr int NC = 1 << N;
  constexpr int G = NC < 64 ? NC : 64;
  constexpr int LG = N < 6 ? N : 6;
  constexpr int PPW = 64 / G;
  constexpr int CPL = NC / G;
  constexpr int D = 2 * N + 1;
  constexpr int NH = N - LG;
  constexpr int NS = LG + 2 * NH; // per-lane accumulator slots
  constexpr int DS = wick_stride(N);
  __shared__ double2 W[PPW][2 * N * DS];
  __shared__ double2 Acc[64][NS];
  __shared__ double Us[PPW][N];
  __shared__ unsigned long long Ssc[PPW]; // per point: max |Re|+|Im| of the vertex block (bits of a double >= 0)
  extern __shared__ double2 sbins[]; // [nbins][2] when binning (dynamic LDS)

  const bool binned = p.mode == MODE_RUN && p.obs == OBS_KERNEL && p.nbins > 0;
  if (binned) {
    for (int i = threadIdx.x; i < 2 * p.nbins; i += 64) sbins[i] = make_double2(0.0, 0.0);
    __syncthreads();
  }
  const int64_t nbatch = (p.n_points + 63) / 64;
  double acc_re = 0.0, acc_im = 0.0, acc_dom = 0.0, acc_fb = 0.0;
  constexpr int GPB = 64 / PPW;
  const int64_t nitems = nbatch * GPB;
  for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    const int64_t grp = it / nbatch;
    const int64_t b = it - grp * nbatch;
    const int g0i = (int)grp * PPW;
    const int cnt = buf.cnt[b];
    if (g0i >= cnt) continue;
    const int ng = min(PPW, cnt - g0i);
    int lane = threadIdx.x;
    __asm__ volatile("" : "+v"(lane));
    const int64_t slot0 = b * 64 + g0i;
    if (lane < ng * N) Us[lane / N][lane % N] = buf.u[slot0 * N + lane];
    if (lane < PPW) Ssc[lane] = 0ull;
    __syncthreads();
    // vertex rows of the Wick matrix (wick_kernel.cpp:80-94), row/column 2v+k
    // for vertex v at Keldysh index k, then the external point X (split N) in
    // column 2N; row stride wick_stride(N)
    // (wick_item: the vertex pairs, then the external column -- items N^2 + 2N ..)
    constexpr int NI = N * N + 2 * N;
    for (int e = lane; e < ng * NI; e += 64) {
      const int q = e / NI, it = e - q * NI;
      const double wm = wick_item<N>(p, Us[q], W[q], DS, it < N * N ? it : it + 2 * N);
      if (wm > 0.0) atomicMax(&Ssc[q], (unsigned long long)__double_as_longlong(wm));
    }
    __syncthreads();

    const int q = lane / G, lq = lane % G;
    const int64_t slot = slot0 + (q < ng ? q : 0);
    const double jac = buf.jac[slot];
    const int64_t idx = b * 64 + buf.idx[slot];
    // per-lane sums in LDS during the configuration loop (registers are the
    // LU's): slot i for vertex i < LG, LG + 2 (i - LG) + index for the others
    double2 *acc = Acc[lane];
    static_for<0, NS>([&](auto I) { acc[(int)I] = make_double2(0.0, 0.0); });
    int nfb = 0;
    if (q < ng) {
      const double2 *Wq = PPW == 1 ? W[0] : W[q];
      // entry of vertex i (configuration lq + G m) into the per-lane sums
      auto add = [&](auto I, double2 x, int m) {
        constexpr int i = I;
        const int sl = i < LG ? i : LG + 2 * (i - LG) + ((m >> (i - LG)) & 1);
        acc[sl] = cadd(acc[sl], x);
      };
      auto accumulate = [&](const double2 *vals, int m, int cfg) {
        static_for<0, N>([&](auto I) { add(I, vals[(int)I], m); });
        if constexpr (EV) {
          double *o = p.out_vals + 2 * ((idx * NC + cfg) * N);
          static_for<0, N>([&](auto I) {
            constexpr int i = I;
            o[2 * i] = vals[i].x * jac;
            o[2 * i + 1] = vals[i].y * jac;
          });
        }
      };
      const double sc = __longlong_as_double((long long)Ssc[PPW == 1 ? 0 : q]);
      uint64_t badmask = 0;
#pragma unroll 1
      for (int m = 0; m < CPL; m++) {
        const int cfg = lq + G * m;
        bool bad;
        if constexpr (EV) {
          double2 v[N];
          cofactor_config<N>(Wq, cfg, sc, bad, [&](auto I, double2 x) { v[(int)I] = x; });
          if (!bad) accumulate(v, m, cfg);
        } else {
          cofactor_config<N>(Wq, cfg, sc, bad, [&](auto I, double2 x) { add(I, x, m); });
        }
        badmask |= (uint64_t)bad << m;
      }
      while (badmask) {
        const int m = __builtin_ctzll(badmask);
        badmask &= badmask - 1;
        const int cfg = lq + G * m;
        double2 vf[N];
        cofactor_config_pivoted(Wq, cfg, N, DS, vf);
        accumulate(vf, m, cfg);
        nfb++;
      }
    }
    acc_fb += (double)nfb;
    double2 alo[LG];
    double2 ahi[NH > 0 ? NH : 1][2];
    static_for<0, LG>([&](auto I) { alo[I] = acc[(int)I]; });
    static_for<0, NH>([&](auto I) {
      ahi[I][0] = acc[LG + 2 * (int)I];
      ahi[I][1] = acc[LG + 2 * (int)I + 1];
    });

    if (p.obs == OBS_KERNEL_OMEGA) {
      // sum_i (K_i0 + K_i1) exp(i omega t_i) over the point's lanes
      double2 s = make_double2(0.0, 0.0);
      if (q < ng) {
        static_for<0, N>([&](auto I) {
          constexpr int i = I;
          double sn, cs;
          sincos(p.omega * Us[q][i], &sn, &cs);
          const double2 ph = make_double2(cs, sn);
          if constexpr (i < LG)
            s = cadd(s, cmul<double>(alo[i], ph));
          else
            s = cadd(s, cmul<double>(cadd(ahi[i - LG][0], ahi[i - LG][1]), ph));
        });
      }
#pragma unroll
      for (int off = 1; off < G; off <<= 1) s = cadd(s, shfl_xor2(s, off));
      if (lq == 0 && q < ng) {
        s.x *= jac;
        s.y *= jac;
        if (p.mode == MODE_RUN) {
          acc_re += s.x;
          acc_im += s.y;
          acc_dom += 1.0;
        } else {
          p.out_vals[2 * idx] = s.x;
          p.out_vals[2 * idx + 1] = s.y;
          if (p.out_dom) p.out_dom[idx] = 1;
          if (p.out_jac) p.out_jac[idx] = jac;
        }
      }
    } else if (p.mode == MODE_RUN) {
      // total of all entries (the accumulator's re, im)
      double2 tot = make_double2(0.0, 0.0);
      static_for<0, LG>([&](auto I) { tot = cadd(tot, alo[I]); });
      static_for<0, NH>([&](auto I) { tot = cadd(tot, cadd(ahi[I][0], ahi[I][1])); });
#pragma unroll
      for (int off = 1; off < G; off <<= 1) tot = cadd(tot, shfl_xor2(tot, off));
      // per (vertex, Keldysh index): sums over the lanes that share the index
      static_for<0, LG>([&](auto I) {
        constexpr int i = I;
#pragma unroll
        for (int off = 1; off < G; off <<= 1)
          if (off != (1 << i)) alo[i] = cadd(alo[i], shfl_xor2(alo[i], off));
      });
      static_for<0, NH>([&](auto I) {
#pragma unroll
        for (int off = 1; off < G; off <<= 1) {
          ahi[I][0] = cadd(ahi[I][0], shfl_xor2(ahi[I][0], off));
          ahi[I][1] = cadd(ahi[I][1], shfl_xor2(ahi[I][1], off));
        }
      });
      if (q < ng) {
        if (lq == 0) {
          acc_re += tot.x * jac;
          acc_im += tot.y * jac;
          acc_dom += 1.0;
        }
        if (binned) {
          static_for<0, N>([&](auto I) {
            constexpr int i = I;
            const int bi = bin_index(p, Us[q][i]);
            if (bi >= 0) {
              if constexpr (i < LG) {
                if (lq == 0 || lq == (1 << i)) {
                  const int k = (lq >> i) & 1;
                  atomicAdd(&sbins[2 * bi + k].x, alo[i].x * jac);
                  atomicAdd(&sbins[2 * bi + k].y, alo[i].y * jac);
                }
              } else if (lq == 0) {
                atomicAdd(&sbins[2 * bi].x, ahi[i - LG][0].x * jac);
                atomicAdd(&sbins[2 * bi].y, ahi[i - LG][0].y * jac);
                atomicAdd(&sbins[2 * bi + 1].x, ahi[i - LG][1].x * jac);
                atomicAdd(&sbins[2 * bi + 1].y, ahi[i - LG][1].y * jac);
              }
            }
          });
        }
      }
    } else if (lq == 0 && q < ng) {
      if (p.out_dom) p.out_dom[idx] = 1;
      if (p.out_jac) p.out_jac[idx] = jac;
    }
    __syncthreads();
  }

  if (p.mode == MODE_RUN) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      acc_re += __shfl_xor(acc_re, off, 64);
      acc_im += __shfl_xor(acc_im, off, 64);
      acc_dom += __shfl_xor(acc_dom, off, 64);
      acc_fb += __shfl_xor(acc_fb, off, 64);
    }
    if (threadIdx.x == 0) {
      double *o = p.partials + 4 * (int64_t)blockIdx.x;
      o[0] = acc_re;
      o[1] = acc_im;
      o[2] = acc_dom;
      o[3] = acc_fb;
    }
    if (binned) {
      __syncthreads();
      for (int i = threadIdx.x; i < 2 * p.nbins; i += 64)
        p.bin_part[(int64_t)blockIdx.x * 2 * p.nbins + i] = sbins[i];
    }
  }
}

template <int N> static hipError_t launch_cofactor_n(const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st) {
  if constexpr (N < 1) {
    return hipErrorInvalidValue;
  } else {
    const bool binned = p.mode == MODE_RUN && p.obs == OBS_KERNEL && p.nbins > 0;
    const size_t shm = binned ? sizeof(double2) * 2 * (size_t)p.nbins : 0;
    if (p.mode != MODE_RUN && p.obs == OBS_KERNEL)
      hipLaunchKernelGGL((cofactor_kernel<N, true>), dim3(grid), dim3(64), shm, st, p, buf);
    else
      hipLaunchKernelGGL((cofactor_kernel<N, false>), dim3(grid), dim3(64), shm, st, p, buf);
    return hipGetLastError();
  }
}

hipError_t launch_cofactor(int order, const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st) {
  switch (order) {
  case 1: return launch_cofactor_n<1>(p, buf, grid, st);
  case 2: return launch_cofactor_n<2>(p, buf, grid, st);
  case 3: return launch_cofactor_n<3>(p, buf, grid, st);
  case 4: return launch_cofactor_n<4>(p, buf, grid, st);
  case 5: return launch_cofactor_n<5>(p, buf, grid, st);
  case 6: return launch_cofactor_n<6>(p, buf, grid, st);
  case 7: return launch_cofactor_n<7>(p, buf, grid, st);
  case 8: return launch_cofactor_n<8>(p, buf, grid, st);
  case 9: return launch_cofactor_n<9>(p, buf, grid, st);
  case 10: return launch_cofactor_n<10>(p, buf, grid, st);
  case 11: return launch_cofactor_n<11>(p, buf, grid, st);
  case 12: return launch_cofactor_n<12>(p, buf, grid, st);
  default: return hipErrorInvalidValue;
  }
}

} // namespace kd
