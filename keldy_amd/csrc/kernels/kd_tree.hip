// kd_tree.hip -- the Keldysh sum by a Schur-complement tree (orders 6..12).
//
// The 2^n configurations sigma of one point share their leading vertex
// choices.  Eliminating the vertices in the order 0, 1, ..., n-1 (Gaussian
// elimination without row exchanges, as the per-configuration kernel does),
// the Schur complement after the first k vertices depends only on sigma_0..
// sigma_{k-1}, and it is a matrix over the REMAINING doubled index set: rows
// and columns (i, s) for vertices i >= k and both Keldysh indices s, plus the
// external border.  So one table per tree node serves every configuration
// below it:
//     level-k table  T_k[(i,s)][(j,t)],  index 2(i-k)+s, border 2(n-k),
//                    size D_k = 2(n-k)+1 (T_0 = the Wick matrix, corner 0)
//     child with sigma_k = s:  T_{k+1}[a][b] = T_k[a+2][b+2]
//                                - T_k[a+2][s] * T_k[s][b+2] / T_k[s][s]
// and det A_sigma = the product of the pivots T_k[s][s] along the path, while
// the bordered Schur complement -c A^{-1} b is the corner left at the end.
//
// Work per point at order 8: ~8.5k complex multiply-adds against 52k for one
// bordered LU per configuration.  Layout of the traversal for one point, one
// 64-lane wave:
//   - levels 1 .. A = n-6: depth first, one node per level kept in LDS
//     (entry-parallel over the wave); the 2^A nodes of level A are the roots
//     of the slabs, visited in an order that recomputes only the levels whose
//     choice changed;
//   - per slab, levels A+1 .. A+3 breadth first (2, 4, 8 nodes of sizes
//     11^2, 9^2, 7^2, entry-parallel);
//   - bottom: each lane takes one of the 8 level-(n-3) nodes and one of the 8
//     choices of the last three vertices, and finishes with a 4x4 bordered LU
//     in registers.
// Pivots are checked as in the per-configuration kernel (|pivot| against the
// original diagonal entry); a configuration with a flagged pivot anywhere on
// its path is recomputed by partial-pivoting FP64 LU from the Wick matrix.
#include "kd_device.hpp"
#include "kd_kernels.hpp"
#include "kd_lu.hpp"

namespace kd {

template <int N> struct TreeShape {
  static constexpr int A = N - 6; // depth-first levels 1..A
  static constexpr int K = N - 3; // last slab level; the bottom eliminates vertices K..N-1
  static constexpr int D0 = 2 * N + 1;
  static constexpr int Dk(int k) { return 2 * (N - k) + 1; }
  static constexpr int off(int k) { // offset of level k (1..A) in the depth-first stack
    int s = 0;
    for (int j = 1; j < k; j++) s += Dk(j) * Dk(j);
    return s;
  }
  static constexpr int stack() { return off(A + 1) > 0 ? off(A + 1) : 1; }
};

template <typename T> __device__ __forceinline__ cx<T> ldc(const double2 &v) { return cmk<T>((T)v.x, (T)v.y); }
template <typename T> __device__ __forceinline__ cx<T> ldc(const float2 &v) { return cmk<T>((T)v.x, (T)v.y); }

// tabs: PivotTol times the largest |Re|+|Im| of the point's vertex block
template <typename T> __device__ __forceinline__ bool bad_pivot(cx<T> piv, double2 wdiag, T tabs) {
  const T pm = cabs1<T>(piv);
  return !(pm >= PivotTol<T>::v * (T)(fabs(wdiag.x) + fabs(wdiag.y))) || !(pm >= tabs) || pm == T(0);
}

// NCH children from NPAR parent tables (Dp x Dp each, contiguous), one child
// row per lane: child m eliminates vertex choice s = m / NPAR of parent
// m % NPAR (depth first: NCH = NPAR = 1 and s = s_dfs).  The lane holding
// row 0 of a child also writes the child's path determinant (parent's times
// the pivot) and pivot flag; wbase = 2k locates the pivot's original diagonal
// entry in the level-0 matrix W (D0 x D0).
template <typename T, typename PT, int Dp, int NPAR, int NCH, int D0>
__device__ __forceinline__ void tree_rows(const PT *__restrict__ par, cx<T> *__restrict__ ch,
                                          const double2 *__restrict__ dpar, const int *__restrict__ fpar,
                                          double2 *__restrict__ dch, int *__restrict__ fch,
                                          const double2 *__restrict__ W, int wbase, int s_dfs, int lane, T tabs) {
  constexpr int Dc = Dp - 2;
  for (int r = lane; r < NCH * Dc; r += 64) {
    const int m = r / Dc, a = r - m * Dc;
    const int s = NCH == 1 ? s_dfs : m / NPAR;
    const int pm = NCH == 1 ? 0 : m % NPAR;
    const PT *P = par + pm * (Dp * Dp);
    const cx<T> piv = ldc<T>(P[s * Dp + s]);
    const cx<T> l = cmul<T>(ldc<T>(P[(a + 2) * Dp + s]), cinv<T>(piv));
    const PT *Pa = P + (a + 2) * Dp + 2;
    const PT *Ps = P + s * Dp + 2;
    cx<T> *C = ch + m * (Dc * Dc) + a * Dc;
#pragma unroll
    for (int b = 0; b < Dc; b++) C[b] = cfms<T>(ldc<T>(Pa[b]), l, ldc<T>(Ps[b]));
    if (a == 0) {
      const double2 d = dpar[pm];
      dch[m] = make_double2(d.x * (double)piv.x - d.y * (double)piv.y, d.x * (double)piv.y + d.y * (double)piv.x);
      const int wi = wbase + s;
      fch[m] = fpar[pm] | (int)bad_pivot<T>(piv, W[wi * D0 + wi], tabs);
    }
  }
  __syncthreads();
}

// Partial-pivoting FP64 LU of one configuration from the level-0 (interleaved)
// Wick matrix: vertex i with Keldysh index s at 2i+s, border 2N.  The LAPACK
// zgetf2 rule (max |Re|+|Im|, first on ties); the permutation sign cancels in
// det(A)^2.
__device__ __noinline__ double2 tree_config_pivoted(const double2 *__restrict__ W, uint32_t cfg, int N) {
  const int D = 2 * N + 1, M = N + 1;
  double2 a[(KD_MAX_ORDER + 1) * (KD_MAX_ORDER + 1)];
  int pk[KD_MAX_ORDER + 1];
  for (int i = 0; i < N; i++) pk[i] = 2 * i + ((cfg >> i) & 1);
  pk[N] = 2 * N;
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) a[i * M + j] = W[pk[i] * D + pk[j]];
  a[N * M + N] = make_double2(0.0, 0.0);
  double2 det = make_double2(1.0, 0.0);
  for (int k = 0; k < N; k++) {
    int p = k;
    double best = cabs1<double>(a[k * M + k]);
    for (int r = k + 1; r < N; r++) {
      const double v = cabs1<double>(a[r * M + k]);
      if (v > best) {
        best = v;
        p = r;
      }
    }
    if (best == 0.0) return make_double2(0.0, 0.0);
    if (p != k)
      for (int j = 0; j < M; j++) {
        const double2 t = a[k * M + j];
        a[k * M + j] = a[p * M + j];
        a[p * M + j] = t;
      }
    const double2 piv = a[k * M + k];
    det = cmul<double>(det, piv);
    const double2 inv = cinv<double>(piv);
    for (int r = k + 1; r < M; r++) {
      const double2 l = cmul<double>(a[r * M + k], inv);
      for (int j = k + 1; j < M; j++) a[r * M + j] = cfms<double>(a[r * M + j], l, a[k * M + j]);
    }
  }
  return cmul<double>(cmul<double>(det, det), a[N * M + N]);
}

template <int N, typename T>
__global__ __launch_bounds__(64) void integrand_tree_kernel(KdParams p, PrepBuf buf) {
  using S = TreeShape<N>;
  constexpr int A = S::A, K = S::K, D0 = S::D0;
  constexpr int DA = S::Dk(A); // 13: the slab roots
  __shared__ double2 W[D0 * D0];
  __shared__ cx<T> stk[S::stack()];
  __shared__ cx<T> X[8 * 49], Y[4 * 81];
  __shared__ double2 dd[A + 1], dn[4][8]; // path determinants: depth-first levels, slab levels
  __shared__ int ff[A + 1], fl[4][8];     // pivot flags, same layout
  __shared__ double Us[N];
  __shared__ unsigned long long Ssc; // max |Re|+|Im| of the vertex block (bits of a double >= 0)
  extern __shared__ double2 sbins[];

  if (p.nbins > 0) {
    for (int i = threadIdx.x; i < p.nbins; i += 64) sbins[i] = make_double2(0.0, 0.0);
    __syncthreads();
  }
  const int64_t nbatch = (p.n_points + 63) / 64;
  double acc_re = 0.0, acc_im = 0.0, acc_dom = 0.0, acc_fb = 0.0;
  // one compacted point per item, numbered slot-major (see integrand_kernel)
  const int64_t nitems = nbatch * 64;
  for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    const int64_t grp = it / nbatch;
    const int64_t b = it - grp * nbatch;
    const int g0i = (int)grp;
    if (g0i >= buf.cnt[b]) continue;
    int lane = threadIdx.x;
    __asm__ volatile("" : "+v"(lane));
    const int64_t slot = b * 64 + g0i;
    if (lane < N) Us[lane] = buf.u[slot * N + lane];
    if (lane == 0) Ssc = 0ull;
    __syncthreads();
    // Wick matrix in level-0 order (wick_direct.cpp:92-101 picks, interleaved),
    // one vertex pair or external entry per item (wick_item)
    for (int q = lane; q < wick_items<N>(); q += 64) {
      const double wm = wick_item<N>(p, Us, W, D0, q);
      if (wm > 0.0) atomicMax(&Ssc, (unsigned long long)__double_as_longlong(wm));
    }
    __syncthreads();
    const T tabs = (T)(PivotTol<T>::v * __longlong_as_double((long long)Ssc));

    double sre = 0.0, sim = 0.0;
    int nfb = 0;
    if (lane == 0) {
      dd[0] = make_double2(1.0, 0.0);
      ff[0] = 0;
    }
    const int m = lane & 7, bb3 = lane >> 3;
    for (int r = 0; r < (1 << A); r++) {
      // depth first: recompute the levels below the highest changed choice
      const int k0 = r == 0 ? 0 : A - 1 - (31 - __clz((unsigned)(r ^ (r - 1))));
      static_for<0, A>([&](auto KI) {
        constexpr int k = KI;
        if (k >= k0) {
          const int s = (r >> (A - 1 - k)) & 1;
          if constexpr (k == 0)
            tree_rows<T, double2, S::Dk(0), 1, 1, D0>(W, stk + S::off(1), dd, ff, dd + 1, ff + 1, W, 0, s, lane, tabs);
          else
            tree_rows<T, cx<T>, S::Dk(k), 1, 1, D0>(stk + S::off(k), stk + S::off(k + 1), dd + k, ff + k, dd + k + 1,
                                                   ff + k + 1, W, 2 * k, s, lane, tabs);
        }
      });
      // slab: levels A+1 (2 nodes), A+2 (4), A+3 = K (8), breadth first
      if constexpr (A == 0)
        tree_rows<T, double2, DA, 1, 2, D0>(W, X, dd, ff, dn[1], fl[1], W, 0, 0, lane, tabs);
      else
        tree_rows<T, cx<T>, DA, 1, 2, D0>(stk + S::off(A), X, dd + A, ff + A, dn[1], fl[1], W, 2 * A, 0, lane, tabs);
      tree_rows<T, cx<T>, 11, 2, 4, D0>(X, Y, dn[1], fl[1], dn[2], fl[2], W, 2 * (A + 1), 0, lane, tabs);
      tree_rows<T, cx<T>, 9, 4, 8, D0>(Y, X, dn[2], fl[2], dn[3], fl[3], W, 2 * (A + 2), 0, lane, tabs);
      const double2 dK = dn[3][m];
      const int fK = fl[3][m];

      // bottom: node m, choices bb3 of vertices K, K+1, K+2: 4x4 bordered LU
      {
        const cx<T> *Sm = X + m * 49;
        int pk[4];
        pk[0] = (bb3 & 1);
        pk[1] = 2 + ((bb3 >> 1) & 1);
        pk[2] = 4 + ((bb3 >> 2) & 1);
        pk[3] = 6;
        cx<T> Uh[4][4];
        cx<T> det = cmk<T>(T(1), T(0));
        cx<T> sc = cmk<T>(T(0), T(0));
        bool flag = fK != 0;
        static_for<0, 4>([&](auto I) {
          constexpr int i = I;
          cx<T> row[4];
          static_for<0, 4>([&](auto J) {
            constexpr int j = J;
            row[j] = Sm[pk[i] * 7 + pk[j]];
          });
          static_for<0, (i < 3 ? i : 3)>([&](auto KB) {
            constexpr int kb = KB;
            const cx<T> l = row[kb];
            static_for<kb + 1, 4>([&](auto J) {
              constexpr int j = J;
              row[j] = cfms<T>(row[j], l, Uh[kb][j]);
            });
          });
          if constexpr (i < 3) {
            const cx<T> piv = row[i];
            const int wi = 2 * (K + i) + ((bb3 >> i) & 1);
            flag = flag || bad_pivot<T>(piv, W[wi * D0 + wi], tabs);
            det = cmul<T>(det, piv);
            const cx<T> inv = cinv<T>(piv);
            static_for<i + 1, 4>([&](auto J) {
              constexpr int j = J;
              Uh[i][j] = cmul<T>(row[j], inv);
            });
          } else {
            sc = row[3];
          }
        });
        // det A = (path pivots) * (bottom pivots); value = det^2 * s
        const double2 dt = make_double2(dK.x * (double)det.x - dK.y * (double)det.y,
                                        dK.x * (double)det.y + dK.y * (double)det.x);
        double2 v = cmul<double>(cmul<double>(dt, dt), make_double2((double)sc.x, (double)sc.y));
        // configuration bits: sigma_0..sigma_{A-1} from r (sigma_{A-1} lowest),
        // sigma_A..A+2 = bits of m, sigma_K..K+2 = bits of bb3
        if (flag) {
          uint32_t cfg = ((uint32_t)m << A) | ((uint32_t)bb3 << K);
          for (int k = 0; k < A; k++) cfg |= (uint32_t)((r >> (A - 1 - k)) & 1) << k;
          v = tree_config_pivoted(W, cfg, N);
          nfb++;
        }
        const double sg = ((__popc((unsigned)r) + __popc((unsigned)m) + __popc((unsigned)bb3)) & 1) ? -1.0 : 1.0;
        sre += sg * v.x;
        sim += sg * v.y;
      }
      __syncthreads(); // X, Y, dn reused by the next slab
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      sre += __shfl_xor(sre, off, 64);
      sim += __shfl_xor(sim, off, 64);
    }
    acc_fb += (double)nfb;
    if (lane == 0) {
      if (sqrt(sre * sre + sim * sim) < p.cutoff) {
        sre = 0.0;
        sim = 0.0;
      }
      const double jj = buf.jac[slot];
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
    if (p.nbins > 0) {
      __syncthreads();
      for (int i = threadIdx.x; i < p.nbins; i += 64) p.bin_part[(int64_t)blockIdx.x * p.nbins + i] = sbins[i];
    }
  }
}

template <int N, typename T>
static hipError_t launch_tree_n(const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st) {
  const size_t shm = p.nbins > 0 ? sizeof(double2) * (size_t)p.nbins : 0;
  hipLaunchKernelGGL((integrand_tree_kernel<N, T>), dim3(grid), dim3(64), shm, st, p, buf);
  return hipGetLastError();
}

hipError_t launch_tree(int order, int precision, const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st) {
  const bool f32 = lu_is_fp32(order, precision);
#define KD_TREE_CASE(n)                                                                                           \
  case n:                                                                                                         \
    return f32 ? launch_tree_n<n, float>(p, buf, grid, st) : launch_tree_n<n, double>(p, buf, grid, st);
  switch (order) {
    KD_TREE_CASE(6)
    KD_TREE_CASE(7)
    KD_TREE_CASE(8)
    KD_TREE_CASE(9)
    KD_TREE_CASE(10)
    KD_TREE_CASE(11)
    KD_TREE_CASE(12)
  default:
    return hipErrorInvalidValue;
  }
#undef KD_TREE_CASE
}

} // namespace kd
