// kd_split_lu.hpp -- the per-configuration bordered LUs of the split Keldysh
// sum (kd_split.hip, kd_wsplit.hip): after the leading vertices have been
// eliminated once per node of the tree of their Keldysh choices, each
// configuration finishes with a bordered LU of its remaining vertices taken
// from its node's table.
#pragma once

#include "kd_device.hpp"
#include "kd_lu.hpp"

namespace kd {

template <typename T> __device__ __forceinline__ cx<T> ldx(const double2 &v) { return cmk<T>((T)v.x, (T)v.y); }
template <typename T> __device__ __forceinline__ cx<T> ldx(const float2 &v) { return cmk<T>((T)v.x, (T)v.y); }

// det(A)^2 s of one configuration from its level-K node table Tn (2M+1 rows at
// stride wick_stride(M)): loc holds sigma_K..sigma_{n-1} in bits 0..M-1; thr
// the pivot thresholds from index 2K on; dp the node's path determinant.
template <int M, typename T, typename PT>
__device__ __forceinline__ double2 split_config(const PT *__restrict__ Tn, int loc, const T *__restrict__ thr,
                                                double2 dp, bool &bad) {
  constexpr int S = wick_stride(M);
  int pr[M + 1], pc[M + 1];
  static_for<0, M>([&](auto I) {
    constexpr int i = I;
    const int pi = 2 * i + ((loc >> i) & 1);
    pr[i] = pi * S;
    pc[i] = pi;
  });
  pr[M] = 2 * M * S;
  pc[M] = 2 * M;
  cx<T> Uh[M + 1][M + 1];
  cx<T> det = cmk<T>(T(1), T(0));
  cx<T> s = cmk<T>(T(0), T(0));
  bool flag = false;
  static_for<0, M + 1>([&](auto I) {
    constexpr int i = I;
    cx<T> row[M + 1];
    // opaque row offset: the row's addresses are formed here, not all up front
    int ri = pr[i];
    __asm__ volatile("" : "+v"(ri));
    static_for<0, M + 1>([&](auto J) {
      constexpr int j = J;
      row[j] = ldx<T>(Tn[ri + pc[j]]);
    });
    static_for<0, i>([&](auto KB) {
      constexpr int kb = KB;
      const cx<T> l = row[kb];
      static_for<kb + 1, M + 1>([&](auto J) {
        constexpr int j = J;
        row[j] = cfms<T>(row[j], l, Uh[kb][j]);
      });
    });
    if constexpr (i < M) {
      const cx<T> piv = row[i];
      flag |= !(cabs1<T>(piv) >= thr[pc[i]]);
      if constexpr (i == 0)
        det = piv;
      else
        det = cmul<T>(det, piv);
      const cx<T> inv = cinv<T>(piv);
      static_for<i + 1, M + 1>([&](auto J) {
        constexpr int j = J;
        Uh[i][j] = cmul<T>(row[j], inv);
      });
    } else {
      s = row[M];
    }
    __builtin_amdgcn_sched_barrier(0);
  });
  bad = flag;
  const double2 dt = make_double2(dp.x * (double)det.x - dp.y * (double)det.y,
                                  dp.x * (double)det.y + dp.y * (double)det.x);
  return cmul<double>(cmul<double>(dt, dt), make_double2((double)s.x, (double)s.y));
}

// det(A)^2 s of the two configurations that differ only in the Keldysh index of
// the last vertex (local bit M-1 of loc: 0 for v[0], 1 for v[1]), from the same
// node table.  The first M-1 vertex rows are eliminated once over the columns
// of both choices of the last vertex (and the border), the border row likewise;
// only the last vertex row and the corner update are per choice -- at M = 5,
// 70 instead of 110 complex multiply-adds and 4 instead of 10 reciprocals per
// pair (the last pivot enters without one).
// PF: row q+1 is loaded while row q is eliminated (one row of registers more).
template <int M, typename T, typename PT, bool PF = true>
__device__ __forceinline__ void split_config_pair(const PT *__restrict__ Tn, int loc, const T *__restrict__ thr,
                                                  double2 dp, double2 (&v)[2], bool (&bad)[2]) {
  constexpr int S = wick_stride(M);
  constexpr int MC = M + 2; // columns: vertices 0..M-2, last vertex (index 0), last vertex (index 1), border
  int pc[MC];
  static_for<0, M - 1>([&](auto I) {
    constexpr int i = I;
    pc[i] = 2 * i + ((loc >> i) & 1);
  });
  pc[M - 1] = 2 * M - 2;
  pc[M] = 2 * M - 1;
  pc[M + 1] = 2 * M;
  // The rows in the order they are eliminated: the M-1 vertex rows (q < M-1),
  // the border row (q = M-1), then the last vertex's row for Keldysh index 0
  // and 1 (q = M, M+1; only the other vertices' columns, their own and the
  // border's are needed).  With PF, row q+1 is loaded while row q is
  // eliminated, so that one LDS round trip per row overlaps the arithmetic.
  constexpr int NQ = M + 2;
  auto used = [](int q, int j) { return q < M || j < M - 1 || j == q - 1 || j == MC - 1; };
  auto load_row = [&](auto Q, cx<T>(&dst)[MC]) {
    constexpr int q = Q;
    int ri = (q < M - 1 ? pc[q] : q == M - 1 ? 2 * M : pc[q - 1]) * S;
    __asm__ volatile("" : "+v"(ri));
    static_for<0, MC>([&](auto J) {
      constexpr int j = J;
      if constexpr (used(q, j)) dst[j] = ldx<T>(Tn[ri + pc[j]]);
    });
  };
  T th[M + 1]; // pivot thresholds of the vertex rows and of the last vertex's two rows
  static_for<0, M + 1>([&](auto I) { th[I] = thr[pc[I]]; });
  cx<T> Uh[M - 1][MC];
  cx<T> rb[MC];
  cx<T> det = cmk<T>(T(1), T(0));
  double2 D2;
  bool flag = false;
  cx<T> nxt[MC];
  if constexpr (PF) load_row(std::integral_constant<int, 0>{}, nxt);
  static_for<0, NQ>([&](auto Q) {
    constexpr int q = Q;
    cx<T> row[MC];
    if constexpr (PF) {
      static_for<0, MC>([&](auto J) {
        if constexpr (used(q, (int)J)) row[J] = nxt[J];
      });
      if constexpr (q + 1 < NQ) load_row(std::integral_constant<int, q + 1>{}, nxt);
    } else {
      load_row(Q, row);
    }
    static_for<0, (q < M - 1 ? q : M - 1)>([&](auto KB) {
      constexpr int kb = KB;
      const cx<T> l = row[kb];
      static_for<kb + 1, MC>([&](auto J) {
        constexpr int j = J;
        if constexpr (used(q, j)) row[j] = cfms<T>(row[j], l, Uh[kb][j]);
      });
    });
    if constexpr (q < M - 1) {
      const cx<T> piv = row[q];
      flag |= !(cabs1<T>(piv) >= th[q]);
      if constexpr (q == 0)
        det = piv;
      else
        det = cmul<T>(det, piv);
      const cx<T> inv = cinv<T>(piv);
      static_for<q + 1, MC>([&](auto J) {
        constexpr int j = J;
        Uh[q][j] = cmul<T>(row[j], inv);
      });
    } else if constexpr (q == M - 1) {
      // border row over the same columns (its border entry is the node's corner)
      static_for<0, MC>([&](auto J) { rb[J] = row[J]; });
    } else {
      // det(A)^2 s with the last pivot p and the border's Schur complement
      // s = rb_B - rb_c row_B / p:  (dp det p)^2 s = (dp det)^2 p (p rb_B - rb_c row_B),
      // no reciprocal of the last pivot; (dp det)^2 is shared by the two rows
      constexpr int t = q - M, ct = M - 1 + t; // Keldysh index of the last vertex, its column
      if constexpr (t == 0) {
        const double2 D =
            make_double2(dp.x * (double)det.x - dp.y * (double)det.y, dp.x * (double)det.y + dp.y * (double)det.x);
        D2 = cmul<double>(D, D);
      }
      const cx<T> piv = row[ct];
      bad[t] = flag | !(cabs1<T>(piv) >= th[M - 1 + t]);
      const cx<T> x = cfms<T>(cmul<T>(piv, rb[MC - 1]), rb[ct], row[MC - 1]);
      const cx<T> y = cmul<T>(piv, x);
      v[t] = cmul<double>(D2, make_double2((double)y.x, (double)y.y));
    }
    __builtin_amdgcn_sched_barrier(0);
  });
}

} // namespace kd
