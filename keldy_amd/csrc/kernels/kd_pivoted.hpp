// kd_pivoted.hpp -- the partial-pivoting FP64 fallback of the Keldysh-sum
// kernels, for configurations whose unpivoted LU fails the pivot check.
#pragma once

#include "kd_device.hpp"

namespace kd {

// Fallback: bordered LU with partial pivoting over the rows of the
// configuration's (N+1) x (N+1) matrix a (pivot = max |Re|+|Im|, first on ties:
// the LAPACK zgetf2 rule), FP64; returns det(vertex block) * det(a), i.e.
// det(A)^2 s.  The row permutation's sign cancels in det(A)^2.
__device__ __forceinline__ double2 pivoted_value(double2 *__restrict__ a, int N) {
  const int M = N + 1;
  double2 det = make_double2(1.0, 0.0);
  for (int k = 0; k < N; k++) {
    int p = k;
    double best = cabs1<double>(a[k * M + k]);
    for (int r = k + 1; r < N; r++) {
      double v = cabs1<double>(a[r * M + k]);
      if (v > best) {
        best = v;
        p = r;
      }
    }
    if (best == 0.0) return make_double2(0.0, 0.0);
    if (p != k)
      for (int j = 0; j < M; j++) {
        double2 t = a[k * M + j];
        a[k * M + j] = a[p * M + j];
        a[p * M + j] = t;
      }
    double2 piv = a[k * M + k];
    det = cmul<double>(det, piv);
    double2 inv = cinv<double>(piv);
    for (int r = k + 1; r < M; r++) {
      double2 l = cmul<double>(a[r * M + k], inv);
      for (int j = k + 1; j < M; j++) a[r * M + j] = cfms<double>(a[r * M + j], l, a[k * M + j]);
    }
  }
  double2 d2 = cmul<double>(det, det);
  return cmul<double>(d2, a[N * M + N]);
}

// The configuration cfg (bit i: Keldysh index of vertex i) from the interleaved
// Wick matrix Wq (row/column 2i+s, 2N the external point, row stride DS).
__device__ __forceinline__ double2 config_value_pivoted_body(const double2 *__restrict__ Wq, int cfg, int N, int DS) {
  const int M = N + 1;
  double2 a[(KD_MAX_ORDER + 1) * (KD_MAX_ORDER + 1)];
  int pk[KD_MAX_ORDER + 1];
  for (int i = 0; i < N; i++) pk[i] = 2 * i + ((cfg >> i) & 1);
  pk[N] = 2 * N;
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) a[i * M + j] = Wq[pk[i] * DS + pk[j]];
  a[N * M + N] = make_double2(0.0, 0.0);
  return pivoted_value(a, N);
}

// The configuration cfg of the time-ordered vertex-level table Wv (vertices
// sorted by time; Wv[(i (N+1) + j) 2 + s]: entry (i, j) when the earlier of i,
// j -- the lower index -- has Keldysh index s; index N the external points;
// the diagonal is the constant D).
static __device__ __noinline__ double2 config_value_pivoted_sorted(const double2 *__restrict__ Wv, int cfg, int N,
                                                                   double2 D) {
  const int M = N + 1;
  double2 a[(KD_MAX_ORDER + 1) * (KD_MAX_ORDER + 1)];
  for (int i = 0; i < N; i++) {
    const int si = (cfg >> i) & 1;
    for (int j = 0; j < N; j++) {
      const int e = i < j ? i : j;
      a[i * M + j] = i == j ? D : Wv[(i * M + j) * 2 + ((cfg >> e) & 1)];
    }
    a[i * M + N] = Wv[(i * M + N) * 2 + si];
    a[N * M + i] = Wv[(N * M + i) * 2 + si];
  }
  a[N * M + N] = make_double2(0.0, 0.0);
  return pivoted_value(a, N);
}

// Called out of line by the kernels whose hot loop must not carry its code.
static __device__ __noinline__ double2 config_value_pivoted(const double2 *__restrict__ Wq, int cfg, int N, int DS) {
  return config_value_pivoted_body(Wq, cfg, N, DS);
}

} // namespace kd
