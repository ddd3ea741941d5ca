// kd_pivoted.hpp -- the partial-pivoting FP64 fallback of the Keldysh-sum
// kernels, for configurations whose unpivoted LU fails the pivot check.
#pragma once

#include "kd_device.hpp"

namespace kd {

// Fallback: bordered LU with partial pivoting over the rows of A (pivot =
// max |Re|+|Im|, first on ties: the LAPACK zgetf2 rule), FP64.  The row
// permutation's sign cancels in det(A)^2.
__device__ __forceinline__ double2 config_value_pivoted_body(const double2 *__restrict__ Wq, int cfg, int N, int DS) {
  const int M = N + 1;
  double2 a[(KD_MAX_ORDER + 1) * (KD_MAX_ORDER + 1)];
  int pk[KD_MAX_ORDER + 1];
  for (int i = 0; i < N; i++) pk[i] = 2 * i + ((cfg >> i) & 1);
  pk[N] = 2 * N;
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) a[i * M + j] = Wq[pk[i] * DS + pk[j]];
  a[N * M + N] = make_double2(0.0, 0.0);
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

// Called out of line by the kernels whose hot loop must not carry its code.
static __device__ __noinline__ double2 config_value_pivoted(const double2 *__restrict__ Wq, int cfg, int N, int DS) {
  return config_value_pivoted_body(Wq, cfg, N, DS);
}

} // namespace kd
