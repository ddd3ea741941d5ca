// kd_lu.hpp -- compile-time helpers shared by the Keldysh-sum kernels.
#pragma once

#include <type_traits>

namespace kd {

// f(integral_constant<int, i>) for i = B .. E-1, unrolled at compile time.
template <int B, int E, typename F> __device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Pivot check of the unpivoted LU: a pivot below this fraction of the largest
// |Re|+|Im| of the point's vertex block flags the configuration for the
// partial-pivoting FP64 fallback.  It covers cancellation during the
// elimination (the pivot's own diagonal entry is part of the block) and a
// diagonal that is small from the start (g^<(0) - i alpha ~ 0 at half filling
// with alpha = 1/2), where the unpivoted LU would divide by rounding noise.
// The kernels form the threshold as max(|diagonal|, block maximum): the same
// value, in a form that keeps their register allocation.
template <typename T> struct PivotTol;
// floor: a positive lower bound of the threshold's scale, so that a zero
// pivot fails the check without a separate test.
template <> struct PivotTol<double> {
  static constexpr double v = 1.0 / 4096.0;
  static constexpr double floor = 1e-290;
};
template <> struct PivotTol<float> {
  static constexpr float v = 1.0f / 256.0f;
  static constexpr float floor = 1e-30f;
};

// LDS layout of a point's Wick matrix: row/column 2i+s for vertex i with
// Keldysh index s, 2N for the external point; row stride wick_stride(N) >= 2N+1,
// padded so that the four candidates (2i+{0,1}) x (2j+{0,1}) of an entry --
// what the lanes of a wave read at once, one per configuration -- fall in
// different LDS banks (row stride mod 8 not in {0, 1, 7}, in 16-byte units).
constexpr int wick_stride(int N) {
  int s = 2 * N + 1;
  while (s % 8 == 0 || s % 8 == 1 || s % 8 == 7) s++;
  return s;
}

} // namespace kd
