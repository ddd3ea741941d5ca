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
template <> struct PivotTol<double> { static constexpr double v = 1.0 / 4096.0; };
template <> struct PivotTol<float> { static constexpr float v = 1.0f / 256.0f; };

} // namespace kd
