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

// Pivot check of the unpivoted LU: a pivot that lost more than this fraction
// of its original diagonal entry (|Re|+|Im|) flags the configuration for the
// partial-pivoting FP64 fallback.
template <typename T> struct PivotTol;
template <> struct PivotTol<double> { static constexpr double v = 1.0 / 4096.0; };
template <> struct PivotTol<float> { static constexpr float v = 1.0f / 256.0f; };

} // namespace kd
