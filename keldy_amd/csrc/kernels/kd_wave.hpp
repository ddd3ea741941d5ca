// kd_wave.hpp -- synchronisation and reductions inside one 64-lane wavefront,
// shared by the one-point-per-wave Keldysh-sum kernels (kd_wsplit.hip,
// kd_ctree.hip).
#pragma once

#include "kd_device.hpp"

namespace kd {

// LDS writes of the wave visible to its other lanes: a wave's LDS operations
// complete in issue order, so the only thing to stop is the compiler moving
// memory operations across this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave reductions without LDS traffic: DPP moves within each row of 16 lanes
// (quad permutations, then the half-row and row mirrors), then the four row
// results read out as scalars.  Every lane gets the same value.
template <int CTRL> __device__ __forceinline__ double dpp_f64(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_f64(double x, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l),
                          __builtin_amdgcn_readlane(__double2loint(x), l));
}
// a wave-uniform lane's value (the lane index uniform: v_readlane, no LDS trip)
__device__ __forceinline__ int rdlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_f64<0xB1>(x);  // quad_perm [1,0,3,2]
  x += dpp_f64<0x4E>(x);  // quad_perm [2,3,0,1]
  x += dpp_f64<0x141>(x); // row_half_mirror
  x += dpp_f64<0x140>(x); // row_mirror
  return (lane_f64(x, 0) + lane_f64(x, 16)) + (lane_f64(x, 32) + lane_f64(x, 48));
}
__device__ __forceinline__ double wave_max(double x) {
  x = fmax(x, dpp_f64<0xB1>(x));
  x = fmax(x, dpp_f64<0x4E>(x));
  x = fmax(x, dpp_f64<0x141>(x));
  x = fmax(x, dpp_f64<0x140>(x));
  return fmax(fmax(lane_f64(x, 0), lane_f64(x, 16)), fmax(lane_f64(x, 32), lane_f64(x, 48)));
}

// The same reductions over each half (32 lanes) of the wave: every lane gets
// the value of its half (upper: lanes 32..63).
__device__ __forceinline__ double half_sum(double x, bool upper) {
  x += dpp_f64<0xB1>(x);
  x += dpp_f64<0x4E>(x);
  x += dpp_f64<0x141>(x);
  x += dpp_f64<0x140>(x);
  const double lo = lane_f64(x, 0) + lane_f64(x, 16), hi = lane_f64(x, 32) + lane_f64(x, 48);
  return upper ? hi : lo;
}
__device__ __forceinline__ double half_max(double x, bool upper) {
  x = fmax(x, dpp_f64<0xB1>(x));
  x = fmax(x, dpp_f64<0x4E>(x));
  x = fmax(x, dpp_f64<0x141>(x));
  x = fmax(x, dpp_f64<0x140>(x));
  const double lo = fmax(lane_f64(x, 0), lane_f64(x, 16)), hi = fmax(lane_f64(x, 32), lane_f64(x, 48));
  return upper ? hi : lo;
}

} // namespace kd
