// common.hpp -- basic types of the host API (reference: c++/keldy/common.hpp:30-75,
// common.cpp:47-57).
#pragma once

#include <cmath>
#include <complex>
#include <limits>
#include <stdexcept>
#include <string>

namespace keldy {

using dcomplex = std::complex<double>;
using time_real_t = double;
using orbital_t = int;

enum spin_t { up = 0, down = 1 };
enum keldysh_idx_t { forward = 0, backward = 1 };

// Errors of the host API: the reference raises TRIQS_RUNTIME_ERROR, which its
// Python bindings turn into RuntimeError; std::runtime_error maps the same way.
struct keldy_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

/// Point of the Keldysh contour (time, Keldysh index, time-splitting order).
struct contour_pt_t {
  time_real_t time{};
  keldysh_idx_t k_idx = forward;
  int timesplit_n = 0;
  bool operator==(const contour_pt_t &o) const {
    return time == o.time && k_idx == o.k_idx && timesplit_n == o.timesplit_n;
  }
};

/// Three-way contour ordering: forward branch before backward branch, forward
/// branch ordered by time, backward branch anti-ordered, ties split by
/// timesplit_n (common.cpp:47-57).
inline int compare_3way(const contour_pt_t &a, const contour_pt_t &b) {
  const int kd = int(a.k_idx) - int(b.k_idx);
  if (kd != 0) return kd;
  const double dt = a.time - b.time;
  if (dt != 0.0) return (1 - 2 * int(a.k_idx)) * (std::signbit(dt) ? -1 : 1);
  return (1 - 2 * int(a.k_idx)) * (a.timesplit_n - b.timesplit_n);
}

/// Fermi function for real or complex argument, beta >= 0 (common.hpp:64-75).
template <typename T> inline T n_fermi(T omega, double beta) {
  if (omega == T(0.)) return T(0.5);
  if (std::isinf(beta)) {
    // beta = +infinity: step function (the reference reaches the same limit
    // through exp(+-inf))
    return std::real(omega) > 0 ? T(0.) : T(1.);
  }
  if (std::real(omega) > 0) {
    T y = std::exp(-beta * omega);
    return y / (1. + y);
  }
  return T(1.0) / (std::exp(beta * omega) + 1.);
}

} // namespace keldy
