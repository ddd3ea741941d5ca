// qrng.hpp -- Sobol generator of the host API (reference: c++/keldy/qrng.hpp:49-114,
// qrng_details/digitalseq_b2g.hpp).  The host object owns the (scrambled,
// shifted) generating matrices; points for the QMC driver are generated on the
// device from the same tables (kd_set_sobol), the host operator() serves the
// reference's per-point Python API.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace keldy {

/// Joe-Kuo (2008) generating matrices: Cs[j*m + c] is column c of dimension j
/// as an integer with the least significant bit in the top row.  dim <= 32.
std::vector<uint32_t> jk2008_matrices(int dim, int m);

class sobol {
 public:
  static constexpr int max_dim = 32;
  static constexpr int m = 31; // columns; 2^31 points

  /// Unshifted, unscrambled sequence ("sobol_unshifted", qrng.hpp:69-75).
  sobol(int dim, int rng_state_seed, int log_max_points = 31);
  /// Digitally shifted and/or linearly scrambled sequence (qrng.hpp:77-109).
  sobol(int dim, int rng_state_seed, bool do_shift, bool do_scramble, int rng_seed_shift,
        int log_max_points = 31);

  /// Next point (first call returns point k = seed + 1).
  std::vector<double> operator()();
  void seed(int k) { set_state(uint32_t(k)); }
  void discard(int nr_discard) {
    if (nr_discard != 0) set_state(k_ + uint32_t(nr_discard));
  }

  int dim() const { return dim_; }
  uint32_t state() const { return k_; }
  void set_state(uint32_t k);
  bool past_end() const { return k_ >= (uint32_t(1) << m); }
  /// Bit-reversed generating matrices [dim][31] and shift [dim] (device format).
  const std::vector<uint32_t> &csr() const { return csr_; }
  const std::vector<uint32_t> &shfr() const { return shfr_; }

 private:
  void init(int dim, uint32_t seed, bool do_shift, bool do_scramble, int rng_seed_shift, int log_max_points);
  int dim_ = 0;
  uint32_t k_ = 0;
  std::vector<uint32_t> csr_, shfr_, cur_;
};

} // namespace keldy
