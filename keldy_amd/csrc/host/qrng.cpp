// qrng.cpp -- Sobol set-up (reference: qrng.hpp:49-114; digitalseq_b2g.hpp:130-237,
// 295-547).  Direction numbers are generated from the Joe-Kuo (2008)
// new-joe-kuo-6.21201 primitive polynomials and initial values; the
// reference stores the same matrices as a literal table
// (digitalseq_b2g.hpp:598-861).
#include "qrng.hpp"

#include <algorithm>
#include <random>

namespace keldy {

namespace {

// degree s, coefficients a, initial m_1..m_s for dimensions 2..32 (dimension 1
// is the van der Corput sequence)
const int jk_s[sobol::max_dim] = {0, 1, 2, 3, 3, 4, 4, 5, 5, 5, 5, 5, 5, 6, 6, 6,
                                  6, 6, 6, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7};
const int jk_a[sobol::max_dim] = {0, 0, 1, 1, 2, 1, 4, 2, 4, 7, 11, 13, 14, 1, 13, 16,
                                  19, 22, 25, 1, 4, 7, 8, 14, 19, 21, 28, 31, 32, 37, 41, 42};
const int jk_m[sobol::max_dim][7] = {
    {0}, {1}, {1, 3}, {1, 3, 1}, {1, 1, 1}, {1, 1, 3, 3}, {1, 3, 5, 13}, {1, 1, 5, 5, 17},
    {1, 1, 5, 5, 5}, {1, 1, 7, 11, 19}, {1, 1, 5, 1, 1}, {1, 1, 1, 3, 11}, {1, 3, 5, 5, 31},
    {1, 3, 3, 9, 7, 49}, {1, 1, 1, 15, 21, 21}, {1, 3, 1, 13, 27, 49}, {1, 1, 1, 15, 7, 5},
    {1, 3, 1, 15, 13, 25}, {1, 1, 5, 5, 19, 61}, {1, 3, 7, 11, 23, 15, 103}, {1, 3, 7, 13, 13, 15, 69},
    {1, 1, 3, 13, 7, 35, 63}, {1, 3, 5, 9, 1, 25, 53}, {1, 3, 1, 13, 9, 35, 107}, {1, 3, 1, 5, 27, 61, 31},
    {1, 1, 5, 11, 19, 41, 61}, {1, 3, 5, 3, 3, 13, 69}, {1, 1, 7, 13, 1, 19, 1}, {1, 3, 7, 5, 13, 19, 59},
    {1, 1, 3, 9, 25, 29, 41}, {1, 3, 5, 13, 23, 1, 55}, {1, 3, 7, 3, 13, 59, 17}};

uint32_t bitrev32(uint32_t v) {
  v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
  v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
  v = ((v >> 4) & 0x0F0F0F0Fu) | ((v & 0x0F0F0F0Fu) << 4);
  v = ((v >> 8) & 0x00FF00FFu) | ((v & 0x00FF00FFu) << 8);
  return (v >> 16) | (v << 16);
}

} // namespace

std::vector<uint32_t> jk2008_matrices(int dim, int m) {
  if (dim < 1 || dim > sobol::max_dim) throw keldy_error("Sobol dimension must be in [1, 32]");
  std::vector<uint32_t> Cs(size_t(dim) * m);
  for (int j = 0; j < dim; j++) {
    std::vector<uint64_t> mv(m);
    const int s = jk_s[j];
    if (j == 0) {
      for (int c = 0; c < m; c++) mv[c] = 1;
    } else {
      for (int c = 0; c < s && c < m; c++) mv[c] = uint64_t(jk_m[j][c]);
      // Sobol recurrence m_k = 2 a_1 m_{k-1} ^ ... ^ 2^{s-1} a_{s-1} m_{k-s+1} ^ 2^s m_{k-s} ^ m_{k-s}
      for (int c = s; c < m; c++) {
        uint64_t v = mv[c - s] ^ (mv[c - s] << s);
        for (int q = 1; q < s; q++)
          if ((jk_a[j] >> (s - 1 - q)) & 1) v ^= mv[c - q] << q;
        mv[c] = v;
      }
    }
    for (int c = 0; c < m; c++) {
      uint32_t r = 0;
      for (int b = 0; b <= c; b++)
        if ((mv[c] >> b) & 1) r |= 1u << (c - b);
      Cs[size_t(j) * m + c] = r;
    }
  }
  return Cs;
}

sobol::sobol(int dim, int rng_state_seed, int log_max_points) {
  init(dim, uint32_t(rng_state_seed), false, false, 0, log_max_points);
}

sobol::sobol(int dim, int rng_state_seed, bool do_shift, bool do_scramble, int rng_seed_shift,
             int log_max_points) {
  init(dim, uint32_t(rng_state_seed), do_shift, do_scramble, rng_seed_shift, log_max_points);
}

void sobol::init(int dim, uint32_t seed, bool do_shift, bool do_scramble, int rng_seed_shift,
                 int log_max_points) {
  if (log_max_points > 31) throw keldy_error("To many points requested from Sobol generator.");
  if (log_max_points != 31)
    throw keldy_error("only log_max_points = 31 is supported (the generator keeps 31 columns)");
  std::vector<uint32_t> Cs = jk2008_matrices(dim, m);
  std::vector<uint32_t> shifts(dim, 0u), scr = Cs;
  if (do_shift || do_scramble) {
    // The reference hands its engine by value to std::generate_n and to the
    // scramble generator, so both draw from the freshly seeded state.
    std::mt19937_64 draw_shift(rng_seed_shift);
    for (int j = 0; j < dim; j++) shifts[j] = uint32_t(draw_shift());
    if (!do_shift) std::fill(shifts.begin(), shifts.end(), 0u);
    if (do_scramble) {
      uint32_t mx = *std::max_element(Cs.begin(), Cs.end());
      unsigned t = 31; // bit depth of the columns (digitalseq_b2g.hpp:130-138)
      while (t && !((1u << (t - 1)) & mx)) t--;
      std::mt19937_64 draw(rng_seed_shift);
      // lower-triangular unit-diagonal scramble matrices, columns as integers
      std::vector<uint32_t> M(size_t(dim) * t);
      for (int j = 0; j < dim; j++)
        for (unsigned k = 0; k < t; k++) {
          uint64_t r = draw();
          const uint32_t hb = 1u << (t - k - 1);
          r &= uint64_t((hb - 1u) | hb);
          r |= 1u;
          r <<= k;
          M[size_t(j) * t + k] = uint32_t(r);
        }
      for (int j = 0; j < dim; j++)
        for (int c = 0; c < m; c++) {
          uint32_t out = 0;
          for (unsigned i = 0; i < t; i++)
            if ((Cs[size_t(j) * m + c] >> i) & 1u) out ^= M[size_t(j) * t + i];
          scr[size_t(j) * m + c] = out;
        }
    }
  }
  dim_ = dim;
  csr_.resize(scr.size());
  for (size_t i = 0; i < scr.size(); i++) csr_[i] = bitrev32(scr[i]);
  shfr_.resize(dim);
  for (int j = 0; j < dim; j++) shfr_[j] = bitrev32(shifts[j]);
  cur_.assign(dim, 0u);
  set_state(seed);
}

void sobol::set_state(uint32_t k) {
  k_ = k;
  const uint32_t gk = (k >> 1) ^ k;
  for (int j = 0; j < dim_; j++) {
    uint32_t c = 0;
    if (k != 0)
      for (int i = 0; i < m; i++)
        if ((1u << i) & gk) c ^= csr_[size_t(j) * m + i];
    cur_[j] = c;
  }
}

std::vector<double> sobol::operator()() {
  if (past_end()) throw keldy_error("More sample points requested then generator can hold.");
  k_ += 1;
  const unsigned c = unsigned(__builtin_ctz(k_)) % unsigned(m);
  std::vector<double> x(dim_);
  for (int j = 0; j < dim_; j++) {
    cur_[j] ^= csr_[size_t(j) * m + c];
    x[j] = double(cur_[j] ^ shfr_[j]) * (1.0 / 4294967296.0);
  }
  return x;
}

} // namespace keldy
