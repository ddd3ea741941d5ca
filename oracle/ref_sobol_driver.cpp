// Drives the reference's own Sobol header (digitalseq_b2g.hpp, included from
// /root/reference at build time) through the generator set-up that
// keldy::sobol performs (qrng.hpp:69-109), and prints points for the golden
// vectors.  Test infrastructure only; the set-up steps below are written here
// because qrng.hpp itself needs TRIQS and Boost headers that are not present.
//
// usage: sobol_ref dim state_seed do_shift do_scramble rng_seed_shift n_points
//        sobol_ref mt64 seed count      (std::mt19937_64 reference stream)
#include "digitalseq_b2g.hpp"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

int main(int argc, char **argv) {
  if (argc >= 4 && std::strcmp(argv[1], "mt64") == 0) {
    std::mt19937_64 g(std::strtoull(argv[2], nullptr, 10));
    long cnt = std::atol(argv[3]);
    for (long i = 0; i < cnt; i++) std::printf("%llu\n", (unsigned long long)g());
    return 0;
  }
  if (argc < 7) return 2;
  int dim = std::atoi(argv[1]);
  int seed = std::atoi(argv[2]);
  bool do_shift = std::atoi(argv[3]) != 0;
  bool do_scramble = std::atoi(argv[4]) != 0;
  int rng_seed_shift = std::atoi(argv[5]);
  long npts = std::atol(argv[6]);
  const int m = 31;
  std::vector<std::uint32_t> Cs(dim * m);
  qmc::JK2008_sobol_matrices(Cs.begin(), dim, m);
  auto t = qmc::bit_depth(Cs.begin(), Cs.end());
  unsigned t2 = t;
  std::vector<std::uint32_t> shifts(dim);
  std::vector<std::uint32_t> scrambled = Cs;
  std::vector<std::uint32_t> M(dim * t);
  if (do_shift || do_scramble) {
    std::mt19937_64 rng(rng_seed_shift);
    std::generate_n(shifts.begin(), dim, rng);
    if (!do_shift) std::fill_n(shifts.begin(), dim, 0);
    if (do_scramble) {
      qmc::generate_random_linear_scramble(dim, t2, t, M.begin(), rng);
      qmc::scramble(dim, t, m, M.begin(), Cs.begin(), scrambled.begin());
    }
  }
  qmc::digitalseq_b2g<double, std::uint32_t> gen(dim, m, scrambled.begin(), shifts.begin());
  gen.set_state(std::uint32_t(seed));
  for (long i = 0; i < npts; i++) {
    if (gen.past_end()) return 3;
    auto x = gen.next();
    for (int j = 0; j < dim; j++) std::printf(j ? " %.17g" : "%.17g", x[j]);
    std::printf("\n");
  }
  return 0;
}
