// quadrature.hpp -- numerical special functions and adaptive quadrature used to
// tabulate g0 on the host.
#pragma once

#include <complex>
#include <functional>

namespace keldy::numeric {

using dcomplex = std::complex<double>;

/// exp(x) * E1(x) for x > 0 (E1 the exponential integral).
double exp_E1(double x);
/// exp(-x) * Ei(x) for x > 0.
double expm_Ei(double x);

struct quad_result {
  dcomplex value;
  double abserr;
  int intervals;
};

/// Adaptive bisection quadrature of a complex function on [a, b]: each panel
/// is integrated with 20- and 40-point Gauss-Legendre rules and the panel with
/// the largest |I40 - I20| is split until the total falls below
/// max(abstol, reltol |I|) or max_intervals is reached.
quad_result integrate(const std::function<dcomplex(double)> &f, double a, double b, double abstol,
                      double reltol, int max_intervals = 2000);
/// Same on [0, +inf) through x = s / (1 - s).
quad_result integrate_semi_infinite(const std::function<dcomplex(double)> &f, double abstol, double reltol,
                                    int max_intervals = 2000);

} // namespace keldy::numeric
