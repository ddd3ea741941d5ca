// quadrature.cpp -- exponential integrals (power series, Lentz continued
// fraction and the asymptotic series) and Gauss-Legendre adaptive quadrature.
#include "quadrature.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <queue>
#include <vector>

namespace keldy::numeric {

namespace {
constexpr double euler_gamma = 0.57721566490153286060651209008240243;

// E1(x) for 0 < x <= 1: -gamma - ln x - sum_k (-x)^k / (k k!)
double E1_series(double x) {
  double sum = 0.0, term = 1.0;
  for (int k = 1; k < 200; k++) {
    term *= -x / k;
    const double add = term / k;
    sum += add;
    if (std::abs(add) < 1e-18 * std::abs(sum)) break;
  }
  return -euler_gamma - std::log(x) - sum;
}

// exp(x) E1(x) for x > 1 by the continued fraction
// E1(x) = e^{-x} / (x + 1 - 1/(x + 3 - 4/(x + 5 - ...))) (modified Lentz).
double expE1_cf(double x) {
  const double tiny = 1e-300;
  double b = x + 1.0, c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int i = 1; i < 1000; i++) {
    const double an = -double(i) * double(i);
    b += 2.0;
    d = 1.0 / (an * d + b);
    c = b + an / c;
    const double del = c * d;
    h *= del;
    if (std::abs(del - 1.0) < 1e-16) break;
  }
  return h;
}

// Ei(x) for 0 < x <= 40: gamma + ln x + sum_k x^k / (k k!)
double Ei_series(double x) {
  double sum = 0.0, term = 1.0;
  for (int k = 1; k < 400; k++) {
    term *= x / k;
    const double add = term / k;
    sum += add;
    if (add < 1e-18 * sum) break;
  }
  return euler_gamma + std::log(x) + sum;
}

// e^{-x} Ei(x) for x > 40: (1/x) sum_k k! / x^k, truncated at its smallest term
double expmEi_asym(double x) {
  double sum = 1.0, term = 1.0;
  for (int k = 1; k < 200; k++) {
    const double next = term * k / x;
    if (next > term) break;
    term = next;
    sum += term;
    if (term < 1e-18 * sum) break;
  }
  return sum / x;
}

struct gl_rule {
  std::vector<double> x, w;
};

// Gauss-Legendre nodes and weights on [-1, 1] by Newton iteration on P_n.
gl_rule make_gl(int n) {
  gl_rule r;
  r.x.resize(n);
  r.w.resize(n);
  for (int i = 0; i < n; i++) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; it++) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      const double pp = n * (z * p1 - p2) / (z * z - 1.0);
      const double z1 = z;
      z = z1 - p1 / pp;
      if (std::abs(z - z1) < 1e-16) {
        r.w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
        break;
      }
      r.w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
    r.x[i] = z;
  }
  return r;
}

const gl_rule &gl20() {
  static const gl_rule r = make_gl(20);
  return r;
}
const gl_rule &gl40() {
  static const gl_rule r = make_gl(40);
  return r;
}

struct panel {
  double a, b;
  dcomplex val;
  double err;
  bool operator<(const panel &o) const { return err < o.err; }
};

panel eval_panel(const std::function<dcomplex(double)> &f, double a, double b) {
  const double c = 0.5 * (a + b), h = 0.5 * (b - a);
  dcomplex s20 = 0.0, s40 = 0.0;
  const gl_rule &r20 = gl20(), &r40 = gl40();
  for (size_t i = 0; i < r20.x.size(); i++) s20 += r20.w[i] * f(c + h * r20.x[i]);
  for (size_t i = 0; i < r40.x.size(); i++) s40 += r40.w[i] * f(c + h * r40.x[i]);
  s20 *= h;
  s40 *= h;
  return panel{a, b, s40, std::abs(s40 - s20)};
}
} // namespace

double exp_E1(double x) {
  if (x <= 1.0) return std::exp(x) * E1_series(x);
  return expE1_cf(x);
}

double expm_Ei(double x) {
  if (x <= 40.0) return std::exp(-x) * Ei_series(x);
  return expmEi_asym(x);
}

quad_result integrate(const std::function<dcomplex(double)> &f, double a, double b, double abstol, double reltol,
                      int max_intervals) {
  std::priority_queue<panel> q;
  panel p0 = eval_panel(f, a, b);
  dcomplex total = p0.val;
  double err = p0.err;
  q.push(p0);
  int n = 1;
  while (err > std::max(abstol, reltol * std::abs(total)) && n < max_intervals) {
    panel p = q.top();
    q.pop();
    const double m = 0.5 * (p.a + p.b);
    panel l = eval_panel(f, p.a, m), r = eval_panel(f, m, p.b);
    total += l.val + r.val - p.val;
    err += l.err + r.err - p.err;
    q.push(l);
    q.push(r);
    n++;
  }
  // re-sum to avoid the drift of the running updates
  total = 0.0;
  err = 0.0;
  while (!q.empty()) {
    total += q.top().val;
    err += q.top().err;
    q.pop();
  }
  return quad_result{total, err, n};
}

quad_result integrate_semi_infinite(const std::function<dcomplex(double)> &f, double abstol, double reltol,
                                    int max_intervals) {
  auto g = [&f](double s) -> dcomplex {
    if (s >= 1.0) return 0.0;
    const double om = 1.0 - s;
    return f(s / om) / (om * om);
  };
  return integrate(g, 0.0, 1.0, abstol, reltol, max_intervals);
}

} // namespace keldy::numeric
