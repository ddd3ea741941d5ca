// model.cpp -- bath hybridisations, g0(omega) and its tabulation in time.
// Reference: c++/keldy/impurity_oneband/model.cpp:63-466, model.hpp:100-153.
#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <iostream>

#include "quadrature.hpp"

namespace keldy::impurity_oneband {

using namespace std::complex_literals;

bool operator<(gf_index_t const &a, gf_index_t const &b) {
  const int c = compare_3way(a.contour, b.contour);
  if (c != 0) return c < 0;
  if (a.orbital != b.orbital) return a.orbital < b.orbital;
  return a.spin < b.spin;
}

// --------------------------------------------------------------------------
g0_model_omega::g0_model_omega(model_param_t const &parameters) : param_(parameters) {
  if (param_.beta < 0) {
    std::cout << "WARNING: beta < 0. Interpret this as beta = +infinity." << std::endl;
    param_.beta = std::numeric_limits<double>::infinity();
  }
  if (param_.nr_time_points_gf % 2 == 0) {
    std::cout << "WARNING: number of time points has been increased to make it odd." << std::endl;
    param_.nr_time_points_gf += 1;
  }
  const double G = param_.Gamma, D = param_.half_bandwidth;
  if (param_.bath_type == "semicircle") {
    hyb_ = [G, D](dcomplex omega) -> dcomplex {
      if (std::imag(omega) != 0.)
        throw keldy_error("Semicircular hybridization function out of the real axis is not implemented.");
      if (std::abs(std::real(omega)) < D) return 0.5 * (G / D) * (omega - 1.0i * std::sqrt(D * D - omega * omega));
      const double sgn = std::signbit(std::real(omega)) ? -1.0 : 1.0;
      return 0.5 * (G / D) * (omega - sgn * std::sqrt(omega * omega - D * D));
    };
  } else if (param_.bath_type == "flatband") {
    hyb_ = [G](dcomplex) -> dcomplex { return -1.0i * G / 2.; };
  } else if (param_.bath_type == "lorentzian") {
    // half_bandwidth is the half width at half maximum
    hyb_ = [G, D](dcomplex omega) -> dcomplex { return G * D / 2. / (omega + 1.0i * D); };
  } else {
    throw keldy_error("bath_type not defined");
  }
}

dcomplex g0_model_omega::g0_dot_R(dcomplex omega) const {
  return 1.0 / (omega - param_.eps_d - bath_hybrid_R_left(omega) - bath_hybrid_R_right(omega));
}
dcomplex g0_model_omega::g0_dot_A(dcomplex omega) const {
  return 1.0 / (omega - param_.eps_d - bath_hybrid_A_left(omega) - bath_hybrid_A_right(omega));
}

dcomplex g0_model_omega::g0_dot(dcomplex omega, bool is_lesser) const {
  const dcomplex hl = bath_hybrid_R_left(omega) - bath_hybrid_A_left(omega);
  const dcomplex hr = bath_hybrid_R_right(omega) - bath_hybrid_A_right(omega);
  const double sign = is_lesser ? 1.0 : -1.0;
  const dcomplex out = -sign * g0_dot_R(omega) *
                       (n_fermi<dcomplex>(sign * (omega - mu_left()), param_.beta) * hl +
                        n_fermi<dcomplex>(sign * (omega - mu_right()), param_.beta) * hr) *
                       g0_dot_A(omega);
  if (std::imag(omega) == 0) return 1.0i * std::imag(out);
  return out;
}

dcomplex g0_model_omega::g0_rightlead_dot(dcomplex omega, bool is_lesser) const {
  const dcomplex hr = bath_hybrid_R_right(omega) - bath_hybrid_A_right(omega);
  const double sign = is_lesser ? 1.0 : -1.0;
  return -sign * n_fermi<dcomplex>(sign * (omega - mu_right()), param_.beta) * hr * g0_dot_A(omega) +
         bath_hybrid_R_right(omega) * g0_dot(omega, is_lesser);
}

dcomplex g0_model_omega::g0_dot_keldysh(keldysh_idx_t a, keldysh_idx_t b, dcomplex omega) const {
  if (a < b) return g0_dot(omega, true);
  if (a > b) return g0_dot(omega, false);
  if (a == forward) return g0_dot(omega, true) + g0_dot_R(omega);
  return g0_dot(omega, false) - g0_dot_R(omega);
}

dcomplex g0_model_omega::bath_hybrid_K_left(dcomplex omega) const {
  return -(2. * n_fermi<dcomplex>(omega - mu_left(), param_.beta) - 1.) *
         (bath_hybrid_R_left(omega) - bath_hybrid_A_left(omega));
}
dcomplex g0_model_omega::bath_hybrid_K_right(dcomplex omega) const {
  return -(2. * n_fermi<dcomplex>(omega - mu_right(), param_.beta) - 1.) *
         (bath_hybrid_R_right(omega) - bath_hybrid_A_right(omega));
}

// --------------------------------------------------------------------------
void g0_model::allocate() {
  const size_t sz = size_t(n_orb) * n_orb * n_pts;
  lesser.assign(sz, 0.);
  greater.assign(sz, 0.);
  lesser_ft_error.assign(sz, 0.);
  greater_ft_error.assign(sz, 0.);
}

g0_model::g0_model(g0_model_omega model_omega_, bool make_dot_lead_, int n_orb_, double time_max_,
                   std::vector<dcomplex> lesser_, std::vector<dcomplex> greater_)
   : model_omega(std::move(model_omega_)), make_dot_lead(make_dot_lead_), n_orb(n_orb_), time_max(time_max_),
     lesser(std::move(lesser_)), greater(std::move(greater_)) {
  if (n_orb != 1 && n_orb != 2) throw keldy_error("n_orb must be 1 or 2");
  if (lesser.size() != greater.size() || lesser.size() % size_t(n_orb * n_orb) != 0)
    throw keldy_error("g0 tables have inconsistent sizes");
  n_pts = int(lesser.size() / size_t(n_orb * n_orb));
  if (n_pts < 3 || n_pts % 2 == 0) throw keldy_error("g0 tables need an odd number (>= 3) of time points");
}

g0_model::g0_model(g0_model_omega model_omega_, bool make_dot_lead_)
   : model_omega(std::move(model_omega_)), make_dot_lead(make_dot_lead_) {
  const model_param_t p = model_omega.get_param();
  if (make_dot_lead && p.bath_type == "flatband")
    throw keldy_error("The dot-lead g0 for flatband bath is singular and not implemented accurately.");
  n_orb = make_dot_lead ? 2 : 1;
  n_pts = p.nr_time_points_gf;
  time_max = p.time_max;
  if (!(time_max > 0)) throw keldy_error("time_max must be > 0");
  if (p.ft_method == "fft" || p.ft_method == "contour") {
    // Both are computed by quadrature of the frequency integral: "contour"
    // as the reference does (model.cpp:151-165); "fft" (a TRIQS FFT on the
    // adjoint mesh in the reference, model.cpp:179-212) by the same
    // quadrature, which is the limit that FFT approximates.
    if (p.ft_method == "fft")
      std::cout << "WARNING: ft_method 'fft' is computed by quadrature of the frequency integral (the limit the "
                   "FFT approximates); the tables differ from an FFT's by its discretisation error."
                << std::endl;
    const double half_bias = std::abs(p.bias_V / 2);
    if (p.bath_type == "semicircle") {
      if (half_bias < p.half_bandwidth)
        make_g0_by_finite_contour({-p.half_bandwidth, -half_bias, half_bias, p.half_bandwidth});
      else
        make_g0_by_finite_contour({-p.half_bandwidth, p.half_bandwidth});
    } else {
      const double margin = std::abs(std::max(p.Gamma, 1. / p.beta));
      const double left = std::min(-half_bias, p.eps_d) - margin;
      const double right = std::max(half_bias, p.eps_d) + margin;
      make_g0_by_contour(left, right);
    }
  } else if (p.ft_method == "analytic") {
    if (p.bath_type == "flatband")
      make_flat_band_analytic();
    else
      throw keldy_error("analytic ft only works for flatband");
  } else {
    throw keldy_error("ft_method '" + p.ft_method + "' is not defined");
  }
}

void g0_model::fill_negative_times() {
  const int L = n_pts;
  for (int i = 0; i < L; i++) {
    const int ii = L - i - 1;
    if (time(i) < 0) {
      lesser[i] = -std::conj(lesser[ii]);
      greater[i] = -std::conj(greater[ii]);
    }
    if (make_dot_lead) {
      // (0,1)(t) = -conj((1,0)(-t))
      lesser[1 * n_pts + i] = -std::conj(lesser[2 * n_pts + ii]);
      greater[1 * n_pts + i] = -std::conj(greater[2 * n_pts + ii]);
    }
  }
}

// make_flat_band_analytic (model.cpp:428-466):
// g^<(t) = [e^{Gt} Ei(-Gt) - e^{-Gt} Ei(Gt)] / (2 pi) + (i/2) e^{-|Gt|}, g^> = conj(g^<)
void g0_model::make_flat_band_analytic() {
  const model_param_t p = model_omega.get_param();
  if (p.beta != std::numeric_limits<double>::infinity() || p.bias_V != 0. || p.eps_d != 0.)
    throw keldy_error(
        "Analytic flatband covers only the following parameters: beta=infinity (zero temperature), bias_V=0, eps_d=0.");
  if (make_dot_lead) throw keldy_error("No analytic formula for dot-lead Green's functions.");
  allocate();
  for (int k = 0; k < n_pts; k++) {
    const double t = time(k);
    dcomplex val;
    if (t == 0.0) {
      val = 0.5i;
    } else {
      const double Gt = p.Gamma * t;
      const double x = std::abs(Gt);
      const double r = (numeric::exp_E1(x) + numeric::expm_Ei(x)) / (2 * M_PI);
      val = dcomplex(Gt > 0 ? -r : r, 0.5 * std::exp(-x));
    }
    lesser[k] = val;
    greater[k] = std::conj(val);
  }
}

namespace {
constexpr double kAbsTol = 1e-14;
constexpr double kRelTol = 1e-11;
} // namespace

// make_g0_by_contour (model.cpp:326-426) with the three-segment contour of
// contour_integral.hpp:59-82.
void g0_model::make_g0_by_contour(double left_turn_pt, double right_turn_pt) {
  const model_param_t p = model_omega.get_param();
  if (!(left_turn_pt < -std::abs(p.bias_V / 2) && right_turn_pt > std::abs(p.bias_V / 2)))
    throw keldy_error("Contour is wrong regarding Fermi functions poles.");
  if (!(left_turn_pt < right_turn_pt)) throw keldy_error("The order left_turn_pt < right_turn_pt should be respected.");
  allocate();
  const g0_model_omega &mo = model_omega;

  auto contour = [&](const std::function<dcomplex(dcomplex)> &f, dcomplex ldir, dcomplex rdir, double t,
                     double &err) -> dcomplex {
    auto f1 = [&](double x) -> dcomplex {
      const dcomplex w = ldir * x + left_turn_pt;
      return -ldir * std::exp(-1.i * w * t) * f(w);
    };
    auto f2 = [&](double x) -> dcomplex { return std::exp(-1.i * x * t) * f(dcomplex(x, 0.)); };
    auto f3 = [&](double x) -> dcomplex {
      const dcomplex w = right_turn_pt + rdir * x;
      return rdir * std::exp(-1.i * w * t) * f(w);
    };
    auto r1 = numeric::integrate_semi_infinite(f1, kAbsTol, kRelTol);
    auto r2 = numeric::integrate(f2, left_turn_pt, right_turn_pt, kAbsTol, kRelTol);
    auto r3 = numeric::integrate_semi_infinite(f3, kAbsTol, kRelTol);
    err = r1.abserr + r2.abserr + r3.abserr;
    return r1.value + r2.value + r3.value;
  };
  const auto les = [&mo](dcomplex w) { return mo.g0_dot_lesser(w); };
  const auto gre = [&mo](dcomplex w) { return mo.g0_dot_greater(w); };
  const auto les_l = [&mo](dcomplex w) { return mo.g0_rightlead_dot_lesser(w); };
  const auto gre_l = [&mo](dcomplex w) { return mo.g0_rightlead_dot_greater(w); };
  const double inv2pi = 1.0 / (2 * M_PI);

#pragma omp parallel for schedule(dynamic, 8)
  for (int k = 0; k < n_pts; k++) {
    const double t = time(k);
    const int sign_of_t = (t > 0) - (t < 0);
    const dcomplex d1 = (t == 0) ? dcomplex(-1.) : -1.0i * double(sign_of_t);
    dcomplex d2 = std::isinf(p.beta) ? dcomplex(1.) : dcomplex(p.beta, -t);
    d2 /= std::abs(d2);
    double err;
    if (t >= 0) {
      lesser[k] = contour(les, d1, d2, t, err) * inv2pi;
      lesser_ft_error[k] = err * inv2pi;
      greater[k] = contour(gre, -std::conj(d2), -std::conj(d1), t, err) * inv2pi;
      greater_ft_error[k] = err * inv2pi;
    }
    if (make_dot_lead) {
      lesser[2 * n_pts + k] = contour(les_l, d1, d2, t, err) * inv2pi;
      lesser_ft_error[2 * n_pts + k] = err * inv2pi;
      greater[2 * n_pts + k] = contour(gre_l, -std::conj(d2), -std::conj(d1), t, err) * inv2pi;
      greater_ft_error[2 * n_pts + k] = err * inv2pi;
    }
  }
  fill_negative_times();
}

// make_g0_by_finite_contour (model.cpp:220-319): real-axis integral over the
// support [pts.front(), pts.back()] split at the interior points.
void g0_model::make_g0_by_finite_contour(std::vector<double> pts) {
  std::sort(pts.begin(), pts.end());
  pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
  allocate();
  const g0_model_omega &mo = model_omega;
  const double inv2pi = 1.0 / (2 * M_PI);
  auto qagp = [&](const std::function<dcomplex(double)> &f, double &err) -> dcomplex {
    dcomplex s = 0.;
    err = 0.;
    for (size_t i = 0; i + 1 < pts.size(); i++) {
      auto r = numeric::integrate(f, pts[i], pts[i + 1], kAbsTol, kRelTol);
      s += r.value;
      err += r.abserr;
    }
    return s;
  };
#pragma omp parallel for schedule(dynamic, 8)
  for (int k = 0; k < n_pts; k++) {
    const double t = time(k);
    double err;
    if (t >= 0) {
      lesser[k] = qagp([&](double w) { return std::exp(-1.i * w * t) * mo.g0_dot_lesser(dcomplex(w, 0.)); }, err) *
                  inv2pi;
      lesser_ft_error[k] = err * inv2pi;
      greater[k] = qagp([&](double w) { return std::exp(-1.i * w * t) * mo.g0_dot_greater(dcomplex(w, 0.)); }, err) *
                   inv2pi;
      greater_ft_error[k] = err * inv2pi;
    }
    if (make_dot_lead) {
      lesser[2 * n_pts + k] =
          qagp([&](double w) { return std::exp(-1.i * w * t) * mo.g0_rightlead_dot_lesser(dcomplex(w, 0.)); }, err) *
          inv2pi;
      greater[2 * n_pts + k] =
          qagp([&](double w) { return std::exp(-1.i * w * t) * mo.g0_rightlead_dot_greater(dcomplex(w, 0.)); }, err) *
          inv2pi;
    }
  }
  fill_negative_times();
}

dcomplex g0_model::eval(const std::vector<dcomplex> &tab, int oa, int ob, double t) const {
  const dcomplex *g = tab.data() + size_t(oa * n_orb + ob) * n_pts;
  const double a = (t + time_max) / dt();
  long i = long(std::floor(a));
  i = std::min<long>(std::max<long>(i, 0), n_pts - 2);
  const double w = a - double(i);
  return dcomplex((1.0 - w) * g[i].real() + w * g[i + 1].real(), (1.0 - w) * g[i].imag() + w * g[i + 1].imag());
}

// --------------------------------------------------------------------------
dcomplex g0_keldysh_contour_t::operator()(gf_index_t const &a, gf_index_t const &b, bool internal_point) const {
  if (a.spin != b.spin) return 0.0;
  if (a.orbital >= model.n_orb || b.orbital >= model.n_orb)
    throw keldy_error("Need g0_model with make_dot_lead == true to calculate off_diagonal gf.");
  const int order = compare_3way(a.contour, b.contour);
  if (order > 0) return model.eval(model.greater, a.orbital, b.orbital, a.contour.time - b.contour.time);
  if (order < 0) return model.eval(model.lesser, a.orbital, b.orbital, a.contour.time - b.contour.time);
  dcomplex v = model.eval(model.lesser, a.orbital, b.orbital, 0.0);
  if (a.orbital == b.orbital) v -= double(internal_point) * 1.0i * model.model_omega.get_param_alpha();
  return v;
}

} // namespace keldy::impurity_oneband
