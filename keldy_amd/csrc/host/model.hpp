// model.hpp -- single-orbital Anderson impurity: parameters, non-interacting
// Green function g0 in frequency and on the Keldysh contour.
// Reference: c++/keldy/impurity_oneband/model.hpp:37-268, model.cpp:79-466.
#pragma once

#include <functional>
#include <string>
#include <vector>

#include "common.hpp"

namespace keldy::impurity_oneband {

/// Model parameters with the reference's defaults (model.hpp:37-50).
struct model_param_t {
  double beta = 1.0;
  double bias_V = 0.0;
  double eps_d = 0.0;
  double Gamma = 1.0;
  double alpha = 0.0;
  double half_bandwidth = 2.;
  std::string bath_type = "semicircle";
  double time_max = +100.0;
  int nr_time_points_gf = 1000;
  std::string ft_method = "fft"; // or "contour" or "analytic"
};

/// Index of the Keldysh Green function (model.hpp:61-76).
struct gf_index_t {
  contour_pt_t contour{};
  spin_t spin = up;
  orbital_t orbital = 0;
  gf_index_t() = default;
  gf_index_t(time_real_t t, int spin_, int k_idx) : contour{t, keldysh_idx_t(k_idx), 0}, spin(spin_t(spin_)) {}
  gf_index_t(time_real_t t, int spin_, int k_idx, int timesplit)
     : contour{t, keldysh_idx_t(k_idx), timesplit}, spin(spin_t(spin_)) {}
  gf_index_t(time_real_t t, int spin_, int k_idx, int timesplit, int orbital_)
     : contour{t, keldysh_idx_t(k_idx), timesplit}, spin(spin_t(spin_)), orbital(orbital_) {}
};

/// Total order on gf_index_t: contour, then orbital, then spin (model.cpp:63-75).
bool operator<(gf_index_t const &a, gf_index_t const &b);

/// g0 in frequency from the bath hybridisation (model.hpp:94-174, model.cpp:79-123).
class g0_model_omega {
 public:
  g0_model_omega() = default;
  explicit g0_model_omega(model_param_t const &parameters);

  model_param_t get_param() const { return param_; }
  double get_param_alpha() const { return param_.alpha; }
  double mu_left() const { return -param_.bias_V / 2; }
  double mu_right() const { return +param_.bias_V / 2; }

  dcomplex g0_dot_R(dcomplex omega) const;
  dcomplex g0_dot_A(dcomplex omega) const;
  dcomplex g0_dot_lesser(dcomplex omega) const { return g0_dot(omega, true); }
  dcomplex g0_dot_greater(dcomplex omega) const { return g0_dot(omega, false); }
  dcomplex g0_dot_keldysh(keldysh_idx_t a, keldysh_idx_t b, dcomplex omega) const;
  dcomplex bath_hybrid_R_left(dcomplex omega) const { return hyb_(omega); }
  dcomplex bath_hybrid_A_left(dcomplex omega) const { return std::conj(hyb_(std::conj(omega))); }
  dcomplex bath_hybrid_K_left(dcomplex omega) const;
  dcomplex bath_hybrid_R_right(dcomplex omega) const { return hyb_(omega); }
  dcomplex bath_hybrid_A_right(dcomplex omega) const { return std::conj(hyb_(std::conj(omega))); }
  dcomplex bath_hybrid_K_right(dcomplex omega) const;
  dcomplex g0_rightlead_dot_lesser(dcomplex omega) const { return g0_rightlead_dot(omega, true); }
  dcomplex g0_rightlead_dot_greater(dcomplex omega) const { return g0_rightlead_dot(omega, false); }

 private:
  dcomplex g0_dot(dcomplex omega, bool is_lesser) const;
  dcomplex g0_rightlead_dot(dcomplex omega, bool is_lesser) const;
  model_param_t param_{};
  std::function<dcomplex(dcomplex)> hyb_ = [](dcomplex) { return dcomplex(0.); };
};

/// g^<(t), g^>(t) tabulated on the mesh t_k = -time_max + k dt, k < n_pts
/// (n_pts odd), as [n_orb][n_orb][n_pts] complex arrays; n_orb = 2 with the
/// dot-lead components (model.cpp:139-466).  Spin up and down are identical.
class g0_model {
 public:
  g0_model_omega model_omega{};
  bool make_dot_lead = false;
  int n_orb = 1;
  int n_pts = 0;
  double time_max = 0.0;
  std::vector<dcomplex> lesser, greater;
  std::vector<dcomplex> lesser_ft_error, greater_ft_error;

  g0_model() = default;
  g0_model(g0_model_omega model_omega_, bool make_dot_lead_);
  /// A model from externally computed tables (same layout).
  g0_model(g0_model_omega model_omega_, bool make_dot_lead_, int n_orb_, double time_max_,
           std::vector<dcomplex> lesser_, std::vector<dcomplex> greater_);

  static std::string hdf5_format() { return "KELDY_G0Model"; }
  double dt() const { return 2.0 * time_max / double(n_pts - 1); }
  double time(int k) const { return -time_max + k * dt(); }
  /// Linear interpolation of a table component at time t.
  dcomplex eval(const std::vector<dcomplex> &tab, int oa, int ob, double t) const;

 private:
  void make_flat_band_analytic();
  void make_g0_by_contour(double left_turn_pt, double right_turn_pt);
  void make_g0_by_finite_contour(std::vector<double> pts);
  void allocate();
  void fill_negative_times();
};

/// g0 on the Keldysh contour (model.hpp:238-268).
struct g0_keldysh_contour_t {
  g0_model model;
  g0_keldysh_contour_t() = default;
  explicit g0_keldysh_contour_t(g0_model model_) : model(std::move(model_)) {}
  dcomplex operator()(gf_index_t const &a, gf_index_t const &b, bool internal_point = true) const;
  double get_time_max() const { return model.time_max; }
};

} // namespace keldy::impurity_oneband
