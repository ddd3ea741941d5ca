// warpers.hpp -- coordinate maps l in [0,1]^n -> u (importance sampling).
// Reference: c++/keldy/warpers/{identity,plasma_uv,product_1d_simple,product_1d,
// make_warper_from_proj,warpers}.hpp.  Each warper evaluates on the host for the
// per-point API and exports a device descriptor (kd_warper_desc / kd_map1d)
// that the QMC kernel applies in registers.
#pragma once

#include <atomic>
#include <functional>
#include <memory>
#include <utility>
#include <variant>
#include <vector>

#include "../../../include/keldy_amd.h"
#include "common.hpp"

namespace keldy::warpers {

using vec = std::vector<double>;
using map_pair = std::pair<vec, double>;

/// Device form of a warper chain.
struct device_chain {
  std::vector<kd_warper_desc> warpers;
  std::vector<kd_map1d> maps;
  std::vector<double> tables;
  // learned flow warper (last in the chain), if any
  int flow_order = 0, flow_layers = 0, flow_nc = 0;
  double flow_cond = 0.0;
  std::vector<float> flow_params;
  uint64_t flow_stamp = 0;
};

/// One-dimensional inverse-CDF map u = F^{-1}(l) with density f = F'.
class map1d {
 public:
  enum kind_t { uniform, exponential, inverse, inverse_square, inverse_power, callable, nearest, hybrid };

  static map1d make_uniform(double u_max);
  static map1d make_exponential(double u_max, double w);
  static map1d make_inverse(double u_max, double w);
  static map1d make_inverse_square(double u_max, double w);
  static map1d make_inverse_power(int power, double u_max, double w);
  static map1d make_callable(std::function<double(double)> f1, std::function<double(double)> F,
                             std::function<double(double)> Finv, double u_max, bool do_domain_checks);
  static map1d make_nearest(vec times, vec f1);
  static map1d make_nearest_from_function(const std::function<double(double)> &f1, double u_max, int n);
  /// hybrid: Simpson-integrated CDF with Steffen splines; power > 0 records that
  /// f1(t) = w^power / (w + t)^power (needed on the device)
  static map1d make_hybrid(std::function<double(double)> f1, double u_max, int n, int power = 0, double w = 0);

  double f1(double u) const;            // density in u
  double F(double u) const;             // l from u
  double Finv(double l) const;          // u from l
  double normalization() const { return norm_; }
  double u_max() const { return u_max_; }
  kind_t kind() const { return kind_; }
  bool device_ok() const;
  kd_map1d to_device(std::vector<double> &tables) const;
  // tabulated maps: knot times, normalised f1 knots, normalised negated CDF knots
  vec const &knot_times() const { return t_; }
  vec const &knot_f1() const { return f_; }
  vec const &knot_F() const { return Fk_; }
  int cut_index() const { return cut_; }

 private:
  kind_t kind_ = uniform;
  int power_ = 1;
  double u_max_ = 1.0, w_ = 1.0;
  long double norm_ld_ = 1.0L; // analytic makers' normalisation (long double, as the reference)
  double norm_ = 1.0;          // f1_integrated_normalization
  std::function<double(double)> f1_fn_, F_fn_, Finv_fn_;
  // tabulated maps
  vec t_, f_, Fk_;
  int cut_ = 0;
  vec inv_coef_, fwd_coef_; // Steffen coefficients (hybrid)
};

class warper_identity_t {
 public:
  map_pair map_reverse(vec const &l) const { return {l, 1.0}; }
  map_pair map_forward(vec const &u) const { return {u, 1.0}; }
  vec ui_from_li(vec const &l) const { return l; }
  vec li_from_ui(vec const &u) const { return u; }
  double jacobian_reverse(vec const &) const { return 1.0; }
  double jacobian_forward(vec const &) const { return 1.0; }
  void to_device(device_chain &c) const;
};

vec vi_from_ui(double t_max, vec const &u_times);
vec ui_from_vi(double t_max, vec const &v_times);

class warper_plasma_uv_t {
 public:
  double t_max = 0.0;
  warper_plasma_uv_t() = default;
  explicit warper_plasma_uv_t(double t) : t_max(t) {}
  map_pair map_reverse(vec const &l) const { return {ui_from_li(l), 1.0}; }
  map_pair map_forward(vec const &u) const { return {li_from_ui(u), 1.0}; }
  vec ui_from_li(vec const &l) const { return ui_from_vi(t_max, l); }
  vec li_from_ui(vec const &u) const { return vi_from_ui(t_max, u); }
  double jacobian_reverse(vec const &) const { return 1.0; }
  double jacobian_forward(vec const &) const { return 1.0; }
  void to_device(device_chain &c) const;
};

/// The same 1D map on every axis (product_1d_simple.hpp).
class warper_product_1d_simple_base {
 public:
  map1d m;
  warper_product_1d_simple_base() : m(map1d::make_uniform(1.0)) {}
  explicit warper_product_1d_simple_base(map1d mm) : m(std::move(mm)) {}
  map_pair map_reverse(vec const &l) const;
  map_pair map_forward(vec const &u) const;
  vec ui_from_li(vec const &l) const;
  vec li_from_ui(vec const &u) const;
  double jacobian_reverse(vec const &l) const;
  double jacobian_forward(vec const &u) const;
  double operator()(vec const &u) const; // product of normalised densities
  void to_device(device_chain &c) const;
};

struct warper_product_1d_simple_t : warper_product_1d_simple_base {
  using warper_product_1d_simple_base::warper_product_1d_simple_base;
  warper_product_1d_simple_t() = default;
  warper_product_1d_simple_t(std::function<double(double)> f1, std::function<double(double)> F,
                             std::function<double(double)> Finv, double u_max, bool do_domain_checks = true)
     : warper_product_1d_simple_base(map1d::make_callable(std::move(f1), std::move(F), std::move(Finv), u_max,
                                                          do_domain_checks)) {}
};

struct warper_product_1d_simple_interp_nearest_t : warper_product_1d_simple_base {
  using warper_product_1d_simple_base::warper_product_1d_simple_base;
  warper_product_1d_simple_interp_nearest_t() = default;
  warper_product_1d_simple_interp_nearest_t(vec times, vec f1)
     : warper_product_1d_simple_base(map1d::make_nearest(std::move(times), std::move(f1))) {}
  warper_product_1d_simple_interp_nearest_t(std::function<double(double)> const &f1, double u_max, int n)
     : warper_product_1d_simple_base(map1d::make_nearest_from_function(f1, u_max, n)) {}
};

struct warper_product_1d_simple_interp_hybrid_t : warper_product_1d_simple_base {
  using warper_product_1d_simple_base::warper_product_1d_simple_base;
  warper_product_1d_simple_interp_hybrid_t() = default;
  warper_product_1d_simple_interp_hybrid_t(std::function<double(double)> f1, double u_max, int n)
     : warper_product_1d_simple_base(map1d::make_hybrid(std::move(f1), u_max, n)) {}
};

/// One 1D map per axis (product_1d.hpp:123-252).
template <typename W> class warper_product_1d_base {
 public:
  std::vector<W> warpers_dims;
  void reserve(size_t n) { warpers_dims.reserve(n); }
  int size() const { return int(warpers_dims.size()); }
  void emplace_back(W w) { warpers_dims.push_back(std::move(w)); }
  map_pair map_reverse(vec const &l) const {
    vec x = l;
    double j = 1.0;
    for (size_t i = 0; i < x.size(); i++) {
      auto r = dim(i).map_reverse(vec{x[i]});
      x[i] = r.first.at(0);
      j *= r.second;
    }
    return {x, j};
  }
  map_pair map_forward(vec const &u) const {
    vec x = u;
    double j = 1.0;
    for (size_t i = 0; i < x.size(); i++) {
      auto r = dim(i).map_forward(vec{x[i]});
      x[i] = r.first.at(0);
      j *= r.second;
    }
    return {x, j};
  }
  vec ui_from_li(vec const &l) const { return map_reverse(l).first; }
  vec li_from_ui(vec const &u) const { return map_forward(u).first; }
  double jacobian_reverse(vec const &l) const {
    double j = 1.0;
    for (size_t i = 0; i < l.size(); i++) j *= dim(i).jacobian_reverse(vec{l[i]});
    return j;
  }
  double jacobian_forward(vec const &u) const {
    double j = 1.0;
    for (size_t i = 0; i < u.size(); i++) j *= dim(i).jacobian_forward(vec{u[i]});
    return j;
  }
  double operator()(vec const &u) const {
    double r = 1.0;
    for (size_t i = 0; i < u.size(); i++) r *= dim(i)(vec{u[i]});
    return r;
  }
  W get_1d_warper(int axis) const {
    if (axis < 0 || axis >= size()) throw keldy_error("out of range axis");
    return warpers_dims[axis];
  }
  void to_device(device_chain &c) const {
    if (warpers_dims.empty()) throw keldy_error("empty product warper");
    kd_warper_desc d{};
    d.kind = KD_WARP_PRODUCT;
    d.n_maps = int32_t(warpers_dims.size());
    d.map_offset = int64_t(c.maps.size());
    for (auto const &w : warpers_dims) c.maps.push_back(w.m.to_device(c.tables));
    c.warpers.push_back(d);
  }

 private:
  W const &dim(size_t i) const {
    if (i >= warpers_dims.size()) throw keldy_error("product warper has fewer axes than the point");
    return warpers_dims[i];
  }
};

using warper_product_1d_t = warper_product_1d_base<warper_product_1d_simple_t>;
using warper_product_1d_interp_nearest_t = warper_product_1d_base<warper_product_1d_simple_interp_nearest_t>;
using warper_product_1d_interp_hybrid_t = warper_product_1d_base<warper_product_1d_simple_interp_hybrid_t>;

// makers (product_1d_simple.hpp:560-630, product_1d.hpp:256-310)
warper_product_1d_simple_t make_product_1d_simple_exponential(double domain_u_max, double w_scale);
warper_product_1d_simple_t make_product_1d_simple_inverse(double domain_u_max, double w_scale);
warper_product_1d_simple_t make_product_1d_simple_inverse_square(double domain_u_max, double w_scale);
warper_product_1d_simple_t make_product_1d_simple_inverse_power(int power, double domain_u_max, double w_scale);
/// uniform density 1/u_max on [0, u_max] (the analytic form of the reference's
/// identity-on-[0,t_max] test warper, product_1d_simple_test.cpp:10-23)
warper_product_1d_simple_t make_product_1d_simple_uniform(double domain_u_max);
warper_product_1d_t make_product_1d_inverse_cube_alternate(int order, double time, double warper_scale);
warper_product_1d_interp_nearest_t make_product_1d_inverse_cube_alternate_interp_nearest(int order, double time,
                                                                                       double warper_scale, int n);
warper_product_1d_interp_hybrid_t make_product_1d_inverse_cube_alternate_interp_hybrid(int order, double time,
                                                                                     double warper_scale, int n);

/// Pilot fit (make_warper_from_proj.hpp:52-153): per-axis histograms of |f| at
/// nr_samples unshifted Sobol points, log-space local-linear kernel smoothing,
/// one nearest-interpolated 1D map per axis on [0, 1].
warper_product_1d_interp_nearest_t make_warper_from_proj(std::function<dcomplex(vec)> const &warped_integrand,
                                                         int order, int num_bins, int nr_samples, double sigma);
/// Same fit from precomputed samples: l [nr_samples][order] (the first
/// nr_samples unshifted Sobol points) and |f| at them.
warper_product_1d_interp_nearest_t make_warper_from_proj_samples(std::vector<vec> const &l, vec const &abs_values,
                                                                 int order, int num_bins, double sigma);
/// The fit from per-axis histograms: sums of |f| and sample counts per bin of
/// [0, 1] (the device computes them, kd_pilot_histograms).
warper_product_1d_interp_nearest_t make_warper_from_proj_histograms(std::vector<vec> sums,
                                                                    std::vector<std::vector<long>> const &counts,
                                                                    int order, int num_bins, double sigma);
namespace proj_details {
vec kernel_smoothing(vec const &y_in, double sigma);
/// weighted least squares y = c0 + c1 x (weights <= 0 skipped), running means
std::pair<double, double> fit_wlinear(const double *x, const double *w, const double *y, size_t n);
} // namespace proj_details

/// Projection warper (projection.hpp:176-382): per axis, the mean |f| of the
/// pilot samples in num_bins bins of [-d/2, 1+d/2] (d = 1/(num_bins-1)),
/// smoothed by local-linear regression in log space (width sigma, optionally
/// chosen by Brent minimisation of the leave-one-out error); fn on the uniform
/// mesh of [0, 1] (num_bins points), its midpoint-rule CDF normalised to 1, and
/// the CDF's inverse by Akima interpolation tabulated on 1 + 5(num_bins-1)
/// points.  The maps evaluate these tables by linear interpolation on their
/// meshes: u = Finv(l), J = norm / fn(u) per axis.
class warper_projection_t {
 public:
  warper_projection_t() = default;
  /// pilot: |warped_integrand| at the first nr_samples unshifted Sobol points
  warper_projection_t(std::function<dcomplex(vec)> const &warped_integrand, int order, int num_bins, int nr_samples,
                      double sigma, bool optimize_sigma = true);
  /// the same construction from given pilot points l [n][order] and |f| at them
  static warper_projection_t from_samples(std::vector<vec> const &l, vec const &abs_values, int order, int num_bins,
                                          double sigma, bool optimize_sigma = true);
  vec get_xi(int axis) const;                      // per-bin mean |f| of the pilot
  std::vector<long> get_xi_counts(int axis) const; // pilot samples per bin
  vec get_xi_bin_coord(int axis) const;            // bin centres
  vec get_sigmas() const;                          // smoothing widths per axis (natural units)
  vec get_fi(int axis) const;                      // fn on the mesh of [0, 1]
  vec get_fi_integrated(int axis) const;           // normalised CDF on the same mesh
  vec get_fi_integrated_inverse(int axis) const;   // its inverse on the finer mesh
  vec ui_from_li(vec const &l) const;
  vec li_from_ui(vec const &u) const;
  double jacobian_reverse(vec const &l) const;
  double jacobian_forward(vec const &u) const;
  map_pair map_reverse(vec const &l) const { return {ui_from_li(l), jacobian_reverse(l)}; }
  map_pair map_forward(vec const &u) const { return {li_from_ui(u), jacobian_forward(u)}; }
  void to_device(device_chain &c) const;
  int order() const { return order_; }
  int num_bins() const { return num_bins_; }

 private:
  void build(std::vector<vec> const &l, vec const &abs_values, double sigma, bool optimize_sigma);
  void check_axis(int axis) const;
  int order_ = 0, num_bins_ = 0;
  double bin_size_ = 0.0;
  std::vector<vec> xi_, fn_, F_, Finv_;
  std::vector<std::vector<long>> xi_count_;
  vec norm_, sigmas_; // sigmas_ in bin units
};

namespace proj_details {
/// the leave-one-out smoothing width of projection.hpp:80-121 (bin units):
/// Brent minimisation on [-upper, upper] from sigma, |result| >= 0.3
double optimize_sigma(vec const &y, double sigma, double upper);
/// Akima spline through (x, y) (GSL's non-periodic form): per interval b, c, d
void akima_coeffs(vec const &x, vec const &y, vec &b, vec &c, vec &d);
double akima_eval(vec const &x, vec const &y, vec const &b, vec const &c, vec const &d, double xv);
} // namespace proj_details

inline uint64_t next_stamp() {
  static std::atomic<uint64_t> counter{1};
  return counter.fetch_add(1);
}

/// Learned affine-coupling flow on (0,1)^n: an importance map trained on
/// |integrand| samples (no counterpart in keldy).  Map, Jacobian and parameter
/// layout are those of the device kernels (kd_flow_device.hpp): y = logit(l),
/// coupling layers y_B <- y_B exp(s(y_A)) + t(y_A), x = sigmoid(y).  The host
/// methods evaluate the same network (bf16-rounded operands where the device
/// uses bf16 MFMA, FP64 accumulation); runs use the device kernels.  Must be
/// the last warper of a train (it is applied first to the Sobol point).
/// With n_cond = 1 the flow is conditioned on one value (the time-resolved
/// runs: a scaled external time): the map is T_cond, one member of the family
/// the parameters hold; with_cond gives the same parameters at another value.
class warper_flow_t {
 public:
  warper_flow_t() = default;
  warper_flow_t(int order, int n_layers, std::vector<float> params, int n_cond = 0, double cond = 0.0);
  warper_flow_t with_cond(double cond) const { return warper_flow_t(order_, layers_, params_, nc_, cond); }
  map_pair map_reverse(vec const &l) const; // l -> x, dx/dl
  map_pair map_forward(vec const &x) const; // x -> l, dl/dx
  vec ui_from_li(vec const &l) const { return map_reverse(l).first; }
  vec li_from_ui(vec const &u) const { return map_forward(u).first; }
  double jacobian_reverse(vec const &l) const { return map_reverse(l).second; }
  double jacobian_forward(vec const &u) const { return map_forward(u).second; }
  void to_device(device_chain &c) const;
  int order() const { return order_; }
  int n_layers() const { return layers_; }
  int n_cond() const { return nc_; }
  double cond() const { return cond_; }
  std::vector<float> const &params() const { return params_; }
  uint64_t stamp() const { return stamp_; }

 private:
  void net(int k, const double *y, double *s, double *t) const;
  int order_ = 0, layers_ = 0, nc_ = 0;
  double cond_ = 0.0;
  std::vector<float> params_;
  uint64_t stamp_ = 0;
};

using warper_variant =
    std::variant<warper_identity_t, warper_plasma_uv_t, warper_product_1d_simple_interp_nearest_t,
                 warper_product_1d_simple_t, warper_product_1d_t, warper_product_1d_interp_nearest_t,
                 warper_product_1d_interp_hybrid_t, warper_product_1d_simple_interp_hybrid_t, warper_projection_t,
                 warper_flow_t>;

/// Ordered chain of warpers; map_reverse applies them last-added first
/// (warpers.hpp:41-137).

class warper_train_t {
 public:
  std::vector<warper_variant> warpers;
  template <typename W> void emplace_back(W w) {
    warpers.emplace_back(std::move(w));
    stamp = next_stamp();
  }
  void reserve(size_t n) { warpers.reserve(n); }
  void clear() {
    warpers.clear();
    stamp = next_stamp();
  }
  size_t size() const { return warpers.size(); }
  map_pair map_reverse(vec const &l, int start_domain_nr, int end_domain_nr) const;
  map_pair map_forward(vec const &u, int start_domain_nr, int end_domain_nr) const;
  map_pair map_reverse(vec const &l) const { return map_reverse(l, int(size()), 0); }
  map_pair map_forward(vec const &u) const { return map_forward(u, 0, int(size())); }
  double jacobian_reverse(vec const &l) const { return map_reverse(l).second; }
  double jacobian_forward(vec const &u) const { return map_forward(u).second; }
  vec ui_from_li(vec const &l) const;
  vec li_from_ui(vec const &u) const;
  device_chain to_device() const;
  // Unique per content change across all trains; the integrator re-uploads
  // the device chain when the stamp of its train differs from the uploaded one.
  uint64_t stamp = next_stamp();
};

} // namespace keldy::warpers
