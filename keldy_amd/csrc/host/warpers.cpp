// warpers.cpp -- host side of the warpers (reference: c++/keldy/warpers/*.hpp).
#include "warpers.hpp"

#include "../kernels/kd_flow_device.hpp"

#include <algorithm>
#include <cmath>
#include <sstream>

#include "qrng.hpp"

namespace keldy::warpers {

namespace {

// gsl_interp_bsearch over [0, n-1]: largest i in [0, n-2] with xa[i] <= x
size_t knot_find(const double *xa, size_t n, double x) {
  size_t ilo = 0, ihi = n - 1;
  while (ihi > ilo + 1) {
    const size_t i = (ihi + ilo) / 2;
    if (xa[i] > x)
      ihi = i;
    else
      ilo = i;
  }
  return ilo;
}

void check_domain(const double *xa, size_t n, double x) {
  if (x < xa[0] || x > xa[n - 1]) throw keldy_error("interpolation error: point outside the tabulated domain");
}

// linear spline (gsl_interp_linear) through (xa, ya)
double linear_eval(const double *xa, const double *ya, size_t n, double x) {
  check_domain(xa, n, x);
  const size_t i = knot_find(xa, n, x);
  const double dx = xa[i + 1] - xa[i];
  if (!(dx > 0.0)) return 0.0;
  return ya[i] + (x - xa[i]) / dx * (ya[i + 1] - ya[i]);
}

double steffen_sign(double x, double y) { return ((x < 0 && y > 0) || (x > 0 && y < 0)) ? -x : x; }

// Steffen (1990) monotone cubic: slopes from eq. (11), one-sided end slopes,
// coefficients (a, b, c, d) per interval, y = d + dx (c + dx (b + dx a)).
vec steffen_coeffs(const vec &x, const vec &y, size_t n) {
  vec yp(n), co(4 * (n - 1));
  yp[0] = (y[1] - y[0]) / (x[1] - x[0]);
  for (size_t i = 1; i + 1 < n; i++) {
    const double hi = x[i + 1] - x[i], him1 = x[i] - x[i - 1];
    const double si = (y[i + 1] - y[i]) / hi, sim1 = (y[i] - y[i - 1]) / him1;
    const double pi = (sim1 * hi + si * him1) / (him1 + hi);
    yp[i] = (steffen_sign(1.0, sim1) + steffen_sign(1.0, si)) *
            std::min(std::abs(sim1), std::min(std::abs(si), 0.5 * std::abs(pi)));
  }
  yp[n - 1] = (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]);
  for (size_t i = 0; i + 1 < n; i++) {
    const double hi = x[i + 1] - x[i];
    const double si = (y[i + 1] - y[i]) / hi;
    co[4 * i + 0] = (yp[i] + yp[i + 1] - 2 * si) / hi / hi;
    co[4 * i + 1] = (3 * si - 2 * yp[i] - yp[i + 1]) / hi;
    co[4 * i + 2] = yp[i];
    co[4 * i + 3] = y[i];
  }
  return co;
}

double steffen_eval(const double *x, const double *co, size_t n, double xv) {
  check_domain(x, n, xv);
  const size_t i = knot_find(x, n, xv);
  const double d = xv - x[i];
  return co[4 * i + 3] + d * (co[4 * i + 2] + d * (co[4 * i + 1] + d * co[4 * i + 0]));
}

} // namespace

// --------------------------------------------------------------------------
// map1d
map1d map1d::make_uniform(double u_max) {
  if (!(0.0 < u_max)) throw keldy_error("coord_domain_u coordinates must be ordered.");
  map1d m;
  m.kind_ = uniform;
  m.u_max_ = u_max;
  return m;
}

map1d map1d::make_exponential(double u_max, double w) {
  map1d m;
  m.kind_ = exponential;
  m.u_max_ = u_max;
  m.w_ = w;
  m.norm_ld_ = -std::expm1(-static_cast<long double>(u_max) / w);
  return m;
}

map1d map1d::make_inverse(double u_max, double w) {
  map1d m;
  m.kind_ = inverse;
  m.u_max_ = u_max;
  m.w_ = w;
  m.norm_ld_ = std::log1p(static_cast<long double>(u_max) / w);
  return m;
}

map1d map1d::make_inverse_square(double u_max, double w) {
  map1d m;
  m.kind_ = inverse_square;
  m.u_max_ = u_max;
  m.w_ = w;
  m.norm_ld_ = static_cast<long double>(u_max) / (w + u_max);
  return m;
}

map1d map1d::make_inverse_power(int power, double u_max, double w) {
  if (power == 1) return make_inverse(u_max, w);
  map1d m;
  m.kind_ = inverse_power;
  m.power_ = power;
  m.u_max_ = u_max;
  m.w_ = w;
  const long double um = u_max;
  m.norm_ld_ = (power - 1) * std::pow(um + w, power - 1) / (std::pow(um / w + 1, power - 1) - 1);
  return m;
}

map1d map1d::make_callable(std::function<double(double)> f1, std::function<double(double)> F,
                           std::function<double(double)> Finv, double u_max, bool do_domain_checks) {
  map1d m;
  m.kind_ = callable;
  m.u_max_ = u_max;
  m.f1_fn_ = std::move(f1);
  m.F_fn_ = std::move(F);
  m.Finv_fn_ = std::move(Finv);
  if (!(0.0 < u_max)) throw keldy_error("coord_domain_u coordinates must be ordered.");
  // product_1d_simple.hpp:74-104
  if (do_domain_checks) {
    const double eps = 1e-12;
    const bool ok = std::abs(m.F_fn_(0.0) - 0.0) < eps && std::abs(m.F_fn_(u_max) - 1.0) < eps &&
                    std::abs(m.Finv_fn_(0.0) - 0.0) < eps && std::abs(m.Finv_fn_(1.0) - u_max) < eps;
    if (!ok) throw keldy_error("f1_integrated and f1_integrated_inverse do not map domain boundaries correctly.");
    for (double lp : {0., 0.125, 0.25, 0.375, 0.5, 0.625, 0.75, 0.875, 1.}) {
      if (std::abs(m.F_fn_(m.Finv_fn_(lp)) - lp) > 1e-12) {
        std::ostringstream os;
        os << "Inconsistent functions: f1_integrated_inverse should be the inverse of f1_integrated ("
           << m.F_fn_(m.Finv_fn_(lp)) << " != " << lp << ")";
        throw keldy_error(os.str());
      }
    }
  }
  return m;
}

// warper_product_1d_simple_interp_nearest_t (product_1d_simple.hpp:219-287)
map1d map1d::make_nearest(vec times, vec f1) {
  const size_t n = times.size();
  if (f1.size() != n) throw keldy_error("times_u_pts and f1_pts should have the same length.");
  if (n < 3) throw keldy_error("Expect nr_sample_points to be >= 3. Given was " + std::to_string(n));
  if (times[0] != 0.) throw keldy_error("times_u_pts should start at 0");
  map1d m;
  m.kind_ = nearest;
  m.u_max_ = times[n - 1];
  vec F(n, 0.0);
  size_t cut = n - 1;
  bool found = false;
  for (size_t ii = n - 1; ii-- > 0;) {
    F[ii] = F[ii + 1] - 0.5 * (f1[ii + 1] + f1[ii]) * (times[ii + 1] - times[ii]);
    found = found || (F[ii] != 0);
    if (!found) cut = ii;
  }
  if (!found) throw keldy_error("Model function is flat zero.");
  const double norm = -F[0];
  for (auto &v : F) v /= norm;
  for (auto &v : f1) v /= norm;
  F[0] = -1;
  m.norm_ = norm;
  m.t_ = std::move(times);
  m.f_ = std::move(f1);
  m.Fk_ = std::move(F);
  m.cut_ = int(cut);
  return m;
}

map1d map1d::make_nearest_from_function(const std::function<double(double)> &f1, double u_max, int n) {
  if (n < 3) throw keldy_error("Expect nr_sample_points to be >= 3. Given was " + std::to_string(n));
  vec t(n), f(n);
  for (int i = 0; i < n; i++) t[i] = static_cast<double>(i) / (n - 1) * u_max;
  t[n - 1] = u_max;
  for (int i = 0; i < n; i++) f[i] = f1(t[i]);
  return make_nearest(std::move(t), std::move(f));
}

// warper_product_1d_simple_interp_hybrid_t (product_1d_simple.hpp:424-491)
map1d map1d::make_hybrid(std::function<double(double)> f1, double u_max, int n, int power, double w) {
  if (n < 3) throw keldy_error("Expect nr_sample_points to be >= 3. Given was " + std::to_string(n));
  map1d m;
  m.kind_ = hybrid;
  m.u_max_ = u_max;
  m.power_ = power;
  m.w_ = w;
  const double delta = u_max / (n - 1);
  vec t(n), F(n, 0.0);
  for (int i = 0; i < n; i++) t[i] = delta * i;
  t[n - 1] = u_max;
  int cut = n - 1;
  bool found = false;
  double f2d = f1(t[n - 1]);
  for (int i = n - 2; i >= 0; i--) {
    const double f0d = f2d;
    const double f1d = f1(0.5 * (t[i + 1] + t[i]));
    f2d = f1(t[i]);
    const double dl = delta / 6. * (f0d + 4 * f1d + f2d);
    F[i] = F[i + 1] - dl;
    if (dl < 1e-100)
      throw keldy_error("dl=" + std::to_string(dl) + " < 1e-100 is too small for a cubic interpolator in double precision.");
    found = found || (F[i] != 0);
    if (!found) cut = i;
  }
  if (!found) throw keldy_error("Model function is flat zero.");
  const double norm = -F[0];
  for (auto &v : F) v /= norm;
  F[0] = -1;
  m.norm_ = norm;
  m.f1_fn_ = std::move(f1);
  m.t_ = std::move(t);
  m.Fk_ = std::move(F);
  m.cut_ = cut;
  m.fwd_coef_ = steffen_coeffs(m.t_, m.Fk_, size_t(n));
  vec Fi(m.Fk_.begin(), m.Fk_.begin() + cut + 1), ti(m.t_.begin(), m.t_.begin() + cut + 1);
  m.inv_coef_ = steffen_coeffs(Fi, ti, size_t(cut + 1));
  return m;
}

double map1d::f1(double u) const {
  const double w = w_;
  switch (kind_) {
  case uniform: return 1.0 / u_max_;
  case exponential: return double(std::exp(-u / w) / (w * norm_ld_));
  case inverse: return double(1.0 / (norm_ld_ * (w + u)));
  case inverse_square: return double(w / ((w + u) * (w + u) * norm_ld_));
  case inverse_power: return double(norm_ld_ / std::pow(w + u, power_));
  case callable: return f1_fn_(u);
  case nearest: {
    const size_t i = knot_find(t_.data(), t_.size(), u);
    if (i == t_.size() - 1) return f_[i];
    return 0.5 * (f_[i] + f_[i + 1]);
  }
  case hybrid: return f1_fn_(u) / norm_;
  }
  return 0.0;
}

double map1d::F(double u) const {
  const double w = w_, um = u_max_;
  switch (kind_) {
  case uniform: return u / um;
  case exponential: return std::expm1((um - u) / w) / std::expm1(um / w);
  case inverse: return double(std::log1p(u / w) / norm_ld_);
  case inverse_square: return double(u / ((w + u) * norm_ld_));
  case inverse_power: {
    const long double umax = um;
    const long double n = std::pow(umax + w, power_ - 1) / (std::pow(umax / w + 1, power_ - 1) - 1);
    return double(n * (std::pow(u / w + 1, power_ - 1) - 1) / std::pow(u + w, power_ - 1));
  }
  case callable: return F_fn_(u);
  case nearest: return -linear_eval(t_.data(), Fk_.data(), t_.size(), u);
  case hybrid: return -steffen_eval(t_.data(), fwd_coef_.data(), t_.size(), u);
  }
  return 0.0;
}

double map1d::Finv(double l) const {
  const double w = w_, um = u_max_;
  switch (kind_) {
  case uniform: return um * l;
  case exponential: return um - w * std::log1p(std::expm1(um / w) * l);
  case inverse: return double(w * std::expm1(l * norm_ld_));
  case inverse_square: return double(norm_ld_ * l * w / (1.0 - norm_ld_ * l));
  case inverse_power: {
    const long double umax = um;
    const long double denom =
        std::pow(1 + l * (std::pow(w / (umax + w), power_ - 1) - 1), 1.0 / (power_ - 1.0));
    return double(w * (1 - denom) / denom);
  }
  case callable: return Finv_fn_(l);
  case nearest: return linear_eval(Fk_.data(), t_.data(), size_t(cut_ + 1), -l);
  case hybrid: return steffen_eval(Fk_.data(), inv_coef_.data(), size_t(cut_ + 1), -l);
  }
  return 0.0;
}

bool map1d::device_ok() const {
  if (kind_ == callable) return false;
  if (kind_ == hybrid) return power_ > 0;
  return true;
}

kd_map1d map1d::to_device(std::vector<double> &tables) const {
  if (!device_ok())
    throw keldy_error(
        "this warper is defined by host callables and has no device form; use the makers "
        "(make_product_1d_simple_*, interp_nearest/interp_hybrid tables) to run it on the GPU");
  kd_map1d d{};
  d.power = power_;
  d.u_max = u_max_;
  d.w = w_;
  d.norm = double(norm_ld_);
  switch (kind_) {
  case uniform: d.kind = KD_MAP_UNIFORM; break;
  case exponential: d.kind = KD_MAP_EXPONENTIAL; break;
  case inverse: d.kind = KD_MAP_INVERSE; break;
  case inverse_square: d.kind = KD_MAP_INVERSE_SQ; break;
  case inverse_power: d.kind = KD_MAP_INVERSE_POW; break;
  case nearest:
  case hybrid: {
    d.kind = kind_ == nearest ? KD_MAP_NEAREST : KD_MAP_HYBRID;
    d.n_pts = int32_t(t_.size());
    d.n_inv = int32_t(cut_ + 1);
    d.t_off = int64_t(tables.size());
    tables.insert(tables.end(), t_.begin(), t_.end());
    d.F_off = int64_t(tables.size());
    tables.insert(tables.end(), Fk_.begin(), Fk_.end());
    if (kind_ == nearest) {
      d.f_off = int64_t(tables.size());
      tables.insert(tables.end(), f_.begin(), f_.end());
    } else {
      d.c_off = int64_t(tables.size());
      tables.insert(tables.end(), inv_coef_.begin(), inv_coef_.end());
      d.norm = norm_;
    }
    break;
  }
  default: break;
  }
  return d;
}

// --------------------------------------------------------------------------
void warper_identity_t::to_device(device_chain &c) const {
  kd_warper_desc d{};
  d.kind = KD_WARP_IDENTITY;
  c.warpers.push_back(d);
}

// plasma_uv.hpp:35-51
vec vi_from_ui(double t_max, vec const &u_times) {
  vec v = u_times;
  std::sort(v.begin(), v.end(), std::greater<>());
  const int n = int(v.size());
  for (int i = n - 1; i > 0; i--) v[i] = v[i - 1] - v[i];
  if (n) v[0] = t_max - v[0];
  return v;
}

vec ui_from_vi(double t_max, vec const &v_times) {
  vec u = v_times;
  if (u.empty()) return u;
  u[0] = t_max - u[0];
  for (size_t i = 1; i < u.size(); i++) u[i] = u[i - 1] - u[i];
  return u;
}

void warper_plasma_uv_t::to_device(device_chain &c) const {
  kd_warper_desc d{};
  d.kind = KD_WARP_PLASMA;
  d.t_max = t_max;
  c.warpers.push_back(d);
}

// --------------------------------------------------------------------------
map_pair warper_product_1d_simple_base::map_reverse(vec const &l) const {
  vec x = l;
  double j = 1.0;
  for (auto &xi : x) {
    xi = m.Finv(xi);
    j *= 1.0 / m.f1(xi);
  }
  return {x, j};
}

map_pair warper_product_1d_simple_base::map_forward(vec const &u) const {
  vec x = u;
  double j = 1.0;
  for (auto &xi : x) {
    j *= m.f1(xi);
    xi = m.F(xi);
  }
  return {x, j};
}

vec warper_product_1d_simple_base::ui_from_li(vec const &l) const {
  vec x = l;
  for (auto &xi : x) xi = m.Finv(xi);
  return x;
}

vec warper_product_1d_simple_base::li_from_ui(vec const &u) const {
  vec x = u;
  for (auto &xi : x) xi = m.F(xi);
  return x;
}

double warper_product_1d_simple_base::jacobian_reverse(vec const &l) const {
  double r = 1.0;
  for (double li : l) r *= 1.0 / m.f1(m.Finv(li));
  return r;
}

double warper_product_1d_simple_base::jacobian_forward(vec const &u) const {
  double r = 1.0;
  for (double ui : u) r *= m.f1(ui);
  return r;
}

double warper_product_1d_simple_base::operator()(vec const &u) const {
  double r = 1.0;
  for (double ui : u) r *= m.f1(ui) / m.normalization();
  return r;
}

void warper_product_1d_simple_base::to_device(device_chain &c) const {
  kd_warper_desc d{};
  d.kind = KD_WARP_PRODUCT;
  d.n_maps = 1;
  d.map_offset = int64_t(c.maps.size());
  c.maps.push_back(m.to_device(c.tables));
  c.warpers.push_back(d);
}

// --------------------------------------------------------------------------
warper_product_1d_simple_t make_product_1d_simple_exponential(double u_max, double w) {
  return warper_product_1d_simple_t(map1d::make_exponential(u_max, w));
}
warper_product_1d_simple_t make_product_1d_simple_inverse(double u_max, double w) {
  return warper_product_1d_simple_t(map1d::make_inverse(u_max, w));
}
warper_product_1d_simple_t make_product_1d_simple_inverse_square(double u_max, double w) {
  return warper_product_1d_simple_t(map1d::make_inverse_square(u_max, w));
}
warper_product_1d_simple_t make_product_1d_simple_inverse_power(int power, double u_max, double w) {
  return warper_product_1d_simple_t(map1d::make_inverse_power(power, u_max, w));
}
warper_product_1d_simple_t make_product_1d_simple_uniform(double u_max) {
  return warper_product_1d_simple_t(map1d::make_uniform(u_max));
}

warper_product_1d_t make_product_1d_inverse_cube_alternate(int order, double time, double ws) {
  warper_product_1d_t r;
  r.reserve(order);
  for (int n = 1; n <= order; ++n)
    r.emplace_back(n % 2 == 0 ? make_product_1d_simple_inverse_power(3, time, ws)
                              : make_product_1d_simple_inverse(time, ws));
  return r;
}

warper_product_1d_interp_nearest_t make_product_1d_inverse_cube_alternate_interp_nearest(int order, double time,
                                                                                       double ws, int n_pts) {
  auto f1 = [ws](double t) -> double { return ws / (ws + t); };
  auto f2 = [ws](double t) -> double { return ws * ws * ws / ((ws + t) * (ws + t) * (ws + t)); };
  warper_product_1d_interp_nearest_t r;
  r.reserve(order);
  for (int n = 1; n <= order; ++n)
    r.emplace_back(warper_product_1d_simple_interp_nearest_t(n % 2 == 0 ? std::function<double(double)>(f2)
                                                                        : std::function<double(double)>(f1),
                                                             time, n_pts));
  return r;
}

warper_product_1d_interp_hybrid_t make_product_1d_inverse_cube_alternate_interp_hybrid(int order, double time,
                                                                                     double ws, int n_pts) {
  auto f1 = [ws](double t) -> double { return ws / (ws + t); };
  auto f2 = [ws](double t) -> double { return ws * ws * ws / ((ws + t) * (ws + t) * (ws + t)); };
  warper_product_1d_interp_hybrid_t r;
  r.reserve(order);
  for (int n = 1; n <= order; ++n) {
    const bool even = n % 2 == 0;
    r.emplace_back(warper_product_1d_simple_interp_hybrid_t(
        map1d::make_hybrid(even ? std::function<double(double)>(f2) : std::function<double(double)>(f1), time, n_pts,
                           even ? 3 : 1, ws)));
  }
  return r;
}

// --------------------------------------------------------------------------
// pilot fit (make_warper_from_proj.hpp:52-153, gsl_fit_linear_wrap.hpp:66-113)
namespace proj_details {

std::pair<double, double> fit_wlinear(const double *x, const double *w, const double *y, size_t n) {
  double W = 0, wm_x = 0, wm_y = 0, wm_dx2 = 0, wm_dxdy = 0;
  for (size_t i = 0; i < n; i++) {
    const double wi = w[i];
    if (wi > 0) {
      W += wi;
      wm_x += (x[i] - wm_x) * (wi / W);
      wm_y += (y[i] - wm_y) * (wi / W);
    }
  }
  W = 0;
  for (size_t i = 0; i < n; i++) {
    const double wi = w[i];
    if (wi > 0) {
      const double dx = x[i] - wm_x, dy = y[i] - wm_y;
      W += wi;
      wm_dx2 += (dx * dx - wm_dx2) * (wi / W);
      wm_dxdy += (dx * dy - wm_dxdy) * (wi / W);
    }
  }
  const double b = wm_dxdy / wm_dx2;
  const double a = wm_y - wm_x * b;
  return {a, b};
}

vec kernel_smoothing(vec const &y_in, double sigma) {
  const long N = long(y_in.size());
  vec x(N), logy(N), out(N);
  std::vector<long> dead;
  for (long i = 0; i < N; i++) {
    x[i] = double(i);
    logy[i] = std::log(y_in[i]);
    if (y_in[i] <= 0) dead.push_back(i);
  }
  const long H = std::max(std::min(long(3 * sigma), N), 1L);
  const long K = 2 * H + 1;
  vec kernel(K);
  for (long j = 0; j < K; j++) kernel[j] = std::exp(-std::pow((double(j) - double(H)) / sigma, 2) / 2);
  vec kc;
  for (long i = 0; i < N; i++) {
    const double *kp = kernel.data();
    if (!dead.empty()) {
      kc = kernel;
      for (long j : dead) {
        const long dp = j + H - i;
        if (dp >= 0 && dp < K) kc[dp] = 0;
      }
      kp = kc.data();
    }
    const long start_ker = std::max(0L, H - i);
    const long end_ker = std::min(N + H - i - 1, K - 1);
    const long size_fit = end_ker - start_ker + 1;
    const long start_x = start_ker - H + i;
    auto c = fit_wlinear(x.data() + start_x, kp + start_ker, logy.data() + start_x, size_t(size_fit));
    out[i] = std::exp(c.first + x[i] * c.second);
  }
  return out;
}

} // namespace proj_details

warper_product_1d_interp_nearest_t make_warper_from_proj_samples(std::vector<vec> const &l, vec const &abs_values,
                                                                 int order, int num_bins, double sigma) {
  if (l.size() != abs_values.size()) throw keldy_error("sample and value counts differ");
  if (num_bins < 2) throw keldy_error("num_bins must be >= 2");
  const double bin_size = 1.0 / num_bins;
  std::vector<vec> data(order, vec(num_bins, 0.0));
  std::vector<std::vector<long>> count(order, std::vector<long>(num_bins, 0));
  for (size_t s = 0; s < l.size(); s++) {
    if (int(l[s].size()) != order) throw keldy_error("sample of wrong dimension");
    for (int a = 0; a < order; a++) {
      const double xv = l[s][a];
      if (!(0.0 <= xv && xv <= 1.0)) continue;  // binner drops out-of-range values
      const long b = (xv == 1.0) ? num_bins - 1 : long(xv / bin_size);
      data[a][b] += abs_values[s];
      count[a][b]++;
    }
  }
  return make_warper_from_proj_histograms(std::move(data), count, order, num_bins, sigma);
}

warper_product_1d_interp_nearest_t make_warper_from_proj_histograms(std::vector<vec> data,
                                                                    std::vector<std::vector<long>> const &count,
                                                                    int order, int num_bins, double sigma) {
  if (num_bins < 2) throw keldy_error("num_bins must be >= 2");
  if (int(data.size()) != order || int(count.size()) != order) throw keldy_error("histograms of wrong dimension");
  const double bin_size = 1.0 / num_bins;
  warper_product_1d_interp_nearest_t warper;
  vec time_pts(num_bins + 1), f1_pts(num_bins + 1);
  const double delta = 1.0 / num_bins;
  for (int i = 0; i <= num_bins; i++) time_pts[i] = i * delta;
  time_pts[num_bins] = 1.;
  for (int a = 0; a < order; a++) {
    for (int b = 0; b < num_bins; b++)
      if (count[a][b] > 0) data[a][b] /= double(count[a][b]);
    const vec sm = proj_details::kernel_smoothing(data[a], sigma / bin_size);
    f1_pts[0] = 0;
    for (int i = 0; i < num_bins; i++) f1_pts[i + 1] = 2 * sm[i] - f1_pts[i];
    warper.emplace_back(warper_product_1d_simple_interp_nearest_t(time_pts, f1_pts));
  }
  return warper;
}

warper_product_1d_interp_nearest_t make_warper_from_proj(std::function<dcomplex(vec)> const &warped_integrand,
                                                         int order, int num_bins, int nr_samples, double sigma) {
  sobol g(order, 0);
  std::vector<vec> l;
  vec vals;
  l.reserve(nr_samples);
  vals.reserve(nr_samples);
  for (int i = 0; i < nr_samples; i++) {
    l.push_back(g());
    vals.push_back(std::abs(warped_integrand(l.back())));
  }
  return make_warper_from_proj_samples(l, vals, order, num_bins, sigma);
}

// --------------------------------------------------------------------------
// projection warper (projection.hpp:53-382)
namespace proj_details {

// Akima (1970) spline, GSL's non-periodic variant: the interval slopes m_i,
// extended by m_-1 = 2 m_0 - m_1, m_-2 = 3 m_0 - 2 m_1 and the mirror at the
// far end; knot slope t_i = ((1 - a) m_{i-1} + a m_i) with a = |m_{i-1} -
// m_{i-2}| / (|m_{i+1} - m_i| + |m_{i-1} - m_{i-2}|), m_i where both jumps
// vanish; y = y_i + dx (b_i + dx (c_i + dx d_i)).
void akima_coeffs(vec const &x, vec const &y, vec &b, vec &c, vec &d) {
  const size_t n = x.size();
  if (n < 3 || y.size() != n) throw keldy_error("Akima interpolation needs at least 3 points");
  vec mm(n + 3); // mm[k + 2] = slope of interval k, k = -2 .. n
  for (size_t i = 0; i + 1 < n; i++) mm[i + 2] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
  mm[0] = 3.0 * mm[2] - 2.0 * mm[3];
  mm[1] = 2.0 * mm[2] - mm[3];
  mm[n + 1] = 2.0 * mm[n] - mm[n - 1];
  mm[n + 2] = 3.0 * mm[n] - 2.0 * mm[n - 1];
  b.assign(n - 1, 0.0);
  c.assign(n - 1, 0.0);
  d.assign(n - 1, 0.0);
  for (size_t i = 0; i + 1 < n; i++) {
    const double *m = mm.data() + i + 2; // m[k] = slope of interval i + k
    const double ne = std::abs(m[1] - m[0]) + std::abs(m[-1] - m[-2]);
    if (ne == 0.0) {
      b[i] = m[0];
      continue;
    }
    const double h = x[i + 1] - x[i];
    const double ne_next = std::abs(m[2] - m[1]) + std::abs(m[0] - m[-1]);
    const double a_i = std::abs(m[-1] - m[-2]) / ne;
    double t_next;
    if (ne_next == 0.0) {
      t_next = m[0];
    } else {
      const double a_n = std::abs(m[0] - m[-1]) / ne_next;
      t_next = (1.0 - a_n) * m[0] + a_n * m[1];
    }
    b[i] = (1.0 - a_i) * m[-1] + a_i * m[0];
    c[i] = (3.0 * m[0] - 2.0 * b[i] - t_next) / h;
    d[i] = (b[i] + t_next - 2.0 * m[0]) / (h * h);
  }
}

double akima_eval(vec const &x, vec const &y, vec const &b, vec const &c, vec const &d, double xv) {
  check_domain(x.data(), x.size(), xv);
  const size_t i = knot_find(x.data(), x.size(), xv);
  const double dx = xv - x[i];
  return y[i] + dx * (b[i] + dx * (c[i] + dx * d[i]));
}

namespace {

// local_linear_reg (gsl_fit_linear_wrap.hpp:66-113) of y at x = 0..N-1 with the
// kernel (odd size) centred on each point; dead points weighted zero
vec local_linear(vec const &y, vec const &kernel, std::vector<long> const &dead) {
  const long N = long(y.size()), K = long(kernel.size()), H = K / 2;
  vec x(N), out(N), kc;
  for (long i = 0; i < N; i++) x[i] = double(i);
  for (long i = 0; i < N; i++) {
    const double *kp = kernel.data();
    if (!dead.empty()) {
      kc = kernel;
      for (long j : dead) {
        const long dp = j + H - i;
        if (dp >= 0 && dp < K) kc[dp] = 0;
      }
      kp = kc.data();
    }
    const long start_ker = std::max(0L, H - i);
    const long end_ker = std::min(N + H - i - 1, K - 1);
    const long start_x = start_ker - H + i;
    auto cf = fit_wlinear(x.data() + start_x, kp + start_ker, y.data() + start_x, size_t(end_ker - start_ker + 1));
    out[i] = cf.first + x[i] * cf.second;
  }
  return out;
}

// Brent's minimisation (golden section with parabolic steps) of f on
// [lo, hi] from the guess x0, which must lie below both ends; stops when the
// bracket is narrower than abstol or after max_iter steps.
template <typename F> double brent_minimize(F const &f, double lo, double x0, double hi, double abstol, int max_iter) {
  const double golden = 0.3819660112501051; // (3 - sqrt 5) / 2
  const double f_lo = f(lo), f_x0 = f(x0), f_hi = f(hi);
  if (!std::isfinite(f_lo) || !std::isfinite(f_x0) || !std::isfinite(f_hi))
    throw keldy_error("Function returned Nan or Inf at the minimisation bracket");
  if (!(f_x0 < f_lo && f_x0 < f_hi))
    throw keldy_error("Endpoints do not enclose a minimum. Make sure f(x_lower) > f(x_guess) < f(x_upper)");
  double z = x0, fz = f_x0;
  double v = lo + golden * (hi - lo), fv = f(v), w = v, fw = fv;
  double d = 0.0, e = 0.0;
  for (int it = 0; it < max_iter; it++) {
    const double w_lower = z - lo, w_upper = hi - z;
    const double tol = 1.4901161193847656e-08 * std::abs(z); // sqrt(DBL_EPSILON) |z|
    const double midpoint = 0.5 * (lo + hi);
    double p = 0.0, q = 0.0, r = 0.0;
    if (std::abs(e) > tol) { // parabola through (z, fz), (w, fw), (v, fv)
      r = (z - w) * (fz - fv);
      q = (z - v) * (fz - fw);
      p = (z - v) * q - (z - w) * r;
      q = 2.0 * (q - r);
      if (q > 0.0)
        p = -p;
      else
        q = -q;
      r = e;
      e = d;
    }
    double u;
    if (std::abs(p) < std::abs(0.5 * q * r) && p < q * w_lower && p < q * w_upper) {
      const double t2 = 2.0 * tol;
      d = p / q;
      u = z + d;
      if ((u - lo) < t2 || (hi - u) < t2) d = (z < midpoint) ? tol : -tol;
    } else {
      e = (z < midpoint) ? hi - z : -(z - lo);
      d = golden * e;
    }
    u = std::abs(d) >= tol ? z + d : z + ((d > 0) ? tol : -tol);
    const double fu = f(u);
    if (!std::isfinite(fu)) throw keldy_error("Function returned Nan or Inf.");
    if (fu <= fz) {
      if (u < z)
        hi = z;
      else
        lo = z;
      v = w;
      fv = fw;
      w = z;
      fw = fz;
      z = u;
      fz = fu;
    } else {
      if (u < z)
        lo = u;
      else
        hi = u;
      if (fu <= fw || w == z) {
        v = w;
        fv = fw;
        w = u;
        fw = fu;
      } else if (fu <= fv || v == z || v == w) {
        v = u;
        fv = fu;
      }
    }
    if (std::abs(hi - lo) < abstol) break;
  }
  return z;
}

double gaussian(double x, double x0, double sigma) { return std::exp(-std::pow((x - x0) / sigma, 2) / 2); }

} // namespace

double optimize_sigma(vec const &y, double sigma, double upper) {
  const long N = long(y.size());
  const double sigma_min = 0.3; // narrow gaussians make the local fits unstable
  vec log_y(N);
  std::vector<long> dead;
  for (long i = 0; i < N; i++) {
    if (y[i] <= 0) {
      dead.push_back(i);
      log_y[i] = 0;
    } else {
      log_y[i] = std::log(y[i]);
    }
  }
  const long H = std::max(std::min(long(3 * upper), N), 1L);
  vec kernel(2 * H + 1);
  // the leave-one-out error of the smoothing (the centre weight is zero),
  // symmetric in sigma: searched on [-upper, upper] so that a noiseless
  // projection still has a bracketed minimum
  auto cost = [&](double s) {
    s = std::max(std::abs(s), sigma_min);
    for (long j = 0; j < long(kernel.size()); j++) kernel[j] = gaussian(double(j), double(H), s);
    kernel[H] = 0;
    const double k1 = kernel[H + 1];
    for (double &kv : kernel) kv /= k1;
    const vec out = local_linear(log_y, kernel, dead);
    double r = 0.0;
    for (long i = 0; i < N; i++) r += std::pow(std::exp(out[i]) - y[i], 2);
    if (std::isnan(r))
      throw keldy_error("Smoothing procedure returned a Nan. This may be due to a lack of data, try to increase the "
                        "number of samples.");
    return r;
  };
  return std::max(std::abs(brent_minimize(cost, -upper, sigma, upper, 1e-8, 20)), sigma_min);
}

} // namespace proj_details

namespace {
// TRIQS gf<retime> evaluation on the uniform mesh of [0, 1] with n points:
// linear interpolation between the neighbouring mesh points
double mesh_eval(vec const &v, double t) {
  const size_t n = v.size();
  if (!(t >= -1e-12 && t <= 1.0 + 1e-12)) throw keldy_error("point outside the [0, 1] mesh of the projection warper");
  const double delta = 1.0 / double(n - 1);
  const double a = t / delta;
  long i = long(std::floor(a));
  i = std::min<long>(std::max<long>(i, 0), long(n) - 2);
  const double w = a - double(i);
  return (1.0 - w) * v[size_t(i)] + w * v[size_t(i) + 1];
}
} // namespace

warper_projection_t::warper_projection_t(std::function<dcomplex(vec)> const &warped_integrand, int order, int num_bins,
                                         int nr_samples, double sigma, bool optimize_sigma) {
  if (order < 1) throw keldy_error("order must be >= 1");
  if (nr_samples < 0) throw keldy_error("nr_samples must be >= 0");
  sobol g(order, 0);
  std::vector<vec> l;
  vec vals;
  l.reserve(size_t(nr_samples));
  vals.reserve(size_t(nr_samples));
  for (int i = 0; i < nr_samples; i++) {
    l.push_back(g());
    vals.push_back(std::abs(warped_integrand(l.back())));
  }
  order_ = order;
  num_bins_ = num_bins;
  build(l, vals, sigma, optimize_sigma);
}

warper_projection_t warper_projection_t::from_samples(std::vector<vec> const &l, vec const &abs_values, int order,
                                                      int num_bins, double sigma, bool optimize_sigma) {
  if (order < 1) throw keldy_error("order must be >= 1");
  if (l.size() != abs_values.size()) throw keldy_error("sample and value counts differ");
  warper_projection_t w;
  w.order_ = order;
  w.num_bins_ = num_bins;
  w.build(l, abs_values, sigma, optimize_sigma);
  return w;
}

void warper_projection_t::build(std::vector<vec> const &l, vec const &vals, double sigma, bool optimize_sigma) {
  const int nb = num_bins_;
  if (nb < 3) throw keldy_error("num_bins must be >= 3");
  const double delta = 1.0 / (nb - 1.0);
  // gather_data (projection.hpp:189-222): binner_t<1, 0, double> on [-d/2, 1 + d/2]
  const double xmin = -delta / 2, xmax = 1 + delta / 2;
  bin_size_ = (xmax - xmin) / nb;
  xi_.assign(size_t(order_), vec(size_t(nb), 0.0));
  xi_count_.assign(size_t(order_), std::vector<long>(size_t(nb), 0));
  for (size_t s = 0; s < l.size(); s++) {
    if (int(l[s].size()) != order_) throw keldy_error("sample of wrong dimension");
    for (int a = 0; a < order_; a++) {
      const double x = l[s][size_t(a)];
      if (!(xmin <= x && x <= xmax)) continue;
      const long b = x == xmax ? nb - 1 : long((x - xmin) / bin_size_);
      xi_[size_t(a)][size_t(b)] += vals[s];
      xi_count_[size_t(a)][size_t(b)]++;
    }
  }
  for (int a = 0; a < order_; a++)
    for (int b = 0; b < nb; b++)
      if (xi_count_[size_t(a)][size_t(b)] > 0) xi_[size_t(a)][size_t(b)] /= double(xi_count_[size_t(a)][size_t(b)]);
  // populate_sigmas (projection.hpp:247-274)
  sigmas_.assign(size_t(order_), sigma / bin_size_);
  if (optimize_sigma) {
    if (bin_size_ >= 2) throw keldy_error("delta too large");
    for (int a = 0; a < order_; a++)
      sigmas_[size_t(a)] = proj_details::optimize_sigma(xi_[size_t(a)], sigmas_[size_t(a)], 1.0 / bin_size_);
  }
  // populate_functions (projection.hpp:276-317)
  const int n_inv = 1 + 5 * (nb - 1);
  vec mesh(static_cast<size_t>(nb)), times(static_cast<size_t>(nb));
  for (int k = 0; k < nb; k++) mesh[size_t(k)] = k * delta;
  for (int k = 0; k < nb; k++) times[size_t(k)] = xmin + (k + 0.5) * bin_size_;
  times[0] = 0;      // snap boundaries
  times[nb - 1] = 1; // snap boundaries
  fn_.assign(size_t(order_), vec());
  F_.assign(size_t(order_), vec());
  Finv_.assign(size_t(order_), vec());
  norm_.assign(size_t(order_), 0.0);
  for (int a = 0; a < order_; a++) {
    const vec sm = proj_details::kernel_smoothing(xi_[size_t(a)], sigmas_[size_t(a)]);
    vec &fn = fn_[size_t(a)];
    fn.resize(size_t(nb));
    for (int k = 0; k < nb; k++) fn[size_t(k)] = linear_eval(times.data(), sm.data(), size_t(nb), mesh[size_t(k)]);
    // midpoint rule on the mesh: F_k = sum_{j <= k} fn(t_j - d/2) d, F_0 = 0
    vec &F = F_[size_t(a)];
    F.assign(size_t(nb), 0.0);
    for (int k = 1; k < nb; k++) F[size_t(k)] = mesh_eval(fn, mesh[size_t(k)] - delta / 2) * delta;
    for (int k = 1; k < nb; k++) F[size_t(k)] += F[size_t(k) - 1];
    norm_[size_t(a)] = F[size_t(nb) - 1];
    for (double &x : F) x /= norm_[size_t(a)];
    // the inverse CDF by Akima interpolation through (F, mesh), on 1 + 5(nb - 1) points
    vec b, c, d;
    proj_details::akima_coeffs(F, mesh, b, c, d);
    vec &Fi = Finv_[size_t(a)];
    Fi.resize(size_t(n_inv));
    const double dl = 1.0 / double(n_inv - 1);
    for (int j = 0; j < n_inv; j++) Fi[size_t(j)] = proj_details::akima_eval(F, mesh, b, c, d, j * dl);
  }
}

void warper_projection_t::check_axis(int axis) const {
  if (axis < 0 || axis >= order_) throw keldy_error("this axis does not exist.");
}

vec warper_projection_t::get_xi(int axis) const {
  check_axis(axis);
  return xi_[size_t(axis)];
}

std::vector<long> warper_projection_t::get_xi_counts(int axis) const {
  check_axis(axis);
  return xi_count_[size_t(axis)];
}

vec warper_projection_t::get_xi_bin_coord(int axis) const {
  check_axis(axis);
  vec t(static_cast<size_t>(num_bins_));
  const double xmin = -0.5 / (num_bins_ - 1.0);
  for (int k = 0; k < num_bins_; k++) t[size_t(k)] = xmin + (k + 0.5) * bin_size_;
  return t;
}

vec warper_projection_t::get_sigmas() const {
  vec out(sigmas_.size());
  for (size_t i = 0; i < sigmas_.size(); i++) out[i] = sigmas_[i] * bin_size_;
  return out;
}

vec warper_projection_t::get_fi(int axis) const {
  check_axis(axis);
  return fn_[size_t(axis)];
}

vec warper_projection_t::get_fi_integrated(int axis) const {
  check_axis(axis);
  return F_[size_t(axis)];
}

vec warper_projection_t::get_fi_integrated_inverse(int axis) const {
  check_axis(axis);
  return Finv_[size_t(axis)];
}

vec warper_projection_t::ui_from_li(vec const &l) const {
  if (int(l.size()) > order_) throw keldy_error("point of higher dimension than the projection warper");
  vec u(l.size());
  for (size_t i = 0; i < l.size(); i++) u[i] = mesh_eval(Finv_[i], l[i]);
  return u;
}

vec warper_projection_t::li_from_ui(vec const &u) const {
  if (int(u.size()) > order_) throw keldy_error("point of higher dimension than the projection warper");
  vec l(u.size());
  for (size_t i = 0; i < u.size(); i++) l[i] = mesh_eval(F_[i], u[i]);
  return l;
}

double warper_projection_t::jacobian_reverse(vec const &l) const {
  if (int(l.size()) > order_) throw keldy_error("point of higher dimension than the projection warper");
  double j = 1.0;
  for (size_t i = 0; i < l.size(); i++) j *= norm_[i] / mesh_eval(fn_[i], mesh_eval(Finv_[i], l[i]));
  return j;
}

double warper_projection_t::jacobian_forward(vec const &u) const {
  if (int(u.size()) > order_) throw keldy_error("point of higher dimension than the projection warper");
  double j = 1.0;
  for (size_t i = 0; i < u.size(); i++) j *= mesh_eval(fn_[i], u[i]);
  return j;
}

void warper_projection_t::to_device(device_chain &c) const {
  if (order_ < 1) throw keldy_error("empty projection warper");
  kd_warper_desc d{};
  d.kind = KD_WARP_PRODUCT;
  d.n_maps = int32_t(order_);
  d.map_offset = int64_t(c.maps.size());
  for (int a = 0; a < order_; a++) {
    kd_map1d m{};
    m.kind = KD_MAP_PROJECTION;
    m.u_max = 1.0;
    m.w = 1.0;
    m.norm = norm_[size_t(a)];
    m.n_pts = int32_t(fn_[size_t(a)].size());
    m.n_inv = int32_t(Finv_[size_t(a)].size());
    m.t_off = int64_t(c.tables.size());
    c.tables.insert(c.tables.end(), Finv_[size_t(a)].begin(), Finv_[size_t(a)].end());
    m.f_off = int64_t(c.tables.size());
    c.tables.insert(c.tables.end(), fn_[size_t(a)].begin(), fn_[size_t(a)].end());
    c.maps.push_back(m);
  }
  c.warpers.push_back(d);
}

// --------------------------------------------------------------------------
map_pair warper_train_t::map_reverse(vec const &l, int start, int end) const {
  if (start < 0 || end < 0 || start > int(size()) || end > int(size()) || end > start)
    throw keldy_error("invalid warper domain range");
  double j = 1.0;
  vec x = l;
  for (int q = start - 1; q >= end; q--) {
    auto r = std::visit([&x](auto const &w) { return w.map_reverse(x); }, warpers[q]);
    x = std::move(r.first);
    j *= r.second;
  }
  return {x, j};
}

map_pair warper_train_t::map_forward(vec const &u, int start, int end) const {
  if (start < 0 || end < 0 || start > int(size()) || end > int(size()) || start > end)
    throw keldy_error("invalid warper domain range");
  double j = 1.0;
  vec x = u;
  for (int q = start; q < end; q++) {
    auto r = std::visit([&x](auto const &w) { return w.map_forward(x); }, warpers[q]);
    x = std::move(r.first);
    j *= r.second;
  }
  return {x, j};
}

vec warper_train_t::ui_from_li(vec const &l) const {
  vec x = l;
  for (auto it = warpers.crbegin(); it != warpers.crend(); ++it)
    x = std::visit([&x](auto const &w) { return w.ui_from_li(x); }, *it);
  return x;
}

vec warper_train_t::li_from_ui(vec const &u) const {
  vec x = u;
  for (auto const &w : warpers) x = std::visit([&x](auto const &ww) { return ww.li_from_ui(x); }, w);
  return x;
}

device_chain warper_train_t::to_device() const {
  device_chain c;
  for (auto const &w : warpers) std::visit([&c](auto const &ww) { ww.to_device(c); }, w);
  return c;
}

} // namespace keldy::warpers

// --------------------------------------------------------------------------
// learned flow warper (host evaluation of the device network)
namespace keldy::warpers {

namespace {
constexpr int kH = KD_FLOW_HIDDEN;
constexpr double kFlowClampHost = 1.0 / 1099511627776.0; // 2^-40
int fl_nA(int n, int k) { return kd::flow_nA(n, k); }
float bf16_round(float f) { return kd::bf2f(kd::f2bf(f)); }
double log_sig_prime(double y) {
  const double ay = std::abs(y);
  return -ay - 2.0 * std::log1p(std::exp(-ay));
}
} // namespace

warper_flow_t::warper_flow_t(int order, int n_layers, std::vector<float> params, int n_cond, double cond)
   : order_(order), layers_(n_layers), nc_(n_cond), cond_(cond), params_(std::move(params)), stamp_(next_stamp()) {
  if (order < 1 || order > KD_MAX_ORDER) throw keldy_error("flow dimension must be in [1, 12]");
  if (n_layers < 1 || n_layers > KD_FLOW_MAX_LAYERS) throw keldy_error("flow layers must be in [1, 16]");
  if (n_cond < 0 || n_cond > 1) throw keldy_error("n_cond must be 0 or 1");
  if (!std::isfinite(cond)) throw keldy_error("non-finite conditioning value");
  if (int64_t(params_.size()) != kd::flow_n_params(order, n_layers, n_cond))
    throw keldy_error("flow parameter count does not match (order, layers, conditioning inputs)");
}

void warper_flow_t::net(int k, const double *y, double *s, double *t) const {
  const int n = order_, nA = fl_nA(n, k), nB = n - nA;
  const float *W1 = params_.data() + kd::flow_layer_offset(n, k, nc_);
  const float *b1 = W1 + kH * nA, *W2 = b1 + kH, *b2 = W2 + kH * kH, *W3 = b2 + kH, *b3 = W3 + 2 * nB * kH;
  const float *Wc = b3 + 2 * nB; // nc_ = 1 only
  const float cv = float(cond_);
  float a[KD_MAX_ORDER];
  int na = 0;
  for (int i = 0; i < n; i++)
    if (((i + k) & 1) == 0) a[na++] = float(y[i]);
  float h1[kH], h2[kH];
  for (int j = 0; j < kH; j++) {
    float z = b1[j];
    if (nc_) z = std::fmaf(Wc[j], cv, z);
    for (int i = 0; i < nA; i++) z = std::fmaf(W1[j * nA + i], a[i], z);
    h1[j] = bf16_round(std::max(z, 0.f));
  }
  for (int j = 0; j < kH; j++) {
    double z = 0.0;
    for (int i = 0; i < kH; i++) z += double(bf16_round(W2[j * kH + i])) * double(h1[i]);
    h2[j] = bf16_round(std::max(float(z) + b2[j], 0.f));
  }
  for (int q = 0; q < 2 * nB; q++) {
    double z = 0.0;
    for (int i = 0; i < kH; i++) z += double(bf16_round(W3[q * kH + i])) * double(h2[i]);
    const float o = float(z) + b3[q];
    if (q < nB)
      s[q] = double(kd::flow_s_scale(layers_) * std::tanh(o));
    else
      t[q - nB] = double(o);
  }
}

map_pair warper_flow_t::map_reverse(vec const &l) const {
  if (int(l.size()) != order_) throw keldy_error("flow warper: point dimension differs from the flow's");
  double y[KD_MAX_ORDER], s[KD_MAX_ORDER], t[KD_MAX_ORDER];
  double ldj = 0.0;
  for (int i = 0; i < order_; i++) {
    const double li = std::min(std::max(l[i], kFlowClampHost), 1.0 - kFlowClampHost);
    y[i] = std::log(li) - std::log1p(-li);
    ldj -= std::log(li) + std::log1p(-li);
  }
  for (int k = 0; k < layers_; k++) {
    net(k, y, s, t);
    int nb = 0;
    for (int i = 0; i < order_; i++)
      if (((i + k) & 1) == 1) {
        y[i] = y[i] * std::exp(s[nb]) + t[nb];
        ldj += s[nb];
        nb++;
      }
  }
  vec x(order_);
  for (int i = 0; i < order_; i++) {
    x[i] = 1.0 / (1.0 + std::exp(-y[i]));
    ldj += log_sig_prime(y[i]);
  }
  return {x, std::exp(ldj)};
}

map_pair warper_flow_t::map_forward(vec const &xin) const {
  if (int(xin.size()) != order_) throw keldy_error("flow warper: point dimension differs from the flow's");
  double v[KD_MAX_ORDER], s[KD_MAX_ORDER], t[KD_MAX_ORDER];
  double ldj = 0.0; // log dl/dx
  for (int i = 0; i < order_; i++) {
    const double xi = std::min(std::max(xin[i], kFlowClampHost), 1.0 - kFlowClampHost);
    v[i] = std::log(xi) - std::log1p(-xi);
    ldj -= std::log(xi) + std::log1p(-xi);
  }
  for (int k = layers_ - 1; k >= 0; k--) {
    net(k, v, s, t);
    int nb = 0;
    for (int i = 0; i < order_; i++)
      if (((i + k) & 1) == 1) {
        v[i] = (v[i] - t[nb]) * std::exp(-s[nb]);
        ldj -= s[nb];
        nb++;
      }
  }
  vec l(order_);
  for (int i = 0; i < order_; i++) {
    l[i] = 1.0 / (1.0 + std::exp(-v[i]));
    ldj += log_sig_prime(v[i]);
  }
  return {l, std::exp(ldj)};
}

void warper_flow_t::to_device(device_chain &c) const {
  if (params_.empty()) throw keldy_error("flow warper has no parameters");
  kd_warper_desc d{};
  d.kind = KD_WARP_FLOW;
  c.warpers.push_back(d);
  c.flow_order = order_;
  c.flow_layers = layers_;
  c.flow_nc = nc_;
  c.flow_cond = cond_;
  c.flow_params = params_;
  c.flow_stamp = stamp_;
}

} // namespace keldy::warpers
