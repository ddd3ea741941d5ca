// integrator.cpp -- host side of the integrand and the QMC driver.
// Reference: wick_direct.hpp:48-60 / wick_direct.cpp:43-228, integrator.hpp:68-131,
// compute_obs.hpp:51-80.
#include "integrator.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <random>

namespace keldy {

device_ctx::device_ctx() {
  kd_ctx *c = nullptr;
  const int rc = kd_create(&c, 0);
  if (rc == KD_ERR_NODEVICE) throw keldy_error("keldy_amd: no HIP device available (the integrand runs on the GPU)");
  if (rc != KD_OK) throw keldy_error("keldy_amd: kd_create failed (" + std::to_string(rc) + ")");
  ctx_ = std::shared_ptr<kd_ctx>(c, [](kd_ctx *p) { kd_destroy(p); });
}

void device_ctx::check(int rc) const {
  if (rc != KD_OK) throw keldy_error(kd_last_error(ctx_.get()));
}

namespace binner {
std::vector<double> binner_1d_1_t::get_bin_coord() const {
  std::vector<double> c(static_cast<size_t>(nr_bins));
  for (long i = 0; i < nr_bins; i++) c[size_t(i)] = xmin + (double(i) + 0.5) * get_bin_size();
  return c;
}
// entries at equal coordinates merged first, in order of appearance (binner.hpp:95-114)
double sparse_binner_1d_1_t::sum_moduli() const {
  std::map<std::pair<double, long>, size_t> where;
  std::vector<dcomplex> merged;
  for (auto const &d : data) {
    auto it = where.find(d.first);
    if (it == where.end()) {
      where.emplace(d.first, merged.size());
      merged.push_back(d.second);
    } else {
      merged[it->second] += d.second;
    }
  }
  double s = 0.0;
  for (auto const &v : merged) s += std::abs(v);
  return s;
}
std::vector<double> binner_1d_t::get_bin_coord() const {
  std::vector<double> c(static_cast<size_t>(nr_bins));
  for (long i = 0; i < nr_bins; i++) c[size_t(i)] = xmin + (double(i) + 0.5) * get_bin_size();
  return c;
}
double sparse_binner_1d_t::sum_moduli() const {
  double s = 0.0;
  for (auto const &d : data) s += std::abs(d.second);
  return s;
}
} // namespace binner

namespace impurity_oneband {

namespace {
kd_gf_index to_kd(gf_index_t const &g) {
  kd_gf_index k{};
  k.time = g.contour.time;
  k.k_idx = int32_t(g.contour.k_idx);
  k.spin = int32_t(g.spin);
  k.orbital = int32_t(g.orbital);
  return k;
}
} // namespace

integrand_g_direct::integrand_g_direct(g0_keldysh_contour_t g0_in, gf_index_t A, gf_index_t B, double cutoff_)
   : g0_(std::move(g0_in)), external_A(A), external_B(B), cutoff(cutoff_) {
  if (external_A.contour.time < 0. || external_B.contour.time < 0.)
    throw keldy_error("An external point has negative time.");
  const double t_max = g0_.get_time_max();
  if (external_A.contour.time > t_max || external_B.contour.time > t_max)
    throw keldy_error("An external point is out of g0 time window.");
  if ((external_A.orbital != 0 || external_B.orbital != 0) && !g0_.model.make_dot_lead)
    throw keldy_error("Need g0_model with make_dot_lead == true to calculate off_diagonal gf.");
}

integrand_g_direct integrand_g_direct::kernel_method(g0_keldysh_contour_t g0_in, gf_index_t X, int obs,
                                                     double omega) {
  integrand_g_direct f(std::move(g0_in), X, X, 0.0);
  f.obs_ = obs;
  f.omega_ = omega;
  return f;
}

void integrand_g_direct::bind(device_ctx const &ctx, int order) const {
  const g0_model &m = g0_.model;
  ctx.check(kd_set_g0_table(ctx.get(), reinterpret_cast<const double *>(m.lesser.data()),
                            reinterpret_cast<const double *>(m.greater.data()), m.n_orb, m.n_pts, m.time_max,
                            m.model_omega.get_param_alpha()));
  const kd_gf_index a = to_kd(external_A), b = to_kd(external_B);
  if (obs_ == KD_OBS_KERNEL)
    ctx.check(kd_set_kernel_integrand(ctx.get(), order, &a));
  else if (obs_ == KD_OBS_KERNEL_OMEGA)
    ctx.check(kd_set_kernel_omega_integrand(ctx.get(), order, &a, omega_));
  else
    ctx.check(kd_set_integrand(ctx.get(), order, &a, &b, cutoff));
}

void integrand_g_direct::ensure_bound(int order) const {
  if (order > KD_MAX_ORDER)
    throw keldy_error("order " + std::to_string(order) + " exceeds the device maximum " + std::to_string(KD_MAX_ORDER));
  if (obs_ != KD_OBS_DIRECT && order < 1) throw keldy_error("Order 0 is not implemented");
  if (!ctx_) ctx_ = std::make_shared<device_ctx>();
  if (bound_order_ != order) {
    bind(*ctx_, order);
    bound_order_ = order;
  }
}

void integrand_g_direct::eval_entries(const double *u, long n, int order, bool keep_u_hypercube, dcomplex *out,
                                      int *in_domain) const {
  if (obs_ != KD_OBS_KERNEL) throw keldy_error("entries exist for the kernel method only");
  ensure_bound(order);
  ctx_->check(kd_eval_kernel_batch(ctx_->get(), u, n, 0, keep_u_hypercube ? 1 : 0, reinterpret_cast<double *>(out),
                                   in_domain, nullptr));
  check_window(out, n * (long(1) << order) * order);
}

// The device yields NaN for g0 lookups outside the tabulated window; the
// reference's gf evaluation throws there, and so does the host API.
void check_window(const dcomplex *out, long n) {
  for (long i = 0; i < n; i++)
    if (std::isnan(out[i].real()) || std::isnan(out[i].imag()))
      throw keldy_error("g0 evaluated outside its time window [-time_max, time_max]");
}

void integrand_g_direct::eval_batch(const double *u, long n, int order, bool keep_u_hypercube, dcomplex *out,
                                    int *in_domain) const {
  if (obs_ == KD_OBS_KERNEL) throw keldy_error("the binned kernel method has entries, not one value per point");
  ensure_bound(order);
  ctx_->check(kd_eval_batch(ctx_->get(), u, n, keep_u_hypercube ? 1 : 0, reinterpret_cast<double *>(out),
                            in_domain));
  check_window(out, n);
}

std::pair<dcomplex, int> integrand_g_direct::operator()(std::vector<double> const &times, bool keep_u_hypercube) const {
  dcomplex v;
  int dom = 0;
  eval_batch(times.data(), 1, int(times.size()), keep_u_hypercube, &v, &dom);
  return {v, dom};
}

// Orders 0-2 by explicit Wick contractions (wick_direct.cpp:161-228).
std::pair<dcomplex, int> integrand_g_direct::eval_no_det(std::vector<double> const &times, bool keep_u_hypercube) const {
  if (external_A.spin != external_B.spin) return {0.0, 0};
  if (keep_u_hypercube && std::any_of(times.begin(), times.end(), [](double t) { return t < 0.0; })) return {0.0, 0};
  const int n = int(times.size());
  gf_index_t a = external_A, b = external_B;
  a.contour.timesplit_n = n;
  b.contour.timesplit_n = n;
  const g0_keldysh_contour_t &g = g0_;
  if (n == 0) return {g(a, b, false), 1};
  dcomplex res = 0.0;
  if (n == 1) {
    for (int k0 : {0, 1}) {
      auto pt = [&](int sp) { return gf_index_t(times[0], sp, k0, 0); };
      const double sign = k0 == 0 ? 1. : -1.;
      res += -sign * g(pt(down), pt(down)) * g(a, pt(up)) * g(pt(up), b);
    }
  } else if (n == 2) {
    for (int k0 : {0, 1})
      for (int k1 : {0, 1}) {
        auto pt = [&](int i, int sp) { return gf_index_t(times[i], sp, i == 0 ? k0 : k1, i); };
        const double sign = ((k0 + k1) % 2 == 0) ? 1. : -1.;
        res += sign *
               (g(pt(0, up), pt(0, up)) * g(a, pt(1, up)) * g(pt(1, up), b) +
                g(a, pt(0, up)) * g(pt(1, up), pt(1, up)) * g(pt(0, up), b)) *
               g(pt(1, down), pt(0, down)) * g(pt(0, down), pt(1, down));
        res += sign *
               (g(a, pt(0, up)) * g(pt(0, up), pt(1, up)) * g(pt(1, up), b) +
                g(pt(1, up), pt(0, up)) * g(a, pt(1, up)) * g(pt(0, up), b)) *
               (g(pt(0, down), pt(0, down)) * g(pt(1, down), pt(1, down)) -
                g(pt(1, down), pt(0, down)) * g(pt(0, down), pt(1, down)));
      }
  } else {
    throw keldy_error("Not supported for order > 2");
  }
  if (std::abs(res) < cutoff) res = 0.;
  return {res, 1};
}

// --------------------------------------------------------------------------
integrator_direct::integrator_direct(integrand_g_direct integrand_, int dimension_, std::string rng_name,
                                     int rng_state_seed)
   : dimension(dimension_), integrand(std::move(integrand_)) {
  if (dimension <= 0) throw keldy_error("Dimension of the integration space must be > 0.");
  if (dimension > KD_MAX_ORDER)
    throw keldy_error("order " + std::to_string(dimension) + " exceeds the device maximum " +
                      std::to_string(KD_MAX_ORDER));
  integrand.bind(ctx_, dimension);
  reset_rng(std::move(rng_name), rng_state_seed);
}

void integrator_direct::set_precision(int precision) {
  ctx_.check(kd_set_precision(ctx_.get(), precision));
  precision_ = precision;
}

void integrator_direct::sync_warper() {
  if (warper.stamp == uploaded_stamp_) return;
  const warpers::device_chain c = warper.to_device();
  if (c.flow_layers > 0 && c.flow_stamp != uploaded_flow_stamp_) {
    ctx_.check(kd_flow_set_conditioning(ctx_.get(), c.flow_nc, c.flow_cond));
    ctx_.check(kd_flow_set_params(ctx_.get(), c.flow_order, c.flow_layers, c.flow_params.data(),
                                  int64_t(c.flow_params.size())));
    uploaded_flow_stamp_ = c.flow_stamp;
  }
  ctx_.check(kd_set_warper_chain(ctx_.get(), c.warpers.data(), int32_t(c.warpers.size()), c.maps.data(),
                                 int32_t(c.maps.size()), c.tables.data(), int64_t(c.tables.size())));
  uploaded_stamp_ = warper.stamp;
}

std::pair<warpers::warper_flow_t, std::vector<double>>
integrator_direct::train_flow_warper(int n_layers, long nr_samples, int steps, int batch, double lr, double lr_final,
                                     int rng_seed_shift, uint64_t init_seed, double weight_power) {
  sync_warper();
  const sobol gen(dimension, 0, true, true, rng_seed_shift);
  ctx_.check(kd_set_sobol(ctx_.get(), gen.dim(), gen.csr().data(), gen.shfr().data(), gen.state()));
  rng_dirty_ = true; // the next run re-uploads the integrator's own generator
  double w_mean = 0.0;
  ctx_.check(kd_flow_collect(ctx_.get(), int64_t(nr_samples), weight_power, &w_mean));
  ctx_.check(kd_flow_set_conditioning(ctx_.get(), 0, 0.0));
  ctx_.check(kd_flow_init(ctx_.get(), dimension, n_layers, init_seed));
  std::vector<double> losses(size_t(std::max(steps, 0)));
  ctx_.check(kd_flow_train(ctx_.get(), steps, batch, float(lr), float(lr_final), losses.data()));
  int64_t np = 0;
  ctx_.check(kd_flow_n_params(dimension, n_layers, &np));
  std::vector<float> p(static_cast<size_t>(np));
  ctx_.check(kd_flow_get_params(ctx_.get(), p.data(), np));
  warpers::warper_flow_t w(dimension, n_layers, std::move(p));
  // the device now holds these parameters; any flow already in the chain must
  // be uploaded again before it is used
  uploaded_flow_stamp_ = w.stamp();
  uploaded_stamp_ = 0;
  return {std::move(w), std::move(losses)};
}

std::pair<std::vector<double>, std::vector<double>>
integrator_direct::collect_flow_samples(long nr_samples, int rng_seed_shift, double weight_power) {
  sync_warper();
  const sobol gen(dimension, 0, true, true, rng_seed_shift);
  ctx_.check(kd_set_sobol(ctx_.get(), gen.dim(), gen.csr().data(), gen.shfr().data(), gen.state()));
  rng_dirty_ = true; // the next run re-uploads the integrator's own generator
  double w_mean = 0.0;
  ctx_.check(kd_flow_collect(ctx_.get(), int64_t(nr_samples), weight_power, &w_mean));
  std::vector<double> x(size_t(nr_samples) * size_t(dimension)), w(static_cast<size_t>(nr_samples));
  ctx_.check(kd_flow_get_samples(ctx_.get(), x.data(), w.data(), int64_t(nr_samples)));
  return {std::move(x), std::move(w)};
}

std::pair<warpers::warper_flow_t, std::vector<double>> integrator_direct::train_flow_warper_conditioned(
    std::vector<integrator_direct *> const &computers, std::vector<double> const &conds, int n_layers,
    long samples_per_computer, int steps, int batch, double lr, double lr_final, int rng_seed_shift,
    uint64_t init_seed, double weight_power) {
  if (computers.empty() || computers.size() != conds.size())
    throw keldy_error("one conditioning value per computer is needed");
  if (batch < 64 || batch % 64) throw keldy_error("batch must be a positive multiple of 64");
  const int dim = computers[0]->dimension;
  std::vector<double> x, w, c;
  for (size_t i = 0; i < computers.size(); i++) {
    integrator_direct &comp = *computers[i];
    if (comp.dimension != dim) throw keldy_error("the computers' dimensions differ");
    if (!std::isfinite(conds[i])) throw keldy_error("non-finite conditioning value");
    auto [xi, wi] = comp.collect_flow_samples(samples_per_computer, rng_seed_shift, weight_power);
    double s = 0.0;
    for (double v : wi) s += v;
    if (!(s > 0.0)) continue; // no weight at this value: nothing to fit
    const double scale = double(wi.size()) / s;
    for (size_t p = 0; p < wi.size(); p++) {
      if (!(wi[p] > 0.0)) continue;
      x.insert(x.end(), xi.begin() + long(p) * dim, xi.begin() + long(p + 1) * dim);
      w.push_back(wi[p] * scale);
      c.push_back(conds[i]);
    }
  }
  const size_t n_all = w.size();
  const size_t n = n_all / size_t(batch) * size_t(batch);
  if (n < size_t(batch)) throw keldy_error("fewer nonzero training samples than one batch");
  // a fixed shuffle, so that every batch mixes the conditioning values
  std::vector<size_t> perm(n_all);
  for (size_t i = 0; i < n_all; i++) perm[i] = i;
  std::mt19937_64 eng(init_seed ^ 0x9E3779B97F4A7C15ull);
  std::shuffle(perm.begin(), perm.end(), eng);
  std::vector<double> xs(n * size_t(dim)), ws(n), cs(n);
  for (size_t q = 0; q < n; q++) {
    const size_t p = perm[q];
    std::copy(x.begin() + long(p) * dim, x.begin() + long(p + 1) * dim, xs.begin() + long(q) * dim);
    ws[q] = w[p];
    cs[q] = c[p];
  }
  integrator_direct &tr = *computers[0];
  device_ctx &ctx = tr.ctx_;
  ctx.check(kd_flow_set_conditioning(ctx.get(), 1, conds[0]));
  ctx.check(kd_flow_init(ctx.get(), dim, n_layers, init_seed));
  ctx.check(kd_flow_set_samples(ctx.get(), xs.data(), ws.data(), cs.data(), int64_t(n)));
  std::vector<double> losses(size_t(std::max(steps, 0)));
  ctx.check(kd_flow_train(ctx.get(), steps, batch, float(lr), float(lr_final), losses.data()));
  int64_t np = 0;
  ctx.check(kd_flow_n_params_cond(dim, n_layers, 1, &np));
  std::vector<float> prm(static_cast<size_t>(np));
  ctx.check(kd_flow_get_params(ctx.get(), prm.data(), np));
  warpers::warper_flow_t fw(dim, n_layers, std::move(prm), 1, conds[0]);
  tr.uploaded_flow_stamp_ = fw.stamp();
  tr.uploaded_stamp_ = 0;
  return {std::move(fw), std::move(losses)};
}

std::pair<dcomplex, double> integrator_direct::evaluate_warped_integrand(std::vector<double> const &li_vec,
                                                                       int start_domain_nr, bool keep_u_hypercube) {
  if (start_domain_nr == int(warper.size())) {
    dcomplex v;
    int dom = 0;
    double jac = 0;
    evaluate_warped_integrand_batch(li_vec.data(), 1, &v, &dom, &jac, keep_u_hypercube);
    return {v, double(dom)};
  }
  // partial chain: the host maps the point, the device evaluates the integrand
  auto [ui, jac] = warper.map_reverse(li_vec, start_domain_nr, 0);
  auto [v, dom] = integrand(ui, keep_u_hypercube);
  return {v * jac, double(dom)};
}

void integrator_direct::evaluate_warped_integrand_batch(const double *l, long n, dcomplex *out, int *in_domain,
                                                        double *jac, bool keep_u_hypercube) {
  sync_warper();
  ctx_.check(kd_eval_warped_batch(ctx_.get(), l, n, keep_u_hypercube ? 1 : 0, reinterpret_cast<double *>(out),
                                  in_domain, jac));
  check_window(out, n);
}

void integrator_direct::prepare_device_run() {
  sync_warper();
  if (rng_dirty_) {
    ctx_.check(kd_set_sobol(ctx_.get(), rng_->dim(), rng_->csr().data(), rng_->shfr().data(), rng_->state()));
    rng_dirty_ = false;
  }
}

void integrator_direct::finish_device_run(long nr_steps) {
  n_points_ += uint64_t(nr_steps);
  uint32_t k = 0;
  ctx_.check(kd_get_sobol_state(ctx_.get(), &k));
  rng_->set_state(k);
}

void integrator_direct::run(long nr_steps) {
  if (nr_steps <= 0) return;
  prepare_device_run();
  kd_accum acc{};
  ctx_.check(kd_run(ctx_.get(), uint64_t(nr_steps), &acc));
  finish_device_run(nr_steps);
}

warpers::warper_product_1d_interp_nearest_t
integrator_direct::make_warper_from_proj_device(int num_bins, long nr_samples, double sigma, bool keep_u_hypercube) {
  sync_warper();
  const sobol g(dimension, 0);
  const size_t nb = size_t(dimension) * size_t(std::max(num_bins, 0));
  std::vector<double> sums(nb);
  std::vector<uint64_t> cnt(nb);
  ctx_.check(kd_pilot_histograms(ctx_.get(), g.csr().data(), g.shfr().data(), int64_t(nr_samples), num_bins,
                                 keep_u_hypercube ? 1 : 0, sums.data(), cnt.data()));
  rng_dirty_ = true; // the device generator was restored; re-upload ours before the next run anyway
  std::vector<std::vector<double>> data(static_cast<size_t>(dimension));
  std::vector<std::vector<long>> count(static_cast<size_t>(dimension));
  for (int a = 0; a < dimension; a++) {
    data[a].assign(sums.begin() + size_t(a) * num_bins, sums.begin() + size_t(a + 1) * num_bins);
    count[a].resize(size_t(num_bins));
    for (int b = 0; b < num_bins; b++) count[a][b] = long(cnt[size_t(a) * num_bins + b]);
  }
  return warpers::make_warper_from_proj_histograms(std::move(data), count, dimension, num_bins, sigma);
}

warpers::warper_projection_t integrator_direct::make_warper_projection_device(int num_bins, long nr_samples,
                                                                             double sigma, bool optimize_sigma,
                                                                             bool keep_u_hypercube) {
  if (nr_samples < 0) throw keldy_error("nr_samples must be >= 0");
  sobol g(dimension, 0);
  std::vector<std::vector<double>> l(static_cast<size_t>(nr_samples));
  std::vector<double> flat(static_cast<size_t>(nr_samples) * static_cast<size_t>(dimension));
  for (long i = 0; i < nr_samples; i++) {
    l[size_t(i)] = g();
    std::copy(l[size_t(i)].begin(), l[size_t(i)].end(), flat.begin() + i * dimension);
  }
  std::vector<dcomplex> v(static_cast<size_t>(nr_samples));
  std::vector<int> dom(static_cast<size_t>(nr_samples));
  std::vector<double> jac(static_cast<size_t>(nr_samples));
  if (nr_samples > 0) evaluate_warped_integrand_batch(flat.data(), nr_samples, v.data(), dom.data(), jac.data(),
                                                      keep_u_hypercube);
  std::vector<double> a(static_cast<size_t>(nr_samples));
  for (long i = 0; i < nr_samples; i++) a[size_t(i)] = std::abs(v[size_t(i)]);
  return warpers::warper_projection_t::from_samples(l, a, dimension, num_bins, sigma, optimize_sigma);
}

std::vector<double> integrator_direct::autotune(long sample_points) {
  prepare_device_run();
  int32_t lanes = 0, wpc = 0;
  double ms = 0;
  ctx_.check(kd_autotune(ctx_.get(), int64_t(sample_points), &lanes, &wpc, &ms));
  return {double(lanes), double(wpc), ms};
}

double run_batch(std::vector<integrator_direct *> const &its, std::vector<long> const &steps) {
  if (its.size() != steps.size()) throw keldy_error("run_batch: one point count per integrator");
  if (its.empty()) return 0.0;
  std::vector<kd_ctx *> ctxs;
  std::vector<uint64_t> ns;
  for (size_t i = 0; i < its.size(); i++) {
    if (steps[i] < 0) throw keldy_error("run_batch: negative point count");
    its[i]->prepare_device_run();
    ctxs.push_back(its[i]->ctx().get());
    ns.push_back(uint64_t(steps[i]));
  }
  double ms = 0;
  its[0]->ctx().check(kd_run_batch(ctxs.data(), ns.data(), int32_t(ctxs.size()), nullptr, &ms));
  for (size_t i = 0; i < its.size(); i++) its[i]->finish_device_run(steps[i]);
  return ms;
}

void integrator_direct::reset_rng(std::string rng_name, int rng_state_seed, bool do_shift, bool do_scramble,
                                  int rng_seed_shift) {
  if (rng_name == "sobol_unshifted") {
    rng_ = std::make_unique<sobol>(dimension, rng_state_seed);
  } else if (rng_name == "sobol") {
    rng_ = std::make_unique<sobol>(dimension, rng_state_seed, do_shift, do_scramble, rng_seed_shift);
  } else if (rng_name == "mt19937_64" || rng_name == "ranlux48" || rng_name == "taus88") {
    throw keldy_error("rng '" + rng_name + "' is a pseudo-random generator; keldy_amd generates Sobol points on the "
                      "device only ('sobol', 'sobol_unshifted')");
  } else {
    throw keldy_error("No other rng available.");
  }
  rng_dirty_ = true;
}

dcomplex integrator_direct::reduce_result() const {
  kd_accum acc{};
  ctx_.check(kd_get_accum(ctx_.get(), &acc));
  return dcomplex(acc.re, acc.im) * (1.0 / double(n_points_));
}

uint64_t integrator_direct::reduce_nr_points_in_domain() const {
  kd_accum acc{};
  ctx_.check(kd_get_accum(ctx_.get(), &acc));
  return acc.n_in_domain;
}

uint64_t integrator_direct::reduce_nr_fallback() const {
  kd_accum acc{};
  ctx_.check(kd_get_accum(ctx_.get(), &acc));
  return acc.n_fallback;
}

void integrator_direct::set_timing(int level) { ctx_.check(kd_set_timing(ctx_.get(), level)); }

std::vector<double> integrator_direct::last_timing() const {
  double t = 0, lu = 0, pr = 0;
  int32_t n = 0;
  ctx_.check(kd_last_timing(ctx_.get(), &t, &lu, &pr, &n));
  return {t, lu, pr, double(n)};
}

double integrator_direct::flops_per_point() const {
  double f = 0;
  ctx_.check(kd_flops_per_point(ctx_.get(), &f));
  return f;
}

// --------------------------------------------------------------------------
compute_charge_Q_direct::compute_charge_Q_direct(g0_model model, double time, int order, double cutoff_integrand)
   : integrator_direct(integrand_g_direct(g0_keldysh_contour_t(std::move(model)), gf_index_t(time, up, forward),
                                          gf_index_t(time, up, backward), cutoff_integrand),
                       order, "sobol_unshifted", 0) {}

// --------------------------------------------------------------------------
// time-binned charge
namespace {
std::pair<binner::sparse_binner_1d_t, int> binned(dcomplex value, int in_domain, std::vector<double> const &times) {
  binner::sparse_binner_1d_t r;
  r.accumulate(value, *std::min_element(times.begin(), times.end()));
  return {r, in_domain};
}
} // namespace

std::pair<binner::sparse_binner_1d_t, int> integrand_g_direct_time::operator()(std::vector<double> const &times,
                                                                              bool keep_u_hypercube) const {
  if (times.empty()) throw keldy_error("Order 0 is not implemented");
  auto [v, dom] = direct(times, keep_u_hypercube);
  return binned(v, dom, times);
}

compute_charge_Q_direct_time::compute_charge_Q_direct_time(g0_model model, double time, int order,
                                                           int nr_time_slices, double cutoff_integrand)
   : integrator_direct(integrand_g_direct(g0_keldysh_contour_t(model), gf_index_t(time, up, forward),
                                          gf_index_t(time, up, backward), cutoff_integrand),
                       order, "sobol_unshifted", 0),
     integrand_time(g0_keldysh_contour_t(std::move(model)), gf_index_t(time, up, forward),
                    gf_index_t(time, up, backward), cutoff_integrand),
     time_(time), nr_time_slices_(nr_time_slices) {
  if (nr_time_slices < 1) throw keldy_error("nr_time_slices must be >= 1");
  if (!(0.0 < time))
    throw keldy_error("Boundaries of binning range are in wrong order (0 >= " + std::to_string(time) + ")");
  ctx_.check(kd_set_time_bins(ctx_.get(), 0.0, time, nr_time_slices));
}

binner::binner_1d_t compute_charge_Q_direct_time::reduce_result_binned() const {
  binner::binner_1d_t b;
  b.xmin = 0.0;
  b.xmax = time_;
  b.nr_bins = nr_time_slices_;
  std::vector<double> v(2 * size_t(nr_time_slices_));
  std::vector<uint64_t> cnt(static_cast<size_t>(nr_time_slices_));
  uint64_t dropped = 0;
  ctx_.check(kd_get_time_bins(ctx_.get(), v.data(), cnt.data(), &dropped));
  const double inv = n_points_ ? 1.0 / double(n_points_) : 0.0;
  for (int i = 0; i < nr_time_slices_; i++) {
    b.data.emplace_back(v[2 * i] * inv, v[2 * i + 1] * inv);
    b.nr_values_added.push_back(long(cnt[i]));
  }
  b.nr_values_dropped = long(dropped);
  return b;
}

std::pair<binner::sparse_binner_1d_t, double>
compute_charge_Q_direct_time::evaluate_warped_integrand_time(std::vector<double> const &li_vec, int start_domain_nr,
                                                             bool keep_u_hypercube) {
  auto [v, dom] = evaluate_warped_integrand(li_vec, start_domain_nr, keep_u_hypercube);
  auto ui = warper.map_reverse(li_vec, start_domain_nr, 0).first;
  auto r = binned(v, int(dom), ui);
  return {r.first, dom};
}

compute_current_J_direct::compute_current_J_direct(g0_model model, double time, int order, double cutoff_integrand)
   : integrator_direct(integrand_g_direct(g0_keldysh_contour_t(std::move(model)), gf_index_t(time, up, forward, 0, 1),
                                          gf_index_t(time, up, backward, 0, 0), cutoff_integrand),
                       order, "sobol_unshifted", 0) {}

} // namespace impurity_oneband
} // namespace keldy

namespace keldy::impurity_oneband {

// --------------------------------------------------------------------------
// kernel method
namespace {
binner::sparse_binner_1d_1_t entries_to_sparse(const dcomplex *e, int order, std::vector<double> const &times) {
  binner::sparse_binner_1d_1_t r;
  const long nc = long(1) << order;
  r.data.reserve(size_t(nc * order));
  for (long c = 0; c < nc; c++)
    for (int i = 0; i < order; i++) r.accumulate(e[c * order + i], times[size_t(i)], (c >> i) & 1);
  return r;
}
} // namespace

std::pair<binner::sparse_binner_1d_1_t, int> integrand_g_kernel::operator()(std::vector<double> const &times,
                                                                           bool keep_u_hypercube) const {
  const int n = int(times.size());
  if (n == 0) throw keldy_error("Order 0 is not implemented");
  std::vector<dcomplex> e(size_t(n) << n);
  int dom = 0;
  eval_entries(times.data(), 1, n, keep_u_hypercube, e.data(), &dom);
  if (!dom) return {binner::sparse_binner_1d_1_t{}, 0};
  return {entries_to_sparse(e.data(), n, times), 1};
}

compute_gf_kernel::compute_gf_kernel(g0_model model, double time, int order, int nr_bins)
   : integrator_direct(integrand_g_direct::kernel_method(g0_keldysh_contour_t(model), gf_index_t(time, up, forward),
                                                         KD_OBS_KERNEL),
                       order, "sobol_unshifted", 0),
     kernel(g0_keldysh_contour_t(std::move(model)), gf_index_t(time, up, forward)), time_(time), nr_bins_(nr_bins) {
  if (nr_bins < 1) throw keldy_error("nr_bins must be >= 1");
  if (!(0.0 < time))
    throw keldy_error("Boundaries of binning range are in wrong order (0 >= " + std::to_string(time) + ")");
  ctx_.check(kd_set_kernel_bins(ctx_.get(), 0.0, time, nr_bins));
}

binner::binner_1d_1_t compute_gf_kernel::reduce_result_binned() const {
  binner::binner_1d_1_t b;
  b.xmin = 0.0;
  b.xmax = time_;
  b.nr_bins = nr_bins_;
  std::vector<double> v(4 * size_t(nr_bins_));
  std::vector<uint64_t> cnt(2 * size_t(nr_bins_));
  uint64_t dropped = 0;
  ctx_.check(kd_get_kernel_bins(ctx_.get(), v.data(), cnt.data(), &dropped));
  const double inv = n_points_ ? 1.0 / double(n_points_) : 0.0;
  for (size_t i = 0; i < 2 * size_t(nr_bins_); i++) {
    b.data.emplace_back(v[2 * i] * inv, v[2 * i + 1] * inv);
    b.nr_values_added.push_back(long(cnt[i]));
  }
  b.nr_values_dropped = long(dropped);
  return b;
}

void compute_gf_kernel::evaluate_warped_entries_batch(const double *l, long n, dcomplex *out, int *in_domain,
                                                      double *jac, bool keep_u_hypercube) {
  sync_warper();
  ctx_.check(kd_eval_kernel_batch(ctx_.get(), l, n, 1, keep_u_hypercube ? 1 : 0, reinterpret_cast<double *>(out),
                                  in_domain, jac));
  check_window(out, n * (long(1) << dimension) * dimension);
}

std::pair<binner::sparse_binner_1d_1_t, double>
compute_gf_kernel::evaluate_warped_integrand_kernel(std::vector<double> const &li_vec, int start_domain_nr,
                                                    bool keep_u_hypercube) {
  const int n = dimension;
  if (int(li_vec.size()) != n) throw keldy_error("point dimension differs from the integrand order");
  std::vector<dcomplex> e(size_t(n) << n);
  int dom = 0;
  auto [ui, jac] = warper.map_reverse(li_vec, start_domain_nr, 0);
  if (start_domain_nr == int(warper.size())) {
    double j = 0;
    evaluate_warped_entries_batch(li_vec.data(), 1, e.data(), &dom, &j, keep_u_hypercube);
  } else {
    kernel.eval_entries(ui.data(), 1, n, keep_u_hypercube, e.data(), &dom);
    for (auto &x : e) x *= jac;
  }
  if (!dom) return {binner::sparse_binner_1d_1_t{}, 0.0};
  return {entries_to_sparse(e.data(), n, ui), 1.0};
}

compute_gf_kernel_single_omega::compute_gf_kernel_single_omega(g0_model model, double time, double omega, int order)
   : integrator_direct(integrand_g_direct::kernel_method(g0_keldysh_contour_t(model), gf_index_t(time, up, forward),
                                                         KD_OBS_KERNEL_OMEGA, omega),
                       order, "sobol_unshifted", 0),
     kernel(g0_keldysh_contour_t(std::move(model)), gf_index_t(time, up, forward), omega) {}

} // namespace keldy::impurity_oneband
