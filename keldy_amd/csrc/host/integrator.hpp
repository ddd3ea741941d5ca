// integrator.hpp -- integrand and QMC driver of the host API, backed by the
// HIP path through the C-ABI (include/keldy_amd.h).
// Reference: c++/keldy/impurity_oneband/wick_direct.hpp:34-65,
// c++/keldy/integrator.hpp:42-136, impurity_oneband/compute_obs.hpp:51-80.
#pragma once

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/keldy_amd.h"
#include "model.hpp"
#include "qrng.hpp"
#include "warpers.hpp"

namespace keldy {

/// Shared owner of one device context (one HIP stream on device 0).
class device_ctx {
 public:
  device_ctx();
  kd_ctx *get() const { return ctx_.get(); }
  /// throw keldy_error with the C-ABI's message if rc != KD_OK
  void check(int rc) const;

 private:
  std::shared_ptr<kd_ctx> ctx_;
};

namespace binner {

/// One-axis binner (binner.hpp:147-352, binner_t<1>): sums and counts per bin
/// of [xmin, xmax], nr_bins equal bins (xmax falls in the last one).
struct binner_1d_t {
  double xmin = 0.0, xmax = 1.0;
  long nr_bins = 1;
  std::vector<dcomplex> data;
  std::vector<long> nr_values_added;
  long nr_values_dropped = 0;
  double get_bin_size() const { return (xmax - xmin) / double(nr_bins); }
  std::vector<double> get_bin_coord() const;
};

/// Sparse binner (binner.hpp:59-140, sparse_binner_t<1>): (coordinate, value) pairs.
struct sparse_binner_1d_t {
  std::vector<std::pair<double, dcomplex>> data;
  void accumulate(dcomplex value, double x) { data.emplace_back(x, value); }
  double sum_moduli() const;
};

/// binner_t<1,1> (binner.hpp:147-352) with one continuous axis [xmin, xmax]
/// (nr_bins bins, xmax in the last one) and one discrete axis of 2 bins (the
/// Keldysh index): data and counts [nr_bins][2], row-major.
struct binner_1d_1_t {
  double xmin = 0.0, xmax = 1.0;
  long nr_bins = 1;
  long nr_disc = 2;
  std::vector<dcomplex> data;
  std::vector<long> nr_values_added;
  long nr_values_dropped = 0;
  double get_bin_size() const { return (xmax - xmin) / double(nr_bins); }
  std::vector<double> get_bin_coord() const;
};

/// Sparse binner (binner.hpp:59-140, sparse_binner_t<1,1>): ((time, index), value).
struct sparse_binner_1d_1_t {
  std::vector<std::pair<std::pair<double, long>, dcomplex>> data;
  void accumulate(dcomplex value, double x, long k) { data.push_back({{x, k}, value}); }
  double sum_moduli() const;
};

} // namespace binner

namespace impurity_oneband {

void check_window(const dcomplex *out, long n);

/// Keldysh-sum integrand (wick_direct.hpp:34-65).  Evaluation runs on the GPU.
class integrand_g_direct {
 public:
  using result_t = dcomplex;

  integrand_g_direct(g0_keldysh_contour_t g0_, gf_index_t external_A_, gf_index_t external_B_,
                     double cutoff_ = 0.);

  /// (value, in_domain) at one vector of vertex times (wick_direct.cpp:43-158)
  std::pair<result_t, int> operator()(std::vector<double> const &times, bool keep_u_hypercube = true) const;
  /// values at many points, row-major [n][order]
  void eval_batch(const double *u, long n, int order, bool keep_u_hypercube, dcomplex *out, int *in_domain) const;
  /// closed forms of orders 0-2 by explicit Wick contractions (wick_direct.cpp:161-228)
  std::pair<result_t, int> eval_no_det(std::vector<double> const &times, bool keep_u_hypercube = true) const;

  /// The kernel method on the same device path (wick_kernel.hpp:43-60,
  /// wick_kernel_single_omega.hpp:45-58): obs = KD_OBS_KERNEL (entries per
  /// configuration and vertex) or KD_OBS_KERNEL_OMEGA (their phase-weighted
  /// sum at omega), with external point X.
  static integrand_g_direct kernel_method(g0_keldysh_contour_t g0_, gf_index_t X, int obs, double omega = 0.0);
  /// KD_OBS_KERNEL: entries [n][2^order][order] at vertex times u [n][order]
  void eval_entries(const double *u, long n, int order, bool keep_u_hypercube, dcomplex *out, int *in_domain) const;
  int observable() const { return obs_; }
  double omega() const { return omega_; }

  g0_keldysh_contour_t const &g0() const { return g0_; }
  gf_index_t const &A() const { return external_A; }
  gf_index_t const &B() const { return external_B; }
  double get_cutoff() const { return cutoff; }
  /// upload table + externals + order to a device context
  void bind(device_ctx const &ctx, int order) const;

 private:
  g0_keldysh_contour_t g0_;
  gf_index_t external_A, external_B;
  double cutoff;
  mutable std::shared_ptr<device_ctx> ctx_; // lazily created for per-point calls
  mutable int bound_order_ = -1;
  int obs_ = KD_OBS_DIRECT;
  double omega_ = 0.0;
  void ensure_bound(int order) const;
};

/// Kernel method integrand (wick_kernel.hpp:43-60): per configuration and
/// vertex, the cofactor entry at (vertex time, Keldysh index).  On the GPU.
class integrand_g_kernel {
 public:
  using result_t = binner::sparse_binner_1d_1_t;
  integrand_g_kernel(g0_keldysh_contour_t g0_, gf_index_t g_idx_X_)
     : base(integrand_g_direct::kernel_method(std::move(g0_), g_idx_X_, KD_OBS_KERNEL)) {}
  /// the reference's sparse binner: 2^n n entries in its accumulate order
  std::pair<result_t, int> operator()(std::vector<double> const &times, bool keep_u_hypercube = true) const;
  void eval_entries(const double *u, long n, int order, bool keep_u_hypercube, dcomplex *out, int *in_domain) const {
    base.eval_entries(u, n, order, keep_u_hypercube, out, in_domain);
  }
  integrand_g_direct const &device() const { return base; }

 private:
  integrand_g_direct base;
};

/// Kernel method at one frequency (wick_kernel_single_omega.hpp:45-58).
class integrand_g_kernel_single_omega {
 public:
  using result_t = dcomplex;
  integrand_g_kernel_single_omega(g0_keldysh_contour_t g0_, gf_index_t g_idx_X_, double omega_)
     : base(integrand_g_direct::kernel_method(std::move(g0_), g_idx_X_, KD_OBS_KERNEL_OMEGA, omega_)) {}
  std::pair<result_t, int> operator()(std::vector<double> const &times, bool keep_u_hypercube = true) const {
    return base(times, keep_u_hypercube);
  }
  void eval_batch(const double *u, long n, int order, bool keep_u_hypercube, dcomplex *out, int *in_domain) const {
    base.eval_batch(u, n, order, keep_u_hypercube, out, in_domain);
  }

 private:
  integrand_g_direct base;
};

/// QMC driver: Sobol points -> warper chain -> integrand -> accumulation,
/// all on the device (integrator.hpp:42-136).
class integrator_direct {
 public:
  warpers::warper_train_t warper{};

  integrator_direct(integrand_g_direct integrand_, int dimension_, std::string rng_name, int rng_state_seed);

  std::pair<dcomplex, double> evaluate_warped_integrand(std::vector<double> const &li_vec, int start_domain_nr,
                                                        bool keep_u_hypercube = true);
  std::pair<dcomplex, double> evaluate_warped_integrand(std::vector<double> const &li_vec) {
    return evaluate_warped_integrand(li_vec, int(warper.size()), true);
  }
  /// batched warped evaluation with the full chain: values and in-domain flags
  void evaluate_warped_integrand_batch(const double *l, long n, dcomplex *out, int *in_domain, double *jac,
                                       bool keep_u_hypercube = true);

  void run(long nr_steps);
  void reset_rng(std::string rng_name, int rng_state_seed, bool do_shift = false, bool do_scramble = false,
                 int rng_seed_shift = 0);
  dcomplex reduce_result() const;
  uint64_t reduce_nr_points_run() const { return n_points_; }
  uint64_t reduce_nr_points_in_domain() const;
  uint64_t reduce_nr_fallback() const;

  integrand_g_direct get_integrand() const { return integrand; }
  warpers::warper_train_t get_warper() const { return warper; }
  int get_dimension() const { return dimension; }
  void set_precision(int precision);
  int get_precision() const { return precision_; }
  /// timing of the last device call (total, LU kernels, prep kernels; ms)
  // KD_TIMING_NONE / _TOTAL / _KERNELS (the default): what last_timing measures
  void set_timing(int level);
  std::vector<double> last_timing() const;
  double flops_per_point() const;
  device_ctx const &ctx() const { return ctx_; }
  /// upload the warper chain if it changed since the last upload
  void sync_warper();
  /// Keldysh-sum kernel launch shape tuned on sample_points points of the
  /// current chain (generator and accumulator untouched): (lanes per point,
  /// waves per CU, LU ms per 2^20 points).
  std::vector<double> autotune(long sample_points);
  /// make_warper_from_proj with the pilot on the device: the warped integrand
  /// (current chain) at the first nr_samples unshifted Sobol points and the
  /// per-axis histograms are computed there; the host smooths and fits.
  /// warper_projection_t with the pilot on the device: |warped integrand|
  /// (current chain) at the first nr_samples unshifted Sobol points, the
  /// construction on the host
  warpers::warper_projection_t make_warper_projection_device(int num_bins, long nr_samples, double sigma,
                                                             bool optimize_sigma, bool keep_u_hypercube);
  warpers::warper_product_1d_interp_nearest_t make_warper_from_proj_device(int num_bins, long nr_samples, double sigma,
                                                                          bool keep_u_hypercube);
  /// device state before / after a run launched elsewhere (run_batch)
  void prepare_device_run();
  void finish_device_run(long nr_steps);
  /// Learned flow warper (no counterpart in keldy): collects nr_samples points
  /// of the current chain (scrambled Sobol, seed shift rng_seed_shift, the
  /// flow warper excluded), then trains an n_layers coupling flow on the
  /// device for `steps` Adam steps of `batch` points, learning rate cosine-
  /// decayed from lr to lr_final, toward the density |integrand|^weight_power.
  /// Returns the warper (append it last) and the per-step loss.
  std::pair<warpers::warper_flow_t, std::vector<double>> train_flow_warper(int n_layers, long nr_samples, int steps,
                                                                          int batch, double lr, double lr_final,
                                                                          int rng_seed_shift, uint64_t init_seed,
                                                                          double weight_power = 1.0);
  /// Training samples of the current chain without a flow warper: nr_samples
  /// scrambled Sobol points (seed shift rng_seed_shift) [n][dimension] and
  /// their weights |warped integrand|^weight_power.
  std::pair<std::vector<double>, std::vector<double>> collect_flow_samples(long nr_samples, int rng_seed_shift,
                                                                           double weight_power);
  /// The time-conditioned flow (no counterpart in keldy): one coupling flow
  /// with one conditioning input, trained on the samples of every computer
  /// (same dimension), computer i's samples tagged with conds[i] and their
  /// weights scaled to mean 1 per computer (so every conditional density is
  /// fitted alike); zero-weight samples dropped, the set shuffled and cut to a
  /// multiple of batch.  Trains on computers[0]'s device context.  Returns the
  /// flow at conds[0] (with_cond gives the others) and the per-step loss.
  static std::pair<warpers::warper_flow_t, std::vector<double>>
  train_flow_warper_conditioned(std::vector<integrator_direct *> const &computers, std::vector<double> const &conds,
                                int n_layers, long samples_per_computer, int steps, int batch, double lr,
                                double lr_final, int rng_seed_shift, uint64_t init_seed, double weight_power);

 protected:
  int dimension = 1;
  integrand_g_direct integrand;
  device_ctx ctx_;
  std::unique_ptr<sobol> rng_;
  bool rng_dirty_ = true;
  uint64_t n_points_ = 0;
  int precision_ = KD_PREC_FP64;
  uint64_t uploaded_stamp_ = 0;
  uint64_t uploaded_flow_stamp_ = 0;
};

/// Keldysh-sum integrand binned by the earliest vertex time
/// (wick_direct_time.hpp:35-60, wick_direct_time.cpp:42-163).
class integrand_g_direct_time {
 public:
  using result_t = binner::sparse_binner_1d_t;
  integrand_g_direct_time(g0_keldysh_contour_t g0_, gf_index_t external_A_, gf_index_t external_B_,
                          double cutoff_ = 0.)
     : direct(std::move(g0_), external_A_, external_B_, cutoff_) {}
  std::pair<result_t, int> operator()(std::vector<double> const &times, bool keep_u_hypercube = true) const;
  integrand_g_direct const &base() const { return direct; }

 private:
  integrand_g_direct direct;
};

/// Time-resolved charge: the charge integrand binned by the earliest vertex
/// time into nr_time_slices bins of [0, time] (compute_obs.hpp:85-101).
class compute_charge_Q_direct_time : public integrator_direct {
 public:
  compute_charge_Q_direct_time(g0_model model, double time, int order, int nr_time_slices, double cutoff_integrand);
  compute_charge_Q_direct_time(model_param_t params, double time, int order, int nr_time_slices,
                               double cutoff_integrand)
     : compute_charge_Q_direct_time(g0_model{g0_model_omega{params}, false}, time, order, nr_time_slices,
                                    cutoff_integrand) {}
  /// binner_t<1> of the estimate: bin sums divided by the points run
  binner::binner_1d_t reduce_result_binned() const;
  integrand_g_direct_time get_integrand_time() const { return integrand_time; }
  std::pair<binner::sparse_binner_1d_t, double> evaluate_warped_integrand_time(std::vector<double> const &li_vec,
                                                                              int start_domain_nr,
                                                                              bool keep_u_hypercube = true);

 private:
  integrand_g_direct_time integrand_time;
  double time_;
  int nr_time_slices_;
};

/// Kernel method: G(X) kernel binned by vertex time and Keldysh index
/// (compute_obs.hpp:105-117): binner_t<1,1> over [0, time] x {forward, backward}.
class compute_gf_kernel : public integrator_direct {
 public:
  compute_gf_kernel(g0_model model, double time, int order, int nr_bins = 100);
  compute_gf_kernel(model_param_t params, double time, int order, int nr_bins = 100)
     : compute_gf_kernel(g0_model{g0_model_omega{params}, false}, time, order, nr_bins) {}
  /// binner_t<1,1> of the estimate: bin sums divided by the points run
  binner::binner_1d_1_t reduce_result_binned() const;
  integrand_g_kernel get_integrand_kernel() const { return kernel; }
  std::pair<binner::sparse_binner_1d_1_t, double> evaluate_warped_integrand_kernel(std::vector<double> const &li_vec,
                                                                                  int start_domain_nr,
                                                                                  bool keep_u_hypercube = true);
  /// entries [n][2^order][order], in-domain flags and Jacobians of the full chain at l [n][order]
  void evaluate_warped_entries_batch(const double *l, long n, dcomplex *out, int *in_domain, double *jac,
                                     bool keep_u_hypercube = true);

 private:
  integrand_g_kernel kernel;
  double time_;
  int nr_bins_;
};

/// Kernel method at one frequency (compute_obs.hpp:120-133).
class compute_gf_kernel_single_omega : public integrator_direct {
 public:
  compute_gf_kernel_single_omega(g0_model model, double time, double omega, int order);
  compute_gf_kernel_single_omega(model_param_t params, double time, double omega, int order)
     : compute_gf_kernel_single_omega(g0_model{g0_model_omega{params}, false}, time, omega, order) {}
  integrand_g_kernel_single_omega get_integrand_omega() const { return kernel; }

 private:
  integrand_g_kernel_single_omega kernel;
};

/// Several integrators (e.g. orders 1..N with uneven point counts) run as one
/// hipGraph launch; each advances and accumulates as by its own run().
/// Returns the launch's wall time in ms.  No counterpart in keldy.
double run_batch(std::vector<integrator_direct *> const &integrators, std::vector<long> const &nr_steps);

/// Charge Q = G^<_{up,up}(t, t) (compute_obs.hpp:51-64).
class compute_charge_Q_direct : public integrator_direct {
 public:
  compute_charge_Q_direct(g0_model model, double time, int order, double cutoff_integrand);
  compute_charge_Q_direct(model_param_t params, double time, int order, double cutoff_integrand)
     : compute_charge_Q_direct(g0_model{g0_model_omega{params}, false}, time, order, cutoff_integrand) {}
};

/// Current J ~ Re G^<_{lead,dot}(t, t) (compute_obs.hpp:67-80).
class compute_current_J_direct : public integrator_direct {
 public:
  compute_current_J_direct(g0_model model, double time, int order, double cutoff_integrand);
  compute_current_J_direct(model_param_t params, double time, int order, double cutoff_integrand)
     : compute_current_J_direct(g0_model{g0_model_omega{params}, true}, time, order, cutoff_integrand) {}
};

} // namespace impurity_oneband
} // namespace keldy
