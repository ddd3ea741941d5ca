/*
 * keldy_amd.h -- C-ABI of the MI355X Keldysh-sum integrand and QMC driver.
 *
 * Plain pointers and sizes only.  One kd_ctx owns one HIP device, one HIP
 * stream and the device buffers of one integrator; it is not thread-safe
 * (neither is keldy's integrator).  Every call returns KD_OK (0) or a
 * negative KD_ERR_*; kd_last_error() gives the message.  Calls are
 * synchronous at return unless stated otherwise.
 *
 * Reference interfaces replaced (macekmar/keldy, paths relative to its root):
 *   kd_set_g0_table        g0_keldysh_contour_t (c++/keldy/impurity_oneband/model.hpp:238-268)
 *                          + TRIQS gf<retime> evaluation of g0_lesser/g0_greater
 *   kd_set_integrand       integrand_g_direct ctor (impurity_oneband/wick_direct.hpp:48-60)
 *   kd_eval_batch          integrand_g_direct::operator() (impurity_oneband/wick_direct.cpp:43-158)
 *   kd_set_warper_chain    warper_train_t (c++/keldy/warpers/warpers.hpp:41-137)
 *   kd_eval_warped_batch   integrator::evaluate_warped_integrand (c++/keldy/integrator.hpp:68-79)
 *   kd_set_sobol           integrator::reset_rng / keldy::sobol (integrator.hpp:97-112, qrng.hpp:49-114)
 *   kd_run                 integrator::run (integrator.hpp:81-95)
 *   kd_get_accum/kd_reset_accum  reduce_result / reduce_nr_points_* (integrator.hpp:123-131)
 *   kd_set_kernel_integrand integrand_g_kernel ctor (impurity_oneband/wick_kernel.hpp:56)
 *   kd_eval_kernel_batch   integrand_g_kernel::operator() (wick_kernel.cpp:42-138)
 *   kd_set_kernel_bins / kd_get_kernel_bins  compute_gf_kernel's binner_t<1,1>
 *                          (compute_obs.hpp:105-117, binner.hpp:147-352)
 *   kd_set_kernel_omega_integrand  integrand_g_kernel_single_omega ctor and
 *                          operator() (wick_kernel_single_omega.hpp:45-58, .cpp:41-140)
 */
#ifndef KELDY_AMD_H
#define KELDY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KD_OK 0
#define KD_ERR_ARG (-1)      /* invalid argument (maps to RuntimeError in the host API) */
#define KD_ERR_HIP (-2)      /* HIP runtime failure */
#define KD_ERR_STATE (-3)    /* required set-up call missing */
#define KD_ERR_RANGE (-4)    /* Sobol generator exhausted (more than 2^31 points) */
#define KD_ERR_NODEVICE (-5) /* no HIP device */

#define KD_MAX_ORDER 12

/* LU precision of the Keldysh sum; the sum itself and the QMC accumulation are
 * always FP64.  KD_PREC_FP32_LU keeps the FP64 LU from order 10 on (and for
 * configurations flagged by the pivot check at any order); KD_PREC_FP32_LU_ALL
 * uses FP32 at every order. */
#define KD_PREC_FP64 0
#define KD_PREC_FP32_LU 1
#define KD_PREC_FP32_LU_ALL 2

/* Keldysh-sum algorithm.  KD_ALGO_DEFAULT (a new ctx's setting): KD_ALGO_CTREE
 * at orders 6..10, KD_ALGO_CONFIG otherwise -- the fastest on MI355X for
 * full-size launches; kd_autotune measures every candidate instead.
 * KD_ALGO_CONFIG: one bordered LU per
 * configuration sigma (2^n of them, sharing the staged Wick matrix).
 * KD_ALGO_TREE (orders >= 6; lower orders use KD_ALGO_CONFIG): the
 * configurations share the Schur complements of their common leading vertex
 * choices -- one table per node of the binary tree of choices -- so order 8
 * needs ~5x fewer flops, at lower issue efficiency; kd_autotune picks between
 * them.  Same arithmetic up to reassociation (FP64 results agree to ~1e-13). */
#define KD_ALGO_DEFAULT (-1)
#define KD_ALGO_CONFIG 0
#define KD_ALGO_TREE 1
/* KD_ALGO_PAIR (orders >= 7 at 64 lanes per point; otherwise KD_ALGO_CONFIG):
 * the per-configuration LU with the two Keldysh indices of the last vertex
 * sharing the elimination of the other vertex rows -- a third fewer
 * multiply-adds at order 8, same arithmetic per configuration. */
#define KD_ALGO_PAIR 2
/* KD_ALGO_SPLIT (orders 6..10; otherwise KD_ALGO_CONFIG): the tree of Schur
 * complements over the leading 2 (orders 6, 7) or 3 vertices, shared by the
 * configurations below each node, then one bordered LU per configuration over
 * the remaining vertices (at orders 7 and 8 two configurations per thread,
 * sharing the elimination of all but the last vertex); one point per
 * workgroup.  Order 8: ~3.9x fewer multiply-adds than KD_ALGO_CONFIG, same
 * arithmetic up to reassociation (FP64 results agree to ~1e-13). */
#define KD_ALGO_SPLIT 3
/* KD_ALGO_WSPLIT (orders 6..10; otherwise KD_ALGO_CONFIG): the split
 * algorithm with one point per wavefront and no workgroup barriers; the tree
 * covers the leading n - 5 vertices from order 8 (2 below), its last two
 * levels formed in one step, so each configuration finishes with a bordered
 * LU over five vertices (four at order 6), the two choices of the last vertex
 * paired.  Same arithmetic up to reassociation (FP64 results agree to
 * ~1e-13). */
#define KD_ALGO_WSPLIT 4
/* KD_ALGO_CTREE (orders 6..10; otherwise KD_ALGO_CONFIG): the time-ordered
 * tree.  With the vertices sorted by time, compare_3way (common.cpp:47-57)
 * makes each Wick entry between two vertices depend only on the Keldysh index
 * of the earlier one, so eliminating the vertices in time order subtracts from
 * every remaining entry a correction that does not depend on the later
 * vertices' indices: one value per vertex pair and tree node instead of four.
 * Orders 8-10: 2.2-2.7x fewer flops than KD_ALGO_WSPLIT; same determinants up
 * to reassociation (FP64 results agree to ~1e-13 at order 8, 1e-12 at 10). */
#define KD_ALGO_CTREE 5

typedef struct kd_ctx kd_ctx;

/* gf_index_t (model.hpp:61-76) */
typedef struct {
  double time;
  int32_t k_idx;   /* 0 forward, 1 backward */
  int32_t spin;    /* 0 up, 1 down */
  int32_t orbital; /* 0 dot, 1 lead (needs a 2-orbital table) */
  int32_t pad;
} kd_gf_index;

/* One-dimensional inverse-CDF maps (warpers/product_1d_simple.hpp, product_1d.hpp) */
#define KD_MAP_UNIFORM 0     /* f = 1/u_max */
#define KD_MAP_EXPONENTIAL 1 /* make_product_1d_simple_exponential (product_1d_simple.hpp:560-576) */
#define KD_MAP_INVERSE 2     /* make_product_1d_simple_inverse (578-592) */
#define KD_MAP_INVERSE_SQ 3  /* make_product_1d_simple_inverse_square (594-608) */
#define KD_MAP_INVERSE_POW 4 /* make_product_1d_simple_inverse_power (610-630) */
#define KD_MAP_NEAREST 5     /* warper_product_1d_simple_interp_nearest_t (174-379) */
#define KD_MAP_HYBRID 6      /* warper_product_1d_simple_interp_hybrid_t (381-554), f = w^p/(w+u)^p */
/* warper_projection_t (warpers/projection.hpp:176-382), one axis: u = Finv(l) and
 * f = fn(u) / norm, Finv (n_inv values at t_off) and fn (n_pts values at f_off)
 * on uniform meshes of [0, 1], each evaluated by linear interpolation */
#define KD_MAP_PROJECTION 7

typedef struct {
  int32_t kind;
  int32_t power;
  double u_max, w;
  double norm;   /* analytic maps: the maker's normalisation; HYBRID: f1 normalisation */
  int32_t n_pts; /* tabulated maps: knots of the forward spline */
  int32_t n_inv; /* tabulated maps: knots of the inverse spline (cut_index + 1) */
  int64_t t_off; /* offsets (in doubles) into the chain's table buffer */
  int64_t f_off; /* NEAREST: normalised f1 at knots */
  int64_t F_off; /* integrated, normalised, negated CDF knots in [-1, 0] */
  int64_t c_off; /* HYBRID: Steffen coefficients of F^-1, (n_inv-1) x 4 */
} kd_map1d;

#define KD_WARP_IDENTITY 0 /* warper_identity_t (warpers/identity.hpp:30-49) */
#define KD_WARP_PLASMA 1   /* warper_plasma_uv_t (warpers/plasma_uv.hpp:35-81) */
#define KD_WARP_PRODUCT 2  /* product of 1D maps: n_maps == 1 (same map on every axis) or == order */
#define KD_WARP_FLOW 3     /* learned affine-coupling flow (no reference counterpart) */

typedef struct {
  int32_t kind;
  int32_t n_maps;
  double t_max;
  int64_t map_offset; /* PRODUCT: index of the first kd_map1d of this warper */
} kd_warper_desc;

typedef struct {
  double re, im;         /* sum of warped integrand values */
  uint64_t n_points;     /* points run */
  uint64_t n_in_domain;  /* points inside the integration domain */
  uint64_t n_fallback;   /* Keldysh configurations recomputed by the pivoted FP64 LU */
} kd_accum;

int kd_create(kd_ctx **out, int device);
int kd_destroy(kd_ctx *ctx);
const char *kd_last_error(const kd_ctx *ctx);
int kd_device_count(int *count);

/* g0 tables on the time mesh t_k = -time_max + k*dt, dt = 2 time_max/(n_pts-1).
 * lesser/greater: [n_orb][n_orb][n_pts] complex as interleaved (re, im) doubles.
 * alpha: the alpha shift of equal-time internal contractions (model.hpp:263-264). */
int kd_set_g0_table(kd_ctx *ctx, const double *lesser, const double *greater, int32_t n_orb,
                    int32_t n_pts, double time_max, double alpha);

/* External points A, B and cutoff of integrand_g_direct; order = number of vertices. */
int kd_set_integrand(kd_ctx *ctx, int32_t order, const kd_gf_index *A, const kd_gf_index *B,
                     double cutoff);
int kd_set_precision(kd_ctx *ctx, int32_t precision);

/* Integrand at n_points vertex-time vectors u (row-major n_points x order, host memory).
 * out: n_points x (re, im); in_domain: n_points (may be NULL).  A g0 lookup
 * outside [-time_max, time_max] yields NaN (the reference throws). */
int kd_eval_batch(kd_ctx *ctx, const double *u, int64_t n_points, int32_t keep_u_hypercube,
                  double *out, int32_t *in_domain);
/* Same with device pointers (inputs already in HBM); asynchronous on the ctx stream. */
int kd_eval_batch_device(kd_ctx *ctx, const double *d_u, int64_t n_points, int32_t keep_u_hypercube,
                         double *d_out, int32_t *d_in_domain);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): kd_eval_batch_device with the same arguments and results,
 * but its compacted-point buffer is a static device array of the library
 * (KD_NOALLOC_MAX_POINTS points at order KD_MAX_ORDER, about 3.5 MB, held by
 * each device the library runs on until the process ends) instead of ctx
 * memory; the call allocates nothing.  n_points above KD_NOALLOC_MAX_POINTS
 * returns KD_ERR_RANGE.  Every call on one device shares that array: calls
 * that overlap on a device (other contexts or streams) race.  Asynchronous on
 * the ctx stream, as kd_eval_batch_device. */
#define KD_NOALLOC_MAX_POINTS 32768
int kd_eval_batch_device_noalloc(kd_ctx *ctx, const double *d_u, int64_t n_points, int32_t keep_u_hypercube,
                                 double *d_out, int32_t *d_in_domain);

/* Warper chain, in the order warpers were added (applied last-to-first).
 * maps: all kd_map1d of PRODUCT warpers; tables: the doubles they index.
 * At most KD_MAX_WARPERS warpers and KD_MAX_MAPS maps (KD_ERR_ARG beyond):
 * the device keeps a chain's descriptors in on-chip memory. */
#define KD_MAX_WARPERS 16
#define KD_MAX_MAPS 64
int kd_set_warper_chain(kd_ctx *ctx, const kd_warper_desc *warpers, int32_t n_warpers,
                        const kd_map1d *maps, int32_t n_maps, const double *tables,
                        int64_t n_tables);
/* Warped integrand (times Jacobian) at n_points hypercube points l (host memory).
 * keep_u_hypercube = 0 evaluates also at negative vertex times (as the
 * reference's evaluate_warped_integrand(l, start, false)).  A g0 lookup
 * outside [-time_max, time_max] yields NaN (the reference throws). */
int kd_eval_warped_batch(kd_ctx *ctx, const double *l, int64_t n_points, int32_t keep_u_hypercube,
                         double *out, int32_t *in_domain, double *jacobian);

/* Sobol generator: Csr = bit-reversed (scrambled) generating matrices [dim][31],
 * shfr = bit-reversed digital shift [dim]; state k (next point is k+1). */
int kd_set_sobol(kd_ctx *ctx, int32_t dim, const uint32_t *Csr, const uint32_t *shfr, uint32_t k);
int kd_get_sobol_state(const kd_ctx *ctx, uint32_t *k);

/* Run n_steps Sobol points through warper chain and integrand, accumulating on
 * the device; returns the accumulator totals since the last kd_reset_accum. */
int kd_run(kd_ctx *ctx, uint64_t n_steps, kd_accum *out);
int kd_get_accum(kd_ctx *ctx, kd_accum *out);
int kd_reset_accum(kd_ctx *ctx);

/* Per-point samples of the warped integrand at the next n_steps Sobol points
 * (advances the generator): l (n x order), value (n x 2). Device pointers. */
int kd_sample_device(kd_ctx *ctx, uint64_t n_steps, double *d_l, double *d_value);

/* Timing of the last kd_run / kd_eval_* / kd_sample_device call, from HIP
 * events on the ctx stream: whole call, the Keldysh-sum (LU) kernels and the
 * point-preparation kernels, and the number of kernel launches made.  What is
 * timed is set by kd_set_timing (no counterpart in the reference):
 * KD_TIMING_KERNELS (the default) all three; KD_TIMING_TOTAL the whole call
 * only (lu_ms = prep_ms = 0); KD_TIMING_NONE nothing (all 0): each timing
 * event is a queue packet, and on a run of ~5e4 points the events of
 * KD_TIMING_KERNELS cost ~20 us of wall time. */
#define KD_TIMING_NONE 0
#define KD_TIMING_TOTAL 1
#define KD_TIMING_KERNELS 2
int kd_set_timing(kd_ctx *ctx, int32_t level);
int kd_last_timing(const kd_ctx *ctx, double *total_ms, double *lu_ms, double *prep_ms, int32_t *launches);
/* Executed FP operations per point of the last configured order/precision
 * (the algorithm's own count, used for roofline figures). */
int kd_flops_per_point(const kd_ctx *ctx, double *flops);

/* ---- pilot projection on the device ---------------------------------------
 * make_warper_from_proj's data gathering (warpers/make_warper_from_proj.hpp:80-114):
 * the warped integrand with the ctx's chain at the first n_samples points of
 * the pilot generator (Csr, shfr: normally the unshifted sobol(order, 0)),
 * then per axis a and bin b of [0, 1] (num_bins bins) the sum of |value| over
 * the samples whose coordinate a falls in b, and their count:
 * sums, counts [order][num_bins].  The ctx's own generator is restored. */
int kd_pilot_histograms(kd_ctx *ctx, const uint32_t *Csr, const uint32_t *shfr, int64_t n_samples, int32_t num_bins,
                        int32_t keep_u_hypercube, double *sums, uint64_t *counts);

/* ---- time binning --------------------------------------------------------
 * Replaces integrand_g_direct_time (impurity_oneband/wick_direct_time.cpp:42-163)
 * with the integrator's binner_t<1> result (binner.hpp:147-352): in kd_run,
 * each point's warped value is also added to the bin of its earliest vertex
 * time (bins of width (xmax - xmin) / nbins on [xmin, xmax], xmax in the last
 * bin), and every point, in domain or not, counts one value added to that bin
 * or one dropped value when its earliest time is outside [xmin, xmax].
 * nbins in [0, 3072]; 0 switches binning off.  values: nbins x (re, im) sums (divide by
 * the points run for the estimate); counts: nbins. */
int kd_set_time_bins(kd_ctx *ctx, double xmin, double xmax, int32_t nbins);
int kd_get_time_bins(kd_ctx *ctx, double *values, uint64_t *counts, uint64_t *dropped);

/* ---- launch shape and batched runs ---------------------------------------
 * No counterpart in the reference (its integrator runs one order per object).
 * kd_autotune times the Keldysh-sum kernel on sample_points Sobol points of
 * the ctx's chain (from the current generator state, which does not advance;
 * the accumulator is untouched) for every candidate -- the split, paired and
 * tree algorithms and the per-configuration kernel at 64, 32 or 16 lanes per
 * point, each at 16, 32, 64 or 128 waves per CU -- and keeps the fastest for
 * this ctx (lanes = 0 when the tree or the split algorithm wins).  Without
 * tuning or kd_set_launch_shape (waves_per_cu in [1, 128]) the grid holds 128
 * waves per CU for KD_ALGO_CTREE, 32 for KD_ALGO_WSPLIT, 16 otherwise.  kd_run_batch runs several
 * ctxs (e.g. orders 1..10, uneven point counts) as ONE hipGraph launch: one
 * graph branch per ctx (its persistent kernels on its own stream), forked from
 * and joined to ctxs[0]'s stream; each ctx advances and accumulates as by
 * kd_run.  ms = wall time of the graph launch. */
int kd_autotune(kd_ctx *ctx, int64_t sample_points, int32_t *lanes, int32_t *waves_per_cu, double *ms_per_mpoint);
int kd_set_launch_shape(kd_ctx *ctx, int32_t lanes, int32_t waves_per_cu);
int kd_set_algorithm(kd_ctx *ctx, int32_t algo);
int kd_get_algorithm(const kd_ctx *ctx, int32_t *algo);
int kd_get_launch_shape(const kd_ctx *ctx, int32_t *lanes, int32_t *waves_per_cu);
int kd_run_batch(kd_ctx *const *ctxs, const uint64_t *n_steps, int32_t count, kd_accum *out, double *ms);

/* ---- learned coupling-flow warper (KD_WARP_FLOW) -------------------------
 * No counterpart in the reference: an importance map learned from |integrand|
 * samples, used as the last warper of a chain (applied first to the Sobol
 * point).  Map, Jacobian, parameter layout and loss: keldy_amd/csrc/kernels/
 * kd_flow_device.hpp.  One flow per ctx, of dimension `order` with n_layers
 * affine-coupling layers of hidden width 64. */
int kd_flow_n_params(int32_t order, int32_t n_layers, int64_t *n);
/* identity-initialised flow (W3 = b3 = 0; He-uniform W1, W2 from seed); resets Adam */
int kd_flow_init(kd_ctx *ctx, int32_t order, int32_t n_layers, uint64_t seed);
int kd_flow_set_params(kd_ctx *ctx, int32_t order, int32_t n_layers, const float *params, int64_t n);
int kd_flow_get_params(kd_ctx *ctx, float *params, int64_t n);
/* x = T(l) and dx/dl at n points (host memory, row-major n x order) */
int kd_flow_map(kd_ctx *ctx, const double *l, int64_t n, double *x, double *jac);
/* loss = mean_p (w_p / w_mean) (-log q(x_p)) over the n <= 2^18 points and its
 * gradient (n_params floats), q the flow's density; host memory */
int kd_flow_loss_grad(kd_ctx *ctx, const double *x, const double *w, int64_t n, double w_mean, float *grad,
                      double *loss);
/* Training set: the next n_samples Sobol points of the ctx's generator (which
 * advances) pushed through the warper chain WITHOUT the flow; keeps the points
 * and the weights w = |warped integrand|^weight_power on the device (1: the
 * flow's density targets |integrand|; below 1: a tempered target); returns
 * the mean weight (0 when every sample is zero: kd_flow_train then fails). */
int kd_flow_collect(kd_ctx *ctx, int64_t n_samples, double weight_power, double *w_mean);
/* Adam steps on consecutive aligned batches of the training set (batch | n_samples,
 * 64 | batch, batch <= 2^16), learning rate cosine-decayed from lr to lr_final; per-step loss
 * into losses (steps doubles, may be NULL). */
int kd_flow_train(kd_ctx *ctx, int32_t steps, int32_t batch, float lr, float lr_final, double *losses);
/* The time-conditioned flow (one learned warper for the runs at many external
 * times, the time an extra input feature; no reference counterpart): with
 * n_cond = 1 every coupling layer's conditioner also takes one value c,
 * h1 = relu(W1 a + Wc c + b1) (parameter layout in kd_flow_device.hpp), so one
 * parameter set holds a family of maps T_c.  Set before kd_flow_init /
 * kd_flow_set_params (the parameter count depends on n_cond); `value` is the c
 * the ctx's map and chain use (may be changed at any time). */
int kd_flow_set_conditioning(kd_ctx *ctx, int32_t n_cond, double value);
int kd_flow_n_params_cond(int32_t order, int32_t n_layers, int32_t n_cond, int64_t *n);
/* kd_flow_loss_grad with one conditioning value per point (cond: n doubles) */
int kd_flow_loss_grad_cond(kd_ctx *ctx, const double *x, const double *w, const double *cond, int64_t n,
                           double w_mean, float *grad, double *loss);
/* The training set of the last kd_flow_collect, copied to host memory
 * (x: n x order, w: n; n = the collected count). */
int kd_flow_get_samples(kd_ctx *ctx, double *x, double *w, int64_t n);
/* A training set from host memory, replacing the collected one: points x
 * (n x order, in [0, 1]), weights w >= 0 (w_mean = their mean), and per point
 * its conditioning value (cond, n doubles; NULL: the ctx's value). */
int kd_flow_set_samples(kd_ctx *ctx, const double *x, const double *w, const double *cond, int64_t n);

/* ---- kernel method ---------------------------------------------------------
 * Observables of the integrator besides the Keldysh sum of kd_set_integrand.
 * KD_OBS_KERNEL (integrand_g_kernel, wick_kernel.cpp:42-138): per point and
 * Keldysh configuration c of the `order` vertices, one complex entry per vertex
 * i -- the cofactor of the external row of the same-spin Wick matrix (external
 * point X first, the last vertex's diagonal entry set to zero as in the
 * reference), times c's parity and the opposite-spin determinant.  Entry (c, i)
 * belongs to vertex time u_i and Keldysh index bit i of c.  kd_run adds every
 * entry times the warper Jacobian to the bin (time u_i, index) of
 * kd_set_kernel_bins (binner_t<1,1>: time bins of width (xmax - xmin) / nbins,
 * xmax in the last one, and 2 index bins); the accumulator's re, im hold the
 * sum of all entries.  KD_OBS_KERNEL_OMEGA (integrand_g_kernel_single_omega):
 * the scalar sum over entries of value * exp(i omega u_i); every scalar entry
 * point (kd_eval_batch, kd_eval_warped_batch, kd_run, kd_sample_device) works
 * with it as with the direct integrand.  Both are FP64 (kd_set_precision does
 * not apply) and use the pivot check / pivoted fallback of the Keldysh sum.
 * kd_set_integrand switches back to the direct integrand. */
#define KD_OBS_DIRECT 0
#define KD_OBS_KERNEL 1
#define KD_OBS_KERNEL_OMEGA 2

/* external point X (its split is set to `order`), order in [1, KD_MAX_ORDER];
 * clears the bins */
int kd_set_kernel_integrand(kd_ctx *ctx, int32_t order, const kd_gf_index *X);
int kd_set_kernel_omega_integrand(kd_ctx *ctx, int32_t order, const kd_gf_index *X, double omega);
/* nbins in [0, 1536] (0: no binning) */
int kd_set_kernel_bins(kd_ctx *ctx, double xmin, double xmax, int32_t nbins);
/* values [nbins][2] (re, im) sums since the last kd_reset_accum (divide by the
 * points run for the estimate); counts [nbins][2] entries added; dropped:
 * entries whose vertex time was outside [xmin, xmax] */
int kd_get_kernel_bins(kd_ctx *ctx, double *values, uint64_t *counts, uint64_t *dropped);
/* Entries at n_points points (host memory): pts are vertex times u (warped =
 * 0) or hypercube points l pushed through the warper chain (warped = 1, the
 * entries then carry the Jacobian).  out: [n_points][2^order][order] (re, im),
 * zero for points out of the domain; in_domain, jacobian [n_points] (may be
 * NULL).  At most 2^26 entries per call. */
int kd_eval_kernel_batch(kd_ctx *ctx, const double *pts, int64_t n_points, int32_t warped, int32_t keep_u_hypercube,
                         double *out, int32_t *in_domain, double *jacobian);

#ifdef __cplusplus
}
#endif
#endif
