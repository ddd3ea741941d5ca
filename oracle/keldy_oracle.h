/*
 * keldy_oracle.h -- CPU oracle for the Keldysh-sum integrand and QMC driver.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, serial restatement of
 * the reference algorithm (macekmar/keldy), used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * The product (keldy_amd) never links or calls it.
 *
 * Each function cites the reference file:line it restates.  Where the
 * reference delegates to a third-party library (TRIQS determinant -> LAPACK
 * zgetrf, TRIQS gf<retime> evaluation, GSL splines) the restated algorithm is
 * named in the .c file.
 */
#ifndef KELDY_ORACLE_H
#define KELDY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KO_MAX_DIM 32
#define KO_SOBOL_M 31

/* ---- mt19937_64 (std::mt19937_64 semantics) ---------------------------- */
typedef struct {
  uint64_t mt[312];
  int mti;
} ko_mt64;
void ko_mt64_seed(ko_mt64 *g, uint64_t seed);
uint64_t ko_mt64_next(ko_mt64 *g);

/* ---- Sobol (qrng.hpp:49-114, digitalseq_b2g.hpp:295-547) --------------- */
typedef struct {
  int dim;
  uint32_t k;
  uint32_t Csr[KO_MAX_DIM * KO_SOBOL_M]; /* bit-reversed columns */
  uint32_t shfr[KO_MAX_DIM];             /* bit-reversed digital shift */
  uint32_t cur[KO_MAX_DIM];
} ko_sobol;

/* Unscrambled JK2008 generating matrices: Cs[j*m + c], LSB = top row. */
int ko_jk2008_matrices(uint32_t *Cs, int dim, int m);
int ko_sobol_init(ko_sobol *s, int dim, uint32_t state_seed, int do_shift, int do_scramble,
                  int64_t rng_seed_shift);
void ko_sobol_set_state(ko_sobol *s, uint32_t k);
/* returns 0, or -1 when the generator is past its end (k >= 2^31) */
int ko_sobol_next(ko_sobol *s, double *x);

/* ---- g0 on the Keldysh contour (model.hpp:238-268, common.cpp:47-57) --- */
typedef struct {
  double time;
  int k_idx;     /* 0 forward, 1 backward */
  int timesplit; /* contour point splitting order */
  int spin;      /* 0 up, 1 down */
  int orbital;   /* 0 dot, 1 lead */
} ko_gf_index;

typedef struct {
  int n_pts;       /* odd number of time mesh points */
  int n_orb;       /* 1 (dot only) or 2 (dot + lead) */
  double t_min;    /* = -time_max */
  double dt;       /* = 2 time_max / (n_pts - 1) */
  double alpha;
  const double *lesser;  /* [n_orb][n_orb][n_pts][2] (re, im) */
  const double *greater; /* same layout */
} ko_g0;

int ko_compare_3way(const ko_gf_index *a, const ko_gf_index *b);
void ko_g0_eval(const ko_g0 *g, const ko_gf_index *a, const ko_gf_index *b, int internal_point,
                double out[2]);

/* ---- warpers (warpers/*.hpp) ------------------------------------------- */
enum {
  KO_MAP_UNIFORM = 0,  /* f = 1/u_max (product_1d_simple_test.cpp:10-23 form) */
  KO_MAP_EXPONENTIAL,  /* make_product_1d_simple_exponential */
  KO_MAP_INVERSE,      /* make_product_1d_simple_inverse */
  KO_MAP_INVERSE_SQ,   /* make_product_1d_simple_inverse_square */
  KO_MAP_INVERSE_POW,  /* make_product_1d_simple_inverse_power */
  KO_MAP_NEAREST,      /* warper_product_1d_simple_interp_nearest_t */
  KO_MAP_HYBRID        /* warper_product_1d_simple_interp_hybrid_t */
};

typedef struct {
  int kind;
  int power;
  double u_max, w;
  /* tabulated maps (NEAREST / HYBRID); owned by the caller */
  int n_pts;          /* knots of the forward spline */
  int n_inv;          /* knots of the inverse spline (cut_index + 1) */
  const double *t;    /* knot times [n_pts] */
  const double *f;    /* NEAREST: normalised f1 at knots [n_pts] */
  const double *F;    /* integrated, normalised, in [-1, 0] [n_pts] */
  const double *inv_c;/* HYBRID: Steffen coeffs of F^-1 [n_inv-1][4] (a,b,c,d) */
  double norm;        /* HYBRID: f1 normalisation; f1(u) = w^p/(w+u)^p / norm */
} ko_map1d;

enum { KO_WARP_IDENTITY = 0, KO_WARP_PLASMA = 1, KO_WARP_PRODUCT = 2 };
typedef struct {
  int kind;
  double t_max;           /* PLASMA */
  const ko_map1d *maps;   /* PRODUCT: one per dimension */
  int n_maps;
} ko_warper;

/* warper_train_t::map_reverse (warpers.hpp:72-85): applies w[count-1] ... w[0] */
double ko_warp_reverse(const ko_warper *w, int count, int n, const double *l, double *u);
double ko_map1d_reverse(const ko_map1d *m, double l, double *u);
/* Steffen spline coefficients (GSL interp_steffen restated), coeffs[n-1][4] */
void ko_steffen_coeffs(const double *x, const double *y, int n, double *coeffs);
double ko_steffen_eval(const double *x, const double *coeffs, int n, double xv);

/* ---- integrand (wick_direct.cpp:43-158) -------------------------------- */
typedef struct {
  ko_g0 g0;
  ko_gf_index A, B;
  double cutoff;
} ko_integrand;

int ko_integrand_eval(const ko_integrand *f, int n, const double *times, int keep_u_hypercube,
                      double out[2]);
int ko_integrand_eval_no_det(const ko_integrand *f, int n, const double *times,
                             int keep_u_hypercube, double out[2]);
/* complex LU determinant with partial pivoting (LAPACK zgetf2 pivot rule) */
void ko_det(double *a /* m*m*2, destroyed */, int m, double out[2]);

/* ---- kernel method (wick_kernel.cpp:42-138, det_expansion.hpp:44-74) ---- */
int ko_first_row_expansion(double *a /* m*m*2, destroyed */, int m, double *x /* m*2 */, double det[2]);
/* entries [2^n][n] (re, im): configuration-major, vertex-minor */
int ko_kernel_eval(const ko_g0 *g, const ko_gf_index *X, int n, const double *times, int keep_u_hypercube,
                   double *out);
int ko_kernel_omega_eval(const ko_g0 *g, const ko_gf_index *X, int n, const double *times, int keep_u_hypercube,
                         double omega, double out[2]);

/* ---- integrator (integrator.hpp:68-131) -------------------------------- */
typedef struct {
  double re, im;
  uint64_t n_points, n_in_domain;
} ko_accum;

int ko_eval_warped(const ko_integrand *f, const ko_warper *w, int nw, int n, const double *l,
                   int keep_u_hypercube, double out[2]);
/* integrator::run over Sobol points k = s->k+1 ... ; returns 0 or -1 on exhaustion */
int ko_run(const ko_integrand *f, const ko_warper *w, int nw, ko_sobol *s, uint64_t n_steps,
           ko_accum *acc);
/* same sum, points split over `threads` contiguous Sobol ranges (OpenMP) */
/* integrator::run with OpenMP over each point's Wick rows and configurations
 * (the reference's own structure, wick_direct.cpp:81,93,110). */
int ko_run_intrapoint(const ko_integrand *f, const ko_warper *w, int nw, ko_sobol *s, uint64_t n_steps,
                      int threads, ko_accum *acc);
int ko_run_parallel(const ko_integrand *f, const ko_warper *w, int nw, const ko_sobol *s0,
                    uint64_t n_steps, int threads, ko_accum *acc);

/* compute_gf_kernel's run (compute_obs.hpp:105-117) into binner_t<1,1>:
 * bins [nbins][2] (re, im), counts [nbins][2] */
int ko_kernel_run(const ko_g0 *g, const ko_gf_index *X, const ko_warper *w, int nw, ko_sobol *s, uint64_t n_steps,
                  double xmax, int nbins, double *bins, uint64_t *counts, uint64_t *dropped, uint64_t *n_in_domain);

#ifdef __cplusplus
}
#endif
#endif
