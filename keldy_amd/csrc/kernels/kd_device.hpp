// kd_device.hpp -- device-side building blocks of the Keldysh-sum kernel (gfx950).
//
// Complex arithmetic on double2/float2, the g0 contour lookup, the Sobol point
// generator and the warper chain.  Reference semantics cited per function.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/keldy_amd.h"

namespace kd {

// ---------------------------------------------------------------------------
// complex helpers
template <typename T> struct cvec;
template <> struct cvec<double> { using type = double2; };
template <> struct cvec<float> { using type = float2; };
template <typename T> using cx = typename cvec<T>::type;

template <typename T> __device__ __forceinline__ cx<T> cmk(T re, T im) { cx<T> z; z.x = re; z.y = im; return z; }
template <typename T> __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return cmk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a - l*b
template <typename T> __device__ __forceinline__ cx<T> cfms(cx<T> a, cx<T> l, cx<T> b) {
  return cmk<T>(a.x - l.x * b.x + l.y * b.y, a.y - l.x * b.y - l.y * b.x);
}
// 1/d from the hardware reciprocal estimate refined by Newton steps
// r <- r + r (1 - d r) (two in FP64, one in FP32): within an ulp or two of the
// correctly rounded quotient, without the scaling and fix-up sequence of IEEE
// division.  For the LU pivots here: nonzero, finite, far from the
// denormal range (flagged otherwise).
__device__ __forceinline__ double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ float rcp_nr(float d) {
  float r = __builtin_amdgcn_rcpf(d);
  const float e = fmaf(-d, r, 1.0f);
  return fmaf(r, e, r);
}
template <typename T> __device__ __forceinline__ cx<T> cinv(cx<T> a) {
  T d = a.x * a.x + a.y * a.y;
  T r = rcp_nr(d);
  return cmk<T>(a.x * r, -a.y * r);
}
template <typename T> __device__ __forceinline__ T cabs1(cx<T> a) { return fabs(a.x) + fabs(a.y); }

// ---------------------------------------------------------------------------
// Kernel parameters (by value)
struct KdGf {
  double time;
  int k_idx, spin, orbital;
};

// Guide of a knot search over a non-decreasing array xa[0..n-1]: x falls in
// bucket b = clamp(floor((x - a) * s), 0, G - 1) of a uniform partition of
// [xa[0], xa[n-1]], and entry off + b of the guide buffer holds bounds
// (lo, hi) with lo <= ans(x) < hi for every x of that bucket, where ans is the
// bisection's answer over the whole array.  G = 0: no guide (plain bisection).
struct KnotGuide {
  double a, s;
  int off, G;
};
// A one-dimensional map of a product warper with the guides of its knot
// searches (the inverse CDF's knots F, the forward spline's knots t).
struct MapDev {
  kd_map1d m;
  KnotGuide gF, gt;
  // nearest maps: interval records in the knot buffer -- {F[i], F[i+1], t[i],
  // t[i+1]} at rec + 4 i and {f[j], f[j+1]} at fpair + 2 j (in doubles) -- so
  // that an interval's knots come in one aligned load; F[0] and F[n_inv - 1]
  // (the range of the inverse CDF)
  int64_t rec, fpair;
  double F0, Fl;
};
// A warper of the chain; for a product warper, map_kind is the kind every one
// of its maps has (-1: kinds differ), so that the axes run as one straight
// sequence per kind instead of one branch per axis.
struct WarperDev {
  kd_warper_desc w;
  int map_kind, pad;
};

struct KdParams {
  // g0 table (model.hpp:238-268): [n_orb][n_orb][n_pts] complex
  const double2 *lesser;
  const double2 *greater;
  int n_pts, n_orb;
  double t_min, dt, inv_dt, alpha; // inv_dt = RN(1 / dt)
  // integrand (wick_direct.hpp:34-60)
  KdGf A, B;
  double cutoff;
  int keep_u;
  // warper chain (warpers.hpp:41-137)
  const WarperDev *warpers;
  const MapDev *maps;
  const double *tables;
  const int2 *guides;   // bisection bounds of the knot searches (KnotGuide)
  const double *knots;  // interval records of the nearest maps (MapDev::rec, fpair)
  int n_warpers;       // warpers applied (<= KD_MAX_WARPERS)
  int n_maps;          // maps of the chain (<= KD_MAX_MAPS)
  // Sobol (qrng.hpp:49-114): bit-reversed matrices [dim][31], shifts [dim]
  const uint32_t *Csr;
  const uint32_t *shfr;
  uint32_t k0; // points k0+1 ... k0+n_points
  // per-point I/O
  int mode;
  const double *in_pts; // MODE_EVAL_U: u, MODE_EVAL_L: l
  int64_t n_points;
  double *out_vals; // n x 2
  int *out_dom;
  double *out_jac;
  double *out_l; // MODE_SAMPLE
  // accumulation (MODE_RUN): per block {re, im, n_dom, n_fallback}
  double *partials;
  // optional per-point initial Jacobian (output of the flow stage), or null
  const double *in_jac;
  // time binning by the earliest vertex time (binner_t<1>, binner.hpp:147-352):
  // nbins > 0 in MODE_RUN; bin_part [grid][nbins] per-block sums,
  // bin_counts [nbins + 1] (values added per bin, then the dropped ones)
  int nbins;
  double bin_min, bin_max, bin_size;
  double2 *bin_part;
  unsigned long long *bin_counts;
  // observable: OBS_DIRECT (the Keldysh sum of integrand_g_direct), or the
  // kernel method with external point B (OBS_KERNEL: per-vertex cofactor sums
  // binned by vertex time and Keldysh index, bins [nbins][2]; OBS_KERNEL_OMEGA:
  // the same entries summed with the phase exp(i omega t))
  int obs;
  double omega;
};

enum { OBS_DIRECT = 0, OBS_KERNEL = 1, OBS_KERNEL_OMEGA = 2 };

// Bin of coordinate x (binner.hpp coord2index / in_bounds): -1 when outside
// [xmin, xmax]; x == xmax falls in the last bin.
__device__ __forceinline__ int bin_index(const KdParams &p, double x) {
  if (!(x >= p.bin_min && x <= p.bin_max)) return -1;
  if (x == p.bin_max) return p.nbins - 1;
  const int i = (int)((x - p.bin_min) / p.bin_size);
  return i < p.nbins ? i : p.nbins - 1;
}

// Compacted in-domain points between the two kernel stages (batch b of 64
// points owns slots [64 b, 64 b + cnt[b])).
struct PrepBuf {
  double *u;   // [n][order]
  double *jac; // [n]
  int *idx;    // lane of the point inside its batch
  int *cnt;    // [n / 64] in-domain points per batch
};

enum { MODE_EVAL_U = 0, MODE_EVAL_L = 1, MODE_RUN = 2, MODE_SAMPLE = 3 };

// ---------------------------------------------------------------------------
// g0 lookup: g0_keldysh_contour_t::operator() (model.hpp:244-265) with
// compare_3way (common.cpp:47-57) and linear interpolation on the time mesh.
struct CPt {
  double time;
  int k_idx, split, orb;
};

__device__ __forceinline__ int compare_3way(const CPt &a, const CPt &b) {
  int kd = a.k_idx - b.k_idx;
  if (kd != 0) return kd;
  double dt = a.time - b.time;
  if (dt != 0.0) return (1 - 2 * a.k_idx) * (signbit(dt) ? -1 : 1);
  return (1 - 2 * a.k_idx) * (a.split - b.split);
}

// table_eval in two halves: the position of t on the mesh (a 32-bit index),
// then the interpolation.
struct TabPos {
  int i;     // left mesh point, clamped to [0, n_pts - 2]
  double w;  // weight of the right one
  bool out;  // outside the mesh by more than rounding: NaN (the reference's gf evaluation throws)
};

// The mesh coordinate a = (t - t_min) / dt, correctly rounded (so that mesh
// points land on their index exactly), by Markstein's correction of the
// reciprocal product: q = x (1/dt), r = x - q dt exactly (FMA),
// a = RN(q + r (1/dt)) -- equal to the IEEE quotient when 1/dt is the
// correctly rounded reciprocal (Markstein, IBM J. Res. Dev. 34 (1990) 111),
// three operations instead of the scaled division sequence.  The index in
// 32-bit arithmetic from a clamped copy (no out-of-range conversion).
__device__ __forceinline__ TabPos table_pos(const KdParams &p, double t) {
  const double x = t - p.t_min;
  const double q0 = x * p.inv_dt;
  const double a = fma(fma(-q0, p.dt, x), p.inv_dt, q0);
  TabPos q;
  q.out = !(a >= -1e-9) || !(a <= (double)(p.n_pts - 1) + 1e-9);
  const double ac = fmin(fmax(a, 0.0), (double)(p.n_pts - 2));
  const int i = (int)floor(ac);
  q.i = i;
  q.w = a - (double)i;
  return q;
}

__device__ __forceinline__ double2 table_mix(double2 g0, double2 g1, const TabPos &q) {
  double2 r;
  r.x = (1.0 - q.w) * g0.x + q.w * g1.x;
  r.y = (1.0 - q.w) * g0.y + q.w * g1.y;
  if (q.out) r.x = r.y = __builtin_nan("");
  return r;
}

__device__ __forceinline__ double2 table_eval(const KdParams &p, const double2 *tab, int oa, int ob,
                                              double t) {
  const double2 *g = tab + (size_t)(oa * p.n_orb + ob) * p.n_pts;
  const TabPos q = table_pos(p, t);
  return table_mix(g[q.i], g[q.i + 1], q);
}

__device__ __forceinline__ double2 g0_eval(const KdParams &p, const CPt &a, const CPt &b, bool internal) {
  int ord = compare_3way(a, b);
  if (ord > 0) return table_eval(p, p.greater, a.orb, b.orb, a.time - b.time);
  if (ord < 0) return table_eval(p, p.lesser, a.orb, b.orb, a.time - b.time);
  double2 v = table_eval(p, p.lesser, a.orb, b.orb, 0.0);
  if (internal && a.orb == b.orb) v.y -= p.alpha;
  return v;
}

// One item of a point's Wick matrix build (wick_direct.cpp:92-101 picks), stored
// interleaved by vertex: row/column 2i+s for vertex i with Keldysh index s, 2N
// for the external points, row stride S.  Items q < N^2 are the vertex pairs
// (i, j) = (q / N, q mod N): their four entries share the time difference
// d = t_i - t_j, and compare_3way makes them g0^>(d) or g0^<(d) only --
// (0,1): lesser; (1,0): greater; (0,0): greater when (i,0) is later on the
// contour (d > 0, or d = 0 and i > j), else lesser; (1,1): the reverse; equal
// points (i = j) give lesser(0) - i alpha -- so a pair needs two table
// lookups instead of four.  Items N^2 .. N^2+4N are the external row (q - N^2
// < 2N), the external column, and the corner (0; at order 0 g0(A, B)).
// wick_store and wick_item return the largest |Re|+|Im| of the vertex-block
// entries written (0 for the others).
//
// The build item in two halves: wick_fetch loads the table entries the item
// needs from its two times (ta = the row side's, A's for the external row; tb =
// the column side's, B's for the external column), wick_store interpolates and
// writes the entries; wick_item does both.
struct WickPre {
  double2 l0, l1, g0, g1; // pair: lesser and greater entries around the mesh position; external: the table's
  TabPos pos;
  int o; // pair: the (0,0) order (compare_3way) and equal points; external: the alpha correction applies
};

template <int N> __device__ __forceinline__ void wick_item_vertices(int q, int &ia, int &ib) {
  if (q < N * N) {
    ia = q / N;
    ib = q - ia * N;
    return;
  }
  const int e = q - N * N;
  ia = e < 2 * N ? -1 : e < 4 * N ? (e - 2 * N) >> 1 : -1;
  ib = e < 2 * N ? e >> 1 : -1;
}

// Where item q's table entries are: w gets the mesh position and the order
// flag; the entries are a[0], a[1] (pair items: lesser; external items: the
// table compare_3way picks) and, for pair items, b[0], b[1] (greater).  The
// corner item (no entries) returns a = b = nullptr.
struct WickLoc {
  const double2 *a, *b;
};

template <int N>
__device__ __forceinline__ WickLoc wick_locate(const KdParams &p, int q, double ta, double tb, WickPre &w) {
  if (q < N * N) {
    const int i = q / N, j = q - i * N;
    const double d = ta - tb;
    w.pos = table_pos(p, d);
    w.o = d > 0.0 ? 1 : d < 0.0 ? -1 : (i > j) - (i < j);
    return WickLoc{p.lesser + w.pos.i, p.greater + w.pos.i};
  }
  const int e = q - N * N;
  if (e >= 4 * N) return WickLoc{nullptr, nullptr};
  CPt a, b;
  if (e < 2 * N) {
    a = CPt{ta, p.A.k_idx, N, p.A.orbital};
    b = CPt{tb, e & 1, e >> 1, 0};
  } else {
    const int r = e - 2 * N;
    a = CPt{ta, r & 1, r >> 1, 0};
    b = CPt{tb, p.B.k_idx, N, p.B.orbital};
  }
  // g0_eval(p, a, b, internal = true)
  const int ord = compare_3way(a, b);
  const double2 *g = (ord > 0 ? p.greater : p.lesser) + (size_t)(a.orb * p.n_orb + b.orb) * p.n_pts;
  w.pos = table_pos(p, ord == 0 ? 0.0 : a.time - b.time);
  w.o = ord == 0 && a.orb == b.orb;
  return WickLoc{g + w.pos.i, nullptr};
}

template <int N>
__device__ __forceinline__ void wick_fetch(const KdParams &p, int q, double ta, double tb, WickPre &w) {
  const WickLoc L = wick_locate<N>(p, q, ta, tb, w);
  if (q < N * N) {
    w.l0 = L.a[0];
    w.l1 = L.a[1];
    w.g0 = L.b[0];
    w.g1 = L.b[1];
  } else if (q - N * N < 4 * N) {
    w.l0 = L.a[0];
    w.l1 = L.a[1];
  }
}

template <int N>
__device__ __forceinline__ double wick_store(const KdParams &p, double2 *__restrict__ W, int S, int q,
                                             const WickPre &w) {
  if (q < N * N) {
    const int i = q / N, j = q - i * N;
    const double2 gl = table_mix(w.l0, w.l1, w.pos);
    const double2 gg = table_mix(w.g0, w.g1, w.pos);
    double2 g00 = w.o > 0 ? gg : gl;
    double2 g11 = w.o < 0 ? gg : gl;
    if (w.o == 0) {
      g00.y -= p.alpha;
      g11.y -= p.alpha;
    }
    double2 *d = W + 2 * i * S + 2 * j;
    d[0] = g00;
    d[1] = gl;
    d[S] = gg;
    d[S + 1] = g11;
    return fmax(fmax(fabs(g00.x) + fabs(g00.y), fabs(gl.x) + fabs(gl.y)),
                fmax(fabs(gg.x) + fabs(gg.y), fabs(g11.x) + fabs(g11.y)));
  }
  const int e = q - N * N;
  if (e < 4 * N) {
    double2 v = table_mix(w.l0, w.l1, w.pos);
    if (w.o) v.y -= p.alpha;
    if (e < 2 * N)
      W[2 * N * S + e] = v;
    else
      W[(e - 2 * N) * S + 2 * N] = v;
  } else if constexpr (N > 0) {
    W[2 * N * S + 2 * N] = make_double2(0.0, 0.0);
  } else { // order 0: the value is g0(A, B) itself
    W[0] = g0_eval(p, CPt{p.A.time, p.A.k_idx, 0, p.A.orbital}, CPt{p.B.time, p.B.k_idx, 0, p.B.orbital}, false);
  }
  return 0.0;
}

template <int N>
__device__ __forceinline__ double wick_item(const KdParams &p, const double *__restrict__ Us, double2 *__restrict__ W,
                                            int S, int q) {
  int ia, ib;
  wick_item_vertices<N>(q, ia, ib);
  WickPre w;
  wick_fetch<N>(p, q, ia >= 0 ? Us[ia] : p.A.time, ib >= 0 ? Us[ib] : p.B.time, w);
  return wick_store<N>(p, W, S, q, w);
}
// items of the build above
template <int N> constexpr int wick_items() { return N * N + 4 * N + 1; }

// ---------------------------------------------------------------------------
// Sobol point k (digitalseq_b2g.hpp:415-426 / set_state 455-471): the state
// after k steps is the XOR of the columns selected by the Gray code of k.
// k = 2^31 follows next()'s column rule (ctz(k) % 31 == 0).
__device__ __forceinline__ double sobol_coord(const uint32_t *__restrict__ Csr, const uint32_t *__restrict__ shfr,
                                              int j, uint32_t k) {
  uint32_t g = k ^ (k >> 1);
  uint32_t cur = 0;
  const uint32_t *col = Csr + j * 31;
  for (int i = 0; i < 31; i++) cur ^= col[i] & (0u - ((g >> i) & 1u));
  if (k == 0x80000000u) cur = col[30] ^ col[0];
  return (double)(cur ^ shfr[j]) * (1.0 / 4294967296.0);
}

// ---------------------------------------------------------------------------
// GSL-compatible interval search (gsl_interp_bsearch over [0, n-1]), between
// bounds lo < hi that hold the answer: the largest i with xa[i] <= x, clamped to
// [0, n-2] (n-2 for NaN x, which no comparison stops).
__device__ __forceinline__ int bisect_knots(const double *xa, int2 r, double x) {
  int ilo = r.x, ihi = r.y;
  while (ihi > ilo + 1) {
    int i = (ihi + ilo) >> 1;
    if (xa[i] > x)
      ihi = i;
    else
      ilo = i;
  }
  return ilo;
}
__device__ __forceinline__ int bsearch_knots(const double *xa, int n, double x) {
  return bisect_knots(xa, make_int2(0, n - 1), x);
}

// Bounds of the search of x over xa[0..n-1] from the array's guide: usually
// hi = lo + 1, so that no bisection step remains and a search costs one guide
// load instead of log2(n) dependent knot loads.  Without a guide (G = 0, off =
// 0: the buffer's first entry is a placeholder), the whole array; for NaN x,
// the interval the full bisection ends in.  Branch-free, so that the loads of
// several searches issue together.
__device__ __forceinline__ int2 guide_bounds(int n, const int2 *__restrict__ gd, const KnotGuide &g, double x) {
  const double tb = (x - g.a) * g.s;
  int b = tb >= 0.0 ? (tb < (double)g.G ? (int)tb : g.G - 1) : 0; // NaN: 0
  b = g.G == 0 ? 0 : b;
  int2 r = gd[g.off + b];
  if (g.G == 0) r = make_int2(0, n - 1);
  if (x != x) r = make_int2(n - 2, n - 1);
  return r;
}

// v [n] on the uniform mesh of [0, 1], at t: linear interpolation between the
// neighbouring mesh points (index clamped to the first and last intervals, so
// any t, NaN included, reads inside v)
__device__ __forceinline__ double mesh_lin(const double *__restrict__ v, int n, double t) {
  const double delta = 1.0 / (double)(n - 1);
  const double a = t / delta;
  int i = (int)floor(a);
  i = i > n - 2 ? n - 2 : i;
  i = i < 0 ? 0 : i;
  const double w = a - (double)i;
  return (1.0 - w) * v[i] + w * v[i + 1];
}

// x = F^{-1}(l) of one map kind; returns 1/f(x)   (product_1d_simple.hpp:109-117 and
// makers 560-630, interp nearest 198-217, interp hybrid 403-415)
template <int K>
__device__ __forceinline__ double map_reverse_k(const MapDev &md, const double *__restrict__ tab,
                                                const int2 *__restrict__ gd, double l, double *u) {
  const kd_map1d &m = md.m;
  double x, f;
  const double w = m.w, um = m.u_max;
  if constexpr (K == KD_MAP_UNIFORM) {
    x = um * l;
    f = 1.0 / um;
  } else if constexpr (K == KD_MAP_EXPONENTIAL) {
    x = um - w * log1p(expm1(um / w) * l);
    f = exp(-x / w) / (w * m.norm);
  } else if constexpr (K == KD_MAP_INVERSE) {
    x = w * expm1(l * m.norm);
    f = 1.0 / (m.norm * (w + x));
  } else if constexpr (K == KD_MAP_INVERSE_SQ) {
    x = m.norm * l * w / (1.0 - m.norm * l);
    f = w / ((w + x) * (w + x) * m.norm);
  } else if constexpr (K == KD_MAP_INVERSE_POW) {
    const int pw = m.power;
    double base = pow(w / (um + w), (double)(pw - 1));
    double denom = pow(1.0 + l * (base - 1.0), 1.0 / (pw - 1.0));
    x = w * (1.0 - denom) / denom;
    f = m.norm / pow(w + x, (double)pw);
  } else if constexpr (K == KD_MAP_NEAREST) {
    const double *F = tab + m.F_off, *t = tab + m.t_off, *fk = tab + m.f_off;
    double y = -l;
    int i = bisect_knots(F, guide_bounds(m.n_inv, gd, md.gF, y), y);
    double dx = F[i + 1] - F[i];
    x = dx > 0.0 ? t[i] + (y - F[i]) / dx * (t[i + 1] - t[i]) : 0.0;
    if (y < F[0] || y > F[m.n_inv - 1]) x = __builtin_nan("");
    int j = bisect_knots(t, guide_bounds(m.n_pts, gd, md.gt, x), x);
    f = 0.5 * (fk[j] + fk[j + 1]);
  } else if constexpr (K == KD_MAP_HYBRID) {
    const double *F = tab + m.F_off, *c = tab + m.c_off;
    double y = -l;
    int i = bisect_knots(F, guide_bounds(m.n_inv, gd, md.gF, y), y);
    double d = y - F[i];
    x = c[4 * i + 3] + d * (c[4 * i + 2] + d * (c[4 * i + 1] + d * c[4 * i + 0]));
    if (y < F[0] || y > F[m.n_inv - 1]) x = __builtin_nan("");
    double num = w, den = w + x, dd = den;
    for (int q = 1; q < m.power; q++) num *= w;
    for (int q = 1; q < m.power; q++) dd *= den;
    f = (num / dd) / m.norm;
  } else {
    static_assert(K == KD_MAP_PROJECTION, "unknown map kind");
    // tables on the uniform meshes of [0, 1], evaluated by linear interpolation
    // (the TRIQS gf<retime> evaluation of projection.hpp:331-357)
    x = mesh_lin(tab + m.t_off, m.n_inv, l);
    f = mesh_lin(tab + m.f_off, m.n_pts, x) / m.norm;
    if (!(l >= 0.0 && l <= 1.0)) x = __builtin_nan("");
  }
  *u = x;
  return 1.0 / f;
}

__device__ inline double map1d_reverse(const MapDev &md, const double *__restrict__ tab, const int2 *__restrict__ gd,
                                       double l, double *u) {
  switch (md.m.kind) {
  case KD_MAP_UNIFORM: return map_reverse_k<KD_MAP_UNIFORM>(md, tab, gd, l, u);
  case KD_MAP_EXPONENTIAL: return map_reverse_k<KD_MAP_EXPONENTIAL>(md, tab, gd, l, u);
  case KD_MAP_INVERSE: return map_reverse_k<KD_MAP_INVERSE>(md, tab, gd, l, u);
  case KD_MAP_INVERSE_SQ: return map_reverse_k<KD_MAP_INVERSE_SQ>(md, tab, gd, l, u);
  case KD_MAP_INVERSE_POW: return map_reverse_k<KD_MAP_INVERSE_POW>(md, tab, gd, l, u);
  case KD_MAP_NEAREST: return map_reverse_k<KD_MAP_NEAREST>(md, tab, gd, l, u);
  case KD_MAP_HYBRID: return map_reverse_k<KD_MAP_HYBRID>(md, tab, gd, l, u);
  case KD_MAP_PROJECTION: return map_reverse_k<KD_MAP_PROJECTION>(md, tab, gd, l, u);
  default:
    *u = __builtin_nan("");
    return __builtin_nan("");
  }
}

// N axes of a product of nearest-interpolation maps (map_reverse_k<KD_MAP_NEAREST>
// on every axis, the same results): each step runs over all axes before the
// next starts, so the axes' dependent loads are in flight together.  The f pair
// of interval i of the inverse-CDF search is loaded with that interval's record:
// when the mapped time x lies in [t[i], t[i+1]), the search of x over the time
// knots ends in i (the one interval of a sorted array that holds x), and for a
// point outside [F[0], F[n_inv-1]] (x NaN) it ends in n_pts - 2.  The guided
// time search runs only on the lanes where neither holds (x rounded onto
// t[i+1], a NaN l), so a point costs two memory round trips (guide, record and
// f pair) instead of four.  inv_f[a] = 1/f of axis a.
template <int N>
__device__ __forceinline__ void nearest_axes(const MapDev *__restrict__ ms, bool one, const double *__restrict__ tab,
                                             const int2 *__restrict__ gd, const double *__restrict__ kn,
                                             double (&x)[N], double (&inv_f)[N]) {
  double y[N];
  int2 r[N];
  int k[N];
#pragma unroll
  for (int a = 0; a < N; a++) {
    const MapDev &md = ms[one ? 0 : a];
    y[a] = -x[a];
    r[a] = guide_bounds(md.m.n_inv, gd, md.gF, y[a]);
  }
#pragma unroll
  for (int a = 0; a < N; a++) k[a] = bisect_knots(tab + ms[one ? 0 : a].m.F_off, r[a], y[a]);
  // the axes' interval records and f pairs, in groups of four axes whose loads
  // are all issued before the first use (registers for the records of one group)
  constexpr int GA = 4;
  unsigned slow = 0; // bit a: axis a needs the time search
#pragma unroll
  for (int a0 = 0; a0 < N; a0 += GA) {
    double4 rec[GA];
    double2 fp[GA];
    bool out[GA];
#pragma unroll
    for (int a = a0; a < a0 + GA && a < N; a++) {
      const MapDev &md = ms[one ? 0 : a];
      out[a - a0] = (y[a] < md.F0) | (y[a] > md.Fl);
      const int kf = out[a - a0] ? md.m.n_pts - 2 : k[a];
      rec[a - a0] = *reinterpret_cast<const double4 *>(kn + md.rec + 4 * (int64_t)k[a]); // F[i], F[i+1], t[i], t[i+1]
      fp[a - a0] = *reinterpret_cast<const double2 *>(kn + md.fpair + 2 * (int64_t)kf);  // f[i], f[i+1]
    }
#pragma unroll
    for (int a = a0; a < a0 + GA && a < N; a++) {
      const double4 r = rec[a - a0];
      const double dx = r.y - r.x;
      // on every lane (inf or NaN, and unused, unless dx > 0): kept out of a
      // branch, which would hold the loads of the axes after it back
      double xa = r.z + (y[a] - r.x) / dx * (r.w - r.z);
      __asm__ volatile("" : "+v"(xa));
      xa = dx > 0.0 ? xa : 0.0;
      x[a] = out[a - a0] ? __builtin_nan("") : xa;
      slow |= (unsigned)!(out[a - a0] || (r.z <= xa && xa < r.w)) << a;
      inv_f[a] = 1.0 / (0.5 * (fp[a - a0].x + fp[a - a0].y));
    }
  }
  if (slow) {
    // the guided search of x over the time knots, one axis after the other
#pragma unroll
    for (int a = 0; a < N; a++) {
      if (slow & (1u << a)) {
        const MapDev &md = ms[one ? 0 : a];
        const int j = bisect_knots(tab + md.m.t_off, guide_bounds(md.m.n_pts, gd, md.gt, x[a]), x[a]);
        const double2 f = *reinterpret_cast<const double2 *>(kn + md.fpair + 2 * (int64_t)j);
        inv_f[a] = 1.0 / (0.5 * (f.x + f.y));
      }
    }
  }
}

} // namespace kd
