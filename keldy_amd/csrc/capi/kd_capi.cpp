// kd_capi.cpp -- implementation of the C-ABI in include/keldy_amd.h.
//
// Owns device memory, one HIP stream (default priority) and the launch
// geometry.  All validation that the reference performs when constructing or
// running an integrator is repeated here so that a C caller gets the same
// errors as the C++/Python host API.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/keldy_amd.h"
#include "../kernels/kd_kernels.hpp"

struct kd_ctx {
  int device = 0;
  int n_cu = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> ev; // per-chunk timing events: prep start/end, LU start/end
  std::string err;
  // g0 table
  double2 *d_lesser = nullptr, *d_greater = nullptr;
  int n_pts = 0, n_orb = 0;
  double time_max = 0.0, alpha = 0.0;
  bool has_g0 = false;
  // integrand
  int order = -1;
  kd_gf_index A{}, B{};
  double cutoff = 0.0;
  int precision = KD_PREC_FP64;
  int obs = KD_OBS_DIRECT; // observable (kd_set_integrand / kd_set_kernel_integrand)
  double omega = 0.0;
  // warper chain
  kd::WarperDev *d_warpers = nullptr;
  kd::MapDev *d_maps = nullptr;
  double *d_tables = nullptr;
  int2 *d_guides = nullptr;   // knot-search guides of the maps
  double *d_knots = nullptr;  // interval records of the nearest maps
  int n_warpers = 0, n_maps = 0;
  std::vector<kd_warper_desc> h_warpers;
  // learned flow warper (KD_WARP_FLOW): FP32 parameters, Adam moments,
  // gradient, bf16 operand copies
  int flow_order = -1, flow_layers = 0, flow_step = 0;
  int flow_nc = 0, flow_alloc_nc = 0; // conditioning inputs (requested; of the allocated flow)
  double flow_cond = 0.0;             // the map's conditioning value
  float *d_fw = nullptr, *d_fm = nullptr, *d_fv = nullptr, *d_fgrad = nullptr;
  uint16_t *d_fpk = nullptr;
  bool flow_in_chain = false, flow_suppressed = false;
  double *d_fx = nullptr; // flow-stage output per chunk: x [kChunk][order], jac [kChunk]
  size_t fx_cap = 0;
  // flow training set and scratch
  double *d_tx = nullptr, *d_tw = nullptr, *d_tc = nullptr; // points, weights, conditioning values
  int64_t n_train = 0;
  double w_mean = 0.0;
  float *d_fpart = nullptr;
  size_t fpart_cap = 0;
  double *d_floss = nullptr; // [loss_part blocks | per-step losses]
  size_t floss_cap = 0;
  // Sobol: one device buffer for the largest dimension (d_shfr points into it),
  // filled from the pinned staging buffer h_sobol by copies on the ctx stream
  uint32_t *d_Csr = nullptr, *d_shfr = nullptr;
  uint32_t *h_sobol = nullptr;
  hipEvent_t ev_sobol = nullptr; // recorded after the last staging copy
  bool sobol_pending = false;
  int sobol_dim = 0;
  uint32_t k = 0;
  bool has_sobol = false;
  // accumulators
  double *d_acc = nullptr; // re, im, n_dom, n_fallback
  double *h_acc = nullptr; // pinned copy of d_acc
  double *d_part = nullptr;
  int part_cap = 0;
  uint64_t n_points = 0;
  // staging for host-memory batch calls
  void *d_stage = nullptr;
  size_t stage_cap = 0;
  // compacted points between the two kernel stages
  void *d_prep = nullptr;
  size_t prep_cap = 0, prep_pts = 0;
  // runs of several chunks: a second record buffer and stream, so that the
  // next chunk's points are prepared while the Keldysh sum of this one runs
  void *d_prep2 = nullptr;
  size_t prep2_pts = 0;
  int prep2_order = -1;
  hipStream_t stream2 = nullptr;
  // [2 c]: chunk c prepared (on stream2; chunk 0 on the ctx stream), [2 c + 1]:
  // its records read by the Keldysh sum (ctx stream)
  std::vector<hipEvent_t> ev_sync;
  int prep_order = -1;
  float last_ms = 0.f, last_lu_ms = 0.f, last_prep_ms = 0.f;
  int last_launches = 0, last_chunks = 0;
  int timing = KD_TIMING_KERNELS; // what kd_run / kd_eval_* / kd_sample_device time (kd_set_timing)
  bool last_kernels_timed = false;
  // launch shape of the Keldysh-sum kernel (kd_autotune): maximum lanes per
  // point and waves per CU of the persistent grid (0: the algorithm's default)
  int lu_lanes_max = 64, lu_waves_per_cu = 0;
  int algo = KD_ALGO_DEFAULT; // Keldysh-sum algorithm (KD_ALGO_DEFAULT: time-ordered tree at orders 6..10)
  float tune_ms_per_mpt = 0.f;
  bool no_reduce = false; // autotuning launches leave the accumulator alone
  // time binning (kd_set_time_bins)
  int nbins = 0;
  int bin_k = 1; // bins per time bin: 1 (time binning), 2 (kernel method: Keldysh index)
  double bin_min = 0.0, bin_max = 0.0;
  double2 *d_bins = nullptr, *d_bin_part = nullptr;
  unsigned long long *d_bin_counts = nullptr;
  int bin_part_cap = 0;
};

// Scratch of kd_eval_batch_device_noalloc: the compacted-point buffer
// (u [n][order], jac [n], idx [n], cnt [n / 64]) for up to
// KD_NOALLOC_MAX_POINTS points at order KD_MAX_ORDER, held by every device the
// library runs on until the process ends.  Host code never uses the symbol's
// name as an address: kd_create resolves the device's own copy once per device
// (hipGetSymbolAddress) and the entry point passes that address to the kernels.
constexpr size_t kNoallocScratchBytes =
    (size_t)KD_NOALLOC_MAX_POINTS * (sizeof(double) * KD_MAX_ORDER + sizeof(double) + sizeof(int)) +
    (size_t)KD_NOALLOC_MAX_POINTS / 64 * sizeof(int) + 256;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};       // device address of `scratch`, per device
hipError_t g_scratch_rc[kMaxDevices] = {};   // result of its lookup (hipSuccess before the first)
bool g_scratch_done[kMaxDevices] = {};

// Resolves `scratch` on `device` (the current device) once per process.
void resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_done[device]) return;
  void *ptr = nullptr;
  g_scratch_rc[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_ptr[device] = g_scratch_rc[device] == hipSuccess ? ptr : nullptr;
  g_scratch_done[device] = true;
}

int fail(kd_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define KD_HIP(c, expr)                                                                            \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return fail((c), KD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename P> void dfree(P *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

int ensure_stage(kd_ctx *c, size_t bytes) {
  if (bytes <= c->stage_cap) return KD_OK;
  dfree(c->d_stage);
  c->stage_cap = 0;
  KD_HIP(c, hipMalloc(&c->d_stage, bytes));
  c->stage_cap = bytes;
  return KD_OK;
}

bool use_tree(const kd_ctx *c) { return c->algo == KD_ALGO_TREE && c->order >= 6 && c->obs == KD_OBS_DIRECT; }
// the paired per-configuration LU (orders >= 7 at 64 lanes per point)
bool use_pair(const kd_ctx *c) {
  return c->algo == KD_ALGO_PAIR && c->order >= 7 && c->obs == KD_OBS_DIRECT && c->lu_lanes_max == 64;
}

// orders where the time-ordered tree is the default algorithm (all it supports:
// the fastest there on MI355X)
bool ctree_default(int order) { return kd::ctree_supported(order); }

// the split tree + per-configuration LU with one point per wave (orders 6..10)
bool use_wsplit(const kd_ctx *c) {
  const bool chosen =
      c->algo == KD_ALGO_WSPLIT || (c->algo == KD_ALGO_DEFAULT && !ctree_default(c->order));
  return chosen && c->obs == KD_OBS_DIRECT && kd::wsplit_supported(c->order);
}

// the time-ordered tree of vertex-level corrections, one point per wave (orders
// 6..10; the default there)
bool use_ctree(const kd_ctx *c) {
  const bool chosen = c->algo == KD_ALGO_CTREE || (c->algo == KD_ALGO_DEFAULT && ctree_default(c->order));
  return chosen && c->obs == KD_OBS_DIRECT && kd::ctree_supported(c->order);
}

// the split tree + per-configuration LU with one point per workgroup (orders 6..10)
bool use_split(const kd_ctx *c) {
  return c->algo == KD_ALGO_SPLIT && c->obs == KD_OBS_DIRECT && kd::split_threads(c->order) > 0;
}

// complex values per point of an evaluation output
int64_t vals_per_point(const kd_ctx *c) {
  return c->obs == KD_OBS_KERNEL ? (int64_t(1) << c->order) * c->order : 1;
}

// Waves per CU of the persistent Keldysh-sum grid: the tuned or set value, else
// the algorithm's default.  The one-point-per-wave kernels give each wave a
// fixed share of the points, so a grid of many times the waves that fit at
// once lets the dispatcher balance the shares across the CUs (order 8, the
// time-ordered tree, 2^21 points: 4.07 ms at 128 waves per CU, 4.18 at 64,
// 4.36 at 32; the wave-per-point split: 9.7 ms at 24-32, 9.9 ms at 16).
int grid_waves_per_cu(const kd_ctx *c) {
  if (c->lu_waves_per_cu > 0) return c->lu_waves_per_cu;
  if (use_ctree(c)) return 128;
  if (use_wsplit(c)) return 32;
  return 16;
}

// LU-kernel grid: one workgroup per group of points (64 / lanes points per
// group), capped at lu_waves_per_cu waves per CU -- the kernel is persistent
// over the rest.
int grid_for(const kd_ctx *c, int64_t n_points) {
  const bool per_point = use_tree(c) || use_split(c) || use_wsplit(c) || use_ctree(c); // one workgroup item per point
  const int64_t per_batch = per_point ? 64 : kd::lu_lanes(c->order > 0 ? c->order : 0, c->lu_lanes_max);
  int64_t nb = (n_points + 63) / 64 * per_batch;
  // waves per CU -> workgroups per CU (the split kernel's workgroups hold several waves)
  const int wg_waves = use_split(c) ? std::max(1, kd::split_threads(c->order) / 64) : 1;
  int64_t cap = (int64_t)(c->n_cu > 0 ? c->n_cu : 256) * std::max(1, grid_waves_per_cu(c) / wg_waves);
  // the time-ordered tree below the cap: at least three candidate points per
  // wave (order 8, 54,875 points: 121 us at 3.3 per wave, 124 at 1.7), six with
  // two points per wave, so that each wave has a few pairs (that kernel at
  // order 8: 146 us at 8 points per wave, 180 at 3)
  if (use_ctree(c)) nb = std::min(nb, kd::ctree_pairs(c->order) ? (n_points + 5) / 6 : (n_points + 2) / 3);
  if (nb > cap) nb = cap;
  return (int)(nb < 1 ? 1 : nb);
}

int ensure_bin_partials(kd_ctx *c, int grid) {
  if (c->nbins <= 0 || grid <= c->bin_part_cap) return KD_OK;
  dfree(c->d_bin_part);
  c->bin_part_cap = 0;
  KD_HIP(c, hipMalloc(&c->d_bin_part, sizeof(double2) * (size_t)c->nbins * c->bin_k * (size_t)grid));
  c->bin_part_cap = grid;
  return KD_OK;
}

// per-workgroup partial sums (4 per workgroup), then the reduction's block
// sums and its ticket, which starts at zero
int ensure_partials(kd_ctx *c, int grid) {
  if (grid <= c->part_cap) return KD_OK;
  dfree(c->d_part);
  c->part_cap = 0;
  const size_t n = kd::partials_doubles(grid);
  KD_HIP(c, hipMalloc(&c->d_part, sizeof(double) * n));
  KD_HIP(c, hipMemset(c->d_part + n - 1, 0, sizeof(double)));
  c->part_cap = grid;
  return KD_OK;
}

kd::KdParams base_params(const kd_ctx *c) {
  kd::KdParams p{};
  p.lesser = c->d_lesser;
  p.greater = c->d_greater;
  p.n_pts = c->n_pts;
  p.n_orb = c->n_orb;
  p.t_min = -c->time_max;
  p.dt = 2.0 * c->time_max / (double)(c->n_pts - 1);
  p.inv_dt = 1.0 / p.dt;
  p.alpha = c->alpha;
  p.A = kd::KdGf{c->A.time, c->A.k_idx, c->A.spin, c->A.orbital};
  p.B = kd::KdGf{c->B.time, c->B.k_idx, c->B.spin, c->B.orbital};
  p.cutoff = c->cutoff;
  p.keep_u = 1;
  p.warpers = c->d_warpers;
  p.maps = c->d_maps;
  p.tables = c->d_tables;
  p.guides = c->d_guides;
  p.knots = c->d_knots;
  p.n_warpers = c->n_warpers;
  p.n_maps = c->n_maps;
  p.Csr = c->d_Csr;
  p.shfr = c->d_shfr;
  p.k0 = c->k;
  p.nbins = c->nbins;
  p.bin_min = c->bin_min;
  p.bin_max = c->bin_max;
  p.bin_size = c->nbins > 0 ? (c->bin_max - c->bin_min) / c->nbins : 0.0;
  p.bin_counts = c->d_bin_counts;
  p.obs = c->obs;
  p.omega = c->omega;
  return p;
}

int need_integrand(kd_ctx *c) {
  if (!c->has_g0) return fail(c, KD_ERR_STATE, "g0 table not set (kd_set_g0_table)");
  if (c->order < 0) return fail(c, KD_ERR_STATE, "integrand not set (kd_set_integrand)");
  return KD_OK;
}

// Points per launch pair (prep + integrand); bounds the compacted-point buffer.
constexpr int64_t kChunk = int64_t(1) << 21;

// Compacted-point buffer for chunks of up to n_points points (at most kChunk):
// u [n][order], jac [n], idx [n], cnt [n / 64].  Grows on demand only, so many
// small contexts (one per external time, say) stay small.
int ensure_prep(kd_ctx *c, int order, int64_t n_points, kd::PrepBuf *buf) {
  const size_t n = (size_t)((std::min<int64_t>(std::max<int64_t>(n_points, 64), kChunk) + 63) / 64 * 64);
  const size_t per_pt = sizeof(double) * (size_t)(order > 0 ? order : 1) + sizeof(double) + sizeof(int);
  if (n > c->prep_pts || c->prep_order != order) {
    const size_t m = std::max(n, c->prep_order == order ? c->prep_pts : (size_t)0);
    const size_t bytes = m * per_pt + (m / 64) * sizeof(int) + 256;
    dfree(c->d_prep);
    c->prep_cap = 0;
    KD_HIP(c, hipMalloc(&c->d_prep, bytes));
    c->prep_cap = bytes;
    c->prep_pts = m;
    c->prep_order = order;
  }
  const size_t np = c->prep_pts;
  char *base = (char *)c->d_prep;
  buf->u = (double *)base;
  base += np * sizeof(double) * (size_t)(order > 0 ? order : 1);
  buf->jac = (double *)base;
  base += np * sizeof(double);
  buf->idx = (int *)base;
  base += np * sizeof(int);
  buf->cnt = (int *)base;
  return KD_OK;
}

// The second record buffer, stream and synchronisation events of a run of
// n_points > kChunk (nothing otherwise).
int ensure_pipeline(kd_ctx *c, int order, int64_t n_points, kd::PrepBuf *buf2) {
  if (n_points <= kChunk) return KD_OK;
  const size_t n = (size_t)kChunk;
  const size_t per_pt = sizeof(double) * (size_t)(order > 0 ? order : 1) + sizeof(double) + sizeof(int);
  if (n > c->prep2_pts || c->prep2_order != order) {
    dfree(c->d_prep2);
    c->prep2_pts = 0;
    KD_HIP(c, hipMalloc(&c->d_prep2, n * per_pt + (n / 64) * sizeof(int) + 256));
    c->prep2_pts = n;
    c->prep2_order = order;
  }
  if (!c->stream2) KD_HIP(c, hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
  const size_t nev = 2 * (size_t)((n_points + kChunk - 1) / kChunk);
  while (c->ev_sync.size() < nev) {
    hipEvent_t e;
    KD_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    c->ev_sync.push_back(e);
  }
  char *base = (char *)c->d_prep2;
  buf2->u = (double *)base;
  base += n * sizeof(double) * (size_t)(order > 0 ? order : 1);
  buf2->jac = (double *)base;
  base += n * sizeof(double);
  buf2->idx = (int *)base;
  base += n * sizeof(int);
  buf2->cnt = (int *)base;
  return KD_OK;
}

int ensure_events(kd_ctx *c, size_t n) {
  while (c->ev.size() < n) {
    hipEvent_t e;
    KD_HIP(c, hipEventCreate(&e));
    c->ev.push_back(e);
  }
  return KD_OK;
}

int ensure_flow_stage(kd_ctx *c, int order) {
  const size_t bytes = sizeof(double) * (size_t)kChunk * (size_t)(order + 1);
  if (bytes <= c->fx_cap) return KD_OK;
  dfree(c->d_fx);
  c->fx_cap = 0;
  KD_HIP(c, hipMalloc(&c->d_fx, bytes));
  c->fx_cap = bytes;
  return KD_OK;
}

kd::FlowParams flow_params(const kd_ctx *c) {
  return kd::FlowParams{c->d_fw, c->d_fpk, c->flow_layers, kd::flow_s_scale(c->flow_layers), c->flow_alloc_nc,
                        (float)c->flow_cond};
}

int flow_grid(const kd_ctx *c, int64_t n) {
  const int64_t nb = (n + 63) / 64, cap = (int64_t)(c->n_cu > 0 ? c->n_cu : 256) * 16;
  return (int)std::max<int64_t>(1, std::min(nb, cap));
}

// timing: KD_TIMING_NONE, _TOTAL (events around the launches) or _KERNELS
// (also around each chunk's prep and Keldysh-sum kernels)
int launch_chunks(kd_ctx *c, const kd::KdParams &p0, int timing, const kd::PrepBuf *ext = nullptr);
int launch_timed(kd_ctx *c, const kd::KdParams &p0, int /*unused*/) { return launch_chunks(c, p0, c->timing); }

// Allocate every buffer a launch of p0 needs (before a graph capture).
int prepare_launch(kd_ctx *c, const kd::KdParams &p0) {
  kd::PrepBuf buf, buf2;
  int rc = ensure_prep(c, c->order, p0.n_points, &buf);
  if (rc) return rc;
  if ((rc = ensure_pipeline(c, c->order, p0.n_points, &buf2))) return rc;
  if (c->flow_in_chain && p0.mode != kd::MODE_EVAL_U && c->order > 0 && (rc = ensure_flow_stage(c, c->order)))
    return rc;
  const int grid = grid_for(c, std::min<int64_t>(kChunk, p0.n_points));
  if ((rc = ensure_partials(c, grid))) return rc;
  return p0.mode == kd::MODE_RUN ? ensure_bin_partials(c, grid) : KD_OK;
}

// Prep + integrand kernels over p0.n_points points in chunks (and, in
// MODE_RUN, the deterministic partial reduction into d_acc); with `timed`
// each kernel is bracketed by HIP events on the ctx stream (kd_last_timing).
// `ext`: a compacted-point buffer of the caller's for up to kChunk points (no
// allocation of the ctx's own); only for evaluation modes without the flow stage.
int launch_chunks(kd_ctx *c, const kd::KdParams &p0, int timing, const kd::PrepBuf *ext) {
  const bool timed = timing != KD_TIMING_NONE;
  const bool kernels = timing == KD_TIMING_KERNELS;
  kd::PrepBuf bufs[2];
  int rc = KD_OK;
  if (ext)
    bufs[0] = *ext;
  else if ((rc = ensure_prep(c, c->order, p0.n_points, &bufs[0])))
    return rc;
  // several chunks: chunk c's points prepared on stream2 into bufs[c % 2]
  // while the Keldysh sum of chunk c - 1 runs on the ctx stream
  const bool pipe = !ext && p0.n_points > kChunk;
  if (pipe && (rc = ensure_pipeline(c, c->order, p0.n_points, &bufs[1]))) return rc;
  const int N = c->order;
  // The flow warper, when last in the chain, is applied first: its own stage
  // maps l -> x (and the Jacobian) before the rest of the chain.
  const bool flow = c->flow_in_chain && !c->flow_suppressed && p0.mode != kd::MODE_EVAL_U && N > 0;
  if (flow) {
    if (c->flow_order != N) return fail(c, KD_ERR_STATE, "flow warper dimension differs from the integrand order");
    if ((rc = ensure_flow_stage(c, N))) return rc;
  }
  const int64_t nchunks = (p0.n_points + kChunk - 1) / kChunk;
  if (kernels && (rc = ensure_events(c, size_t(4 * nchunks)))) return rc;
  int launches = 0;
  if (timed) KD_HIP(c, hipEventRecord(c->ev0, c->stream));
  int64_t ci = 0;
  for (int64_t off = 0; off < p0.n_points; off += kChunk, ci++) {
    // the stream that prepares this chunk's points, and their buffer
    hipStream_t sp = pipe && ci > 0 ? c->stream2 : c->stream;
    const kd::PrepBuf &buf = bufs[pipe ? ci & 1 : 0];
    if (pipe && ci > 0) {
      // stream2 starts after chunk 0's prep (the flow stage's buffer and any
      // earlier work on the ctx stream), and reuses a buffer only after the
      // Keldysh sum two chunks back has read it
      if (ci == 1) KD_HIP(c, hipStreamWaitEvent(sp, c->ev_sync[0], 0));
      if (ci >= 2) KD_HIP(c, hipStreamWaitEvent(sp, c->ev_sync[2 * (ci - 2) + 1], 0));
    }
    kd::KdParams p = p0;
    const int64_t m = std::min<int64_t>(kChunk, p0.n_points - off);
    p.n_points = m;
    p.k0 = p0.k0 + (uint32_t)off;
    if (p.in_pts) p.in_pts = p0.in_pts + off * N;
    if (p.out_vals) p.out_vals = p0.out_vals + 2 * off * vals_per_point(c);
    if (p.out_dom) p.out_dom = p0.out_dom + off;
    if (p.out_jac) p.out_jac = p0.out_jac + off;
    if (p.out_l) p.out_l = p0.out_l + off * N;
    const int grid = grid_for(c, m);
    const int64_t nb = (m + 255) / 256;
    const int grid_prep = (int)std::min<int64_t>(nb, (int64_t)(c->n_cu > 0 ? c->n_cu : 256) * 8);
    if (p.mode == kd::MODE_RUN) {
      if ((rc = ensure_partials(c, grid))) return rc;
      p.partials = c->d_part;
      if ((rc = ensure_bin_partials(c, grid))) return rc;
      p.bin_part = c->d_bin_part;
    } else {
      p.nbins = 0;
    }
    if (kernels) KD_HIP(c, hipEventRecord(c->ev[4 * ci + 0], sp));
    if (flow) {
      kd::FlowMapIO io{};
      io.Csr = p.Csr;
      io.shfr = p.shfr;
      io.k0 = p.k0;
      io.in_l = p.mode == kd::MODE_EVAL_L ? p.in_pts : nullptr;
      io.n_points = m;
      io.out_x = c->d_fx;
      io.out_jac = c->d_fx + (size_t)kChunk * N;
      io.out_l = p.mode == kd::MODE_SAMPLE ? p.out_l : nullptr;
      KD_HIP(c, kd::launch_flow_map(N, flow_params(c), io, flow_grid(c, m), sp));
      launches += 1;
      p.in_pts = io.out_x;
      p.in_jac = io.out_jac;
      p.out_l = nullptr;
      p.n_warpers = c->n_warpers - 1;
    } else if (c->flow_in_chain) {
      p.n_warpers = c->n_warpers - 1; // flow suppressed: the chain without it
    }
    KD_HIP(c, kd::launch_prep(N, p, buf, grid_prep, sp));
    if (kernels) KD_HIP(c, hipEventRecord(c->ev[4 * ci + 1], sp));
    if (pipe) {
      if (ci == 0) {
        KD_HIP(c, hipEventRecord(c->ev_sync[0], c->stream)); // chunk 0 prepared: stream2 may start
      } else {
        KD_HIP(c, hipEventRecord(c->ev_sync[2 * ci], sp));
        KD_HIP(c, hipStreamWaitEvent(c->stream, c->ev_sync[2 * ci], 0));
      }
    }
    if (kernels) KD_HIP(c, hipEventRecord(c->ev[4 * ci + 2], c->stream));
    if (c->obs != KD_OBS_DIRECT)
      KD_HIP(c, kd::launch_cofactor(N, p, buf, grid, c->stream));
    else if (use_tree(c))
      KD_HIP(c, kd::launch_tree(N, c->precision, p, buf, grid, c->stream));
    else if (use_split(c))
      KD_HIP(c, kd::launch_split(N, c->precision, p, buf, grid, c->stream));
    else if (use_ctree(c))
      KD_HIP(c, kd::launch_ctree(N, c->precision, p, buf, grid, c->stream));
    else if (use_wsplit(c))
      KD_HIP(c, kd::launch_wsplit(N, c->precision, p, buf, grid, c->stream));
    else
      KD_HIP(c, kd::launch_lu(N, c->precision, c->lu_lanes_max, use_pair(c), p, buf, grid, c->stream));
    if (kernels) KD_HIP(c, hipEventRecord(c->ev[4 * ci + 3], c->stream));
    // this chunk's records read: its buffer may be refilled (chunk ci + 2)
    if (pipe) KD_HIP(c, hipEventRecord(c->ev_sync[2 * ci + 1], c->stream));
    launches += 2;
    if (p.mode == kd::MODE_RUN && !c->no_reduce) {
      KD_HIP(c, kd::launch_reduce(c->d_part, grid, c->part_cap, c->d_acc, c->stream));
      launches += 1;
      if (p.nbins > 0) {
        KD_HIP(c, kd::launch_reduce_bins(c->d_bin_part, grid, p.nbins * c->bin_k, c->d_bins, c->stream));
        launches += 1;
      }
    }
  }
  if (timed) KD_HIP(c, hipEventRecord(c->ev1, c->stream));
  c->last_launches = launches;
  c->last_chunks = timed ? (int)nchunks : 0;
  c->last_kernels_timed = kernels;
  return KD_OK;
}

// Bin storage shared by the time binning (k = 1 value per bin) and the
// kernel method (k = 2: the Keldysh index): values [nbins][k], counts
// [nbins + 1] (values added per time bin, then the dropped ones).
int set_bins(kd_ctx *c, double xmin, double xmax, int nbins, int k) {
  if (nbins > 0 && !(xmin < xmax)) {
    return fail(c, KD_ERR_ARG, "Boundaries of binning range are in wrong order (" + std::to_string(xmin) +
                                   " >= " + std::to_string(xmax) + ")");
  }
  KD_HIP(c, hipSetDevice(c->device));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  dfree(c->d_bins);
  dfree(c->d_bin_part);
  dfree(c->d_bin_counts);
  c->bin_part_cap = 0;
  c->nbins = 0;
  c->bin_k = k;
  if (nbins == 0) return KD_OK;
  KD_HIP(c, hipMalloc(&c->d_bins, sizeof(double2) * nbins * k));
  KD_HIP(c, hipMalloc(&c->d_bin_counts, sizeof(unsigned long long) * (nbins + 1)));
  KD_HIP(c, hipMemset(c->d_bins, 0, sizeof(double2) * nbins * k));
  KD_HIP(c, hipMemset(c->d_bin_counts, 0, sizeof(unsigned long long) * (nbins + 1)));
  c->nbins = nbins;
  c->bin_min = xmin;
  c->bin_max = xmax;
  return KD_OK;
}

int rc_bins_clear(kd_ctx *c) { return set_bins(c, 0.0, 1.0, 0, 1); }

int finish_timing(kd_ctx *c) {
  if (c->last_chunks == 0) { // untimed (KD_TIMING_NONE) or nothing launched
    c->last_ms = c->last_lu_ms = c->last_prep_ms = 0.f;
    return KD_OK;
  }
  KD_HIP(c, hipEventSynchronize(c->ev1));
  float ms = 0.f, lu = 0.f, prep = 0.f, x = 0.f;
  KD_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  for (int i = 0; i < (c->last_kernels_timed ? c->last_chunks : 0); i++) {
    KD_HIP(c, hipEventElapsedTime(&x, c->ev[4 * i + 0], c->ev[4 * i + 1]));
    prep += x;
    KD_HIP(c, hipEventElapsedTime(&x, c->ev[4 * i + 2], c->ev[4 * i + 3]));
    lu += x;
  }
  c->last_ms = ms;
  c->last_lu_ms = lu;
  c->last_prep_ms = prep;
  return KD_OK;
}

} // namespace

extern "C" {

int kd_device_count(int *count) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return KD_OK;
}

int kd_create(kd_ctx **out, int device) {
  if (!out) return KD_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return KD_ERR_NODEVICE;
  if (device < 0 || device >= n) return KD_ERR_ARG;
  kd_ctx *c = new kd_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return KD_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->n_cu = prop.multiProcessorCount;
  resolve_scratch(device);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipMalloc(&c->d_acc, 4 * sizeof(double)) != hipSuccess ||
      hipHostMalloc(&c->h_acc, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
      hipMemset(c->d_acc, 0, 4 * sizeof(double)) != hipSuccess) {
    kd_destroy(c);
    return KD_ERR_HIP;
  }
  *out = c;
  return KD_OK;
}

int kd_destroy(kd_ctx *c) {
  if (!c) return KD_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dfree(c->d_lesser);
  dfree(c->d_greater);
  dfree(c->d_warpers);
  dfree(c->d_maps);
  dfree(c->d_tables);
  dfree(c->d_guides);
  dfree(c->d_knots);
  dfree(c->d_fw);
  dfree(c->d_fm);
  dfree(c->d_fv);
  dfree(c->d_fgrad);
  dfree(c->d_fpk);
  dfree(c->d_fx);
  dfree(c->d_tx);
  dfree(c->d_tw);
  dfree(c->d_tc);
  dfree(c->d_fpart);
  dfree(c->d_floss);
  dfree(c->d_bins);
  dfree(c->d_bin_part);
  dfree(c->d_bin_counts);
  dfree(c->d_Csr);
  c->d_shfr = nullptr;
  if (c->h_sobol) (void)hipHostFree(c->h_sobol);
  c->h_sobol = nullptr;
  if (c->h_acc) (void)hipHostFree(c->h_acc);
  c->h_acc = nullptr;
  if (c->ev_sobol) (void)hipEventDestroy(c->ev_sobol);
  c->ev_sobol = nullptr;
  dfree(c->d_acc);
  dfree(c->d_part);
  dfree(c->d_stage);
  dfree(c->d_prep);
  dfree(c->d_prep2);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->ev_sync) (void)hipEventDestroy(e);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return KD_OK;
}

const char *kd_last_error(const kd_ctx *c) { return c ? c->err.c_str() : "null context"; }

int kd_set_g0_table(kd_ctx *c, const double *lesser, const double *greater, int32_t n_orb,
                    int32_t n_pts, double time_max, double alpha) {
  if (!c) return KD_ERR_ARG;
  if (!lesser || !greater) return fail(c, KD_ERR_ARG, "null g0 table");
  if (n_orb != 1 && n_orb != 2) return fail(c, KD_ERR_ARG, "n_orb must be 1 or 2");
  if (n_pts < 2) return fail(c, KD_ERR_ARG, "g0 table needs at least 2 time points");
  if (!(time_max > 0.0)) return fail(c, KD_ERR_ARG, "time_max must be > 0");
  KD_HIP(c, hipSetDevice(c->device));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  dfree(c->d_lesser);
  dfree(c->d_greater);
  size_t bytes = sizeof(double2) * (size_t)n_orb * n_orb * n_pts;
  KD_HIP(c, hipMalloc(&c->d_lesser, bytes));
  KD_HIP(c, hipMalloc(&c->d_greater, bytes));
  KD_HIP(c, hipMemcpy(c->d_lesser, lesser, bytes, hipMemcpyHostToDevice));
  KD_HIP(c, hipMemcpy(c->d_greater, greater, bytes, hipMemcpyHostToDevice));
  c->n_pts = n_pts;
  c->n_orb = n_orb;
  c->time_max = time_max;
  c->alpha = alpha;
  c->has_g0 = true;
  return KD_OK;
}

int kd_set_integrand(kd_ctx *c, int32_t order, const kd_gf_index *A, const kd_gf_index *B,
                     double cutoff) {
  if (!c) return KD_ERR_ARG;
  if (!A || !B) return fail(c, KD_ERR_ARG, "null external point");
  if (order < 0 || order > KD_MAX_ORDER)
    return fail(c, KD_ERR_ARG, "order must be in [0, " + std::to_string(KD_MAX_ORDER) + "]");
  if (!c->has_g0) return fail(c, KD_ERR_STATE, "g0 table not set (kd_set_g0_table)");
  // wick_direct.hpp:50-59
  if (A->time < 0.0 || B->time < 0.0) return fail(c, KD_ERR_ARG, "An external point has negative time.");
  if (A->time > c->time_max || B->time > c->time_max)
    return fail(c, KD_ERR_ARG, "An external point is out of g0 time window.");
  if ((A->orbital != 0 || B->orbital != 0) && c->n_orb < 2)
    return fail(c, KD_ERR_ARG, "Need g0_model with make_dot_lead == true to calculate off_diagonal gf.");
  if (A->orbital < 0 || A->orbital > 1 || B->orbital < 0 || B->orbital > 1)
    return fail(c, KD_ERR_ARG, "orbital must be 0 or 1");
  if (c->obs != KD_OBS_DIRECT && c->nbins > 0 && (rc_bins_clear(c) != KD_OK)) return KD_ERR_HIP;
  c->order = order;
  c->A = *A;
  c->B = *B;
  c->cutoff = cutoff;
  c->obs = KD_OBS_DIRECT;
  return KD_OK;
}

int kd_set_precision(kd_ctx *c, int32_t precision) {
  if (!c) return KD_ERR_ARG;
  if (precision != KD_PREC_FP64 && precision != KD_PREC_FP32_LU && precision != KD_PREC_FP32_LU_ALL)
    return fail(c, KD_ERR_ARG, "unknown precision");
  c->precision = precision;
  return KD_OK;
}

int kd_eval_batch_device(kd_ctx *c, const double *d_u, int64_t n_points, int32_t keep_u_hypercube,
                         double *d_out, int32_t *d_in_domain) {
  if (!c) return KD_ERR_ARG;
  int rc = need_integrand(c);
  if (rc) return rc;
  if (c->obs == KD_OBS_KERNEL) return fail(c, KD_ERR_STATE, "kernel-method entries: use kd_eval_kernel_batch");
  if (n_points < 0) return fail(c, KD_ERR_ARG, "negative point count");
  if (n_points == 0) return KD_OK;
  KD_HIP(c, hipSetDevice(c->device));
  kd::KdParams p = base_params(c);
  p.mode = kd::MODE_EVAL_U;
  p.keep_u = keep_u_hypercube ? 1 : 0;
  p.in_pts = d_u;
  p.n_points = n_points;
  p.out_vals = d_out;
  p.out_dom = d_in_domain;
  return launch_timed(c, p, grid_for(c, n_points));
}

int kd_eval_batch_device_noalloc(kd_ctx *c, const double *d_u, int64_t n_points, int32_t keep_u_hypercube,
                                 double *d_out, int32_t *d_in_domain) {
  if (!c) return KD_ERR_ARG;
  int rc = need_integrand(c);
  if (rc) return rc;
  if (c->obs == KD_OBS_KERNEL) return fail(c, KD_ERR_STATE, "kernel-method entries: use kd_eval_kernel_batch");
  if (n_points < 0) return fail(c, KD_ERR_ARG, "negative point count");
  if (n_points > KD_NOALLOC_MAX_POINTS)
    return fail(c, KD_ERR_RANGE, "kd_eval_batch_device_noalloc: at most KD_NOALLOC_MAX_POINTS (" +
                                     std::to_string(KD_NOALLOC_MAX_POINTS) + ") points per call, got " +
                                     std::to_string(n_points));
  if (n_points == 0) return KD_OK;
  if ((!d_u && c->order > 0) || !d_out) return fail(c, KD_ERR_ARG, "null buffer");
  if (c->device < 0 || c->device >= kMaxDevices || !g_scratch_ptr[c->device])
    return fail(c, KD_ERR_HIP, std::string("scratch of kd_eval_batch_device_noalloc unavailable: ") +
                                   hipGetErrorString(c->device >= 0 && c->device < kMaxDevices
                                                         ? g_scratch_rc[c->device]
                                                         : hipErrorInvalidDevice));
  KD_HIP(c, hipSetDevice(c->device));
  // the compacted-point buffer carved from this device's scratch, laid out as
  // ensure_prep lays out the ctx's own (every slot read is written first by
  // the prep kernel of the same call, so nothing needs clearing)
  const size_t np = ((size_t)n_points + 63) / 64 * 64;
  const int N = c->order > 0 ? c->order : 1;
  char *base = (char *)g_scratch_ptr[c->device];
  kd::PrepBuf buf;
  buf.u = (double *)base;
  base += np * sizeof(double) * (size_t)N;
  buf.jac = (double *)base;
  base += np * sizeof(double);
  buf.idx = (int *)base;
  base += np * sizeof(int);
  buf.cnt = (int *)base;
  kd::KdParams p = base_params(c);
  p.mode = kd::MODE_EVAL_U;
  p.keep_u = keep_u_hypercube ? 1 : 0;
  p.in_pts = d_u;
  p.n_points = n_points;
  p.out_vals = d_out;
  p.out_dom = d_in_domain;
  return launch_chunks(c, p, KD_TIMING_NONE, &buf);
}

int kd_eval_batch(kd_ctx *c, const double *u, int64_t n_points, int32_t keep_u_hypercube, double *out,
                  int32_t *in_domain) {
  if (!c) return KD_ERR_ARG;
  int rc = need_integrand(c);
  if (rc) return rc;
  if (c->obs == KD_OBS_KERNEL) return fail(c, KD_ERR_STATE, "kernel-method entries: use kd_eval_kernel_batch");
  if (n_points <= 0) return n_points == 0 ? KD_OK : fail(c, KD_ERR_ARG, "negative point count");
  if ((!u && c->order > 0) || !out) return fail(c, KD_ERR_ARG, "null buffer");
  const size_t nu = (size_t)n_points * (size_t)c->order;
  const size_t bytes = sizeof(double) * (nu + 2 * (size_t)n_points) + sizeof(int32_t) * (size_t)n_points + 64;
  KD_HIP(c, hipSetDevice(c->device));
  if ((rc = ensure_stage(c, bytes))) return rc;
  double *du = (double *)c->d_stage;
  double *dout = du + nu;
  int32_t *ddom = (int32_t *)(dout + 2 * n_points);
  if (nu) KD_HIP(c, hipMemcpyAsync(du, u, sizeof(double) * nu, hipMemcpyHostToDevice, c->stream));
  if ((rc = kd_eval_batch_device(c, du, n_points, keep_u_hypercube, dout, ddom))) return rc;
  KD_HIP(c, hipMemcpyAsync(out, dout, sizeof(double) * 2 * n_points, hipMemcpyDeviceToHost, c->stream));
  if (in_domain)
    KD_HIP(c, hipMemcpyAsync(in_domain, ddom, sizeof(int32_t) * n_points, hipMemcpyDeviceToHost, c->stream));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  return finish_timing(c);
}

// The interval bsearch_knots (the device's GSL bisection over [0, n-1]) finds
// for x: max{i : xa[i] <= x} clamped to [0, n-2] when xa is non-decreasing.
static int knot_interval(const double *xa, int n, double x) {
  int ilo = 0, ihi = n - 1;
  while (ihi > ilo + 1) {
    const int i = (ihi + ilo) >> 1;
    if (xa[i] > x)
      ihi = i;
    else
      ilo = i;
  }
  return ilo;
}

// Guide of the knot search over xa[0..n-1] (kd::KnotGuide): 2n uniform buckets
// over [xa[0], xa[n-1]], each holding the intervals of its edges widened by
// 1e-6 of a bucket, far beyond the rounding of the device's bucket index, so
// that the interval of every x of the bucket lies in [lo, hi).  Arrays that are
// not finite and non-decreasing, or whose span is lost in the rounding of x - a,
// get no guide (G = 0: plain bisection).
static kd::KnotGuide knot_guide(const double *xa, int n, std::vector<int2> &out) {
  kd::KnotGuide g{0.0, 0.0, 0, 0};
  if (n < 3) return g;
  for (int i = 0; i < n; i++)
    if (!std::isfinite(xa[i]) || (i > 0 && xa[i] < xa[i - 1])) return g;
  const double a = xa[0], span = xa[n - 1] - a;
  const int G = 2 * n;
  const double width = span / G, margin = 1e-6 * width;
  if (!(span > 0.0) || !(margin > 1e-12 * (std::fabs(a) + std::fabs(xa[n - 1])))) return g;
  g.a = a;
  g.s = G / span;
  g.off = (int)out.size();
  g.G = G;
  for (int b = 0; b < G; b++) {
    const int lo = b == 0 ? 0 : knot_interval(xa, n, a + b * width - margin);
    const int hi = b == G - 1 ? n - 1 : knot_interval(xa, n, a + (b + 1) * width + margin) + 1;
    out.push_back(int2{lo, hi});
  }
  return g;
}

int kd_set_warper_chain(kd_ctx *c, const kd_warper_desc *w, int32_t n_w, const kd_map1d *maps,
                        int32_t n_maps, const double *tables, int64_t n_tables) {
  if (!c) return KD_ERR_ARG;
  if (n_w < 0 || n_maps < 0 || n_tables < 0) return fail(c, KD_ERR_ARG, "negative size");
  if (n_w > KD_MAX_WARPERS || n_maps > KD_MAX_MAPS)
    return fail(c, KD_ERR_ARG, "a warper chain holds at most " + std::to_string(KD_MAX_WARPERS) + " warpers and " +
                                   std::to_string(KD_MAX_MAPS) + " maps");
  if ((n_w && !w) || (n_maps && !maps) || (n_tables && !tables)) return fail(c, KD_ERR_ARG, "null buffer");
  for (int i = 0; i < n_w; i++) {
    const kd_warper_desc &d = w[i];
    if (d.kind == KD_WARP_PRODUCT) {
      if (d.n_maps < 1 || d.map_offset < 0 || d.map_offset + d.n_maps > n_maps)
        return fail(c, KD_ERR_ARG, "product warper map range out of bounds");
    } else if (d.kind == KD_WARP_FLOW) {
      if (!c->d_fw) return fail(c, KD_ERR_STATE, "flow warper used before its parameters were set");
      if (i != n_w - 1)
        return fail(c, KD_ERR_ARG, "the flow warper must be the last warper of the chain (it is applied first)");
    } else if (d.kind != KD_WARP_IDENTITY && d.kind != KD_WARP_PLASMA) {
      return fail(c, KD_ERR_ARG, "unknown warper kind " + std::to_string(d.kind));
    }
  }
  for (int i = 0; i < n_maps; i++) {
    const kd_map1d &m = maps[i];
    if (m.kind < KD_MAP_UNIFORM || m.kind > KD_MAP_PROJECTION) return fail(c, KD_ERR_ARG, "unknown map kind");
    if (m.kind == KD_MAP_PROJECTION) {
      auto in = [&](int64_t off, int64_t len) { return off >= 0 && off + len <= n_tables; };
      if (m.n_pts < 2 || m.n_inv < 2) return fail(c, KD_ERR_ARG, "bad mesh sizes");
      if (!in(m.t_off, m.n_inv) || !in(m.f_off, m.n_pts)) return fail(c, KD_ERR_ARG, "map table out of bounds");
      if (!(m.norm > 0.0)) return fail(c, KD_ERR_ARG, "projection map normalisation must be positive");
    }
    if (m.kind == KD_MAP_NEAREST || m.kind == KD_MAP_HYBRID) {
      if (m.n_pts < 2 || m.n_inv < 2 || m.n_inv > m.n_pts) return fail(c, KD_ERR_ARG, "bad knot counts");
      auto in = [&](int64_t off, int64_t len) { return off >= 0 && off + len <= n_tables; };
      if (!in(m.t_off, m.n_pts) || !in(m.F_off, m.n_pts)) return fail(c, KD_ERR_ARG, "map table out of bounds");
      if (m.kind == KD_MAP_NEAREST && !in(m.f_off, m.n_pts)) return fail(c, KD_ERR_ARG, "map table out of bounds");
      if (m.kind == KD_MAP_HYBRID && !in(m.c_off, 4 * (int64_t)(m.n_inv - 1)))
        return fail(c, KD_ERR_ARG, "map table out of bounds");
    }
  }
  std::vector<kd::WarperDev> wdev(n_w);
  for (int i = 0; i < n_w; i++) {
    wdev[i].w = w[i];
    wdev[i].map_kind = -1;
    wdev[i].pad = 0;
    if (w[i].kind == KD_WARP_PRODUCT) {
      const kd_map1d *m = maps + w[i].map_offset;
      wdev[i].map_kind = m[0].kind;
      for (int k = 1; k < w[i].n_maps; k++)
        if (m[k].kind != m[0].kind) wdev[i].map_kind = -1;
    }
  }
  std::vector<kd::MapDev> dev(n_maps);
  std::vector<int2> guides(1, int2{0, 1}); // entry 0: what maps without a guide read (and ignore)
  std::vector<double> knots(4, 0.0); // records start 32-byte aligned (the buffer is)
  for (int i = 0; i < n_maps; i++) {
    const kd_map1d &m = maps[i];
    dev[i].m = m;
    dev[i].gF = dev[i].gt = kd::KnotGuide{0.0, 0.0, 0, 0};
    dev[i].rec = dev[i].fpair = 0;
    dev[i].F0 = dev[i].Fl = 0.0;
    if (m.kind == KD_MAP_NEAREST || m.kind == KD_MAP_HYBRID) dev[i].gF = knot_guide(tables + m.F_off, m.n_inv, guides);
    if (m.kind == KD_MAP_NEAREST) {
      dev[i].gt = knot_guide(tables + m.t_off, m.n_pts, guides);
      // interval records: i in [0, n_inv - 1) for the inverse CDF, j in
      // [0, n_pts - 1) for f (the searches return intervals in those ranges)
      const double *F = tables + m.F_off, *t = tables + m.t_off, *f = tables + m.f_off;
      dev[i].F0 = F[0];
      dev[i].Fl = F[m.n_inv - 1];
      dev[i].rec = (int64_t)knots.size();
      for (int k = 0; k + 1 < m.n_inv; k++) knots.insert(knots.end(), {F[k], F[k + 1], t[k], t[k + 1]});
      dev[i].fpair = (int64_t)knots.size();
      for (int k = 0; k + 1 < m.n_pts; k++) knots.insert(knots.end(), {f[k], f[k + 1]});
      if (knots.size() & 3) knots.insert(knots.end(), 4 - (knots.size() & 3), 0.0);
    }
  }
  KD_HIP(c, hipSetDevice(c->device));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  dfree(c->d_warpers);
  dfree(c->d_maps);
  dfree(c->d_tables);
  dfree(c->d_guides);
  dfree(c->d_knots);
  if (n_w) {
    KD_HIP(c, hipMalloc(&c->d_warpers, sizeof(kd::WarperDev) * n_w));
    KD_HIP(c, hipMemcpy(c->d_warpers, wdev.data(), sizeof(kd::WarperDev) * n_w, hipMemcpyHostToDevice));
  }
  if (n_maps) {
    KD_HIP(c, hipMalloc(&c->d_maps, sizeof(kd::MapDev) * n_maps));
    KD_HIP(c, hipMemcpy(c->d_maps, dev.data(), sizeof(kd::MapDev) * n_maps, hipMemcpyHostToDevice));
  }
  if (n_maps) {
    KD_HIP(c, hipMalloc(&c->d_knots, sizeof(double) * knots.size()));
    KD_HIP(c, hipMemcpy(c->d_knots, knots.data(), sizeof(double) * knots.size(), hipMemcpyHostToDevice));
    KD_HIP(c, hipMalloc(&c->d_guides, sizeof(int2) * guides.size()));
    KD_HIP(c, hipMemcpy(c->d_guides, guides.data(), sizeof(int2) * guides.size(), hipMemcpyHostToDevice));
  }
  if (n_tables) {
    KD_HIP(c, hipMalloc(&c->d_tables, sizeof(double) * n_tables));
    KD_HIP(c, hipMemcpy(c->d_tables, tables, sizeof(double) * n_tables, hipMemcpyHostToDevice));
  }
  c->n_warpers = n_w;
  c->n_maps = n_maps;
  c->h_warpers.assign(w, w + n_w);
  c->flow_in_chain = n_w > 0 && w[n_w - 1].kind == KD_WARP_FLOW;
  return KD_OK;
}

int kd_eval_warped_batch(kd_ctx *c, const double *l, int64_t n_points, int32_t keep_u_hypercube, double *out,
                         int32_t *in_domain, double *jacobian) {
  if (!c) return KD_ERR_ARG;
  int rc = need_integrand(c);
  if (rc) return rc;
  if (c->obs == KD_OBS_KERNEL) return fail(c, KD_ERR_STATE, "kernel-method entries: use kd_eval_kernel_batch");
  if (n_points <= 0) return n_points == 0 ? KD_OK : fail(c, KD_ERR_ARG, "negative point count");
  if ((!l && c->order > 0) || !out) return fail(c, KD_ERR_ARG, "null buffer");
  const size_t nl = (size_t)n_points * (size_t)c->order;
  const size_t bytes = sizeof(double) * (nl + 3 * (size_t)n_points) + sizeof(int32_t) * (size_t)n_points + 64;
  KD_HIP(c, hipSetDevice(c->device));
  if ((rc = ensure_stage(c, bytes))) return rc;
  double *dl = (double *)c->d_stage;
  double *dout = dl + nl;
  double *djac = dout + 2 * n_points;
  int32_t *ddom = (int32_t *)(djac + n_points);
  if (nl) KD_HIP(c, hipMemcpyAsync(dl, l, sizeof(double) * nl, hipMemcpyHostToDevice, c->stream));
  kd::KdParams p = base_params(c);
  p.mode = kd::MODE_EVAL_L;
  p.keep_u = keep_u_hypercube ? 1 : 0;
  p.in_pts = dl;
  p.n_points = n_points;
  p.out_vals = dout;
  p.out_dom = ddom;
  p.out_jac = djac;
  if ((rc = launch_timed(c, p, grid_for(c, n_points)))) return rc;
  KD_HIP(c, hipMemcpyAsync(out, dout, sizeof(double) * 2 * n_points, hipMemcpyDeviceToHost, c->stream));
  if (in_domain)
    KD_HIP(c, hipMemcpyAsync(in_domain, ddom, sizeof(int32_t) * n_points, hipMemcpyDeviceToHost, c->stream));
  if (jacobian)
    KD_HIP(c, hipMemcpyAsync(jacobian, djac, sizeof(double) * n_points, hipMemcpyDeviceToHost, c->stream));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  return finish_timing(c);
}

constexpr int kSobolMaxDim = 32;
constexpr size_t kSobolWords = 31 * kSobolMaxDim + kSobolMaxDim;

int kd_set_sobol(kd_ctx *c, int32_t dim, const uint32_t *Csr, const uint32_t *shfr, uint32_t k) {
  if (!c) return KD_ERR_ARG;
  if (dim <= 0) return fail(c, KD_ERR_ARG, "Dimension of the integration space must be > 0.");
  if (dim > kSobolMaxDim) return fail(c, KD_ERR_ARG, "Sobol dimension must be <= 32");
  if (!Csr || !shfr) return fail(c, KD_ERR_ARG, "null Sobol table");
  KD_HIP(c, hipSetDevice(c->device));
  // No host wait for the device: the new tables go in by copies on the ctx
  // stream, behind every launch that reads the old ones.  Only the staging
  // buffer's previous copy must have been consumed before it is refilled.
  if (!c->d_Csr) {
    KD_HIP(c, hipMalloc(&c->d_Csr, sizeof(uint32_t) * kSobolWords));
    c->d_shfr = c->d_Csr + 31 * kSobolMaxDim;
  }
  if (!c->h_sobol) KD_HIP(c, hipHostMalloc(&c->h_sobol, sizeof(uint32_t) * kSobolWords, hipHostMallocDefault));
  if (!c->ev_sobol) KD_HIP(c, hipEventCreateWithFlags(&c->ev_sobol, hipEventDisableTiming));
  if (c->sobol_pending) KD_HIP(c, hipEventSynchronize(c->ev_sobol));
  std::memcpy(c->h_sobol, Csr, sizeof(uint32_t) * 31 * dim);
  std::memcpy(c->h_sobol + 31 * kSobolMaxDim, shfr, sizeof(uint32_t) * dim);
  KD_HIP(c, hipMemcpyAsync(c->d_Csr, c->h_sobol, sizeof(uint32_t) * 31 * dim, hipMemcpyHostToDevice, c->stream));
  KD_HIP(c, hipMemcpyAsync(c->d_shfr, c->h_sobol + 31 * kSobolMaxDim, sizeof(uint32_t) * dim,
                           hipMemcpyHostToDevice, c->stream));
  KD_HIP(c, hipEventRecord(c->ev_sobol, c->stream));
  c->sobol_pending = true;
  c->sobol_dim = dim;
  c->k = k;
  c->has_sobol = true;
  return KD_OK;
}

int kd_get_sobol_state(const kd_ctx *c, uint32_t *k) {
  if (!c || !k) return KD_ERR_ARG;
  *k = c->k;
  return KD_OK;
}

static int run_common(kd_ctx *c, uint64_t n_steps, int mode, double *d_l, double *d_value) {
  int rc = need_integrand(c);
  if (rc) return rc;
  if (mode == kd::MODE_SAMPLE && c->obs == KD_OBS_KERNEL)
    return fail(c, KD_ERR_STATE, "per-point samples need a scalar integrand (not the binned kernel method)");
  if (c->nbins > 0 && c->order == 0) return fail(c, KD_ERR_ARG, "Order 0 is not implemented");
  if (!c->has_sobol) return fail(c, KD_ERR_STATE, "rng not set (kd_set_sobol)");
  if (c->sobol_dim != c->order) return fail(c, KD_ERR_ARG, "Sobol dimension differs from the integrand order");
  if ((uint64_t)c->k + n_steps > (1ull << 31))
    return fail(c, KD_ERR_RANGE, "More sample points requested then generator can hold.");
  if (n_steps == 0) return KD_OK;
  KD_HIP(c, hipSetDevice(c->device));
  kd::KdParams p = base_params(c);
  p.mode = mode;
  p.n_points = (int64_t)n_steps;
  if (mode != kd::MODE_RUN) {
    p.out_l = d_l;
    p.out_vals = d_value;
  }
  if ((rc = launch_timed(c, p, 0))) return rc;
  c->k += (uint32_t)n_steps;
  if (mode == kd::MODE_RUN) c->n_points += n_steps;
  return KD_OK;
}

int kd_get_accum(kd_ctx *c, kd_accum *out) {
  if (!c || !out) return KD_ERR_ARG;
  double h[4];
  KD_HIP(c, hipSetDevice(c->device));
  KD_HIP(c, hipMemcpyAsync(h, c->d_acc, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  out->re = h[0];
  out->im = h[1];
  out->n_points = c->n_points;
  out->n_in_domain = (uint64_t)h[2];
  out->n_fallback = (uint64_t)h[3];
  return KD_OK;
}

int kd_run(kd_ctx *c, uint64_t n_steps, kd_accum *out) {
  if (!c) return KD_ERR_ARG;
  int rc = run_common(c, n_steps, kd::MODE_RUN, nullptr, nullptr);
  if (rc) return rc;
  if (!out) return n_steps ? finish_timing(c) : KD_OK;
  // the accumulator's copy queued behind the launches: one host wait for both
  KD_HIP(c, hipMemcpyAsync(c->h_acc, c->d_acc, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  if (n_steps && (rc = finish_timing(c))) return rc;
  out->re = c->h_acc[0];
  out->im = c->h_acc[1];
  out->n_points = c->n_points;
  out->n_in_domain = (uint64_t)c->h_acc[2];
  out->n_fallback = (uint64_t)c->h_acc[3];
  return KD_OK;
}

int kd_reset_accum(kd_ctx *c) {
  if (!c) return KD_ERR_ARG;
  KD_HIP(c, hipSetDevice(c->device));
  KD_HIP(c, hipMemsetAsync(c->d_acc, 0, 4 * sizeof(double), c->stream));
  if (c->nbins > 0) {
    KD_HIP(c, hipMemsetAsync(c->d_bins, 0, sizeof(double2) * c->nbins * c->bin_k, c->stream));
    KD_HIP(c, hipMemsetAsync(c->d_bin_counts, 0, sizeof(unsigned long long) * (c->nbins + 1), c->stream));
  }
  KD_HIP(c, hipStreamSynchronize(c->stream));
  c->n_points = 0;
  return KD_OK;
}

int kd_sample_device(kd_ctx *c, uint64_t n_steps, double *d_l, double *d_value) {
  if (!c) return KD_ERR_ARG;
  if (!d_l || !d_value) return fail(c, KD_ERR_ARG, "null buffer");
  int rc = run_common(c, n_steps, kd::MODE_SAMPLE, d_l, d_value);
  if (rc) return rc;
  if (n_steps) return finish_timing(c);
  return KD_OK;
}

int kd_set_timing(kd_ctx *c, int32_t level) {
  if (!c) return KD_ERR_ARG;
  if (level != KD_TIMING_NONE && level != KD_TIMING_TOTAL && level != KD_TIMING_KERNELS)
    return fail(c, KD_ERR_ARG, "timing level must be KD_TIMING_NONE, KD_TIMING_TOTAL or KD_TIMING_KERNELS");
  c->timing = level;
  return KD_OK;
}

int kd_last_timing(const kd_ctx *c, double *total_ms, double *lu_ms, double *prep_ms, int32_t *launches) {
  if (!c) return KD_ERR_ARG;
  if (total_ms) *total_ms = c->last_ms;
  if (lu_ms) *lu_ms = c->last_lu_ms;
  if (prep_ms) *prep_ms = c->last_prep_ms;
  if (launches) *launches = c->last_launches;
  return KD_OK;
}

// Executed flops per point of the tree algorithm (kd_tree.hip): a child table
// of size Dc costs 8 Dc^2 (entries, one complex multiply-add each) + 13 Dc
// (scaled pivot row: reciprocal 7 + multiply 6); the bottom 4x4 bordered LU
// per configuration 14 complex multiply-adds (112) + 3 reciprocals (21) + 6
// row scalings (36) + 3 determinant products (18), then path x bottom
// determinant (12), det^2 * s (12) and the signed FP64 sum (2).
// Executed flops per point of the kernel method (kd_cofactor.hip), per
// configuration: the row-wise LU of the n vertex rows carried over the extra
// column (complex multiply-adds sum_{i<n} sum_{k<i} (n - k), 8 flops each; n-1
// reciprocals, 7; sum_{i<n-1} (n - i) row scalings and n-1 pivot products, 6
// each), p' (2), the scaled back substitution (n-1 products, sum_{i<n-1}
// (n-1-i) multiply-adds), the coefficient (3 products) and n entries (6 each),
// plus the accumulation (2 per entry).
static double cofactor_flops(int N) {
  if (N < 1) return 0.0;
  double macs = 0.0, muls = 0.0;
  for (int i = 0; i < N; i++)
    for (int k = 0; k < i; k++) macs += N - k;
  for (int i = 0; i < N - 1; i++) muls += (N - i) + 1;
  for (int i = 0; i < N - 1; i++) macs += N - 1 - i;
  muls += (N - 1) + 3 + N;
  const double per_cfg = 8 * macs + 6 * muls + 7 * (N - 1) + 2 + 4.0 * N;
  return std::ldexp(per_cfg, N);
}

// Executed flops per point of the paired LU (config_value_pair), per pair of
// configurations: the N-1 shared vertex rows over N+2 columns and the external
// row (complex multiply-add 8, multiply 6, reciprocal 7), the last vertex row
// and the corner per Keldysh index, det^2 * s, and the two signed FP64 sums.
static double pair_flops(int N) {
  const int M = N + 2;
  double macs = 0.0, muls = 0.0, invs = 0.0;
  for (int i = 0; i < N - 1; i++) {
    for (int k = 0; k < i; k++) macs += M - 1 - k;
    invs += 1;
    muls += (M - 1 - i) + 1;
  }
  for (int k = 0; k < N - 1; k++) macs += M - 1 - k;
  for (int t = 0; t < 2; t++) {
    for (int k = 0; k < N - 1; k++) macs += (N - 2 - k) + 2;
    macs += 1;
    invs += 1;
    muls += 4;
  }
  const double per_pair = 8 * macs + 6 * muls + 7 * invs + 4;
  return std::ldexp(per_pair, N - 1);
}

static double tree_flops(int N) {
  auto node = [](double dc) { return 8.0 * dc * dc + 13.0 * dc; };
  const int A = N - 6;
  double f = 0.0;
  for (int k = 1; k <= A; k++) f += std::ldexp(node(2.0 * (N - k) + 1), k);
  f += std::ldexp(2 * node(11) + 4 * node(9) + 8 * node(7), A);
  f += std::ldexp(112.0 + 21 + 36 + 18 + 12 + 12 + 2, N);
  return f;
}

// split: per child node of the tree levels 1..K one pivot reciprocal (7),
// Dc multipliers (6 each), Dc^2 multiply-adds (8 each) and the path
// determinant (6); per configuration the bordered LU over M = n - K vertices,
// det^2 s, the path determinant (6) and the signed FP64 sum -- or, where the
// kernel pairs the two choices of the last vertex, the paired LU's count per
// pair (as pair_flops over M vertices) plus two path determinants
static double split_like_flops(int N, int K, bool paired) {
  const int M = N - K;
  double f = 0.0;
  for (int k = 0; k < K; k++) {
    const double dc = 2.0 * (N - k) - 1;
    f += std::ldexp(8 * dc * dc + 6 * dc + 13, k + 1);
  }
  if (paired) {
    const int MC = M + 2;
    double macs = 0, muls = 0, invs = 0;
    for (int i = 0; i < M - 1; i++) {
      for (int k = 0; k < i; k++) macs += MC - 1 - k;
      invs += 1;
      muls += (MC - 1 - i) + 1;
    }
    for (int k = 0; k < M - 1; k++) macs += MC - 1 - k;
    for (int t = 0; t < 2; t++) {
      for (int k = 0; k < M - 1; k++) macs += (M - 2 - k) + 2;
      macs += 1;
      muls += 3; // p rb_B, p x, (dp det)^2 times it
    }
    muls += 2; // (dp det)^2, shared by the two rows
    return f + std::ldexp(8 * macs + 6 * muls + 7 * invs + 4, N - 1);
  }
  const double m = M;
  const double macs = m * (m + 1) * (2 * m + 1) / 6.0;
  const double muls = m * (m + 1) / 2.0 + m;
  f += std::ldexp(8 * macs + 6 * muls + 7 * m + 14 + 6, N);
  return f;
}

static double split_flops(int N) { return split_like_flops(N, N <= 7 ? 2 : 3, kd::split_paired(N)); }
// the wave-per-point split: K = n - 5 tree levels from order 8 (2 below), the
// last two formed in one step from level K - 2 -- per level-K node and row of
// its Dc = 2(n-K)+1, the 2x2 pivot block's determinant and inverse (21) and
// the row's two coefficients (40), two multiply-adds per entry, and per node
// the second pivot and the path determinant (19) -- then the five remaining
// vertices paired (order 6: four, unpaired)
static double wsplit_flops(int N) {
  const int K = N <= 7 ? 2 : N - 5;
  double f = split_like_flops(N, K, N >= 7);
  for (int k = K - 2; k < K; k++) { // the two single steps split_like_flops counted
    const double dc = 2.0 * (N - k) - 1;
    f -= std::ldexp(8 * dc * dc + 6 * dc + 13, k + 1);
  }
  const double dc = 2.0 * (N - K) + 1;
  return f + std::ldexp(16 * dc * dc + 61 * dc + 19, K);
}

// The time-ordered tree, as executed (complex multiply-add 8, multiply 6,
// subtraction 2, reciprocal 7, FP64 complex product 6): levels 1..3 breadth
// first, each (child, row) item with its own pivot reciprocal and row factor
// (the root's pivot reciprocal is a constant), a subtraction and a
// multiply-add per entry (a multiply at level 1); per lane, its forced levels
// from 3 (pivot, reciprocal, the factors u and w, the path determinant; every
// entry of the levels after 3 one multiply-add) down to its own node -- level
// 6, or at order 10 level 7 in two passes over vertex 6's index -- then the
// subtree below it; per leaf the last pivot, the border factors,
// x = p (p c_ee + u_e v_e) and its signed FP64 sum, per pair of leaves dp^2 and
// its product with that sum.
static double ctree_flops(int N) {
  // levels 2 and 3 breadth first: per item a pivot (sub, reciprocal), the row
  // factor, a subtraction and a multiply-add per entry and the path product;
  // the level-2 items also form their parent's level-1 entries from Wv
  // (2 rc + 4 complex products: level 1 has no pass of its own)
  double f = 0.0;
  for (int k = 1; k < 3; k++) {
    const double rc = N - k, items = std::ldexp(rc, k + 1);
    double per_item = 2 + 2 + 7 + 6 + rc * (2 + 8) + 6;
    if (k == 1) per_item += (2 * rc + 4) * 6;
    f += items * per_item;
  }
  const double leaf = 2 + 4 + 20 + 4, leaf_pair = 6 + 6 + 2;
  auto step = [&](int k) { const double r = N - k + 1; return 2 + 7 + (r - 1) * (2 + 2 + 6) + 6; };
  // each lane's path from its level-3 node to its own node at level kf (the
  // entries of levels 4..kf), then its subtree; lanes per point: 64, or 32
  // when two points share a wave (the lane's node at level 5)
  const bool pairs = kd::ctree_pairs(N);
  const int kf = pairs ? 5 : N <= 9 ? 6 : 7, passes = N == 10 ? 2 : 1;
  double lane = 0.0;
  if (!pairs && N == 6) {
    lane = step(3) + step(4) + 8.0 * (9 + 4) + leaf + leaf_pair; // levels 4 and 5 entries, the lane's one leaf
  } else {
    double path = 0.0;
    for (int k = 3; k < kf; k++) path += step(k);
    for (int k = 4; k <= kf; k++) path += 8.0 * (N - k + 1) * (N - k + 1); // entries of levels 4..kf
    double sub = 0.0;
    for (int k = kf; k < N - 1; k++) { // internal subtree levels: 2^(k-kf+1) children
      const double r = N - k + 1;
      sub += std::ldexp(step(k) + 8 * (r - 1) * (r - 1), k - kf + 1);
    }
    sub += std::ldexp(leaf_pair, N - 1 - kf) + std::ldexp(leaf, N - kf);
    lane = passes * (path + sub);
  }
  return f + (pairs ? 32.0 : 64.0) * lane;
}

int kd_flops_per_point(const kd_ctx *c, double *flops) {
  if (!c || !flops) return KD_ERR_ARG;
  if (c->obs != KD_OBS_DIRECT) {
    *flops = cofactor_flops(c->order);
    return KD_OK;
  }
  if (use_tree(c)) {
    *flops = tree_flops(c->order);
    return KD_OK;
  }
  if (use_pair(c)) {
    *flops = pair_flops(c->order);
    return KD_OK;
  }
  if (use_split(c)) {
    *flops = split_flops(c->order);
    return KD_OK;
  }
  if (use_ctree(c)) {
    *flops = ctree_flops(c->order);
    return KD_OK;
  }
  if (use_wsplit(c)) {
    *flops = wsplit_flops(c->order);
    return KD_OK;
  }
  const double N = c->order < 0 ? 0 : c->order;
  // per configuration: row-wise bordered LU (complex MAC = 8, complex mul = 6,
  // complex reciprocal = 7), det^2 * s (12) and the signed FP64 sum (2)
  const double macs = N * (N + 1) * (2 * N + 1) / 6.0;
  const double muls = N * (N + 1) / 2.0 + N;
  const double per_cfg = 8 * macs + 6 * muls + 7 * N + 14;
  *flops = std::ldexp(per_cfg, (int)N);
  return KD_OK;
}

// ---------------------------------------------------------------------------
// learned flow warper

static uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int kd_flow_n_params(int32_t order, int32_t n_layers, int64_t *n) { return kd_flow_n_params_cond(order, n_layers, 0, n); }

int kd_flow_n_params_cond(int32_t order, int32_t n_layers, int32_t n_cond, int64_t *n) {
  if (!n || order < 1 || order > KD_MAX_ORDER || n_layers < 1 || n_layers > KD_FLOW_MAX_LAYERS || n_cond < 0 ||
      n_cond > 1)
    return KD_ERR_ARG;
  *n = kd::flow_n_params(order, n_layers, n_cond);
  return KD_OK;
}

int kd_flow_set_conditioning(kd_ctx *c, int32_t n_cond, double value) {
  if (!c) return KD_ERR_ARG;
  if (n_cond < 0 || n_cond > 1) return fail(c, KD_ERR_ARG, "n_cond must be 0 or 1");
  if (!std::isfinite(value)) return fail(c, KD_ERR_ARG, "non-finite conditioning value");
  c->flow_nc = n_cond;
  c->flow_cond = n_cond ? value : 0.0;
  return KD_OK;
}

static int flow_alloc(kd_ctx *c, int32_t order, int32_t n_layers) {
  if (order < 1 || order > KD_MAX_ORDER) return fail(c, KD_ERR_ARG, "flow dimension must be in [1, 12]");
  if (n_layers < 1 || n_layers > KD_FLOW_MAX_LAYERS) return fail(c, KD_ERR_ARG, "flow layers must be in [1, 16]");
  KD_HIP(c, hipSetDevice(c->device));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  if (c->flow_order != order || c->flow_layers != n_layers || c->flow_alloc_nc != c->flow_nc || !c->d_fw) {
    dfree(c->d_fw);
    dfree(c->d_fm);
    dfree(c->d_fv);
    dfree(c->d_fgrad);
    dfree(c->d_fpk);
    c->flow_order = -1;
    const size_t np = (size_t)kd::flow_n_params(order, n_layers, c->flow_nc);
    KD_HIP(c, hipMalloc(&c->d_fw, np * sizeof(float)));
    KD_HIP(c, hipMalloc(&c->d_fm, np * sizeof(float)));
    KD_HIP(c, hipMalloc(&c->d_fv, np * sizeof(float)));
    KD_HIP(c, hipMalloc(&c->d_fgrad, np * sizeof(float)));
    KD_HIP(c, hipMalloc(&c->d_fpk, sizeof(uint16_t) * (size_t)kd::kFlowPackPerLayer * n_layers));
    c->flow_order = order;
    c->flow_layers = n_layers;
    c->flow_alloc_nc = c->flow_nc;
  }
  return KD_OK;
}

// upload params, reset Adam, rebuild the bf16 copies
static int flow_upload(kd_ctx *c, const float *params) {
  const size_t np = (size_t)kd::flow_n_params(c->flow_order, c->flow_layers, c->flow_alloc_nc);
  KD_HIP(c, hipMemcpy(c->d_fw, params, np * sizeof(float), hipMemcpyHostToDevice));
  KD_HIP(c, hipMemsetAsync(c->d_fm, 0, np * sizeof(float), c->stream));
  KD_HIP(c, hipMemsetAsync(c->d_fv, 0, np * sizeof(float), c->stream));
  KD_HIP(c, kd::launch_flow_pack(c->d_fw, c->d_fpk, c->flow_order, c->flow_layers, c->flow_alloc_nc, c->stream));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  c->flow_step = 0;
  return KD_OK;
}

int kd_flow_init(kd_ctx *c, int32_t order, int32_t n_layers, uint64_t seed) {
  if (!c) return KD_ERR_ARG;
  int rc = flow_alloc(c, order, n_layers);
  if (rc) return rc;
  // He-uniform W1, W2 (fan-in nA and H) and Wc (as W1); zero biases; W3 = b3 = 0: the identity map
  const int nc = c->flow_alloc_nc;
  std::vector<float> w((size_t)kd::flow_n_params(order, n_layers, nc), 0.f);
  uint64_t st = seed;
  auto unif = [&]() { return (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0; };
  for (int k = 0; k < n_layers; k++) {
    const int nA = kd::flow_nA(order, k), H = KD_FLOW_HIDDEN;
    float *W1 = w.data() + kd::flow_layer_offset(order, k, nc);
    float *W2 = W1 + H * nA + H;
    const double s1 = std::sqrt(6.0 / std::max(nA, 1)), s2 = std::sqrt(6.0 / H);
    for (int i = 0; i < H * nA; i++) W1[i] = (float)(unif() * s1);
    for (int i = 0; i < H * H; i++) W2[i] = (float)(unif() * s2);
    if (nc) {
      float *Wc = W1 + kd::flow_layer_size(order, k, nc) - H;
      for (int i = 0; i < H; i++) Wc[i] = (float)(unif() * s1);
    }
  }
  return flow_upload(c, w.data());
}

int kd_flow_set_params(kd_ctx *c, int32_t order, int32_t n_layers, const float *params, int64_t n) {
  if (!c) return KD_ERR_ARG;
  if (!params) return fail(c, KD_ERR_ARG, "null buffer");
  if (order < 1 || order > KD_MAX_ORDER || n_layers < 1 || n_layers > KD_FLOW_MAX_LAYERS ||
      n != kd::flow_n_params(order, n_layers, c->flow_nc))
    return fail(c, KD_ERR_ARG, "flow parameter count does not match (order, layers, conditioning inputs)");
  for (int64_t i = 0; i < n; i++)
    if (!std::isfinite(params[i])) return fail(c, KD_ERR_ARG, "non-finite flow parameter");
  int rc = flow_alloc(c, order, n_layers);
  if (rc) return rc;
  return flow_upload(c, params);
}

int kd_flow_get_params(kd_ctx *c, float *params, int64_t n) {
  if (!c) return KD_ERR_ARG;
  if (!c->d_fw) return fail(c, KD_ERR_STATE, "flow not initialised (kd_flow_init)");
  if (!params || n != kd::flow_n_params(c->flow_order, c->flow_layers, c->flow_alloc_nc))
    return fail(c, KD_ERR_ARG, "flow parameter buffer size does not match");
  KD_HIP(c, hipSetDevice(c->device));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  KD_HIP(c, hipMemcpy(params, c->d_fw, sizeof(float) * n, hipMemcpyDeviceToHost));
  return KD_OK;
}

int kd_flow_map(kd_ctx *c, const double *l, int64_t n, double *x, double *jac) {
  if (!c) return KD_ERR_ARG;
  if (!c->d_fw) return fail(c, KD_ERR_STATE, "flow not initialised (kd_flow_init)");
  if (n < 0) return fail(c, KD_ERR_ARG, "negative point count");
  if (n == 0) return KD_OK;
  if (!l || !x || !jac) return fail(c, KD_ERR_ARG, "null buffer");
  const int N = c->flow_order;
  const size_t nl = (size_t)n * N;
  KD_HIP(c, hipSetDevice(c->device));
  int rc = ensure_stage(c, sizeof(double) * (2 * nl + (size_t)n) + 64);
  if (rc) return rc;
  double *dl = (double *)c->d_stage, *dx = dl + nl, *dj = dx + nl;
  KD_HIP(c, hipMemcpyAsync(dl, l, sizeof(double) * nl, hipMemcpyHostToDevice, c->stream));
  kd::FlowMapIO io{};
  io.in_l = dl;
  io.n_points = n;
  io.out_x = dx;
  io.out_jac = dj;
  KD_HIP(c, kd::launch_flow_map(N, flow_params(c), io, flow_grid(c, n), c->stream));
  KD_HIP(c, hipMemcpyAsync(x, dx, sizeof(double) * nl, hipMemcpyDeviceToHost, c->stream));
  KD_HIP(c, hipMemcpyAsync(jac, dj, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  return KD_OK;
}

static int flow_scratch(kd_ctx *c, int blocks, int extra_losses) {
  const size_t np = (size_t)kd::flow_n_params(c->flow_order, c->flow_layers, c->flow_alloc_nc);
  const size_t pb = np * (size_t)blocks * sizeof(float);
  if (pb > c->fpart_cap) {
    dfree(c->d_fpart);
    c->fpart_cap = 0;
    KD_HIP(c, hipMalloc(&c->d_fpart, pb));
    c->fpart_cap = pb;
  }
  const size_t lb = sizeof(double) * ((size_t)blocks + (size_t)extra_losses + 1);
  if (lb > c->floss_cap) {
    dfree(c->d_floss);
    c->floss_cap = 0;
    KD_HIP(c, hipMalloc(&c->d_floss, lb));
    c->floss_cap = lb;
  }
  return KD_OK;
}

// gradient of the weighted loss over points [offset, offset + 64 blocks) of
// (x, w) into d_fgrad; the loss into d_floss[blocks + loss_slot]
static int flow_grad_launch(kd_ctx *c, const double *dx, const double *dw, const double *dc, int64_t n,
                            int64_t offset, int blocks, double inv_norm, int loss_slot) {
  kd::FlowTrainIO io{};
  io.x = dx;
  io.w = dw;
  io.c = c->flow_alloc_nc ? dc : nullptr;
  io.n_points = n;
  io.offset = offset;
  io.inv_norm = inv_norm;
  io.part = c->d_fpart;
  io.loss_part = c->d_floss;
  io.n_params = kd::flow_n_params(c->flow_order, c->flow_layers, c->flow_alloc_nc);
  KD_HIP(c, kd::launch_flow_train(c->flow_order, flow_params(c), io, blocks, c->stream));
  KD_HIP(c, kd::launch_flow_grad_reduce(c->d_fpart, blocks, io.n_params, c->d_fgrad, c->d_floss,
                                        c->d_floss + blocks + loss_slot, c->stream));
  return KD_OK;
}

int kd_flow_loss_grad(kd_ctx *c, const double *x, const double *w, int64_t n, double w_mean, float *grad,
                      double *loss) {
  return kd_flow_loss_grad_cond(c, x, w, nullptr, n, w_mean, grad, loss);
}

int kd_flow_loss_grad_cond(kd_ctx *c, const double *x, const double *w, const double *cond, int64_t n,
                           double w_mean, float *grad, double *loss) {
  if (!c) return KD_ERR_ARG;
  if (!c->d_fw) return fail(c, KD_ERR_STATE, "flow not initialised (kd_flow_init)");
  if (n <= 0 || n > (int64_t(1) << 18)) return fail(c, KD_ERR_ARG, "point count must be in [1, 2^18]");
  if (!x || !w || !grad || !loss) return fail(c, KD_ERR_ARG, "null buffer");
  if (!(w_mean > 0.0)) return fail(c, KD_ERR_ARG, "w_mean must be > 0");
  const int N = c->flow_order;
  const size_t nx = (size_t)n * N;
  KD_HIP(c, hipSetDevice(c->device));
  int rc = ensure_stage(c, sizeof(double) * (nx + 2 * (size_t)n) + 64);
  if (rc) return rc;
  const int blocks = (int)((n + 63) / 64);
  if ((rc = flow_scratch(c, blocks, 0))) return rc;
  double *dx = (double *)c->d_stage, *dw = dx + nx, *dc = dw + n;
  KD_HIP(c, hipMemcpyAsync(dx, x, sizeof(double) * nx, hipMemcpyHostToDevice, c->stream));
  KD_HIP(c, hipMemcpyAsync(dw, w, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  if (cond) KD_HIP(c, hipMemcpyAsync(dc, cond, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  if ((rc = flow_grad_launch(c, dx, dw, cond ? dc : nullptr, n, 0, blocks, 1.0 / (w_mean * (double)n), 0)))
    return rc;
  const size_t np = (size_t)kd::flow_n_params(N, c->flow_layers, c->flow_alloc_nc);
  KD_HIP(c, hipMemcpyAsync(grad, c->d_fgrad, sizeof(float) * np, hipMemcpyDeviceToHost, c->stream));
  KD_HIP(c, hipMemcpyAsync(loss, c->d_floss + blocks, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  return KD_OK;
}

int kd_flow_collect(kd_ctx *c, int64_t n_samples, double weight_power, double *w_mean) {
  if (!c) return KD_ERR_ARG;
  int rc = need_integrand(c);
  if (rc) return rc;
  if (n_samples < 64 || n_samples > (int64_t(1) << 24))
    return fail(c, KD_ERR_ARG, "training sample count must be in [64, 2^24]");
  if (!(weight_power > 0.0) || weight_power > 4.0) return fail(c, KD_ERR_ARG, "weight_power must be in (0, 4]");
  const int N = c->order;
  KD_HIP(c, hipSetDevice(c->device));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  dfree(c->d_tx);
  dfree(c->d_tw);
  dfree(c->d_tc);
  c->n_train = 0;
  KD_HIP(c, hipMalloc(&c->d_tx, sizeof(double) * (size_t)n_samples * N));
  KD_HIP(c, hipMalloc(&c->d_tw, sizeof(double) * (size_t)n_samples));
  double *dvals = nullptr;
  KD_HIP(c, hipMalloc(&dvals, sizeof(double) * 2 * (size_t)n_samples));
  // samples of the chain without the flow warper
  c->flow_suppressed = true;
  rc = run_common(c, (uint64_t)n_samples, kd::MODE_SAMPLE, c->d_tx, dvals);
  c->flow_suppressed = false;
  if (!rc) rc = finish_timing(c);
  const int grid = 1024;
  double *dpart = nullptr;
  if (!rc && hipMalloc(&dpart, sizeof(double) * grid) != hipSuccess) rc = fail(c, KD_ERR_HIP, "hipMalloc failed");
  std::vector<double> part(grid);
  if (!rc && kd::launch_flow_abs(dvals, n_samples, weight_power, c->d_tw, dpart, grid, c->stream) != hipSuccess)
    rc = fail(c, KD_ERR_HIP, "flow_abs launch failed");
  if (!rc && hipMemcpyAsync(part.data(), dpart, sizeof(double) * grid, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    rc = fail(c, KD_ERR_HIP, "copy failed");
  if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, KD_ERR_HIP, "sync failed");
  (void)hipFree(dvals);
  if (dpart) (void)hipFree(dpart);
  if (rc) return rc;
  double s = 0.0;
  for (double v : part) s += v;
  if (!std::isfinite(s)) return fail(c, KD_ERR_ARG, "non-finite integrand samples (g0 window too small?)");
  // an all-zero set is kept (kd_flow_get_samples may still read it); training on it fails
  c->n_train = n_samples;
  c->w_mean = s / (double)n_samples;
  if (w_mean) *w_mean = c->w_mean;
  return KD_OK;
}

int kd_flow_get_samples(kd_ctx *c, double *x, double *w, int64_t n) {
  if (!c) return KD_ERR_ARG;
  if (c->n_train <= 0) return fail(c, KD_ERR_STATE, "no training samples (kd_flow_collect)");
  if (n != c->n_train) return fail(c, KD_ERR_ARG, "sample count differs from the training set's");
  if (!x || !w) return fail(c, KD_ERR_ARG, "null buffer");
  KD_HIP(c, hipSetDevice(c->device));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  KD_HIP(c, hipMemcpy(x, c->d_tx, sizeof(double) * (size_t)n * c->order, hipMemcpyDeviceToHost));
  KD_HIP(c, hipMemcpy(w, c->d_tw, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
  return KD_OK;
}

int kd_flow_set_samples(kd_ctx *c, const double *x, const double *w, const double *cond, int64_t n) {
  if (!c) return KD_ERR_ARG;
  int rc = need_integrand(c);
  if (rc) return rc;
  if (n < 64 || n > (int64_t(1) << 24)) return fail(c, KD_ERR_ARG, "training sample count must be in [64, 2^24]");
  if (!x || !w) return fail(c, KD_ERR_ARG, "null buffer");
  const int N = c->order;
  double s = 0.0;
  for (int64_t i = 0; i < n; i++) {
    if (!(w[i] >= 0.0) || !std::isfinite(w[i])) return fail(c, KD_ERR_ARG, "weights must be finite and >= 0");
    if (cond && !std::isfinite(cond[i])) return fail(c, KD_ERR_ARG, "non-finite conditioning value");
    for (int d = 0; d < N; d++)
      if (!(x[i * N + d] >= 0.0 && x[i * N + d] <= 1.0)) return fail(c, KD_ERR_ARG, "training points must lie in [0, 1]");
    s += w[i];
  }
  if (!(s > 0.0)) return fail(c, KD_ERR_ARG, "all training samples are zero");
  KD_HIP(c, hipSetDevice(c->device));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  dfree(c->d_tx);
  dfree(c->d_tw);
  dfree(c->d_tc);
  c->n_train = 0;
  KD_HIP(c, hipMalloc(&c->d_tx, sizeof(double) * (size_t)n * N));
  KD_HIP(c, hipMalloc(&c->d_tw, sizeof(double) * (size_t)n));
  KD_HIP(c, hipMemcpy(c->d_tx, x, sizeof(double) * (size_t)n * N, hipMemcpyHostToDevice));
  KD_HIP(c, hipMemcpy(c->d_tw, w, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
  if (cond) {
    KD_HIP(c, hipMalloc(&c->d_tc, sizeof(double) * (size_t)n));
    KD_HIP(c, hipMemcpy(c->d_tc, cond, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
  }
  c->n_train = n;
  c->w_mean = s / (double)n;
  return KD_OK;
}

int kd_flow_train(kd_ctx *c, int32_t steps, int32_t batch, float lr, float lr_final, double *losses) {
  if (!c) return KD_ERR_ARG;
  if (!c->d_fw) return fail(c, KD_ERR_STATE, "flow not initialised (kd_flow_init)");
  if (c->n_train <= 0) return fail(c, KD_ERR_STATE, "no training samples (kd_flow_collect)");
  if (!(c->w_mean > 0.0)) return fail(c, KD_ERR_ARG, "all training samples are zero");
  if (c->flow_order != c->order) return fail(c, KD_ERR_STATE, "flow dimension differs from the integrand order");
  if (steps < 0) return fail(c, KD_ERR_ARG, "negative step count");
  if (batch < 64 || batch % 64 || batch > (1 << 16) || c->n_train % batch)
    return fail(c, KD_ERR_ARG, "batch must be a multiple of 64 that divides the training sample count");
  if (!(lr > 0.f) || !(lr_final > 0.f)) return fail(c, KD_ERR_ARG, "learning rates must be > 0");
  if (steps == 0) return KD_OK;
  KD_HIP(c, hipSetDevice(c->device));
  const int blocks = batch / 64;
  int rc = flow_scratch(c, blocks, steps);
  if (rc) return rc;
  const int np = kd::flow_n_params(c->flow_order, c->flow_layers, c->flow_alloc_nc);
  const double inv_norm = 1.0 / (c->w_mean * (double)batch);
  KD_HIP(c, hipEventRecord(c->ev0, c->stream));
  for (int s = 0; s < steps; s++) {
    const int64_t off = ((int64_t)s * batch) % c->n_train;
    if ((rc = flow_grad_launch(c, c->d_tx, c->d_tw, c->d_tc, c->n_train, off, blocks, inv_norm, s))) return rc;
    // cosine decay from lr to lr_final
    const float lrs = lr_final + 0.5f * (lr - lr_final) * (1.f + std::cos(3.14159265358979f * (float)s / (float)steps));
    c->flow_step++;
    KD_HIP(c, kd::launch_flow_adam(c->d_fw, c->d_fm, c->d_fv, c->d_fgrad, np, lrs, 0.9f, 0.999f, 1e-8f,
                                   c->flow_step, c->stream));
    KD_HIP(c, kd::launch_flow_pack(c->d_fw, c->d_fpk, c->flow_order, c->flow_layers, c->flow_alloc_nc, c->stream));
  }
  KD_HIP(c, hipEventRecord(c->ev1, c->stream));
  if (losses)
    KD_HIP(c, hipMemcpyAsync(losses, c->d_floss + blocks, sizeof(double) * steps, hipMemcpyDeviceToHost, c->stream));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  KD_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  c->last_lu_ms = 0.f;
  c->last_prep_ms = 0.f;
  c->last_launches = 4 * steps;
  c->last_chunks = 0;
  return KD_OK;
}

// ---------------------------------------------------------------------------
// launch-shape autotuner and the batched multi-integrator run

int kd_autotune(kd_ctx *c, int64_t sample_points, int32_t *lanes, int32_t *waves_per_cu, double *ms_per_mpoint) {
  if (!c) return KD_ERR_ARG;
  int rc = need_integrand(c);
  if (rc) return rc;
  if (!c->has_sobol) return fail(c, KD_ERR_STATE, "rng not set (kd_set_sobol)");
  if (c->sobol_dim != c->order) return fail(c, KD_ERR_ARG, "Sobol dimension differs from the integrand order");
  if (sample_points < 64 || sample_points > kChunk) return fail(c, KD_ERR_ARG, "sample size must be in [64, 2^21]");
  if ((uint64_t)c->k + (uint64_t)sample_points > (1ull << 31))
    return fail(c, KD_ERR_RANGE, "More sample points requested then generator can hold.");
  KD_HIP(c, hipSetDevice(c->device));
  const int lane_opts[3] = {64, 32, 16}, wave_opts[4] = {16, 32, 64, 128};
  const int old_l = c->lu_lanes_max, old_w = c->lu_waves_per_cu, old_a = c->algo;
  int best_l = old_l, best_w = old_w, best_a = old_a;
  float best = 1e30f;
  kd::KdParams p = base_params(c);
  p.mode = kd::MODE_RUN;
  p.n_points = sample_points;
  p.nbins = 0; // no bin counts from tuning launches
  c->no_reduce = true;
  // candidates: the split tree + per-configuration LU (orders 6..10), the
  // paired LU (orders >= 7), the tree (orders >= 6; these at 64 lanes) and the
  // per-configuration kernel with 64, 32 or 16 lanes per point, each at 16, 32,
  // 64 or 128 waves per CU
  const bool direct = c->obs == KD_OBS_DIRECT;
  const int li0 = !direct                           ? 0
                  : kd::ctree_supported(c->order)    ? -5
                  : kd::wsplit_supported(c->order)   ? -4
                  : kd::split_threads(c->order) > 0 ? -3
                  : c->order >= 7                   ? -2
                  : c->order >= 6                   ? -1
                                                    : 0;
  for (int li = li0; li < (direct ? 3 : 1) && !rc; li++) {
    if (li == -4 && !kd::wsplit_supported(c->order)) continue;
    if (li == -3 && kd::split_threads(c->order) == 0) continue;
    if (li == -2 && c->order < 7) continue;
    const int L = li < 0 ? 64 : lane_opts[li];
    if (li > 0 && kd::lu_lanes(c->order, L) == kd::lu_lanes(c->order, lane_opts[li - 1])) continue;
    for (int wi = 0; wi < 4 && !rc; wi++) {
      c->algo = li == -5   ? KD_ALGO_CTREE
                : li == -4 ? KD_ALGO_WSPLIT
                : li == -3 ? KD_ALGO_SPLIT
                : li == -2 ? KD_ALGO_PAIR
                : li == -1 ? KD_ALGO_TREE
                           : KD_ALGO_CONFIG;
      c->lu_lanes_max = L;
      c->lu_waves_per_cu = wave_opts[wi];
      float t = 1e30f;
      for (int rep = 0; rep < 3 && !rc; rep++) { // first repetition warms up
        if ((rc = launch_chunks(c, p, KD_TIMING_KERNELS)) || (rc = finish_timing(c))) break;
        if (rep > 0) t = std::min(t, c->last_lu_ms);
      }
      if (!rc && t < best) {
        best = t;
        best_l = L;
        best_w = wave_opts[wi];
        best_a = c->algo;
      }
    }
  }
  c->no_reduce = false;
  if (rc) {
    c->lu_lanes_max = old_l;
    c->lu_waves_per_cu = old_w;
    c->algo = old_a;
    return rc;
  }
  c->lu_lanes_max = best_l;
  c->lu_waves_per_cu = best_w;
  c->algo = best_a;
  c->tune_ms_per_mpt = best * 1048576.f / (float)sample_points;
  if (lanes) *lanes = use_tree(c) || use_split(c) || use_wsplit(c) || use_ctree(c) ? 0 : kd::lu_lanes(c->order, best_l);
  if (waves_per_cu) *waves_per_cu = best_w;
  if (ms_per_mpoint) *ms_per_mpoint = c->tune_ms_per_mpt;
  return KD_OK;
}

int kd_set_launch_shape(kd_ctx *c, int32_t lanes, int32_t waves_per_cu) {
  if (!c) return KD_ERR_ARG;
  if (lanes != 16 && lanes != 32 && lanes != 64) return fail(c, KD_ERR_ARG, "lanes must be 16, 32 or 64");
  if (waves_per_cu < 1 || waves_per_cu > 128) return fail(c, KD_ERR_ARG, "waves_per_cu must be in [1, 128]");
  c->lu_lanes_max = lanes;
  c->lu_waves_per_cu = waves_per_cu;
  return KD_OK;
}

int kd_set_algorithm(kd_ctx *c, int32_t algo) {
  if (!c) return KD_ERR_ARG;
  if (algo != KD_ALGO_DEFAULT && algo != KD_ALGO_CONFIG && algo != KD_ALGO_TREE && algo != KD_ALGO_PAIR &&
      algo != KD_ALGO_SPLIT && algo != KD_ALGO_WSPLIT && algo != KD_ALGO_CTREE)
    return fail(c, KD_ERR_ARG, "unknown Keldysh-sum algorithm");
  c->algo = algo;
  return KD_OK;
}

int kd_get_algorithm(const kd_ctx *c, int32_t *algo) {
  if (!c || !algo) return KD_ERR_ARG;
  *algo = use_tree(c)     ? KD_ALGO_TREE
          : use_pair(c)   ? KD_ALGO_PAIR
          : use_split(c)  ? KD_ALGO_SPLIT
          : use_ctree(c)  ? KD_ALGO_CTREE
          : use_wsplit(c) ? KD_ALGO_WSPLIT
                          : KD_ALGO_CONFIG;
  return KD_OK;
}

int kd_get_launch_shape(const kd_ctx *c, int32_t *lanes, int32_t *waves_per_cu) {
  if (!c) return KD_ERR_ARG;
  if (lanes) *lanes = kd::lu_lanes(c->order > 0 ? c->order : 0, c->lu_lanes_max);
  if (waves_per_cu) *waves_per_cu = grid_waves_per_cu(c);
  return KD_OK;
}

int kd_run_batch(kd_ctx *const *ctxs, const uint64_t *n_steps, int32_t count, kd_accum *out, double *ms) {
  if (!ctxs || !n_steps || count < 1) return KD_ERR_ARG;
  kd_ctx *c0 = ctxs[0];
  if (!c0) return KD_ERR_ARG;
  for (int i = 0; i < count; i++) {
    kd_ctx *c = ctxs[i];
    if (!c) return fail(c0, KD_ERR_ARG, "null context in batch");
    for (int j = 0; j < i; j++)
      if (ctxs[j] == c) return fail(c0, KD_ERR_ARG, "a context appears twice in the batch");
    if (c->device != c0->device) return fail(c0, KD_ERR_ARG, "batched contexts must share one device");
    int rc = need_integrand(c);
    if (rc) return fail(c0, rc, c->err);
    if (!c->has_sobol) return fail(c0, KD_ERR_STATE, "rng not set (kd_set_sobol)");
    if (c->sobol_dim != c->order) return fail(c0, KD_ERR_ARG, "Sobol dimension differs from the integrand order");
    if ((uint64_t)c->k + n_steps[i] > (1ull << 31))
      return fail(c0, KD_ERR_RANGE, "More sample points requested then generator can hold.");
  }
  KD_HIP(c0, hipSetDevice(c0->device));
  std::vector<kd::KdParams> ps(count);
  for (int i = 0; i < count; i++) {
    kd_ctx *c = ctxs[i];
    ps[i] = base_params(c);
    ps[i].mode = kd::MODE_RUN;
    ps[i].n_points = (int64_t)n_steps[i];
    if (n_steps[i] == 0) continue;
    int rc = prepare_launch(c, ps[i]);
    if (rc) return fail(c0, rc, c->err);
    KD_HIP(c0, hipStreamSynchronize(c->stream));
  }
  // one graph: fork from ctx 0's stream to every ctx stream, join back
  std::vector<hipEvent_t> evs(count + 1, nullptr);
  for (auto &e : evs) KD_HIP(c0, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int rc = KD_OK;
  hipStream_t origin = c0->stream;
  if (hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    rc = fail(c0, KD_ERR_HIP, "hipStreamBeginCapture failed");
  } else {
    if (hipEventRecord(evs[count], origin) != hipSuccess) rc = fail(c0, KD_ERR_HIP, "hipEventRecord failed");
    for (int i = 0; i < count && !rc; i++) {
      kd_ctx *c = ctxs[i];
      if (n_steps[i] == 0) continue;
      if (i > 0 && hipStreamWaitEvent(c->stream, evs[count], 0) != hipSuccess) {
        rc = fail(c0, KD_ERR_HIP, "hipStreamWaitEvent failed");
        break;
      }
      if ((rc = launch_chunks(c, ps[i], KD_TIMING_NONE))) {
        rc = fail(c0, rc, c->err);
        break;
      }
      if (i > 0 && (hipEventRecord(evs[i], c->stream) != hipSuccess ||
                    hipStreamWaitEvent(origin, evs[i], 0) != hipSuccess))
        rc = fail(c0, KD_ERR_HIP, "join failed");
    }
    if (hipStreamEndCapture(origin, &graph) != hipSuccess && !rc) rc = fail(c0, KD_ERR_HIP, "hipStreamEndCapture failed");
  }
  if (!rc && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
    rc = fail(c0, KD_ERR_HIP, "hipGraphInstantiate failed");
  float t_ms = 0.f;
  if (!rc) {
    if (hipEventRecord(c0->ev0, origin) != hipSuccess || hipGraphLaunch(exec, origin) != hipSuccess ||
        hipEventRecord(c0->ev1, origin) != hipSuccess || hipStreamSynchronize(origin) != hipSuccess ||
        hipEventElapsedTime(&t_ms, c0->ev0, c0->ev1) != hipSuccess)
      rc = fail(c0, KD_ERR_HIP, "graph launch failed");
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (auto &e : evs) (void)hipEventDestroy(e);
  if (rc) return rc;
  for (int i = 0; i < count; i++) {
    kd_ctx *c = ctxs[i];
    c->k += (uint32_t)n_steps[i];
    c->n_points += n_steps[i];
    c->last_ms = t_ms;
    c->last_lu_ms = c->last_prep_ms = 0.f;
    if (out && (rc = kd_get_accum(c, &out[i]))) return fail(c0, rc, c->err);
  }
  if (ms) *ms = t_ms;
  return KD_OK;
}

// ---------------------------------------------------------------------------
// time binning (integrand_g_direct_time + binner_t<1>)

int kd_set_time_bins(kd_ctx *c, double xmin, double xmax, int32_t nbins) {
  if (!c) return KD_ERR_ARG;
  if (c->obs == KD_OBS_KERNEL) return fail(c, KD_ERR_STATE, "the kernel method bins with kd_set_kernel_bins");
  // the per-workgroup histogram lives in LDS next to the Wick matrices
  if (nbins < 0 || nbins > 3072) return fail(c, KD_ERR_ARG, "number of bins must be in [0, 3072]");
  return set_bins(c, xmin, xmax, nbins, 1);
}

int kd_get_time_bins(kd_ctx *c, double *values, uint64_t *counts, uint64_t *dropped) {
  if (!c) return KD_ERR_ARG;
  if (c->nbins <= 0) return fail(c, KD_ERR_STATE, "time bins not set (kd_set_time_bins)");
  KD_HIP(c, hipSetDevice(c->device));
  std::vector<unsigned long long> cnt(c->nbins + 1);
  if (values)
    KD_HIP(c, hipMemcpyAsync(values, c->d_bins, sizeof(double2) * c->nbins, hipMemcpyDeviceToHost, c->stream));
  KD_HIP(c, hipMemcpyAsync(cnt.data(), c->d_bin_counts, sizeof(unsigned long long) * (c->nbins + 1),
                           hipMemcpyDeviceToHost, c->stream));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  if (counts)
    for (int i = 0; i < c->nbins; i++) counts[i] = cnt[i];
  if (dropped) *dropped = cnt[c->nbins];
  return KD_OK;
}

// ---------------------------------------------------------------------------
// pilot projection on the device (make_warper_from_proj.hpp:80-114)

int kd_pilot_histograms(kd_ctx *c, const uint32_t *Csr, const uint32_t *shfr, int64_t n_samples, int32_t num_bins,
                        int32_t keep_u_hypercube, double *sums, uint64_t *counts) {
  if (!c) return KD_ERR_ARG;
  int rc = need_integrand(c);
  if (rc) return rc;
  if (!Csr || !shfr || !sums || !counts) return fail(c, KD_ERR_ARG, "null buffer");
  if (c->order < 1) return fail(c, KD_ERR_ARG, "Dimension of the integration space must be > 0.");
  if (num_bins < 2 || num_bins > 65536) return fail(c, KD_ERR_ARG, "num_bins must be in [2, 65536]");
  if (c->obs == KD_OBS_KERNEL)
    return fail(c, KD_ERR_STATE, "per-point samples need a scalar integrand (not the binned kernel method)");
  if (n_samples < 1 || n_samples > (int64_t(1) << 24)) return fail(c, KD_ERR_ARG, "pilot sample count must be in [1, 2^24]");
  const int N = c->order;
  KD_HIP(c, hipSetDevice(c->device));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  // the pilot generator replaces the ctx's for this call; its tables are restored after
  std::vector<uint32_t> sCsr, sShfr;
  const bool had = c->has_sobol;
  const int sdim = c->sobol_dim;
  const uint32_t sk = c->k;
  if (had) {
    sCsr.resize(31 * (size_t)sdim);
    sShfr.resize(sdim);
    KD_HIP(c, hipMemcpy(sCsr.data(), c->d_Csr, sizeof(uint32_t) * sCsr.size(), hipMemcpyDeviceToHost));
    KD_HIP(c, hipMemcpy(sShfr.data(), c->d_shfr, sizeof(uint32_t) * sShfr.size(), hipMemcpyDeviceToHost));
  }
  if ((rc = kd_set_sobol(c, N, Csr, shfr, 0))) return rc;
  double *dl = nullptr, *dv = nullptr, *ds = nullptr;
  unsigned long long *dc = nullptr;
  const size_t nb = (size_t)N * num_bins;
  if (hipMalloc(&dl, sizeof(double) * (size_t)n_samples * N) != hipSuccess ||
      hipMalloc(&dv, sizeof(double) * 2 * (size_t)n_samples) != hipSuccess ||
      hipMalloc(&ds, sizeof(double) * nb) != hipSuccess || hipMalloc(&dc, sizeof(unsigned long long) * nb) != hipSuccess)
    rc = fail(c, KD_ERR_HIP, "hipMalloc failed");
  if (!rc) {
    // warped integrand at the pilot points (the flow, if last, is part of the chain)
    kd::KdParams p = base_params(c);
    p.mode = kd::MODE_SAMPLE;
    p.keep_u = keep_u_hypercube ? 1 : 0;
    p.n_points = n_samples;
    p.out_l = dl;
    p.out_vals = dv;
    rc = launch_timed(c, p, 0);
    if (!rc) rc = finish_timing(c);
  }
  if (!rc && kd::launch_pilot_hist(dl, dv, n_samples, N, num_bins, ds, dc, c->stream) != hipSuccess)
    rc = fail(c, KD_ERR_HIP, "pilot histogram launch failed");
  if (!rc && (hipMemcpyAsync(sums, ds, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
              hipMemcpyAsync(counts, dc, sizeof(uint64_t) * nb, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
              hipStreamSynchronize(c->stream) != hipSuccess))
    rc = fail(c, KD_ERR_HIP, "copy failed");
  for (void *q : {(void *)dl, (void *)dv, (void *)ds, (void *)dc})
    if (q) (void)hipFree(q);
  // restore the ctx's generator
  if (had) {
    const int r2 = kd_set_sobol(c, sdim, sCsr.data(), sShfr.data(), sk);
    if (!rc) rc = r2;
  } else {
    c->has_sobol = false;
    c->k = 0;
  }
  if (!rc) {
    for (size_t i = 0; i < nb; i++)
      if (!std::isfinite(sums[i])) return fail(c, KD_ERR_ARG, "g0 evaluated outside its time window [-time_max, time_max]");
  }
  return rc;
}

} // extern "C"

// ---------------------------------------------------------------------------
// kernel method (wick_kernel.cpp:42-138, wick_kernel_single_omega.cpp:41-140)

extern "C" {

static int set_kernel(kd_ctx *c, int32_t order, const kd_gf_index *X, int obs, double omega) {
  if (!c) return KD_ERR_ARG;
  if (!X) return fail(c, KD_ERR_ARG, "null external point");
  if (order < 1 || order > KD_MAX_ORDER)
    return fail(c, KD_ERR_ARG, "order must be in [1, " + std::to_string(KD_MAX_ORDER) + "]");
  if (!c->has_g0) return fail(c, KD_ERR_STATE, "g0 table not set (kd_set_g0_table)");
  if (X->orbital < 0 || X->orbital > 1 || (X->orbital != 0 && c->n_orb < 2))
    return fail(c, KD_ERR_ARG, "orbital must be 0, or 1 with a 2-orbital table");
  if (X->spin < 0 || X->spin > 1 || X->k_idx < 0 || X->k_idx > 1) return fail(c, KD_ERR_ARG, "bad external point");
  if (!std::isfinite(omega)) return fail(c, KD_ERR_ARG, "omega must be finite");
  int rc;
  if (c->nbins > 0 && (rc = rc_bins_clear(c))) return rc;
  c->order = order;
  c->A = *X;
  c->B = *X;
  c->cutoff = 0.0;
  c->obs = obs;
  c->omega = omega;
  return KD_OK;
}

int kd_set_kernel_integrand(kd_ctx *c, int32_t order, const kd_gf_index *X) {
  return set_kernel(c, order, X, KD_OBS_KERNEL, 0.0);
}

int kd_set_kernel_omega_integrand(kd_ctx *c, int32_t order, const kd_gf_index *X, double omega) {
  return set_kernel(c, order, X, KD_OBS_KERNEL_OMEGA, omega);
}

int kd_set_kernel_bins(kd_ctx *c, double xmin, double xmax, int32_t nbins) {
  if (!c) return KD_ERR_ARG;
  if (c->obs != KD_OBS_KERNEL) return fail(c, KD_ERR_STATE, "kernel integrand not set (kd_set_kernel_integrand)");
  if (nbins < 0 || nbins > 1536) return fail(c, KD_ERR_ARG, "number of bins must be in [0, 1536]");
  return set_bins(c, xmin, xmax, nbins, 2);
}

int kd_get_kernel_bins(kd_ctx *c, double *values, uint64_t *counts, uint64_t *dropped) {
  if (!c) return KD_ERR_ARG;
  if (c->obs != KD_OBS_KERNEL || c->nbins <= 0) return fail(c, KD_ERR_STATE, "kernel bins not set (kd_set_kernel_bins)");
  KD_HIP(c, hipSetDevice(c->device));
  std::vector<unsigned long long> cnt(c->nbins + 1);
  if (values)
    KD_HIP(c, hipMemcpyAsync(values, c->d_bins, sizeof(double2) * 2 * c->nbins, hipMemcpyDeviceToHost, c->stream));
  KD_HIP(c, hipMemcpyAsync(cnt.data(), c->d_bin_counts, sizeof(unsigned long long) * (c->nbins + 1),
                           hipMemcpyDeviceToHost, c->stream));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  if (counts)
    for (int i = 0; i < c->nbins; i++) counts[2 * i] = counts[2 * i + 1] = cnt[i];
  if (dropped) *dropped = cnt[c->nbins];
  return KD_OK;
}

int kd_eval_kernel_batch(kd_ctx *c, const double *pts, int64_t n_points, int32_t warped, int32_t keep_u_hypercube,
                         double *out, int32_t *in_domain, double *jacobian) {
  if (!c) return KD_ERR_ARG;
  int rc = need_integrand(c);
  if (rc) return rc;
  if (c->obs != KD_OBS_KERNEL) return fail(c, KD_ERR_STATE, "kernel integrand not set (kd_set_kernel_integrand)");
  if (n_points <= 0) return n_points == 0 ? KD_OK : fail(c, KD_ERR_ARG, "negative point count");
  if (!pts || !out) return fail(c, KD_ERR_ARG, "null buffer");
  const int64_t vpp = vals_per_point(c);
  if (n_points > (int64_t(1) << 26) / vpp) return fail(c, KD_ERR_ARG, "too many entries for one call");
  const size_t np = (size_t)n_points * (size_t)c->order;
  const size_t nv = 2 * (size_t)n_points * (size_t)vpp;
  const size_t bytes = sizeof(double) * (np + nv + (size_t)n_points) + sizeof(int32_t) * (size_t)n_points + 64;
  KD_HIP(c, hipSetDevice(c->device));
  if ((rc = ensure_stage(c, bytes))) return rc;
  double *dp = (double *)c->d_stage;
  double *dout = dp + np;
  double *djac = dout + nv;
  int32_t *ddom = (int32_t *)(djac + n_points);
  KD_HIP(c, hipMemcpyAsync(dp, pts, sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
  KD_HIP(c, hipMemsetAsync(dout, 0, sizeof(double) * nv, c->stream));
  kd::KdParams p = base_params(c);
  p.mode = warped ? kd::MODE_EVAL_L : kd::MODE_EVAL_U;
  p.keep_u = keep_u_hypercube ? 1 : 0;
  p.in_pts = dp;
  p.n_points = n_points;
  p.out_vals = dout;
  p.out_dom = ddom;
  p.out_jac = djac;
  if ((rc = launch_timed(c, p, 0))) return rc;
  KD_HIP(c, hipMemcpyAsync(out, dout, sizeof(double) * nv, hipMemcpyDeviceToHost, c->stream));
  if (in_domain)
    KD_HIP(c, hipMemcpyAsync(in_domain, ddom, sizeof(int32_t) * n_points, hipMemcpyDeviceToHost, c->stream));
  if (jacobian)
    KD_HIP(c, hipMemcpyAsync(jacobian, djac, sizeof(double) * n_points, hipMemcpyDeviceToHost, c->stream));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  return finish_timing(c);
}

} // extern "C"
