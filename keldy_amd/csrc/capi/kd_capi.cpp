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
  // warper chain
  kd_warper_desc *d_warpers = nullptr;
  kd_map1d *d_maps = nullptr;
  double *d_tables = nullptr;
  int n_warpers = 0;
  std::vector<kd_warper_desc> h_warpers;
  // flow parameters (KD_WARP_FLOW)
  float *d_flow = nullptr;
  int flow_layers = 0, flow_hidden = 0;
  // Sobol
  uint32_t *d_Csr = nullptr, *d_shfr = nullptr;
  int sobol_dim = 0;
  uint32_t k = 0;
  bool has_sobol = false;
  // accumulators
  double *d_acc = nullptr; // re, im, n_dom, n_fallback
  double *d_part = nullptr;
  int part_cap = 0;
  uint64_t n_points = 0;
  // staging for host-memory batch calls
  void *d_stage = nullptr;
  size_t stage_cap = 0;
  // compacted points between the two kernel stages
  void *d_prep = nullptr;
  size_t prep_cap = 0;
  float last_ms = 0.f, last_lu_ms = 0.f, last_prep_ms = 0.f;
  int last_launches = 0, last_chunks = 0;
};

namespace {

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

int grid_for(const kd_ctx *c, int64_t n_points) {
  int64_t nb = (n_points + 63) / 64;
  int64_t cap = (int64_t)(c->n_cu > 0 ? c->n_cu : 256) * 16;
  if (nb > cap) nb = cap;
  return (int)(nb < 1 ? 1 : nb);
}

int ensure_partials(kd_ctx *c, int grid) {
  if (grid <= c->part_cap) return KD_OK;
  dfree(c->d_part);
  c->part_cap = 0;
  KD_HIP(c, hipMalloc(&c->d_part, sizeof(double) * 4 * (size_t)grid));
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
  p.alpha = c->alpha;
  p.A = kd::KdGf{c->A.time, c->A.k_idx, c->A.spin, c->A.orbital};
  p.B = kd::KdGf{c->B.time, c->B.k_idx, c->B.spin, c->B.orbital};
  p.cutoff = c->cutoff;
  p.keep_u = 1;
  p.warpers = c->d_warpers;
  p.maps = c->d_maps;
  p.tables = c->d_tables;
  p.n_warpers = c->n_warpers;
  p.Csr = c->d_Csr;
  p.shfr = c->d_shfr;
  p.k0 = c->k;
  p.flow = c->d_flow;
  p.flow_layers = c->flow_layers;
  p.flow_hidden = c->flow_hidden;
  return p;
}

int need_integrand(kd_ctx *c) {
  if (!c->has_g0) return fail(c, KD_ERR_STATE, "g0 table not set (kd_set_g0_table)");
  if (c->order < 0) return fail(c, KD_ERR_STATE, "integrand not set (kd_set_integrand)");
  return KD_OK;
}

// Points per launch pair (prep + integrand); bounds the compacted-point buffer.
constexpr int64_t kChunk = int64_t(1) << 21;

int ensure_prep(kd_ctx *c, int order, kd::PrepBuf *buf) {
  const size_t n = (size_t)kChunk;
  const size_t bytes = n * (sizeof(double) * (size_t)(order > 0 ? order : 1) + sizeof(double) + sizeof(int)) +
                       (n / 64) * sizeof(int) + 256;
  if (bytes > c->prep_cap) {
    dfree(c->d_prep);
    c->prep_cap = 0;
    KD_HIP(c, hipMalloc(&c->d_prep, bytes));
    c->prep_cap = bytes;
  }
  char *base = (char *)c->d_prep;
  buf->u = (double *)base;
  base += n * sizeof(double) * (size_t)(order > 0 ? order : 1);
  buf->jac = (double *)base;
  base += n * sizeof(double);
  buf->idx = (int *)base;
  base += n * sizeof(int);
  buf->cnt = (int *)base;
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

// Launch prep + integrand kernels over p.n_points points in chunks (and, in
// MODE_RUN, the deterministic partial reduction into d_acc).  Each kernel is
// bracketed by HIP events on the ctx stream (kd_last_timing).
int launch_timed(kd_ctx *c, const kd::KdParams &p0, int /*unused*/) {
  kd::PrepBuf buf;
  int rc = ensure_prep(c, c->order, &buf);
  if (rc) return rc;
  const int N = c->order;
  const int64_t nchunks = (p0.n_points + kChunk - 1) / kChunk;
  if ((rc = ensure_events(c, size_t(4 * nchunks)))) return rc;
  int launches = 0;
  KD_HIP(c, hipEventRecord(c->ev0, c->stream));
  int64_t ci = 0;
  for (int64_t off = 0; off < p0.n_points; off += kChunk, ci++) {
    kd::KdParams p = p0;
    const int64_t m = std::min<int64_t>(kChunk, p0.n_points - off);
    p.n_points = m;
    p.k0 = p0.k0 + (uint32_t)off;
    if (p.in_pts) p.in_pts = p0.in_pts + off * N;
    if (p.out_vals) p.out_vals = p0.out_vals + 2 * off;
    if (p.out_dom) p.out_dom = p0.out_dom + off;
    if (p.out_jac) p.out_jac = p0.out_jac + off;
    if (p.out_l) p.out_l = p0.out_l + off * N;
    const int grid = grid_for(c, m);
    const int64_t nb = (m + 255) / 256;
    const int grid_prep = (int)std::min<int64_t>(nb, (int64_t)(c->n_cu > 0 ? c->n_cu : 256) * 8);
    if (p.mode == kd::MODE_RUN) {
      if ((rc = ensure_partials(c, grid))) return rc;
      p.partials = c->d_part;
    }
    KD_HIP(c, hipEventRecord(c->ev[4 * ci + 0], c->stream));
    KD_HIP(c, kd::launch_prep(N, p, buf, grid_prep, c->stream));
    KD_HIP(c, hipEventRecord(c->ev[4 * ci + 1], c->stream));
    KD_HIP(c, hipEventRecord(c->ev[4 * ci + 2], c->stream));
    KD_HIP(c, kd::launch_lu(N, c->precision, p, buf, grid, c->stream));
    KD_HIP(c, hipEventRecord(c->ev[4 * ci + 3], c->stream));
    launches += 2;
    if (p.mode == kd::MODE_RUN) {
      KD_HIP(c, kd::launch_reduce(c->d_part, grid, c->d_acc, c->stream));
      launches += 1;
    }
  }
  KD_HIP(c, hipEventRecord(c->ev1, c->stream));
  c->last_launches = launches;
  c->last_chunks = (int)nchunks;
  return KD_OK;
}

int finish_timing(kd_ctx *c) {
  KD_HIP(c, hipEventSynchronize(c->ev1));
  float ms = 0.f, lu = 0.f, prep = 0.f, x = 0.f;
  KD_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  for (int i = 0; i < c->last_chunks; i++) {
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
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipMalloc(&c->d_acc, 4 * sizeof(double)) != hipSuccess ||
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
  dfree(c->d_flow);
  dfree(c->d_Csr);
  dfree(c->d_shfr);
  dfree(c->d_acc);
  dfree(c->d_part);
  dfree(c->d_stage);
  dfree(c->d_prep);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
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
  c->order = order;
  c->A = *A;
  c->B = *B;
  c->cutoff = cutoff;
  return KD_OK;
}

int kd_set_precision(kd_ctx *c, int32_t precision) {
  if (!c) return KD_ERR_ARG;
  if (precision != KD_PREC_FP64 && precision != KD_PREC_FP32_LU)
    return fail(c, KD_ERR_ARG, "unknown precision");
  c->precision = precision;
  return KD_OK;
}

int kd_eval_batch_device(kd_ctx *c, const double *d_u, int64_t n_points, int32_t keep_u_hypercube,
                         double *d_out, int32_t *d_in_domain) {
  if (!c) return KD_ERR_ARG;
  int rc = need_integrand(c);
  if (rc) return rc;
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

int kd_eval_batch(kd_ctx *c, const double *u, int64_t n_points, int32_t keep_u_hypercube, double *out,
                  int32_t *in_domain) {
  if (!c) return KD_ERR_ARG;
  int rc = need_integrand(c);
  if (rc) return rc;
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

int kd_set_warper_chain(kd_ctx *c, const kd_warper_desc *w, int32_t n_w, const kd_map1d *maps,
                        int32_t n_maps, const double *tables, int64_t n_tables) {
  if (!c) return KD_ERR_ARG;
  if (n_w < 0 || n_maps < 0 || n_tables < 0) return fail(c, KD_ERR_ARG, "negative size");
  if ((n_w && !w) || (n_maps && !maps) || (n_tables && !tables)) return fail(c, KD_ERR_ARG, "null buffer");
  for (int i = 0; i < n_w; i++) {
    const kd_warper_desc &d = w[i];
    if (d.kind == KD_WARP_PRODUCT) {
      if (d.n_maps < 1 || d.map_offset < 0 || d.map_offset + d.n_maps > n_maps)
        return fail(c, KD_ERR_ARG, "product warper map range out of bounds");
    } else if (d.kind == KD_WARP_FLOW) {
      if (!c->d_flow) return fail(c, KD_ERR_STATE, "flow warper used before its parameters were set");
    } else if (d.kind != KD_WARP_IDENTITY && d.kind != KD_WARP_PLASMA) {
      return fail(c, KD_ERR_ARG, "unknown warper kind " + std::to_string(d.kind));
    }
  }
  for (int i = 0; i < n_maps; i++) {
    const kd_map1d &m = maps[i];
    if (m.kind < KD_MAP_UNIFORM || m.kind > KD_MAP_HYBRID) return fail(c, KD_ERR_ARG, "unknown map kind");
    if (m.kind == KD_MAP_NEAREST || m.kind == KD_MAP_HYBRID) {
      if (m.n_pts < 2 || m.n_inv < 2 || m.n_inv > m.n_pts) return fail(c, KD_ERR_ARG, "bad knot counts");
      auto in = [&](int64_t off, int64_t len) { return off >= 0 && off + len <= n_tables; };
      if (!in(m.t_off, m.n_pts) || !in(m.F_off, m.n_pts)) return fail(c, KD_ERR_ARG, "map table out of bounds");
      if (m.kind == KD_MAP_NEAREST && !in(m.f_off, m.n_pts)) return fail(c, KD_ERR_ARG, "map table out of bounds");
      if (m.kind == KD_MAP_HYBRID && !in(m.c_off, 4 * (int64_t)(m.n_inv - 1)))
        return fail(c, KD_ERR_ARG, "map table out of bounds");
    }
  }
  KD_HIP(c, hipSetDevice(c->device));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  dfree(c->d_warpers);
  dfree(c->d_maps);
  dfree(c->d_tables);
  if (n_w) {
    KD_HIP(c, hipMalloc(&c->d_warpers, sizeof(kd_warper_desc) * n_w));
    KD_HIP(c, hipMemcpy(c->d_warpers, w, sizeof(kd_warper_desc) * n_w, hipMemcpyHostToDevice));
  }
  if (n_maps) {
    KD_HIP(c, hipMalloc(&c->d_maps, sizeof(kd_map1d) * n_maps));
    KD_HIP(c, hipMemcpy(c->d_maps, maps, sizeof(kd_map1d) * n_maps, hipMemcpyHostToDevice));
  }
  if (n_tables) {
    KD_HIP(c, hipMalloc(&c->d_tables, sizeof(double) * n_tables));
    KD_HIP(c, hipMemcpy(c->d_tables, tables, sizeof(double) * n_tables, hipMemcpyHostToDevice));
  }
  c->n_warpers = n_w;
  c->h_warpers.assign(w, w + n_w);
  return KD_OK;
}

int kd_eval_warped_batch(kd_ctx *c, const double *l, int64_t n_points, int32_t keep_u_hypercube, double *out,
                         int32_t *in_domain, double *jacobian) {
  if (!c) return KD_ERR_ARG;
  int rc = need_integrand(c);
  if (rc) return rc;
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

int kd_set_sobol(kd_ctx *c, int32_t dim, const uint32_t *Csr, const uint32_t *shfr, uint32_t k) {
  if (!c) return KD_ERR_ARG;
  if (dim <= 0) return fail(c, KD_ERR_ARG, "Dimension of the integration space must be > 0.");
  if (dim > 32) return fail(c, KD_ERR_ARG, "Sobol dimension must be <= 32");
  if (!Csr || !shfr) return fail(c, KD_ERR_ARG, "null Sobol table");
  KD_HIP(c, hipSetDevice(c->device));
  KD_HIP(c, hipStreamSynchronize(c->stream));
  dfree(c->d_Csr);
  dfree(c->d_shfr);
  KD_HIP(c, hipMalloc(&c->d_Csr, sizeof(uint32_t) * 31 * dim));
  KD_HIP(c, hipMalloc(&c->d_shfr, sizeof(uint32_t) * dim));
  KD_HIP(c, hipMemcpy(c->d_Csr, Csr, sizeof(uint32_t) * 31 * dim, hipMemcpyHostToDevice));
  KD_HIP(c, hipMemcpy(c->d_shfr, shfr, sizeof(uint32_t) * dim, hipMemcpyHostToDevice));
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
  if (n_steps && (rc = finish_timing(c))) return rc;
  if (out) return kd_get_accum(c, out);
  return KD_OK;
}

int kd_reset_accum(kd_ctx *c) {
  if (!c) return KD_ERR_ARG;
  KD_HIP(c, hipSetDevice(c->device));
  KD_HIP(c, hipMemsetAsync(c->d_acc, 0, 4 * sizeof(double), c->stream));
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

int kd_last_timing(const kd_ctx *c, double *total_ms, double *lu_ms, double *prep_ms, int32_t *launches) {
  if (!c) return KD_ERR_ARG;
  if (total_ms) *total_ms = c->last_ms;
  if (lu_ms) *lu_ms = c->last_lu_ms;
  if (prep_ms) *prep_ms = c->last_prep_ms;
  if (launches) *launches = c->last_launches;
  return KD_OK;
}

int kd_flops_per_point(const kd_ctx *c, double *flops) {
  if (!c || !flops) return KD_ERR_ARG;
  const double N = c->order < 0 ? 0 : c->order;
  // per configuration: row-wise bordered LU (complex MAC = 8, complex mul = 6,
  // complex reciprocal = 7), det^2 * s (12) and the signed FP64 sum (2)
  const double macs = N * (N + 1) * (2 * N + 1) / 6.0;
  const double muls = N * (N + 1) / 2.0 + N;
  const double per_cfg = 8 * macs + 6 * muls + 7 * N + 14;
  *flops = std::ldexp(per_cfg, (int)N);
  return KD_OK;
}

} // extern "C"
