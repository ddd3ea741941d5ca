// kd_kernels.hpp -- launchers of the device kernels (internal to libkeldy_amd).
#pragma once

#include <hip/hip_runtime.h>

#include "../../../include/keldy_amd.h"
#include "kd_device.hpp"
#include "kd_flow_device.hpp"

namespace kd {

// prep_kernel<order>: grid blocks of 256 threads over p.n_points points.
hipError_t launch_prep(int order, const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st);
// LU precision actually used: FP32 for KD_PREC_FP32_LU below order 10 and
// for KD_PREC_FP32_LU_ALL; FP64 otherwise.
inline bool lu_is_fp32(int order, int precision) {
  return precision == KD_PREC_FP32_LU_ALL || (precision == KD_PREC_FP32_LU && order < 10);
}
// lanes per point of the Keldysh-sum kernel for a tuned maximum
inline int lu_lanes(int order, int lanes_max) {
  const int nc = 1 << order;
  return nc < lanes_max ? nc : lanes_max;
}
// integrand_kernel<order, T, lanes> (the Keldysh sum): grid blocks of 64
// threads; lanes = the maximum lanes per point (64, 32 or 16).
// pair: the two choices of the last vertex share their elimination (orders >= 7, lanes 64).
hipError_t launch_lu(int order, int precision, int lanes, bool pair, const KdParams &p, const PrepBuf &buf, int grid,
                     hipStream_t st);
// integrand_tree_kernel<order, T> (orders 6..12): the Keldysh sum by the
// Schur-complement tree, one point per 64-thread workgroup item.
hipError_t launch_tree(int order, int precision, const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st);
// integrand_split_kernel<order, T> (orders 6..10): tree over the leading
// vertices, then one bordered LU per configuration; one point per workgroup
// of split_threads(order) threads (0: order not supported).
int split_threads(int order);
// whether it runs the configurations as pairs sharing the last vertex's elimination
bool split_paired(int order);
hipError_t launch_split(int order, int precision, const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st);
// integrand_wsplit_kernel<order, T> (orders 6..10): the split algorithm with
// one point per 64-thread workgroup (one wave), no workgroup barriers.
bool wsplit_supported(int order);
hipError_t launch_wsplit(int order, int precision, const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st);
// integrand_ctree_kernel<order, T> (orders 6..10): the time-ordered tree of
// vertex-level corrections, 64-thread workgroups (one wave) of one point each,
// or of two points (one per half) where ctree_pairs(order).
bool ctree_supported(int order);
// whether the time-ordered tree runs two points per wave (one per half)
bool ctree_pairs(int order);
hipError_t launch_ctree(int order, int precision, const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st);
// cofactor_kernel<order> (kernel method, orders 1..12, FP64): grid blocks of
// 64 threads, same work items as the Keldysh-sum kernel.
hipError_t launch_cofactor(int order, const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st);
// acc[0..3] += sum over blocks of part[4*blk + 0..3] (fixed order).  `part`
// holds partials_doubles(cap) doubles: 4 cap partials, then the reduction's
// block sums and its ticket, zero before the first launch (the reduction leaves
// it zero); nblk <= cap.
constexpr int kRedThreads = 256, kRedSlice = 4 * kRedThreads;
constexpr int reduce_blocks(int nblk) { return (nblk + kRedSlice - 1) / kRedSlice; }
constexpr size_t partials_doubles(int cap) { return 4 * (size_t)cap + 4 * (size_t)reduce_blocks(cap) + 1; }
hipError_t launch_reduce(double *part, int nblk, int cap, double *acc, hipStream_t st);
// per-(axis, bin) sums of |vals| and counts over n samples l [n][order] (pilot projection)
hipError_t launch_pilot_hist(const double *l, const double *vals, int64_t n, int order, int nbins, double *sums,
                             unsigned long long *counts, hipStream_t st);
// acc[b] += sum over blocks of part[blk][b], b < nbins (fixed order).
hipError_t launch_reduce_bins(const double2 *part, int nblk, int nbins, double2 *acc, hipStream_t st);

// ---- learned flow warper (kd_flow.hip) ----
struct FlowMapIO {
  const uint32_t *Csr, *shfr; // Sobol source (when in_l is null): points k0+1 ...
  uint32_t k0;
  const double *in_l; // or caller points [n][order]
  int64_t n_points;
  double *out_x;   // [n][order]
  double *out_jac; // [n]
  double *out_l;   // optional: the l points used [n][order]
};
struct FlowTrainIO {
  const double *x; // training points [n][order]
  const double *w; // weights |integrand| [n]
  const double *c; // per-point conditioning values [n] (nc = 1), or null: FlowParams::cond
  int64_t n_points, offset;
  double inv_norm; // 1 / (mean(w) * batch)
  float *part;     // per-block gradients [grid][n_params]
  double *loss_part;
  int n_params;
};
// flow_map_kernel<order>: 64-thread blocks, grid-stride over 64-point batches.
hipError_t launch_flow_map(int order, const FlowParams &fp, const FlowMapIO &io, int grid, hipStream_t st);
// flow_train_kernel<order>: one 64-point batch per 64-thread block.
hipError_t launch_flow_train(int order, const FlowParams &fp, const FlowTrainIO &io, int grid, hipStream_t st);
hipError_t launch_flow_grad_reduce(const float *part, int nblk, int n_params, float *grad, const double *loss_part,
                                   double *loss, hipStream_t st);
hipError_t launch_flow_adam(float *w, float *m, float *v, const float *g, int n, float lr, float b1, float b2,
                            float eps, int step, hipStream_t st);
hipError_t launch_flow_pack(const float *w, uint16_t *pk, int order, int layers, int nc, hipStream_t st);
hipError_t launch_flow_abs(const double *vals, int64_t n, double power, double *w, double *part, int grid,
                           hipStream_t st);

} // namespace kd
