// kd_integrand.hip -- the fused Keldysh-sum kernel for gfx950.
//
// One 64-lane wavefront per workgroup.  Per pass the wave takes 64 candidate
// points (one per lane): Sobol point or caller-supplied point, warper chain in
// registers, domain check (wick_direct.cpp:53-57).  In-domain points are
// compacted with a ballot.  For each group of compacted points the wave stages
// the (2n+1)^2 complex Wick matrix of g0 values in LDS
// (wick_direct.cpp:92-101; the opposite-spin matrix equals its leading 2n x 2n
// block because g0 is spin independent, model.cpp:463-465).  Each lane then
// owns one or more Keldysh configurations sigma (wick_direct.cpp:110-149) and
// computes
//     (-1)^|sigma| det(S1_sigma) det(S2_sigma) = (-1)^|sigma| det(A)^2 s,
// with A = S2_sigma, S1_sigma = [[A, b], [c, 0]] and s = -c A^{-1} b the
// bordered Schur complement, by a register-resident row-wise LU of the
// bordered matrix.  The LU runs without row exchanges; a pivot that lost more
// than a set number of bits to cancellation flags the configuration, which is
// then recomputed by a partial-pivoting FP64 LU.  The alternating sum over
// sigma is reduced with wavefront shuffles in FP64 and accumulated in FP64.
#include <type_traits>

#include "kd_device.hpp"
#include "kd_kernels.hpp"
#include "kd_lu.hpp"
#include "kd_pivoted.hpp"

namespace kd {


// det(A)^2 * s for one Keldysh configuration, unpivoted row-wise LU in T.
template <int N, typename T>
__device__ __forceinline__ double2 config_value(const double2 *__restrict__ Wq, int cfg, T sc, bool &bad) {
  constexpr int DS = wick_stride(N);
  int pr[N + 1], pc[N + 1];
  static_for<0, N>([&](auto I) {
    constexpr int i = I;
    int pi = 2 * i + ((cfg >> i) & 1);
    pr[i] = pi * DS;
    pc[i] = pi;
  });
  pr[N] = 2 * N * DS;
  pc[N] = 2 * N;
  cx<T> Uh[N + 1][N + 1];
  cx<T> det = cmk<T>(T(1), T(0));
  cx<T> s = cmk<T>(T(0), T(0));
  bool flag = false;
  static_for<0, N + 1>([&](auto I) {
    constexpr int i = I;
    cx<T> row[N + 1];
    static_for<0, N + 1>([&](auto J) {
      constexpr int j = J;
      if constexpr (i == N && j == N) {
        row[j] = cmk<T>(T(0), T(0));
      } else {
        double2 w = Wq[pr[i] + pc[j]];
        row[j] = cmk<T>((T)w.x, (T)w.y);
      }
    });
    T diag0 = T(0);
    if constexpr (i < N) diag0 = fmax(fmax(cabs1<T>(row[i]), sc), PivotTol<T>::floor); // > 0: a zero pivot fails
    constexpr int KE = i < N ? i : N;
    static_for<0, KE>([&](auto K) {
      constexpr int k = K;
      cx<T> l = row[k];
      static_for<k + 1, N + 1>([&](auto J) {
        constexpr int j = J;
        row[j] = cfms<T>(row[j], l, Uh[k][j]);
      });
    });
    if constexpr (i < N) {
      cx<T> piv = row[i];
      T pm = cabs1<T>(piv);
      flag = flag || !(pm >= PivotTol<T>::v * diag0);
      if constexpr (i == 0)
        det = piv;
      else
        det = cmul<T>(det, piv);
      cx<T> inv = cinv<T>(piv);
      static_for<i + 1, N + 1>([&](auto J) {
        constexpr int j = J;
        Uh[i][j] = cmul<T>(row[j], inv);
      });
    } else {
      s = row[N];
    }
    __builtin_amdgcn_sched_barrier(0);
  });
  bad = flag;
  cx<T> d2 = cmul<T>(det, det);
  cx<T> v = cmul<T>(d2, s);
  return make_double2((double)v.x, (double)v.y);
}

// det(A)^2 * s for the two configurations that differ only in the Keldysh
// index of the last vertex (cfg0 has it 0; v[t], bad[t] for index t).  Their
// bordered LUs share the elimination of the first N-1 vertex rows: those rows
// are eliminated once over the columns of both choices of vertex N-1 (and the
// border), the external row likewise, and only the last vertex row and the
// final corner update are done per choice -- 268 instead of 408 complex
// multiply-adds per pair at order 8, same arithmetic per configuration.
template <int N, typename T>
__device__ __forceinline__ void config_value_pair(const double2 *__restrict__ Wq, int cfg0, T sc, double2 (&v)[2],
                                                  bool (&bad)[2]) {
  constexpr int DS = wick_stride(N);
  constexpr int M = N + 2; // columns: vertices 0..N-2, vertex N-1 (index 0), vertex N-1 (index 1), border
  int pr[N - 1], pc[M];
  static_for<0, N - 1>([&](auto I) {
    constexpr int i = I;
    const int pi = 2 * i + ((cfg0 >> i) & 1);
    pr[i] = pi * DS;
    pc[i] = pi;
  });
  pc[N - 1] = 2 * N - 2;
  pc[N] = 2 * N - 1;
  pc[N + 1] = 2 * N;
  cx<T> Uh[N - 1][M];
  cx<T> det = cmk<T>(T(1), T(0));
  bool flag = false;
  static_for<0, N - 1>([&](auto I) {
    constexpr int i = I;
    cx<T> row[M];
    static_for<0, M>([&](auto J) {
      constexpr int j = J;
      const double2 w = Wq[pr[i] + pc[j]];
      row[j] = cmk<T>((T)w.x, (T)w.y);
    });
    const T diag0 = fmax(cabs1<T>(row[i]), sc);
    static_for<0, i>([&](auto K) {
      constexpr int k = K;
      const cx<T> l = row[k];
      static_for<k + 1, M>([&](auto J) {
        constexpr int j = J;
        row[j] = cfms<T>(row[j], l, Uh[k][j]);
      });
    });
    const cx<T> piv = row[i];
    const T pm = cabs1<T>(piv);
    flag = flag || !(pm >= PivotTol<T>::v * diag0) || pm == T(0);
    det = cmul<T>(det, piv);
    const cx<T> inv = cinv<T>(piv);
    static_for<i + 1, M>([&](auto J) {
      constexpr int j = J;
      Uh[i][j] = cmul<T>(row[j], inv);
    });
    __builtin_amdgcn_sched_barrier(0);
  });
  // external row over the same columns (its border entry, the corner, is 0)
  cx<T> rb[M];
  static_for<0, M>([&](auto J) {
    constexpr int j = J;
    if constexpr (j == M - 1) {
      rb[j] = cmk<T>(T(0), T(0));
    } else {
      const double2 w = Wq[2 * N * DS + pc[j]];
      rb[j] = cmk<T>((T)w.x, (T)w.y);
    }
  });
  static_for<0, N - 1>([&](auto K) {
    constexpr int k = K;
    const cx<T> l = rb[k];
    static_for<k + 1, M>([&](auto J) {
      constexpr int j = J;
      rb[j] = cfms<T>(rb[j], l, Uh[k][j]);
    });
  });
  __builtin_amdgcn_sched_barrier(0);
  // the last vertex row for each Keldysh index t
  static_for<0, 2>([&](auto TT) {
    constexpr int t = TT;
    constexpr int ct = N - 1 + t; // its column
    cx<T> row[M];
    static_for<0, M>([&](auto J) {
      constexpr int j = J;
      if constexpr (j < N - 1 || j == ct || j == M - 1) {
        const double2 w = Wq[pc[ct] * DS + pc[j]];
        row[j] = cmk<T>((T)w.x, (T)w.y);
      }
    });
    const T diag0 = fmax(cabs1<T>(row[ct]), sc);
    static_for<0, N - 1>([&](auto K) {
      constexpr int k = K;
      const cx<T> l = row[k];
      static_for<k + 1, M>([&](auto J) {
        constexpr int j = J;
        if constexpr (j < N - 1 || j == ct || j == M - 1) row[j] = cfms<T>(row[j], l, Uh[k][j]);
      });
    });
    const cx<T> piv = row[ct];
    const T pm = cabs1<T>(piv);
    bad[t] = flag || !(pm >= PivotTol<T>::v * diag0) || pm == T(0);
    const cx<T> dt = cmul<T>(det, piv);
    const cx<T> u = cmul<T>(row[M - 1], cinv<T>(piv));
    const cx<T> s = cfms<T>(rb[M - 1], rb[ct], u);
    const cx<T> d2 = cmul<T>(dt, dt);
    const cx<T> val = cmul<T>(d2, s);
    v[t] = make_double2((double)val.x, (double)val.y);
    __builtin_amdgcn_sched_barrier(0);
  });
}

// Warper chain, applied last-added first (warpers.hpp:72-85).
template <int N>
__device__ __forceinline__ double apply_warpers(const KdParams &p, const WarperDev *W, const MapDev *M,
                                                double (&x)[N]) {
  double jac = 1.0;
  for (int q = p.n_warpers - 1; q >= 0; q--) {
    const kd_warper_desc wd = W[q].w;
    if (wd.kind == KD_WARP_PLASMA) {
      // ui_from_vi (plasma_uv.hpp:46-51)
      x[0] = wd.t_max - x[0];
#pragma unroll
      for (int i = 1; i < N; i++) x[i] = x[i - 1] - x[i];
    } else if (wd.kind == KD_WARP_PRODUCT) {
      // axes in order 0..N-1, the Jacobian multiplied in that order whatever the path
      const MapDev *ms = M + wd.map_offset;
      const bool one = wd.n_maps == 1;
      const int kind = W[q].map_kind;
      double inv_f[N];
      if (kind == KD_MAP_NEAREST) {
        nearest_axes<N>(ms, one, p.tables, p.guides, p.knots, x, inv_f);
      } else if (kind == KD_MAP_INVERSE) {
#pragma unroll
        for (int i = 0; i < N; i++) inv_f[i] = map_reverse_k<KD_MAP_INVERSE>(ms[one ? 0 : i], p.tables, p.guides, x[i], &x[i]);
      } else {
#pragma unroll
        for (int i = 0; i < N; i++) inv_f[i] = map1d_reverse(ms[one ? 0 : i], p.tables, p.guides, x[i], &x[i]);
      }
#pragma unroll
      for (int i = 0; i < N; i++) jac *= inv_f[i];
    }
  }
  return jac;
}

// Stage 1: candidate points.  One thread per point; a 64-point batch b is
// one wavefront.  l (Sobol point or caller input) -> warper chain -> u, J and
// the domain check (wick_direct.cpp:53-57).  In-domain points of batch b are
// compacted, in lane order, to slots [64 b, 64 b + cnt[b]).
// three waves per SIMD up to order 8 (the point chains are latency-bound: more
// waves in flight), two above (whose coordinates would spill at three)
template <int N> constexpr int prep_waves_per_eu() { return N <= 8 ? 3 : 2; }
template <int N>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(prep_waves_per_eu<N>()))) void
prep_kernel(KdParams p, PrepBuf buf) {
  constexpr int NU = N > 0 ? N : 1;
  __shared__ uint32_t sC[NU * 31];
  __shared__ uint32_t sS[NU];
  // bin counts of this workgroup (nbins + 1: the last one counts the dropped
  // values), flushed to the global counters once at the end
  extern __shared__ unsigned long long scnt[];
  // the chain's descriptors: read from LDS, their loads do not queue behind
  // the knot-table loads of the axes before them
  __shared__ WarperDev sW[KD_MAX_WARPERS];
  __shared__ MapDev sM[KD_MAX_MAPS];
  // Sobol points unless the caller (or the flow stage) supplies the points
  const bool sob = (p.mode == MODE_RUN || p.mode == MODE_SAMPLE) && p.in_pts == nullptr;
  if (sob) {
    for (int i = threadIdx.x; i < N * 31; i += blockDim.x) sC[i] = p.Csr[i];
    if (threadIdx.x < N) sS[threadIdx.x] = p.shfr[threadIdx.x];
  }
  if (p.mode != MODE_EVAL_U) {
    const int nw = p.n_warpers * (int)(sizeof(WarperDev) / 4), nm = p.n_maps * (int)(sizeof(MapDev) / 4);
    for (int i = threadIdx.x; i < nw; i += blockDim.x)
      reinterpret_cast<uint32_t *>(sW)[i] = reinterpret_cast<const uint32_t *>(p.warpers)[i];
    for (int i = threadIdx.x; i < nm; i += blockDim.x)
      reinterpret_cast<uint32_t *>(sM)[i] = reinterpret_cast<const uint32_t *>(p.maps)[i];
  }
  for (int i = threadIdx.x; i <= p.nbins && p.nbins > 0; i += blockDim.x) scnt[i] = 0ull;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t n = p.n_points;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const bool spin_ok = p.A.spin == p.B.spin;
  for (int64_t pi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pi - lane < n; pi += stride) {
    // keep the descriptors' LDS reads inside the loop (no hoisting into registers)
    __asm__ volatile("" ::: "memory");
    const bool valid = pi < n;
    double x[NU];
    double jac = 1.0;
    bool dom = valid && spin_ok;
    if (valid) {
      if (p.in_jac) jac = p.in_jac[pi];
      if (sob) {
        const uint32_t k = p.k0 + 1u + (uint32_t)pi;
#pragma unroll
        for (int i = 0; i < N; i++) x[i] = sobol_coord(sC, sS, i, k);
        if (p.mode == MODE_SAMPLE) {
#pragma unroll
          for (int i = 0; i < N; i++) p.out_l[pi * N + i] = x[i];
        }
      } else {
#pragma unroll
        for (int i = 0; i < N; i++) x[i] = p.in_pts[pi * N + i];
      }
      if constexpr (N > 0) {
        if (p.mode != MODE_EVAL_U) jac *= apply_warpers<N>(p, sW, sM, x);
      }
      if (p.keep_u) {
#pragma unroll
        for (int i = 0; i < N; i++) dom = dom && !(x[i] < 0.0);
      }
    }
    if (p.nbins > 0 && valid && p.obs == OBS_DIRECT) {
      // binner accumulate(value or 0, min time): one count per point
      double tmin = x[0];
#pragma unroll
      for (int i = 1; i < N; i++) tmin = fmin(tmin, x[i]);
      const int bi = bin_index(p, tmin);
      atomicAdd(scnt + (bi >= 0 ? bi : p.nbins), 1ull);
    }
    if (p.nbins > 0 && dom && p.obs == OBS_KERNEL) {
      // an in-domain point adds 2^N entries per vertex (2^(N-1) per Keldysh
      // index) at that vertex's time; counted per time bin, or dropped
#pragma unroll
      for (int i = 0; i < N; i++) {
        const int bi = bin_index(p, x[i]);
        atomicAdd(scnt + (bi >= 0 ? bi : p.nbins), bi >= 0 ? (1ull << (N > 0 ? N - 1 : 0)) : (1ull << N));
      }
    }
    const uint64_t mask = __ballot(dom);
    const int64_t b = pi >> 6;
    if (dom) {
      const int pos = __popcll(mask & ((1ull << lane) - 1ull));
      const int64_t slot = b * 64 + pos;
#pragma unroll
      for (int i = 0; i < N; i++) buf.u[slot * N + i] = x[i];
      buf.jac[slot] = jac;
      buf.idx[slot] = lane;
    }
    if (lane == 0) buf.cnt[b] = __popcll(mask);
    if (valid && !dom && p.mode != MODE_RUN) {
      // (the kernel method's per-configuration entries are cleared by the host)
      if (p.obs != OBS_KERNEL) {
        p.out_vals[2 * pi] = 0.0;
        p.out_vals[2 * pi + 1] = 0.0;
      }
      if (p.out_dom) p.out_dom[pi] = 0;
      if (p.out_jac) p.out_jac[pi] = jac;
    }
  }
  if (p.nbins > 0) {
    __syncthreads();
    for (int i = threadIdx.x; i <= p.nbins; i += blockDim.x)
      if (scnt[i]) atomicAdd(p.bin_counts + i, scnt[i]);
  }
}

// Stage 2: the Keldysh sum of the compacted points; 64-lane workgroups,
// grid-stride over groups of points.
// PR: the two choices of the last vertex's Keldysh index share their
// elimination (config_value_pair), for 2^N / lanes >= 2.
template <int N, typename T, int GL, bool PR = false>
__global__ __launch_bounds__(64) void integrand_kernel(KdParams p, PrepBuf buf) {
  constexpr int NC = 1 << N;            // Keldysh configurations per point
  constexpr int G = NC < GL ? NC : GL;  // lanes per point (GL: the tuned maximum)
  constexpr int PPW = 64 / G;           // points per group
  constexpr int CPL = NC / G;           // configurations per lane
  constexpr int D = 2 * N + 1;
  constexpr int NU = N > 0 ? N : 1;
  constexpr int DS = wick_stride(N);
  __shared__ double2 W[PPW][D * DS];
  __shared__ double Us[PPW][NU];
  __shared__ unsigned long long Ssc[PPW]; // per point: max |Re|+|Im| of the vertex block (bits of a double >= 0)
  extern __shared__ double2 sbins[]; // p.nbins entries when time-binning (dynamic LDS)

  const int lane = threadIdx.x;
  if (p.nbins > 0) {
    for (int i = threadIdx.x; i < p.nbins; i += 64) sbins[i] = make_double2(0.0, 0.0);
    __syncthreads();
  }
  const int64_t nbatch = (p.n_points + 63) / 64;
  double acc_re = 0.0, acc_im = 0.0, acc_dom = 0.0, acc_fb = 0.0;

  // Work item = one group of PPW compacted points of a 64-point batch, so that
  // a launch of few points still spreads over every CU.  Items are numbered
  // group-major (all batches' first groups, then their second groups, ...):
  // the compacted points sit at the front of each batch, so the items that
  // hold work come first and are dealt evenly over the workgroups.
  constexpr int GPB = 64 / PPW; // groups per batch
  // item = grp * nbatch + b, stepped by the grid size without a division per item
  int grp_i = nbatch > 0 ? (int)(blockIdx.x / nbatch) : GPB; // no points: no items
  int64_t b_i = blockIdx.x - (int64_t)grp_i * nbatch;
  while (grp_i < GPB) {
    const int64_t grp = grp_i;
    const int64_t b = b_i;
    b_i += gridDim.x;
    while (b_i >= nbatch) {
      b_i -= nbatch;
      grp_i++;
    }
    const int g0i = (int)grp * PPW;
    const int cnt = buf.cnt[b];
    if (g0i >= cnt) continue;
    {
      const int ng = min(PPW, cnt - g0i);
      // Opaque per-iteration copy of the lane id: keeps the lane-only index
      // math of the matrix build from being hoisted into registers.
      int lane = threadIdx.x;
      __asm__ volatile("" : "+v"(lane));
      const int64_t slot0 = b * 64 + g0i;
      if (lane < ng * N) Us[lane / NU][lane % NU] = buf.u[slot0 * N + lane];
      if (lane < PPW) Ssc[lane] = 0ull;
      __syncthreads();
      // Stage the Wick matrices of the group's points in LDS (wick_direct.cpp:92-101),
      // one vertex pair or external entry per item (wick_item).
      constexpr int NI = wick_items<N>();
      for (int e = lane; e < ng * NI; e += 64) {
        const int q = e / NI;
        const double wm = wick_item<N>(p, Us[q], W[q], DS, e - q * NI);
        if (wm > 0.0) atomicMax(&Ssc[q], (unsigned long long)__double_as_longlong(wm));
      }
      __syncthreads();

      const int q = lane / G;
      double sre = 0.0, sim = 0.0;
      int nfb = 0;
      if (q < ng) {
        if constexpr (N == 0) {
          const double2 v = W[q][0];
          sre = v.x;
          sim = v.y;
        } else {
          // Flags of the lane's configurations, one 64-bit word per 64 of
          // them (CPL reaches 256 at order 12 with 16 lanes per point); each
          // word's flagged configurations are recomputed before the next word.
          constexpr int CPW = CPL < 64 ? CPL : 64; // configurations per word
#pragma unroll 1
          for (int w0 = 0; w0 < CPL; w0 += CPW) {
            uint64_t badmask = 0;
            if constexpr (PR && CPL >= 2) {
              // bit 2 (mp - w0 / 2) + t: configuration mp + t CPL / 2, mp in [w0 / 2, (w0 + CPW) / 2)
#pragma unroll 1
              for (int mp = w0 / 2; mp < (w0 + CPW) / 2; mp++) {
                const int cfg0 = (lane % G) + G * mp; // last vertex at index 0; + 2^(N-1) for index 1
                double2 v[2];
                bool bad[2];
                config_value_pair<N, T>(PPW == 1 ? W[0] : W[q], cfg0,
                                        (T)__longlong_as_double((long long)Ssc[PPW == 1 ? 0 : q]), v, bad);
                const double sg = (__popc(cfg0) & 1) ? -1.0 : 1.0;
                sre += bad[0] ? 0.0 : sg * v[0].x;
                sim += bad[0] ? 0.0 : sg * v[0].y;
                sre += bad[1] ? 0.0 : -sg * v[1].x;
                sim += bad[1] ? 0.0 : -sg * v[1].y;
                const int bit = 2 * (mp - w0 / 2);
                badmask |= ((uint64_t)bad[0] << bit) | ((uint64_t)bad[1] << (bit + 1));
              }
            } else {
#pragma unroll 1
              for (int m = w0; m < w0 + CPW; m++) {
                const int cfg = (lane % G) + G * m;
                bool bad;
                const double2 v = config_value<N, T>(PPW == 1 ? W[0] : W[q], cfg,
                                                     (T)__longlong_as_double((long long)Ssc[PPW == 1 ? 0 : q]), bad);
                // branch-free: a divergent branch here doubles the register footprint
                const double sg = (__popc(cfg) & 1) ? -1.0 : 1.0;
                sre += bad ? 0.0 : sg * v.x;
                sim += bad ? 0.0 : sg * v.y;
                badmask |= (uint64_t)bad << (m - w0);
              }
            }
            // Flagged configurations: partial-pivoting FP64 LU, outside the hot loop.
            while (badmask) {
              const int bit = __builtin_ctzll(badmask);
              badmask &= badmask - 1;
              const int m = (PR && CPL >= 2) ? w0 / 2 + (bit >> 1) + (bit & 1) * (CPL / 2) : w0 + bit;
              const int cfg = (lane % G) + G * m;
              const double2 v = config_value_pivoted(W[q], cfg, N, DS);
              const double sg = (__popc(cfg) & 1) ? -1.0 : 1.0;
              sre += sg * v.x;
              sim += sg * v.y;
              nfb++;
            }
          }
        }
      }
#pragma unroll
      for (int off = 1; off < G; off <<= 1) {
        sre += __shfl_xor(sre, off, 64);
        sim += __shfl_xor(sim, off, 64);
      }
      acc_fb += (double)nfb;
      if ((lane % G) == 0 && q < ng) {
        const int64_t slot = slot0 + q;
        if (sqrt(sre * sre + sim * sim) < p.cutoff) {
          sre = 0.0;
          sim = 0.0;
        }
        const double jj = buf.jac[slot];
        sre *= jj;
        sim *= jj;
        if (p.mode == MODE_RUN) {
          acc_re += sre;
          acc_im += sim;
          acc_dom += 1.0;
          if (p.nbins > 0) {
            double tmin = Us[q][0];
#pragma unroll
            for (int i = 1; i < N; i++) tmin = fmin(tmin, Us[q][i]);
            const int bi = bin_index(p, tmin);
            if (bi >= 0) {
              atomicAdd(&sbins[bi].x, sre);
              atomicAdd(&sbins[bi].y, sim);
            }
          }
        } else {
          const int64_t idx = b * 64 + buf.idx[slot];
          p.out_vals[2 * idx] = sre;
          p.out_vals[2 * idx + 1] = sim;
          if (p.out_dom) p.out_dom[idx] = 1;
          if (p.out_jac) p.out_jac[idx] = jj;
        }
      }
      __syncthreads();
    }
  }

  if (p.mode == MODE_RUN) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      acc_re += __shfl_xor(acc_re, off, 64);
      acc_im += __shfl_xor(acc_im, off, 64);
      acc_dom += __shfl_xor(acc_dom, off, 64);
      acc_fb += __shfl_xor(acc_fb, off, 64);
    }
    if (lane == 0) {
      double *o = p.partials + 4 * (int64_t)blockIdx.x;
      o[0] = acc_re;
      o[1] = acc_im;
      o[2] = acc_dom;
      o[3] = acc_fb;
    }
    if (p.nbins > 0) {
      __syncthreads();
      for (int i = threadIdx.x; i < p.nbins; i += 64) p.bin_part[(int64_t)blockIdx.x * p.nbins + i] = sbins[i];
    }
  }
}

// Pilot projection (make_warper_from_proj.hpp:80-114 gather_data): per axis a
// and bin b of [0, 1], the sum of |value| over the samples whose coordinate a
// falls in b, and their count.  One thread per (axis, bin) scans the samples in
// order, so the sums are those of a sequential loop, bit for bit.
__global__ __launch_bounds__(256) void pilot_hist_kernel(const double *__restrict__ l, const double *__restrict__ vals,
                                                         int64_t n, int order, int nbins, double *__restrict__ sums,
                                                         unsigned long long *__restrict__ counts) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= order * nbins) return;
  const int a = t / nbins, b = t - a * nbins;
  const double bs = 1.0 / nbins;
  double s = 0.0;
  unsigned long long c = 0;
  for (int64_t i = 0; i < n; i++) {
    const double x = l[i * order + a];
    if (!(x >= 0.0 && x <= 1.0)) continue;
    const int bi = x == 1.0 ? nbins - 1 : (int)(x / bs);
    if (bi == b) {
      s += hypot(vals[2 * i], vals[2 * i + 1]);
      c++;
    }
  }
  sums[t] = s;
  counts[t] = c;
}

// acc[b] += sum over blocks of part[blk][b]: one 256-thread workgroup per bin,
// strided partial sums then a tree in LDS (a fixed order: deterministic).
__global__ __launch_bounds__(256) void reduce_bins_kernel(const double2 *__restrict__ part, int nblk, int nbins,
                                                          double2 *__restrict__ acc) {
  __shared__ double2 sh[256];
  const int b = blockIdx.x, t = threadIdx.x;
  double2 s = make_double2(0.0, 0.0);
  for (int k = t; k < nblk; k += 256) {
    const double2 v = part[(int64_t)k * nbins + b];
    s.x += v.x;
    s.y += v.y;
  }
  sh[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      sh[t].x += sh[t + w].x;
      sh[t].y += sh[t + w].y;
    }
    __syncthreads();
  }
  if (t == 0) {
    acc[b].x += sh[0].x;
    acc[b].y += sh[0].y;
  }
}

// Deterministic sum of per-block partials into acc[0..3] (+=), spread over
// reduce_blocks(nblk) workgroups of kRedThreads threads.  Block k sums partials
// [k kRedSlice, (k + 1) kRedSlice): thread t the four at t + j kRedThreads in
// that order, then a fixed tree; it stores its four sums in slice[4k..4k+3] and
// takes a ticket.  The block that takes the last ticket adds the slices in block
// order into acc and re-arms the ticket (zero between launches; launches of one
// accumulator are stream-ordered).  Which block finishes last decides only who
// adds, not the order of the additions: the same sum for the same partials on
// every run.
__global__ __launch_bounds__(kRedThreads) void reduce_partials_kernel(const double *__restrict__ part, int nblk,
                                                                      double *slice, unsigned *ticket,
                                                                      double *acc) {
  __shared__ double sh[4][kRedThreads];
  const int t = threadIdx.x;
  const int i0 = (int)blockIdx.x * kRedSlice + t;
  double2 v[4][2];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int i = i0 + u * kRedThreads;
    if (i < nblk) {
      v[u][0] = reinterpret_cast<const double2 *>(part)[2 * i];
      v[u][1] = reinterpret_cast<const double2 *>(part)[2 * i + 1];
    } else {
      v[u][0] = v[u][1] = make_double2(0.0, 0.0);
    }
  }
  double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int u = 0; u < 4; u++) {
    s[0] += v[u][0].x;
    s[1] += v[u][0].y;
    s[2] += v[u][1].x;
    s[3] += v[u][1].y;
  }
#pragma unroll
  for (int c = 0; c < 4; c++) sh[c][t] = s[c];
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if (t < w)
#pragma unroll
      for (int c = 0; c < 4; c++) sh[c][t] += sh[c][t + w];
    __syncthreads();
  }
  __shared__ int last;
  if (t == 0) {
#pragma unroll
    for (int c = 0; c < 4; c++) slice[4 * blockIdx.x + c] = sh[c][0];
    __threadfence(); // release: this block's slice before its ticket
    last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence(); // acquire: every block's slice
  const int G = (int)gridDim.x;
  if (G <= kRedThreads) {
    // the slices loaded at once, then added in block order
    if (t < G)
#pragma unroll
      for (int c = 0; c < 4; c++) sh[c][t] = slice[4 * t + c];
    __syncthreads();
    if (t < 4) {
      double a = 0.0;
      for (int k = 0; k < G; k++) a += sh[t][k];
      acc[t] += a;
    }
  } else if (t < 4) {
    double a = 0.0;
    for (int k = 0; k < G; k++) a += slice[4 * k + t];
    acc[t] += a;
  }
  if (t == 0) *ticket = 0u;
}

template <int N> static hipError_t launch_prep_n(const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st) {
  const size_t shm = p.nbins > 0 ? sizeof(unsigned long long) * ((size_t)p.nbins + 1) : 0;
  hipLaunchKernelGGL((prep_kernel<N>), dim3(grid), dim3(256), shm, st, p, buf);
  return hipGetLastError();
}

template <int N, typename T>
static hipError_t launch_lu_n(int lanes, bool pair, const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st) {
  // lane-count variants only where they differ (2^N > lanes)
  const size_t shm = p.nbins > 0 ? sizeof(double2) * (size_t)p.nbins : 0;
  if (pair && N >= 7 && lanes == 64)
    hipLaunchKernelGGL((integrand_kernel<N, T, 64, true>), dim3(grid), dim3(64), shm, st, p, buf);
  else if (lanes == 16 && N >= 5)
    hipLaunchKernelGGL((integrand_kernel<N, T, 16>), dim3(grid), dim3(64), shm, st, p, buf);
  else if (lanes == 32 && N >= 6)
    hipLaunchKernelGGL((integrand_kernel<N, T, 32>), dim3(grid), dim3(64), shm, st, p, buf);
  else
    hipLaunchKernelGGL((integrand_kernel<N, T, 64>), dim3(grid), dim3(64), shm, st, p, buf);
  return hipGetLastError();
}

#define KD_ORDER_SWITCH(ORDER, CALL)                                                                           \
  switch (ORDER) {                                                                                             \
  case 0: return CALL(0);                                                                                      \
  case 1: return CALL(1);                                                                                      \
  case 2: return CALL(2);                                                                                      \
  case 3: return CALL(3);                                                                                      \
  case 4: return CALL(4);                                                                                      \
  case 5: return CALL(5);                                                                                      \
  case 6: return CALL(6);                                                                                      \
  case 7: return CALL(7);                                                                                      \
  case 8: return CALL(8);                                                                                      \
  case 9: return CALL(9);                                                                                      \
  case 10: return CALL(10);                                                                                    \
  case 11: return CALL(11);                                                                                    \
  case 12: return CALL(12);                                                                                    \
  default: return hipErrorInvalidValue;                                                                        \
  }

hipError_t launch_prep(int order, const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st) {
#define KD_PREP_CALL(n) launch_prep_n<n>(p, buf, grid, st)
  KD_ORDER_SWITCH(order, KD_PREP_CALL)
#undef KD_PREP_CALL
}

hipError_t launch_lu(int order, int precision, int lanes, bool pair, const KdParams &p, const PrepBuf &buf, int grid,
                     hipStream_t st) {
  if (lu_is_fp32(order, precision)) {
#define KD_LU32_CALL(n) launch_lu_n<n, float>(lanes, pair, p, buf, grid, st)
    KD_ORDER_SWITCH(order, KD_LU32_CALL)
#undef KD_LU32_CALL
  }
#define KD_LU64_CALL(n) launch_lu_n<n, double>(lanes, pair, p, buf, grid, st)
  KD_ORDER_SWITCH(order, KD_LU64_CALL)
#undef KD_LU64_CALL
}

hipError_t launch_reduce(double *part, int nblk, int cap, double *acc, hipStream_t st) {
  if (nblk < 1 || nblk > cap) return hipErrorInvalidValue;
  double *slice = part + 4 * (size_t)cap;
  unsigned *ticket = reinterpret_cast<unsigned *>(slice + 4 * (size_t)reduce_blocks(cap));
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(reduce_blocks(nblk)), dim3(kRedThreads), 0, st, part, nblk, slice,
                     ticket, acc);
  return hipGetLastError();
}

hipError_t launch_pilot_hist(const double *l, const double *vals, int64_t n, int order, int nbins, double *sums,
                             unsigned long long *counts, hipStream_t st) {
  const int tot = order * nbins;
  hipLaunchKernelGGL(pilot_hist_kernel, dim3((tot + 255) / 256), dim3(256), 0, st, l, vals, n, order, nbins, sums,
                     counts);
  return hipGetLastError();
}

hipError_t launch_reduce_bins(const double2 *part, int nblk, int nbins, double2 *acc, hipStream_t st) {
  hipLaunchKernelGGL(reduce_bins_kernel, dim3(nbins), dim3(256), 0, st, part, nblk, nbins, acc);
  return hipGetLastError();
}

} // namespace kd
