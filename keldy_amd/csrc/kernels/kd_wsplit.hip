// kd_wsplit.hip -- the split Keldysh sum with one point per wavefront (orders
// 6..10): no workgroup barriers, every synchronisation inside the wave.
//
// The algorithm is that of kd_split.hip: the leading K vertices are eliminated
// once per node of the tree of their Keldysh choices (level-k tables
// T_k[(i,s)][(j,t)] over the remaining doubled index set, index 2(i-k)+s,
// border 2(n-k); child with sigma_k = s: T_{k+1}[a][b] = T_k[a+2][b+2] -
// T_k[a+2][s] T_k[s][b+2] / T_k[s][s]), and each configuration finishes with a
// bordered LU of its M = n-K remaining vertices from its level-K node table;
// det A is the product of the path pivots and its own.  Here K = n - 5 from
// order 8 on (M = 5, the two choices of the last vertex paired) and K = 2 at
// orders 6-7 (M = 4, 5), so that the per-configuration LUs stay at five
// vertices and the deeper levels carry the sharing.
//
// A 64-lane wave walks the tree depth first over its first SD = K - 2 levels
// (one node per level, recomputed only where the prefix sigma_0..sigma_{SD-1}
// changed) and forms the last two levels in one step: per prefix, the four
// level-K grandchildren of the level-SD node by a rank-2 update (wlevel2), whose
// 128 configurations (64 at order 6) are one LU (pair) per lane.  LDS per wave
// at order 8, FP64: the Wick matrix 4.9 KB, the level-1 node 4.3 KB, the four
// level-3 tables 7.7 KB -- 17.5 KB, so eight waves (two per SIMD) fit on a CU
// with no workgroup barrier between a point's phases: the waves of a CU
// overlap each other's latency freely.
//
// Pivot check and fallback as in kd_split.hip: every pivot on a
// configuration's path (the tree's and its own) against PivotTol times
// max(|its original diagonal entry|, the vertex block's largest |Re|+|Im|); a
// flagged configuration is recomputed by partial-pivoting FP64 LU from the
// Wick matrix.
#include "kd_device.hpp"
#include "kd_kernels.hpp"
#include "kd_lu.hpp"
#include "kd_pivoted.hpp"
#include "kd_split_lu.hpp"
#include "kd_wave.hpp"

namespace kd {

template <int N> struct WShape {
  static constexpr int K = N <= 7 ? 2 : N - 5; // tree levels
  static constexpr int M = N - K;              // vertices per configuration LU
  static constexpr int SD = K - 2;             // depth-first levels
  static constexpr bool PAIR = M == 5;         // one pair (5 vertices) per lane; order 6: one configuration
  // row strides (complex entries) of the Wick matrix and the level tables; at
  // order 8 the Wick matrix and the level-1 table use 17 (272 B: consecutive
  // rows of a column 16 B apart across the banks), measured 1.4% faster there
  // than the general choice
  static constexpr int S0 = N == 8 ? 17 : wick_stride(N);
  static constexpr int Dk(int k) { return 2 * (N - k) + 1; }
  static constexpr int Sk(int k) { return k == 0 ? S0 : (N == 8 && k == 1) ? 17 : wick_stride(N - k); }
  static constexpr int nodes(int k) { return k <= SD ? 1 : 1 << (k - SD); } // per prefix
  static constexpr int tab(int k) { return Dk(k) * Sk(k); }                 // entries per node table
  static constexpr int size(int k) { return nodes(k) * tab(k); }
  // levels 1..SD in their own storage, one after the other; level K (from
  // level SD in one step) in Y
  static constexpr int doff(int k) { return k <= 1 ? 0 : doff(k - 1) + size(k - 1); }
  static constexpr int dsize() { return SD >= 1 ? doff(SD + 1) : 1; }
  static constexpr int ysize() { return size(K); }
  static constexpr int NI = wick_items<N>(), NR = (NI + 63) / 64; // Wick items, rounds of 64
};

// Children of the NPAR parent tables at level k (local index pm, at stride
// Dp * Sp in `par`) into `ch` (stride Dc * Sc): with ONE, the single child of
// parent 0 along the Keldysh index s1; otherwise both children of every
// parent, local index pm + NPAR s.  Work item = (child, row, column chunk) over
// the 64 lanes.  The item of a child's row 0, chunk 0 also writes its path
// determinant and pivot flag into the heap (node of level k with Keldysh bits
// g at (1 << k) + g); gp0 is the bits of parent 0, and parent pm has gp0 +
// (pm << sd).
template <typename T, typename PT, int Dp, int Sp, int Sc, int NPAR, bool ONE>
__device__ __forceinline__ void wlevel(const PT *__restrict__ par, cx<T> *__restrict__ ch,
                                       double2 *__restrict__ dpath, int *__restrict__ fpath,
                                       const T *__restrict__ thr, int k, int gp0, int sd, int s1, int lane) {
  constexpr int Dc = Dp - 2, NCH = ONE ? 1 : 2 * NPAR;
  constexpr int CH0 = 64 / (NCH * Dc);
  constexpr int CH = CH0 < 1 ? 1 : CH0 > Dc ? Dc : CH0; // column chunks per row
  constexpr int CW = (Dc + CH - 1) / CH;                // columns per chunk
  for (int r = lane; r < NCH * Dc * CH; r += 64) {
    const int q = r / CH, c0 = (r - q * CH) * CW;
    const int m = q / Dc, a = q - m * Dc;
    const int pm = ONE ? 0 : (m & (NPAR - 1));
    const int s = ONE ? s1 : m / NPAR;
    const PT *P = par + pm * (Dp * Sp);
    const PT *Pa = P + (a + 2) * Sp + 2 + c0;
    const PT *Ps = P + s * Sp + 2 + c0;
    // the pivot and the row's multiplier input first: LDS returns in issue
    // order, so the reciprocal's dependent chain runs while the columns
    // stream in.  Every operand of the item is loaded before any result is
    // stored: the compiler cannot move these loads above the stores into `ch`
    // by itself, and a load-wait-store round trip per column would serialise
    // the item
    const cx<T> piv = ldx<T>(P[s * Sp + s]);
    const cx<T> la = ldx<T>(P[(a + 2) * Sp + s]);
    cx<T> pa[CW], ps[CW];
#pragma unroll
    for (int b = 0; b < CW; b++)
      if (Dc % CW == 0 || c0 + b < Dc) {
        pa[b] = ldx<T>(Pa[b]);
        ps[b] = ldx<T>(Ps[b]);
      }
    const cx<T> l = cmul<T>(la, cinv<T>(piv));
    cx<T> *C = ch + m * (Dc * Sc) + a * Sc + c0;
#pragma unroll
    for (int b = 0; b < CW; b++)
      if (Dc % CW == 0 || c0 + b < Dc) C[b] = cfms<T>(pa[b], l, ps[b]);
    if (a == 0 && c0 == 0) {
      const int gp = gp0 + (pm << sd);
      const int hp = (1 << k) + gp, hc = (2 << k) + gp + (s << k);
      const double2 d = dpath[hp];
      dpath[hc] = make_double2(d.x * (double)piv.x - d.y * (double)piv.y, d.x * (double)piv.y + d.y * (double)piv.x);
      fpath[hc] = fpath[hp] | (int)!(cabs1<T>(piv) >= thr[2 * k + s]);
    }
  }
  wave_sync();
}

// The four grandchildren (Keldysh indices s1 of vertex k, s2 of vertex k+1;
// local index s1 + 2 s2) of one level-k table P in a single step: eliminating
// the two vertices one after the other leaves the Schur complement of P's
// 2x2 block B = P[{s1, 2+s2}][{s1, 2+s2}],
//     G[a][b] = P[a+4][b+4] - (P[a+4][s1], P[a+4][2+s2]) B^-1 (P[s1][b+4]; P[2+s2][b+4]),
// the same matrix as two single steps in exact arithmetic, with one LDS round
// trip and no level-(k+1) table.  The path pivots are B[0][0] and det B /
// B[0][0] (their product det B goes into the path determinant).  Work item =
// (grandchild, row, column chunk); gp0 = the Keldysh bits of P, grandchild m
// has gp0 + (m << k).
template <typename T, typename PT, int Dp, int Sp, int Sc>
__device__ __forceinline__ void wlevel2(const PT *__restrict__ P, cx<T> *__restrict__ ch, double2 *__restrict__ dpath,
                                        int *__restrict__ fpath, const T *__restrict__ thr, int k, int gp0,
                                        int lane) {
  constexpr int Dc = Dp - 4, NCH = 4;
  constexpr int CH0 = 64 / (NCH * Dc);
  constexpr int CH = CH0 < 1 ? 1 : CH0 > Dc ? Dc : CH0; // column chunks per row
  constexpr int CW = (Dc + CH - 1) / CH;                // columns per chunk
  for (int r = lane; r < NCH * Dc * CH; r += 64) {
    const int q = r / CH, c0 = (r - q * CH) * CW;
    const int m = q / Dc, a = q - m * Dc;
    const int s1 = m & 1, r2 = 2 + (m >> 1); // the two eliminated rows / columns of P
    const PT *Pa = P + (a + 4) * Sp + 4 + c0;
    const PT *P1 = P + s1 * Sp + 4 + c0;
    const PT *P2 = P + r2 * Sp + 4 + c0;
    // the pivot block and the row's two entries first (their dependent chain
    // runs while the columns stream in), then every column operand before any
    // store
    const cx<T> b00 = ldx<T>(P[s1 * Sp + s1]), b01 = ldx<T>(P[s1 * Sp + r2]);
    const cx<T> b10 = ldx<T>(P[r2 * Sp + s1]), b11 = ldx<T>(P[r2 * Sp + r2]);
    const cx<T> ra = ldx<T>(P[(a + 4) * Sp + s1]), rb = ldx<T>(P[(a + 4) * Sp + r2]);
    cx<T> pa[CW], p1[CW], p2[CW];
#pragma unroll
    for (int b = 0; b < CW; b++)
      if (Dc % CW == 0 || c0 + b < Dc) {
        pa[b] = ldx<T>(Pa[b]);
        p1[b] = ldx<T>(P1[b]);
        p2[b] = ldx<T>(P2[b]);
      }
    const cx<T> det = cfms<T>(cmul<T>(b00, b11), b01, b10);
    const cx<T> idet = cinv<T>(det);
    // (c1, c2) = (ra, rb) B^-1,  B^-1 = (b11, -b01; -b10, b00) / det B
    const cx<T> c1 = cmul<T>(cfms<T>(cmul<T>(ra, b11), rb, b10), idet);
    const cx<T> c2 = cmul<T>(cfms<T>(cmul<T>(rb, b00), ra, b01), idet);
    cx<T> *C = ch + m * (Dc * Sc) + a * Sc + c0;
#pragma unroll
    for (int b = 0; b < CW; b++)
      if (Dc % CW == 0 || c0 + b < Dc) C[b] = cfms<T>(cfms<T>(pa[b], c1, p1[b]), c2, p2[b]);
    if (a == 0 && c0 == 0) {
      const int hp = (1 << k) + gp0, hc = (4 << k) + gp0 + (m << k);
      const double2 d = dpath[hp];
      dpath[hc] = make_double2(d.x * (double)det.x - d.y * (double)det.y, d.x * (double)det.y + d.y * (double)det.x);
      // the second pivot, det B / B[0][0], tested without the division:
      // |Re|+|Im| of det B conj(B[0][0]) against the threshold times |B[0][0]|^2
      const cx<T> q2 = cmul<T>(det, cmk<T>(b00.x, -b00.y));
      const T n00 = b00.x * b00.x + b00.y * b00.y;
      fpath[hc] = fpath[hp] | (int)!(cabs1<T>(b00) >= thr[2 * k + s1]) |
                  (int)!(cabs1<T>(q2) >= thr[2 * (k + 1) + (m >> 1)] * n00);
    }
  }
  wave_sync();
}

#ifndef KD_WSPLIT_WAVES_PER_EU
#define KD_WSPLIT_WAVES_PER_EU 2
#endif
template <int N, typename T>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(KD_WSPLIT_WAVES_PER_EU))) void
integrand_wsplit_kernel(KdParams p, PrepBuf buf) {
  using S = WShape<N>;
  constexpr int K = S::K, M = S::M, SD = S::SD;
  constexpr int D0 = 2 * N + 1, S0 = S::S0;
  constexpr bool PAIR = S::PAIR;
  __shared__ double2 W[D0 * S0];
  __shared__ cx<T> Dl[S::dsize()], Y[S::ysize()];
  __shared__ double2 dpath[2 << K];
  __shared__ int fpath[2 << K];
  __shared__ T thr[2 * N];
  __shared__ double Us[N];
  extern __shared__ double2 sbins[];

  const int lane = threadIdx.x;
  if (p.nbins > 0) {
    for (int i = lane; i < p.nbins; i += 64) sbins[i] = make_double2(0.0, 0.0);
    wave_sync();
  }
  const int64_t nbatch = (p.n_points + 63) / 64;
  double acc_re = 0.0, acc_im = 0.0, acc_dom = 0.0, acc_fb = 0.0;
  const bool lane_sums = p.mode == MODE_RUN && p.nbins == 0 && !(p.cutoff > 0.0);
  // Work items and the window scan as in kd_split.hip: one compacted point per
  // item, numbered slot-major (item = grp * nbatch + b), every gridDim.x-th
  // item to this wave; lane l of a window loads the batch count of the
  // window's item l and a ballot keeps the live ones.  The next live item's
  // vertex times and Jacobian are loaded while the current one is computed.
  const int64_t G = gridDim.x;
  int wgrp = nbatch > 0 ? (int)(blockIdx.x / nbatch) : 64; // no points: no items (zero partials)
  int64_t wb = blockIdx.x - (int64_t)wgrp * nbatch;
  uint64_t mask = 0;
  int gl = 0, bl = 0;
  double u_n = 0.0, jac_n = 0.0;
  for (;;) {
    if (!mask) {
      if (wgrp >= 64) break;
      const int64_t bw = wb + (int64_t)lane * G;
      gl = wgrp + (int)(bw / nbatch);
      bl = (int)(bw % nbatch);
      mask = __ballot(gl < 64 && gl < buf.cnt[bl]);
      const int64_t adv = wb + 64 * G;
      wgrp += (int)(adv / nbatch);
      wb = adv % nbatch;
      if (!mask) continue;
      const int j1 = __builtin_ctzll(mask);
      const int64_t s1 = (int64_t)rdlane(bl, j1) * 64 + rdlane(gl, j1);
      u_n = buf.u[s1 * N + (lane < N ? lane : N - 1)];
      jac_n = buf.jac[s1];
    }
    const int j = __builtin_ctzll(mask);
    mask &= mask - 1;
    const int g = rdlane(gl, j);
    const int64_t b = rdlane(bl, j);
    const int64_t slot = b * 64 + g;
    const double uc = u_n, jj = jac_n;
    // Opaque per-point copy of the lane id: keeps the lane-only index math of
    // the matrix build, the tree levels and the LUs' LDS addresses from being
    // hoisted out of the point loop into registers.
    int lq = lane;
    __asm__ volatile("" : "+v"(lq));
    if (lq < N) Us[lq] = uc;
    if (lq == 0) {
      dpath[1] = make_double2(1.0, 0.0);
      fpath[1] = 0;
    }
    wave_sync();
    // Wick matrix, interleaved by vertex (wick_direct.cpp:92-101 picks), and the
    // vertex block's largest |Re|+|Im| across the wave: every item's table
    // entries are fetched before any is interpolated, so the rounds' loads
    // are in flight together
    double wmax = 0.0;
    {
      constexpr int NI = S::NI, NR = S::NR;
      WickPre wp[NR];
      static_for<0, NR>([&](auto R) {
        const int q = lq + 64 * (int)R;
        if (q < NI) {
          int ia, ib;
          wick_item_vertices<N>(q, ia, ib);
          wick_fetch<N>(p, q, ia >= 0 ? Us[ia] : p.A.time, ib >= 0 ? Us[ib] : p.B.time, wp[R]);
        }
      });
      static_for<0, NR>([&](auto R) {
        const int q = lq + 64 * (int)R;
        if (q < NI) wmax = fmax(wmax, wick_store<N>(p, W, S0, q, wp[R]));
      });
    }
    wmax = wave_max(wmax);
    // the next live item's vertex times and Jacobian, in flight until the next
    // point (issued after this point's table loads, which would otherwise wait
    // for them)
    // (every lane loads, so that no lane keeps an old value: a partial load
    // would make the compiler copy the register and wait for the load at once)
    if (mask) {
      const int j2 = __builtin_ctzll(mask);
      const int64_t s2 = (int64_t)rdlane(bl, j2) * 64 + rdlane(gl, j2);
      u_n = buf.u[s2 * N + (lane < N ? lane : N - 1)];
      jac_n = buf.jac[s2];
    }
    wave_sync();
    if (lq < 2 * N) {
      const double2 wd = W[lq * S0 + lq];
      thr[lq] = PivotTol<T>::v * fmax(fmax((T)(fabs(wd.x) + fabs(wd.y)), (T)wmax), PivotTol<T>::floor);
    }
    wave_sync();

    double sre = 0.0, sim = 0.0;
    int nfb = 0;
    // prefixes sigma_0..sigma_{SD-1} (sigma_i = bit i of pfx), the deepest
    // choice varying fastest: pfx is the bit reversal of the counter q
#pragma unroll 1
    for (int q = 0; q < (1 << SD); q++) {
      const int pfx = SD >= 1 ? (int)(__builtin_bitreverse32((uint32_t)q) >> (32 - (SD >= 1 ? SD : 1))) : 0;
      // depth-first levels: recompute from the first level whose choice changed
      // (sigma_{SD-1-ctz(q)} and those below it in the counter)
      if constexpr (SD >= 1) {
        const int k0 = q == 0 ? 0 : SD - 1 - __builtin_ctz(q); // levels k0+1 .. SD change
        static_for<0, SD>([&](auto KI) {
          constexpr int k = KI; // parent level; child level k + 1
          if (k >= k0) {
            cx<T> *dst = Dl + S::doff(k + 1);
            const int sk = (pfx >> k) & 1;
            const int gp = pfx & ((1 << k) - 1);
            if constexpr (k == 0)
              wlevel<T, double2, D0, S0, S::Sk(1), 1, true>(W, dst, dpath, fpath, thr, 0, gp, SD, sk, lq);
            else
              wlevel<T, cx<T>, S::Dk(k), S::Sk(k), S::Sk(k + 1), 1, true>(Dl + S::doff(k), dst, dpath, fpath, thr,
                                                                           k, gp, SD, sk, lq);
          }
        });
      }
      // level K: the four grandchildren of the level-SD node (of W at SD = 0)
      if constexpr (SD == 0)
        wlevel2<T, double2, D0, S0, S::Sk(K)>(W, Y, dpath, fpath, thr, 0, 0, lq);
      else
        wlevel2<T, cx<T>, S::Dk(SD), S::Sk(SD), S::Sk(K)>(Dl + S::doff(SD), Y, dpath, fpath, thr, SD, pfx, lq);

      // one LU (pair) per lane: node c >> (M - 1) of level K (bits sigma_SD,
      // sigma_{SD+1}), local bits sigma_K.. from the low bits of c
      const int c = lq;
      bool bad[2] = {false, false};
      int cfg0;
      if constexpr (PAIR) {
        const int loc = c & ((1 << (M - 1)) - 1), node = c >> (M - 1);
        cfg0 = pfx | (node << SD) | (loc << K);
        const int h = (1 << K) + pfx + (node << SD);
        double2 v[2];
        split_config_pair<M, T>(Y + node * S::tab(K), loc, thr + 2 * K, dpath[h], v, bad);
        const bool fp = fpath[h] != 0;
        bad[0] = bad[0] || fp;
        bad[1] = bad[1] || fp;
        const double sg = (__popc(cfg0) & 1) ? -1.0 : 1.0; // the index-1 configuration has the other sign
        sre += bad[0] ? 0.0 : sg * v[0].x;
        sim += bad[0] ? 0.0 : sg * v[0].y;
        sre += bad[1] ? 0.0 : -sg * v[1].x;
        sim += bad[1] ? 0.0 : -sg * v[1].y;
      } else {
        const int loc = c & ((1 << M) - 1), node = c >> M;
        cfg0 = pfx | (node << SD) | (loc << K);
        const int h = (1 << K) + pfx + (node << SD);
        const double2 v = split_config<M, T>(Y + node * S::tab(K), loc, thr + 2 * K, dpath[h], bad[0]);
        bad[0] = bad[0] || fpath[h] != 0;
        const double sg = (__popc(cfg0) & 1) ? -1.0 : 1.0;
        sre += bad[0] ? 0.0 : sg * v.x;
        sim += bad[0] ? 0.0 : sg * v.y;
      }
      // flagged configurations: partial-pivoting FP64 LU from the Wick matrix
      if (bad[0]) {
        const double2 v = config_value_pivoted(W, cfg0, N, S0);
        const double sg = (__popc(cfg0) & 1) ? -1.0 : 1.0;
        sre += sg * v.x;
        sim += sg * v.y;
        nfb++;
      }
      if (PAIR && bad[1]) {
        const int cfg1 = cfg0 | (1 << (N - 1));
        const double2 v = config_value_pivoted(W, cfg1, N, S0);
        const double sg = (__popc(cfg1) & 1) ? -1.0 : 1.0;
        sre += sg * v.x;
        sim += sg * v.y;
        nfb++;
      }
      wave_sync(); // X, Y (and the level-SD node in Y) are rewritten by the next prefix
    }
    acc_fb += (double)nfb;
    if (lane_sums) {
      // no per-point total needed (no cutoff, no time bins): every lane keeps
      // its own running sum, reduced across the wave once at the end
      acc_re += sre * jj;
      acc_im += sim * jj;
      acc_dom += lane == 0 ? 1.0 : 0.0;
      wave_sync(); // W, Us and the heap are rewritten by the next point
      continue;
    }
    sre = wave_sum(sre);
    sim = wave_sum(sim);
    if (lane == 0) {
      if (sqrt(sre * sre + sim * sim) < p.cutoff) {
        sre = 0.0;
        sim = 0.0;
      }
      sre *= jj;
      sim *= jj;
      if (p.mode == MODE_RUN) {
        acc_re += sre;
        acc_im += sim;
        acc_dom += 1.0;
        if (p.nbins > 0) {
          double tmin = Us[0];
#pragma unroll
          for (int i = 1; i < N; i++) tmin = fmin(tmin, Us[i]);
          const int bi = bin_index(p, tmin);
          if (bi >= 0) {
            sbins[bi].x += sre;
            sbins[bi].y += sim;
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
    wave_sync(); // W, Us and the heap are rewritten by the next point
  }

  if (p.mode == MODE_RUN) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc_fb += __shfl_xor(acc_fb, off, 64);
    // lane 0 alone holds the sums on the per-point path (the other lanes add zeros)
    acc_re = wave_sum(acc_re);
    acc_im = wave_sum(acc_im);
    if (lane == 0) {
      double *o = p.partials + 4 * (int64_t)blockIdx.x;
      o[0] = acc_re;
      o[1] = acc_im;
      o[2] = acc_dom;
      o[3] = acc_fb;
    }
    if (p.nbins > 0) {
      wave_sync();
      for (int i = lane; i < p.nbins; i += 64) p.bin_part[(int64_t)blockIdx.x * p.nbins + i] = sbins[i];
    }
  }
}

template <int N, typename T>
static hipError_t launch_wsplit_n(const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st) {
  const size_t shm = p.nbins > 0 ? sizeof(double2) * (size_t)p.nbins : 0;
  hipLaunchKernelGGL((integrand_wsplit_kernel<N, T>), dim3(grid), dim3(64), shm, st, p, buf);
  return hipGetLastError();
}

bool wsplit_supported(int order) { return order >= 6 && order <= 10; }

hipError_t launch_wsplit(int order, int precision, const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st) {
  const bool f32 = lu_is_fp32(order, precision);
#define KD_WSPLIT_CASE(n)                                                                                         \
  case n:                                                                                                         \
    return f32 ? launch_wsplit_n<n, float>(p, buf, grid, st) : launch_wsplit_n<n, double>(p, buf, grid, st);
  switch (order) {
    KD_WSPLIT_CASE(6)
    KD_WSPLIT_CASE(7)
    KD_WSPLIT_CASE(8)
    KD_WSPLIT_CASE(9)
    KD_WSPLIT_CASE(10)
  default:
    return hipErrorInvalidValue;
  }
#undef KD_WSPLIT_CASE
}

} // namespace kd
