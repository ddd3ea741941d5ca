// kd_ctree.hip -- the Keldysh sum by the time-ordered tree (orders 6..10), one
// point per wavefront.
//
// In the contour order of compare_3way (common.cpp:47-57), g0 between two
// vertices depends only on the Keldysh index of the earlier one (earlier time,
// or the lower index at equal times): for vertex i after vertex j,
//     g0(i a; j b) = b ? g0^<(t_i - t_j) : g0^>(t_i - t_j)     for both a,
//     g0(j b; i a) = b ? g0^>(t_j - t_i) : g0^<(t_j - t_i)     for both a,
// and the diagonal g0(i a; i a) = g0^<(0) - i alpha is one constant D.  With
// the vertices sorted by (time, index), entry (i, j) of configuration sigma's
// bordered matrix (wick_direct.cpp:92-101 picks) is W(i, j, sigma_min(i,j)),
// the border column W(i, e, sigma_i), the border row W(e, j, sigma_j).
// Eliminating vertex k first to last, with pivot p = D - C[k][k], subtracts
//     u_i w_j,   u_i = W(i, k, s) - C[i][k],   w_j = (W(k, j, s) - C[k][j]) / p
// from every remaining entry (i, j): u and w do not depend on the Keldysh
// indices of the vertices after k, so the Schur complement after vertices
// 0..k-1 is W - C with a correction C that depends on sigma_0..sigma_{k-1}
// only -- one value per vertex pair instead of the four of the interleaved
// Wick matrix.  The tree of these corrections over the 2^n configurations
// holds one (n-k+1)^2 table per node of level k; a leaf (all n vertices
// eliminated) gives det of the vertex block = the product of its pivots and the
// bordered determinant = that product times 0 - C[e][e], so the configuration's
// term (-1)^|sigma| det(A)^2 s.
//
// Layout per wave (one point at a time): the sorted times, the vertex-level
// table Wv in LDS (FP64); levels 1..3 of the tree breadth first in LDS, one
// (child, row) item per lane and level; below level 3 each lane follows its own
// path -- sigma_0..sigma_5 are the lane's bits -- with the level-4 and level-5
// entries formed on demand from the level-3 table and the rank-1 factors of the
// levels between (each entry once), then the subtree over the vertices after
// the sixth in registers, its leaves summed per lane.
//
// Pivot check and fallback as in the other Keldysh-sum kernels: every pivot on
// a configuration's path against PivotTol times the largest |Re|+|Im| of the
// point's vertex block; a flagged configuration is recomputed by
// partial-pivoting FP64 LU of its matrix from Wv.
#include "kd_device.hpp"
#include "kd_kernels.hpp"
#include "kd_lu.hpp"
#include "kd_pivoted.hpp"
#include "kd_split_lu.hpp"
#include "kd_wave.hpp"

namespace kd {

// a + l b
template <typename T> __device__ __forceinline__ cx<T> cfma(cx<T> a, cx<T> l, cx<T> b) {
  return cmk<T>(a.x + l.x * b.x - l.y * b.y, a.y + l.x * b.y + l.y * b.x);
}
template <typename T> __device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return cmk<T>(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 dmul(double2 d, double2 p) {
  return make_double2(d.x * p.x - d.y * p.y, d.x * p.y + d.y * p.x);
}
template <typename T> __device__ __forceinline__ double2 dcx(cx<T> v) { return make_double2((double)v.x, (double)v.y); }

template <int N> struct CShape {
  static constexpr int WS = N + 1;                           // Wv row stride: vertices, then the border e = N
  static constexpr int R(int k) { return N - k + 1; }        // a level-k table: vertices k..N-1 and the border
  static constexpr int RS(int k) { return R(k); }                // row stride of a level-k table
  static constexpr int NS(int k) { return (R(k) * RS(k)) | 1; }  // node stride (odd: nodes spread over the banks)
  static constexpr int sizeA() { return 8 * NS(3); } // level 3
  static constexpr int sizeB() { return 4 * NS(2); } // level 2
  // global index of local row/column a of a level-k table
  static constexpr int gidx(int k, int a) { return a < R(k) - 1 ? k + a : N; }
};

template <int N> __device__ __forceinline__ const double2 &wv(const double2 *Wv, int i, int j, int s) {
  return Wv[(i * (N + 1) + j) * 2 + s];
}

// Level K+1 of the tree from level K (K = 0: the root, whose correction is 0):
// child m has parent g = m mod 2^K and sigma_K = m >> K; one item per (child,
// group of RPI rows) per lane of the LANES working on the point, all columns of
// its rows (the parent's first row and pivot loaded once per item).  The items
// of a child's row 0 also extend the path determinant and pivot flag (heap
// index (1 << k) + bits for level k).  FROM_W (K = 1): the level-1 parent is
// not read from a table but formed from Wv -- a level-1 correction is the
// rank-1 product C1[i][j] = (W(i, 0, s0) / D) W(0, j, s0), the same products
// the level-1 pass would store -- so that level 1 needs no pass of its own.
template <int N, int K, typename T, int LANES = 64, int RPI = 1, bool FROM_W = false>
__device__ __forceinline__ void clevel(const cx<T> *__restrict__ par, cx<T> *__restrict__ ch,
                                       const double2 *__restrict__ Wv, cx<T> D, cx<T> invD, T thr,
                                       double2 *__restrict__ dpath, int *__restrict__ fpath, int lane) {
  using S = CShape<N>;
  static_assert(!FROM_W || K == 1, "only level 1 is formed from Wv");
  constexpr int r = S::R(K), rc = r - 1, rsp = S::RS(K), rsc = S::RS(K + 1);
  constexpr int NRG = (rc + RPI - 1) / RPI; // row groups per child
  constexpr int NIT = (2 << K) * NRG;
  for (int it = lane; it < NIT; it += LANES) {
    const int m = it / NRG, ag = it - m * NRG;
    const int g = m & ((1 << K) - 1), s = m >> K;
    const cx<T> *P = par + g * S::NS(K);
    // level-1 factors of parent g (FROM_W): l0(i) = W(i, 0, g) / D, w0(j) = W(0, j, g)
    auto l0 = [&](int i) { return cmul<T>(ldx<T>(wv<N>(Wv, S::gidx(1, i), 0, g)), invD); };
    auto w0 = [&](int j) { return ldx<T>(wv<N>(Wv, 0, S::gidx(1, j), g)); };
    cx<T> piv = D, inv = invD, wr[rc];
    cx<T> l00 = cmk<T>(0, 0);
    if constexpr (FROM_W) {
      l00 = l0(0);
      piv = csub<T>(D, cmul<T>(l00, w0(0)));
    } else if constexpr (K > 0) {
      piv = csub<T>(D, P[0]);
    }
    if constexpr (K > 0) inv = cinv<T>(piv);
    static_for<0, rc>([&](auto B) {
      constexpr int b = B;
      wr[b] = ldx<T>(wv<N>(Wv, K, S::gidx(K + 1, b), s));
      if constexpr (FROM_W)
        wr[b] = csub<T>(wr[b], cmul<T>(l00, w0(b + 1)));
      else if constexpr (K > 0)
        wr[b] = csub<T>(wr[b], P[b + 1]);
    });
    static_for<0, RPI>([&](auto RI) {
      const int a = ag * RPI + (int)RI;
      if (RPI == 1 || a < rc) {
        const int ig = a < rc - 1 ? K + 1 + a : N;
        cx<T> u = ldx<T>(wv<N>(Wv, ig, K, s));
        cx<T> la = cmk<T>(0, 0);
        if constexpr (FROM_W) {
          la = l0(a + 1);
          u = csub<T>(u, cmul<T>(la, w0(0)));
        } else if constexpr (K > 0) {
          u = csub<T>(u, P[(a + 1) * rsp]);
        }
        const cx<T> l = cmul<T>(u, inv);
        cx<T> *C = ch + m * S::NS(K + 1) + a * rsc;
        static_for<0, rc>([&](auto B) {
          constexpr int b = B;
          if constexpr (FROM_W)
            C[b] = cfma<T>(cmul<T>(la, w0(b + 1)), l, wr[b]);
          else if constexpr (K > 0)
            C[b] = cfma<T>(P[(a + 1) * rsp + b + 1], l, wr[b]);
          else
            C[b] = cmul<T>(l, wr[b]);
        });
        if (a == 0) {
          const int hp = (1 << K) + g, hc = (2 << K) + m;
          if constexpr (FROM_W) {
            // the level-1 node's path: det D, flagged with D
            dpath[hc] = dmul(dcx<T>(D), dcx<T>(piv));
            fpath[hc] = (int)!(cabs1<T>(D) >= thr) | (int)!(cabs1<T>(piv) >= thr);
          } else {
            dpath[hc] = dmul(dpath[hp], dcx<T>(piv));
            fpath[hc] = fpath[hp] | (int)!(cabs1<T>(piv) >= thr);
          }
        }
      }
    });
  }
  wave_sync();
}

// A lane's path below its level-3 node: the entries of its level-k tables
// (k = 3..3+NL) on demand -- level 3 from LDS, level k > 3 as level k-1's
// entry plus the rank-1 term of vertex k-1 -- and the rank-1 factors u, w of
// the vertices 3..2+NL along the lane's Keldysh choices.
template <int N, typename T, int NL> struct LanePath {
  using S = CShape<N>;
  static constexpr int R3 = S::R(3);
  const cx<T> *P3;
  cx<T> u[NL][R3 - 1], w[NL][R3 - 1];
  template <int K, int A, int B> __device__ __forceinline__ cx<T> at() const {
    if constexpr (K == 3)
      return ldx<T>(P3[A * S::RS(3) + B]);
    else
      return cfma<T>(at<K - 1, A + 1, B + 1>(), u[K - 4][A], w[K - 4][B]);
  }
  // vertex K eliminated with Keldysh index s (K < N - 1): its factors; returns the pivot
  template <int K> __device__ __forceinline__ cx<T> step(const double2 *__restrict__ Wv, cx<T> D, int s) {
    constexpr int r = S::R(K);
    const cx<T> piv = csub<T>(D, at<K, 0, 0>());
    const cx<T> inv = cinv<T>(piv);
    static_for<0, r - 1>([&](auto A) {
      constexpr int a = A;
      constexpr int ig = S::gidx(K + 1, a);
      u[K - 3][a] = csub<T>(ldx<T>(wv<N>(Wv, ig, K, s)), at<K, a + 1, 0>());
      w[K - 3][a] = cmul<T>(csub<T>(ldx<T>(wv<N>(Wv, K, ig, s)), at<K, 0, a + 1>()), inv);
    });
    return piv;
  }
};

// One leaf: vertex K = N-1 eliminated with index s from its 2x2 table
// {(K,K), (K,e), (e,K), (e,e)} = c00, c0e, ce0, cee.  Its term is
//   (-1)^|sigma| (dp p)^2 (0 - cee - ue ve / p) = -(-1)^|sigma| dp^2 x,  x = p (p cee + ue ve);
// returns x, and bad when p fails the pivot check.
template <int N, typename T>
__device__ __forceinline__ cx<T> leafx(const double2 *__restrict__ Wv, cx<T> D, T thr, cx<T> c00, cx<T> c0e,
                                      cx<T> ce0, cx<T> cee, int s, bool &bad) {
  constexpr int K = N - 1;
  const cx<T> piv = csub<T>(D, c00);
  const cx<T> ue = csub<T>(ldx<T>(wv<N>(Wv, N, K, s)), ce0);
  const cx<T> ve = csub<T>(ldx<T>(wv<N>(Wv, K, N, s)), c0e);
  bad = !(cabs1<T>(piv) >= thr);
  return cmul<T>(piv, cfma<T>(cmul<T>(piv, cee), ue, ve));
}

// The subtree of the lane's level-K node (R = N - K + 1, table C) over all
// choices of vertices K..N-1.  cfg holds sigma_0..sigma_{K-1}, the low LB of
// them the lane's bits; a flagged leaf sets bit cfg >> LB of bad.
template <int N, int R, typename T, int LB = 6, bool SERIAL = false>
__device__ __forceinline__ void subtree(const cx<T> (&C)[R][R], const double2 *__restrict__ Wv, cx<T> D, T thr,
                                        double2 dp, int flag, int cfg, double &sre, double &sim, uint32_t &bad) {
  constexpr int K = N + 1 - R;
  if constexpr (R == 2) {
    // the two leaves' x, signed, summed in FP64, then one product with dp^2
    // (a flagged path can make dp non-finite: the product is taken only when a
    // leaf passed, whose path pivots all did)
    double2 xs = make_double2(0.0, 0.0);
    const double sg = (__popc(cfg) & 1) ? 1.0 : -1.0; // -(-1)^|sigma| of the index-0 leaf
    bool good = false;
    static_for<0, 2>([&](auto Sx) {
      constexpr int s = Sx;
      bool b;
      const double2 x = dcx<T>(leafx<N, T>(Wv, D, thr, C[0][0], C[0][1], C[1][0], C[1][1], s, b));
      b = b || flag;
      const double g = s ? -sg : sg;
      xs.x += b ? 0.0 : g * x.x;
      xs.y += b ? 0.0 : g * x.y;
      good = good || !b;
      bad |= (uint32_t)b << ((cfg | (s << K)) >> LB);
    });
    const double2 v = dmul(dmul(dp, dp), xs);
    sre += good ? v.x : 0.0;
    sim += good ? v.y : 0.0;
  } else {
    // vertex K with Keldysh index s: the child's table, then its subtree
    // (SERIAL: one child after the other, the second's registers not live
    // during the first)
    const cx<T> piv = csub<T>(D, C[0][0]);
    const cx<T> inv = cinv<T>(piv);
    const double2 dpc = dmul(dp, dcx<T>(piv));
    const int flagc = flag | (int)!(cabs1<T>(piv) >= thr);
    auto child = [&](int s) {
      cx<T> u[R - 1], w[R - 1], ch[R - 1][R - 1];
      static_for<0, R - 1>([&](auto A) {
        constexpr int a = A;
        constexpr int ig = CShape<N>::gidx(K + 1, a);
        u[a] = csub<T>(ldx<T>(wv<N>(Wv, ig, K, s)), C[a + 1][0]);
        w[a] = cmul<T>(csub<T>(ldx<T>(wv<N>(Wv, K, ig, s)), C[0][a + 1]), inv);
      });
      static_for<0, R - 1>([&](auto A) {
        static_for<0, R - 1>([&](auto B) { ch[A][B] = cfma<T>(C[A + 1][B + 1], u[A], w[B]); });
      });
      subtree<N, R - 1, T, LB>(ch, Wv, D, thr, dpc, flagc, cfg | (s << K), sre, sim, bad);
    };
    if constexpr (SERIAL) {
#pragma unroll 1
      for (int s = 0; s < 2; s++) child(s);
    } else {
      static_for<0, 2>([&](auto Sx) { child(Sx); });
    }
  }
}

// waves per SIMD the register allocation targets: two up to order 9 (at most
// 256 VGPRs; at orders 7 and 8 2-3% faster than three waves of 168 VGPRs), one
// at order 10, whose lane path and subtree then keep their tables in the 512
// VGPRs and AGPRs of a lane instead of spilling to memory (order 10, all points
// in the domain: 18.6 ms per 2^21 points against 29.5 ms at two waves per SIMD)
template <int N> constexpr int ctree_waves_per_eu() { return N <= 9 ? 2 : 1; }
template <int N, typename T>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(ctree_waves_per_eu<N>()))) void
integrand_ctree_kernel(KdParams p, PrepBuf buf) {
  static_assert(N >= 6 && N <= 10, "time-ordered tree: orders 6..10");
  using S = CShape<N>;
  constexpr int WS = S::WS;
  __shared__ double2 Wv[WS * WS * 2];
  __shared__ cx<T> LA[S::sizeA()], LB[S::sizeB()];
  __shared__ double2 dpath[16];
  __shared__ int fpath[16];
  __shared__ double Us[N], Ts[N];
  extern __shared__ double2 sbins[];

  const int lane = threadIdx.x;
  if (p.nbins > 0) {
    for (int i = lane; i < p.nbins; i += 64) sbins[i] = make_double2(0.0, 0.0);
    wave_sync();
  }
  // the diagonal D = g0^<(0) - i alpha and the constant part of the vertex
  // block's scale (|D| and the same-vertex entries g0^<(0), g0^>(0))
  const double2 gl0 = table_eval(p, p.lesser, 0, 0, 0.0), gg0 = table_eval(p, p.greater, 0, 0, 0.0);
  const double2 Dd = make_double2(gl0.x, gl0.y - p.alpha);
  const double wconst = fmax(fmax(cabs1<double>(gl0), cabs1<double>(gg0)), cabs1<double>(Dd));
  const cx<T> D = cmk<T>((T)Dd.x, (T)Dd.y);
  const cx<T> invD = cinv<T>(D);

  // a launch holds at most 2^21 points (the C-ABI launches chunks of that
  // size), so item, batch and slot indices fit 32 bits
  const int nbatch = (int)((p.n_points + 63) / 64);
  double acc_re = 0.0, acc_im = 0.0, acc_dom = 0.0, acc_fb = 0.0;
  const bool lane_sums = p.mode == MODE_RUN && p.nbins == 0 && !(p.cutoff > 0.0);
  // Work items and the window scan as in kd_wsplit.hip: one compacted point per
  // item, numbered slot-major (item = grp * nbatch + b), every gridDim.x-th
  // item to this wave; lane l of a window loads the batch count of the
  // window's item l and a ballot keeps the live ones.
  const int G = (int)gridDim.x;
  int wgrp = nbatch > 0 ? (int)(blockIdx.x / (unsigned)nbatch) : 64;
  int wb = (int)blockIdx.x - wgrp * nbatch;
  uint64_t mask = 0;
  int gl = 0, bl = 0;
  double u_n = 0.0, jac_n = 0.0;
  for (;;) {
    if (!mask) {
      if (wgrp >= 64) break;
      const unsigned bw = (unsigned)(wb + lane * G);
      gl = wgrp + (int)(bw / (unsigned)nbatch);
      bl = (int)(bw % (unsigned)nbatch);
      mask = __ballot(gl < 64 && gl < buf.cnt[bl]);
      const unsigned adv = (unsigned)(wb + 64 * G);
      wgrp += (int)(adv / (unsigned)nbatch);
      wb = (int)(adv % (unsigned)nbatch);
      if (!mask) continue;
      const int j1 = __builtin_ctzll(mask);
      const int s1 = rdlane(bl, j1) * 64 + rdlane(gl, j1);
      u_n = buf.u[s1 * N + (lane < N ? lane : N - 1)];
      jac_n = buf.jac[s1];
    }
    const int j = __builtin_ctzll(mask);
    mask &= mask - 1;
    const int g = rdlane(gl, j);
    const int b = rdlane(bl, j);
    const int slot = b * 64 + g;
    const double uc = u_n, jj = jac_n;
    int lq = lane;
    __asm__ volatile("" : "+v"(lq));
    if (lq < N) Us[lq] = uc;
    if (lq == 0) {
      dpath[1] = make_double2(1.0, 0.0);
      fpath[1] = 0;
    }
    wave_sync();
    // the next live item's record, in flight until the next point
    if (mask) {
      const int j2 = __builtin_ctzll(mask);
      const int s2 = rdlane(bl, j2) * 64 + rdlane(gl, j2);
      u_n = buf.u[s2 * N + (lane < N ? lane : N - 1)];
      jac_n = buf.jac[s2];
    }
    // sort by (time, index): lane a N + c asks whether vertex c comes before
    // vertex a, and a ballot gives each vertex its rank (NaN last, so that the
    // ranks stay a permutation and the lookups yield NaN)
    if constexpr (N * N <= 64) {
      const int a = lq / N, c = lq - a * N;
      const double ta = Us[a < N ? a : 0], tc = Us[c];
      const double ka = isnan(ta) ? __builtin_inf() : ta, kc = isnan(tc) ? __builtin_inf() : tc;
      const uint64_t before = __ballot(lq < N * N && (kc < ka || (kc == ka && c < a)));
      if (lq < N) Ts[__popcll((before >> (lq * N)) & ((1ull << N) - 1))] = uc;
    } else if (lq < N) { // above order 8: each vertex counts the ones before it
      const double kt = isnan(uc) ? __builtin_inf() : uc;
      int rank = 0;
#pragma unroll
      for (int q = 0; q < N; q++) {
        const double tq = Us[q];
        const double kq = isnan(tq) ? __builtin_inf() : tq;
        rank += (kq < kt) || (kq == kt && q < lq);
      }
      Ts[rank] = uc;
    }
    wave_sync();
    // Wv: item (a, c) = a N + c with a <= c, four table lookups at two time
    // differences.  a < c, the vertex pair (a earlier): g0^> and g0^< at
    // d = t_c - t_a (row c, column a: Keldysh index 0 / 1 of a), g0^< and g0^>
    // at -d (row a, column c).  a == c, vertex a's border: column (a s; B) at
    // t_a - t_B and row (A; a s) at t_A - t_a, the table of each s by
    // compare_3way (the vertex's split is below the external points' N, no
    // alpha: the points differ).  One instruction stream for both kinds.
    double wmax = 0.0;
    static_for<0, (N * N + 63) / 64>([&](auto RI) {
      const int q = lq + 64 * (int)RI;
      const int a = q / N, c = q - a * N;
      if (q < N * N && a <= c) {
        const bool pair = a < c;
        const double ta = Ts[a], tc = Ts[c];
        const double d1 = pair ? tc - ta : ta - p.B.time;
        const double d2 = pair ? ta - tc : p.A.time - ta;
        const TabPos p1 = table_pos(p, d1), p2 = table_pos(p, d2);
        // the border's contour order per Keldysh index s of the vertex
        const int kA = p.A.k_idx, kB = p.B.k_idx;
        const int sgC = d1 != 0.0 ? (d1 < 0.0 ? -1 : 1) : -1; // vertex vs B at equal k: B's split is larger
        const int sgR = d2 != 0.0 ? (d2 < 0.0 ? -1 : 1) : 1;  // A vs vertex at equal k
        const double2 *tab[4];
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const int ordC = h != kB ? h - kB : (1 - 2 * h) * sgC;
          const int ordR = kA != h ? kA - h : (1 - 2 * kA) * sgR;
          const double2 *bc = (ordC > 0 ? p.greater : p.lesser) + (size_t)p.B.orbital * p.n_pts;
          const double2 *br = (ordR > 0 ? p.greater : p.lesser) + (size_t)(p.A.orbital * p.n_orb) * p.n_pts;
          tab[h] = pair ? (h == 0 ? p.greater : p.lesser) : bc;
          tab[2 + h] = pair ? (h == 0 ? p.lesser : p.greater) : br;
        }
        double2 e0[4], e1[4];
#pragma unroll
        for (int h = 0; h < 4; h++) {
          const int i = h < 2 ? p1.i : p2.i;
          e0[h] = tab[h][i];
          e1[h] = tab[h][i + 1];
        }
        double2 v[4];
#pragma unroll
        for (int h = 0; h < 4; h++) v[h] = table_mix(e0[h], e1[h], h < 2 ? p1 : p2);
        // pair: (c, a, s) and (a, c, s); border: (a, e, s) and (e, a, s)
        const int i1 = pair ? c * WS + a : a * WS + N, i2 = pair ? a * WS + c : N * WS + a;
        Wv[i1 * 2 + 0] = v[0];
        Wv[i1 * 2 + 1] = v[1];
        Wv[i2 * 2 + 0] = v[2];
        Wv[i2 * 2 + 1] = v[3];
        double m = 0.0;
#pragma unroll
        for (int h = 0; h < 4; h++) m = fmax(m, cabs1<double>(v[h]));
        wmax = pair ? fmax(wmax, m) : wmax;
      }
    });
    wmax = fmax(wave_max(wmax), wconst);
    const T thr = PivotTol<T>::v * fmax((T)wmax, PivotTol<T>::floor);
    wave_sync();

    // levels 2 and 3 breadth first (level 2 in LB, level 3 in LA; level 1
    // formed from Wv inside the level-2 pass)
    clevel<N, 1, T, 64, 1, true>(nullptr, LB, Wv, D, invD, thr, dpath, fpath, lq);
    clevel<N, 2, T>(LB, LA, Wv, D, invD, thr, dpath, fpath, lq);

    // the lane's path: sigma_0..sigma_5 = bits of the lane
    double sre = 0.0, sim = 0.0;
    uint32_t bad = 0;
    {
      const int g3 = lq & 7;
      if constexpr (N <= 9) {
        LanePath<N, T, 3> lp;
        lp.P3 = LA + g3 * S::NS(3);
        double2 dp = dpath[8 + g3];
        int flag = fpath[8 + g3];
        static_for<3, 6>([&](auto KI) {
          constexpr int k = KI;
          if constexpr (k < N - 1) {
            const cx<T> piv = lp.template step<k>(Wv, D, (lq >> k) & 1);
            dp = dmul(dp, dcx<T>(piv));
            flag |= (int)!(cabs1<T>(piv) >= thr);
          }
        });
        if constexpr (N == 6) {
          // vertex 5 is the last: the lane's own configuration
          bool b;
          const double2 x = dcx<T>(leafx<N, T>(Wv, D, thr, lp.template at<5, 0, 0>(), lp.template at<5, 0, 1>(),
                                               lp.template at<5, 1, 0>(), lp.template at<5, 1, 1>(), (lq >> 5) & 1,
                                               b));
          b = b || flag;
          const double2 v = dmul(dmul(dp, dp), x);
          const double sg = (__popc(lq) & 1) ? 1.0 : -1.0;
          sre += b ? 0.0 : sg * v.x;
          sim += b ? 0.0 : sg * v.y;
          bad |= (uint32_t)b;
        } else {
          constexpr int R6 = S::R(6);
          cx<T> C6[R6][R6];
          static_for<0, R6>([&](auto A) {
            static_for<0, R6>([&](auto B) { C6[A][B] = lp.template at<6, A, B>(); });
          });
          subtree<N, R6, T>(C6, Wv, D, thr, dp, flag, lq, sre, sim, bad);
        }
      } else {
        // order 10: two passes over sigma_6, each recomputing the lane's path to
        // its level-7 node (measured faster than keeping the 5x5 level-6 table
        // through both passes: 18.6 against 20.4 ms per 2^21 points)
#pragma unroll 1
        for (int s6 = 0; s6 < 2; s6++) {
          LanePath<N, T, 4> lp;
          lp.P3 = LA + g3 * S::NS(3);
          double2 dp = dpath[8 + g3];
          int flag = fpath[8 + g3];
          static_for<3, 7>([&](auto KI) {
            constexpr int k = KI;
            const cx<T> piv = lp.template step<k>(Wv, D, k < 6 ? (lq >> k) & 1 : s6);
            dp = dmul(dp, dcx<T>(piv));
            flag |= (int)!(cabs1<T>(piv) >= thr);
          });
          constexpr int R7 = S::R(7);
          cx<T> C7[R7][R7];
          static_for<0, R7>([&](auto A) {
            static_for<0, R7>([&](auto B) { C7[A][B] = lp.template at<7, A, B>(); });
          });
          subtree<N, R7, T>(C7, Wv, D, thr, dp, flag, lq | (s6 << 6), sre, sim, bad);
        }
      }
    }
    // flagged configurations: partial-pivoting FP64 LU of their matrices
    int nfb = 0;
    while (bad) {
      const int jb = __builtin_ctz(bad);
      bad &= bad - 1;
      const int cfg = lq | (jb << 6);
      const double2 v = config_value_pivoted_sorted(Wv, cfg, N, Dd);
      const double sg = (__popc(cfg) & 1) ? -1.0 : 1.0;
      sre += sg * v.x;
      sim += sg * v.y;
      nfb++;
    }
    acc_fb += (double)nfb;
    if (lane_sums) {
      acc_re += sre * jj;
      acc_im += sim * jj;
      acc_dom += lane == 0 ? 1.0 : 0.0;
      wave_sync(); // Us, Ts, Wv and the tables are rewritten by the next point
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
          const int bi = bin_index(p, Ts[0]);
          if (bi >= 0) {
            sbins[bi].x += sre;
            sbins[bi].y += sim;
          }
        }
      } else {
        const int idx = b * 64 + buf.idx[slot];
        p.out_vals[2 * idx] = sre;
        p.out_vals[2 * idx + 1] = sim;
        if (p.out_dom) p.out_dom[idx] = 1;
        if (p.out_jac) p.out_jac[idx] = jj;
      }
    }
    wave_sync();
  }

  if (p.mode == MODE_RUN) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc_fb += __shfl_xor(acc_fb, off, 64);
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

// Orders 6..8: two points per wave, one per half (32 lanes).  A lane's bits
// are sigma_0..sigma_4 of its point: its path runs from its level-3 node through
// vertices 3 and 4 to a level-5 table, and its subtree covers vertices 5..N-1.
// Against one point per wave this halves the lanes that re-form the same path
// entries from the level-3 table (fewer VALU and LDS instructions per point) and
// shares every breadth-first pass between the two points.  Work items as in
// the one-point kernel, taken in pairs across windows, so that a half idles
// (on a copy of the other half's point, its results dropped) only for the last
// item of the wave's share.
template <int N, typename T>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void
integrand_ctree2_kernel(KdParams p, PrepBuf buf) {
  static_assert(N >= 6 && N <= 8, "two points per wave: orders 6..8");
  using S = CShape<N>;
  constexpr int WS = S::WS;
  constexpr int NW = WS * WS * 2, NA = S::sizeA(), NB = S::sizeB();
  __shared__ double2 Wv2[2 * NW];
  __shared__ cx<T> LA2[2 * NA], LB2[2 * NB];
  __shared__ double2 dpath2[2 * 16];
  __shared__ int fpath2[2 * 16];
  __shared__ double Us2[2 * N], Ts2[2 * N];
  extern __shared__ double2 sbins[];

  const int lane = threadIdx.x;
  if (p.nbins > 0) {
    for (int i = lane; i < p.nbins; i += 64) sbins[i] = make_double2(0.0, 0.0);
    wave_sync();
  }
  int lq = lane;
  __asm__ volatile("" : "+v"(lq));
  const int h = lq >> 5, l32 = lq & 31;
  const bool upper = h != 0;
  double2 *Wv = Wv2 + h * NW;
  cx<T> *LA = LA2 + h * NA, *LB = LB2 + h * NB;
  double2 *dpath = dpath2 + 16 * h;
  int *fpath = fpath2 + 16 * h;
  double *Us = Us2 + h * N, *Ts = Ts2 + h * N;

  const double2 gl0 = table_eval(p, p.lesser, 0, 0, 0.0), gg0 = table_eval(p, p.greater, 0, 0, 0.0);
  const double2 Dd = make_double2(gl0.x, gl0.y - p.alpha);
  const double wconst = fmax(fmax(cabs1<double>(gl0), cabs1<double>(gg0)), cabs1<double>(Dd));
  const cx<T> D = cmk<T>((T)Dd.x, (T)Dd.y);
  const cx<T> invD = cinv<T>(D);

  // a launch holds at most 2^21 points, so item, batch and slot indices fit 32 bits
  const int nbatch = (int)((p.n_points + 63) / 64);
  double acc_re = 0.0, acc_im = 0.0, acc_dom = 0.0, acc_fb = 0.0;
  const bool lane_sums = p.mode == MODE_RUN && p.nbins == 0 && !(p.cutoff > 0.0);
  const int G = (int)gridDim.x;
  int wgrp = nbatch > 0 ? (int)(blockIdx.x / (unsigned)nbatch) : 64;
  int wb = (int)blockIdx.x - wgrp * nbatch;
  uint64_t mask = 0;
  int gl = 0, bl = 0;
  // the wave's next live item (slot-major numbering, every G-th item), or false
  // when its share is done
  auto next_item = [&](int &slot) -> bool {
    while (!mask) {
      if (wgrp >= 64) return false;
      const unsigned bw = (unsigned)(wb + lane * G);
      gl = wgrp + (int)(bw / (unsigned)nbatch);
      bl = (int)(bw % (unsigned)nbatch);
      mask = __ballot(gl < 64 && gl < buf.cnt[bl]);
      const unsigned adv = (unsigned)(wb + 64 * G);
      wgrp += (int)(adv / (unsigned)nbatch);
      wb = (int)(adv % (unsigned)nbatch);
    }
    const int j = __builtin_ctzll(mask);
    mask &= mask - 1;
    slot = rdlane(bl, j) * 64 + rdlane(gl, j);
    return true;
  };
  int sa = 0, sb = 0;
  bool va = next_item(sa);
  bool vb = va && next_item(sb);
  double u_n = 0.0, jac_n = 0.0;
  if (va) {
    const int s1 = upper && vb ? sb : sa;
    u_n = buf.u[s1 * N + (l32 < N ? l32 : N - 1)];
    jac_n = buf.jac[s1];
  }
  while (va) {
    const bool valid = !upper || vb; // the upper half idles without a second item
    const int slot = upper && vb ? sb : sa;
    const double uc = u_n, jj = jac_n;
    if (l32 < N) Us[l32] = uc;
    if (l32 == 0) {
      dpath[1] = make_double2(1.0, 0.0);
      fpath[1] = 0;
    }
    wave_sync();
    // the next pair's records, in flight until the next iteration
    int na = 0, nb = 0;
    const bool nva = next_item(na);
    const bool nvb = nva && next_item(nb);
    if (nva) {
      const int s2 = upper && nvb ? nb : na;
      u_n = buf.u[s2 * N + (l32 < N ? l32 : N - 1)];
      jac_n = buf.jac[s2];
    }
    // sort by (time, index): each vertex counts the ones before it (NaN last,
    // so that the ranks stay a permutation and the lookups yield NaN)
    if (l32 < N) {
      const double kt = isnan(uc) ? __builtin_inf() : uc;
      int rank = 0;
#pragma unroll
      for (int q = 0; q < N; q++) {
        const double tq = Us[q];
        const double kq = isnan(tq) ? __builtin_inf() : tq;
        rank += (kq < kt) || (kq == kt && q < l32);
      }
      Ts[rank] = uc;
    }
    wave_sync();
    // Wv: item (a, c) = a N + c with a <= c, as in the one-point kernel
    double wmax = 0.0;
    static_for<0, (N * N + 31) / 32>([&](auto RI) {
      const int q = l32 + 32 * (int)RI;
      const int a = q / N, c = q - a * N;
      if (q < N * N && a <= c) {
        const bool pair = a < c;
        const double ta = Ts[a], tc = Ts[c];
        const double d1 = pair ? tc - ta : ta - p.B.time;
        const double d2 = pair ? ta - tc : p.A.time - ta;
        const TabPos p1 = table_pos(p, d1), p2 = table_pos(p, d2);
        const int kA = p.A.k_idx, kB = p.B.k_idx;
        const int sgC = d1 != 0.0 ? (d1 < 0.0 ? -1 : 1) : -1;
        const int sgR = d2 != 0.0 ? (d2 < 0.0 ? -1 : 1) : 1;
        const double2 *tab[4];
#pragma unroll
        for (int hh = 0; hh < 2; hh++) {
          const int ordC = hh != kB ? hh - kB : (1 - 2 * hh) * sgC;
          const int ordR = kA != hh ? kA - hh : (1 - 2 * kA) * sgR;
          const double2 *bc = (ordC > 0 ? p.greater : p.lesser) + (size_t)p.B.orbital * p.n_pts;
          const double2 *br = (ordR > 0 ? p.greater : p.lesser) + (size_t)(p.A.orbital * p.n_orb) * p.n_pts;
          tab[hh] = pair ? (hh == 0 ? p.greater : p.lesser) : bc;
          tab[2 + hh] = pair ? (hh == 0 ? p.lesser : p.greater) : br;
        }
        double2 e0[4], e1[4];
#pragma unroll
        for (int hh = 0; hh < 4; hh++) {
          const int i = hh < 2 ? p1.i : p2.i;
          e0[hh] = tab[hh][i];
          e1[hh] = tab[hh][i + 1];
        }
        double2 v[4];
#pragma unroll
        for (int hh = 0; hh < 4; hh++) v[hh] = table_mix(e0[hh], e1[hh], hh < 2 ? p1 : p2);
        const int i1 = pair ? c * WS + a : a * WS + N, i2 = pair ? a * WS + c : N * WS + a;
        Wv[i1 * 2 + 0] = v[0];
        Wv[i1 * 2 + 1] = v[1];
        Wv[i2 * 2 + 0] = v[2];
        Wv[i2 * 2 + 1] = v[3];
        double m = 0.0;
#pragma unroll
        for (int hh = 0; hh < 4; hh++) m = fmax(m, cabs1<double>(v[hh]));
        wmax = pair ? fmax(wmax, m) : wmax;
      }
    });
    wmax = fmax(half_max(wmax, upper), wconst);
    const T thr = PivotTol<T>::v * fmax((T)wmax, PivotTol<T>::floor);
    wave_sync();

    // levels 2 and 3 breadth first, each half on its point (level 1 formed
    // from Wv inside the level-2 pass; two rows per level-3 item: one pass)
    clevel<N, 1, T, 32, 1, true>(nullptr, LB, Wv, D, invD, thr, dpath, fpath, l32);
    clevel<N, 2, T, 32, 2>(LB, LA, Wv, D, invD, thr, dpath, fpath, l32);

    // the lane's path through vertices 3 and 4, then its subtree
    double sre = 0.0, sim = 0.0;
    uint32_t bad = 0;
    {
      const int g3 = l32 & 7;
      LanePath<N, T, 2> lp;
      lp.P3 = LA + g3 * S::NS(3);
      double2 dp = dpath[8 + g3];
      int flag = fpath[8 + g3];
      static_for<3, 5>([&](auto KI) {
        constexpr int k = KI;
        const cx<T> piv = lp.template step<k>(Wv, D, (l32 >> k) & 1);
        dp = dmul(dp, dcx<T>(piv));
        flag |= (int)!(cabs1<T>(piv) >= thr);
      });
      constexpr int R5 = S::R(5);
      cx<T> C5[R5][R5];
      static_for<0, R5>([&](auto A) {
        static_for<0, R5>([&](auto B) { C5[A][B] = lp.template at<5, A, B>(); });
      });
      subtree<N, R5, T, 5, (N >= 8)>(C5, Wv, D, thr, dp, flag, l32, sre, sim, bad);
    }
    if (!valid) {
      sre = 0.0;
      sim = 0.0;
      bad = 0;
    }
    // flagged configurations: partial-pivoting FP64 LU of their matrices
    int nfb = 0;
    while (bad) {
      const int jb = __builtin_ctz(bad);
      bad &= bad - 1;
      const int cfg = l32 | (jb << 5);
      const double2 v = config_value_pivoted_sorted(Wv, cfg, N, Dd);
      const double sg = (__popc(cfg) & 1) ? -1.0 : 1.0;
      sre += sg * v.x;
      sim += sg * v.y;
      nfb++;
    }
    acc_fb += (double)nfb;
    if (lane_sums) {
      acc_re += sre * jj;
      acc_im += sim * jj;
      acc_dom += l32 == 0 && valid ? 1.0 : 0.0;
    } else {
      sre = half_sum(sre, upper);
      sim = half_sum(sim, upper);
      if (l32 == 0 && valid) {
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
        } else {
          const int idx = (slot >> 6) * 64 + buf.idx[slot];
          p.out_vals[2 * idx] = sre;
          p.out_vals[2 * idx + 1] = sim;
          if (p.out_dom) p.out_dom[idx] = 1;
          if (p.out_jac) p.out_jac[idx] = jj;
        }
      }
      if (p.mode == MODE_RUN && p.nbins > 0) {
        // the halves one after the other: both points may fall in one bin
#pragma unroll
        for (int hh = 0; hh < 2; hh++) {
          if (lq == 32 * hh && valid) {
            const int bi = bin_index(p, Ts[0]);
            if (bi >= 0) {
              sbins[bi].x += sre;
              sbins[bi].y += sim;
            }
          }
          wave_sync();
        }
      }
    }
    wave_sync(); // Us, Ts, Wv and the tables are rewritten by the next pair
    va = nva;
    vb = nvb;
    sa = na;
    sb = nb;
  }

  if (p.mode == MODE_RUN) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc_fb += __shfl_xor(acc_fb, off, 64);
    acc_re = wave_sum(acc_re);
    acc_im = wave_sum(acc_im);
    acc_dom = wave_sum(acc_dom);
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
static hipError_t launch_ctree_n(const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st) {
  const size_t shm = p.nbins > 0 ? sizeof(double2) * (size_t)p.nbins : 0;
  if constexpr (N <= 7)
    hipLaunchKernelGGL((integrand_ctree2_kernel<N, T>), dim3(grid), dim3(64), shm, st, p, buf);
  else
    hipLaunchKernelGGL((integrand_ctree_kernel<N, T>), dim3(grid), dim3(64), shm, st, p, buf);
  return hipGetLastError();
}

bool ctree_supported(int order) { return order >= 6 && order <= 10; }
bool ctree_pairs(int order) { return order >= 6 && order <= 7; }

hipError_t launch_ctree(int order, int precision, const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st) {
  const bool f32 = lu_is_fp32(order, precision);
#define KD_CTREE_CASE(n)                                                                                          \
  case n:                                                                                                         \
    return f32 ? launch_ctree_n<n, float>(p, buf, grid, st) : launch_ctree_n<n, double>(p, buf, grid, st);
  switch (order) {
    KD_CTREE_CASE(6)
    KD_CTREE_CASE(7)
    KD_CTREE_CASE(8)
    KD_CTREE_CASE(9)
    KD_CTREE_CASE(10)
  default:
    return hipErrorInvalidValue;
  }
#undef KD_CTREE_CASE
}

} // namespace kd
