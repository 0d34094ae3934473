// Partitioned block Thomas (a SPIKE-type solver) for block-tridiagonal
// systems whose off-diagonal blocks are scalar multiples of the identity
// (the flamelet Laplacian couplings):
//   L[i-1] x_{i-1} + C_i x_i + U[i] x_{i+1} = r_i
// (L[g] couples row g+1 to x_g, U[g] couples row g to x_{g+1}, as in
// k_block_thomas<NV, true>).  The result is that of the block Thomas
// algorithm of pyrometheus/flamelets/linear_solver.py:6-103; only the order
// of elimination changes.
//
// Every system is cut into partitions [s, e) of a few dozen rows.  Each
// partition q is reduced by two sweeps at once (for NV <= 16 four sweeps
// share a wavefront, for larger NV each sweep is a workgroup of eight
// wavefronts):
//   down: rows s..e-1 with the coupling to x_{s-1} carried as NV extra
//         right-hand-side columns; gives x_{e-1} = a + B x_{s-1} + D x_e and
//         stores [X_U | X_r | X_W] of every row for the recovery,
//   up:   rows e-1..s with the coupling to x_e carried the same way; gives
//         x_s = a' + B' x_{s-1} + D' x_e.
// With b_q = x_{e-1} and t_q = x_s of partition q the two maps read
//   b_q = a_q  + B_q  b_{q-1} + D_q  t_{q+1}   (down map of q)
//   t_q = a'_q + B'_q b_{q-1} + D'_q t_{q+1}   (up map of q),
// a recursion over the partitions of a system with NV x NV blocks
// (k_spike_reduced) gives the interface unknowns Z_j = (b_q, t_{q+1}); then
// every partition recovers its rows from the stored factors:
//   x_i = X_r,i + X_W,i x_{s-1} - X_U,i x_{i+1}.
// The chain each wavefront walks is the partition length plus the number
// of partitions instead of the system length.
#pragma once

#include "pyro_block_thomas.hpp"

namespace pyro_spike {

using pyro_bt::bcast;

// Gauss-Jordan as pyro_bt::gauss_jordan_dpp for four independent systems
// per wave, one per 16-lane row: lanes base..base+NV-1 (base = 16 g) hold
// the rows of system g, gl = lane - base.  The pivot of each row of lanes
// comes from its own DPP max-reduction; the pivot row's entries are read
// with ds_bpermute from lane base + p, p differing between the systems.
template <int NV, int W>
__device__ __forceinline__ void gauss_jordan_grp(double (&a)[W], int lane,
                                                 int& my_k, bool& singular) {
  static_assert(NV <= 16, "one 16-lane row holds the rows");
  const int base = lane & 48;
  const int gl = lane - base;
  const bool active = gl < NV;
  {
    double mx = 0.0;
#pragma unroll
    for (int c = 0; c < NV; ++c) mx = fmax(mx, fabs(a[c]));
    const double sc = mx > 0.0 ? pyro::recip(mx) : 1.0;
#pragma unroll
    for (int c = 0; c < W; ++c) a[c] *= sc;
  }
  unsigned used = 0u;
  constexpr unsigned all = (1u << NV) - 1u;
  my_k = -1;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    // every lane of the 16-lane row holds the row's maximum key
    const unsigned best = pyro_bt::row16_max_u32(
        pyro_bt::pivot_key(a[k], active && !((used >> gl) & 1u), lane));
    const int p = best >= 64u ? 63 - static_cast<int>(best & 63u) - base
                              : __builtin_ctz(~used & all);
    used |= 1u << p;
    const int src = base + p;
    const double piv = __shfl(a[k], src);
    singular = singular || !(fabs(piv) > 0.0);
    const double inv = pyro::recip(piv);
    const double g = (gl == p) ? 1.0 - inv : a[k] * inv;
#pragma unroll
    for (int j = k + 1; j < W; ++j) a[j] = fma(-g, __shfl(a[j], src), a[j]);
    if (gl == p) my_k = k;
  }
}

// The partition sweeps for NV <= 16, four sweeps per wave: sweep task
// t = 4 blockIdx.x + lane / 16 is (partition t / 2, down when t is even),
// the same work as blockIdx t of k_spike_sweep_waves.  Consecutive tasks are the
// two sweeps of neighbouring partitions of one system, so the four sweeps
// of a wave have about the same length.
template <int NV>
__global__ void __launch_bounds__(64)
k_spike_sweep_grp(const double* __restrict__ L, const double* __restrict__ C,
                  const double* __restrict__ U, const double* __restrict__ R,
                  const int64_t* __restrict__ off,
                  const int64_t* __restrict__ pstart,
                  const int64_t* __restrict__ pend,
                  const int* __restrict__ psys, int64_t nparts,
                  double* __restrict__ fact, double* __restrict__ maps,
                  int* __restrict__ n_singular) {
  constexpr int W = 3 * NV + 1;
  constexpr int MW = 2 * NV + 1;
  const int lane = threadIdx.x;
  const int gl = lane & 15;
  const int64_t task = static_cast<int64_t>(blockIdx.x) * 4 + (lane >> 4);
  if (task >= 2 * nparts) return;  // whole 16-lane rows leave together
  const int64_t part = task >> 1;
  const bool down = (task & 1) == 0;
  const bool active = gl < NV;
  const int r = active ? gl : 0;
  const int64_t s = pstart[part], e = pend[part];
  const int sy = psys[part];
  const int64_t s0 = off[sy], s1 = off[sy + 1];
  const int64_t m = e - s;
  if (!down && s == s0) return;  // never used
  bool singular = false;
  auto row = [&](int64_t t) { return down ? s + t : e - 1 - t; };
  auto slot_coef = [&](int64_t i) -> double {
    if (down) return i < s1 - 1 ? U[i] : 0.0;
    return i > s0 ? L[i - 1] : 0.0;
  };
  auto back_coef = [&](int64_t i) -> double {
    return down ? L[i] : U[i - 1];
  };
  const int base = lane & 48;
  double a[W];
  int my_k = gl;
  {
    const int64_t i = row(0);
#pragma unroll
    for (int c = 0; c < NV; ++c) a[c] = C[(i * NV + r) * NV + c];
    const double sc = slot_coef(i);
#pragma unroll
    for (int c = 0; c < NV; ++c) a[NV + c] = (c == r) ? sc : 0.0;
    a[2 * NV] = R[i * NV + r];
    double ec = 0.0;
    if (down && s > s0) ec = -L[s - 1];
    if (!down && e < s1) ec = -U[e - 1];
#pragma unroll
    for (int c = 0; c < NV; ++c) a[2 * NV + 1 + c] = (c == r) ? ec : 0.0;
  }
  for (int64_t t = 0; t < m; ++t) {
    const int64_t i = row(t);
    double cpre[NV], rpre = 0.0, bpre = 0.0, spre = 0.0;
    if (t + 1 < m) {
      const int64_t i1 = row(t + 1);
#pragma unroll
      for (int c = 0; c < NV; ++c) cpre[c] = C[(i1 * NV + r) * NV + c];
      rpre = R[i1 * NV + r];
      bpre = back_coef(i);
      spre = slot_coef(i1);
    }
    gauss_jordan_grp<NV, W>(a, lane, my_k, singular);
    if (down && active) {
      double* f = fact + (i * NV + my_k) * MW;
#pragma unroll
      for (int c = 0; c < MW; ++c) f[c] = a[NV + c];
    }
    if (t + 1 == m) break;
    const int k = active ? my_k : 0;
    const int src = base + k;
    const double l = bpre;
#pragma unroll
    for (int c = 0; c < NV; ++c)
      a[c] = fma(-l, a[NV + c], __shfl(cpre[c], src));
    a[2 * NV] = fma(-l, a[2 * NV], __shfl(rpre, src));
#pragma unroll
    for (int c = 0; c < NV; ++c) a[2 * NV + 1 + c] *= -l;
#pragma unroll
    for (int c = 0; c < NV; ++c) a[NV + c] = (c == k) ? spre : 0.0;
  }
  if (active) {
    double* mp = maps + ((part * 2 + (down ? 0 : 1)) * NV + my_k) * MW;
    mp[0] = a[2 * NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      const double ext = a[2 * NV + 1 + c], nxt = -a[NV + c];
      mp[1 + c] = down ? ext : nxt;
      mp[1 + NV + c] = down ? nxt : ext;
    }
  }
  if (singular && gl == 0 && n_singular) atomicAdd(n_singular, 1);
}

// Interface unknowns of one system per wave, lane = row (NV <= 64).
// Forward over the system's partitions q = 0..Q-1:
//   b_0 = beta_0 + G_0 t_1 with beta_0 = a_0, G_0 = D_0;
//   q > 0: substituting b_{q-1} = beta_{q-1} + G_{q-1} t_q into the up map,
//     (I - B'_q G_{q-1}) t_q = a'_q + B'_q beta_{q-1} + D'_q t_{q+1},
//   solved by Gauss-Jordan for t_q = tau_q + H_q t_{q+1}; then the down map
//   gives b_q = beta_q + G_q t_{q+1} with P = B_q G_{q-1},
//     beta_q = a_q + B_q beta_{q-1} + P tau_q,  G_q = P H_q + D_q.
// Back: t_{Q-1} = tau_{Q-1}; for q = Q-2..0, b_q = beta_q + G_q t_{q+1},
//   t_q = tau_q + H_q t_{q+1}.  Z_j of the interface between q and q+1
//   receives (b_q, t_{q+1}).  red holds per partition and row
//   [beta | G row | tau | H row] (2 NV + 2 doubles).
template <int NV>
__global__ void __launch_bounds__(64)
k_spike_reduced(const double* __restrict__ maps,
                const int64_t* __restrict__ roff, double* __restrict__ red,
                double* __restrict__ Z, int* __restrict__ n_singular) {
  constexpr int MW = 2 * NV + 1;
  constexpr int N2 = 2 * NV;
  constexpr int RW = 2 * NV + 2;
  constexpr int WA = 2 * NV + 1;
  const int sy = blockIdx.x;
  const int lane = threadIdx.x;
  const bool active = lane < NV;
  const int r = active ? lane : 0;
  const int64_t j0 = roff[sy], j1 = roff[sy + 1];
  const int64_t nq = j1 - j0 + 1;   // partitions of this system
  if (nq < 2) return;
  const int64_t q0 = j0 + sy;       // its first partition
  // map row r of partition q: [a | B (on b_{q-1}) | D (on t_{q+1})]
  auto mrow = [&](int64_t q, int dir) {
    return maps + ((q * 2 + dir) * NV + r) * MW;
  };
  auto rrow = [&](int64_t q, int row) { return red + (q * NV + row) * RW; };
  bool singular = false;
  double beta, G[NV];
  {
    const double* md = mrow(q0, 0);
    beta = md[0];
#pragma unroll
    for (int c = 0; c < NV; ++c) G[c] = md[1 + NV + c];
    if (active) {
      double* o = rrow(q0, r);
      o[0] = beta;
#pragma unroll
      for (int c = 0; c < NV; ++c) o[1 + c] = G[c];
    }
  }
  for (int64_t q = q0 + 1; q < q0 + nq; ++q) {
    const double* mu = mrow(q, 1);
    const double* md = mrow(q, 0);
    // [I - B' G | a' + B' beta | D'] in row r
    double aug[WA];
#pragma unroll
    for (int c = 0; c < NV; ++c) aug[c] = (c == r) ? 1.0 : 0.0;
    double rhs = mu[0];
#pragma unroll
    for (int mm = 0; mm < NV; ++mm) {
      const double bp = mu[1 + mm];
      rhs = fma(bp, pyro_bt::bcast(beta, mm), rhs);
#pragma unroll
      for (int c = 0; c < NV; ++c)
        aug[c] = fma(-bp, pyro_bt::bcast(G[c], mm), aug[c]);
    }
    aug[NV] = rhs;
#pragma unroll
    for (int c = 0; c < NV; ++c) aug[NV + 1 + c] = mu[1 + NV + c];
    // P = B G_{q-1} (row r) before G is replaced
    double P[NV], bsum = md[0];
#pragma unroll
    for (int c = 0; c < NV; ++c) P[c] = 0.0;
#pragma unroll
    for (int mm = 0; mm < NV; ++mm) {
      const double bm = md[1 + mm];
      bsum = fma(bm, pyro_bt::bcast(beta, mm), bsum);
#pragma unroll
      for (int c = 0; c < NV; ++c)
        P[c] = fma(bm, pyro_bt::bcast(G[c], mm), P[c]);
    }
    int my_k, perm[NV];
    pyro_bt::gauss_jordan_dpp<NV, WA>(aug, lane, my_k, perm, singular);
    // lane perm[m] holds row m of [tau | H]
    if (active) {
      double* o = rrow(q, my_k);
      o[1 + NV] = aug[NV];
#pragma unroll
      for (int c = 0; c < NV; ++c) o[2 + NV + c] = aug[NV + 1 + c];
    }
    // G_{q-1} is no longer needed: G_q = D + P H accumulates in place
#pragma unroll
    for (int c = 0; c < NV; ++c) G[c] = md[1 + NV + c];
#pragma unroll
    for (int mm = 0; mm < NV; ++mm) {
      const int src = perm[mm];
      const double pm = P[mm];
      bsum = fma(pm, pyro_bt::bcast(aug[NV], src), bsum);
#pragma unroll
      for (int c = 0; c < NV; ++c)
        G[c] = fma(pm, pyro_bt::bcast(aug[NV + 1 + c], src), G[c]);
    }
    beta = bsum;
    if (active) {
      double* o = rrow(q, r);
      o[0] = beta;
#pragma unroll
      for (int c = 0; c < NV; ++c) o[1 + c] = G[c];
    }
  }
  // back substitution; t in natural row order, lane r holds t[r]
  const int64_t ql = q0 + nq - 1;
  double t = rrow(ql, r)[1 + NV];
  for (int64_t q = ql - 1; q >= q0; --q) {
    const double* o = rrow(q, r);
    double b = o[0];
#pragma unroll
    for (int mm = 0; mm < NV; ++mm) b = fma(o[1 + mm], pyro_bt::bcast(t, mm), b);
    const int64_t j = j0 + (q - q0);
    if (active) {
      Z[j * N2 + r] = b;
      Z[j * N2 + NV + r] = t;
    }
    if (q > q0) {
      double tn = o[1 + NV];
#pragma unroll
      for (int mm = 0; mm < NV; ++mm)
        tn = fma(o[2 + NV + mm], pyro_bt::bcast(t, mm), tn);
      t = tn;
    }
  }
  if (singular && lane == 0 && n_singular) atomicAdd(n_singular, 1);
}

// Recovery: x_i = X_r,i + X_W,i x_{s-1} - X_U,i x_{i+1}, from the last row
// of the partition up, with x_{s-1} = b_{q-1} and x_e = t_{q+1} taken from
// the interface solution Z (jbefore/jafter: interface index or -1).
template <int NV>
__global__ void __launch_bounds__(64)
k_spike_recover(const double* __restrict__ fact,
                const int64_t* __restrict__ pstart,
                const int64_t* __restrict__ pend,
                const int64_t* __restrict__ jbefore,
                const int64_t* __restrict__ jafter,
                const double* __restrict__ Z, double* __restrict__ X) {
  constexpr int MW = 2 * NV + 1;
  constexpr int N2 = 2 * NV;
  const int part = blockIdx.x;
  const int lane = threadIdx.x;
  const bool active = lane < NV;
  const int r = active ? lane : 0;
  const int64_t s = pstart[part], e = pend[part];
  const int64_t jb = jbefore[part], ja = jafter[part];
  double xs[NV], xe[NV];
#pragma unroll
  for (int m = 0; m < NV; ++m) {
    xs[m] = jb >= 0 ? Z[jb * N2 + m] : 0.0;        // b_{q-1}
    xe[m] = ja >= 0 ? Z[ja * N2 + NV + m] : 0.0;   // t_{q+1}
  }
  double xn[NV];
#pragma unroll
  for (int m = 0; m < NV; ++m) xn[m] = xe[m];
  double fc[MW];
  {
    const double* f = fact + ((e - 1) * NV + r) * MW;
#pragma unroll
    for (int c = 0; c < MW; ++c) fc[c] = f[c];
  }
  for (int64_t i = e - 1; i >= s; --i) {
    double fnext[MW];
    if (i > s) {
      const double* f = fact + ((i - 1) * NV + r) * MW;
#pragma unroll
      for (int c = 0; c < MW; ++c) fnext[c] = f[c];
    }
    double xr = fc[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      xr = fma(fc[NV + 1 + m], xs[m], xr);
      xr = fma(-fc[m], xn[m], xr);
    }
    if (active) X[i * NV + r] = xr;
#pragma unroll
    for (int m = 0; m < NV; ++m) xn[m] = bcast(xr, m);
    if (i > s) {
#pragma unroll
      for (int c = 0; c < MW; ++c) fc[c] = fnext[c];
    }
  }
}

// ---- NV > 16: one workgroup of kWaves wavefronts per sweep or system ----
// The lanes hold the rows (lanes < NV), and every NV-wide block of columns
// is spread over the kWaves waves by column index: wave w holds the columns
// c = kWaves t + w (t < T) of each block; columns c >= NV hold zeros.  One
// extra column ("x", a right-hand side) is carried by wave 0 only; the other
// waves carry zeros there.
// Eight waves (two per SIMD): GRI-3.0 flamelet blocks (nv = 54) 10 %
// faster than four.
constexpr int kWaves = 8;

template <int NV>
struct Tiles {
  static constexpr int T = (NV + kWaves - 1) / kWaves;  // columns per wave
  static constexpr int LD = kWaves * T + 1;  // odd LDS row stride (doubles)
  static constexpr int XC = kWaves * T;      // LDS column of "x"
};

struct GjShared {
  double g[2][64];       // multipliers of pivot step k at [k & 1]
  double smax[kWaves][64];
  int p[2];              // pivot lane of step k at [k & 1]
  int singular;
};

// Gauss-Jordan with scaled partial pivoting as pyro_bt::gauss_jordan_dpp,
// for rows held one per lane and columns spread over the kWaves waves of a
// workgroup.  a0 is the block that holds the pivot columns, ab[NB] further
// blocks, x the extra column.  Pivot step k: the wave owning column k
// (k mod kWaves) finds the pivot lane p by a DPP max-reduction of the scaled
// |a0[k]| over the unused rows and writes the multipliers g (1 - 1/piv on
// the pivot row, a[k]/piv elsewhere) and p to LDS; after one barrier every
// wave updates its columns c > k (and a few dead columns c <= k, whose
// values are never read again) with one FMA each, the pivot row's entries
// read from lane p by v_readlane.  Double-buffered multipliers need one
// barrier per step.  On return lane l holds logical row my_k of [I | X].
template <int NV, int T, int NB>
__device__ __forceinline__ void gj_waves(double (&a0)[T], double (&ab)[NB][T],
                                         double& x, int lane, int wv,
                                         int& my_k, GjShared& sh) {
  static_assert(NV <= 64, "one wavefront holds the rows");
  const bool active = lane < NV;
  {
    double mx = 0.0;
#pragma unroll
    for (int t = 0; t < T; ++t) mx = fmax(mx, fabs(a0[t]));
    sh.smax[wv][lane] = mx;
    __syncthreads();
    mx = sh.smax[0][lane];
#pragma unroll
    for (int q = 1; q < kWaves; ++q) mx = fmax(mx, sh.smax[q][lane]);
    const double sc = mx > 0.0 ? pyro::recip(mx) : 1.0;
#pragma unroll
    for (int t = 0; t < T; ++t) a0[t] *= sc;
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int t = 0; t < T; ++t) ab[b][t] *= sc;
    x *= sc;
  }
  unsigned long long used = 0ull;
  const unsigned long long all = NV == 64 ? ~0ull : ((1ull << NV) - 1ull);
  my_k = -1;
  bool singular = false;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const int tk = k / kWaves;
    if (wv == k % kWaves) {
      const int p = pyro_bt::pivot_lane<NV>(
          pyro_bt::pivot_key(a0[tk], active && !((used >> lane) & 1ull), lane),
          used, all);
      const double piv = bcast(a0[tk], p);
      singular = singular || !(fabs(piv) > 0.0);
      const double inv = pyro::recip(piv);
      sh.g[k & 1][lane] = (lane == p) ? 1.0 - inv : a0[tk] * inv;
      if (lane == 0) sh.p[k & 1] = p;
    }
    __syncthreads();
    const int p = __builtin_amdgcn_readfirstlane(sh.p[k & 1]);
    const double g = sh.g[k & 1][lane];
    used |= 1ull << p;
    if (lane == p) my_k = k;
#pragma unroll
    for (int t = tk; t < T; ++t) a0[t] = fma(-g, bcast(a0[t], p), a0[t]);
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int t = 0; t < T; ++t) ab[b][t] = fma(-g, bcast(ab[b][t], p), ab[b][t]);
    x = fma(-g, bcast(x, p), x);
  }
  if (singular && lane == 0) sh.singular = 1;
}

// One sweep over a partition for NV > 16, one workgroup of kWaves waves per
// sweep task (blockIdx = 2 partition + (0 down, 1 up)).  Lane = logical row
// of [C' | slot | r' | E], spread over the waves (a0 = C', ab[0] = slot,
// ab[1] = E, x = r'): slot holds the coupling to the next row to be
// eliminated (the neighbour outside the partition at the last one), E the
// coefficients of the outside unknown met first (x_{s-1} down, x_e up).
// After each row's Gauss-Jordan the down sweep stores [X_U | X_r | X_W] of
// every row for the recovery; the last row gives the partition's map
//   down: x_{e-1} = X_r + X_W x_{s-1} - X_U x_e
//   up:   x_s     = X_r + X_V x_e     - X_L x_{s-1},
// stored as [a | B (coefficient of x_{s-1}) | D (coefficient of x_e)].
// k_spike_sweep_grp does the same for NV <= 16.
template <int NV>
__global__ void __launch_bounds__(64 * kWaves)
k_spike_sweep_waves(const double* __restrict__ L, const double* __restrict__ C,
                    const double* __restrict__ U, const double* __restrict__ R,
                    const int64_t* __restrict__ off,
                    const int64_t* __restrict__ pstart,
                    const int64_t* __restrict__ pend,
                    const int* __restrict__ psys, double* __restrict__ fact,
                    double* __restrict__ maps, int* __restrict__ n_singular) {
  constexpr int T = Tiles<NV>::T;
  constexpr int MW = 2 * NV + 1;
  __shared__ GjShared sh;
  const int64_t part = blockIdx.x >> 1;
  const bool down = (blockIdx.x & 1) == 0;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool active = lane < NV;
  const int r = active ? lane : 0;
  const int64_t s = pstart[part], e = pend[part];
  const int sy = psys[part];
  const int64_t s0 = off[sy], s1 = off[sy + 1];
  const int64_t m = e - s;
  if (!down && s == s0) return;  // never used; the whole workgroup leaves
  if (threadIdx.x == 0) sh.singular = 0;
  auto row = [&](int64_t t) { return down ? s + t : e - 1 - t; };
  auto slot_coef = [&](int64_t i) -> double {
    if (down) return i < s1 - 1 ? U[i] : 0.0;
    return i > s0 ? L[i - 1] : 0.0;
  };
  auto back_coef = [&](int64_t i) -> double {
    return down ? L[i] : U[i - 1];
  };
  auto col = [&](int t) { return kWaves * t + wv; };
  double a0[T], ab[2][T], x;
  int my_k = lane;
  {
    const int64_t i = row(0);
    const double sc = slot_coef(i);
    double ec = 0.0;
    if (down && s > s0) ec = -L[s - 1];
    if (!down && e < s1) ec = -U[e - 1];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const int c = col(t);
      a0[t] = c < NV ? C[(i * NV + r) * NV + c] : 0.0;
      ab[0][t] = (c == r) ? sc : 0.0;
      ab[1][t] = (c == r) ? ec : 0.0;
    }
    x = wv == 0 ? R[i * NV + r] : 0.0;
  }
  for (int64_t t = 0; t < m; ++t) {
    const int64_t i = row(t);
    double cpre[T], rpre = 0.0, bpre = 0.0, spre = 0.0;
    if (t + 1 < m) {
      const int64_t i1 = row(t + 1);
#pragma unroll
      for (int u = 0; u < T; ++u) {
        const int c = col(u);
        cpre[u] = c < NV ? C[(i1 * NV + r) * NV + c] : 0.0;
      }
      if (wv == 0) rpre = R[i1 * NV + r];
      bpre = back_coef(i);
      spre = slot_coef(i1);
    }
    gj_waves<NV, T, 2>(a0, ab, x, lane, wv, my_k, sh);
    if (down && active) {
      // [X_U | X_r | X_W] of logical row my_k
      double* f = fact + (i * NV + my_k) * MW;
#pragma unroll
      for (int u = 0; u < T; ++u) {
        const int c = col(u);
        if (c < NV) {
          f[c] = ab[0][u];
          f[NV + 1 + c] = ab[1][u];
        }
      }
      if (wv == 0) f[NV] = x;
    }
    if (t + 1 == m) break;
    const int k = active ? my_k : 0;
    const double l = bpre;
#pragma unroll
    for (int u = 0; u < T; ++u) {
      a0[u] = fma(-l, ab[0][u], __shfl(cpre[u], k));
      ab[1][u] *= -l;
      ab[0][u] = (col(u) == k) ? spre : 0.0;
    }
    x = fma(-l, x, __shfl(rpre, k));
  }
  if (active) {
    double* mp = maps + ((part * 2 + (down ? 0 : 1)) * NV + my_k) * MW;
    if (wv == 0) mp[0] = x;
#pragma unroll
    for (int u = 0; u < T; ++u) {
      const int c = col(u);
      if (c < NV) {
        const double ext = ab[1][u], nxt = -ab[0][u];
        mp[1 + c] = down ? ext : nxt;
        mp[1 + NV + c] = down ? nxt : ext;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && sh.singular && n_singular) atomicAdd(n_singular, 1);
}

// k_spike_reduced for NV > 16: one workgroup of kWaves waves per system.
// G_{q-1} (with beta_{q-1} in column XC), P = B G_{q-1} (with
// a + B beta_{q-1} in column XC) and H_q (with tau_q in column XC) live in
// LDS, row-major with an odd stride; the products read one operand as an
// LDS broadcast, the other from the lane's own row.  Per partition:
//   A: lane r, wave w: row r of [I - B' G | a' + B' beta | D'] in the
//      wave's columns, and row r of P to LDS;
//   B: Gauss-Jordan over the waves (gj_waves) gives [tau | H], stored by
//      logical row to LDS and to red;
//   C: row r of G_q = D + P H and beta_q = a + B beta + P tau, to LDS and
//      red.
// The back substitution is k_spike_reduced's, on wave 0.
template <int NV>
__global__ void __launch_bounds__(64 * kWaves)
k_spike_reduced_waves(const double* __restrict__ maps,
                      const int64_t* __restrict__ roff,
                      double* __restrict__ red, double* __restrict__ Z,
                      int* __restrict__ n_singular) {
  constexpr int T = Tiles<NV>::T;
  constexpr int LD = Tiles<NV>::LD;
  constexpr int XC = Tiles<NV>::XC;
  constexpr int MW = 2 * NV + 1;
  constexpr int N2 = 2 * NV;
  constexpr int RW = 2 * NV + 2;
  __shared__ double Gs[NV * LD], Ps[NV * LD], Hs[NV * LD];
  __shared__ GjShared sh;
  const int sy = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool active = lane < NV;
  const int r = active ? lane : 0;
  const int64_t j0 = roff[sy], j1 = roff[sy + 1];
  const int64_t nq = j1 - j0 + 1;
  if (nq < 2) return;  // the whole workgroup leaves
  const int64_t q0 = j0 + sy;
  auto mrow = [&](int64_t q, int dir) {
    return maps + ((q * 2 + dir) * NV + r) * MW;
  };
  auto rrow = [&](int64_t q, int row) { return red + (q * NV + row) * RW; };
  auto col = [&](int t) { return kWaves * t + wv; };
  // the padding columns NV..XC-1 of G and H stay zero
  for (int idx = threadIdx.x; idx < NV * LD; idx += 64 * kWaves) {
    Gs[idx] = 0.0;
    Hs[idx] = 0.0;
  }
  if (threadIdx.x == 0) sh.singular = 0;
  __syncthreads();
  {
    const double* md = mrow(q0, 0);
    if (active) {
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const int c = col(t);
        if (c < NV) {
          const double v = md[1 + NV + c];
          Gs[r * LD + c] = v;
          rrow(q0, r)[1 + c] = v;
        }
      }
      if (wv == 0) {
        Gs[r * LD + XC] = md[0];
        rrow(q0, r)[0] = md[0];
      }
    }
  }
  __syncthreads();
  for (int64_t q = q0 + 1; q < q0 + nq; ++q) {
    const double* mu = mrow(q, 1);
    const double* md = mrow(q, 0);
    double a0[T], ab[1][T], pp[T];
    double x = mu[0], bs = md[0];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      a0[t] = (col(t) == r) ? 1.0 : 0.0;
      pp[t] = 0.0;
    }
#pragma unroll 6
    for (int mm = 0; mm < NV; ++mm) {
      const double bp = mu[1 + mm], bm = md[1 + mm];
      const double* gm = Gs + mm * LD;
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const double gv = gm[col(t)];
        a0[t] = fma(-bp, gv, a0[t]);
        pp[t] = fma(bm, gv, pp[t]);
      }
      const double gb = gm[XC];
      x = fma(bp, gb, x);
      bs = fma(bm, gb, bs);
    }
    if (active) {
#pragma unroll
      for (int t = 0; t < T; ++t)
        if (col(t) < NV) Ps[r * LD + col(t)] = pp[t];
      if (wv == 0) Ps[r * LD + XC] = bs;
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const int c = col(t);
      ab[0][t] = (active && c < NV) ? mu[1 + NV + c] : 0.0;
      if (!active) a0[t] = 0.0;
    }
    if (wv != 0 || !active) x = 0.0;
    int my_k;
    gj_waves<NV, T, 1>(a0, ab, x, lane, wv, my_k, sh);
    if (active) {
      double* o = rrow(q, my_k);
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const int c = col(t);
        if (c < NV) {
          Hs[my_k * LD + c] = ab[0][t];
          o[2 + NV + c] = ab[0][t];
        }
      }
      if (wv == 0) {
        Hs[my_k * LD + XC] = x;
        o[1 + NV] = x;
      }
    }
    __syncthreads();
    double gq[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const int c = col(t);
      gq[t] = c < NV ? md[1 + NV + c] : 0.0;
    }
    double bq = Ps[r * LD + XC];
#pragma unroll 6
    for (int mm = 0; mm < NV; ++mm) {
      const double pm = Ps[r * LD + mm];
      const double* hm = Hs + mm * LD;
#pragma unroll
      for (int t = 0; t < T; ++t) gq[t] = fma(pm, hm[col(t)], gq[t]);
      bq = fma(pm, hm[XC], bq);
    }
    if (active) {
      double* o = rrow(q, r);
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const int c = col(t);
        if (c < NV) {
          Gs[r * LD + c] = gq[t];
          o[1 + c] = gq[t];
        }
      }
      if (wv == 0) {
        Gs[r * LD + XC] = bq;
        o[0] = bq;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && sh.singular && n_singular) atomicAdd(n_singular, 1);
  if (wv != 0) return;
  // back substitution on wave 0 (the rows of red were written by all the
  // waves before the last barrier); t in natural row order, lane r
  const int64_t ql = q0 + nq - 1;
  double t = rrow(ql, r)[1 + NV];
  for (int64_t q = ql - 1; q >= q0; --q) {
    const double* o = rrow(q, r);
    double b = o[0];
#pragma unroll
    for (int mm = 0; mm < NV; ++mm) b = fma(o[1 + mm], bcast(t, mm), b);
    const int64_t j = j0 + (q - q0);
    if (active) {
      Z[j * N2 + r] = b;
      Z[j * N2 + NV + r] = t;
    }
    if (q > q0) {
      double tn = o[1 + NV];
#pragma unroll
      for (int mm = 0; mm < NV; ++mm) tn = fma(o[2 + NV + mm], bcast(t, mm), tn);
      t = tn;
    }
  }
}

}  // namespace pyro_spike
