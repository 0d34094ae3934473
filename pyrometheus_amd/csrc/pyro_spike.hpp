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
// partition q is reduced by two wavefronts at once:
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

// One sweep over a partition, one wavefront, lane = logical row of
// [C' | slot | r' | E] (W = 3 NV + 1 doubles): slot holds the coupling to
// the next row to be eliminated (the neighbour outside the partition at
// the last one), E the coefficients of the outside unknown met first.
template <int NV>
__global__ void __launch_bounds__(64)
k_spike_sweep(const double* __restrict__ L, const double* __restrict__ C,
              const double* __restrict__ U, const double* __restrict__ R,
              const int64_t* __restrict__ off,
              const int64_t* __restrict__ pstart,
              const int64_t* __restrict__ pend,
              const int* __restrict__ psys, double* __restrict__ fact,
              double* __restrict__ maps, int* __restrict__ n_singular) {
  constexpr int W = 3 * NV + 1;
  constexpr int MW = 2 * NV + 1;  // [a | B | D] per logical row
  const int part = blockIdx.x >> 1;
  const bool down = (blockIdx.x & 1) == 0;
  const int lane = threadIdx.x;
  const bool active = lane < NV;
  const int r = active ? lane : 0;
  const int64_t s = pstart[part], e = pend[part];
  const int sy = psys[part];
  const int64_t s0 = off[sy], s1 = off[sy + 1];
  const int64_t m = e - s;
  // the up map of a system's first partition is never used
  if (!down && s == s0) return;
  bool singular = false;
  // row index of step t, its coupling to the next row of the sweep (the
  // "slot"), and the next row's coupling back to it
  auto row = [&](int64_t t) { return down ? s + t : e - 1 - t; };
  auto slot_coef = [&](int64_t i) -> double {
    if (down) return i < s1 - 1 ? U[i] : 0.0;
    return i > s0 ? L[i - 1] : 0.0;
  };
  auto back_coef = [&](int64_t i) -> double {  // next row's coupling to row i
    return down ? L[i] : U[i - 1];
  };
  double a[W];
  int perm[NV];
  int my_k = lane;
  {
    const int64_t i = row(0);
#pragma unroll
    for (int c = 0; c < NV; ++c) a[c] = C[(i * NV + r) * NV + c];
    const double sc = slot_coef(i);
#pragma unroll
    for (int c = 0; c < NV; ++c) a[NV + c] = (c == r) ? sc : 0.0;
    a[2 * NV] = R[i * NV + r];
    // the outside unknown met first: x_{s-1} (down) or x_e (up)
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
    pyro_bt::gauss_jordan_dpp<NV, W>(a, lane, my_k, perm, singular);
    if (down && active) {
      // [X_U | X_r | X_W] of logical row my_k
      double* f = fact + (i * NV + my_k) * MW;
#pragma unroll
      for (int c = 0; c < MW; ++c) f[c] = a[NV + c];
    }
    if (t + 1 == m) break;
    const int k = active ? my_k : 0;
    const double l = bpre;
#pragma unroll
    for (int c = 0; c < NV; ++c)
      a[c] = fma(-l, a[NV + c], __shfl(cpre[c], k));
    a[2 * NV] = fma(-l, a[2 * NV], __shfl(rpre, k));
#pragma unroll
    for (int c = 0; c < NV; ++c) a[2 * NV + 1 + c] *= -l;
#pragma unroll
    for (int c = 0; c < NV; ++c) a[NV + c] = (c == k) ? spre : 0.0;
  }
  // map of the last row of the sweep: x = X_r + X_E x_first - X_slot x_next
  //   down: x_{e-1} = X_r + X_W x_{s-1} - X_U x_e
  //   up:   x_s     = X_r + X_V x_e     - X_L x_{s-1}
  // stored as [a | B (coeff of x_{s-1}) | D (coeff of x_e)]
  if (active) {
    double* mp = maps + ((static_cast<int64_t>(part) * 2 + (down ? 0 : 1)) *
                             NV + my_k) * MW;
    mp[0] = a[2 * NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      const double ext = a[2 * NV + 1 + c], nxt = -a[NV + c];
      mp[1 + c] = down ? ext : nxt;
      mp[1 + NV + c] = down ? nxt : ext;
    }
  }
  if (singular && lane == 0 && n_singular) atomicAdd(n_singular, 1);
}

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
    double key = -1.0;
    if (active && !((used >> gl) & 1u)) {
      const double v = fabs(a[k]);
      key = (v == v) ? v : -0.5;  // a NaN row is taken only as a last resort
    }
    long long kb = __double_as_longlong(key);
    if (key >= 0.0) kb = (kb & ~0x3Fll) | (63 - lane);
    // every lane of the 16-lane row holds the row's maximum
    const double best = pyro_bt::row16_max(__longlong_as_double(kb));
    int p;
    if (best >= 0.0)
      p = 63 - static_cast<int>(__double_as_longlong(best) & 0x3F) - base;
    else
      p = __builtin_ctz(~used & all);
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

// k_spike_sweep for NV <= 16 with four sweeps per wave: sweep task
// t = 4 blockIdx.x + lane / 16 is (partition t / 2, down when t is even),
// the same work as blockIdx t of k_spike_sweep.  Consecutive tasks are the
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

// Host side: the three launches.  All arrays are device arrays prepared by
// the caller (see pyro_flamelet_thomas_partitioned).
template <int NV>
int solve(const double* L, const double* C, const double* U, const double* R,
          double* X, const int64_t* off, const int64_t* pstart,
          const int64_t* pend, const int* psys, int64_t nparts,
          const int64_t* jbefore, const int64_t* jafter, const int64_t* roff,
          int64_t nsys, int64_t nj, double* fact, double* maps, double* red,
          double* Z, int* n_singular, hipStream_t stream) {
  if constexpr (NV <= 16) {
    k_spike_sweep_grp<NV>
        <<<static_cast<unsigned>((2 * nparts + 3) / 4), 64, 0, stream>>>(
            L, C, U, R, off, pstart, pend, psys, nparts, fact, maps,
            n_singular);
  } else {
    k_spike_sweep<NV><<<static_cast<unsigned>(2 * nparts), 64, 0, stream>>>(
        L, C, U, R, off, pstart, pend, psys, fact, maps, n_singular);
  }
  if (nj > 0)
    k_spike_reduced<NV><<<static_cast<unsigned>(nsys), 64, 0, stream>>>(
        maps, roff, red, Z, n_singular);
  k_spike_recover<NV><<<static_cast<unsigned>(nparts), 64, 0, stream>>>(
      fact, pstart, pend, jbefore, jafter, Z, X);
  return pyro::launch_status();
}

}  // namespace pyro_spike
