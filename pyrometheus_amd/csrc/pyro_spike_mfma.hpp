// The partition sweep of pyro_spike.hpp for 16 < NV <= 63 as a blocked
// Gauss-Jordan whose trailing updates run on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64).
//
// One workgroup of four waves per sweep task (two workgroups fit a CU).
// The augmented block row [C' | slot | E | x] is held as 16 x 16 tiles in
// the MFMA accumulator layout: 64 rows (NV used) by 192 columns, C' in
// columns 0..63, slot in 64..127, E in 128..191 with x in column 128 + NV;
// unused columns and rows are zero.  Wave w holds column tile w of each of
// the three blocks (C' tile w, slot tile w + 4, E tile w + 8).  In a tile,
// lane l and register r hold row (l / 16) + 4 r + 16 rt of row tile rt and
// column l % 16 (the accumulator layout of the 16x16x4 FP64 MFMA: A[i][k]
// from lane i + 16 k, B[k][j] from lane j + 16 k, D[(l/16)+4r][l%16] in
// lane l, register r).
//
// Per block row: each row is scaled so that its largest |C'| entry is 1;
// then the C' columns are eliminated in panels of eight.  The wave that
// holds a panel's columns copies them to one row per lane (through LDS) and
// reduces them alone by Gauss-Jordan with scaled partial pivoting and
// virtual pivoting, as pyro_bt::gauss_jordan_dpp (pivot from the DPP
// max-reduction of pivot_key, reciprocal by v_rcp_f64 and two Newton
// steps, multiplier 1 - 1/piv on the pivot row).  The eight elementary
// transforms T_k = I - g_k e_{p_k}^T compose to T = I - W E_P^T (W: 64 x 8,
// E_P: the pivot rows' unit vectors), accumulated as the panel is reduced
// (W[:, k] = g_k, W[:, c<k] -= g_k W[p_k, c]).  Every wave then applies T
// to its tiles, A <- A - W A[P, :], with two MFMAs per 16 x 16 tile: W is
// the A operand (from LDS), the panel's pivot rows of the tile the B
// operand (gathered through LDS).  The wave that holds the next panel
// updates that tile first and reduces the next panel before its other
// tiles, so one workgroup barrier per panel separates the panels.  The
// result is the same [I | X] as the unblocked elimination up to rounding;
// row q ends as logical row lk[q].
#pragma once

#include "pyro_spike.hpp"

namespace pyro_spike {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaWaves = 4;

__device__ __forceinline__ d4 mfma_f64(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Max over each 16-lane row of |v| (DPP rotations within the row).
__device__ __forceinline__ double row16_max_abs(double v) {
  v = fabs(v);
#define PYRO_ROR_FMAX(N)                                                    \
  {                                                                         \
    const long long b = __double_as_longlong(v);                            \
    const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b), 0x120 + N, \
                                            0xF, 0xF, false);               \
    const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32),      \
                                            0x120 + N, 0xF, 0xF, false);    \
    const double o = __longlong_as_double(                                  \
        (static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));    \
    v = fmax(v, o);                                                         \
  }
  PYRO_ROR_FMAX(8)
  PYRO_ROR_FMAX(4)
  PYRO_ROR_FMAX(2)
  PYRO_ROR_FMAX(1)
#undef PYRO_ROR_FMAX
  return v;
}

// Orders this wave's LDS accesses for the compiler (the hardware executes a
// wave's LDS instructions in order; wave-local data needs no barrier).
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

// Workgroup barrier for LDS data only: waits for this wave's LDS accesses,
// not for its global loads, so the next block's prefetch stays in flight.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Per-row tables are kept in the tile order (rt, g, r) of row
// q = 16 rt + g + 4 r, so that a lane's four rows of a row tile are
// contiguous.
__device__ __forceinline__ int trow(int q) {
  return (q & ~15) | ((q & 3) << 2) | ((q >> 2) & 3);
}

// LDS of one sweep workgroup (odd strides keep the lanes of an access on
// distinct banks).
struct MfmaShared {
  double W[2][64 * 9];    // panel transform, [row][k], by panel parity
  double R[4][8 * 17];    // per wave: the pivot rows of one tile, [k][col]
  double pnl[4][64 * 9];  // per wave: a panel as one row per lane, [row][c]
  double pf[4][64 * 17];  // per wave: the next block's C rows, [row][col]
  double xpf[64];         // the next block's right-hand side
  double smax[4][64];     // row maxima of the C' tiles (tile order)
  double scl[64];         // row scales (tile order)
  int kOf[2][64];         // pivot slot of each row in the panel or -1 (tile order)
  int lk[64];             // logical row of each row (tile order)
  int singular;
};

template <int NV>
__global__ void __launch_bounds__(64 * kMfmaWaves, 2)
k_spike_sweep_mfma(const double* __restrict__ L, const double* __restrict__ C,
                   const double* __restrict__ U, const double* __restrict__ R,
                   const int64_t* __restrict__ off,
                   const int64_t* __restrict__ pstart,
                   const int64_t* __restrict__ pend,
                   const int* __restrict__ psys, double* __restrict__ fact,
                   double* __restrict__ maps, int* __restrict__ n_singular) {
  static_assert(NV > 16 && NV <= 63, "x sits in column 128 + NV < 192");
  constexpr int MW = 2 * NV + 1;
  constexpr int NP = (NV + 7) / 8;  // panels
  constexpr int XW = NV / 16;       // the wave holding x (in its E tile)
  __shared__ MfmaShared sh;
  const int64_t part = blockIdx.x >> 1;
  const bool down = (blockIdx.x & 1) == 0;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = lane >> 4, cl = lane & 15;
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
  // this lane's column within each 64-column block (the same in C', slot
  // and E: tile w of each), and its rows
  const int c = 16 * wv + cl;
  auto rowq = [&](int rt, int r) { return 16 * rt + g + 4 * r; };

  // acc[0] C' tile, acc[1] slot tile, acc[2] E tile (with x)
  d4 acc[3][4];
  // block row 0: [C | diag(sc) | diag(ec) | r], logical row = row
  {
    const int64_t i = row(0);
    const double sc = slot_coef(i);
    double ec = 0.0;
    if (down && s > s0) ec = -L[s - 1];
    if (!down && e < s1) ec = -U[e - 1];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = rowq(rt, r);
        const bool live = q < NV;
        acc[0][rt][r] = (live && c < NV) ? C[(i * NV + q) * NV + c] : 0.0;
        acc[1][rt][r] = (live && c == q) ? sc : 0.0;
        acc[2][rt][r] = !live ? 0.0
                        : c == q ? ec : (c == NV ? R[i * NV + q] : 0.0);
      }
  }
  // the loop below reads acc with no wait for these loads
  __builtin_amdgcn_s_waitcnt(0);
  const unsigned long long all = (1ull << NV) - 1ull;
  bool singular = false;

  for (int64_t t = 0; t < m; ++t) {
    const int64_t i = row(t);
    // prefetch of the next block's C rows (by row index; permuted after
    // the elimination) and right-hand side
    d4 cpre[4] = {};
    double xpre = 0.0, bpre = 0.0, spre = 0.0;
    const bool more = t + 1 < m;
    if (more) {
      const int64_t i1 = row(t + 1);
      // every lane loads (from a clamped, valid address: 32-bit offsets
      // from the block's base); entries outside the block are masked where
      // they are used
      const double* cb = C + i1 * NV * NV;
      const int cc = c < NV ? c : NV - 1;
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int q = rowq(rt, r);
          cpre[rt][r] = cb[(q < NV ? q : NV - 1) * NV + cc];
        }
      // one lane per row: lane l fetches row l (used by the wave holding x)
      xpre = R[i1 * NV + (lane < NV ? lane : NV - 1)];
      bpre = back_coef(i);
      spre = slot_coef(i1);
    }

    // ---- row scales: the largest |C'| of each row -> 1 ------------------
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double mx = row16_max_abs(acc[0][rt][r]);
        if (cl == 0) sh.smax[wv][16 * rt + 4 * g + r] = mx;
      }
    lds_barrier();
    if (wv == 0) {
      const int q = lane;  // tile-order index
      const double mx = fmax(fmax(sh.smax[0][q], sh.smax[1][q]),
                             fmax(sh.smax[2][q], sh.smax[3][q]));
      sh.scl[q] = mx > 0.0 ? pyro::recip(mx) : 1.0;
    }
    lds_barrier();
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      d4 sc;
#pragma unroll
      for (int r = 0; r < 4; ++r) sc[r] = sh.scl[16 * rt + 4 * g + r];
#pragma unroll
      for (int j = 0; j < 3; ++j) acc[j][rt] *= sc;
    }

    // ---- blocked Gauss-Jordan over the C' columns -------------------------
    unsigned long long used = 0ull;
    // reduce panel p (held by wave p / 2) and publish its transform
    auto reduce_panel = [&](int p) {
      const int buf = p & 1;
      const int c0 = 8 * (p & 1);  // panel's first column in the tile
      const int np = NV - 8 * p < 8 ? NV - 8 * p : 8;
      double* pn_l = sh.pnl[wv];
      if (cl >= c0 && cl < c0 + 8) {
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            pn_l[rowq(rt, r) * 9 + (cl - c0)] = acc[0][rt][r];
      }
      wave_lds_fence();
      double pn[8], wr[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        pn[k] = pn_l[lane * 9 + k];
        wr[k] = 0.0;
      }
      const bool active = lane < NV;
      int kq = -1;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k < np) {
          const int pv = pyro_bt::pivot_lane<NV>(
              pyro_bt::pivot_key(pn[k], active && !((used >> lane) & 1ull),
                                 lane),
              used, all);
          used |= 1ull << pv;
          const double piv = bcast(pn[k], pv);
          singular = singular || !(fabs(piv) > 0.0);
          const double inv = pyro::recip(piv);
          const double gk = (lane == pv) ? 1.0 - inv : pn[k] * inv;
#pragma unroll
          for (int x = k + 1; x < 8; ++x)
            pn[x] = fma(-gk, bcast(pn[x], pv), pn[x]);
#pragma unroll
          for (int x = 0; x < k; ++x)
            wr[x] = fma(-gk, bcast(wr[x], pv), wr[x]);
          wr[k] = gk;
          if (lane == pv) kq = k;
        }
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) sh.W[buf][lane * 9 + k] = wr[k];
      sh.kOf[buf][trow(lane)] = kq;
      if (kq >= 0) sh.lk[trow(lane)] = 8 * p + kq;
    };
    // apply panel p's transform to tile j of this wave (the operands are
    // read from LDS here, so that nothing of panel p stays live in
    // registers while the next panel is reduced)
    auto apply_panel = [&](int p, int j) {
      const int buf = p & 1;
      double aw[4][2];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
          aw[rt][ks] = sh.W[buf][(16 * rt + cl) * 9 + 4 * ks + g];
      double* rl = sh.R[wv];
      const int np = NV - 8 * p < 8 ? NV - 8 * p : 8;
      if (np < 8 && g == 0) {
        for (int k = np; k < 8; ++k) rl[k * 17 + cl] = 0.0;
      }
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int k = sh.kOf[buf][16 * rt + 4 * g + r];
          if (k >= 0) rl[k * 17 + cl] = acc[j][rt][r];
        }
      wave_lds_fence();
      const double b0 = rl[g * 17 + cl], b1 = rl[(4 + g) * 17 + cl];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
        acc[j][rt] = mfma_f64(-aw[rt][0], b0, acc[j][rt]);
        acc[j][rt] = mfma_f64(-aw[rt][1], b1, acc[j][rt]);
      }
      wave_lds_fence();
    };

    if (wv == 0) reduce_panel(0);
    lds_barrier();
#pragma unroll 1
    for (int p = 0; p < NP; ++p) {
      used |= __ballot(sh.kOf[p & 1][trow(lane)] >= 0);
      // the C' tile: live while it holds a column >= 8 (p + 1); the wave
      // holding panel p + 1 reduces it right after updating that tile
      if (16 * wv + 16 > 8 * (p + 1)) apply_panel(p, 0);
      if (p + 1 < NP && wv == (p + 1) / 2) reduce_panel(p + 1);
      apply_panel(p, 1);
      apply_panel(p, 2);
      lds_barrier();
    }
    if (singular && lane == 0) sh.singular = 1;

    // ---- outputs of the block row ----------------------------------------
    int lkr[16];
#pragma unroll
    for (int x = 0; x < 16; ++x) lkr[x] = sh.lk[16 * (x >> 2) + 4 * g + (x & 3)];
    if (down) {
      // [X_U | X_r | X_W] of each logical row: slot -> columns 0..NV-1,
      // x -> NV, E -> NV+1..2NV
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (rowq(rt, r) >= NV) continue;
          double* f = fact + (i * NV + lkr[4 * rt + r]) * MW;
          if (c < NV) {
            f[c] = acc[1][rt][r];
            f[NV + 1 + c] = acc[2][rt][r];
          } else if (c == NV) {
            f[NV] = acc[2][rt][r];
          }
        }
    }
    if (!more) break;

    // ---- next block row: C_next[lk] - l X_slot | diag(spre) | -l E | ... --
    const double l = bpre;
    double* pf = sh.pf[wv];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) pf[rowq(rt, r) * 17 + cl] = cpre[rt][r];
    if (wv == XW) sh.xpf[lane] = xpre;
    wave_lds_fence();
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const bool live = rowq(rt, r) < NV;
        const int k = lkr[4 * rt + r];
        const double cn = (live && c < NV) ? pf[k * 17 + cl] : 0.0;
        acc[0][rt][r] = fma(-l, acc[1][rt][r], cn);
        acc[1][rt][r] = (live && c == k) ? spre : 0.0;
        if (wv == XW && c == NV)
          acc[2][rt][r] = live ? fma(-l, acc[2][rt][r], sh.xpf[k]) : 0.0;
        else
          acc[2][rt][r] *= -l;
      }
    wave_lds_fence();
    // every row of the next block is a pivot of one of its panels, so lk
    // is rewritten (after the next row-scale barriers) before it is read
  }
  // the partition's map from its last block row:
  //   down: x_{e-1} = X_r + X_W x_{s-1} - X_U x_e
  //   up:   x_s     = X_r + X_V x_e     - X_L x_{s-1}
  {
    int lkr[16];
#pragma unroll
    for (int x = 0; x < 16; ++x) lkr[x] = sh.lk[16 * (x >> 2) + 4 * g + (x & 3)];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (rowq(rt, r) >= NV || c > NV) continue;
        double* mp = maps + ((part * 2 + (down ? 0 : 1)) * NV +
                             lkr[4 * rt + r]) * MW;
        if (c == NV) {
          mp[0] = acc[2][rt][r];
        } else {
          // E: coefficient of the outside unknown; slot: minus the next's
          mp[(down ? 1 : 1 + NV) + c] = acc[2][rt][r];
          mp[(down ? 1 + NV : 1) + c] = -acc[1][rt][r];
        }
      }
  }
  __syncthreads();
  if (threadIdx.x == 0 && sh.singular && n_singular) atomicAdd(n_singular, 1);
}

// k_spike_reduced_waves with its three nv x nv products on the FP64 matrix
// cores: per partition q of a system,
//   A: [Q | .] = B'_q [G_{q-1} | beta_{q-1}] and [P | .] = B_q [G | beta]
//      (16 output tiles each, 14 k-steps of v_mfma_f64_16x16x4f64; B'_q and
//      B_q staged in LDS as the A operands, G read from LDS as the B
//      operand), then the Gauss-Jordan system [I - Q | D'_q | a' + B' beta]
//      row by row as before (gj_waves, eight waves);
//   C: [G_q | beta_q] = [D_q | a_q + B beta] + P [H_q | tau_q] (P from LDS
//      as the A operand, H from LDS as the B operand).
// The same recursion as k_spike_reduced_waves up to rounding.
template <int NV>
__global__ void __launch_bounds__(64 * kWaves)
k_spike_reduced_mfma(const double* __restrict__ maps,
                     const int64_t* __restrict__ roff,
                     double* __restrict__ red, double* __restrict__ Z,
                     int* __restrict__ n_singular) {
  static_assert(NV > 16 && NV <= 64, "rows on lanes, 64-row tiles");
  constexpr int T = Tiles<NV>::T;
  constexpr int LD = Tiles<NV>::LD;
  constexpr int XC = Tiles<NV>::XC;
  static_assert(XC < 64, "beta and tau sit in the fourth column tile");
  constexpr int MW = 2 * NV + 1;
  constexpr int N2 = 2 * NV;
  constexpr int RW = 2 * NV + 2;
  constexpr int KS = (NV + 3) / 4;  // k-steps of 4
  constexpr int LB = NV + 1;        // LDS row stride of B' and B (odd)
  __shared__ double Gs[NV * LD], Ps[NV * LD], Hs[NV * LD];
  __shared__ double Bp[NV * LB], Bm[NV * LB];
  __shared__ GjShared sh;
  const int sy = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = lane >> 4, cl = lane & 15;
  const bool active = lane < NV;
  const int r = active ? lane : 0;
  const int64_t j0 = roff[sy], j1 = roff[sy + 1];
  const int64_t nq = j1 - j0 + 1;
  if (nq < 2) return;  // the whole workgroup leaves
  const int64_t q0 = j0 + sy;
  auto mrow = [&](int64_t q, int dir, int row) {
    return maps + ((q * 2 + dir) * NV + row) * MW;
  };
  auto rrow = [&](int64_t q, int row) { return red + (q * NV + row) * RW; };
  auto col = [&](int t) { return kWaves * t + wv; };
  // the product tiles of this wave: column tile ct, row tiles rt0, rt0 + 1
  const int ct = wv & 3;
  const int rt0 = 2 * (wv >> 2);
  const int tc = 16 * ct + cl;  // this lane's column in the product tiles
  for (int idx = threadIdx.x; idx < NV * LD; idx += 64 * kWaves) {
    Gs[idx] = 0.0;
    Hs[idx] = 0.0;
    Ps[idx] = 0.0;
  }
  if (threadIdx.x == 0) sh.singular = 0;
  __syncthreads();
  {
    const double* md = mrow(q0, 0, r);
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
    // ---- stage B'_q and B_q (rows of the up and down maps) in LDS --------
    for (int idx = threadIdx.x; idx < NV * NV; idx += 64 * kWaves) {
      const int row = idx / NV, k = idx - row * NV;
      Bp[row * LB + k] = mrow(q, 1, row)[1 + k];
      Bm[row * LB + k] = mrow(q, 0, row)[1 + k];
    }
    __syncthreads();
    // ---- A: Q = B' [G | beta], P = B [G | beta] on the matrix cores -------
    {
      d4 qa[2] = {}, pa[2] = {};
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int k = 4 * ks + g;
        const double b = (k < NV && tc <= XC) ? Gs[k * LD + tc] : 0.0;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int i = 16 * (rt0 + u) + cl;
          const bool ok = i < NV && k < NV;
          const double ap = ok ? Bp[i * LB + k] : 0.0;
          const double am = ok ? Bm[i * LB + k] : 0.0;
          qa[u] = mfma_f64(ap, b, qa[u]);
          pa[u] = mfma_f64(am, b, pa[u]);
        }
      }
      // Q to Hs (H of the previous partition is no longer needed), P and
      // a + B beta to Ps
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const int row = 16 * (rt0 + u) + g + 4 * jj;
          if (row < NV && (tc < NV || tc == XC)) {
            Hs[row * LD + tc] = qa[u][jj];
            Ps[row * LD + tc] =
                tc == XC ? mrow(q, 0, row)[0] + pa[u][jj] : pa[u][jj];
          }
        }
    }
    __syncthreads();
    double a0[T], ab[1][T];
    double x = mrow(q, 1, r)[0];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const int c = col(t);
      a0[t] = (active && c < NV) ? ((c == r) ? 1.0 : 0.0) - Hs[r * LD + c]
                                 : 0.0;
      ab[0][t] = (active && c < NV) ? mrow(q, 1, r)[1 + NV + c] : 0.0;
    }
    x = (wv == 0 && active) ? x + Hs[r * LD + XC] : 0.0;
    __syncthreads();  // every wave has read Q before H overwrites it
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
    // ---- C: [G_q | beta_q] = [D | a + B beta] + P [H | tau] -------------
    {
      d4 ga[2];
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const int row = 16 * (rt0 + u) + g + 4 * jj;
          double v = 0.0;
          if (row < NV) {
            if (tc < NV) v = mrow(q, 0, row)[1 + NV + tc];
            if (tc == XC) v = Ps[row * LD + XC];
          }
          ga[u][jj] = v;
        }
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int k = 4 * ks + g;
        const double b = (k < NV && tc <= XC) ? Hs[k * LD + tc] : 0.0;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int i = 16 * (rt0 + u) + cl;
          const double a = (i < NV && k < NV) ? Ps[i * LD + k] : 0.0;
          ga[u] = mfma_f64(a, b, ga[u]);
        }
      }
      __syncthreads();  // every wave has read G_{q-1}'s products (P, H)
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const int row = 16 * (rt0 + u) + g + 4 * jj;
          if (row < NV && (tc < NV || tc == XC)) {
            Gs[row * LD + tc] = ga[u][jj];
            rrow(q, row)[tc == XC ? 0 : 1 + tc] = ga[u][jj];
          }
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

// Form for 16 < NV <= 63: 1 the blocked MFMA sweep and the MFMA recursion
// above (default), 0 k_spike_sweep_waves and k_spike_reduced_waves.
inline int& sweep_form() {
  static int form = 1;
  return form;
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
    bool blocked = false;
    if constexpr (NV <= 63) {
      if (sweep_form() == 1) {
        blocked = true;
        k_spike_sweep_mfma<NV>
            <<<static_cast<unsigned>(2 * nparts), 64 * kMfmaWaves, 0,
               stream>>>(
                L, C, U, R, off, pstart, pend, psys, fact, maps, n_singular);
      }
    }
    if (!blocked)
      k_spike_sweep_waves<NV>
          <<<static_cast<unsigned>(2 * nparts), 64 * kWaves, 0, stream>>>(
              L, C, U, R, off, pstart, pend, psys, fact, maps, n_singular);
  }
  if (nj > 0) {
    if constexpr (NV <= 16)
      k_spike_reduced<NV><<<static_cast<unsigned>(nsys), 64, 0, stream>>>(
          maps, roff, red, Z, n_singular);
    else {
      bool mfma = false;
      if constexpr (NV <= 63) {
        if (sweep_form() == 1) {
          mfma = true;
          k_spike_reduced_mfma<NV>
              <<<static_cast<unsigned>(nsys), 64 * kWaves, 0, stream>>>(
                  maps, roff, red, Z, n_singular);
        }
      }
      if (!mfma)
        k_spike_reduced_waves<NV>
            <<<static_cast<unsigned>(nsys), 64 * kWaves, 0, stream>>>(
                maps, roff, red, Z, n_singular);
    }
  }
  k_spike_recover<NV><<<static_cast<unsigned>(nparts), 64, 0, stream>>>(
      fact, pstart, pend, jbefore, jafter, Z, X);
  return pyro::launch_status();
}

}  // namespace pyro_spike
