// Batched block-tridiagonal solver (block Thomas algorithm) for gfx950.
//
// Solves, for each system s of a batch with ragged sizes N_s,
//   C_0 x_0 + U_0 x_1 = r_0
//   L_{i-1} x_{i-1} + C_i x_i + U_i x_{i+1} = r_i      (0 < i < N-1)
//   L_{N-2} x_{N-2} + C_{N-1} x_{N-1} = r_{N-1}
// with NV x NV blocks: the algorithm of pyrometheus/flamelets/
// linear_solver.py:6-103 (forward elimination C'_i = C_i - L_{i-1} C'^-1 U,
// r'_i = r_i - L_{i-1} C'^-1 r', then back substitution).
//
// Mapping: one wave64 per system.  Lane l (< NV) owns one logical row of the
// augmented block [C' | U | r'] (2 NV + 1 doubles in registers).  Each block
// row is reduced by Gauss-Jordan elimination with scaled partial pivoting in
// which the pivot row is never moved: the lane chosen as pivot for column k
// simply takes logical row k ("virtual pivoting"), the pivot row is
// broadcast with v_readlane, and every lane eliminates with its own
// column-k entry.
// The result [I | C'^-1 U | C'^-1 r'] is stored to a workspace for the back
// substitution, and each lane builds its logical row of the next block from
// the row it already holds, so no row data ever crosses lanes.  With DIAG the
// off-diagonal blocks are scalar multiples of the identity (the flamelet
// Laplacian coupling), given as one scalar per block.
//
// Workspace: for global block row g, rows k=0..NV-1 of
// [C'^-1 U | C'^-1 r'] at work[(g*NV + k)*(NV+1) + c].
#pragma once

#include "pyro_common.hpp"

namespace pyro_bt {

__device__ __forceinline__ double bcast(double v, int lane) {
  // lane must be wave-uniform
  const int lo = __builtin_amdgcn_readlane(
      static_cast<int>(__double_as_longlong(v) & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane(
      static_cast<int>(__double_as_longlong(v) >> 32), lane);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

// Pivot key of a row for column k: the top bits of |a| (exponent and 15
// mantissa bits, inf capped to the largest finite value) above the lane
// index inverted, so that an integer max picks the largest |a| and, among
// rows equal to 2^-15 relative, the lowest lane -- as good a pivot for
// partial pivoting.  Rows that are not candidates, and NaN rows, get 0: a
// maximum below 64 means no usable row is left, and the caller takes the
// lowest unused row.
__device__ __forceinline__ unsigned pivot_key(double v, bool cand, int lane) {
  const double a = fabs(v);
  if (!cand || !(a == a)) return 0u;
  unsigned hi = static_cast<unsigned>(__double_as_longlong(a) >> 32);
  hi = hi < 0x7FEFFFFFu ? hi : 0x7FEFFFFFu;
  return (((hi >> 5) + 1u) << 6) | static_cast<unsigned>(63 - lane);
}

// Max over each 16-lane row of a 32-bit key, in every lane of the row (DPP
// rotations within the row, one v_max_u32 each).
__device__ __forceinline__ unsigned row16_max_u32(unsigned v) {
#define PYRO_ROR_MAXU(N)                                                     \
  {                                                                          \
    const unsigned o = static_cast<unsigned>(__builtin_amdgcn_mov_dpp(       \
        static_cast<int>(v), 0x120 + N, 0xF, 0xF, false));                   \
    v = v > o ? v : o;                                                       \
  }
  PYRO_ROR_MAXU(8)
  PYRO_ROR_MAXU(4)
  PYRO_ROR_MAXU(2)
  PYRO_ROR_MAXU(1)
#undef PYRO_ROR_MAXU
  return v;
}

// Lane of the wave-wide maximum key (wave-uniform), or the lowest unused
// row when no usable row is left.
template <int NV>
__device__ __forceinline__ int pivot_lane(unsigned key,
                                          unsigned long long used,
                                          unsigned long long all) {
  constexpr int kRows = (NV + 15) / 16;
  const unsigned rm = row16_max_u32(key);
  unsigned best = static_cast<unsigned>(
      __builtin_amdgcn_readlane(static_cast<int>(rm), 0));
#pragma unroll
  for (int q = 1; q < kRows; ++q) {
    const unsigned b = static_cast<unsigned>(
        __builtin_amdgcn_readlane(static_cast<int>(rm), 16 * q));
    best = best > b ? best : b;
  }
  const int p = best >= 64u ? 63 - static_cast<int>(best & 63u)
                            : __builtin_ctzll(~used & all);
  return __builtin_amdgcn_readfirstlane(p);
}

// Gauss-Jordan on the augmented rows a[0..W) held one per lane (lanes < NV).
// On return the lane's row is logical row `my_k` of [I | X]; `perm[k]` is
// the lane holding logical row k (wave-uniform).  Each lane first scales
// its row so that its largest |C'| entry is 1 (the solution does not
// change; pivoting becomes scaled partial pivoting).  The pivot is the
// lane of the largest pivot_key(a[k]) (a DPP max-reduction within 16-lane
// rows and a max over the rows' results), the reciprocal of the pivot by
// v_rcp_f64 and two Newton
// steps, and every lane -- the pivot lane included -- updates its row with
// one FMA per column: the pivot lane uses the multiplier 1 - 1/piv, which
// turns a - (1 - 1/piv) a into a/piv with a relative error of |piv - 1| eps,
// small because the rows are scaled.
template <int NV, int W>
__device__ __forceinline__ void gauss_jordan_dpp(double (&a)[W], int lane,
                                                 int& my_k, int (&perm)[NV],
                                                 bool& singular) {
  static_assert(NV <= 64, "one wavefront holds the rows");
  const bool active = lane < NV;
  {
    double mx = 0.0;
#pragma unroll
    for (int c = 0; c < NV; ++c) mx = fmax(mx, fabs(a[c]));
    const double sc = mx > 0.0 ? pyro::recip(mx) : 1.0;
#pragma unroll
    for (int c = 0; c < W; ++c) a[c] *= sc;
  }
  unsigned long long used = 0ull;
  const unsigned long long all =
      NV == 64 ? ~0ull : ((1ull << NV) - 1ull);
  my_k = -1;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const int p = pivot_lane<NV>(
        pivot_key(a[k], active && !((used >> lane) & 1ull), lane), used, all);
    used |= 1ull << p;
    perm[k] = p;
    const double piv = bcast(a[k], p);
    singular = singular || !(fabs(piv) > 0.0);
    const double inv = pyro::recip(piv);
    const double g = (lane == p) ? 1.0 - inv : a[k] * inv;
#pragma unroll
    for (int j = k + 1; j < W; ++j) a[j] = fma(-g, bcast(a[j], p), a[j]);
    if (lane == p) my_k = k;
  }
}

template <int NV, bool DIAG>
__global__ void __launch_bounds__(64)
k_block_thomas(const double* __restrict__ L, const double* __restrict__ C,
               const double* __restrict__ U, const double* __restrict__ R,
               double* __restrict__ X, const int64_t* __restrict__ offsets,
               double* __restrict__ work, int64_t work_rows,
               int* __restrict__ n_singular) {
  constexpr int W = 2 * NV + 1;
  const int s = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t off = offsets[s];
  const int64_t n = offsets[s + 1] - off;
  if (n <= 0) return;
  // rows beyond the workspace (a fixed-size workspace and offsets that
  // exceed the row count the caller stated): no solve, X = NaN, counted
  // as singular
  if (off < 0 || offsets[s + 1] > work_rows) {
    if (lane < NV)
      for (int64_t i = 0; i < n; ++i)
        X[(off + i) * NV + lane] = __builtin_nan("");
    if (lane == 0 && n_singular) atomicAdd(n_singular, 1);
    return;
  }
  const bool active = lane < NV;
  bool singular = false;

  double a[W];
  int perm[NV];
  int my_k = lane;
  // block row 0: [C_0 | U_0 | r_0], logical row = lane
  {
    const int64_t g = off;
    const int r = active ? lane : 0;
#pragma unroll
    for (int c = 0; c < NV; ++c) a[c] = C[(g * NV + r) * NV + c];
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      if constexpr (DIAG) {
        a[NV + c] = (n > 1 && c == r) ? U[g] : 0.0;
      } else {
        a[NV + c] = n > 1 ? U[(g * NV + r) * NV + c] : 0.0;
      }
    }
    a[2 * NV] = R[g * NV + r];
  }
  for (int64_t i = 0; i < n; ++i) {
    const int64_t g = off + i;
    // block row i+1 of C and r, row = lane, loaded before the elimination
    // so that its latency overlaps it (the flamelet form)
    double cpre[NV], rpre = 0.0, lpre = 0.0, upre = 0.0;
    if constexpr (DIAG) {
      if (i + 1 < n) {
        const int64_t g1 = g + 1;
        const int r = active ? lane : 0;
#pragma unroll
        for (int c = 0; c < NV; ++c) cpre[c] = C[(g1 * NV + r) * NV + c];
        rpre = R[g1 * NV + r];
        lpre = L[g];
        if (i + 2 < n) upre = U[g1];
      }
    }
    gauss_jordan_dpp<NV, W>(a, lane, my_k, perm, singular);
    // store logical row my_k of [C'^-1 U | C'^-1 r']
    if (active) {
      double* wrow = work + (g * NV + my_k) * (NV + 1);
#pragma unroll
      for (int c = 0; c <= NV; ++c) wrow[c] = a[NV + c];
    }
    if (i + 1 == n) break;
    // next block row, logical row my_k:
    //   C'_{i+1}[k,:] = C_{i+1}[k,:] - (L_i X_U)[k,:]
    //   r'_{i+1}[k]   = r_{i+1}[k]   - (L_i X_r)[k]
    const int64_t g1 = g + 1;
    const int k = active ? my_k : 0;
    double nxt[W];
    if constexpr (DIAG) {
      // this lane continues with logical row k: take row k of the
      // prefetched block from the lane that loaded it
      const double l = lpre;
#pragma unroll
      for (int c = 0; c < NV; ++c)
        nxt[c] = fma(-l, a[NV + c], __shfl(cpre[c], k));
      nxt[2 * NV] = fma(-l, a[2 * NV], __shfl(rpre, k));
    } else {
#pragma unroll
      for (int c = 0; c < NV; ++c) nxt[c] = C[(g1 * NV + k) * NV + c];
      nxt[2 * NV] = R[g1 * NV + k];
#pragma unroll
      for (int m = 0; m < NV; ++m) {
        const double lkm = L[(g * NV + k) * NV + m];
        const int src = perm[m];  // lane holding logical row m
#pragma unroll
        for (int c = 0; c < NV; ++c)
          nxt[c] = fma(-lkm, bcast(a[NV + c], src), nxt[c]);
        nxt[2 * NV] = fma(-lkm, bcast(a[2 * NV], src), nxt[2 * NV]);
      }
    }
    const bool has_u = (i + 2 < n);
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      if constexpr (DIAG) {
        nxt[NV + c] = (has_u && c == k) ? upre : 0.0;
      } else {
        nxt[NV + c] = has_u ? U[(g1 * NV + k) * NV + c] : 0.0;
      }
    }
#pragma unroll
    for (int c = 0; c < W; ++c) a[c] = nxt[c];
  }
  // the workspace rows were written by other lanes of this wave
  __syncthreads();
  // back substitution, lane = logical row: x_i = X_r,i - X_U,i x_{i+1}
  const int r = active ? lane : 0;
  double x_next[NV];
  {
    const int64_t g = off + n - 1;
    const double xr = work[(g * NV + r) * (NV + 1) + NV];
    if (active) X[g * NV + r] = xr;
#pragma unroll
    for (int m = 0; m < NV; ++m) x_next[m] = bcast(xr, m);
  }
  double wcur[NV + 1];
  if (n >= 2) {
    const double* wrow = work + ((off + n - 2) * NV + r) * (NV + 1);
#pragma unroll
    for (int m = 0; m <= NV; ++m) wcur[m] = wrow[m];
  }
  for (int64_t i = n - 2; i >= 0; --i) {
    const int64_t g = off + i;
    // next row's workspace entries, loaded while this row is reduced
    double wnext[NV + 1];
    if (i > 0) {
      const double* wrow = work + ((g - 1) * NV + r) * (NV + 1);
#pragma unroll
      for (int m = 0; m <= NV; ++m) wnext[m] = wrow[m];
    }
    double xr = wcur[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) xr = fma(-wcur[m], x_next[m], xr);
    if (active) X[g * NV + r] = xr;
#pragma unroll
    for (int m = 0; m < NV; ++m) x_next[m] = bcast(xr, m);
    if (i > 0) {
#pragma unroll
      for (int m = 0; m <= NV; ++m) wcur[m] = wnext[m];
    }
  }
  if (singular && lane == 0 && n_singular) atomicAdd(n_singular, 1);
}

}  // namespace pyro_bt
