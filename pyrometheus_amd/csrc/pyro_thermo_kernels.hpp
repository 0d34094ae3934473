// Mechanism-generic pointwise kernels and their C ABI.
//
// This header is included at the end of every generated mechanism
// translation unit, after the generator has defined namespace `mech` with:
//   NS, NR, NF, kW[NS], kInvW[NS]
//   species_cp_r / species_h_rt / species_s_r / species_g_rt
//   species_dcp_r / species_dh_rt / species_ds_r
//   reaction_quantities(...)   (k_f, -ln K_c, r_f, r_r, r_net per reaction)
//   net_production_rates(...)  (fused omega-dot)
//   creation_destruction(...)
//   net_production_rates_jac(...) (omega-dot and its (rho,T,Y) Jacobian)
// Layout: every per-cell array is SoA; a per-species array is (NS, ld) with
// cell i of species k at [k*ld + i].  Per-cell scalar inputs take a stride
// (0 broadcasts one value, 1 reads one value per cell).
//
// All entry points are asynchronous on the given stream and return a status
// (0 ok, <0 error).  Pointers are device pointers owned by the caller.
#pragma once

#include "pyro_common.hpp"

namespace pyro_kern {

using namespace pyro;

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// species thermo: out[k*ld + i]
// ---------------------------------------------------------------------------
enum SpeciesThermo : int {
  kCpR = 0, kHRT = 1, kSR = 2, kGRT = 3, kDCpR = 4, kDHRT = 5, kDSR = 6,
};

template <int WHICH>
__global__ void __launch_bounds__(kBlock)
k_species_thermo(const double* __restrict__ T, int64_t t_stride,
                 double* __restrict__ out, int64_t ld, int64_t n) {
  const int64_t i = cell_index();
  if (i >= n) return;
  const double t = T[i * t_stride];
  double o[mech::NS];
  if constexpr (WHICH == kCpR) mech::species_cp_r(t, o);
  if constexpr (WHICH == kHRT) mech::species_h_rt(t, 1.0 / t, o);
  if constexpr (WHICH == kSR) mech::species_s_r(t, log(t), o);
  if constexpr (WHICH == kGRT) mech::species_g_rt(t, log(t), 1.0 / t, o);
  if constexpr (WHICH == kDCpR) mech::species_dcp_r(t, o);
  if constexpr (WHICH == kDHRT) mech::species_dh_rt(t, 1.0 / t, o);
  if constexpr (WHICH == kDSR) mech::species_ds_r(t, 1.0 / t, o);
  // streaming stores: the outputs are not read again by this kernel
#pragma unroll
  for (int k = 0; k < mech::NS; ++k)
    __builtin_nontemporal_store(o[k], &out[k * ld + i]);
}

// ---------------------------------------------------------------------------
// mixture thermo (codegen/python.py:287-311)
// ---------------------------------------------------------------------------
enum MixtureThermo : int { kCpMass = 0, kCvMass = 1, kHMass = 2, kEMass = 3 };

// sum_k Y_k * prop_k / W_k
__device__ __forceinline__ double mass_average(const double* __restrict__ Y,
                                               int64_t ldY, int64_t i,
                                               const double (&prop)[mech::NS],
                                               double shift) {
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < mech::NS; ++k)
    acc += Y[k * ldY + i] * (prop[k] - shift) * mech::kInvW[k];
  return acc;
}

template <int WHICH>
__global__ void __launch_bounds__(kBlock)
k_mixture_thermo(const double* __restrict__ T, int64_t t_stride,
                 const double* __restrict__ Y, int64_t ldY,
                 double* __restrict__ out, int64_t n) {
  const int64_t i = cell_index();
  if (i >= n) return;
  const double t = T[i * t_stride];
  double p[mech::NS];
  if constexpr (WHICH == kCpMass || WHICH == kCvMass) {
    mech::species_cp_r(t, p);
    out[i] = kGasConstant *
             mass_average(Y, ldY, i, p, WHICH == kCvMass ? 1.0 : 0.0);
  } else {
    mech::species_h_rt(t, 1.0 / t, p);
    out[i] = kGasConstant * t *
             mass_average(Y, ldY, i, p, WHICH == kEMass ? 1.0 : 0.0);
  }
}

// ---------------------------------------------------------------------------
// mixture state (codegen/python.py:248-270)
// ---------------------------------------------------------------------------
enum MixtureState : int {
  kMixMolWeight = 0, kSpecificGasConstant = 1, kDensity = 2, kPressure = 3,
};

__device__ __forceinline__ double inv_w_sum(const double* __restrict__ Y,
                                            int64_t ldY, int64_t i) {
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < mech::NS; ++k) acc += mech::kInvW[k] * Y[k * ldY + i];
  return acc;
}

template <int WHICH>
__global__ void __launch_bounds__(kBlock)
k_mixture_state(const double* __restrict__ a, int64_t a_stride,
                const double* __restrict__ T, int64_t t_stride,
                const double* __restrict__ Y, int64_t ldY,
                double* __restrict__ out, int64_t n) {
  const int64_t i = cell_index();
  if (i >= n) return;
  const double s = inv_w_sum(Y, ldY, i);
  if constexpr (WHICH == kMixMolWeight) out[i] = 1.0 / s;
  if constexpr (WHICH == kSpecificGasConstant) out[i] = kGasConstant * s;
  if constexpr (WHICH == kDensity) {
    const double mmw = 1.0 / s;
    out[i] = a[i * a_stride] * mmw / (kGasConstant * T[i * t_stride]);
  }
  if constexpr (WHICH == kPressure) {
    const double mmw = 1.0 / s;
    out[i] = a[i * a_stride] * (kGasConstant * T[i * t_stride]) / mmw;
  }
}

// species state: concentrations C_k = invW_k*rho*Y_k, mole fractions
// X_k = invW_k*Y_k*mmw (codegen/python.py:272-285)
enum SpeciesState : int { kConcentrations = 0, kMoleFractions = 1 };

template <int WHICH>
__global__ void __launch_bounds__(kBlock)
k_species_state(const double* __restrict__ a, int64_t a_stride,
                const double* __restrict__ Y, int64_t ldY,
                double* __restrict__ out, int64_t ld, int64_t n) {
  const int64_t i = cell_index();
  if (i >= n) return;
  const double s = a[i * a_stride];
#pragma unroll
  for (int k = 0; k < mech::NS; ++k) {
    if constexpr (WHICH == kConcentrations)
      out[k * ld + i] = mech::kInvW[k] * s * Y[k * ldY + i];
    else
      out[k * ld + i] = mech::kInvW[k] * Y[k * ldY + i] * s;
  }
}

// ---------------------------------------------------------------------------
// kinetics
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_conc(const double* __restrict__ rho,
                                          const double* __restrict__ Y,
                                          int64_t ldY, int64_t i,
                                          double (&C)[mech::NS]) {
  const double r = rho[i];
#pragma unroll
  for (int k = 0; k < mech::NS; ++k) C[k] = mech::kInvW[k] * r * Y[k * ldY + i];
}

// MASK: the outputs computed (1 kf, 2 nlk, 4 rf, 8 rr, 16 rnet), each
// also skipped when its pointer is null.
template <int MASK>
__global__ void __launch_bounds__(kBlock)
k_reaction_quantities(const double* __restrict__ T,
                      const double* __restrict__ C, int64_t ldC,
                      double* __restrict__ kf, double* __restrict__ nlk,
                      double* __restrict__ rf, double* __restrict__ rr,
                      double* __restrict__ rnet, int64_t ldR, int64_t n) {
  const int64_t i = cell_index();
  if (i >= n) return;
  double c[mech::NS];
#pragma unroll
  for (int k = 0; k < mech::NS; ++k) c[k] = C[k * ldC + i];
  mech::reaction_quantities<MASK>(T[i], c, i, ldR, kf, nlk, rf, rr, rnet);
}

// kWdotWaves > 0 asks the compiler for that many waves per SIMD (it then
// budgets registers for that occupancy).
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK, mech::kWdotWaves > 0 ? mech::kWdotWaves : 1)
k_wdot(const double* __restrict__ rho, const double* __restrict__ T,
       const double* __restrict__ Y, int64_t ldY, double* __restrict__ W,
       int64_t ldW, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
  if (i >= n) return;
  mech::net_production_rates_lazy(rho[i], T[i], Y, ldY, i, W, ldW);
}

__global__ void __launch_bounds__(kBlock)
k_creation_destruction(const double* __restrict__ rho,
                       const double* __restrict__ T,
                       const double* __restrict__ Y, int64_t ldY,
                       double* __restrict__ cdot, double* __restrict__ ddot,
                       int64_t ldW, int64_t n) {
  const int64_t i = cell_index();
  if (i >= n) return;
  mech::creation_destruction_lazy(rho[i], T[i], Y, ldY, i, cdot, ddot, ldW);
}

// ---------------------------------------------------------------------------
// derivatives of w(rho, T, Y) with C_k = rho Y_k / W_k
// ---------------------------------------------------------------------------
// JVP: dC_k = (Y_k v_rho + rho v_Yk) / W_k, one forward-mode pass.
__global__ void __launch_bounds__(kBlock, mech::kDerivWaves > 0 ? mech::kDerivWaves : 1)
k_wdot_jvp(const double* __restrict__ rho, const double* __restrict__ T,
           const double* __restrict__ Y, int64_t ldY,
           const double* __restrict__ vr, const double* __restrict__ vT,
           const double* __restrict__ vY, int64_t ldV,
           double* __restrict__ wout, double* __restrict__ dwout, int64_t ldO,
           int64_t n) {
  const int64_t i = cell_index();
  if (i >= n) return;
  mech::net_production_rates_jvp_lazy(rho[i], T[i], Y, ldY, vr ? vr[i] : 0.0,
                                      vT ? vT[i] : 0.0, vY, ldV, i, wout,
                                      dwout, ldO);
}

// VJP: Cb = wb^T dw/dC, Tb = wb^T dw/dT in one reverse pass, then
// g_rho = sum_k Cb_k Y_k / W_k, g_Yk = Cb_k rho / W_k, g_T = Tb.
__global__ void __launch_bounds__(kBlock)
k_wdot_vjp(const double* __restrict__ rho, const double* __restrict__ T,
           const double* __restrict__ Y, int64_t ldY,
           const double* __restrict__ wb, int64_t ldB,
           double* __restrict__ g_rho, double* __restrict__ g_T,
           double* __restrict__ g_Y, int64_t ldG, int64_t n) {
  const int64_t i = cell_index();
  if (i >= n) return;
  mech::net_production_rates_vjp_lazy(rho[i], T[i], Y, ldY, wb, ldB, i, g_rho,
                                      g_T, g_Y, ldG);
}

// Jacobian of the flamelet source S = [vd/rho; W_k w_k/rho] with respect to
// phi = [h; Y] at a state whose temperature T is consistent with h, and
// rho = p W/(R T) (flamelets/equations.py:361-389 differentiated as in
// :124-129).  By the chain rule through T(h, Y) and rho(T, Y), column c
// moves T and rho by
//   c = 0:      dT = 1/cp,              drho = -rho/T dT
//   c = 1 + m:  dT = -h_m/cp,           drho = -rho W/W_m - rho/T dT
// and C_j by (Y_j/W_j) drho + delta_jm rho/W_m.  Two passes per chunk of
// mech::kCacheLd cells: k_reaction_cache stores per reaction k_f, k_r, q,
// dq/dT, dq/d[M] to a workspace (slot s of chunk cell i at
// work[s*kCacheLd + i]), and k_source_jacobian_rows assembles each row
// from the reactions of its species.  Entry (r, c) of cell i goes to
// out[(r*NV + c)*s_rc + i*s_i].
__device__ __forceinline__ void source_state(
    const double* __restrict__ T, const double* __restrict__ p,
    int64_t p_stride, const double* __restrict__ Y, int64_t ldY, int64_t i,
    double& t, double& rho, double& mmw, double (&C)[mech::NS]) {
  t = T[i];
  double sy = 0.0;
#pragma unroll
  for (int k = 0; k < mech::NS; ++k) sy = fma(Y[k * ldY + i], mech::kInvW[k], sy);
  mmw = 1.0 / sy;
  rho = p[i * p_stride] * mmw / (kGasConstant * t);
#pragma unroll
  for (int k = 0; k < mech::NS; ++k) C[k] = mech::kInvW[k] * rho * Y[k * ldY + i];
}

__global__ void __launch_bounds__(kBlock)
k_reaction_cache(const double* __restrict__ T, const double* __restrict__ p,
                 int64_t p_stride, const double* __restrict__ Y, int64_t ldY,
                 double* __restrict__ work, int64_t n) {
  const int64_t i = cell_index();
  if (i >= n) return;
  double t, rho, mmw, C[mech::NS];
  source_state(T, p, p_stride, Y, ldY, i, t, rho, mmw, C);
  mech::reaction_cache(t, C, work, static_cast<unsigned>(i));
}

// Row pass of the source Jacobian.  G = 1: one thread per cell builds
// every row (kBlock cells per block); G = mech::kJacGroups: the G waves of
// a block build one group of rows each (balanced by work) for the same
// kBlock / G cells, sharing the dT table in LDS -- G times the waves for a
// chunk of few cells (a flamelet batch).  A row's arithmetic is the same
// in both forms.
template <int G>
__global__ void __launch_bounds__(kBlock, mech::kJacWaves > 0 ? mech::kJacWaves : 1)
k_source_jacobian_rows(const double* __restrict__ T,
                       const double* __restrict__ p, int64_t p_stride,
                       const double* __restrict__ Y, int64_t ldY,
                       const double* __restrict__ vd,
                       const double* __restrict__ work,
                       double* __restrict__ out, int64_t s_rc, int64_t s_i,
                       double* __restrict__ S, int64_t ldS, int64_t n) {
  constexpr int kCells = kBlock / G;
  static_assert(kCells % 64 == 0, "a row group is whole waves");
  // dT of column 1+m for every cell of the block, read by every row
  __shared__ double dTl[mech::NS * kCells];
  const int cl = threadIdx.x % kCells;
  const int rg = __builtin_amdgcn_readfirstlane(threadIdx.x / kCells);
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kCells + cl;
  // lanes past the end work on the last cell (no early return before the
  // barrier) and store nothing
  const int64_t i = i0 < n ? i0 : n - 1;
  double t, rho, mmw, C[mech::NS];
  source_state(T, p, p_stride, Y, ldY, i, t, rho, mmw, C);
  // one species array live at a time: cp first, then h/RT straight to LDS
  double cp;
  {
    double cpr[mech::NS];
    mech::species_cp_r(t, cpr);
    double sc = 0.0;
#pragma unroll
    for (int k = 0; k < mech::NS; ++k)
      sc = fma(Y[k * ldY + i] * mech::kInvW[k], cpr[k], sc);
    cp = kGasConstant * sc;
  }
  {
    double hrt[mech::NS];
    mech::species_h_rt(t, 1.0 / t, hrt);
    const double kT = -kGasConstant * t / cp;
    // the waves write the species k = rg (mod G)
#pragma unroll
    for (int k = 0; k < mech::NS; ++k)
      if (k % G == rg)
        dTl[k * kCells + cl] = kT * hrt[k] * mech::kInvW[k];
  }
  __syncthreads();
  if (i0 >= n) return;
  if constexpr (G == 1)
    mech::source_jacobian_rows<0, mech::kJacGroups>(
        C, rho, t, cp, mmw, vd ? vd[i] : 0.0, work, static_cast<unsigned>(i),
        (pyro::lptr)(dTl + cl), kCells, out + i * s_i, s_rc,
        S ? S + i : nullptr, ldS);
  else
    mech::source_jacobian_rows_group(
        rg, C, rho, t, cp, mmw, vd ? vd[i] : 0.0, work,
        static_cast<unsigned>(i), (pyro::lptr)(dTl + cl), kCells,
        out + i * s_i, s_rc, S ? S + i : nullptr, ldS);
}

// One-pass row kernel, no workspace: a block of 64 G threads (G waves,
// G dividing mech::kJacFusedGroups) takes 64 cells; its waves share the
// cells' tables in LDS (species-major, 64 cells per row: concentrations C,
// dT of each column, and with mech::kJacFusedGibbs the Gibbs factors
// E = exp(g'), F = exp(-g')), each wave filling the species k = rg
// (mod G), and wave rg then builds the row groups [rg P, (rg + 1) P),
// P = kJacFusedGroups / G, evaluating each row's reactions itself.  The
// mixture sums sum_k Y_k/W_k and sum_k Y_k c_p,k/W_k are taken over eight
// fixed blocks of species (k mod 8) and the blocks added in order, so no
// value depends on G.  Stores by layout: with s_i == 1 (entry-major) a
// wave's store covers 512 contiguous bytes and the entries are written
// with streaming stores (mode 1); with s_rc == 1 (each cell's row
// contiguous, as in per-cell blocks) each wave passes a row through its
// LDS buffer of 8 entries x 64 cells, 8 entries at a time, and stores 8
// cells x 64 contiguous bytes per instruction (mode 2, blocks of 64 cells
// and rows of at least 8 entries; the last, partial block, shorter rows
// and other strides take ordinary stores, mode 0).
constexpr int kJacFusedCells = mech::kJacFusedCells;
constexpr int kJacFusedTables = mech::kJacFusedGibbs ? 4 : 2;

__global__ void __launch_bounds__(64 * 8)
k_source_jacobian_fused(const double* __restrict__ T,
                        const double* __restrict__ p, int64_t p_stride,
                        const double* __restrict__ Y, int64_t ldY,
                        const double* __restrict__ vd,
                        double* __restrict__ out, int64_t s_rc, int64_t s_i,
                        double* __restrict__ S, int64_t ldS, int64_t n) {
  constexpr int kCells = kJacFusedCells;
  constexpr int kTab = mech::NS * kCells;
  __shared__ double tab[kJacFusedTables * kTab];
  __shared__ double part[2 * 8 * kCells];
  __shared__ double stage[8 * 8 * mech::kJacStageLd];
  double* Cl = tab;
  double* dTl = tab + kTab;
  double* El = tab + (kJacFusedTables > 2 ? 2 * kTab : 0);
  double* Fl = tab + (kJacFusedTables > 2 ? 3 * kTab : 0);
  const int cl = threadIdx.x % kCells;
  const int G = __builtin_amdgcn_readfirstlane(blockDim.x / kCells);
  const int rg = __builtin_amdgcn_readfirstlane(threadIdx.x / kCells);
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kCells + cl;
  // lanes past the end work on the last cell (no early return before the
  // barriers) and store nothing
  const int64_t i = i0 < n ? i0 : n - 1;
  const double t = T[i];
  {
    // this wave's blocks b = rg (mod G) of sum Y/W and sum Y c_p/R / W
    double cpr[mech::NS];
    mech::species_cp_r(t, cpr);
    double sy[8] = {}, sc[8] = {};
#pragma unroll
    for (int k = 0; k < mech::NS; ++k)
      if ((k & 7) % G == rg) {
        const double yw = Y[k * ldY + i] * mech::kInvW[k];
        sy[k & 7] += yw;
        sc[k & 7] = fma(yw, cpr[k], sc[k & 7]);
      }
#pragma unroll
    for (int b = 0; b < 8; ++b)
      if (b % G == rg) {
        part[b * kCells + cl] = sy[b];
        part[(8 + b) * kCells + cl] = sc[b];
      }
  }
  __syncthreads();
  double sy = 0.0, sc = 0.0;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    sy += part[b * kCells + cl];
    sc += part[(8 + b) * kCells + cl];
  }
  const double mmw = 1.0 / sy;
  const double rho = p[i * p_stride] * mmw / (kGasConstant * t);
  const double cp = kGasConstant * sc;
  {
    double hrt[mech::NS];
    mech::species_h_rt(t, 1.0 / t, hrt);
    const double kT = -kGasConstant * t / cp;
#pragma unroll
    for (int k = 0; k < mech::NS; ++k)
      if (k % G == rg) {
        Cl[k * kCells + cl] = mech::kInvW[k] * rho * Y[k * ldY + i];
        dTl[k * kCells + cl] = kT * hrt[k] * mech::kInvW[k];
      }
    mech::jac_fused_gibbs(t, rg, G, El + cl, Fl + cl);
  }
  __syncthreads();
  if (i0 >= n) return;
  const int64_t c0 = static_cast<int64_t>(blockIdx.x) * kCells;
  const int mode =
      s_i == 1 ? 1 : ((s_rc == 1 && n - c0 >= kCells && 7 * s_i < (1 << 26)) ? 2 : 0);
  double* stw = stage + rg * (8 * mech::kJacStageLd);
  double ctot = 0.0;
#pragma unroll
  for (int k = 0; k < mech::NS; ++k) ctot += Cl[k * kCells + cl];
  const int per = mech::kJacFusedGroups / G;
  mech::source_jacobian_fused(
      rg * per, (rg + 1) * per, (pyro::lptr)(Cl + cl), ctot, rho, t, cp, mmw,
      vd ? vd[i] : 0.0, (pyro::lptr)(dTl + cl), (pyro::lptr)(El + cl),
      (pyro::lptr)(Fl + cl), (pyro::gmptr)(out + i * s_i), s_rc,
      S ? (pyro::gmptr)(S + i) : nullptr, ldS, mode,
      (pyro::lmptr)(stw + cl),
      (pyro::lptr)(stw + (cl & 7) * mech::kJacStageLd + (cl >> 3)),
      (pyro::gmptr)(out + c0 * s_i),
      static_cast<unsigned>(((cl >> 3) * s_i + (cl & 7)) * 8), 8 * s_i);
}

// ---------------------------------------------------------------------------
// temperature: per-lane Newton on e(T,Y)=e* (do_energy) or h(T,Y)=h*
// (codegen/python.py:375-396; flamelets/make_pyro.py:255-299).
// mode 0: step, then stop when |step| < tol (generated get_temperature)
// mode 1: stop when |step| <= tol without taking it (JAX while_loop form)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock, mech::kTempWaves > 0 ? mech::kTempWaves : 1)
k_temperature(const double* __restrict__ target, int64_t target_stride,
              const double* __restrict__ t_guess, int64_t guess_stride,
              const double* __restrict__ Y, int64_t ldY,
              double* __restrict__ T_out, int64_t n, int do_energy, int mode,
              double tol, int maxit, int* __restrict__ n_fail,
              int* __restrict__ iters) {
  const int64_t i = cell_index();
  if (i >= n) return;
  const double e = target[i * target_stride];
  double t = t_guess[i * guess_stride];
  const double shift = do_energy != 0 ? 1.0 : 0.0;
  // the mixture's NASA-7 coefficients, once per cell
  double M[mech::kMixGroups][2][6], syw;
  {
    double yw[mech::NS];
#pragma unroll
    for (int k = 0; k < mech::NS; ++k) yw[k] = Y[k * ldY + i] * mech::kInvW[k];
    mech::mixture_nasa_coeffs(yw, M, syw);
  }
  int it = 0;
  bool ok = false;
  for (; it < maxit; ++it) {
    // he = e or h (mass), pv = cv or cp (mass), one polynomial per group
    double sh, sc;
    mech::mixture_h_cp_eval(t, 1.0 / t, M, syw, shift, sh, sc);
    const double he = kGasConstant * t * sh;
    const double pv = kGasConstant * sc;
    const double dt = (e - he) / pv;  // = -rhs/deriv with deriv = -pv
    if (mode == 1) {
      if (fabs(dt) <= tol) { ok = true; break; }
      t += dt;
    } else {
      t += dt;
      if (fabs(dt) < tol) { ok = true; ++it; break; }
    }
  }
  T_out[i] = t;
  if (iters) iters[i] = ok ? it : -1;
  if (!ok && n_fail) atomicAdd(n_fail, 1);
}

}  // namespace pyro_kern

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
#define PYRO_API extern "C" __attribute__((visibility("default")))

PYRO_API int pyro_mech_info(int64_t* out, int n) {
  if (!out || n < 4) return pyro::kInvalidArgument;
  out[0] = mech::NS;
  out[1] = mech::NR;
  out[2] = mech::NF;
  out[3] = mech::kAbiVersion;
  return pyro::kOk;
}

PYRO_API int pyro_species_thermo(int which, const double* T, int64_t t_stride,
                                 double* out, int64_t ld, int64_t n,
                                 hipStream_t stream) {
  using namespace pyro_kern;
  if (n == 0) return pyro::kOk;
  if (n < 0 || ld < n || !T || !out) return pyro::kInvalidArgument;
  const dim3 g(pyro::grid_for(n, kBlock)), b(kBlock);
  switch (which) {
    case kCpR: k_species_thermo<kCpR><<<g, b, 0, stream>>>(T, t_stride, out, ld, n); break;
    case kHRT: k_species_thermo<kHRT><<<g, b, 0, stream>>>(T, t_stride, out, ld, n); break;
    case kSR: k_species_thermo<kSR><<<g, b, 0, stream>>>(T, t_stride, out, ld, n); break;
    case kGRT: k_species_thermo<kGRT><<<g, b, 0, stream>>>(T, t_stride, out, ld, n); break;
    case kDCpR: k_species_thermo<kDCpR><<<g, b, 0, stream>>>(T, t_stride, out, ld, n); break;
    case kDHRT: k_species_thermo<kDHRT><<<g, b, 0, stream>>>(T, t_stride, out, ld, n); break;
    case kDSR: k_species_thermo<kDSR><<<g, b, 0, stream>>>(T, t_stride, out, ld, n); break;
    default: return pyro::kInvalidArgument;
  }
  return pyro::launch_status();
}

PYRO_API int pyro_mixture_thermo(int which, const double* T, int64_t t_stride,
                                 const double* Y, int64_t ldY, double* out,
                                 int64_t n, hipStream_t stream) {
  using namespace pyro_kern;
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldY < n || !T || !Y || !out) return pyro::kInvalidArgument;
  const dim3 g(pyro::grid_for(n, kBlock)), b(kBlock);
  switch (which) {
    case kCpMass: k_mixture_thermo<kCpMass><<<g, b, 0, stream>>>(T, t_stride, Y, ldY, out, n); break;
    case kCvMass: k_mixture_thermo<kCvMass><<<g, b, 0, stream>>>(T, t_stride, Y, ldY, out, n); break;
    case kHMass: k_mixture_thermo<kHMass><<<g, b, 0, stream>>>(T, t_stride, Y, ldY, out, n); break;
    case kEMass: k_mixture_thermo<kEMass><<<g, b, 0, stream>>>(T, t_stride, Y, ldY, out, n); break;
    default: return pyro::kInvalidArgument;
  }
  return pyro::launch_status();
}

PYRO_API int pyro_mixture_state(int which, const double* a, int64_t a_stride,
                                const double* T, int64_t t_stride,
                                const double* Y, int64_t ldY, double* out,
                                int64_t n, hipStream_t stream) {
  using namespace pyro_kern;
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldY < n || !Y || !out) return pyro::kInvalidArgument;
  if ((which == kDensity || which == kPressure) && (!a || !T))
    return pyro::kInvalidArgument;
  const dim3 g(pyro::grid_for(n, kBlock)), b(kBlock);
  switch (which) {
    case kMixMolWeight: k_mixture_state<kMixMolWeight><<<g, b, 0, stream>>>(a, a_stride, T, t_stride, Y, ldY, out, n); break;
    case kSpecificGasConstant: k_mixture_state<kSpecificGasConstant><<<g, b, 0, stream>>>(a, a_stride, T, t_stride, Y, ldY, out, n); break;
    case kDensity: k_mixture_state<kDensity><<<g, b, 0, stream>>>(a, a_stride, T, t_stride, Y, ldY, out, n); break;
    case kPressure: k_mixture_state<kPressure><<<g, b, 0, stream>>>(a, a_stride, T, t_stride, Y, ldY, out, n); break;
    default: return pyro::kInvalidArgument;
  }
  return pyro::launch_status();
}

PYRO_API int pyro_species_state(int which, const double* a, int64_t a_stride,
                                const double* Y, int64_t ldY, double* out,
                                int64_t ld, int64_t n, hipStream_t stream) {
  using namespace pyro_kern;
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldY < n || ld < n || !a || !Y || !out)
    return pyro::kInvalidArgument;
  const dim3 g(pyro::grid_for(n, kBlock)), b(kBlock);
  switch (which) {
    case kConcentrations: k_species_state<kConcentrations><<<g, b, 0, stream>>>(a, a_stride, Y, ldY, out, ld, n); break;
    case kMoleFractions: k_species_state<kMoleFractions><<<g, b, 0, stream>>>(a, a_stride, Y, ldY, out, ld, n); break;
    default: return pyro::kInvalidArgument;
  }
  return pyro::launch_status();
}

PYRO_API int pyro_reaction_quantities(const double* T, const double* C,
                                      int64_t ldC, double* kf, double* nlk,
                                      double* rf, double* rr, double* rnet,
                                      int64_t ldR, int64_t n,
                                      hipStream_t stream) {
  using namespace pyro_kern;
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldC < n || ldR < n || !T || !C) return pyro::kInvalidArgument;
  const int mask = (kf ? 1 : 0) | (nlk ? 2 : 0) | (rf ? 4 : 0) |
                   (rr ? 8 : 0) | (rnet ? 16 : 0);
  const unsigned grid = pyro::grid_for(n, kBlock);
  if (mask == 1)
    k_reaction_quantities<1><<<grid, kBlock, 0, stream>>>(
        T, C, ldC, kf, nlk, rf, rr, rnet, ldR, n);
  else if (mask == 2)
    k_reaction_quantities<2><<<grid, kBlock, 0, stream>>>(
        T, C, ldC, kf, nlk, rf, rr, rnet, ldR, n);
  else if (mask == 4)
    k_reaction_quantities<4><<<grid, kBlock, 0, stream>>>(
        T, C, ldC, kf, nlk, rf, rr, rnet, ldR, n);
  else
    k_reaction_quantities<31><<<grid, kBlock, 0, stream>>>(
        T, C, ldC, kf, nlk, rf, rr, rnet, ldR, n);
  return pyro::launch_status();
}

PYRO_API int pyro_wdot(const double* rho, const double* T, const double* Y,
                       int64_t ldY, double* wdot, int64_t ldW, int64_t n,
                       hipStream_t stream) {
  using namespace pyro_kern;
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldY < n || ldW < n || !rho || !T || !Y || !wdot)
    return pyro::kInvalidArgument;
  constexpr int B = mech::kWdotBlock;
  k_wdot<B><<<pyro::grid_for(n, B), B, 0, stream>>>(rho, T, Y, ldY, wdot, ldW,
                                                    n);
  return pyro::launch_status();
}

PYRO_API int pyro_creation_destruction(const double* rho, const double* T,
                                       const double* Y, int64_t ldY,
                                       double* cdot, double* ddot,
                                       int64_t ldW, int64_t n,
                                       hipStream_t stream) {
  using namespace pyro_kern;
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldY < n || ldW < n || !rho || !T || !Y || !cdot || !ddot)
    return pyro::kInvalidArgument;
  k_creation_destruction<<<pyro::grid_for(n, kBlock), kBlock, 0, stream>>>(
      rho, T, Y, ldY, cdot, ddot, ldW, n);
  return pyro::launch_status();
}

PYRO_API int pyro_temperature(const double* target, int64_t target_stride,
                              const double* t_guess, int64_t guess_stride,
                              const double* Y, int64_t ldY, double* T_out,
                              int64_t n, int do_energy, int mode, double tol,
                              int maxit, int* n_fail, int* iters,
                              hipStream_t stream) {
  using namespace pyro_kern;
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldY < n || !target || !t_guess || !Y || !T_out || maxit < 0 ||
      !(tol >= 0.0))
    return pyro::kInvalidArgument;
  k_temperature<<<pyro::grid_for(n, kBlock), kBlock, 0, stream>>>(
      target, target_stride, t_guess, guess_stride, Y, ldY, T_out, n,
      do_energy, mode, tol, maxit, n_fail, iters);
  return pyro::launch_status();
}

PYRO_API int pyro_wdot_jvp(const double* rho, const double* T, const double* Y,
                           int64_t ldY, const double* v_rho, const double* v_T,
                           const double* v_Y, int64_t ldV, double* wdot,
                           double* jvp, int64_t ldO, int64_t n,
                           hipStream_t stream) {
  using namespace pyro_kern;
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldY < n || ldO < n || (v_Y && ldV < n) || !rho || !T || !Y ||
      !jvp)
    return pyro::kInvalidArgument;
  k_wdot_jvp<<<pyro::grid_for(n, kBlock), kBlock, 0, stream>>>(
      rho, T, Y, ldY, v_rho, v_T, v_Y, ldV, wdot, jvp, ldO, n);
  return pyro::launch_status();
}

PYRO_API int pyro_wdot_vjp(const double* rho, const double* T, const double* Y,
                           int64_t ldY, const double* wbar, int64_t ldB,
                           double* g_rho, double* g_T, double* g_Y,
                           int64_t ldG, int64_t n, hipStream_t stream) {
  using namespace pyro_kern;
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldY < n || ldB < n || (g_Y && ldG < n) || !rho || !T || !Y ||
      !wbar)
    return pyro::kInvalidArgument;
  k_wdot_vjp<<<pyro::grid_for(n, kBlock), kBlock, 0, stream>>>(
      rho, T, Y, ldY, wbar, ldB, g_rho, g_T, g_Y, ldG, n);
  return pyro::launch_status();
}

// Chunks of at most this many cells use the split row pass: in the
// one-thread-per-cell form they would leave at least half of MI355X's
// 1,024 SIMDs without a wave (a full chunk gives each SIMD one, and there
// the split form measured slower).
constexpr int64_t kJacSplitCells = 512 * 64;

// Form of pyro_source_jacobian: 1 one pass (k_source_jacobian_fused, no
// workspace; default), 0 the per-reaction cache and its row pass; and the
// waves per 64 cells of the one-pass form (0: kJacFusedGroups, one row
// group per wave; fewer waves share the four LDS tables among fewer
// waves per compute unit).
struct JacForm {
  int form = 1;
  int waves = 0;
};
inline JacForm& jac_form() {
  static JacForm f;
  return f;
}

PYRO_API int64_t pyro_source_jacobian_work_doubles(void) {
  if (jac_form().form == 1) return 0;
  return static_cast<int64_t>(mech::kCacheSlots) * mech::kCacheLd;
}

PYRO_API int pyro_source_jacobian_form(int form, int waves) {
  if (form < 0) return jac_form().form;
  if (form > 1 || waves < 0 || waves > 8 ||
      (waves && mech::kJacFusedGroups % waves != 0))
    return pyro::kInvalidArgument;
  const int prev = jac_form().form;
  jac_form().form = form;
  jac_form().waves = waves;
  return prev;
}

PYRO_API int pyro_source_jacobian(const double* T, const double* p,
                                  int64_t p_stride, const double* Y,
                                  int64_t ldY, const double* vd, double* out,
                                  int64_t s_rc, int64_t s_i, double* S,
                                  int64_t ldS, double* work, int64_t n,
                                  hipStream_t stream) {
  using namespace pyro_kern;
  if (n == 0) return pyro::kOk;
  const bool fused = jac_form().form == 1;
  if (n < 0 || ldY < n || !T || !p || !Y || !out || s_rc <= 0 || s_i <= 0 ||
      (S && ldS < n) || (!work && !fused))
    return pyro::kInvalidArgument;
  if (fused) {
    int g = jac_form().waves;
    if (g == 0) g = mech::kJacFusedGroups;
    const int64_t blocks = (n + kJacFusedCells - 1) / kJacFusedCells;
    if (blocks > 0x7fffffff) return pyro::kInvalidArgument;
    k_source_jacobian_fused<<<static_cast<unsigned>(blocks),
                              kJacFusedCells * g, 0, stream>>>(
        T, p, p_stride, Y, ldY, vd, out, s_rc, s_i, S, ldS, n);
    return pyro::launch_status();
  }
  constexpr int64_t chunk = mech::kCacheLd;
  for (int64_t c0 = 0; c0 < n; c0 += chunk) {
    const int64_t nc = n - c0 < chunk ? n - c0 : chunk;
    k_reaction_cache<<<pyro::grid_for(nc, kBlock), kBlock, 0, stream>>>(
        T + c0, p + c0 * p_stride, p_stride, Y + c0, ldY, work, nc);
    // a small chunk splits every cell's rows over the waves of a block
    constexpr int G = mech::kJacGroups;
    if (G > 1 && nc <= kJacSplitCells)
      k_source_jacobian_rows<G><<<pyro::grid_for(nc, kBlock / G), kBlock, 0,
                                  stream>>>(
          T + c0, p + c0 * p_stride, p_stride, Y + c0, ldY,
          vd ? vd + c0 : nullptr, work, out + c0 * s_i, s_rc, s_i,
          S ? S + c0 : nullptr, ldS, nc);
    else
      k_source_jacobian_rows<1><<<pyro::grid_for(nc, kBlock), kBlock, 0,
                                  stream>>>(
          T + c0, p + c0 * p_stride, p_stride, Y + c0, ldY,
          vd ? vd + c0 : nullptr, work, out + c0 * s_i, s_rc, s_i,
          S ? S + c0 : nullptr, ldS, nc);
    const int st = pyro::launch_status();
    if (st != pyro::kOk) return st;
  }
  return pyro::kOk;
}
