// Flamelet source kernel of a generated mechanism with NS <= 16: T(h, Y),
// S and the NV x NV source Jacobian per point in one pass, with the NS x NS
// kinetics Jacobian held in registers (mech::net_production_rates_jac).
//
// Included after pyro_flamelet_kernels.hpp.  Larger mechanisms form the
// source Jacobian with pyro_source_jacobian instead.
#pragma once

#include "pyro_flamelet_kernels.hpp"

namespace pyro_fl {

// Source term S = [vd/rho ; W_k w_k / rho] of one point
// (flamelets/equations.py:313-389) and, if jout != nullptr, its Jacobian
// dS/dphi (NV x NV, row-major) by the chain rule through T(h, Y) and
// rho(T, Y) = p W/(R T):
//   dT/dh = 1/cp, dT/dY_m = -h_m/cp, drho/dh = -rho/T dT/dh,
//   drho/dY_m = -rho W/W_m - rho/T dT/dY_m, dC_k = (Y_k/W_k) drho + rho/W_k dY_k.
// kRef: T and dT/dphi from the reference's loop (temperature_from_h_tangent)
// instead of the converged Newton and the exact derivative of T(h, Y).
template <bool kRef>
__device__ __forceinline__ bool point_source(const double h,
                                             const double (&Y)[NS],
                                             const double vd,
                                             const double tguess,
                                             const double p, const double tol,
                                             const int maxit, double& T,
                                             double (&S)[NV],
                                             double* __restrict__ jout) {
  double th_ref = 0.0, tY_ref[NS];
  bool ok;
  if constexpr (kRef)
    ok = temperature_from_h_tangent(h, Y, tguess, tol, maxit, T, th_ref,
                                    tY_ref);
  else
    ok = temperature_from_h(h, Y, tguess, tol, maxit, T);
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < NS; ++k) s = fma(mech::kInvW[k], Y[k], s);
  const double mmw = 1.0 / s;
  const double rho = p * mmw / (kGasConstant * T);
  double C[NS], w[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) C[k] = mech::kInvW[k] * rho * Y[k];
  const double irho = 1.0 / rho;
  if (jout == nullptr) {
    mech::net_production_rates(T, C, w);
  } else {
    double J[NS][NS], JT[NS];
    mech::net_production_rates_jac(T, C, w, J, JT);
    double hrt[NS], cpr[NS];
    mech::species_h_rt(T, 1.0 / T, hrt);
    mech::species_cp_r(T, cpr);
    double sc = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k) sc = fma(Y[k] * mech::kInvW[k], cpr[k], sc);
    const double cp = kGasConstant * sc;
    const double dTdh = kRef ? th_ref : 1.0 / cp;
    const double rT = rho / T;
    const double drdh = -rT * dTdh;
    double dTdY[NS], drdY[NS];
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      dTdY[m] = kRef ? tY_ref[m]
                     : -(kGasConstant * T * hrt[m] * mech::kInvW[m]) * dTdh;
      drdY[m] = fma(-rT, dTdY[m], -rho * mmw * mech::kInvW[m]);
    }
    // row 0: d(vd/rho)
    const double c0 = -vd * irho * irho;
    jout[0] = c0 * drdh;
#pragma unroll
    for (int m = 0; m < NS; ++m) jout[1 + m] = c0 * drdY[m];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      double a = 0.0;
#pragma unroll
      for (int j = 0; j < NS; ++j) a = fma(J[k][j], Y[j] * mech::kInvW[j], a);
      const double f = mech::kW[k] * irho;
      const double wr = w[k] * irho;
      const double dwh = fma(a, drdh, JT[k] * dTdh);
      double* row = jout + (1 + k) * NV;
      row[0] = f * fma(-wr, drdh, dwh);
#pragma unroll
      for (int m = 0; m < NS; ++m) {
        const double dwm = fma(J[k][m], rho * mech::kInvW[m],
                               fma(a, drdY[m], JT[k] * dTdY[m]));
        row[1 + m] = f * fma(-wr, drdY[m], dwm);
      }
    }
  }
  S[0] = vd * irho;
#pragma unroll
  for (int k = 0; k < NS; ++k) S[1 + k] = mech::kW[k] * w[k] * irho;
  return ok;
}

template <bool kRef>
__global__ void __launch_bounds__(256)
k_flamelet_source(const double* __restrict__ phi, int64_t ldp,
                  const double* __restrict__ vd,
                  const double* __restrict__ tguess, int64_t tg_stride,
                  const double* __restrict__ p, int64_t p_stride,
                  double* __restrict__ S, int64_t lds,
                  double* __restrict__ T_out, double* __restrict__ J,
                  int64_t n, double tol, int maxit, int* __restrict__ n_fail) {
  const int64_t i = cell_index();
  if (i >= n) return;
  double Y[NS], s[NV], T;
  const double h = phi[i];
#pragma unroll
  for (int k = 0; k < NS; ++k) Y[k] = phi[(1 + k) * ldp + i];
  const bool ok = point_source<kRef>(h, Y, vd ? vd[i] : 0.0,
                                     tguess[i * tg_stride],
                               p[i * p_stride], tol, maxit, T, s,
                               J ? J + i * NV * NV : nullptr);
  // a point whose temperature did not converge poisons its flamelet's
  // update (NaN S), which the caller sees in the update norm
  const double bad = __builtin_nan("");
#pragma unroll
  for (int v = 0; v < NV; ++v) S[v * lds + i] = ok ? s[v] : bad;
  if (T_out) T_out[i] = ok ? T : bad;
  if (!ok && n_fail) atomicAdd(n_fail, 1);
}

}  // namespace pyro_fl

PYRO_API int pyro_flamelet_source(const double* phi, int64_t ldp,
                                  const double* vd, const double* tguess,
                                  int64_t tg_stride, const double* p,
                                  int64_t p_stride, double* S, int64_t lds,
                                  double* T_out, double* J, int64_t n,
                                  double tol, int maxit, int* n_fail,
                                  int semantics, hipStream_t stream) {
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldp < n || lds < n || !phi || !tguess || !p || !S ||
      (semantics != 0 && semantics != 1))
    return pyro::kInvalidArgument;
  if (semantics == 1)
    pyro_fl::k_flamelet_source<true><<<pyro::grid_for(n, 256), 256, 0,
                                       stream>>>(
        phi, ldp, vd, tguess, tg_stride, p, p_stride, S, lds, T_out, J, n,
        tol, maxit, n_fail);
  else
    pyro_fl::k_flamelet_source<false><<<pyro::grid_for(n, 256), 256, 0,
                                        stream>>>(
        phi, ldp, vd, tguess, tg_stride, p, p_stride, S, lds, T_out, J, n,
        tol, maxit, n_fail);
  return pyro::launch_status();
}

