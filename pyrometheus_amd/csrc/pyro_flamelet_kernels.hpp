// Flamelet and derivative kernels of a generated mechanism (NS <= 16: the
// NS x NS kinetics Jacobian is held in registers).
//
// Included at the end of a generated translation unit after
// pyro_thermo_kernels.hpp; uses mech::net_production_rates_jac.
//
// State layout: phi is (NV, ld) with row 0 the mass enthalpy h and rows
// 1..NS the mass fractions (pyrometheus/flamelets/state.py:297-316).
// Jacobian blocks are (P, NV, NV) row-major per point, the layout the block
// Thomas kernel reads.
#pragma once

#include "pyro_block_thomas.hpp"

namespace pyro_fl {

using namespace pyro;
constexpr int NS = mech::NS;
constexpr int NV = NS + 1;

// Newton on h(T, Y) = h: take the step, stop once |step| < tol.
__device__ __forceinline__ bool temperature_from_h(double h,
                                                   const double (&Y)[NS],
                                                   double t, double tol,
                                                   int maxit, double& T) {
  for (int it = 0; it < maxit; ++it) {
    double hrt[NS], cpr[NS];
    mech::species_h_rt(t, 1.0 / t, hrt);
    mech::species_cp_r(t, cpr);
    double sh = 0.0, sc = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const double yw = Y[k] * mech::kInvW[k];
      sh = fma(yw, hrt[k], sh);
      sc = fma(yw, cpr[k], sc);
    }
    const double step = (h - kGasConstant * t * sh) / (kGasConstant * sc);
    t += step;
    if (fabs(step) < tol) {
      T = t;
      return true;
    }
  }
  T = t;
  return false;
}

// Source term S = [vd/rho ; W_k w_k / rho] of one point
// (flamelets/equations.py:313-389) and, if jout != nullptr, its Jacobian
// dS/dphi (NV x NV, row-major) by the chain rule through T(h, Y) and
// rho(T, Y) = p W/(R T):
//   dT/dh = 1/cp, dT/dY_m = -h_m/cp, drho/dh = -rho/T dT/dh,
//   drho/dY_m = -rho W/W_m - rho/T dT/dY_m, dC_k = (Y_k/W_k) drho + rho/W_k dY_k.
__device__ __forceinline__ bool point_source(const double h,
                                             const double (&Y)[NS],
                                             const double vd,
                                             const double tguess,
                                             const double p, const double tol,
                                             const int maxit, double& T,
                                             double (&S)[NV],
                                             double* __restrict__ jout) {
  const bool ok = temperature_from_h(h, Y, tguess, tol, maxit, T);
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
    const double dTdh = 1.0 / cp;
    const double rT = rho / T;
    const double drdh = -rT * dTdh;
    double dTdY[NS], drdY[NS];
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      dTdY[m] = -(kGasConstant * T * hrt[m] * mech::kInvW[m]) * dTdh;
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
  const bool ok = point_source(h, Y, vd ? vd[i] : 0.0, tguess[i * tg_stride],
                               p[i * p_stride], tol, maxit, T, s,
                               J ? J + i * NV * NV : nullptr);
#pragma unroll
  for (int v = 0; v < NV; ++v) S[v * lds + i] = s[v];
  if (T_out) T_out[i] = T;
  if (!ok && n_fail) atomicAdd(n_fail, 1);
}

// Residual and block Jacobian of the steady (mode 0) or Crank-Nicolson
// (mode 1) flamelet equations for a batch of flamelets whose points are
// concatenated; point i belongs to system sys[i] = s with rows
// [off[s], off[s+1]).  (flamelets/equations.py:167-282, solver.py:253-306)
//   R_i = chi_i/2 (a_i phi_{i-1} + b_i phi_i + c_i phi_{i+1}) + m_i S_i - BC_i
//   C_i = chi_i/2 b_i I + m_i J_i,  L_i = chi_{i+1}/2 a_{i+1},  U_i = chi_i/2 c_i
// CN: R' = phi - phi_prev - dt/2 (R + R_prev), C' = I - dt/2 C, L' = -dt/2 L,
// U' = -dt/2 U.  The interior mask m_i is 0 on the two end points.
// Outputs: rhs (P, NV) = -R (AoS), blocks in place of J, L, U (P,), and
// optionally R itself (NV, ldp) when r_out != nullptr.
__global__ void __launch_bounds__(256)
k_flamelet_assemble(const double* __restrict__ phi, int64_t ldp,
                    const double* __restrict__ S, int64_t lds,
                    double* __restrict__ J, const double* __restrict__ chi,
                    const double* __restrict__ la,
                    const double* __restrict__ lb,
                    const double* __restrict__ lc,
                    const int* __restrict__ sys,
                    const int64_t* __restrict__ off,
                    const double* __restrict__ bc,  // (nsys, 2, NV)
                    const int* __restrict__ modes,  // (nsys,): 0 steady, 1 CN
                    const double* __restrict__ phi_prev,
                    const double* __restrict__ r_prev,
                    const double* __restrict__ dt,  // (nsys,)
                    double* __restrict__ rhs, double* __restrict__ L,
                    double* __restrict__ U, double* __restrict__ r_out,
                    int64_t n) {
  const int64_t i = cell_index();
  if (i >= n) return;
  const int s = sys[i];
  const bool first = (i == off[s]);
  const bool last = (i == off[s + 1] - 1);
  const double hc = 0.5 * chi[i];
  const double ai = la[i], bi = lb[i], ci = lc[i];
  const double mask = (first || last) ? 0.0 : 1.0;
  const int mode = modes ? modes[s] : 0;
  const double hdt = (mode == 1) ? 0.5 * dt[s] : 0.0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double lap = bi * phi[v * ldp + i];
    if (!first) lap = fma(ai, phi[v * ldp + i - 1], lap);
    if (!last) lap = fma(ci, phi[v * ldp + i + 1], lap);
    double r = fma(hc, lap, mask * S[v * lds + i]);
    if (first) r -= bc[(s * 2 + 0) * NV + v];
    if (last) r -= bc[(s * 2 + 1) * NV + v];
    if (r_out) r_out[v * ldp + i] = r;
    if (mode == 1)
      r = phi[v * ldp + i] - phi_prev[v * ldp + i] -
          hdt * (r + r_prev[v * ldp + i]);
    rhs[i * NV + v] = -r;
  }
  double* blk = J + i * NV * NV;
#pragma unroll
  for (int rr = 0; rr < NV; ++rr) {
#pragma unroll
    for (int cc = 0; cc < NV; ++cc) {
      double val = mask * blk[rr * NV + cc];
      if (rr == cc) val = fma(hc, bi, val);
      if (mode == 1) val = (rr == cc ? 1.0 : 0.0) - hdt * val;
      blk[rr * NV + cc] = val;
    }
  }
  const double lo = last ? 0.0 : 0.5 * chi[i + 1] * la[i + 1];
  const double up = last ? 0.0 : hc * ci;
  L[i] = (mode == 1) ? -hdt * lo : lo;
  U[i] = (mode == 1) ? -hdt * up : up;
}

// phi += x for points of active systems; x is (P, NV) AoS.
__global__ void __launch_bounds__(256)
k_flamelet_update(double* __restrict__ phi, int64_t ldp,
                  const double* __restrict__ x, const int* __restrict__ sys,
                  const int* __restrict__ active, int64_t n) {
  const int64_t i = cell_index();
  if (i >= n) return;
  if (active && !active[sys[i]]) return;
#pragma unroll
  for (int v = 0; v < NV; ++v) phi[v * ldp + i] += x[i * NV + v];
}

// Deterministic per-system sums of squares of x (P, NV), split into the
// enthalpy row and the species rows: out[2 s] = sum v_h^2, out[2 s + 1] =
// sum v_Y^2.  One block per system, fixed-order strided partial sums and a
// fixed reduction tree, so a system's value does not depend on the batch.
__global__ void __launch_bounds__(256)
k_system_norm2(const double* __restrict__ x, const int64_t* __restrict__ off,
               double* __restrict__ out) {
  __shared__ double part[2][256];
  const int s = blockIdx.x;
  const int64_t b = off[s] * NV, e = off[s + 1] * NV;
  double ah = 0.0, ay = 0.0;
  for (int64_t j = b + threadIdx.x; j < e; j += 256) {
    const double q = x[j] * x[j];
    if ((j - b) % NV == 0)
      ah += q;
    else
      ay += q;
  }
  part[0][threadIdx.x] = ah;
  part[1][threadIdx.x] = ay;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      part[0][threadIdx.x] += part[0][threadIdx.x + w];
      part[1][threadIdx.x] += part[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * s] = part[0][0];
    out[2 * s + 1] = part[1][0];
  }
}

}  // namespace pyro_fl

// ---------------------------------------------------------------------------
// C ABI (mechanisms with NS <= 16)
// ---------------------------------------------------------------------------
PYRO_API int pyro_flamelet_source(const double* phi, int64_t ldp,
                                  const double* vd, const double* tguess,
                                  int64_t tg_stride, const double* p,
                                  int64_t p_stride, double* S, int64_t lds,
                                  double* T_out, double* J, int64_t n,
                                  double tol, int maxit, int* n_fail,
                                  hipStream_t stream) {
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldp < n || lds < n || !phi || !tguess || !p || !S)
    return pyro::kInvalidArgument;
  pyro_fl::k_flamelet_source<<<pyro::grid_for(n, 256), 256, 0, stream>>>(
      phi, ldp, vd, tguess, tg_stride, p, p_stride, S, lds, T_out, J, n, tol,
      maxit, n_fail);
  return pyro::launch_status();
}

PYRO_API int pyro_flamelet_assemble(
    const double* phi, int64_t ldp, const double* S, int64_t lds, double* J,
    const double* chi, const double* la, const double* lb, const double* lc,
    const int* sys, const int64_t* off, const double* bc, const int* modes,
    const double* phi_prev, const double* r_prev, const double* dt,
    double* rhs, double* L, double* U, double* r_out, int64_t n,
    hipStream_t stream) {
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldp < n || lds < n || !phi || !S || !J || !chi || !la || !lb ||
      !lc || !sys || !off || !bc || !rhs || !L || !U ||
      (modes && (!phi_prev || !r_prev || !dt)))
    return pyro::kInvalidArgument;
  pyro_fl::k_flamelet_assemble<<<pyro::grid_for(n, 256), 256, 0, stream>>>(
      phi, ldp, S, lds, J, chi, la, lb, lc, sys, off, bc, modes, phi_prev,
      r_prev, dt, rhs, L, U, r_out, n);
  return pyro::launch_status();
}

PYRO_API int pyro_flamelet_thomas(const double* L, const double* C,
                                  const double* U, const double* R, double* X,
                                  const int64_t* off, int64_t nsys,
                                  double* work, int* n_singular,
                                  hipStream_t stream) {
  if (nsys == 0) return pyro::kOk;
  if (nsys < 0 || !L || !C || !U || !R || !X || !off || !work)
    return pyro::kInvalidArgument;
  pyro_bt::k_block_thomas<pyro_fl::NV, true>
      <<<static_cast<unsigned>(nsys), 64, 0, stream>>>(L, C, U, R, X, off,
                                                       work, n_singular);
  return pyro::launch_status();
}

PYRO_API int pyro_flamelet_update(double* phi, int64_t ldp, const double* x,
                                  const int* sys, const int* active,
                                  const int64_t* off, int64_t nsys,
                                  double* norm2, int64_t n,
                                  hipStream_t stream) {
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldp < n || !phi || !x || !sys) return pyro::kInvalidArgument;
  if (norm2) {
    if (!off || nsys <= 0) return pyro::kInvalidArgument;
    pyro_fl::k_system_norm2<<<static_cast<unsigned>(nsys), 256, 0, stream>>>(
        x, off, norm2);
  }
  pyro_fl::k_flamelet_update<<<pyro::grid_for(n, 256), 256, 0, stream>>>(
      phi, ldp, x, sys, active, n);
  return pyro::launch_status();
}
