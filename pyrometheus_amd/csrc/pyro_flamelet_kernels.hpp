// Flamelet kernels of a generated mechanism that do not depend on how the
// source Jacobian is formed: residual/block assembly, block Thomas with the
// scalar Laplacian couplings, update and per-flamelet norms.
//
// Included at the end of every generated translation unit after
// pyro_thermo_kernels.hpp.
//
// State layout: phi is (NV, ld) with row 0 the mass enthalpy h and rows
// 1..NS the mass fractions (pyrometheus/flamelets/state.py:297-316).
// Jacobian blocks are (P, NV, NV) row-major per point, the layout the block
// Thomas kernel reads.
#pragma once

#include "pyro_block_thomas.hpp"
#include "pyro_spike_mfma.hpp"

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

// Newton on h(T, Y) = h as the reference's JAX loop runs it
// (flamelets/make_pyro.py:255-299, a lax.while_loop): the step is computed,
// and the loop stops before taking it once |step| <= tol.  Alongside, the
// forward-mode tangents th = dT/dh and tY[m] = dT/dY_m are carried through
// the iterations actually taken -- what jax.jacfwd through that loop
// differentiates (equations.py:124-129).  With step = (h - H)/Cp,
// H = R T sum yw_k h_k/RT, Cp = R sum yw_k cp_k/R, one iteration maps
//   th    <- 1/Cp - step dCp/dT th / Cp
//   tY[m] <- -(R T h_m/RT / W_m)/Cp - step (dCp/dT tY[m] + R cp_m/R / W_m)/Cp;
// a guess that already meets the test takes no iteration and leaves the
// tangents 0 (the reference's Jacobian then has no dT/dphi term).
__device__ __forceinline__ bool temperature_from_h_tangent(
    double h, const double (&Y)[NS], double t, double tol, int maxit,
    double& T, double& th, double (&tY)[NS]) {
  th = 0.0;
#pragma unroll
  for (int m = 0; m < NS; ++m) tY[m] = 0.0;
  for (int it = 0; it < maxit; ++it) {
    double hrt[NS], cpr[NS], dcp[NS];
    mech::species_h_rt(t, 1.0 / t, hrt);
    mech::species_cp_r(t, cpr);
    mech::species_dcp_r(t, dcp);
    double sh = 0.0, sc = 0.0, sd = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const double yw = Y[k] * mech::kInvW[k];
      sh = fma(yw, hrt[k], sh);
      sc = fma(yw, cpr[k], sc);
      sd = fma(yw, dcp[k], sd);
    }
    const double cp = kGasConstant * sc;
    const double step = (h - kGasConstant * t * sh) / cp;
    if (!(fabs(step) > tol)) {
      T = t;
      return true;
    }
    const double icp = 1.0 / cp;
    const double sdc = step * kGasConstant * sd * icp;  // step dCp/dT / Cp
    th = fma(-sdc, th, icp);
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      const double dh = kGasConstant * t * hrt[m] * mech::kInvW[m];
      const double dc = kGasConstant * cpr[m] * mech::kInvW[m];
      tY[m] = fma(-sdc, tY[m], -(dh + step * dc) * icp);
    }
    t += step;
  }
  T = t;
  return false;
}


// Residual and block Jacobian of the steady (mode 0) or Crank-Nicolson
// (mode 1) flamelet equations for a batch of flamelets whose points are
// concatenated; point i belongs to system sys[i] = s with rows
// [off[s], off[s+1]).  (flamelets/equations.py:167-282, solver.py:253-306)
//   R_i = chi_i/2 (a_i phi_{i-1} + b_i phi_i + c_i phi_{i+1}) + m_i S_i - BC_i
// evaluated on interior points as a_i (phi_{i-1} - phi_i) + c_i (phi_{i+1} -
// phi_i) (b_i = -(a_i + c_i)): the neighbour differences are exact for a
// smooth profile, so the enthalpy row (h ~ 1e5 J/kg, linear in Z, exact
// Laplacian 0) keeps a rounding floor set by the differences, not by h.
//   C_i = chi_i/2 b_i I + m_i J_i,  L_i = chi_{i+1}/2 a_{i+1},  U_i = chi_i/2 c_i
// CN: R' = phi - phi_prev - dt/2 (R + R_prev), C' = I - dt/2 C, L' = -dt/2 L,
// U' = -dt/2 U.  The interior mask m_i is 0 on the two end points.
// Outputs: rhs (P, NV) = -R (AoS), blocks in place of J, L, U (P,), and
// optionally R itself (NV, ldp) when r_out != nullptr.
__global__ void __launch_bounds__(256)
k_flamelet_assemble(const double* __restrict__ phi, int64_t ldp,
                    const double* __restrict__ S, int64_t lds,
                    double* __restrict__ J, const double* __restrict__ chi,
                    const double* __restrict__ chi0,
                    const double* __restrict__ la,
                    const double* __restrict__ lb,
                    const double* __restrict__ lc,
                    const int* __restrict__ sys,
                    const int64_t* __restrict__ off,
                    const double* __restrict__ bc,  // (nsys, 2, NV)
                    // (nsys,): bit 0 Crank-Nicolson (else steady), bit 1
                    // the dissipation rate chi0 (else chi), bit 2 (with
                    // bit 0) a time step starts here: phi_prev = phi and
                    // r_prev = R(phi) are stored first
                    const int* __restrict__ modes,
                    double* __restrict__ phi_prev,
                    double* __restrict__ r_prev,
                    const double* __restrict__ dt,  // (nsys,)
                    double* __restrict__ rhs, double* __restrict__ L,
                    double* __restrict__ U, double* __restrict__ r_out,
                    int64_t n) {
  const int64_t i = cell_index();
  if (i >= n) return;
  const int s = sys[i];
  const bool first = (i == off[s]);
  const bool last = (i == off[s + 1] - 1);
  const int mode0 = modes ? modes[s] : 0;
  const double* __restrict__ chs = ((mode0 & 2) && chi0) ? chi0 : chi;
  const double hc = 0.5 * chs[i];
  const double ai = la[i], bi = lb[i], ci = lc[i];
  const double mask = (first || last) ? 0.0 : 1.0;
  const int mode = mode0 & 1;
  const bool start = (mode0 & 5) == 5;
  const double hdt = (mode == 1) ? 0.5 * dt[s] : 0.0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double p0 = phi[v * ldp + i];
    double lap;
    if (first || last)
      lap = bi * p0;
    else
      lap = fma(ai, phi[v * ldp + i - 1] - p0,
                ci * (phi[v * ldp + i + 1] - p0));
    double r = fma(hc, lap, mask * S[v * lds + i]);
    if (first) r -= bc[(s * 2 + 0) * NV + v];
    if (last) r -= bc[(s * 2 + 1) * NV + v];
    if (r_out) r_out[v * ldp + i] = r;
    if (mode == 1) {
      double pp, rp;
      if (start) {
        pp = phi[v * ldp + i];
        rp = r;
        phi_prev[v * ldp + i] = pp;
        r_prev[v * ldp + i] = rp;
      } else {
        pp = phi_prev[v * ldp + i];
        rp = r_prev[v * ldp + i];
      }
      r = phi[v * ldp + i] - pp - hdt * (r + rp);
    }
    rhs[i * NV + v] = -r;
  }
  const double lo = last ? 0.0 : 0.5 * chs[i + 1] * la[i + 1];
  const double up = last ? 0.0 : hc * ci;
  L[i] = (mode == 1) ? -hdt * lo : lo;
  U[i] = (mode == 1) ? -hdt * up : up;
}

// Central blocks in place, one thread per entry of J (P, NV, NV) so that
// consecutive threads touch consecutive doubles: J = m J + hc b I, and
// I - dt/2 (that) for Crank-Nicolson, with the point's scalars as in
// k_flamelet_assemble.
__global__ void __launch_bounds__(256)
k_flamelet_blocks(double* __restrict__ J, const double* __restrict__ chi,
                  const double* __restrict__ chi0,
                  const double* __restrict__ lb, const int* __restrict__ sys,
                  const int64_t* __restrict__ off,
                  const int* __restrict__ modes,
                  const double* __restrict__ dt, int64_t n) {
  constexpr int64_t NB = static_cast<int64_t>(NV) * NV;
  const int64_t f = cell_index();
  if (f >= n * NB) return;
  const int64_t i = f / NB;
  const int e = static_cast<int>(f - i * NB);
  const int rr = e / NV, cc = e - rr * NV;
  const int s = sys[i];
  const bool first = (i == off[s]);
  const bool last = (i == off[s + 1] - 1);
  const int mode0 = modes ? modes[s] : 0;
  const double* __restrict__ chs = ((mode0 & 2) && chi0) ? chi0 : chi;
  const double hc = 0.5 * chs[i];
  const double mask = (first || last) ? 0.0 : 1.0;
  const int mode = mode0 & 1;
  const double hdt = (mode == 1) ? 0.5 * dt[s] : 0.0;
  double val = mask * J[f];
  if (rr == cc) val = fma(hc, lb[i], val);
  if (mode == 1) val = (rr == cc ? 1.0 : 0.0) - hdt * val;
  J[f] = val;
}

// phi += x for points of active systems (x is (P, NV) AoS) and, if tg is
// given, tg = T_new for the same points.
__global__ void __launch_bounds__(256)
k_flamelet_update(double* __restrict__ phi, int64_t ldp,
                  const double* __restrict__ x, const int* __restrict__ sys,
                  const int* __restrict__ active,
                  const double* __restrict__ T_new, double* __restrict__ tg,
                  int64_t n) {
  const int64_t i = cell_index();
  if (i >= n) return;
  if (active && !active[sys[i]]) return;
#pragma unroll
  for (int v = 0; v < NV; ++v) phi[v * ldp + i] += x[i * NV + v];
  // the next temperature guess of an updated point is this update's T
  if (tg) tg[i] = T_new[i];
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

// R T (R = R_u / W, the specific gas constant) of every point and its
// gradient with respect to phi = [h; Y] (equations.py:518-535 and the jacfwd
// of it, :140-149), for the compressible-EOS adjoint solves.  T(h, Y) by
// Newton from tguess: kRef the reference's loop with its forward-mode
// tangent (temperature_from_h_tangent), otherwise the converged T and the
// exact dT/dh = 1/cp, dT/dY_m = -h_m/cp.  rtj is (NV, ldj):
//   d(RT)/dh = R dT/dh,  d(RT)/dY_m = R_u T / W_m + R dT/dY_m.
// A point whose Newton did not converge gets NaN.
template <bool kRef>
__global__ void __launch_bounds__(256)
k_flamelet_eos_rt(const double* __restrict__ phi, int64_t ldp,
                  const double* __restrict__ tguess, int64_t tg_stride,
                  double* __restrict__ rt, double* __restrict__ rtj,
                  int64_t ldj, int64_t n, double tol, int maxit,
                  int* __restrict__ n_fail) {
  const int64_t i = cell_index();
  if (i >= n) return;
  double Y[NS];
  const double h = phi[i];
#pragma unroll
  for (int k = 0; k < NS; ++k) Y[k] = phi[(1 + k) * ldp + i];
  double T, th, tY[NS];
  bool ok;
  if constexpr (kRef) {
    ok = temperature_from_h_tangent(h, Y, tguess[i * tg_stride], tol, maxit,
                                    T, th, tY);
  } else {
    ok = temperature_from_h(h, Y, tguess[i * tg_stride], tol, maxit, T);
    double hrt[NS], cpr[NS];
    mech::species_h_rt(T, 1.0 / T, hrt);
    mech::species_cp_r(T, cpr);
    double sc = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k) sc = fma(Y[k] * mech::kInvW[k], cpr[k], sc);
    th = 1.0 / (kGasConstant * sc);
#pragma unroll
    for (int m = 0; m < NS; ++m)
      tY[m] = -(kGasConstant * T * hrt[m] * mech::kInvW[m]) * th;
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < NS; ++k) s = fma(mech::kInvW[k], Y[k], s);
  const double rs = kGasConstant * s;  // R = R_u / W
  const double bad = __builtin_nan("");
  rt[i] = ok ? rs * T : bad;
  if (rtj) {
    rtj[i] = ok ? rs * th : bad;
#pragma unroll
    for (int m = 0; m < NS; ++m)
      rtj[(1 + m) * ldj + i] =
          ok ? fma(rs, tY[m], kGasConstant * T * mech::kInvW[m]) : bad;
  }
  if (!ok && n_fail) atomicAdd(n_fail, 1);
}

// The adjoint operator of the steady flamelet Jacobian (equations.py:
// 451-516) in place of the source-Jacobian blocks J (P, NV, NV): central
// blocks b_i I + 2 J_i^T / chi_i on interior points and I on the two ends
// of each system, the Laplacian's own couplings L_i (row i+1 to x_i) =
// a_{i+1} where row i+1 is interior and U_i (row i to x_{i+1}) = c_i where
// row i is interior (0 elsewhere).  One thread per block entry; the thread
// of (r, c) with r <= c writes both (r, c) and (c, r).
__global__ void __launch_bounds__(256)
k_flamelet_adjoint_blocks(double* __restrict__ J,
                          const double* __restrict__ chi,
                          const double* __restrict__ la,
                          const double* __restrict__ lb,
                          const double* __restrict__ lc,
                          const int* __restrict__ sys,
                          const int64_t* __restrict__ off,
                          double* __restrict__ L, double* __restrict__ U,
                          int64_t n) {
  constexpr int64_t NB = static_cast<int64_t>(NV) * NV;
  const int64_t f = cell_index();
  if (f >= n * NB) return;
  const int64_t i = f / NB;
  const int e = static_cast<int>(f - i * NB);
  const int r = e / NV, c = e - r * NV;
  const int sy = sys[i];
  const bool first = (i == off[sy]);
  const bool last = (i == off[sy + 1] - 1);
  if (e == 0) {
    L[i] = (!last && i + 1 < off[sy + 1] - 1) ? la[i + 1] : 0.0;
    U[i] = (!first && !last) ? lc[i] : 0.0;
  }
  if (r > c) return;
  double* blk = J + i * NB;
  const double d = (r == c) ? 1.0 : 0.0;
  if (first || last) {
    blk[r * NV + c] = d;
    blk[c * NV + r] = d;
    return;
  }
  const double s = 2.0 / chi[i];
  const double jrc = blk[r * NV + c], jcr = blk[c * NV + r];
  blk[r * NV + c] = fma(s, jcr, d * lb[i]);
  if (r != c) blk[c * NV + r] = s * jrc;
}

}  // namespace pyro_fl

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
PYRO_API int pyro_flamelet_assemble(
    const double* phi, int64_t ldp, const double* S, int64_t lds, double* J,
    const double* chi, const double* chi0, const double* la,
    const double* lb, const double* lc,
    const int* sys, const int64_t* off, const double* bc, const int* modes,
    double* phi_prev, double* r_prev, const double* dt,
    double* rhs, double* L, double* U, double* r_out, int64_t n,
    hipStream_t stream) {
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldp < n || lds < n || !phi || !S || !J || !chi || !la || !lb ||
      !lc || !sys || !off || !bc || !rhs || !L || !U ||
      (modes && (!phi_prev || !r_prev || !dt)))
    return pyro::kInvalidArgument;
  pyro_fl::k_flamelet_assemble<<<pyro::grid_for(n, 256), 256, 0, stream>>>(
      phi, ldp, S, lds, J, chi, chi0, la, lb, lc, sys, off, bc, modes,
      phi_prev, r_prev, dt, rhs, L, U, r_out, n);
  pyro_fl::k_flamelet_blocks<<<pyro::grid_for(
                                   n * pyro_fl::NV * pyro_fl::NV, 256),
                               256, 0, stream>>>(J, chi, chi0, lb, sys, off,
                                                 modes, dt, n);
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
      <<<static_cast<unsigned>(nsys), 64, 0, stream>>>(
          L, C, U, R, X, off, work, INT64_MAX, n_singular);
  return pyro::launch_status();
}

PYRO_API int pyro_flamelet_update(double* phi, int64_t ldp, const double* x,
                                  const int* sys, const int* active,
                                  const double* T_new, double* tg,
                                  const int64_t* off, int64_t nsys,
                                  double* norm2, int64_t n,
                                  hipStream_t stream) {
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldp < n || !phi || !x || !sys || (tg && !T_new))
    return pyro::kInvalidArgument;
  if (norm2) {
    if (!off || nsys <= 0) return pyro::kInvalidArgument;
    pyro_fl::k_system_norm2<<<static_cast<unsigned>(nsys), 256, 0, stream>>>(
        x, off, norm2);
  }
  pyro_fl::k_flamelet_update<<<pyro::grid_for(n, 256), 256, 0, stream>>>(
      phi, ldp, x, sys, active, T_new, tg, n);
  return pyro::launch_status();
}

// Partitioned block Thomas for the flamelet blocks (nv <= 64): the same
// system as pyro_flamelet_thomas, with each system cut into partitions
// [pstart[p], pend[p]) of system psys[p] (consecutive within a system; the
// partitions of system s are roff[s] + s .. roff[s+1] + s), the unknowns at
// the interfaces between consecutive partitions found by an nv-sized block
// recursion over the partitions (roff: interface offsets per system;
// jbefore/jafter[p]: interface on either side of partition p or -1).  Work
// arrays: fact P*nv*(2nv+1), maps nparts*2*nv*(2nv+1),
// red nparts*nv*(2nv+2), Z nj*2nv doubles.
PYRO_API int pyro_flamelet_thomas_partitioned(
    const double* L, const double* C, const double* U, const double* R,
    double* X, const int64_t* off, int64_t nsys, const int64_t* pstart,
    const int64_t* pend, const int* psys, int64_t nparts,
    const int64_t* jbefore, const int64_t* jafter, const int64_t* roff,
    int64_t nj, double* fact, double* maps, double* red, double* Z,
    int* n_singular, hipStream_t stream) {
  if (nsys == 0 || nparts == 0) return pyro::kOk;
  if constexpr (pyro_fl::NV > 64) {
    return pyro::kUnsupported;
  } else {
    if (nsys < 0 || nparts < 0 || nj < 0 || !L || !C || !U || !R || !X ||
        !off || !pstart || !pend || !psys || !jbefore || !jafter || !roff ||
        !fact || !maps || (nj > 0 && (!red || !Z)))
      return pyro::kInvalidArgument;
    return pyro_spike::solve<pyro_fl::NV>(
        L, C, U, R, X, off, pstart, pend, psys, nparts, jbefore, jafter,
        roff, nsys, nj, fact, maps, red, Z, n_singular, stream);
  }
}

// Form of the partitioned solve for 16 < NV <= 63: 1 the blocked
// Gauss-Jordan sweep with MFMA trailing updates and the MFMA recursion
// (default), 0 the row-per-lane eight-wave kernels; a negative form only
// queries.  Returns the previous
// form, or -3 (unsupported) when the library's NV has one form.
PYRO_API int pyro_flamelet_thomas_sweep_form(int form) {
  if constexpr (pyro_fl::NV <= 16 || pyro_fl::NV > 63) {
    return pyro::kUnsupported;
  } else {
    const int prev = pyro_spike::sweep_form();
    if (form == 0 || form == 1) pyro_spike::sweep_form() = form;
    return prev;
  }
}

PYRO_API int pyro_flamelet_eos_rt(const double* phi, int64_t ldp,
                                  const double* tguess, int64_t tg_stride,
                                  double* rt, double* rtj, int64_t ldj,
                                  int64_t n, double tol, int maxit,
                                  int* n_fail, int semantics,
                                  hipStream_t stream) {
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldp < n || !phi || !tguess || !rt || (rtj && ldj < n) ||
      (semantics != 0 && semantics != 1))
    return pyro::kInvalidArgument;
  if (semantics == 1)
    pyro_fl::k_flamelet_eos_rt<true><<<pyro::grid_for(n, 256), 256, 0,
                                       stream>>>(
        phi, ldp, tguess, tg_stride, rt, rtj, ldj, n, tol, maxit, n_fail);
  else
    pyro_fl::k_flamelet_eos_rt<false><<<pyro::grid_for(n, 256), 256, 0,
                                        stream>>>(
        phi, ldp, tguess, tg_stride, rt, rtj, ldj, n, tol, maxit, n_fail);
  return pyro::launch_status();
}

PYRO_API int pyro_flamelet_adjoint_blocks(double* J, const double* chi,
                                          const double* la, const double* lb,
                                          const double* lc, const int* sys,
                                          const int64_t* off, double* L,
                                          double* U, int64_t n,
                                          hipStream_t stream) {
  if (n == 0) return pyro::kOk;
  if (n < 0 || !J || !chi || !la || !lb || !lc || !sys || !off || !L || !U)
    return pyro::kInvalidArgument;
  pyro_fl::k_flamelet_adjoint_blocks<<<pyro::grid_for(
                                           n * pyro_fl::NV * pyro_fl::NV,
                                           256),
                                       256, 0, stream>>>(
      J, chi, la, lb, lc, sys, off, L, U, n);
  return pyro::launch_status();
}
