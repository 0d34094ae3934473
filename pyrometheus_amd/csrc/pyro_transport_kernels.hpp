// Transport properties of a generated mechanism (the generated class's
// transport methods, pyrometheus/codegen/python.py:564-641, with the
// expressions of chem_expr.py:190-267).  Included after the generated
// tables:
//   mech::kTrVisc[NS][5], mech::kTrCond[NS][5]  species fits in ln T,
//   mech::kTrDiff[NS*NS][5]                      binary-diffusion fits, the
//                                                symmetric (i, j) table,
//   mech::kTrW[NS]                               molecular weights,
//   mech::kTrWilkeS[NS][NS] = (W_j / W_k)^(1/4),
//   mech::kTrWilkeD[NS][NS] = 1 / sqrt(8 (1 + W_k / W_j)).
//
// One cell per lane.  The double sums over species run with the outer
// species index k uniform across the wave (a runtime loop whose table rows
// come through scalar loads) and the inner index fully unrolled over the
// per-lane arrays (mole fractions and the like), so those stay in
// registers.
#pragma once

#include "pyro_common.hpp"

namespace pyro_tr {

using namespace pyro;
constexpr int NS = mech::NS;

__device__ __forceinline__ double poly5(cptr c, double l) {
  return fma(l, fma(l, fma(l, fma(l, c[4], c[3]), c[2]), c[1]), c[0]);
}

// sum a (ln T)^n of the viscosity fit: mu_k = sqrt(T) q_k^2
// (chem_expr.viscosity_polynomial_expr)
__device__ __forceinline__ double visc_q(int k, double l) {
  return poly5((cptr)&mech::kTrVisc[k][0], l);
}

// sqrt(T) sum b (ln T)^n -- chem_expr.conductivity_polynomial_expr
__device__ __forceinline__ double conductivity(int k, double sqt, double l) {
  return sqt * poly5((cptr)&mech::kTrCond[k][0], l);
}

// T^1.5 sum d (ln T)^n of the pair (a, b) -- diffusivity_polynomial_expr
// (D_ab times the pressure)
__device__ __forceinline__ double pair_diffusivity(int a, int b, double t15,
                                                   double l) {
  return t15 * poly5((cptr)&mech::kTrDiff[a * NS + b][0], l);
}

// which: 0 viscosities, 1 conductivities (out (NS, ld)); 2 binary
// diffusivities (out (NS, NS, ld)).
template <int WHICH>
__global__ void __launch_bounds__(256)
k_species_transport(const double* __restrict__ T, int64_t t_stride,
                    double* __restrict__ out, int64_t ld, int64_t n) {
  const int64_t i = cell_index();
  if (i >= n) return;
  const double t = T[i * t_stride];
  const double l = log(t), sqt = sqrt(t);
  if constexpr (WHICH == 0) {
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const double q = visc_q(k, l);
      out[k * ld + i] = sqt * (q * q);
    }
  } else if constexpr (WHICH == 1) {
#pragma unroll
    for (int k = 0; k < NS; ++k) out[k * ld + i] = conductivity(k, sqt, l);
  } else {
    const double t15 = sqt * t;
    for (int a = 0; a < NS; ++a) {
#pragma unroll
      for (int b = 0; b < NS; ++b)
        out[(static_cast<int64_t>(a) * NS + b) * ld + i] =
            pair_diffusivity(a, b, t15, l);
    }
  }
}

// Mole fractions x_k = (Y_k / W_k) mmw and the mixture molecular weight
// mmw = 1 / sum Y_k / W_k of a cell (the generated class's
// get_mixture_molecular_weight and get_mole_fractions).
__device__ __forceinline__ double mole_fractions(const double* __restrict__ Y,
                                                 int64_t ldY, int64_t i,
                                                 double (&x)[NS]) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    x[k] = mech::kInvW[k] * Y[k * ldY + i];
    s += x[k];
  }
  const double mmw = 1.0 / s;
#pragma unroll
  for (int k = 0; k < NS; ++k) x[k] *= mmw;
  return mmw;
}

// x_k for a wave-uniform k, the same arithmetic as mole_fractions
__device__ __forceinline__ double mole_fraction(const double* __restrict__ Y,
                                                int64_t ldY, int64_t i, int k,
                                                double mmw) {
  return (((cptr)mech::kInvWMem)[k] * Y[k * ldY + i]) * mmw;
}

// which: 0 mixture viscosity (Wilke), 1 mixture conductivity, 2 species
// mixture-averaged mass diffusivities (out (NS, ld)).
template <int WHICH>
__global__ void __launch_bounds__(256)
k_mixture_transport(const double* __restrict__ p, int64_t p_stride,
                    const double* __restrict__ T, int64_t t_stride,
                    const double* __restrict__ Y, int64_t ldY,
                    double* __restrict__ out, int64_t ld, int64_t n) {
  const int64_t i = cell_index();
  if (i >= n) return;
  const double t = T[i * t_stride];
  const double l = log(t), sqt = sqrt(t);
  double x[NS];
  const double mmw = mole_fractions(Y, ldY, i, x);
  if constexpr (WHICH == 0) {
    // sum_k x_k mu_k / phi_k with phi_k = sum_j x_j (1 + sqrt(mu_k / mu_j
    // sqrt(W_j / W_k)))^2 / sqrt(8 (1 + W_k / W_j)); with mu = sqrt(T) q^2
    // the square root is |q_k| (W_j / W_k)^(1/4) / |q_j|.
    double u[NS];  // 1 / |q_j|
#pragma unroll
    for (int j = 0; j < NS; ++j) u[j] = 1.0 / fabs(visc_q(j, l));
    double acc = 0.0;
    for (int k = 0; k < NS; ++k) {
      const double qk = fabs(visc_q(k, l));
      const cptr S = (cptr)&mech::kTrWilkeS[k][0];
      const cptr D = (cptr)&mech::kTrWilkeD[k][0];
      double phi = 0.0;
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const double b = fma(qk * S[j], u[j], 1.0);
        phi = fma(x[j] * b, b * D[j], phi);
      }
      const double xk = mole_fraction(Y, ldY, i, k, mmw);
      acc += xk * (sqt * (qk * qk)) / phi;
    }
    out[i] = acc;
  } else if constexpr (WHICH == 1) {
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const double c = conductivity(k, sqt, l);
      s1 += x[k] * c;
      s2 += x[k] / c;
    }
    out[i] = 0.5 * (s1 + 1.0 / s2);
  } else {
    // (mmw - x_k W_k) / (p mmw sum_{a != k} x_a / D_ak), or D_kk / p where
    // that sum is not positive (a pure species).  The reference forms the
    // sum as the full sum minus the diagonal term; here it is summed
    // directly (the same value without the cancellation; exactly zero for a
    // pure species, as there), in the same order of a, and D_ka = D_ak
    // lets each unordered pair's fit and reciprocal serve both rows:
    // sum_{a != k} x_a / D_ak = (1 / T^1.5) sum_{a != k} x_a / d_ak with
    // d_ak the fit's polynomial.
    const double pr = p[i * p_stride];
    const double t15 = sqt * t;
    double acc[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) acc[k] = 0.0;
    // ln T and the pair's accumulators pass an empty asm every kTrPin pairs:
    // later pairs' fits and reciprocals are not computed ahead (and kept
    // live beside the 2 NS per-species values)
    constexpr int kTrPin = 4;
    double lp = l;
#pragma unroll
    for (int a = 0; a < NS; ++a) {
#pragma unroll
      for (int b = a + 1; b < NS; ++b) {
        if ((b - a - 1) % kTrPin == 0) asm volatile("" : "+v"(lp));
        const double r = recip(poly5((cptr)&mech::kTrDiff[a * NS + b][0], lp));
        acc[a] = fma(x[b], r, acc[a]);
        acc[b] = fma(x[a], r, acc[b]);
        asm volatile("" : "+v"(acc[b]));
      }
      asm volatile("" : "+v"(acc[a]));
    }
    const double it15 = 1.0 / t15;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const double denom = acc[k] * it15;
      const double dkk = pair_diffusivity(k, k, t15, l);
      const double wk = ((cptr)mech::kTrW)[k];
      out[k * ld + i] = denom > 0.0
                            ? (mmw - x[k] * wk) / (pr * mmw * denom)
                            : dkk / pr;
    }
  }
}

}  // namespace pyro_tr

PYRO_API int pyro_species_transport(int which, const double* T,
                                    int64_t t_stride, double* out, int64_t ld,
                                    int64_t n, hipStream_t stream) {
  if (n == 0) return pyro::kOk;
  if (n < 0 || ld < n || !T || !out) return pyro::kInvalidArgument;
  const unsigned g = pyro::grid_for(n, 256);
  switch (which) {
    case 0:
      pyro_tr::k_species_transport<0><<<g, 256, 0, stream>>>(T, t_stride,
                                                             out, ld, n);
      break;
    case 1:
      pyro_tr::k_species_transport<1><<<g, 256, 0, stream>>>(T, t_stride,
                                                             out, ld, n);
      break;
    case 2:
      pyro_tr::k_species_transport<2><<<g, 256, 0, stream>>>(T, t_stride,
                                                             out, ld, n);
      break;
    default:
      return pyro::kInvalidArgument;
  }
  return pyro::launch_status();
}

PYRO_API int pyro_mixture_transport(int which, const double* p,
                                    int64_t p_stride, const double* T,
                                    int64_t t_stride, const double* Y,
                                    int64_t ldY, double* out, int64_t ld,
                                    int64_t n, hipStream_t stream) {
  if (n == 0) return pyro::kOk;
  if (n < 0 || ldY < n || !T || !Y || !out || (which == 2 && (!p || ld < n)))
    return pyro::kInvalidArgument;
  const unsigned g = pyro::grid_for(n, 256);
  switch (which) {
    case 0:
      pyro_tr::k_mixture_transport<0><<<g, 256, 0, stream>>>(
          p, p_stride, T, t_stride, Y, ldY, out, ld, n);
      break;
    case 1:
      pyro_tr::k_mixture_transport<1><<<g, 256, 0, stream>>>(
          p, p_stride, T, t_stride, Y, ldY, out, ld, n);
      break;
    case 2:
      pyro_tr::k_mixture_transport<2><<<g, 256, 0, stream>>>(
          p, p_stride, T, t_stride, Y, ldY, out, ld, n);
      break;
    default:
      return pyro::kInvalidArgument;
  }
  return pyro::launch_status();
}

// The fit coefficients the library was built with (host memory): which 0
// viscosity (NS, 5), 1 conductivity (NS, 5), 2 binary diffusion (NS, NS, 5).
PYRO_API int pyro_transport_fits(int which, double* out, int64_t n) {
  using pyro_tr::NS;
  if (!out) return pyro::kInvalidArgument;
  const int rows = which == 2 ? NS * NS : NS;
  if (which < 0 || which > 2 || n < static_cast<int64_t>(rows) * 5)
    return pyro::kInvalidArgument;
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < 5; ++c)
      out[r * 5 + c] = which == 0   ? mech::kTrViscHost[r][c]
                       : which == 1 ? mech::kTrCondHost[r][c]
                                    : mech::kTrDiffHost[r][c];
  return pyro::kOk;
}
