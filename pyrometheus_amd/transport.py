"""Transport-property fits for the generated class.

The reference's generated ``Thermochemistry`` evaluates per-species
polynomials in ln T that Cantera fits from kinetic theory
(``sol.get_viscosity_polynomial``, ``get_thermal_conductivity_polynomial``,
``get_binary_diff_coeffs_polynomial``; forms in ``chem_expr.py:190-243``):

    mu_k     = sqrt(T) (sum_n a_kn (ln T)^n)^2
    lambda_k = sqrt(T)  sum_n b_kn (ln T)^n
    D_ij p   = T^1.5    sum_n d_ijn (ln T)^n            (n = 0..4)

Cantera is not available here, so the fits are made from the mechanism's
transport data (YAML ``transport`` entries) with published kinetic theory:

* Chapman-Enskog viscosity and binary diffusion coefficient of dilute gases
  (Hirschfelder, Curtiss, Bird, *Molecular Theory of Gases and Liquids*,
  1954; Kee, Coltrin, Glarborg, *Chemically Reacting Flow*, 2003, ch. 12):
      mu   = (5/16) sqrt(pi m k T) / (pi sigma^2 Omega22*)
      D_ij = (3/16) sqrt(2 pi (k T)^3 / m_ij) / (p pi sigma_ij^2 Omega11*)
  with m_ij the reduced mass, sigma_ij = (sigma_i + sigma_j)/2 and
  eps_ij = sqrt(eps_i eps_j);
* the collision integrals by the correlations of Neufeld, Janzen and Aziz
  (J. Chem. Phys. 57, 1100, 1972) in the reduced temperature T* = T/eps,
  with Brokaw's polar correction (Ind. Eng. Chem. Process Des. Dev. 8, 240,
  1969): + 0.2 delta*^2 / T* (Omega22*), + 0.19 delta*^2 / T* (Omega11*),
  delta* = mu_D^2 / (2 eps sigma^3) (Gaussian units);
* the modified Eucken conductivity lambda = (mu / W)(1.32 c_v + 1.77 R)
  (Poling, Prausnitz, O'Connell, *The Properties of Gases and Liquids*, 5th
  ed., 10-3), c_v from the species' NASA-7 fit; 15/4 (R/W) mu for atoms;
* least-squares fits, relative-error weighted, of degree 4 in ln T over
  50 temperatures spanning the mechanism's common NASA-7 range.

These are not Cantera's fits (Cantera uses the Monchick-Mason collision
integral tables and its own conductivity model), so property values differ
from Cantera's at the percent level; every method of the generated class is
evaluated exactly as the reference's template evaluates its fits.
"""

from __future__ import annotations

from typing import Tuple

import numpy as np

from .mechanism import GAS_CONSTANT, Mechanism

K_BOLTZMANN = 1.380649e-23       # J/K
AVOGADRO = 6.02214076e26         # 1/kmol
DEBYE = 3.33564e-30              # C m
EPS0 = 8.8541878128e-12          # F/m


def omega22(tstar, dstar=0.0):
    """Omega(2,2)* (Neufeld et al. 1972) with Brokaw's polar term."""
    t = np.asarray(tstar, dtype=float)
    return (1.16145 * t ** -0.14874 + 0.52487 * np.exp(-0.77320 * t)
            + 2.16178 * np.exp(-2.43787 * t) + 0.2 * dstar ** 2 / t)


def omega11(tstar, dstar=0.0):
    """Omega(1,1)* (Neufeld et al. 1972) with Brokaw's polar term."""
    t = np.asarray(tstar, dtype=float)
    return (1.06036 * t ** -0.15610 + 0.19300 * np.exp(-0.47635 * t)
            + 1.03587 * np.exp(-1.52996 * t)
            + 1.76474 * np.exp(-3.89411 * t) + 0.19 * dstar ** 2 / t)


def _params(mech: Mechanism):
    """Per species: mass [kg], sigma [m], eps [J], delta*, geometry."""
    out = []
    for s in mech.species:
        if s.transport is None:
            raise ValueError(f"species {s.name} has no transport data")
        geom, dia, well, dip, _pol, _rot = s.transport
        sigma = dia * 1e-10
        eps = well * K_BOLTZMANN
        mu_d = dip * DEBYE
        dstar = mu_d ** 2 / (4 * np.pi * EPS0) / (2 * eps * sigma ** 3)
        out.append((s.weight / AVOGADRO, sigma, eps, dstar, geom))
    return out


def species_viscosity(mech: Mechanism, T):
    """Chapman-Enskog viscosity [Pa s] of every species, (ns, len(T))."""
    T = np.atleast_1d(np.asarray(T, dtype=float))
    rows = []
    for m, sigma, eps, dstar, _ in _params(mech):
        om = omega22(K_BOLTZMANN * T / eps, dstar)
        rows.append(5.0 / 16.0 * np.sqrt(np.pi * m * K_BOLTZMANN * T)
                    / (np.pi * sigma ** 2 * om))
    return np.array(rows)


def binary_diffusivity_times_p(mech: Mechanism, T):
    """Chapman-Enskog D_ij p [m^2/s Pa], (ns, ns, len(T)), symmetric."""
    T = np.atleast_1d(np.asarray(T, dtype=float))
    prm = _params(mech)
    ns = len(prm)
    out = np.empty((ns, ns, T.size))
    for i in range(ns):
        for j in range(i, ns):
            mi, si, ei, di, _ = prm[i]
            mj, sj, ej, dj, _ = prm[j]
            mij = mi * mj / (mi + mj)
            sij = 0.5 * (si + sj)
            eij = np.sqrt(ei * ej)
            dij = np.sqrt(di * dj)
            om = omega11(K_BOLTZMANN * T / eij, dij)
            d = (3.0 / 16.0 * np.sqrt(2 * np.pi * (K_BOLTZMANN * T) ** 3
                                      / mij) / (np.pi * sij ** 2 * om))
            out[i, j] = out[j, i] = d
    return out


def _cp_r(mech: Mechanism, T):
    rows = []
    for s in mech.species:
        c = np.where((T > s.t_mid)[None, :], np.array(s.high)[:, None],
                     np.array(s.low)[:, None])
        rows.append(c[0] + T * (c[1] + T * (c[2] + T * (c[3] + T * c[4]))))
    return np.array(rows)


def species_conductivity(mech: Mechanism, T):
    """Modified Eucken conductivity [W/(m K)], (ns, len(T))."""
    T = np.atleast_1d(np.asarray(T, dtype=float))
    mu = species_viscosity(mech, T)
    cv = (_cp_r(mech, T) - 1.0) * GAS_CONSTANT  # J/(kmol K)
    W = mech.molecular_weights[:, None]
    atom = np.array([s.transport[0] == "atom" for s in mech.species])[:, None]
    eucken = mu / W * (1.32 * cv + 1.77 * GAS_CONSTANT)
    return np.where(atom, 3.75 * GAS_CONSTANT * mu / W, eucken)


def fit_range(mech: Mechanism) -> Tuple[float, float]:
    """The temperatures every species' NASA-7 fit covers."""
    return (max(s.t_min for s in mech.species),
            min(s.t_max for s in mech.species))


def _fit(logt, y):
    """Degree-4 polynomial in ln T minimising the relative error;
    coefficients in ascending order, as the generated code reads them."""
    V = np.vander(logt, 5, increasing=True)
    w = 1.0 / np.abs(y)
    c, *_ = np.linalg.lstsq(V * w[:, None], y * w, rcond=None)
    return c


def transport_fits(mech: Mechanism, npts: int = 50):
    """(viscosity (ns, 5), conductivity (ns, 5), diffusivity (ns, ns, 5))
    polynomial coefficients in the forms of ``chem_expr.py:190-243``."""
    t0, t1 = fit_range(mech)
    T = np.linspace(t0, t1, npts)
    logt = np.log(T)
    mu = species_viscosity(mech, T)
    lam = species_conductivity(mech, T)
    dp = binary_diffusivity_times_p(mech, T)
    ns = mech.n_species
    visc = np.array([_fit(logt, np.sqrt(mu[k]) / T ** 0.25)
                     for k in range(ns)])
    cond = np.array([_fit(logt, lam[k] / np.sqrt(T)) for k in range(ns)])
    diff = np.empty((ns, ns, 5))
    for i in range(ns):
        for j in range(i, ns):
            diff[i, j] = diff[j, i] = _fit(logt, dp[i, j] / T ** 1.5)
    return visc, cond, diff


def has_transport(mech: Mechanism) -> bool:
    return all(s.transport is not None for s in mech.species)
