"""Numpy restatement of the generated class's transport methods
(``pyrometheus/codegen/python.py:564-641``) over given fit coefficients.

TEST INFRASTRUCTURE ONLY (see :mod:`oracle`).

The reference template renders one expression per species (or pair) from
``chem_expr.py:190-267``: the viscosity, conductivity and binary-diffusivity
polynomials in ln T, Wilke's mixture rule for the viscosity, the average of
the arithmetic and harmonic means for the conductivity and the
mixture-averaged diffusivity with its pure-species branch.  The coefficients
are arguments here (the fits Cantera would supply; the device library
reports the ones it was built with), so this checks the evaluation, not the
fits.
"""

from __future__ import annotations

import numpy as np


class Transport:
    def __init__(self, mech, visc, cond, diff):
        self.mech = mech
        self.visc = np.asarray(visc)   # (ns, 5)
        self.cond = np.asarray(cond)   # (ns, 5)
        self.diff = np.asarray(diff)   # (ns, ns, 5)
        self.w = mech.molecular_weights
        self.ns = mech.n_species

    @staticmethod
    def _poly(c, t):
        lt = np.log(t)
        return c[0] + c[1] * lt + c[2] * lt ** 2 + c[3] * lt ** 3 + \
            c[4] * lt ** 4

    def get_species_viscosities(self, temperature):
        return np.array([np.sqrt(temperature) * self._poly(c, temperature) ** 2
                         for c in self.visc])

    def get_species_thermal_conductivities(self, temperature):
        return np.array([np.sqrt(temperature) * self._poly(c, temperature)
                         for c in self.cond])

    def get_species_binary_mass_diffusivities(self, temperature):
        return np.array([[np.sqrt(temperature) * temperature
                          * self._poly(self.diff[i, j], temperature)
                          for j in range(self.ns)] for i in range(self.ns)])

    def _mole_fractions(self, mass_fractions):
        mmw = 1.0 / sum(mass_fractions[k] / self.w[k]
                        for k in range(self.ns))
        return mmw, np.array([mass_fractions[k] * mmw / self.w[k]
                              for k in range(self.ns)])

    def get_mixture_viscosity_mixavg(self, temperature, mass_fractions):
        mmw, x = self._mole_fractions(mass_fractions)
        mu = self.get_species_viscosities(temperature)
        w = self.w
        mix_rule_f = np.array([
            sum(x[j] * (1 + np.sqrt((mu[sp] / mu[j])
                                    * np.sqrt(w[j] / w[sp]))) ** 2
                / np.sqrt(8 * (1 + (w[sp] / w[j])))
                for j in range(self.ns))
            for sp in range(self.ns)])
        return sum(x * mu / mix_rule_f)

    def get_mixture_thermal_conductivity_mixavg(self, temperature,
                                                mass_fractions):
        mmw, x = self._mole_fractions(mass_fractions)
        lam = self.get_species_thermal_conductivities(temperature)
        return 0.5 * (sum(x * lam) + 1 / sum(x / lam))

    def get_species_mass_diffusivities_mixavg(self, pressure, temperature,
                                              mass_fractions):
        mmw, x = self._mole_fractions(mass_fractions)
        bd = self.get_species_binary_mass_diffusivities(temperature)
        x_sum = np.array([sum(x[i] / bd[i][j] for i in range(self.ns))
                          for j in range(self.ns)])
        denom = np.array([x_sum[s] - x[s] / bd[s][s]
                          for s in range(self.ns)])
        return np.array([
            np.where(denom[s] > 0,
                     (mmw - x[s] * self.w[s]) / (pressure * mmw * denom[s]),
                     bd[s][s] / pressure)
            for s in range(self.ns)])
