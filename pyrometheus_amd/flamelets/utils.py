"""Problem set-up helpers (``pyrometheus/flamelets/utils.py:9-93``)."""

import numpy as np
from scipy.special import erfcinv


def bell_profile(z):
    """chi(Z) shape exp(-2 erfcinv(2Z)^2)."""
    return np.exp(-2 * erfcinv(2 * np.asarray(z)) ** 2)


def dissipation_profile(domain, z_st, chi_st):
    """chi_st bell(Z)/bell(Z_st) inside, identity-row values at the ends."""
    x = domain.x
    chi = np.empty_like(x)
    chi[1:-1] = chi_st * bell_profile(x[1:-1]) / bell_profile(z_st)
    chi[0], chi[-1] = domain.boundary_dissipation
    return chi


def stoichiometric_mixture_fraction(mech, y_ox, y_fu):
    """Bilger Z_st with the reference's element-weight table."""
    bilger = {"C": 2, "H": 0.5, "O": -1}
    wts_e = {"C": 12.011, "H": 1.00794, "O": 15.999, "Ar": 39.948,
             "N": 28.0133}
    w_s = mech.molecular_weights
    bw = np.array([bilger[e] / wts_e[e] if e in bilger else 0.0
                   for e in mech.elements])
    em = np.array([[mech.n_atoms(s, e) * wts_e.get(mech.elements[e], 1.0)
                    / w_s[s] for s in range(mech.n_species)]
                   for e in range(mech.n_elements)])
    beta_ox = bw @ (em @ np.asarray(y_ox))
    beta_fu = bw @ (em @ np.asarray(y_fu))
    return float(-beta_ox / (beta_fu - beta_ox))


def mole_to_mass(mech, x):
    w = mech.molecular_weights
    return w * x / np.sum(w * x)


def burke_schumann(mech, x, y_ox, y_fu, h_ox, h_fu, fuel="H2", oxidizer="O2",
                   product="H2O", nu_ox=0.5):
    """Closed-form initial state: enthalpy linear in Z; mass fractions from
    complete combustion fuel + nu_ox O2 -> product, linear on each side of
    the stoichiometric point.  Returns ((nv, N) array, Z_st)."""
    idx = mech.species_index
    z_st = stoichiometric_mixture_fraction(mech, y_ox, y_fu)
    w = mech.molecular_weights
    y_st = z_st * np.asarray(y_fu) + (1 - z_st) * np.asarray(y_ox)
    n_f = y_st[idx(fuel)] / w[idx(fuel)]
    n_o = y_st[idx(oxidizer)] / w[idx(oxidizer)]
    react = min(n_f, n_o / nu_ox)
    burnt = y_st.copy()
    burnt[idx(fuel)] -= react * w[idx(fuel)]
    burnt[idx(oxidizer)] -= nu_ox * react * w[idx(oxidizer)]
    burnt[idx(product)] += react * w[idx(product)]
    burnt = np.clip(burnt, 0.0, None)
    burnt /= burnt.sum()
    x = np.asarray(x)
    Y = np.empty((mech.n_species, len(x)))
    lean = x <= z_st
    f = np.where(lean, x / z_st, (x - z_st) / (1 - z_st))
    Y[:, lean] = (1 - f[lean]) * np.asarray(y_ox)[:, None] + f[lean] * burnt[:, None]
    rich = ~lean
    Y[:, rich] = (1 - f[rich]) * burnt[:, None] + f[rich] * np.asarray(y_fu)[:, None]
    h = h_ox + (h_fu - h_ox) * x
    return np.vstack([h[None, :], Y]), z_st
