"""Pins for the CPU oracle (oracle/thermochem_np.py).

The reference's own tests compare against live Cantera, which is absent, so
the oracle is pinned by (a) the reference-held autoignition trajectory,
(b) closed forms written independently from the model equations
(``doc/model.rst``) and the raw YAML numbers, and (c) invariants.  Each check
is chosen so a plausible mistake (dropped factor, wrong sign, wrong unit
order, swapped branch) fails it; ``test_fixture_catches_planted_error``
demonstrates that for the strongest pin.
"""
import dataclasses
import math
import os

import numpy as np
import pytest

from oracle.thermochem_np import Thermochemistry
from pyrometheus_amd.mechanism import GAS_CONSTANT, ONE_ATM, Mechanism, bundled
from tests.helpers import random_states, scaled_err

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "autoignition.npz")


@pytest.fixture(scope="module")
def fixture():
    return np.load(FIXTURE)


@pytest.fixture(scope="module")
def sandiego():
    return Thermochemistry(bundled("sandiego"))


# ---------------------------------------------------------------------------
# (a) reference-held fixture: examples/free_flame/output/autoignition.h5
# ---------------------------------------------------------------------------
def test_fixture_energy_invariant(fixture, sandiego):
    """Constant-volume, adiabatic: e(T_i, Y_i) is the same on every stored row.
    Pins NASA-7 h/RT and the molecular weights."""
    for key, tol in (("explicit_state", 1e-14), ("implicit_state", 1e-12)):
        s = fixture[key]
        e = sandiego.get_mixture_internal_energy_mass(s[:, 9], s[:, :9].T)
        assert (e.max() - e.min()) / abs(e).max() < tol, key


def test_fixture_temperature(fixture, sandiego):
    """get_temperature(e_0, 1200 K, Y_i, do_energy=True) reproduces T_i."""
    s = fixture["explicit_state"]
    e0 = sandiego.get_mixture_internal_energy_mass(1200.0, s[0, :9])
    T = sandiego.get_temperature(e0 * np.ones(len(s)), 1200.0, s[:, :9].T,
                                 do_energy=True)
    assert np.max(np.abs(T - s[:, 9])) < 1e-9


def _rk4_window(gas, y, rho, e0, nsteps=12, h=1.0e-8):
    """One stored window of the fixture's generator: RK4 on
    dY/dt = W * wdot(rho, T(e0, Y), Y) / rho (examples/free_flame/
    reactors.py:115-123, time_integ.py:64-69), T from e0 with guess 1200 K.
    The stored rows sit 12 steps of 1e-8 s apart."""
    def rhs(yy):
        t = gas.get_temperature(e0, 1200.0, yy, do_energy=True)
        return gas.molecular_weights * gas.get_net_production_rates(
            rho, t, yy) / rho
    for _ in range(nsteps):
        k1 = rhs(y)
        k2 = rhs(y + 0.5 * h * k1)
        k3 = rhs(y + 0.5 * h * k2)
        k4 = rhs(y + h * k3)
        y = y + h * (k1 + 2 * (k2 + k3) + k4) / 6
    return y


ROWS = (5, 100, 200, 300, 340, 400, 500)


def _window_errors(gas, fixture):
    s = fixture["explicit_state"]
    rho = gas.get_density(ONE_ATM, 1200.0, s[0, :9])
    e0 = gas.get_mixture_internal_energy_mass(1200.0, s[0, :9])
    out = []
    for i in ROWS:
        y = _rk4_window(gas, s[i, :9].copy(), rho, e0)
        inc = np.max(np.abs(s[i + 1, :9] - s[i, :9]))
        out.append(np.max(np.abs(y - s[i + 1, :9])) / inc)
    return np.array(out)


def test_fixture_rk4_windows(fixture, sandiego):
    """Re-integrating stored windows reproduces the next stored row to
    ~1e-13 of the increment: rate coefficients (Arrhenius, Troe with
    T3 = 1e-30, three-body efficiencies), K_c and the production-rate
    stoichiometry all agree with the reference's generated code."""
    err = _window_errors(sandiego, fixture)
    assert np.all(err < 1e-11), err


def test_fixture_catches_planted_error(fixture):
    """Plant a 1e-4 relative error in one rate constant of the oracle's
    mechanism: the window check must fail."""
    m = bundled("sandiego")
    r = m.reactions[0]
    bad = dataclasses.replace(r, rate=dataclasses.replace(r.rate,
                                                          A=r.rate.A * 1.0001))
    m_bad = Mechanism(name=m.name, source=m.source, elements=m.elements,
                      species=m.species,
                      reactions=(bad,) + tuple(m.reactions[1:]))
    err = _window_errors(Thermochemistry(m_bad), fixture)
    assert np.max(err) > 1e-8


# ---------------------------------------------------------------------------
# (b) closed forms
# ---------------------------------------------------------------------------
def _cs(f, x, h=1e-30):
    return np.imag(f(x + 1j * h)) / h


@pytest.mark.parametrize("name", ["sandiego", "gri30", "synth53"])
def test_nasa_identities(name):
    """d(T h/RT)/dT = cp/R and d(s/R)/dT = cp/(R T): a dropped /2.../5
    factor or swapped coefficient fails.  The *_derivative methods equal the
    complex-step derivatives of the base methods."""
    g = Thermochemistry(bundled(name))
    T = np.linspace(310.0, 3400.0, 97)
    T = T[np.all([np.abs(T - s.t_mid) > 1 for s in g.mech.species], axis=0)]
    cp = g.get_species_specific_heats_r(T)
    d_th = _cs(lambda t: t * g.get_species_enthalpies_rt(t), T.astype(complex))
    d_s = _cs(g.get_species_entropies_r, T.astype(complex))
    assert np.max(np.abs(d_th - cp) / np.abs(cp)) < 1e-11
    assert np.max(np.abs(d_s * T - cp) / np.abs(cp)) < 1e-11
    for base, der in (("specific_heats_r", "specific_heats_r_derivative"),
                      ("enthalpies_rt", "enthalpies_rt_derivative"),
                      ("entropies_r", "entropies_rt_derivative")):
        ref = _cs(getattr(g, f"get_species_{base}"), T.astype(complex))
        got = getattr(g, f"get_species_{der}")(T)
        assert np.max(np.abs(got - ref) / (np.abs(ref) + 1e-3)) < 1e-12


def test_nasa_branch_is_strict(sandiego):
    """T == T_mid takes the LOW coefficients (chem_expr.py:51-57)."""
    s = sandiego.mech.species[0]
    t = s.t_mid
    lo = sum(s.low[j] * t ** j for j in range(5))
    hi = sum(s.high[j] * t ** j for j in range(5))
    assert lo != hi
    assert sandiego.get_species_specific_heats_r(np.array([t]))[0, 0] == \
        pytest.approx(lo, rel=1e-15)


@pytest.mark.parametrize("name", ["sandiego", "gri30", "uiuc"])
def test_equilibrium_constants_closed_form(name):
    """-ln K_c = sum nu'' g - sum nu' g - dnu ln(p0/(R T)), written as a
    stoichiometric-matrix product; irreversible rows hold the placeholder."""
    g = Thermochemistry(bundled(name))
    m = g.mech
    T = np.linspace(300.0, 3000.0, 11)
    G = g.get_species_gibbs_rt(T)
    nu_r = np.array([[m.reactant_stoich_coeff(k, i) for k in range(m.n_species)]
                     for i in range(m.n_reactions)])
    nu_p = np.array([[m.product_stoich_coeff(k, i) for k in range(m.n_species)]
                     for i in range(m.n_reactions)])
    dnu = (nu_p - nu_r).sum(axis=1)
    expect = (nu_p - nu_r) @ G - np.outer(dnu, np.log(ONE_ATM / (GAS_CONSTANT * T)))
    got = g.get_equilibrium_constants(T)
    scale = (np.abs(nu_p) + np.abs(nu_r)) @ np.abs(G) + 1.0
    for i, r in enumerate(m.reactions):
        if r.reversible:
            assert np.max(np.abs(got[i] - expect[i]) / scale[i]) < 1e-14
        else:
            assert np.array_equal(got[i], -0.17364695002734 * T)


def _arr(A_cgs, order, b, Ea_cal, T):
    """Arrhenius from raw cgs/cal YAML numbers (independent conversion)."""
    A = A_cgs * (1e-6 / 1e-3) ** (order - 1)
    return A * T ** b * np.exp(-Ea_cal * 4.184e3 / (GAS_CONSTANT * T))


def test_kf_closed_form_troe_lindemann_threebody():
    """k_f for one reaction of each kind, evaluated from the YAML numbers with
    the textbook formulas (doc/model.rst:25-109): Troe (4-parameter),
    Lindemann and a three-body reaction with efficiencies."""
    g = Thermochemistry(bundled("gri30"))
    m = g.mech
    n = 8
    T, p, Y = random_states(m, n, seed=7)
    rho = g.get_density(p, T, Y)
    C = g.get_concentrations(rho, Y)
    kf = g.get_fwd_rate_coefficients(T, C)
    idx = m.species_index
    Ctot = C.sum(axis=0)

    def eff_sum(effs):
        return Ctot + sum((e - 1.0) * C[idx(s)] for s, e in effs.items())

    # H + CH2 (+M) <=> CH3 (+M): low 6.0e35 b -4.76 Ea 2440; high 6.0e14 0 0;
    # Troe A .562, T3 91, T1 5836, T2 8552; efficiencies below
    i = next(j for j, r in enumerate(m.reactions)
             if r.equation == "H + CH2 (+ M) <=> CH3 (+ M)")
    k0 = _arr(1.04e26, 3, -2.76, 1600.0, T)
    kinf = _arr(6.0e14, 2, 0.0, 0.0, T)
    M = eff_sum({"AR": 0.7, "C2H6": 3, "CH4": 2, "CO": 1.5, "CO2": 2,
                 "H2": 2, "H2O": 6})
    pr = M * k0 / kinf
    fcent = np.log10((1 - 0.562) * np.exp(-T / 91) + 0.562 * np.exp(-T / 5836)
                     + np.exp(-8552 / T))
    c = -0.4 - 0.67 * fcent
    nn = 0.75 - 1.27 * fcent
    f1 = (np.log10(pr) + c) / (nn - 0.14 * (np.log10(pr) + c))
    F = 10 ** (fcent / (1 + f1 ** 2))
    assert np.max(np.abs(kf[i] / (kinf * pr / (1 + pr) * F) - 1)) < 1e-13

    # O + CO (+M) <=> CO2 (+M): Lindemann
    i = next(j for j, r in enumerate(m.reactions)
             if r.equation == "O + CO (+ M) <=> CO2 (+ M)")
    k0 = _arr(6.02e14, 3, 0.0, 3000.0, T)
    kinf = _arr(1.8e10, 2, 0.0, 2385.0, T)
    M = eff_sum({"AR": 0.5, "C2H6": 3, "CH4": 2, "CO": 1.5, "CO2": 3.5,
                 "H2": 2, "H2O": 6, "O2": 6})
    pr = M * k0 / kinf
    assert np.max(np.abs(kf[i] / (kinf * pr / (1 + pr)) - 1)) < 1e-13

    # 2 O + M <=> O2 + M: A 1.2e17 b -1 Ea 0, effs AR .83 C2H6 3 CH4 2 CO 1.75
    # CO2 3.6 H2 2.4 H2O 15.4
    i = next(j for j, r in enumerate(m.reactions)
             if r.equation == "2 O + M <=> O2 + M")
    M = eff_sum({"AR": 0.83, "C2H6": 3, "CH4": 2, "CO": 1.75, "CO2": 3.6,
                 "H2": 2.4, "H2O": 15.4})
    assert np.max(np.abs(kf[i] / (_arr(1.2e17, 3, -1.0, 0.0, T) * M) - 1)) < 1e-13


def test_explicit_collider_and_orders():
    """(+ Ar) falloff uses [Ar] alone (hong); explicit reaction orders use
    C^order (uiuc) and the irreversible rate has no reverse part."""
    h = Thermochemistry(bundled("hong"))
    m = h.mech
    T, p, Y = random_states(m, 5, seed=3)
    C = h.get_concentrations(h.get_density(p, T, Y), Y)
    kf = h.get_fwd_rate_coefficients(T, C)
    k0 = _arr(6.81e18, 3, -1.2, 0.0, T)
    kinf = _arr(5.59e13, 2, 0.2, 0.0, T)
    pr = C[m.species_index("Ar")] * k0 / kinf
    fcent = np.log10(0.3 * np.exp(-T / 1e-30) + 0.7 * np.exp(-T / 1e30)
                     + np.exp(-1e30 / T))
    c = -0.4 - 0.67 * fcent
    nn = 0.75 - 1.27 * fcent
    f1 = (np.log10(pr) + c) / (nn - 0.14 * (np.log10(pr) + c))
    F = 10 ** (fcent / (1 + f1 ** 2))
    assert np.max(np.abs(kf[2] / (kinf * pr / (1 + pr) * F) - 1)) < 1e-13

    u = Thermochemistry(bundled("uiuc"))
    mu = u.mech
    T, p, Y = random_states(mu, 5, seed=4)
    C = u.get_concentrations(u.get_density(p, T, Y), Y)
    rf = u.get_fwd_rates_of_progress(T, C)
    kf = u.get_fwd_rate_coefficients(T, C)
    expect = kf[0] * C[mu.species_index("C2H4")] ** 0.5 \
        * C[mu.species_index("O2")] ** 0.65
    assert np.max(np.abs(rf[0] / expect - 1)) < 1e-14
    assert np.all(u.get_rev_rates_of_progress(T, C)[0] == 0)


# ---------------------------------------------------------------------------
# (c) invariants
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["sandiego", "gri30", "synth53", "hong"])
def test_rate_invariants(name):
    g = Thermochemistry(bundled(name))
    m = g.mech
    inert = "N2" if "N2" in m.species_names else m.species_names[-1]
    T, p, Y = random_states(m, 64, seed=11, inert=inert)
    rho = g.get_density(p, T, Y)
    w = g.get_net_production_rates(rho, T, Y)
    scale = g.production_rate_scale(rho, T, Y)
    W = m.molecular_weights[:, None]
    # mass: sum_k W_k w_k = 0
    assert np.max(np.abs((W * w).sum(0)) / (W * scale).sum(0)) < 1e-14
    # elements: sum_k a_ek w_k = 0
    E = m.element_matrix()
    assert np.max(np.abs(E @ w) / (np.abs(E) @ scale + 1e-300)) < 1e-14
    # creation - destruction = net (test/test.py:276)
    cd, dd = g.get_creation_destruction_rates(rho, T, Y)
    assert scaled_err(cd - dd, w, scale) < 1e-14


def test_temperature_quirks(sandiego):
    T, p, Y = random_states(sandiego.mech, 16, seed=5)
    e = sandiego.get_mixture_internal_energy_mass(T, Y)
    h = sandiego.get_mixture_enthalpy_mass(T, Y)
    # do_energy=0 is not `False`: it selects the energy form
    assert np.allclose(sandiego.get_temperature(e, T + 50, Y, 0), T,
                       rtol=0, atol=1e-9)
    assert np.allclose(sandiego.get_temperature(h, T + 50, Y, False), T,
                       rtol=0, atol=1e-9)
    with pytest.raises(RuntimeError):
        sandiego.get_temperature(e, np.nan, Y, True)
