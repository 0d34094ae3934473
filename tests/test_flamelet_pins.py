"""Pins for the flamelet oracle (oracle/flamelet_np.py, oracle/equilibrium_np.py).

What fixes the flamelet path without importing the reference:

* the reference's own saved run: ``examples/flamelet/flamelet.ipynb`` cell 7
  prints, for sandiego with N = 101 uniform points, T_ox = 500 K air, fuel
  H2:N2 = 0.5:0.5 at 300 K, 1 atm, chi_st = 100 (boundary values 2 dx^2)
  and the HP-equilibrium start of cell 6, the maximum temperature after each
  of 10 Crank-Nicolson steps of 1e-5 s and the update norm |v| of each
  Newton iteration that follows; then the same for a restart at
  chi_st = 1000.  The oracle reproduces them (the Newton iterates only with
  the reference's temperature-loop semantics, whose Jacobian omits dT/dphi
  where the guess is already converged);
* what the mathematics fixes: the 3-point Laplacian is exact on quadratics
  on any grid; at convergence the element mass fractions are linear in Z
  (unity Lewis number, element conservation of the source); the HP
  equilibrium satisfies detailed balance through the kinetics' own
  equilibrium constants and reduces to a closed form for H2 <=> 2 H.

Each check fails a plausible mistake: a factor on chi changes every T_max
(planted below), wrong non-uniform weights break the quadratic, a dropped
source term or mask breaks element linearity, a sign or a missing species
in the equilibrium breaks detailed balance.
"""
import numpy as np
import pytest

from oracle import flamelet_np as F
from oracle.equilibrium_np import Equilibrium
from oracle.thermochem_np import Thermochemistry
from pyrometheus_amd.mechanism import ATOMIC_WEIGHTS, ONE_ATM, bundled

# examples/flamelet/flamelet.ipynb, cell 7 output (printed digits)
NB_TMAX_100 = ["2024.005", "1995.429", "1962.204", "1942.950", "1923.554",
               "1912.424", "1901.378", "1894.576", "1887.830", "1883.539"]
NB_V_100 = [1.0434e-02, 4.5838e-03, 1.6754e-03, 4.9009e-04, 8.9558e-05,
            6.6336e-06, 5.6013e-08, 6.8937e-11]
NB_TMAX_1000 = ["1733.993", "1678.976", "1648.160", "1634.348", "1626.186",
                "1622.389", "1619.933", "1618.749", "1617.898", "1617.466"]
NB_V_1000 = [7.7460e-04, 1.8332e-04, 1.5053e-06, 5.2903e-10]
#: below this an update norm is rounding (the enthalpy row's floor): the
#: printed last iterates (6.9e-11, 5.3e-10) are compared within 10 %
ROUNDING_FLOOR = 1e-9


def notebook_case():
    """Cells 2-6: streams, Z_st, uniform grid, linear mixing and the HP
    equilibrium start (Cantera's equilibrate('HP') restated)."""
    m = bundled("sandiego")
    y_ox, y_fu, h_ox, h_fu = F.stream_compositions(m, t_ox=500.0, t_fu=300.0)
    z_st = F.stoichiometric_mixture_fraction(m, y_ox, y_fu)
    x = F.uniform_grid(101)
    eq = Equilibrium(m)
    h = h_ox + (h_fu - h_ox) * x
    Y = np.empty((m.n_species, len(x)))
    T = np.empty(len(x))
    for i, z in enumerate(x):
        Y[:, i], T[i] = eq.equilibrate(y_ox + (y_fu - y_ox) * z, ONE_ATM,
                                       h=h[i])
    phi = np.vstack([h[None], Y])
    return m, x, y_ox, y_fu, h_ox, h_fu, z_st, phi, T


@pytest.fixture(scope="module")
def case():
    return notebook_case()


def _check_log(log, tmax, vnorms):
    cn = [e for e in log if e[0] == "cn"]
    nw = [e for e in log if e[0] == "newton"]
    assert [f"{e[3]:.3f}" for e in cn] == tmax
    assert len(nw) == len(vnorms)
    for (_, it, v), ref in zip(nw, vnorms):
        if ref > ROUNDING_FLOOR:
            assert f"{v:.4e}" == f"{ref:.4e}", (it, v, ref)
        else:
            assert abs(v - ref) <= 0.1 * ref, (it, v, ref)


def test_notebook_case(case):
    """The oracle with the reference's semantics reproduces cell 7: every
    T_max of both time marches to the printed 3 decimals, every Newton |v|
    above the rounding floor to the printed 5 digits, the same number of
    iterations and one attempt each."""
    m, x, y_ox, y_fu, h_ox, h_fu, z_st, phi, T = case
    assert f"{z_st:.3f}" == "0.304"  # cell 2
    s = F.FlameletSolver(m, x, y_ox, y_fu, semantics="reference")
    vd = np.zeros(len(x))
    for chi_st, tmax, vn in ((100.0, NB_TMAX_100, NB_V_100),
                             (1000.0, NB_TMAX_1000, NB_V_1000)):
        chi = F.dissipation_profile(x, z_st, chi_st)
        assert np.allclose(chi[[0, -1]], 2 * 0.01 ** 2, rtol=1e-12)
        s.log = []
        phi, T, ok = s.solve(10, 1e-8, 20, 1e-9, 1e-5, 10, False, 10, chi, vd,
                             T, ONE_ATM, h_ox, h_fu, phi)
        assert ok
        _check_log(s.log, tmax, vn)


def test_notebook_case_catches_planted_error(case):
    """A factor on the dissipation rate (the 1/2 of 1/2 chi Lap) moves the
    first time step's T_max off the printed value."""
    m, x, y_ox, y_fu, h_ox, h_fu, z_st, phi, T = case
    s = F.FlameletSolver(m, x, y_ox, y_fu, semantics="reference")
    chi = F.dissipation_profile(x, z_st, 100.0)
    chi[1:-1] *= 2.0
    s.flamelet_time_march(20, 1e-9, 1, 1e-5, chi, np.zeros(len(x)), T,
                          ONE_ATM, h_ox, h_fu, phi)
    assert f"{s.log[0][3]:.3f}" != NB_TMAX_100[0]


@pytest.mark.parametrize("beta", [0.0, 3.0])
def test_laplacian_exact_on_quadratics(beta):
    """Interior rows of the 3-point Laplacian give f'' exactly for every
    quadratic on uniform and clustered grids; boundary rows are b f."""
    x = (F.uniform_grid(37) if beta == 0.0
         else F.clustered_grid(37, 0.3, beta))
    abc = F.laplacian_coefficients(x)
    for c2, c1, c0 in ((1.0, 0.0, 0.0), (-3.5, 2.0, 7.0), (0.0, 4.0, -1.0)):
        f = c2 * x ** 2 + c1 * x + c0
        lap = F.apply_laplacian(abc, f)
        assert np.max(np.abs(lap[1:-1] - 2 * c2)) <= 1e-9 * max(1.0, abs(c2))
        assert lap[0] == abc[1][0] * f[0] and lap[-1] == abc[1][-1] * f[-1]


def test_element_mass_fractions_linear(case):
    """At convergence the element mass fractions Z_e = sum_k a_ek W_e/W_k Y_k
    are linear in Z: the source conserves elements, so every interior row of
    1/2 chi Lap(Z_e) vanishes, and the boundary rows pin the end values."""
    m, x, y_ox, y_fu, h_ox, h_fu, z_st, phi, T = case
    xs = F.clustered_grid(41, z_st)
    eqm = Equilibrium(m)
    Y = np.empty((m.n_species, len(xs)))
    hs = h_ox + (h_fu - h_ox) * xs
    for i, z in enumerate(xs):
        Y[:, i], _ = eqm.equilibrate(y_ox + (y_fu - y_ox) * z, ONE_ATM,
                                     h=hs[i])
    gas = Thermochemistry(m)
    T0 = F.temperature_from_enthalpy(gas, hs, Y, 1500.0)
    s = F.FlameletSolver(m, xs, y_ox, y_fu)
    chi = F.dissipation_profile(xs, z_st, 300.0)
    phi_c, _, ok = s.solve(10, 1e-8, 20, 1e-9, 1e-5, 10, False, 10, chi,
                           np.zeros(len(xs)), T0, ONE_ATM, h_ox, h_fu,
                           np.vstack([hs[None], Y]))
    assert ok
    A = m.element_matrix()
    W = m.molecular_weights
    we = np.array([ATOMIC_WEIGHTS[e] for e in m.elements])
    Ze = (A * we[:, None] / W[None, :]) @ phi_c[1:]
    line = Ze[:, :1] + (Ze[:, -1:] - Ze[:, :1]) * xs[None, :]
    assert np.max(np.abs(Ze - line)) <= 1e-9, np.max(np.abs(Ze - line))
    # the species themselves are not linear (a flame burns)
    assert np.max(np.abs(phi_c[1:] - (phi_c[1:, :1] + (
        phi_c[1:, -1:] - phi_c[1:, :1]) * xs))) > 1e-2


def test_equilibrium_detailed_balance():
    """HP equilibrium of linear-mixing states: every reversible reaction's
    forward and reverse rates of progress agree (the kinetics' own
    equilibrium constants, an independent code path), h and every element
    are conserved."""
    m = bundled("sandiego")
    gas = Thermochemistry(m)
    y_ox, y_fu, h_ox, h_fu = F.stream_compositions(m, t_ox=500.0, t_fu=300.0)
    eq = Equilibrium(m)
    A, W = m.element_matrix(), m.molecular_weights
    for z in (0.05, 0.2, 0.304, 0.5, 0.9):
        y0 = y_ox + (y_fu - y_ox) * z
        h = h_ox + (h_fu - h_ox) * z
        y, T = eq.equilibrate(y0, ONE_ATM, h=h)
        # burning raises T above the unburnt mixture's
        assert T > float(gas.get_temperature(h, 400.0, y0)) + 150.0
        assert abs(gas.get_mixture_enthalpy_mass(T, y) - h) <= 1e-12 * abs(
            h_ox - h_fu)
        assert np.allclose(A @ (y / W), A @ (y0 / W), rtol=1e-12, atol=1e-18)
        rho = gas.get_density(ONE_ATM, T, y)
        C = gas.get_concentrations(rho, y)
        rf = np.array(gas.get_fwd_rates_of_progress(T, C), dtype=float)
        rr = np.array(gas.get_rev_rates_of_progress(T, C), dtype=float)
        rel = np.abs(rf - rr) / (np.abs(rf) + np.abs(rr) + 1e-300)
        assert rel.max() <= 1e-10, (z, rel.max())


def test_equilibrium_h2_dissociation_closed_form():
    """Pure hydrogen at fixed T and p: x_H^2 / x_H2 = K_p p0/p with
    K_p = exp(-(2 g_H - g_H2)/RT) from the NASA-7 Gibbs functions, and
    x_H + x_H2 = 1, solved as a quadratic."""
    m = bundled("sandiego")
    gas = Thermochemistry(m)
    eq = Equilibrium(m)
    iH2, iH = m.species_index("H2"), m.species_index("H")
    W = m.molecular_weights
    for T, p in ((3000.0, ONE_ATM), (3500.0, 0.1 * ONE_ATM)):
        y0 = np.zeros(m.n_species)
        y0[iH2] = 1.0
        y, _ = eq.equilibrate(y0, p, T=T)
        g = np.array(gas.get_species_gibbs_rt(T), dtype=float)
        K = np.exp(-(2 * g[iH] - g[iH2])) * ONE_ATM / p
        xh = (-K + np.sqrt(K * K + 4 * K)) / 2
        n = y / W
        x = n / n.sum()
        assert abs(x[iH] - xh) <= 1e-13 and abs(x[iH2] - (1 - xh)) <= 1e-13
        assert np.all(np.delete(x, [iH, iH2]) == 0.0)
