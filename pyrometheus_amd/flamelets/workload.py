"""Synthetic batches of steady flamelets for measurement and serving.

The set-up follows the reference notebook (``examples/flamelet``: fuel
H2:N2 = 0.5:0.5 by mole, oxidiser air, Burke-Schumann start, the bell-shaped
dissipation profile of ``flamelets/utils.py:9-26``) with the spread a batch
needs: grid sizes drawn uniformly in ``[n_min, n_max]``, the grid clustered
about the stoichiometric mixture fraction, and the stoichiometric dissipation
rate drawn log-uniformly.  Stream enthalpies and the initial temperature
guess come from the device ``Thermochemistry``.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import List, Sequence

import numpy as np
import torch

from ..mechanism import ONE_ATM
from .domain import Domain, DomainConfig
from .solver import FlameletProblem
from .utils import (burke_schumann, dissipation_profile, mole_to_mass,
                    stoichiometric_mixture_fraction)


@dataclass
class Streams:
    y_ox: np.ndarray
    y_fu: np.ndarray
    h_ox: float
    h_fu: float
    z_st: float


def streams(gas, t_ox=800.0, t_fu=300.0, fuel=None, oxidizer=None) -> Streams:
    """Oxidiser (default air O2:N2 = 0.21:0.79) at ``t_ox`` and fuel (default
    H2:N2 = 0.5:0.5) at ``t_fu``, both by mole."""
    mech = gas._mech
    xo = np.zeros(mech.n_species)
    xf = np.zeros(mech.n_species)
    for k, v in (oxidizer or {"O2": 0.21, "N2": 0.79}).items():
        xo[mech.species_index(k)] = v
    for k, v in (fuel or {"H2": 0.5, "N2": 0.5}).items():
        xf[mech.species_index(k)] = v
    y_ox, y_fu = mole_to_mass(mech, xo), mole_to_mass(mech, xf)
    dev = gas.device
    T = torch.tensor([t_ox, t_fu], dtype=torch.float64, device=dev)
    Y = torch.tensor(np.stack([y_ox, y_fu], axis=1), dtype=torch.float64,
                     device=dev)
    h = gas.get_mixture_enthalpy_mass(T, Y).cpu().numpy()
    return Streams(y_ox, y_fu, float(h[0]), float(h[1]),
                   stoichiometric_mixture_fraction(mech, y_ox, y_fu))


def make_problem(gas, st: Streams, n: int, chi_st: float, clustered=True,
                 pressure=ONE_ATM, t_guess=1500.0):
    """One flamelet of ``n`` points; returns ``(Domain, FlameletProblem)``."""
    dom = Domain(DomainConfig(num_x=n, x_l=0.0, x_r=1.0,
                              cluster_center=st.z_st if clustered else None))
    phi, _ = burke_schumann(gas._mech, dom.x, st.y_ox, st.y_fu, st.h_ox,
                            st.h_fu)
    chi = dissipation_profile(dom, st.z_st, chi_st)
    dev = gas.device
    ph = torch.as_tensor(phi, dtype=torch.float64, device=dev)
    T0 = gas.get_temperature_from_enthalpy(
        ph[0], ph[1:], torch.full((n,), t_guess, dtype=torch.float64,
                                  device=dev))
    prob = FlameletProblem(x=dom.x, chi=chi, phi=phi, t_guess=T0.cpu().numpy(),
                           y_ox=st.y_ox, y_fu=st.y_fu, h_ox=st.h_ox,
                           h_fu=st.h_fu, pressure=pressure)
    return dom, prob


def flamelet_batch(gas, n_flamelets=256, seed=4, n_min=64, n_max=1024,
                   chi_min=1.0, chi_max=1000.0, t_ox=800.0, t_fu=300.0,
                   clustered=True, pressure=ONE_ATM):
    """``n_flamelets`` problems with N ~ U{n_min..n_max} and
    chi_st = exp(U[ln chi_min, ln chi_max]) from ``seed``; returns
    ``(domains, problems, chi_st)``."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(n_min, n_max + 1, size=n_flamelets)
    chis = np.exp(rng.uniform(np.log(chi_min), np.log(chi_max),
                              size=n_flamelets))
    st = streams(gas, t_ox=t_ox, t_fu=t_fu)
    doms: List[Domain] = []
    probs: List[FlameletProblem] = []
    for n, c in zip(sizes, chis):
        d, p = make_problem(gas, st, int(n), float(c), clustered, pressure)
        doms.append(d)
        probs.append(p)
    return doms, probs, chis


def solve_plain(gas, domain: Domain, prob: FlameletProblem,
                equilibrate_steps=24, newton_norm="reference",
                semantics="exact", **solve_kw):
    """The plain drop-in: one ``FlameletSolver`` per flamelet, equilibrium
    start, then the reference's hybrid ``solve``; returns
    ``(state, temperature, converged)``.  ``semantics`` defaults to the
    batched solver's (``"exact"``), so the two run the same iteration."""
    from .solver import FlameletSolver
    from .state import FlameletState
    kw = dict(verbosity=False, newton_maxiter=10, newton_tol=1e-8,
              bdf_newton_maxiter=20, bdf_newton_tol=1e-9, bdf_time_step=1e-5,
              bdf_maxsteps=10, try_newton=False, max_attempts=10)
    kw.update(solve_kw)
    dev = gas.device
    solver = FlameletSolver(domain, gas, (prob.y_ox, prob.y_fu),
                            newton_norm=newton_norm, semantics=semantics)
    phi = torch.as_tensor(prob.phi, dtype=torch.float64, device=dev)
    state = FlameletState(phi[0].clone(), phi[1:].clone())
    tg = torch.as_tensor(prob.t_guess, dtype=torch.float64, device=dev)
    if equilibrate_steps:
        state, tg = solver.equilibrate(state, tg, prob.pressure, prob.h_ox,
                                       prob.h_fu, steps=equilibrate_steps)
    vd = torch.zeros(len(prob.x), dtype=torch.float64, device=dev)
    chi = torch.as_tensor(prob.chi, dtype=torch.float64, device=dev)
    state, temp = solver.solve(diss_rate=chi, viscous_diss=vd, temp_guess=tg,
                               pressure=prob.pressure, h_ox=prob.h_ox,
                               h_fu=prob.h_fu, state_guess=state, **kw)
    return state, temp, solver.converged


def latency_percentiles(latencies: Sequence[float]):
    a = np.asarray(latencies, dtype=float)
    return float(np.percentile(a, 50)), float(np.percentile(a, 99))
