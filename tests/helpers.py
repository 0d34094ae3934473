"""Shared input generators and error measures for the parity tests."""
import numpy as np

from pyrometheus_amd.mechanism import ONE_ATM


def random_states(mech, n, seed=1, t_lo=300.0, t_hi=2500.0, inert="N2"):
    """T ~ U[t_lo, t_hi], p log-uniform in [0.5, 20] atm,
    Y = 0.5*Dirichlet(0.5) + 0.5*e_inert (the workload's distributions)."""
    rng = np.random.default_rng(seed)
    T = rng.uniform(t_lo, t_hi, n)
    p = ONE_ATM * np.exp(rng.uniform(np.log(0.5), np.log(20.0), n))
    Y = 0.5 * rng.dirichlet(0.5 * np.ones(mech.n_species), n).T
    Y[mech.species_index(inert)] += 0.5
    return T, p, Y


def scaled_err(got, ref, scale):
    return float(np.max(np.abs(got - ref) / (np.abs(scale) + 1e-300)))


def rel_err(got, ref, floor=0.0):
    return float(np.max(np.abs(got - ref) / (np.abs(ref) + floor + 1e-300)))
