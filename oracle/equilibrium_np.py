"""Chemical equilibrium of an ideal-gas mixture at fixed (h, p) -- the
``equilibrate('HP')`` call the reference's flamelet notebook uses for its
initial state (``examples/flamelet/flamelet.ipynb`` cell 6, Cantera's
``Solution.equilibrate``).

TEST INFRASTRUCTURE ONLY (see :mod:`oracle`).  Cantera is absent here, so the
published method is restated: the element-potential formulation of Gibbs
minimisation (W. C. Reynolds, "The element potential method for chemical
equilibrium analysis", STANJAN, Stanford 1986; the method of Cantera's
``ChemEquil``).  At equilibrium the chemical potential of every species is the
sum of its atoms' element potentials,

    mu_k / RT = g0_k(T)/RT + ln x_k + ln(p/p0) = sum_e a_ek lambda_e,

so x_k = exp(sum_e a_ek lambda_e - g0_k/RT - ln(p/p0)).  The unknowns
(lambda_e for the elements present, ln N with N the kmol per kg, and T)
solve, by Newton's method,

    sum_k a_ek N x_k = b_e            (element content of the input Y)
    sum_k x_k        = 1
    sum_k N x_k H_k(T) = h            (H_k molar enthalpy; fixed h).

Species containing an element the mixture lacks are absent (x_k = 0).

Pinned in ``tests/test_oracle_pins.py`` by what the mathematics fixes:
detailed balance (every reversible reaction's net rate of progress vanishes
at the result, through the kinetics' own equilibrium constants -- an
independent code path), the closed form of H2 <=> 2 H at fixed T, and
conservation of h and of every element.
"""

from __future__ import annotations

import numpy as np

from pyrometheus_amd.mechanism import GAS_CONSTANT, ONE_ATM


def _nasa(coeffs_lo, coeffs_hi, t_mid, t):
    """(c_p/R, h/RT, s/R) of every species at scalar T, the strict
    ``T > T_mid`` branch rule of the reference (chem_expr.py:51-57)."""
    c = np.where((t > t_mid)[:, None], coeffs_hi, coeffs_lo)
    cp = c[:, 0] + t * (c[:, 1] + t * (c[:, 2] + t * (c[:, 3] + t * c[:, 4])))
    h = (c[:, 0] + t * (c[:, 1] / 2 + t * (c[:, 2] / 3 + t * (c[:, 3] / 4
         + t * c[:, 4] / 5))) + c[:, 5] / t)
    s = (c[:, 0] * np.log(t) + t * (c[:, 1] + t * (c[:, 2] / 2 + t * (
        c[:, 3] / 3 + t * c[:, 4] / 4))) + c[:, 6])
    return cp, h, s


class Equilibrium:
    """HP (or TP) equilibrium solver for one mechanism."""

    def __init__(self, mech):
        self.mech = mech
        self.ns = mech.n_species
        self.lo = np.array([s.low for s in mech.species])
        self.hi = np.array([s.high for s in mech.species])
        self.t_mid = np.array([s.t_mid for s in mech.species])
        self.W = mech.molecular_weights
        self.A = mech.element_matrix()  # (ne, ns) atoms per molecule

    def _thermo(self, t):
        return _nasa(self.lo, self.hi, self.t_mid, np.full(self.ns, t))

    def _start(self, b, present, sp, t, lnp):
        """Element potentials fitted (least squares, weighted towards the
        major species of an idealised complete-combustion mixture) and N."""
        A = self.A[present][:, sp]
        # a rough mixture: every allowed species at 1e-6, the element
        # content spread over the species by a least-squares fit
        n0, *_ = np.linalg.lstsq(A, b[present], rcond=None)
        n0 = np.maximum(n0, 1e-6 * b[present].sum())
        x0 = n0 / n0.sum()
        _, h, s = self._thermo(t)
        g = (h - s)[sp]
        lam, *_ = np.linalg.lstsq(A.T, np.log(x0) + g + lnp, rcond=None)
        return lam, np.log(n0.sum())

    def equilibrate(self, Y, p, h=None, T=None, t_guess=2000.0, tol=1e-14,
                    maxit=200):
        """Equilibrium mass fractions and temperature of the element content
        of ``Y`` at pressure ``p`` and mass enthalpy ``h`` (HP), or at fixed
        ``T`` (TP) when ``h`` is None."""
        Y = np.asarray(Y, dtype=float)
        b = self.A @ (Y / self.W)  # kmol of each element per kg
        present = b > 1e-300
        sp = np.all((self.A[~present] == 0), axis=0)  # allowed species
        Ap = self.A[present][:, sp]
        bp = b[present]
        ne = int(present.sum())
        lnp = np.log(p / ONE_ATM)
        fixed_t = h is None
        t = float(T if fixed_t else t_guess)
        lam, lnN = self._start(b, present, sp, t, lnp)
        R = GAS_CONSTANT
        converged = False
        for it in range(maxit):
            cp, hrt, s = self._thermo(t)
            cp, hrt, s = cp[sp], hrt[sp], s[sp]
            g = hrt - s
            x = np.exp(np.clip(Ap.T @ lam - g - lnp, -700.0, 700.0))
            N = np.exp(lnN)
            n = N * x
            Hm = R * t * hrt  # molar enthalpy
            F = np.empty(ne + 2)
            F[:ne] = (Ap @ n - bp) / bp
            F[ne] = x.sum() - 1.0
            F[ne + 1] = 0.0 if fixed_t else (n @ Hm - h) / (R * t * bp.sum())
            J = np.zeros((ne + 2, ne + 2))
            dx_dT = x * hrt / t  # d x_k / dT = x_k h_k/(R T^2)
            J[:ne, :ne] = (Ap * n) @ Ap.T / bp[:, None]
            J[:ne, ne] = (Ap @ n) / bp
            J[:ne, ne + 1] = N * (Ap @ dx_dT) / bp
            J[ne, :ne] = Ap @ x
            J[ne, ne + 1] = dx_dT.sum()
            if fixed_t:
                J[ne + 1, ne + 1] = 1.0
            else:
                sc = R * t * bp.sum()
                J[ne + 1, :ne] = (Ap * (n * Hm)) .sum(1) / sc
                J[ne + 1, ne] = (n @ Hm) / sc
                J[ne + 1, ne + 1] = (N * (dx_dT @ Hm) + n @ (R * cp)) / sc
            step = np.linalg.solve(J, -F)
            # keep the step in a trust region: element potentials and ln N
            # by at most 2, T by at most 250 K
            lim = max(np.max(np.abs(step[:ne + 1])) / 2.0,
                      abs(step[ne + 1]) / 250.0, 1.0)
            step /= lim
            lam = lam + step[:ne]
            lnN = lnN + step[ne]
            if not fixed_t:
                t = t + step[ne + 1]
            # done once the residual is at rounding level, or Newton no
            # longer moves the iterate (quadratic convergence has reached
            # the rounding floor of the residual)
            tiny = (np.max(np.abs(step[:ne + 1])) < 1e-13
                    and abs(step[ne + 1]) < 1e-12 * t)
            if lim == 1.0 and (np.max(np.abs(F)) < tol or tiny):
                converged = True
                break
        if not converged:
            raise RuntimeError("equilibrium iteration did not converge")
        _, hrt, s = self._thermo(t)
        x = np.zeros(self.ns)
        x[sp] = np.exp(Ap.T @ lam - (hrt - s)[sp] - lnp)
        y = x * self.W
        return y / y.sum(), t


def equilibrate_hp(mech, Y, p, h, t_guess=2000.0):
    """Equilibrium (Y, T) at fixed mass enthalpy and pressure."""
    return Equilibrium(mech).equilibrate(Y, p, h=h, t_guess=t_guess)
