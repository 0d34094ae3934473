"""Numpy restatement of ``pyrometheus/flamelets/thermodynamic_consistency.py``
(``CompressibleEOS``) and of the adjoint pieces of ``equations.py`` it uses.

TEST INFRASTRUCTURE ONLY (see :mod:`oracle`).

* :func:`adjoint_operator` -- ``FlameletEquations.adjoint_operator``
  (``equations.py:451-516``): central blocks ``Lap_C + 2 J_S^T / chi`` on
  interior points and identity rows at the two ends (the Laplacian's
  ``1/dx^2`` boundary blocks times ``chi/2 = dx^2``), the Laplacian's own
  lower/upper couplings.
* :func:`compressible_eos_rt` (``:518-535``) -- R_u T / W per point, and its
  per-point gradient with respect to phi (complex step through the same
  temperature recovery the flamelet equations use).
* :func:`source_gradient_wrt_pressure` (``:130-135``) -- dS/dp per point at
  fixed (h, Y) (complex step in p).
* :class:`CompressibleEOS` -- ``enthalpy_gradient`` (``:106-166``),
  ``eos_gradient`` (``:168-308``), ``evaluate_flamelet`` (``:310-398``),
  the Gauss-Newton and Picard updates (``:400-540``) and
  ``ensure_consistency`` (``:602-706``).

The boundary traces of the adjoint states enter through the reference's
stencil ``(1.5 l_0 + 2 l_1 - 0.5 l_2) / dx^2`` (and its mirror at the fuel
end); the reference notes that the adjoint states carry an implicit dZ.
"""

from __future__ import annotations

import numpy as np

from oracle import flamelet_np as F


def trapezoidal_rule(integrand):
    """``thermodynamic_consistency.trapezoidal_rule`` (unit spacing)."""
    return 0.5 * (integrand[0] + integrand[-1]) + np.sum(integrand[1:-1])


def adjoint_operator(eqns: F.FlameletEquations, phi, chi, vd, t_guess, p):
    """(lower (N-1,nv,nv), central (N,nv,nv), upper (N-1,nv,nv))."""
    a, b, c = eqns.abc
    nv, n = eqns.nv, phi.shape[1]
    eye = np.eye(nv)
    J = eqns.source_jacobian(phi, vd, t_guess, p)
    src = np.zeros_like(J)
    src[1:-1] = np.transpose(2.0 * J[1:-1] / chi[1:-1, None, None], (0, 2, 1))
    cen = b[:, None, None] * eye + src
    cen[0] = 0.5 * chi[0] * b[0] * eye
    cen[-1] = 0.5 * chi[-1] * b[-1] * eye
    low = a[1:, None, None] * eye
    low[-1] = 0.0
    up = c[:-1, None, None] * eye
    up[0] = 0.0
    return low, cen, up


def _rt(eqns, phi, t_guess, pointwise=True):
    gas = eqns.gas
    h, Y = phi[0], phi[1:]
    if eqns.semantics == "reference":
        T = F.temperature_from_enthalpy_jax(gas, h, Y, t_guess,
                                            pointwise=pointwise)
    else:
        T = F.temperature_from_enthalpy(gas, h, Y, t_guess, extra=3)
    return gas.gas_constant * T * gas._inv_w_sum(Y)


def compressible_eos_rt(eqns, phi, t_guess, p):
    """R T with R = R_u / W per point (``local_compressible_eos_rt``)."""
    return np.real(_rt(eqns, phi, t_guess))


def compressible_eos_rt_jacobian(eqns, phi, t_guess, p, h=1e-30):
    """d(R T)/dphi per point, (N, nv), by complex step."""
    out = np.empty((phi.shape[1], eqns.nv))
    for d in range(eqns.nv):
        pert = phi.astype(complex)
        pert[d] += 1j * h
        out[:, d] = np.imag(_rt(eqns, pert, t_guess)) / h
    return out


def source_gradient_wrt_pressure(eqns, phi, vd, t_guess, p, h=1e-30):
    """dS/dp per point at fixed (h, Y), (N, nv)."""
    S, _ = eqns.source_terms(phi, vd, t_guess, p + 1j * h, extra=3)
    return (np.imag(S) / h).T


class CompressibleEOS:
    """``CompressibleEOS`` over the oracle's flamelet solver."""

    def __init__(self, config, forward_solver: F.FlameletSolver, dx):
        self.config = config
        self.fwd = forward_solver
        self.eqns = forward_solver.eqns
        self.dx = dx
        n = len(self.eqns.x)
        self.unit_h = np.zeros((self.eqns.nv, n))
        self.unit_h[0] = 1.0

    def _trace(self, lam_h):
        d2 = self.dx ** 2
        ox = (1.5 * lam_h[0] + 2 * lam_h[1] - 0.5 * lam_h[2]) / d2
        fu = (1.5 * lam_h[-1] + 2 * lam_h[-2] - 0.5 * lam_h[-3]) / d2
        return ox + fu

    def enthalpy_gradient(self, phi, pdf, chi, vd, t_guess, p):
        rt = compressible_eos_rt(self.eqns, phi, t_guess, p)
        L, C, U = adjoint_operator(self.eqns, phi, chi, vd, t_guess, p)
        rhs = -self.unit_h * pdf
        lam = F.block_thomas(L, C, U, rhs.T).T
        return rt, lam, self._trace(lam[0])

    def eos_gradient(self, phi, pdf, chi, vd, t_guess, p):
        eq = self.eqns
        rt = compressible_eos_rt(eq, phi, t_guess, p)
        rtj = compressible_eos_rt_jacobian(eq, phi, t_guess, p)
        L, C, U = adjoint_operator(eq, phi, chi, vd, t_guess, p)
        ds_dp = source_gradient_wrt_pressure(eq, phi, vd, t_guess, p) / (
            0.5 * chi[:, None])
        rhs_d = (p / rt ** 2) * rtj.T * pdf
        lam_d = F.block_thomas(L, C, U, rhs_d.T).T
        integrand = (1.0 / rt) * pdf + np.einsum("ij,ji->i", ds_dp, lam_d)
        dens = np.array([trapezoidal_rule(integrand), self._trace(lam_d[0])])
        rhs_e = -(self.unit_h - rtj.T) * pdf
        lam_e = F.block_thomas(L, C, U, rhs_e.T).T
        integrand = np.einsum("ij,ji->i", ds_dp, lam_e)
        ener = np.array([trapezoidal_rule(integrand), self._trace(lam_e[0])])
        return (dens, ener), (lam_d, lam_e)

    def _solve(self, params, chi, vd, t_guess, phi_guess):
        c = self.config
        p, h_ox, h_fu = params
        phi, T, ok = self.fwd.solve(
            c["newton"]["maxiter"], c["newton"]["tol"],
            c["bdf"]["newton"]["maxiter"], c["bdf"]["newton"]["tol"],
            c["bdf"]["time_step"], c["bdf"]["maxsteps"], True,
            c["max_attempts"], chi, vd, t_guess, p, h_ox, h_fu, phi_guess)
        return phi, T

    def evaluate_flamelet(self, params, pdf, chi, vd, t_guess, phi_guess):
        p = params[0]
        phi, T = self._solve(params, chi, vd, t_guess, phi_guess)
        (dg, eg), _ = self.eos_gradient(phi, pdf, chi, vd, t_guess, p)
        rt = compressible_eos_rt(self.eqns, phi, T, p)
        density = np.sum((p / rt) * pdf)
        energy = np.sum((phi[0] - rt) * pdf)
        return phi, T, density, energy, dg, eg

    def gauss_newton_update(self, density_sim, energy_sim, params, pdf, chi,
                            vd, t_guess, phi_guess):
        phi, T, dens, ener, dg, eg = self.evaluate_flamelet(
            params, pdf, chi, vd, t_guess, phi_guess)
        G = np.stack((dg, eg))
        res = np.array([dens - density_sim, ener - energy_sim])
        v = np.linalg.solve(G.T @ G, -G.T @ res)
        return phi, T, np.array([v[0], v[1], v[1]]), res

    def picard_update(self, density_sim, energy_sim, params, pdf, chi, vd,
                      t_guess, phi_guess):
        p = params[0]
        phi, T = self._solve(params, chi, vd, t_guess, phi_guess)
        rt, _, dh = self.enthalpy_gradient(phi, pdf, chi, vd, t_guess, p)
        dens = np.sum((p / rt) * pdf)
        ener = np.sum((phi[0] - rt) * pdf)
        dp = density_sim * np.sum(rt * pdf) - p
        dhb = (energy_sim - ener) / dh
        return (phi, T, np.array([dp, dhb, dhb]),
                np.array([density_sim - dens, energy_sim - ener]))

    def ensure_consistency(self, density_sim, energy_sim, params, pdf, chi,
                           vd, t_guess, phi_guess):
        c = self.config["eos"]
        fn = {"gauss_newton": self.gauss_newton_update,
              "picard": self.picard_update}[c["update_method"]]
        phi, T = phi_guess, t_guess
        params = np.asarray(params, dtype=float)
        hist = []
        delta, it = np.inf, 0
        for it in range(c["maxiter"]):
            phi, T, v, res = fn(density_sim, energy_sim, params, pdf, chi,
                                vd, T, phi)
            delta = np.linalg.norm(v)
            params = params + c["update_size"] * v
            hist.append((np.linalg.norm(res) ** 2, delta))
            if delta < c["tol"]:
                break
        return phi, T, params, it, delta, hist
