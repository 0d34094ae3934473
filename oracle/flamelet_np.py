"""Numpy restatement of ``pyrometheus/flamelets`` (steady flamelet solver).

TEST INFRASTRUCTURE ONLY (see :mod:`oracle`).

State layout: ``phi`` is ``(nv, N)`` with row 0 the enthalpy and rows 1..ns the
mass fractions (``flamelets/state.py:297-316``).

* :func:`laplacian_coefficients` -- the 3-point Laplacian with identity
  boundary rows of ``flamelets/onedim_laplacian.py:115-212`` written as
  per-row coefficients ``(a_i, b_i, c_i)``; on a uniform grid
  ``a = c = 1/dx^2``, ``b = -2/dx^2`` and ``b = 1/dx^2`` on the two Dirichlet
  rows.  A non-uniform grid uses the standard 3-point second derivative
  (reduces exactly to the uniform stencil) with boundary rows ``1/h_end^2``.
* :class:`FlameletEquations` -- ``rhs`` (``equations.py:167-226``),
  ``source_terms`` (``:313-359``, T from h by Newton, rho from the EOS),
  ``jac`` (``:228-282``) with the per-point source Jacobian obtained by
  complex-step differentiation of ``local_source_terms`` (``:361-389``; the
  reference uses ``jax.jacfwd``).
* :func:`block_thomas` -- ``linear_solver.py:6-103``: forward elimination
  with two ``solve`` calls per row, back substitution with one, each an LU
  with partial pivoting (``numpy.linalg.solve``, as ``jnp.linalg.solve``).
* :class:`FlameletSolver` -- ``solver.py:63-600``: ``_newton_update``,
  ``_newton_loop_py`` (stop on ``|v| < tol``, fail when ``|v|`` grows),
  ``_newton_loop_lax``, the Crank-Nicolson ``flamelet_time_step``,
  ``flamelet_time_march`` and the hybrid ``solve``.

The temperature recovery here is per point (each point iterates until its own
step is below ``tol``), which is what makes a point's result independent of
the rest of the batch.
"""

from __future__ import annotations

import numpy as np
from scipy.special import erfcinv

from oracle.thermochem_np import Thermochemistry


# ---------------------------------------------------------------------------
# grid, Laplacian, dissipation profile
# ---------------------------------------------------------------------------
def uniform_grid(num_x, x_l=0.0, x_r=1.0):
    """``Domain`` (flamelets/domain.py:73-86)."""
    return np.linspace(x_l, x_r, num_x)


def clustered_grid(num_x, z_c, beta=3.0):
    """Points on [0, 1] clustered about ``z_c`` by the Vinokur/Roberts
    interior-clustering transform (Anderson, Computational Fluid Dynamics,
    Sec. 5.6): x = z_c [1 + sinh(beta (s - A)) / sinh(beta A)] with
    A = ln((1 + (e^beta - 1) z_c) / (1 + (e^-beta - 1) z_c)) / (2 beta)."""
    s = np.linspace(0.0, 1.0, num_x)
    A = np.log((1 + (np.exp(beta) - 1) * z_c)
               / (1 + (np.exp(-beta) - 1) * z_c)) / (2 * beta)
    x = z_c * (1 + np.sinh(beta * (s - A)) / np.sinh(beta * A))
    x[0], x[-1] = 0.0, 1.0
    return x


def laplacian_coefficients(x):
    """Per-row ``(a, b, c)`` with ``Lap_i = a_i f_{i-1} + b_i f_i + c_i f_{i+1}``.
    Boundary rows are identity rows scaled by ``1/h^2`` (``a_0 = c_0 = 0``,
    ``a_{N-1} = c_{N-1} = 0``)."""
    n = len(x)
    h = np.diff(x)
    a = np.zeros(n)
    b = np.zeros(n)
    c = np.zeros(n)
    hm, hp = h[:-1], h[1:]
    a[1:-1] = 2.0 / (hm * (hm + hp))
    c[1:-1] = 2.0 / (hp * (hm + hp))
    b[1:-1] = -(a[1:-1] + c[1:-1])
    b[0] = 1.0 / h[0] ** 2
    b[-1] = 1.0 / h[-1] ** 2
    return a, b, c


def apply_laplacian(abc, f):
    """Interior rows as ``a (f_{i-1} - f_i) + c (f_{i+1} - f_i)`` (the same
    operator, ``b = -(a + c)``): the neighbour differences are exact for a
    smooth f, so the rounding floor scales with the differences rather
    than with f -- which matters for the enthalpy row, where f ~ 1e5 J/kg
    and the exact Lap(h) is 0.  The boundary rows are ``b f``."""
    a, b, c = abc
    out = np.empty(np.shape(f), dtype=np.result_type(f, float))
    out[..., 0] = b[0] * f[..., 0]
    out[..., -1] = b[-1] * f[..., -1]
    out[..., 1:-1] = (a[1:-1] * (f[..., :-2] - f[..., 1:-1])
                      + c[1:-1] * (f[..., 2:] - f[..., 1:-1]))
    return out


def bell_profile(z):
    """``utils.bell_profile`` (flamelets/utils.py:9-26)."""
    return np.exp(-2 * erfcinv(2 * z) ** 2)


def dissipation_profile(x, z_st, chi_st):
    """chi(Z) = chi_st bell(Z)/bell(Z_st) (notebook cell 3) with the two
    boundary values set to 2/b_end so the Dirichlet rows become identity
    rows (the reference's 2 dx^2 on a uniform grid)."""
    a, b, c = laplacian_coefficients(x)
    chi = np.empty_like(x)
    chi[1:-1] = chi_st * bell_profile(x[1:-1]) / bell_profile(z_st)
    chi[0] = 2.0 / b[0]
    chi[-1] = 2.0 / b[-1]
    return chi


def stoichiometric_mixture_fraction(mech, y_ox, y_fu):
    """Bilger Z_st (``flamelets/utils.py:29-93``), including its element-weight
    table (which lists 28.0133 for N; N carries no Bilger weight)."""
    bilger = {"C": 2, "H": 0.5, "O": -1}
    wts_e = {"C": 12.011, "H": 1.00794, "O": 15.999, "Ar": 39.948,
             "N": 28.0133}
    w_s = mech.molecular_weights
    bw = np.array([bilger.get(e, 0) / wts_e[e] if e in bilger else 0.0
                   for e in mech.elements])
    em = np.array([[mech.n_atoms(s, e) * wts_e[mech.elements[e]] / w_s[s]
                    for s in range(mech.n_species)]
                   for e in range(mech.n_elements)])
    beta_ox = bw @ (em @ y_ox)
    beta_fu = bw @ (em @ y_fu)
    return -beta_ox / (beta_fu - beta_ox)


def mole_to_mass(mech, x):
    w = mech.molecular_weights
    return w * x / np.sum(w * x)


def stream_compositions(mech, t_ox=800.0, t_fu=300.0, fuel=None, ox=None):
    """Oxidiser air O2:N2 = 0.21:0.79 and fuel H2:N2 = 0.5:0.5 by mole."""
    idx = mech.species_index
    xo = np.zeros(mech.n_species)
    xf = np.zeros(mech.n_species)
    for k, v in (ox or {"O2": 0.21, "N2": 0.79}).items():
        xo[idx(k)] = v
    for k, v in (fuel or {"H2": 0.5, "N2": 0.5}).items():
        xf[idx(k)] = v
    gas = Thermochemistry(mech)
    y_ox, y_fu = mole_to_mass(mech, xo), mole_to_mass(mech, xf)
    h_ox = float(gas.get_mixture_enthalpy_mass(t_ox, y_ox))
    h_fu = float(gas.get_mixture_enthalpy_mass(t_fu, y_fu))
    return y_ox, y_fu, h_ox, h_fu


def burke_schumann(mech, x, y_ox, y_fu, h_ox, h_fu):
    """Closed-form initial state: h linear in Z; Y from the complete-combustion
    (H2 + 1/2 O2 -> H2O) limit, linear on each side of Z_st."""
    idx = mech.species_index
    z_st = stoichiometric_mixture_fraction(mech, y_ox, y_fu)
    w = mech.molecular_weights
    y_st = z_st * y_fu + (1 - z_st) * y_ox
    n_h2 = y_st[idx("H2")] / w[idx("H2")]
    n_o2 = y_st[idx("O2")] / w[idx("O2")]
    burnt = y_st.copy()
    react = min(n_h2, 2 * n_o2)
    burnt[idx("H2")] -= react * w[idx("H2")]
    burnt[idx("O2")] -= 0.5 * react * w[idx("O2")]
    burnt[idx("H2O")] += react * w[idx("H2O")]
    burnt = np.clip(burnt, 0.0, None)
    burnt /= burnt.sum()
    Y = np.empty((mech.n_species, len(x)))
    lean = x <= z_st
    f = np.where(lean, x / z_st, (x - z_st) / (1 - z_st))
    Y[:, lean] = (1 - f[lean]) * y_ox[:, None] + f[lean] * burnt[:, None]
    rich = ~lean
    Y[:, rich] = (1 - f[rich]) * burnt[:, None] + f[rich] * y_fu[:, None]
    h = h_ox + (h_fu - h_ox) * x
    return np.vstack([h[None, :], Y]), z_st


# ---------------------------------------------------------------------------
# temperature and source terms
# ---------------------------------------------------------------------------
def temperature_from_enthalpy(gas, h, Y, t_guess, tol=1e-10, maxit=100,
                              extra=0):
    """Per-point Newton on h(T, Y) = h: take the step, stop once
    ``|Re(step)| < tol`` (plus ``extra`` further steps, which complex-step
    derivatives need to converge the imaginary part)."""
    T = np.array(t_guess * np.ones_like(np.real(h)), dtype=np.result_type(h, Y,
                                                                         float))
    active = np.ones(T.shape, dtype=bool)
    remaining = np.full(T.shape, extra)
    for _ in range(maxit):
        f = h - gas.get_mixture_enthalpy_mass(T, Y)
        cp = gas.get_mixture_specific_heat_cp_mass(T, Y)
        step = f / cp
        T = np.where(active, T + step, T)
        done = np.abs(np.real(step)) < tol
        remaining = np.where(active & done, remaining - 1, remaining)
        active = active & ~(done & (remaining < 0))
        if not active.any():
            return T
    raise RuntimeError("Temperature iteration failed to converge")


def temperature_from_enthalpy_jax(gas, h, Y, t_guess, tol=1e-10,
                                  maxit=500, pointwise=True):
    """``PyroJAX.get_temperature_from_enthalpy`` (``flamelets/make_pyro.py:
    255-299``) as the reference runs it: a ``while_loop`` that computes the
    step, stops *before* taking it once its norm is ``<= tol``, and carries
    forward-mode tangents through exactly the iterations it takes.  Under
    ``vmap`` (the per-point source Jacobian, ``equations.py:124-129``) the
    norm is per point (``pointwise``); on a whole field (``rhs``) it is the
    2-norm over the field and every point steps until that is small.

    Complex inputs carry the tangent in their imaginary part: where the guess
    already satisfies the test no iteration runs and dT/dphi is exactly 0,
    as ``jacfwd`` through the reference's loop gives."""
    T = np.array(t_guess * np.ones_like(np.real(h)),
                 dtype=np.result_type(h, Y, float))
    active = np.ones(T.shape, dtype=bool)
    for _ in range(maxit):
        f = h - gas.get_mixture_enthalpy_mass(T, Y)
        step = f / gas.get_mixture_specific_heat_cp_mass(T, Y)
        if pointwise:
            active = active & (np.abs(np.real(step)) > tol)
        elif np.linalg.norm(np.real(np.atleast_1d(step))) <= tol:
            return T
        if not active.any():
            return T
        T = np.where(active, T + step, T)
    raise RuntimeError("Temperature iteration failed to converge")


class FlameletEquations:
    """Residual and block Jacobian of the steady flamelet equations.

    ``semantics`` selects the temperature recovery inside them:

    * ``"exact"`` (default): per-point Newton to ``1e-10`` K, and a source
      Jacobian that is the exact derivative of S (the derivative of T(h, Y)
      fully included), which is what the device kernels compute;
    * ``"reference"``: the reference's loops as JAX runs them
      (:func:`temperature_from_enthalpy_jax`): the field's T by the
      batch-norm loop and the Jacobian by forward-mode through the per-point
      loop started from ``t_guess`` -- so where ``t_guess`` is already the
      converged temperature of a point (the first Newton iteration after a
      time march, and the first iteration of every Crank-Nicolson step)
      the Jacobian omits dT/dphi there.  This reproduces the reference's
      Newton iterates, not only its converged state."""

    def __init__(self, mech, x, y_ox, y_fu, semantics="exact"):
        if semantics not in ("exact", "reference"):
            raise ValueError("semantics must be 'exact' or 'reference'")
        self.semantics = semantics
        self.mech = mech
        self.gas = Thermochemistry(mech)
        self.x = np.asarray(x, dtype=float)
        self.abc = laplacian_coefficients(self.x)
        self.nv = mech.n_species + 1
        self.y_ox = np.asarray(y_ox)
        self.y_fu = np.asarray(y_fu)
        n = len(self.x)
        self.mask = np.ones(n)
        self.mask[0] = self.mask[-1] = 0.0

    def boundary_conditions(self, h_ox, h_fu):
        n = len(self.x)
        bc = np.zeros((self.nv, n))
        bc[0, 0], bc[0, -1] = h_ox, h_fu
        bc[1:, 0], bc[1:, -1] = self.y_ox, self.y_fu
        return bc

    def source_terms(self, phi, vd, t_guess, p, extra=0, pointwise=True):
        """S = [vd/rho; W_k wdot_k / rho] (equations.py:313-359)."""
        h, Y = phi[0], phi[1:]
        if self.semantics == "reference":
            T = temperature_from_enthalpy_jax(self.gas, h, Y, t_guess,
                                              pointwise=pointwise)
        else:
            T = temperature_from_enthalpy(self.gas, h, Y, t_guess,
                                          extra=extra)
        rho = self.gas.get_density(p, T, Y)
        w = self.gas.get_net_production_rates(rho, T, Y)
        S = np.empty(phi.shape, dtype=np.result_type(phi, T, rho))
        S[0] = vd / rho
        S[1:] = self.gas.molecular_weights[:, None] * w / rho
        return S, T

    def rhs(self, phi, chi, vd, t_guess, p, h_ox, h_fu):
        """R = 1/2 chi Lap(phi) + S mask - phi_BC (equations.py:167-226)."""
        S, _ = self.source_terms(phi, vd, t_guess, p, pointwise=False)
        return (0.5 * chi * apply_laplacian(self.abc, phi) + S * self.mask
                - self.boundary_conditions(h_ox, h_fu))

    def source_jacobian(self, phi, vd, t_guess, p, h=1e-30):
        """dS/dphi per point, (N, nv, nv), by complex step."""
        n = phi.shape[1]
        J = np.empty((n, self.nv, self.nv))
        for d in range(self.nv):
            pert = phi.astype(complex)
            pert[d] += 1j * h
            S, _ = self.source_terms(pert, vd, t_guess, p, extra=3)
            J[:, :, d] = (np.imag(S) / h).T
        return J

    def jac(self, phi, chi, vd, t_guess, p):
        """(L (N-1,nv,nv), C (N,nv,nv), U (N-1,nv,nv)) (equations.py:228-282)."""
        a, b, c = self.abc
        eye = np.eye(self.nv)
        J = self.source_jacobian(phi, vd, t_guess, p)
        J[0] = 0.0
        J[-1] = 0.0
        C = (0.5 * chi * b)[:, None, None] * eye + J
        L = (0.5 * chi[1:] * a[1:])[:, None, None] * eye
        U = (0.5 * chi[:-1] * c[:-1])[:, None, None] * eye
        return L, C, U


# ---------------------------------------------------------------------------
# block Thomas (linear_solver.py:6-103)
# ---------------------------------------------------------------------------
def block_thomas(lower, central, upper, rhs):
    n = central.shape[0]
    c_tmp = np.empty_like(central)
    r_tmp = np.empty_like(rhs)
    c_tmp[0], r_tmp[0] = central[0], rhs[0]
    for i in range(1, n):
        s = np.linalg.solve(c_tmp[i - 1], upper[i - 1])
        t = np.linalg.solve(c_tmp[i - 1], r_tmp[i - 1])
        c_tmp[i] = central[i] - lower[i - 1] @ s
        r_tmp[i] = rhs[i] - lower[i - 1] @ t
    x = np.empty_like(rhs)
    x[-1] = np.linalg.solve(c_tmp[-1], r_tmp[-1])
    for i in range(n - 2, -1, -1):
        x[i] = np.linalg.solve(c_tmp[i], r_tmp[i] - upper[i] @ x[i + 1])
    return x


def assemble_dense(lower, central, upper):
    n, nv = central.shape[0], central.shape[1]
    A = np.zeros((n * nv, n * nv))
    for i in range(n):
        A[i * nv:(i + 1) * nv, i * nv:(i + 1) * nv] = central[i]
        if i < n - 1:
            A[i * nv:(i + 1) * nv, (i + 1) * nv:(i + 2) * nv] = upper[i]
            A[(i + 1) * nv:(i + 2) * nv, i * nv:(i + 1) * nv] = lower[i]
    return A


# ---------------------------------------------------------------------------
# solver (solver.py:63-600)
# ---------------------------------------------------------------------------
class FlameletSolver:
    def __init__(self, mech, x, y_ox, y_fu, semantics="exact"):
        self.eqns = FlameletEquations(mech, x, y_ox, y_fu, semantics)
        self.nv = self.eqns.nv
        self.log = []

    def _newton_update(self, fn, jac, phi, *args):
        r = fn(phi, *args)
        L, C, U = jac(phi, *args)
        return block_thomas(L, C, U, -r.T).T

    def flamelet_newton_step(self, phi, chi, vd, t_guess, p, h_ox, h_fu):
        return self._newton_update(
            lambda s, *a: self.eqns.rhs(s, *a),
            lambda s, chi_, vd_, tg, pp, *_: self.eqns.jac(s, chi_, vd_, tg, pp),
            phi, chi, vd, t_guess, p, h_ox, h_fu)

    def _newton_loop_py(self, phi, maxiter, tol, *args):
        err_prev = np.inf
        delta = np.inf
        it = 0
        for it in range(maxiter):
            v = self.flamelet_newton_step(phi, *args)
            phi = phi + v
            delta = np.linalg.norm(v)
            self.log.append(("newton", it, delta))
            if delta < tol:
                return phi, it, delta, True
            if err_prev < delta:
                return phi, it, delta, False
            err_prev = delta
        return phi, it, delta, False

    def flamelet_time_step(self, phi, newton_maxiter, newton_tol, phi_prev,
                           dt, chi, vd, t_guess, p, h_ox, h_fu):
        r_prev = self.eqns.rhs(phi_prev, chi, vd, t_guess, p, h_ox, h_fu)

        def fn(s):
            return s - phi_prev - 0.5 * dt * (
                self.eqns.rhs(s, chi, vd, t_guess, p, h_ox, h_fu) + r_prev)

        def jac(s):
            L, C, U = self.eqns.jac(s, chi, vd, t_guess, p)
            return -0.5 * dt * L, np.eye(self.nv) - 0.5 * dt * C, -0.5 * dt * U

        it, delta = 0, np.inf
        while delta > newton_tol and it < newton_maxiter:
            r = fn(phi)
            L, C, U = jac(phi)
            v = block_thomas(L, C, U, -r.T).T
            phi = phi + v
            it += 1
            delta = np.linalg.norm(v)
        return phi, it, delta

    def flamelet_time_march(self, newton_maxiter, newton_tol, maxsteps, dt,
                            chi, vd, t_guess, p, h_ox, h_fu, phi):
        temp = t_guess
        for step in range(maxsteps):
            phi_prev = phi
            phi, it, err = self.flamelet_time_step(
                phi, newton_maxiter, newton_tol, phi_prev, dt, chi, vd, temp,
                p, h_ox, h_fu)
            temp = temperature_from_enthalpy(self.eqns.gas, phi[0], phi[1:],
                                             t_guess)
            self.log.append(("cn", step, err, float(np.max(np.real(temp)))))
        return phi, temp

    def equilibrate(self, phi, t_guess, p, h_ox, h_fu, steps=24, dt0=1e-8,
                    growth=2.5, newton_maxiter=20, newton_tol=1e-9):
        """Crank-Nicolson steps with zero interior dissipation and a
        geometrically growing step: relaxes each interior point to chemical
        equilibrium at its (h, p) (the device solver's ``equilibrate``)."""
        a, b, c = self.eqns.abc
        chi0 = np.zeros(len(self.eqns.x))
        chi0[0], chi0[-1] = 2.0 / b[0], 2.0 / b[-1]
        vd = np.zeros(len(self.eqns.x))
        dt = dt0
        for _ in range(steps):
            phi, it, err = self.flamelet_time_step(
                phi, newton_maxiter, newton_tol, phi, dt, chi0, vd, t_guess, p,
                h_ox, h_fu)
            dt *= growth
        T = temperature_from_enthalpy(self.eqns.gas, phi[0], phi[1:], t_guess)
        return phi, T

    def solve(self, newton_maxiter, newton_tol, bdf_newton_maxiter,
              bdf_newton_tol, bdf_time_step, bdf_maxsteps, try_newton,
              max_attempts, chi, vd, t_guess, p, h_ox, h_fu, phi_guess):
        gas = self.eqns.gas
        if try_newton:
            phi, it, err, ok = self._newton_loop_py(
                phi_guess, newton_maxiter, newton_tol, chi, vd, t_guess, p,
                h_ox, h_fu)
            if ok:
                T = temperature_from_enthalpy(gas, phi[0], phi[1:], t_guess)
                return phi, T, True
        phi, temp = phi_guess, t_guess
        for _ in range(max_attempts):
            phi, temp = self.flamelet_time_march(
                bdf_newton_maxiter, bdf_newton_tol, bdf_maxsteps,
                bdf_time_step, chi, vd, temp, p, h_ox, h_fu, phi)
            phi_t, it, err, ok = self._newton_loop_py(
                phi, newton_maxiter, newton_tol, chi, vd, temp, p, h_ox, h_fu)
            T_t = temperature_from_enthalpy(gas, phi_t[0], phi_t[1:], t_guess)
            if ok:
                return phi_t, T_t, True
        return phi, temp, False
