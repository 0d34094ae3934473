"""Steady flamelet equations and solvers on the device.

``FlameletEquations`` and ``FlameletSolver`` keep the reference's names,
argument orders and control flow (``pyrometheus/flamelets/equations.py``,
``solver.py:16-600``): a host loop per Newton iteration that reads back the
update norm, a Crank-Nicolson time march as globalisation, and the hybrid
``solve`` with its restart attempts.  Every evaluation (temperature recovery,
source terms and their Jacobian, assembly, block-Thomas solve, update) is a
HIP kernel; tensors are float64 on the ROCm device.

``solve_batch`` solves many flamelets (uneven grid sizes, different
dissipation rates, pressures or boundary states) at once: every flamelet runs
the same state machine as ``solve`` on its own, and one launch sequence per
iteration advances all unfinished flamelets, whichever phase each is in.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from .engine import FlameletBatch
from .state import FlameletState, _array_to_state, _state_to_array


def _t(x, dev, n=None):
    if isinstance(x, torch.Tensor):
        t = x.to(device=dev, dtype=torch.float64)
    else:
        t = torch.as_tensor(np.asarray(x, dtype=float), dtype=torch.float64,
                            device=dev)
    if n is not None and t.numel() == 1:
        t = t.reshape(1).expand(n)
    return t.contiguous()


SEMANTICS = ("reference", "exact")


class FlameletEquations:
    """Residual and block Jacobian of the steady flamelet equations.

    ``semantics="reference"`` recovers T(h, Y) with the reference's loop and
    differentiates through it as ``jax.jacfwd`` does (``equations.py:
    124-129`` over ``make_pyro.py:255-299``): where the temperature guess
    already meets the loop's test, no iteration runs and the source Jacobian
    has no dT/dphi term.  That is the case at the first Newton iteration
    after a time march and at the first iteration of every Crank-Nicolson
    step, so it shapes the reference's Newton iterates (not its converged
    state).  ``"exact"`` uses the converged T and its exact derivative; it
    needs no register-Jacobian kernel and serves mechanisms of any size."""

    def __init__(self, pyro_gas, domain, y_ox, y_fu, semantics="exact"):
        self.pyro_gas = pyro_gas
        self.domain = domain
        self.mol_wts = pyro_gas.molecular_weights
        self.batch = FlameletBatch(pyro_gas, [domain.x],
                                   np.asarray(_cpu(y_ox)),
                                   np.asarray(_cpu(y_fu)),
                                   semantics=semantics)
        self.num_v = pyro_gas.num_species + 1

    def _args(self, state, diss_rate, viscous_diss, temp_guess, pressure):
        b, dev = self.batch, self.batch.dev
        phi = _state_to_array(state).to(dev).contiguous()
        n = b.P
        chi = _t(diss_rate, dev, n)
        vd = _t(viscous_diss, dev, n)
        tg = _t(temp_guess, dev, n)
        p = _t(pressure, dev, n)
        return phi, chi, vd, tg, p

    def boundary_conditions(self, h_ox, h_fu):
        n, ns = self.batch.P, self.pyro_gas.num_species
        dev = self.batch.dev
        h = torch.zeros(n, dtype=torch.float64, device=dev)
        y = torch.zeros((ns, n), dtype=torch.float64, device=dev)
        h[0], h[-1] = float(h_ox), float(h_fu)
        y[:, 0] = torch.as_tensor(self.batch.y_ox[0], dtype=torch.float64,
                                  device=dev)
        y[:, -1] = torch.as_tensor(self.batch.y_fu[0], dtype=torch.float64,
                                   device=dev)
        return FlameletState(h, y)

    def equation_of_state(self, state, temp_guess, pressure):
        phi, _, vd, tg, p = self._args(state, 0.0, 0.0, temp_guess, pressure)
        self.batch.source(phi, vd, tg, p, jac=False)
        T = self.batch.T.clone()
        rho = self.pyro_gas.get_density(p, T, phi[1:])
        return rho, T

    def source_terms(self, state, viscous_diss, temp_guess, pressure):
        phi, _, vd, tg, p = self._args(state, 0.0, viscous_diss, temp_guess,
                                       pressure)
        self.batch.source(phi, vd, tg, p, jac=False)
        return _array_to_state(self.batch.S.clone())

    def source_jacobian(self, state, viscous_diss, temp_guess, pressure):
        """dS/dphi per point (N, nv, nv) (``vmap(jacfwd(local_source_terms))``,
        equations.py:124-129)."""
        phi, _, vd, tg, p = self._args(state, 0.0, viscous_diss, temp_guess,
                                       pressure)
        self.batch.source(phi, vd, tg, p, jac=True)
        return self.batch.J.clone()

    def rhs(self, state, diss_rate, viscous_diss, temp_guess, pressure, h_ox,
            h_fu):
        phi, chi, vd, tg, p = self._args(state, diss_rate, viscous_diss,
                                         temp_guess, pressure)
        self.batch.set_boundary_enthalpy(h_ox, h_fu)
        out = torch.empty_like(phi)
        self.batch.residual(phi, chi, vd, tg, p, out)
        return _array_to_state(out)

    def jac(self, state, diss_rate, viscous_diss, temp_guess, pressure):
        """(lower (N-1,nv,nv), central (N,nv,nv), upper (N-1,nv,nv))."""
        phi, chi, vd, tg, p = self._args(state, diss_rate, viscous_diss,
                                         temp_guess, pressure)
        b = self.batch
        b.source(phi, vd, tg, p, jac=True)
        b.assemble(phi, chi)
        eye = torch.eye(self.num_v, dtype=torch.float64, device=b.dev)
        lower = b.L[:-1, None, None] * eye
        upper = b.U[:-1, None, None] * eye
        return lower, b.J.clone(), upper


def _cpu(x):
    return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else x


NEWTON_NORMS = ("reference", "scaled")


def enthalpy_scale(h_ox, h_fu):
    """Magnitude of the enthalpy row, max(|h_ox|, |h_fu|, 1 J/kg): the scale
    of its rounding error."""
    return np.maximum(np.maximum(np.abs(h_ox), np.abs(h_fu)), 1.0)


def update_norm(sums, kind="reference", h_scale=None):
    """Newton update norm per flamelet from the sums of squares of v over
    the enthalpy row and over the species rows (``sums`` (B, 2)).

    ``"reference"``: |v|_2 over every row, the criterion of
    ``solver.py:136-139``.  With h in J/kg its floor in double precision
    rises steeply with N (the rounding error of the Laplacian of h grows
    like N^2 |h|), so on fine grids it is never reached.  ``"scaled"``: the
    enthalpy row is divided by ``h_scale`` (``enthalpy_scale``), the species
    rows are kept as they are."""
    sums = np.asarray(sums, dtype=float).reshape(-1, 2)
    if kind == "reference":
        return np.sqrt(sums[:, 0] + sums[:, 1])
    if kind == "scaled":
        return np.sqrt(sums[:, 0] / np.asarray(h_scale, dtype=float) ** 2
                       + sums[:, 1])
    raise ValueError(f"unknown Newton norm {kind!r}; use one of "
                     f"{NEWTON_NORMS}")


class FlameletSolver:
    """Plain drop-in: one flamelet, host loop per Newton iteration.

    ``semantics`` (``FlameletEquations``): the default ``"reference"``
    reproduces the reference's Newton iterates for mechanisms of at most 16
    species (the register-Jacobian kernel); larger mechanisms use
    ``"exact"``."""

    def __init__(self, domain, pyro_gas, mass_frac_bcs,
                 newton_norm="reference", semantics=None):
        y_ox, y_fu = mass_frac_bcs
        self.domain = domain
        if semantics is None:
            semantics = ("reference" if pyro_gas._lib.has(
                "pyro_flamelet_source") else "exact")
        if semantics not in SEMANTICS:
            raise ValueError(f"unknown semantics {semantics!r}")
        self.semantics = semantics
        self.gov_eqns = FlameletEquations(pyro_gas, domain, y_ox, y_fu,
                                          semantics)
        self.num_v = pyro_gas.num_species + 1
        if newton_norm not in NEWTON_NORMS:
            raise ValueError(f"unknown Newton norm {newton_norm!r}")
        self.newton_norm = newton_norm
        self.log = []
        self.failure = None

    # one Newton update v = -J^-1 R (solver.py:63-101)
    def _update(self, state, diss_rate, viscous_diss, temp_guess, pressure,
                h_ox, h_fu, mode=0, state_prev=None, time_step=0.0):
        eq = self.gov_eqns
        b = eq.batch
        phi, chi, vd, tg, p = eq._args(state, diss_rate, viscous_diss,
                                       temp_guess, pressure)
        b.set_boundary_enthalpy(h_ox, h_fu)
        if mode == 1:
            phi_prev = _state_to_array(state_prev).to(b.dev).contiguous()
            r_prev = torch.empty_like(phi_prev)
            b.residual(phi_prev, chi, vd, tg, p, r_prev)
            b.modes.fill_(1)
            b.dt.fill_(float(time_step))
            b.newton_update(phi, chi, vd, tg, p, b.modes, phi_prev, r_prev,
                            apply=False)
        else:
            b.newton_update(phi, chi, vd, tg, p, apply=False)
        v = b.x.T.contiguous()
        delta = update_norm(b.norm2.cpu().numpy(), self.newton_norm,
                            enthalpy_scale(h_ox, h_fu))
        return _array_to_state(v), float(delta[0])

    def flamelet_newton_step(self, state_guess, diss_rate, viscous_diss,
                             temp_guess, pressure, h_ox, h_fu):
        return self._update(state_guess, diss_rate, viscous_diss, temp_guess,
                            pressure, h_ox, h_fu)[0]

    def _newton_loop_py(self, verbosity, state, maxiter, tol, *args):
        err_prev = np.inf
        delta = np.inf
        it = 0
        for it in range(maxiter):
            t0 = time.time()
            v, delta = self._update(state, *args)
            state = state + v
            self.log.append(("newton", it, delta))
            if verbosity:
                print(f"Newton iter {it}: time: {(time.time() - t0):.4e} s, "
                      f"|v| = {delta:.4e}")
            if not np.isfinite(delta):
                # a singular block or a temperature that did not converge:
                # the reference would iterate on NaN; stop instead
                self.failure = "non-finite update"
                return state, it, delta, False
            if delta < tol:
                return state, it, delta, True
            if err_prev < delta:
                return state, it, delta, False
            err_prev = delta
        return state, it, delta, False

    def flamelet_time_step(self, state, newton_maxiter, newton_tol, state_prev,
                           time_step, diss_rate, viscous_diss, temp_guess,
                           pressure, h_ox, h_fu):
        it, delta = 0, np.inf
        while delta > newton_tol and it < newton_maxiter:
            v, delta = self._update(state, diss_rate, viscous_diss, temp_guess,
                                    pressure, h_ox, h_fu, mode=1,
                                    state_prev=state_prev,
                                    time_step=time_step)
            state = state + v
            it += 1
            if not np.isfinite(delta):
                self.failure = "non-finite update"
                break
        return state, it, delta

    def flamelet_time_march(self, verbosity, newton_maxiter, newton_tol,
                            maxsteps, time_step, diss_rate, viscous_diss,
                            temp_guess, pressure, h_ox, h_fu, state_guess):
        state, temp = state_guess, temp_guess
        for step in range(maxsteps):
            state_prev = state
            t0 = time.time()
            state, it, err = self.flamelet_time_step(
                state, newton_maxiter, newton_tol, state_prev, time_step,
                diss_rate, viscous_diss, temp, pressure, h_ox, h_fu)
            if self.failure is not None:
                break
            temp = self.gov_eqns.pyro_gas.get_temperature_from_enthalpy(
                state.enthalpy, state.mass_fractions,
                _t(temp_guess, state.enthalpy.device, state.enthalpy.numel()))
            self.log.append(("cn", step, err, float(temp.max())))
            if verbosity:
                print(f"BDF time step {step}: time: {time.time() - t0:.4e} s, "
                      f"T_max = {float(temp.max()):.3f} K ")
        return state, temp

    def equilibrate(self, state, temp_guess, pressure, h_ox, h_fu,
                    steps=24, dt0=1e-8, growth=2.5, newton_maxiter=20,
                    newton_tol=1e-9):
        """Relax each interior point towards chemical equilibrium at its own
        (h, p): Crank-Nicolson steps of the flamelet equations with the
        interior dissipation rate set to zero and the step growing
        geometrically.  Gives the burning initial state that the reference
        notebook obtains from Cantera's ``equilibrate('HP')``."""
        n = state.enthalpy.numel()
        dev = state.enthalpy.device
        chi0 = torch.zeros(n, dtype=torch.float64, device=dev)
        e0, e1 = self.domain.boundary_dissipation
        chi0[0], chi0[-1] = e0, e1
        vd = torch.zeros(n, dtype=torch.float64, device=dev)
        dt = dt0
        self.failure = None
        for _ in range(steps):
            state, it, err = self.flamelet_time_step(
                state, newton_maxiter, newton_tol, state, dt, chi0, vd,
                temp_guess, pressure, h_ox, h_fu)
            if self.failure is not None:
                raise RuntimeError("equilibration failed: " + self.failure)
            dt *= growth
        temp = self.gov_eqns.pyro_gas.get_temperature_from_enthalpy(
            state.enthalpy, state.mass_fractions, _t(temp_guess, dev, n))
        return state, temp

    def warmup(self, warmup_fn_name, *args):
        if warmup_fn_name == "flamelet_newton_step":
            return self.flamelet_newton_step(*args)
        if warmup_fn_name == "flamelet_time_step":
            return self.flamelet_time_step(*args)
        raise ValueError(f"Unknown warmup function name {warmup_fn_name}")

    def solve(self, verbosity, newton_maxiter, newton_tol, bdf_newton_maxiter,
              bdf_newton_tol, bdf_time_step, bdf_maxsteps, try_newton,
              max_attempts, diss_rate, viscous_diss, temp_guess, pressure,
              h_ox, h_fu, state_guess):
        """Hybrid Newton / time-march solve (solver.py:476-600); returns
        ``(state, temperature)`` and never raises on non-convergence."""
        gas = self.gov_eqns.pyro_gas
        dev = state_guess.enthalpy.device
        n = state_guess.enthalpy.numel()
        self.converged = False
        self.failure = None
        if try_newton:
            state, it, err, ok = self._newton_loop_py(
                verbosity, state_guess, newton_maxiter, newton_tol, diss_rate,
                viscous_diss, temp_guess, pressure, h_ox, h_fu)
            if ok:
                self.converged = True
                return state, gas.get_temperature_from_enthalpy(
                    state.enthalpy, state.mass_fractions,
                    _t(temp_guess, dev, n))
        state, temp = state_guess, temp_guess
        # a Newton try that failed (even on a non-finite update) falls back
        # to the time march from the guess, as in the reference
        self.failure = None
        for i in range(max_attempts):
            if verbosity:
                print(f"Attempt {i} of BDF time marching")
            state, temp = self.flamelet_time_march(
                verbosity, bdf_newton_maxiter, bdf_newton_tol, bdf_maxsteps,
                bdf_time_step, diss_rate, viscous_diss, temp, pressure, h_ox,
                h_fu, state)
            if self.failure is not None:
                break
            state_t, it, err, ok = self._newton_loop_py(
                verbosity, state, newton_maxiter, newton_tol, diss_rate,
                viscous_diss, temp, pressure, h_ox, h_fu)
            if self.failure is not None:
                break
            temp_t = gas.get_temperature_from_enthalpy(
                state_t.enthalpy, state_t.mass_fractions,
                _t(temp_guess, dev, n))
            if ok:
                self.converged = True
                return state_t, temp_t
        if verbosity:
            print("Flamelet solution did not converge")
        return state, temp


# ---------------------------------------------------------------------------
# batched solve
# ---------------------------------------------------------------------------
@dataclass
class FlameletProblem:
    """One steady flamelet: grid ``x`` (N,), dissipation ``chi`` (N,) with the
    identity-row values at the ends, initial state ``phi`` (nv, N) and
    temperature guess (N,), boundary streams, pressure."""

    x: np.ndarray
    chi: np.ndarray
    phi: np.ndarray
    t_guess: np.ndarray
    y_ox: np.ndarray
    y_fu: np.ndarray
    h_ox: float
    h_fu: float
    pressure: float = 101325.0
    vd: Optional[np.ndarray] = None


@dataclass
class FlameletResult:
    state: FlameletState
    temperature: torch.Tensor
    converged: bool
    newton_iterations: int
    time_steps: int
    attempts: int
    finished_at: float = 0.0
    #: why the solve failed (None when it converged)
    failure: Optional[str] = None


def solve_batch(pyro_gas, problems: Sequence[FlameletProblem],
                newton_maxiter=10, newton_tol=1e-8, bdf_newton_maxiter=20,
                bdf_newton_tol=1e-9, bdf_time_step=1e-5, bdf_maxsteps=10,
                try_newton=False, max_attempts=10, verbosity=False,
                equilibrate_steps=0, equilibrate_dt0=1e-8,
                equilibrate_growth=2.5, newton_norm="reference",
                kernel_events: Optional[dict] = None,
                stats: Optional[dict] = None,
                jacobian: str = "auto",
                graphs: Optional[bool] = None,
                device_control: Optional[bool] = None,
                sync_every: int = 8) -> List[FlameletResult]:
    """Solve every problem with the control flow of ``FlameletSolver.solve``
    applied to each flamelet independently (preceded, when
    ``equilibrate_steps`` > 0, by ``FlameletSolver.equilibrate``); one launch
    sequence per iteration advances every unfinished flamelet
    (``FlameletStepper``).  ``newton_norm`` is the update norm of
    ``update_norm``; ``kernel_events``, if a dict, collects HIP events around
    every kernel (``FlameletBatch.kernel_ms``); ``stats``, if a dict,
    receives the iteration count and the sum of batch points over them;
    ``jacobian`` picks the source-Jacobian kernels (``FlameletBatch``:
    "dense", "general" or "auto");
    ``graphs`` replays each iteration's launches from a captured HIP graph
    (default: on a GPU unless ``kernel_events`` is given);
    ``device_control`` (default: with ``graphs``) runs every flamelet's
    control flow on the device and reads it back once per ``sync_every``
    iterations (the same results bit for bit)."""
    from .stepper import FlameletStepper, SolveSettings
    settings = SolveSettings(
        newton_maxiter=newton_maxiter, newton_tol=newton_tol,
        bdf_newton_maxiter=bdf_newton_maxiter, bdf_newton_tol=bdf_newton_tol,
        bdf_time_step=bdf_time_step, bdf_maxsteps=bdf_maxsteps,
        try_newton=try_newton, max_attempts=max_attempts,
        equilibrate_steps=equilibrate_steps, equilibrate_dt0=equilibrate_dt0,
        equilibrate_growth=equilibrate_growth, newton_norm=newton_norm,
        jacobian=jacobian)
    st = FlameletStepper(pyro_gas, settings, verbosity=verbosity,
                         kernel_events=kernel_events, graphs=graphs,
                         device_control=device_control,
                         sync_every=sync_every)
    handles = st.add(problems)
    results = st.run()
    if stats is not None:
        stats.update(iterations=st.iterations,
                     point_iterations=st.point_iterations)
    return [results[h] for h in handles]
