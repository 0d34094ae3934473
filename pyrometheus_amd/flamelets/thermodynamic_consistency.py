"""Compressible-EOS consistency on the device (``CompressibleEOS``).

The reference's layer over the flamelet solver
(``pyrometheus/flamelets/thermodynamic_consistency.py:44-706``) finds the
boundary parameters ``(p, h_ox, h_fu)`` whose flamelet reproduces a
filtered density and internal energy ``(rho_sim, e_sim)`` under a presumed
mixture-fraction PDF, by Gauss-Newton (density and energy gradients from two
adjoint solves) or Picard (enthalpy gradient only) updates.  Same class,
method names, arguments, configuration schema and return values here; every
array is a float64 tensor on the ROCm device and every step runs HIP
kernels:

* the source Jacobian (``pyro_flamelet_source`` or ``pyro_source_jacobian``) turned
  in place into the adjoint operator (``pyro_flamelet_adjoint_blocks``:
  ``equations.py:451-516``),
* the adjoint solves through the same block-Thomas kernels as the forward
  Newton update (one call per right-hand side),
* R T and its gradient (``pyro_flamelet_eos_rt``: ``equations.py:518-535``),
* dS/dp (``equations.py:130-135``) from the production-rate JVP in rho,
  dS_k/dp = (W_k / rho) (d omega_k/d rho (rho / p) - omega_k / p),
  dS_0/dp = -vd / (rho p),

and the PDF reductions are device reductions.  The boundary traces use the
reference's stencil ``(1.5 l_0 + 2 l_1 - 0.5 l_2) / dx^2`` with the domain's
``dx``, as written for its uniform grids.
"""

from __future__ import annotations

import time
from typing import Dict

import numpy as np
import torch

from .state import FlameletState, _array_to_state, _state_to_array


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def trapezoidal_rule(integrand: torch.Tensor) -> torch.Tensor:
    """Composite trapezoidal sum on a unit-spacing grid
    (``thermodynamic_consistency.py:30-41``)."""
    return 0.5 * (integrand[0] + integrand[-1]) + integrand[1:-1].sum()


class CompressibleEOS:
    """``CompressibleEOS(config, forward_solver)`` with a device
    ``FlameletSolver``; ``config`` has the reference's schema
    (``verbosity``, ``newton``, ``bdf``, ``max_attempts``, ``eos``)."""

    def __init__(self, config: Dict, forward_solver):
        self.config = config
        self.fwd_solver = forward_solver
        eq = forward_solver.gov_eqns
        self.batch = eq.batch
        if self.batch.B != 1:
            raise ValueError("the forward solver must hold one flamelet")
        self.gas = eq.pyro_gas
        self.num_v = self.gas.num_species + 1
        self.num_x = self.batch.P
        dev = self.batch.dev
        self.unit_h = torch.zeros((self.num_v, self.num_x),
                                  dtype=torch.float64, device=dev)
        self.unit_h[0] = 1.0
        self._rt = torch.empty(self.num_x, dtype=torch.float64, device=dev)
        self._rtj = torch.empty((self.num_v, self.num_x), dtype=torch.float64,
                                device=dev)
        self._fail = torch.zeros(1, dtype=torch.int32, device=dev)

    def update_config_option(self, option_path: str, option_val):
        """``config[a][b]...[z] = option_val`` for ``"a/b/.../z"``."""
        keys = option_path.split("/")
        current = self.config
        for k in keys[:-1]:
            if k not in current:
                current[k] = {}
            if not isinstance(current[k], dict):
                raise TypeError(f"Path segment {k} is not a dict")
            current = current[k]
        current[keys[-1]] = option_val

    # -- device pieces ------------------------------------------------------
    def _fields(self, state, diss_rate, viscous_diss, temp_guess, pressure):
        return self.fwd_solver.gov_eqns._args(state, diss_rate, viscous_diss,
                                              temp_guess, pressure)

    def _eos_rt(self, phi, tg, jac: bool):
        b = self.batch
        self._fail.zero_()
        b.lib.call("pyro_flamelet_eos_rt", phi.data_ptr(), phi.stride(0),
                   tg.data_ptr(), 1, self._rt.data_ptr(),
                   self._rtj.data_ptr() if jac else None, self.num_x,
                   self.num_x, 1e-10, 100, self._fail.data_ptr(),
                   1 if b.semantics == "reference" else 0, _stream())
        return self._rt.clone(), (self._rtj.clone() if jac else None)

    def _adjoint_operator(self, phi, chi, vd, tg, p):
        """Source Jacobian at (phi, tg), then the adjoint operator in place
        of it (batch J, L, U)."""
        b = self.batch
        b.source(phi, vd, tg, p, jac=True)
        b.lib.call("pyro_flamelet_adjoint_blocks", b.J.data_ptr(),
                   chi.data_ptr(), b.la.data_ptr(), b.lb.data_ptr(),
                   b.lc.data_ptr(), b.sys.data_ptr(), b.off.data_ptr(),
                   b.L.data_ptr(), b.U.data_ptr(), b.P, _stream())

    def _adjoint_solve(self, rhs: torch.Tensor) -> torch.Tensor:
        """Solve the assembled adjoint operator for ``rhs`` (nv, N)."""
        b = self.batch
        b.rhs.copy_(rhs.T)
        b.thomas()
        return b.x.T.contiguous()

    def _source_gradient_wrt_pressure(self, phi, vd, p):
        """dS/dp per point at fixed (h, Y) and the batch's T, (N, nv)."""
        from ..adjoint import wdot_jvp
        b, gas = self.batch, self.gas
        T = b.T
        Y = phi[1:]
        rho = gas.get_density(p, T, Y)
        w, jv = wdot_jvp(gas, rho, T, Y, rho / p, None, None)
        out = torch.empty((self.num_x, self.num_v), dtype=torch.float64,
                          device=b.dev)
        out[:, 0] = -vd / (rho * p)
        out[:, 1:] = (gas.molecular_weights[:, None]
                      * (jv - w / p) / rho).T
        return out

    def _trace(self, lam_h: torch.Tensor) -> torch.Tensor:
        d2 = self.fwd_solver.domain.dx ** 2
        ox = (1.5 * lam_h[0] + 2 * lam_h[1] - 0.5 * lam_h[2]) / d2
        fu = (1.5 * lam_h[-1] + 2 * lam_h[-2] - 0.5 * lam_h[-3]) / d2
        return ox + fu

    # -- reference methods ----------------------------------------------------
    def enthalpy_gradient(self, state: FlameletState, mixture_fraction_pdf,
                          diss_rate, viscous_diss, temp_guess, pressure):
        """``(rt, adjoint_state_h, h_gradient)``
        (``thermodynamic_consistency.py:106-166``)."""
        phi, chi, vd, tg, p = self._fields(state, diss_rate, viscous_diss,
                                           temp_guess, pressure)
        pdf = _field(mixture_fraction_pdf, phi)
        rt, _ = self._eos_rt(phi, tg, jac=False)
        self._adjoint_operator(phi, chi, vd, tg, p)
        lam = self._adjoint_solve(-self.unit_h * pdf)
        return rt, _array_to_state(lam), self._trace(lam[0])

    def eos_gradient(self, state: FlameletState, mixture_fraction_pdf,
                     diss_rate, viscous_diss, temp_guess, pressure):
        """``((density_gradient, energy_gradient), (adjoint_state_d,
        adjoint_state_e))``, each gradient ``(d/dp, d/dh_boundary)``
        (``thermodynamic_consistency.py:168-308``)."""
        phi, chi, vd, tg, p = self._fields(state, diss_rate, viscous_diss,
                                           temp_guess, pressure)
        pdf = _field(mixture_fraction_pdf, phi)
        rt, rtj = self._eos_rt(phi, tg, jac=True)
        self._adjoint_operator(phi, chi, vd, tg, p)
        ds_dp = self._source_gradient_wrt_pressure(phi, vd, p) / (
            0.5 * chi[:, None])
        lam_d = self._adjoint_solve((p / rt ** 2) * rtj * pdf)
        integrand = (1.0 / rt) * pdf + (ds_dp * lam_d.T).sum(1)
        dens = torch.stack((trapezoidal_rule(integrand),
                            self._trace(lam_d[0])))
        lam_e = self._adjoint_solve(-(self.unit_h - rtj) * pdf)
        integrand = (ds_dp * lam_e.T).sum(1)
        ener = torch.stack((trapezoidal_rule(integrand),
                            self._trace(lam_e[0])))
        return ((dens, ener), (_array_to_state(lam_d), _array_to_state(lam_e)))

    def _solve(self, params, diss_rate, viscous_diss, temp_guess,
               state_guess):
        c = self.config
        p, h_ox, h_fu = (float(v) for v in params)
        return self.fwd_solver.solve(
            c["verbosity"], c["newton"]["maxiter"], c["newton"]["tol"],
            c["bdf"]["newton"]["maxiter"], c["bdf"]["newton"]["tol"],
            c["bdf"]["time_step"], c["bdf"]["maxsteps"], True,
            c["max_attempts"], diss_rate, viscous_diss, temp_guess, p, h_ox,
            h_fu, state_guess)

    def evaluate_flamelet(self, params, mixture_fraction_pdf, diss_rate,
                          viscous_diss, temp_guess, state_guess):
        """``(state, temperature, density, energy, density_gradient,
        energy_gradient)`` (``thermodynamic_consistency.py:310-398``)."""
        pressure = float(params[0])
        t0 = time.time()
        state, temp = self._solve(params, diss_rate, viscous_diss, temp_guess,
                                  state_guess)
        if self.config["verbosity"]:
            print(f"solve time: {(time.time() - t0):.4e} s")
        t0 = time.time()
        (dg, eg), _ = self.eos_gradient(state, mixture_fraction_pdf,
                                        diss_rate, viscous_diss, temp_guess,
                                        pressure)
        if self.config["verbosity"]:
            torch.cuda.synchronize()
            print(f"adjoint time: {(time.time() - t0):.4e} s")
        phi = _state_to_array(state).contiguous()
        tg = self.fwd_solver.gov_eqns._args(state, 0.0, 0.0, temp, pressure)[3]
        rt, _ = self._eos_rt(phi, tg, jac=False)
        pdf = _field(mixture_fraction_pdf, phi)
        density = ((pressure / rt) * pdf).sum()
        energy = ((state.enthalpy - rt) * pdf).sum()
        return state, temp, density, energy, dg, eg

    def _gauss_newton_update(self, density_sim, energy_sim, params,
                             mixture_fraction_pdf, diss_rate, viscous_diss,
                             temp_guess, state_guess):
        """``(state, temperature, update, residual)``
        (``thermodynamic_consistency.py:400-450``): the normal equations of
        the 2x2 density/energy Jacobian, the boundary-enthalpy step applied
        to both h_ox and h_fu."""
        state, temp, dens, ener, dg, eg = self.evaluate_flamelet(
            params, mixture_fraction_pdf, diss_rate, viscous_diss, temp_guess,
            state_guess)
        G = torch.stack((dg, eg))
        res = torch.stack((dens - density_sim, ener - energy_sim))
        v = torch.linalg.solve(G.T @ G, -G.T @ res)
        return state, temp, torch.stack((v[0], v[1], v[1])), res

    def _picard_update(self, density_sim, energy_sim, params,
                       mixture_fraction_pdf, diss_rate, viscous_diss,
                       temp_guess, state_guess):
        """``(state, temperature, update, residual)``
        (``thermodynamic_consistency.py:452-540``): p from rho <R T> = p,
        the boundary enthalpy from the enthalpy-only adjoint gradient."""
        pressure = float(params[0])
        state, temp = self._solve(params, diss_rate, viscous_diss, temp_guess,
                                  state_guess)
        rt, _, dh = self.enthalpy_gradient(state, mixture_fraction_pdf,
                                           diss_rate, viscous_diss,
                                           temp_guess, pressure)
        pdf = _field(mixture_fraction_pdf, rt[None])
        density = ((pressure / rt) * pdf).sum()
        energy = ((state.enthalpy - rt) * pdf).sum()
        dp = density_sim * (rt * pdf).sum() - pressure
        dhb = (energy_sim - energy) / dh
        update = torch.stack((dp, dhb, dhb))
        residual = torch.stack((density_sim - density, energy_sim - energy))
        return state, temp, update, residual

    def warmup(self, params, mixture_fraction_pdf, diss_rate, viscous_diss,
               temp_wmp, state_wmp):
        """One forward solve and one of each adjoint solve (the reference
        warms its JIT caches; here it loads the libraries and allocates)."""
        self._solve(params, diss_rate, viscous_diss, temp_wmp, state_wmp)
        self.enthalpy_gradient(state_wmp, mixture_fraction_pdf, diss_rate,
                               viscous_diss, temp_wmp, float(params[0]))
        self.eos_gradient(state_wmp, mixture_fraction_pdf, diss_rate,
                          viscous_diss, temp_wmp, float(params[0]))

    def ensure_consistency(self, density_sim, energy_sim, params,
                           mixture_fraction_pdf, diss_rate, viscous_diss,
                           temp_guess, state_guess):
        """``(state, temperature, params, it, delta, residual, history,
        False)`` (``thermodynamic_consistency.py:602-706``)."""
        c = self.config["eos"]
        a = c["update_size"]
        method = c["update_method"]
        if method == "gauss_newton":
            update_fn = self._gauss_newton_update
        elif method == "picard":
            update_fn = self._picard_update
        else:
            raise ValueError(f"Available {method} not implemented")
        params = torch.as_tensor(np.asarray(params, dtype=float),
                                 dtype=torch.float64, device=self.batch.dev)
        state, temp = state_guess, temp_guess
        residual = np.full(c["maxiter"], np.nan)
        history = np.full((c["maxiter"], 2), np.nan)
        delta, it = np.inf, 0
        pdf = None
        for it in range(c["maxiter"]):
            t0 = time.time()
            state, temp, v, res = update_fn(density_sim, energy_sim, params,
                                            mixture_fraction_pdf, diss_rate,
                                            viscous_diss, temp, state)
            cost_val = float(torch.linalg.norm(res) ** 2)
            delta = float(torch.linalg.norm(v))
            if self.config["verbosity"]:
                print(f"EOS iteration {it}: time: {(time.time() - t0):.4e} s"
                      f", residual = {cost_val:.4e}, |v| = {delta:.4e}")
            params = params + a * v
            residual[it] = cost_val
            if pdf is None:
                pdf = _field(mixture_fraction_pdf, state.enthalpy[None])
            history[it] = [float((state.enthalpy * pdf).sum()),
                           float((temp * pdf).sum())]
            if delta < c["tol"]:
                break
        return state, temp, params, it, delta, residual, history, False


def _field(x, like: torch.Tensor) -> torch.Tensor:
    """A per-point weight as a float64 device tensor of the grid's length."""
    if isinstance(x, torch.Tensor):
        return x.to(device=like.device, dtype=torch.float64)
    return torch.as_tensor(np.asarray(x, dtype=float), dtype=torch.float64,
                           device=like.device)
