"""Iteration-level scheduler for batches of flamelets whose membership can
change between iterations (continuous batching).

Every flamelet runs the control flow of the reference's hybrid solve
(``flamelets/solver.py:476-600``: optional Newton try, Crank-Nicolson time
march of ``bdf_maxsteps`` steps as globalisation, Newton, up to
``max_attempts`` restarts), optionally preceded by an equilibration phase.
``step()`` advances every live flamelet by one Newton update with one launch
sequence (``FlameletBatch.newton_update``), whichever phase each one is in,
then retires the flamelets that finished.  ``add()`` admits new flamelets;
they join at the next step.  When membership changes the live flamelets are
re-packed into a new contiguous batch; a flamelet's arithmetic never depends
on its neighbours (per-point kernels, per-flamelet norms and control flow),
so its result is the same bit for bit alone, in any batch, and whenever it
joined.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from .engine import FlameletBatch

NEWTON_TRY, MARCH, NEWTON, DONE_OK, DONE_FAIL, EQUIL = range(6)


@dataclass
class SolveSettings:
    """The reference's ``solve`` arguments (``solver.py:476-492``) plus the
    equilibration phase and the update norm."""
    newton_maxiter: int = 10
    newton_tol: float = 1e-8
    bdf_newton_maxiter: int = 20
    bdf_newton_tol: float = 1e-9
    bdf_time_step: float = 1e-5
    bdf_maxsteps: int = 10
    try_newton: bool = False
    max_attempts: int = 10
    equilibrate_steps: int = 0
    equilibrate_dt0: float = 1e-8
    equilibrate_growth: float = 2.5
    newton_norm: str = "reference"
    jacobian: str = "auto"  # FlameletBatch source-Jacobian form


# per-flamelet integer/float control state, in this order
_CTRL = ("phase", "it", "err_prev", "step", "inner", "attempt", "n_newton",
         "n_steps", "eq_step", "fail")
#: ``fail`` codes: why a flamelet ended in DONE_FAIL
FAIL_NONE, FAIL_ITER, FAIL_NONFINITE = 0, 1, 2
FAIL_REASONS = {
    FAIL_ITER: "no attempt converged (max_attempts reached)",
    FAIL_NONFINITE: "update not finite: a singular block or a temperature "
                    "that did not converge",
}
# per-flamelet device arrays: (nv, N) state-shaped and (N,) point-shaped
_STATE = ("phi", "phi_guess", "phi_prev", "r_prev", "phi_save")
_POINT = ("tg", "chi", "chi0", "vd", "pres")


@dataclass
class _Slot:
    handle: int
    problem: object
    ctrl: Dict[str, float]
    arrays: Dict[str, torch.Tensor]
    t_added: float
    finished_at: float = 0.0


#: rows of the device control state, (len(_DCTL), B) int32, in this order
_DCTL = ("phase", "it", "step", "inner", "attempt", "n_newton", "n_steps",
         "eq_step", "fail")


def _eq_dt_table(s: SolveSettings, n: int) -> np.ndarray:
    """Equilibration time steps dt0 growth^k for k < n, each the product of
    k factors taken left to right (the device control kernel computes the
    same products, so both give the same bits)."""
    d = [float(s.equilibrate_dt0)]
    for _ in range(n - 1):
        d.append(d[-1] * float(s.equilibrate_growth))
    return np.array(d)


class FlameletStepper:
    """Advances a batch of flamelets, one Newton update of every live one
    per iteration (see the module docstring).

    ``device_control`` (default: with ``graphs``): every flamelet's control
    flow runs on the device after each update (``pyro_flamelet_control``),
    so ``launch()`` queues ``sync_every`` iterations back to back and the
    host reads the control state only once per ``sync_every`` iterations.
    The iterates are those of the host control flow bit for bit; finished
    flamelets are masked on the device at once and retired by the host at
    the next sync."""

    def __init__(self, gas, settings: Optional[SolveSettings] = None,
                 verbosity=False, device=None,
                 kernel_events: Optional[dict] = None,
                 graphs: Optional[bool] = None,
                 device_control: Optional[bool] = None,
                 sync_every: int = 8):
        from .solver import NEWTON_NORMS
        self.gas = gas
        self.s = settings or SolveSettings()
        if self.s.newton_norm not in NEWTON_NORMS:
            raise ValueError(f"unknown Newton norm {self.s.newton_norm!r}")
        self.dev = torch.device(device) if device is not None else gas.device
        self.nv = gas.num_species + 1
        self.verbosity = verbosity
        self._slots: List[_Slot] = []    # live flamelets, batch order
        self._pending: List[_Slot] = []  # admitted, not yet packed
        self._next = 0
        self.batch: Optional[FlameletBatch] = None
        self.iterations = 0
        self.point_iterations = 0  # sum over iterations of the batch points
        self.t0 = time.time()
        # when a dict is given, every batch records HIP events into it
        self.kernel_events = kernel_events
        # replay the update's launch sequence from a captured HIP graph
        # (captured again whenever the batch is re-packed); off while HIP
        # events are recorded per kernel
        if graphs is None:
            graphs = self.dev.type == "cuda" and kernel_events is None
        if graphs and kernel_events is not None:
            raise ValueError("graph replay and per-kernel events exclude "
                             "each other")
        self.graphs = bool(graphs)
        if device_control is None:
            device_control = self.graphs
        self.device_control = bool(device_control)
        self.sync_every = max(1, int(sync_every))
        self._dt_eq = _eq_dt_table(self.s, max(1, self.s.equilibrate_steps))
        self._graph = None
        self._inflight = None  # (event, live, eq, cn) between launch/complete
        self.graph_captures = 0
        # when a list is given, step() appends the (B, 2) update norms
        self.norm_log: Optional[list] = None

    # -- membership -------------------------------------------------------
    def add(self, problems: Sequence) -> List[int]:
        """Admit flamelets (``FlameletProblem``); returns their handles.  The
        problems' arrays go to the device in one copy per kind; each
        flamelet's arrays are views of those."""
        s, dev, dd = self.s, self.dev, dict(dtype=torch.float64,
                                            device=self.dev)
        problems = list(problems)
        if not problems:
            return []
        first = NEWTON_TRY if s.try_newton else MARCH
        sizes = [len(p.x) for p in problems]
        offs = np.concatenate([[0], np.cumsum(sizes)])
        host = np.concatenate([np.asarray(p.phi, dtype=float)
                               for p in problems], axis=1)
        phi_all = torch.as_tensor(host, **dd)
        guess_all = phi_all.clone()
        spare = torch.empty((3,) + tuple(phi_all.shape), **dd)
        pts = np.stack([
            np.concatenate([np.asarray(p.t_guess, dtype=float)
                            for p in problems]),
            np.concatenate([np.asarray(p.chi, dtype=float) for p in problems]),
            np.concatenate([np.zeros(n) if p.vd is None
                            else np.asarray(p.vd, dtype=float)
                            for p, n in zip(problems, sizes)]),
            np.concatenate([np.full(n, float(p.pressure))
                            for p, n in zip(problems, sizes)])])
        chi0 = pts[1].copy()
        for a, e in zip(offs[:-1], offs[1:]):
            chi0[a + 1:e - 1] = 0.0  # equilibration: chi only at the ends
        pt_all = torch.as_tensor(np.vstack([pts, chi0[None]]), **dd)
        out = []
        for j, p in enumerate(problems):
            a, e = int(offs[j]), int(offs[j + 1])
            arrays = {
                "phi": phi_all[:, a:e], "phi_guess": guess_all[:, a:e],
                "phi_prev": spare[0, :, a:e], "r_prev": spare[1, :, a:e],
                "phi_save": spare[2, :, a:e],
                "tg": pt_all[0, a:e], "chi": pt_all[1, a:e],
                "vd": pt_all[2, a:e], "pres": pt_all[3, a:e],
                "chi0": pt_all[4, a:e],
            }
            ctrl = dict(phase=EQUIL if s.equilibrate_steps > 0 else first,
                        it=0, err_prev=np.inf, step=0, inner=0, attempt=0,
                        n_newton=0, n_steps=0, eq_step=0, fail=FAIL_NONE)
            slot = _Slot(self._next, p, ctrl, arrays, time.time())
            self._next += 1
            self._pending.append(slot)
            out.append(slot.handle)
        return out

    def _live_slots(self):
        if self.batch is None:
            return list(self._slots)
        ph = self.C["phase"]
        return [sl for j, sl in enumerate(self._slots)
                if ph[j] not in (DONE_OK, DONE_FAIL)]

    @property
    def n_live(self) -> int:
        return len(self._live_slots()) + len(self._pending)

    @property
    def points(self) -> int:
        return sum(len(sl.problem.x)
                   for sl in self._live_slots() + self._pending)

    @staticmethod
    def bytes_per_point(nv: int) -> int:
        """Device bytes a flamelet point holds while it is live."""
        return 8 * (nv * nv + nv * (nv + 1) + (len(_STATE) + 4) * nv
                    + len(_POINT) + 10)

    def _pack(self):
        """Build a new batch from the live flamelets of the current one (their
        state gathered from the packed arrays in one index_select per array)
        followed by the newly admitted ones."""
        kept, gather = [], None
        if self.batch is not None:
            ph = self.C["phase"]
            keep = [j for j in range(len(self._slots))
                    if ph[j] not in (DONE_OK, DONE_FAIL)]
            for j in keep:
                sl = self._slots[j]
                for k in _CTRL:
                    sl.ctrl[k] = self.C[k][j]
                kept.append(sl)
            if keep:
                idx = np.concatenate([np.arange(self.off[j], self.off[j + 1])
                                      for j in keep])
                gather = (self.A, torch.as_tensor(idx, device=self.dev))
        else:
            kept = [sl for sl in self._slots
                    if sl.ctrl["phase"] not in (DONE_OK, DONE_FAIL)]
        self.batch = None
        new = self._pending
        self._slots = kept + new
        self._pending = []
        if not self._slots:
            return
        probs = [sl.problem for sl in self._slots]
        b = FlameletBatch(self.gas, [p.x for p in probs],
                          np.stack([p.y_ox for p in probs]),
                          np.stack([p.y_fu for p in probs]), device=self.dev,
                          jacobian=self.s.jacobian)
        h_ox = np.array([p.h_ox for p in probs], dtype=float)
        h_fu = np.array([p.h_fu for p in probs], dtype=float)
        b.set_boundary_enthalpy(h_ox, h_fu)
        if self.kernel_events is not None:
            b.events = self.kernel_events
        from .solver import enthalpy_scale
        self.h_scale = enthalpy_scale(h_ox, h_fu)
        self.batch = b
        self.off = b.offsets_np
        def packed(k, dim):
            pieces = []
            if gather is not None:
                pieces.append(gather[0][k].index_select(dim, gather[1]))
            pieces += [sl.arrays[k] for sl in new]
            return torch.cat(pieces, dim=dim).contiguous()
        self.A = {k: packed(k, 1) for k in _STATE}
        self.A.update({k: packed(k, 0) for k in _POINT})
        self.C = {k: np.array([sl.ctrl[k] for sl in self._slots],
                              dtype=float if k == "err_prev" else np.int64)
                  for k in _CTRL}
        self._graph = None
        # control block [modes | active | dt] of 4B int32 words, pinned on
        # the host; the previous upload from it has finished when step()
        # returns (step() reads the norms back before it returns)
        pinned = self.dev.type == "cuda"
        self._ctl_pin = torch.zeros(4 * b.B, dtype=torch.int32,
                                    pin_memory=pinned)
        self._ctl_np = self._ctl_pin.numpy()
        self._ctl_dev = torch.zeros(4 * b.B, dtype=torch.int32,
                                    device=self.dev)
        b.dt = self._ctl_dev[2 * b.B:].view(torch.float64)
        # the per-flamelet norms come back through pinned host memory
        self._norm_pin = torch.zeros((b.B, 2), dtype=torch.float64,
                                     pin_memory=pinned)
        self._norm_np = self._norm_pin.numpy()
        self.sys = b.sys.long()
        for sl in self._slots:  # the packed arrays now own the state
            sl.arrays = {}
        self._controls_uploaded = False
        if self.device_control:
            from ..runtime import load_core_library
            self._core = load_core_library()
            rows = np.stack([self.C[k] for k in _DCTL]).astype(np.int32)
            self._dctl = torch.as_tensor(rows, device=self.dev).contiguous()
            self._derr = torch.as_tensor(self.C["err_prev"],
                                         dtype=torch.float64,
                                         device=self.dev).contiguous()
            self._action = torch.zeros(b.B, dtype=torch.int32,
                                       device=self.dev)
            self._hscale = torch.as_tensor(
                np.broadcast_to(self.h_scale, (b.B,)).copy(),
                dtype=torch.float64, device=self.dev)
            self._dctl_pin = torch.zeros(tuple(rows.shape), dtype=torch.int32,
                                         pin_memory=pinned)
            self._derr_pin = torch.zeros(b.B, dtype=torch.float64,
                                         pin_memory=pinned)
            s = self.s
            self._settings = np.array(
                [s.newton_maxiter, s.newton_tol, s.bdf_newton_maxiter,
                 s.bdf_newton_tol, s.bdf_time_step, s.bdf_maxsteps,
                 int(bool(s.try_newton)), s.max_attempts,
                 s.equilibrate_steps, s.equilibrate_dt0,
                 s.equilibrate_growth, int(s.newton_norm == "scaled")],
                dtype=np.float64)
            # flamelets already returned to the caller
            self._retired = np.zeros(b.B, dtype=bool)

    # -- one iteration ----------------------------------------------------
    def _pts(self, mask_sys):
        m = torch.as_tensor(mask_sys, dtype=torch.bool, device=self.dev)
        return m[self.sys]

    def step(self):
        """Advance every live flamelet by one update; returns the list of
        ``(handle, FlameletResult)`` of the flamelets that finished."""
        self.launch()
        return self.complete()

    def launch(self) -> bool:
        """First half of ``step``: enqueue one update of every live flamelet
        on the current stream and the copy of its norms to host memory,
        without waiting; ``complete`` finishes the step.  Between the two
        the host may serve other work (another batch's launch)."""
        if self._inflight is not None:
            raise RuntimeError("complete() the previous step first")
        if self._pending or self.batch is None:
            self._pack()
        if self.batch is None:
            return False
        if self.device_control:
            return self._launch_device()
        live, eq, cn = self._upload_controls()
        s, b, A = self.s, self.batch, self.A
        phi, tg, chi, chi0, vd, pres = (A["phi"], A["tg"], A["chi"],
                                        A["chi0"], A["vd"], A["pres"])
        nb = b.B
        modes, active = self._ctl_dev[:nb], self._ctl_dev[nb:2 * nb]
        def update():
            b.newton_update(phi, chi, vd, tg, pres, modes, A["phi_prev"],
                            A["r_prev"], apply=True, active=active,
                            chi0=chi0, refresh_guess=True)
        if self.graphs:
            if self._graph is None:
                self._graph = self._capture(update)
            self._graph.replay()
        else:
            update()
        self._norm_pin[:nb].copy_(b.norm2, non_blocking=True)
        ev = torch.cuda.Event() if self.dev.type == "cuda" else None
        if ev is not None:
            ev.record(torch.cuda.current_stream(self.dev))
        self._inflight = (ev, live, eq, cn)
        return True

    def _upload_controls(self):
        """Launch controls of every flamelet from the host control state,
        uploaded in one copy; returns the (live, eq, cn) masks."""
        s, b, C = self.s, self.batch, self.C
        phase = C["phase"]
        live = (phase != DONE_OK) & (phase != DONE_FAIL)
        eq = phase == EQUIL
        cn = (phase == MARCH) | eq
        k = np.minimum(C["eq_step"], len(self._dt_eq) - 1)
        dts = np.where(eq, self._dt_eq[k], s.bdf_time_step)
        # the per-flamelet controls travel in one upload: modes (bit 0
        # Crank-Nicolson, bit 1 the equilibration dissipation rate, bit 2
        # start of a time step), the active flags and the time steps (b.dt
        # is a view of the upload)
        # a flamelet starting a time step (bit 2) has the assembly kernel
        # store phi_prev = phi and r_prev = R(phi) before it uses them
        nb, host = b.B, self._ctl_np
        starting = cn & (C["inner"] == 0)
        host[:nb] = (cn.astype(np.int32) | (eq.astype(np.int32) << 1)
                     | (starting.astype(np.int32) << 2))
        host[nb:2 * nb] = live
        host[2 * nb:4 * nb].view(np.float64)[:] = dts
        self._ctl_dev.copy_(self._ctl_pin, non_blocking=True)
        return live, eq, cn

    def _launch_device(self) -> bool:
        """``launch`` with the control flow on the device: ``sync_every``
        iterations (update, control decision, state copies) replayed back
        to back, then the control state copied to host memory."""
        b, A = self.batch, self.A
        nb = b.B
        if not self._controls_uploaded:
            self._upload_controls()
            self._controls_uploaded = True
        phi, tg, chi, chi0, vd, pres = (A["phi"], A["tg"], A["chi"],
                                        A["chi0"], A["vd"], A["pres"])
        modes, active = self._ctl_dev[:nb], self._ctl_dev[nb:2 * nb]
        core, st = self._core, lambda: torch.cuda.current_stream(
            self.dev).cuda_stream
        def iteration():
            b.newton_update(phi, chi, vd, tg, pres, modes, A["phi_prev"],
                            A["r_prev"], apply=True, active=active,
                            chi0=chi0, refresh_guess=True)
            core.call("pyro_flamelet_control",
                      self._settings.ctypes.data, self._dctl.data_ptr(),
                      self._derr.data_ptr(), b.norm2.data_ptr(),
                      self._hscale.data_ptr(), modes.data_ptr(),
                      active.data_ptr(), b.dt.data_ptr(),
                      self._action.data_ptr(), nb, st())
            core.call("pyro_flamelet_state_copy", phi.data_ptr(),
                      A["phi_guess"].data_ptr(), A["phi_save"].data_ptr(),
                      phi.stride(0), self.nv, b.sys.data_ptr(),
                      self._action.data_ptr(), b.P, st())
        k = self.sync_every
        if self.graphs:
            if self._graph is None:
                self._graph = self._capture(iteration)
            for _ in range(k):
                self._graph.replay()
        else:
            for _ in range(k):
                iteration()
        self._dctl_pin.copy_(self._dctl, non_blocking=True)
        self._derr_pin.copy_(self._derr, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.dev))
        self._inflight = (ev, k)
        return True

    def _complete_device(self):
        """``complete`` with the control flow on the device: read the
        control state back, return the flamelets that finished since the
        last sync."""
        from .solver import FlameletResult
        from .state import _array_to_state
        ev, k = self._inflight
        self._inflight = None
        ev.synchronize()
        b, A, C = self.batch, self.A, self.C
        rows = self._dctl_pin.numpy()
        for j, name in enumerate(_DCTL):
            C[name][:] = rows[j]
        C["err_prev"][:] = self._derr_pin.numpy()
        self.iterations += k
        self.point_iterations += k * b.P
        now = time.time()
        phase = C["phase"]
        done = ((phase == DONE_OK) | (phase == DONE_FAIL)) & ~self._retired
        if not done.any():
            return []
        self._retired |= done
        phi, tg, vd, pres = A["phi"], A["tg"], A["vd"], A["pres"]
        b.source(phi, vd, tg, pres, jac=False)
        out = []
        for j in np.nonzero(done)[0]:
            sl = self._slots[j]
            a, e = int(self.off[j]), int(self.off[j + 1])
            sl.finished_at = now
            out.append((sl.handle, FlameletResult(
                _array_to_state(phi[:, a:e].clone()), b.T[a:e].clone(),
                bool(phase[j] == DONE_OK), int(C["n_newton"][j]),
                int(C["n_steps"][j]), int(C["attempt"][j]),
                now - sl.t_added, FAIL_REASONS.get(int(C["fail"][j])))))
        if self._retired.all():
            self._slots, self.batch = [], None
        else:
            sizes = np.diff(self.off)
            if sizes[self._retired].sum() * 2 > sizes.sum():
                self._pack()
        return out

    def complete(self):
        """Second half of ``step``: wait for the launched update, run every
        flamelet's control flow on its norms, retire the finished ones;
        returns their ``(handle, FlameletResult)``."""
        from .solver import FlameletResult, update_norm
        from .state import _array_to_state
        if self._inflight is None:
            return []
        if self.device_control:
            return self._complete_device()
        ev, live, eq, cn = self._inflight
        self._inflight = None
        if ev is not None:
            ev.synchronize()
        s, b, A, C = self.s, self.batch, self.A, self.C
        phase = C["phase"]
        phi, tg, vd, pres = A["phi"], A["tg"], A["vd"], A["pres"]
        sums = self._norm_np[:b.B].copy()
        if self.norm_log is not None:
            self.norm_log.append(sums.copy())
        delta = update_norm(sums, s.newton_norm, self.h_scale)
        self.iterations += 1
        self.point_iterations += b.P
        now = time.time()
        it, err_prev = C["it"], C["err_prev"]
        # a non-finite update cannot recover (a singular block, or a point
        # whose temperature did not converge): stop that flamelet now -- but
        # a failed Newton try still falls back to the time march from the
        # guess, as any failed try does
        bad_try = live & ~np.isfinite(delta) & (phase == NEWTON_TRY)
        nonfinite = live & ~np.isfinite(delta) & ~bad_try
        if nonfinite.any():
            phase[nonfinite] = DONE_FAIL
            C["fail"][nonfinite] = FAIL_NONFINITE
            live = live & ~nonfinite
            eq, cn = eq & ~nonfinite, cn & ~nonfinite
        newton = live & ((phase == NEWTON_TRY) | (phase == NEWTON))
        C["n_newton"] += newton
        ok = newton & (delta < s.newton_tol)
        bad = newton & ~ok & ((err_prev < delta) | (it + 1 >= s.newton_maxiter)
                              | bad_try)
        cont = newton & ~ok & ~bad
        phase[ok] = DONE_OK
        err_prev[cont] = delta[cont]
        it[cont] += 1
        f0 = bad & (phase == NEWTON_TRY)
        if f0.any():
            pm = self._pts(f0)
            phi[:, pm] = A["phi_guess"][:, pm]
            phase[f0], C["step"][f0], C["inner"][f0] = MARCH, 0, 0
        f1 = bad & (phase == NEWTON)
        if f1.any():
            pm = self._pts(f1)
            phi[:, pm] = A["phi_save"][:, pm]
            C["attempt"][f1] += 1
            again = f1 & (C["attempt"] < s.max_attempts)
            phase[again], C["step"][again], C["inner"][again] = MARCH, 0, 0
            phase[f1 & ~again] = DONE_FAIL
            C["fail"][f1 & ~again] = FAIL_ITER
        march = live & cn
        C["inner"][march] += 1
        step_done = march & ((delta <= s.bdf_newton_tol)
                             | (C["inner"] >= s.bdf_newton_maxiter))
        C["inner"][step_done] = 0
        eq_done = step_done & eq
        C["eq_step"][eq_done] += 1
        eq_finished = eq_done & (C["eq_step"] >= s.equilibrate_steps)
        if eq_finished.any():
            pm = self._pts(eq_finished)
            A["phi_guess"][:, pm] = phi[:, pm]
            phase[eq_finished] = NEWTON_TRY if s.try_newton else MARCH
            it[eq_finished], err_prev[eq_finished] = 0, np.inf
        step_done = step_done & ~eq
        C["step"][step_done] += 1
        C["n_steps"] += step_done
        to_newton = step_done & (C["step"] >= s.bdf_maxsteps)
        if to_newton.any():
            pm = self._pts(to_newton)
            A["phi_save"][:, pm] = phi[:, pm]
            phase[to_newton] = NEWTON
            it[to_newton], err_prev[to_newton] = 0, np.inf
        if self.verbosity:
            print(f"batch iteration: live {int(live.sum())}, "
                  f"max |v| {np.max(delta[live]):.3e}")
        done = (live | nonfinite) & ((phase == DONE_OK)
                                     | (phase == DONE_FAIL))
        live = live | nonfinite
        if not done.any():
            return []
        # final temperature of the finished flamelets from their last guess
        b.source(phi, vd, tg, pres, jac=False)
        out = []
        for j in np.nonzero(done)[0]:
            sl = self._slots[j]
            a, e = int(self.off[j]), int(self.off[j + 1])
            sl.finished_at = now
            st = _array_to_state(phi[:, a:e].clone())
            out.append((sl.handle, FlameletResult(
                st, b.T[a:e].clone(), bool(phase[j] == DONE_OK),
                int(C["n_newton"][j]), int(C["n_steps"][j]),
                int(C["attempt"][j]), now - sl.t_added,
                FAIL_REASONS.get(int(C["fail"][j])))))
        if done.sum() == live.sum():
            self._slots, self.batch = [], None
        else:
            # finished flamelets stay in the batch, masked out of every
            # update, until they hold half its points (or flamelets join)
            dead = ~live | done
            sizes = np.diff(self.off)
            if sizes[dead].sum() * 2 > sizes.sum():
                self._pack()
        return out

    def _capture(self, fn):
        """Record the kernels ``fn`` launches into a HIP graph (nothing runs
        during the capture).  Every array the kernels touch belongs to the
        packed batch, so the graph is valid until the next re-pack; the
        per-flamelet controls are read from device memory at replay."""
        g = torch.cuda.CUDAGraph()
        cur = torch.cuda.current_stream(self.dev)
        side = torch.cuda.Stream(self.dev)  # default priority
        side.wait_stream(cur)
        with torch.cuda.stream(side):
            g.capture_begin()
            try:
                fn()
            finally:
                g.capture_end()
        cur.wait_stream(side)
        self.graph_captures += 1
        return g

    def run(self) -> Dict[int, object]:
        """Step until every admitted flamelet has finished."""
        results = {}
        while self.n_live:
            for h, r in self.step():
                results[h] = r
        return results
