"""Device engine shared by the plain and batched flamelet solvers.

A :class:`FlameletBatch` holds B flamelets whose grid points are concatenated
(P = sum N_j points): state ``phi`` (nv, P) with row 0 the enthalpy, the
per-point Laplacian coefficients and dissipation rate, and the scratch the
kernels need.  One Newton update for every flamelet of the batch is one
launch sequence:

1. ``pyro_flamelet_source`` -- per point: T(h, Y) by Newton, rho, S and the
   (nv x nv) source Jacobian (chain rule through the analytic kinetics
   Jacobian held in registers; mechanisms with at most 16 species), or, for
   larger mechanisms, ``pyro_temperature`` for T(h, Y) followed by
   ``pyro_source_jacobian`` (each row evaluated in one pass from T and
   the concentrations);
2. ``pyro_flamelet_assemble`` -- residual and block-tridiagonal Jacobian,
   steady or Crank-Nicolson per flamelet;
3. ``pyro_flamelet_thomas_partitioned`` -- block Thomas with the Laplacian
   coupling blocks as scalar multiples of the identity, each flamelet cut
   into partitions swept from both ends and joined through an nv-sized
   interface recursion, nv <= 64 (``pyro_flamelet_thomas``, one wavefront
   per flamelet, for larger nv or when asked);
4. ``pyro_flamelet_update`` -- phi += v for the selected flamelets and the
   per-flamelet sums of squares of v over the enthalpy row and over the
   species rows (fixed-order reductions, so a flamelet's result does not
   depend on what else is in the batch).
"""

from __future__ import annotations

from typing import List, Optional, Sequence

import math

import numpy as np
import torch

from ..runtime import HipError
from .domain import laplacian_coefficients


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _ptr(t):
    return None if t is None else t.data_ptr()


def partition_table(sizes, part_len=None, alpha=None):
    """Cut systems of the given sizes into consecutive partitions of about
    ``part_len`` rows (default alpha sqrt(N) per system, at least 8, with
    alpha = sqrt 2 unless given) of equal length up to one row: returns
    numpy arrays pstart, pend, psys, jpart, jbefore, jafter, roff (see
    ``pyro_flamelet_thomas_partitioned``)."""
    a = math.sqrt(2.0) if alpha is None else float(alpha)
    n = np.asarray(sizes, dtype=np.int64)
    if part_len:
        m = np.full(n.shape, int(part_len), dtype=np.int64)
    else:
        m = np.maximum(8, np.round(a * np.sqrt(n)).astype(np.int64))
    npart = np.maximum(1, -(-n // m))
    offs = np.concatenate([[0], np.cumsum(n)])
    first = np.concatenate([[0], np.cumsum(npart)])
    psys = np.repeat(np.arange(n.size), npart)
    q = np.arange(first[-1]) - first[psys]
    nq, npq, o = n[psys], npart[psys], offs[psys]
    pstart = o + (q * nq) // npq
    pend = o + ((q + 1) * nq) // npq
    roff = np.concatenate([[0], np.cumsum(npart - 1)])
    inner = q < npq - 1
    jbefore = np.where(q > 0, roff[psys] + q - 1, -1)
    jafter = np.where(inner, roff[psys] + q, -1)
    jpart = np.nonzero(inner)[0]
    i64 = lambda v: np.asarray(v, dtype=np.int64)
    i32 = lambda v: np.asarray(v, dtype=np.int32)
    return (i64(pstart), i64(pend), i32(psys), i32(jpart), i64(jbefore),
            i64(jafter), i64(roff))


def _rows(t, name, writable=False):
    """A contiguous (rows, points) array, the layout the kernels read with
    one leading dimension for all state arrays; read-only inputs in another
    layout are copied, arrays the kernels write must already conform."""
    if t is None or t.is_contiguous():
        return t
    if writable:
        raise ValueError(f"{name} must be a contiguous (rows, points) array")
    return t.contiguous()


class FlameletBatch:
    def __init__(self, gas, grids: Sequence[np.ndarray], y_ox, y_fu,
                 device=None, jacobian: str = "auto",
                 thomas: str = "auto", part_len: Optional[int] = None,
                 semantics: str = "exact"):
        lib = gas._lib
        if not lib.has("pyro_flamelet_assemble"):
            raise HipError(f"library of {gas._mech.name} has no flamelet "
                           "kernels")
        self.gas, self.lib = gas, lib
        # "exact": T(h, Y) converged and its exact derivative in the source
        # Jacobian; "reference": the reference's JAX loop and the tangent
        # jacfwd carries through it (pyro_flamelet_source, semantics 1)
        if semantics not in ("exact", "reference"):
            raise ValueError(f"unknown semantics {semantics!r}")
        # "dense": the register Jacobian of pyro_flamelet_source (at most 16
        # species); "general": pyro_source_jacobian after pyro_temperature
        # (any mechanism; "cached" names it too, after its first form, which
        # went through a per-reaction cache); "auto" takes the first
        # available
        if jacobian == "cached":
            jacobian = "general"
        if jacobian not in ("auto", "dense", "general"):
            raise ValueError(f"unknown jacobian form {jacobian!r}")
        if semantics == "reference" and (jacobian == "general" or not lib.has(
                "pyro_flamelet_source")):
            raise HipError("reference semantics need the register-Jacobian "
                           "source kernel (at most 16 species)")
        self.semantics = semantics
        if jacobian == "dense" and not lib.has("pyro_flamelet_source"):
            raise HipError(f"{gas._mech.name}: no register-Jacobian kernel "
                           "(more than 16 species)")
        self.dense = (jacobian != "general"
                      and lib.has("pyro_flamelet_source"))
        self._jac_work = None
        self._iters = None
        self.dev = torch.device(device) if device is not None else gas.device
        self.ns = gas.num_species
        self.nv = self.ns + 1
        self.sizes = [len(x) for x in grids]
        if min(self.sizes) < 3:
            raise ValueError("each flamelet needs at least 3 grid points")
        self.B = len(self.sizes)
        offs = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.offsets_np = offs
        self.P = int(offs[-1])
        dd = dict(dtype=torch.float64, device=self.dev)
        abc = [laplacian_coefficients(np.asarray(x, dtype=float)) for x in grids]
        self.la = torch.tensor(np.concatenate([a for a, _, _ in abc]), **dd)
        self.lb = torch.tensor(np.concatenate([b for _, b, _ in abc]), **dd)
        self.lc = torch.tensor(np.concatenate([c for _, _, c in abc]), **dd)
        self.off = torch.tensor(offs, dtype=torch.int64, device=self.dev)
        self.sys = torch.tensor(np.repeat(np.arange(self.B), self.sizes),
                                dtype=torch.int32, device=self.dev)
        y_ox = np.broadcast_to(np.asarray(y_ox, dtype=float), (self.B, self.ns))
        y_fu = np.broadcast_to(np.asarray(y_fu, dtype=float), (self.B, self.ns))
        self.y_ox, self.y_fu = y_ox.copy(), y_fu.copy()
        self.bc = torch.zeros((self.B, 2, self.nv), **dd)
        self.bc[:, 0, 1:] = torch.tensor(self.y_ox, **dd)
        self.bc[:, 1, 1:] = torch.tensor(self.y_fu, **dd)
        P, nv = self.P, self.nv
        self.S = torch.empty((nv, P), **dd)
        self.J = torch.empty((P, nv, nv), **dd)
        self.rhs = torch.empty((P, nv), **dd)
        self.x = torch.empty((P, nv), **dd)
        self.L = torch.empty(P, **dd)
        self.U = torch.empty(P, **dd)
        self.work = torch.empty(P * nv * (nv + 1), **dd)
        self.norm2 = torch.empty((self.B, 2), **dd)
        self.T = torch.empty(P, **dd)
        self.n_fail = torch.zeros(1, dtype=torch.int32, device=self.dev)
        self.modes = torch.zeros(self.B, dtype=torch.int32, device=self.dev)
        self.dt = torch.zeros(self.B, **dd)
        self.active = torch.ones(self.B, dtype=torch.int32, device=self.dev)
        self.none_active = torch.zeros(self.B, dtype=torch.int32,
                                       device=self.dev)
        # kernel name -> list of (start, end) HIP events when timing is on
        self.events = None
        # block-Thomas schedule: "auto" takes the one tuned for this block
        # size and device (flamelets.autotune), "sequential" or
        # "partitioned" force one
        if thomas not in ("auto", "sequential", "partitioned"):
            raise ValueError(f"unknown block-Thomas schedule {thomas!r}")
        alpha, sweep = None, "mfma"
        if thomas == "auto":
            from .autotune import schedule
            sch = schedule(self.nv, self.dev)
            thomas, alpha, sweep = sch["kind"], sch["alpha"], sch["sweep"]
        self.partitioned = False
        self.set_thomas(thomas, alpha, part_len, sweep=sweep)

    def supports_partitioned(self) -> bool:
        return (self.lib.has("pyro_flamelet_thomas_partitioned")
                and self.nv <= 64)

    def has_sweep_forms(self) -> bool:
        """Whether the partitioned sweep has the blocked MFMA and the
        row-per-lane forms for this block size (16 < nv <= 63)."""
        return (self.lib.has("pyro_flamelet_thomas_sweep_form")
                and 16 < self.nv <= 63)

    def set_thomas(self, kind: str, alpha: Optional[float] = None,
                   part_len: Optional[int] = None, sweep: str = "mfma"):
        """Use the sequential block Thomas or the partitioned one with
        partitions of max(8, round(alpha sqrt(N))) rows (or ``part_len``);
        the partitioned form falls back to the sequential one for nv > 64.
        ``sweep`` ("mfma" or "rowlane") picks the partition sweep where the
        block size has both (a setting of the mechanism's library)."""
        self.partitioned = kind == "partitioned" and self.supports_partitioned()
        self.part_alpha = alpha
        if self.partitioned and self.has_sweep_forms():
            from .autotune import SWEEPS
            self.lib.call("pyro_flamelet_thomas_sweep_form", SWEEPS[sweep])
        if self.partitioned:
            self._setup_partitions(part_len, alpha)

    def _setup_partitions(self, part_len=None, alpha=None):
        tabs = partition_table(self.sizes, part_len, alpha)
        t = lambda a: torch.as_tensor(a, device=self.dev)
        (self.p_start, self.p_end, self.p_sys, self.j_part, self.j_before,
         self.j_after, self.r_off) = (t(a) for a in tabs)
        self.n_parts = len(tabs[0])
        self.n_j = len(tabs[3])
        nv, n2 = self.nv, 2 * self.nv
        dd = dict(dtype=torch.float64, device=self.dev)
        self.fact = torch.empty(self.P * nv * (2 * nv + 1), **dd)
        self.maps = torch.empty(self.n_parts * 2 * nv * (2 * nv + 1), **dd)
        self.red = torch.empty(self.n_parts * nv * (n2 + 2), **dd)
        self.Z = torch.empty(max(self.n_j, 1) * n2, **dd)

    def time_kernels(self, on=True):
        """Record HIP events around every kernel launch (``kernel_ms``)."""
        self.events = {} if on else None

    def kernel_ms(self):
        """Total milliseconds per kernel since ``time_kernels``."""
        torch.cuda.synchronize(self.dev)
        return {k: sum(a.elapsed_time(b) for a, b in v)
                for k, v in (self.events or {}).items()}

    def _call(self, name, *args):
        if self.events is None:
            return self.lib.call(name, *args)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        s = torch.cuda.current_stream(self.dev)
        a.record(s)
        st = self.lib.call(name, *args)
        b.record(s)
        self.events.setdefault(name, []).append((a, b))
        return st

    # -- plumbing --------------------------------------------------------------
    def per_point(self, values) -> torch.Tensor:
        """Expand one value per flamelet to one per point."""
        v = torch.as_tensor(np.asarray(values, dtype=float), dtype=torch.float64)
        v = torch.broadcast_to(v, (self.B,)).to(self.dev)
        return torch.repeat_interleave(v, torch.tensor(self.sizes,
                                                       device=self.dev))

    def set_boundary_enthalpy(self, h_ox, h_fu):
        dd = dict(dtype=torch.float64, device=self.dev)
        self.bc[:, 0, 0] = torch.as_tensor(
            np.broadcast_to(np.asarray(h_ox, dtype=float), (self.B,)).copy(), **dd)
        self.bc[:, 1, 0] = torch.as_tensor(
            np.broadcast_to(np.asarray(h_fu, dtype=float), (self.B,)).copy(), **dd)

    def split(self, t: torch.Tensor) -> List[torch.Tensor]:
        return [t[..., a:b] for a, b in zip(self.offsets_np[:-1],
                                            self.offsets_np[1:])]

    # -- kernels ---------------------------------------------------------------
    def source(self, phi, vd, tguess, p, jac: bool, tol=1e-10, maxit=100):
        phi = _rows(phi, "phi")
        vd, tguess, p = (_rows(x, "input") for x in (vd, tguess, p))
        self.n_fail.zero_()
        if self.dense:
            self._call("pyro_flamelet_source", phi.data_ptr(), phi.stride(0),
                       _ptr(vd), tguess.data_ptr(), 1, p.data_ptr(), 1,
                       self.S.data_ptr(), self.P, self.T.data_ptr(),
                       self.J.data_ptr() if jac else None, self.P, tol, maxit,
                       self.n_fail.data_ptr(),
                       1 if self.semantics == "reference" else 0, _stream())
            return
        # T(h, Y): take the Newton step, stop once |step| < tol, as the
        # one-pass kernel does
        P, ldp = self.P, phi.stride(0)
        y_ptr = phi.data_ptr() + 8 * ldp
        if self._iters is None:
            self._iters = torch.empty(P, dtype=torch.int32, device=self.dev)
        self._call("pyro_temperature", phi.data_ptr(), 1, tguess.data_ptr(),
                   1, y_ptr, ldp, self.T.data_ptr(), P, 0, 0, tol, maxit,
                   self.n_fail.data_ptr(), self._iters.data_ptr(), _stream())
        # as the one-pass kernel does: a point whose temperature did not
        # converge gets T = NaN, so its flamelet's update is not finite
        self.T.masked_fill_(self._iters < 0, float("nan"))
        if jac:
            need = int(self.lib.call("pyro_source_jacobian_work_doubles"))
            if need and (self._jac_work is None
                         or self._jac_work.numel() < need):
                self._jac_work = torch.empty(need, dtype=torch.float64,
                                             device=self.dev)
            nv = self.nv
            self._call("pyro_source_jacobian", self.T.data_ptr(),
                       p.data_ptr(), 1, y_ptr, ldp, _ptr(vd),
                       self.J.data_ptr(), 1, nv * nv, self.S.data_ptr(), P,
                       _ptr(self._jac_work) if need else None, P, _stream())
        else:
            # S alone: rho = p W/(R T), then W_k w_k / rho
            rho = torch.empty(P, dtype=torch.float64, device=self.dev)
            self._call("pyro_mixture_state", 2, p.data_ptr(), 1,
                       self.T.data_ptr(), 1, y_ptr, ldp, rho.data_ptr(), P,
                       _stream())
            w = self.S[1:]
            self._call("pyro_wdot", rho.data_ptr(), self.T.data_ptr(), y_ptr,
                       ldp, w.data_ptr(), self.P, P, _stream())
            w.mul_(self.gas.molecular_weights[:, None] / rho)
            if vd is None:
                self.S[0].zero_()
            else:
                torch.div(vd, rho, out=self.S[0])

    def assemble(self, phi, chi, modes: Optional[torch.Tensor] = None,
                 phi_prev=None, r_prev=None, r_out=None, chi0=None):
        """Residual and Jacobian blocks.  ``modes`` (B,) int32: bit 0 Crank-
        Nicolson with ``phi_prev``, ``r_prev`` and ``self.dt``, bit 1 the
        dissipation rate ``chi0`` in place of ``chi``."""
        phi, chi, phi_prev, r_prev, chi0 = (_rows(x, "input") for x in
                                            (phi, chi, phi_prev, r_prev,
                                             chi0))
        r_out = _rows(r_out, "r_out", writable=True)
        self._call("pyro_flamelet_assemble", phi.data_ptr(), phi.stride(0),
                      self.S.data_ptr(), self.P, self.J.data_ptr(),
                      chi.data_ptr(), _ptr(chi0), self.la.data_ptr(),
                      self.lb.data_ptr(),
                      self.lc.data_ptr(), self.sys.data_ptr(),
                      self.off.data_ptr(), self.bc.data_ptr(), _ptr(modes),
                      _ptr(phi_prev), _ptr(r_prev),
                      self.dt.data_ptr() if modes is not None else None,
                      self.rhs.data_ptr(), self.L.data_ptr(),
                      self.U.data_ptr(), _ptr(r_out), self.P, _stream())

    def thomas(self):
        if self.partitioned:
            self._call("pyro_flamelet_thomas_partitioned", self.L.data_ptr(),
                       self.J.data_ptr(), self.U.data_ptr(),
                       self.rhs.data_ptr(), self.x.data_ptr(),
                       self.off.data_ptr(), self.B, self.p_start.data_ptr(),
                       self.p_end.data_ptr(), self.p_sys.data_ptr(),
                       self.n_parts, self.j_before.data_ptr(),
                       self.j_after.data_ptr(), self.r_off.data_ptr(),
                       self.n_j, self.fact.data_ptr(), self.maps.data_ptr(),
                       self.red.data_ptr(), self.Z.data_ptr(), None,
                       _stream())
            return
        self._call("pyro_flamelet_thomas", self.L.data_ptr(),
                      self.J.data_ptr(), self.U.data_ptr(),
                      self.rhs.data_ptr(), self.x.data_ptr(),
                      self.off.data_ptr(), self.B, self.work.data_ptr(), None,
                      _stream())

    def update(self, phi, active: Optional[torch.Tensor], tg=None):
        """phi += v on the flamelets ``active`` selects and the per-flamelet
        norms of v; with ``tg`` also tg = T on the same points (the next
        temperature guess)."""
        phi = _rows(phi, "phi", writable=True)
        tg = _rows(tg, "tg", writable=True)
        self._call("pyro_flamelet_update", phi.data_ptr(), phi.stride(0),
                      self.x.data_ptr(), self.sys.data_ptr(), _ptr(active),
                      self.T.data_ptr() if tg is not None else None,
                      _ptr(tg), self.off.data_ptr(), self.B,
                      self.norm2.data_ptr(), self.P, _stream())

    def residual(self, phi, chi, vd, tguess, p, out, modes=None,
                 phi_prev=None, r_prev=None, chi0=None):
        """Steady residual R(phi) into ``out`` (nv, P), with the dissipation
        rate chosen per flamelet by bit 1 of ``modes`` as in ``assemble``
        (bit 0 changes only the outputs that are not used here).  The block
        storage is overwritten with values that are not used."""
        self.source(phi, vd, tguess, p, jac=False)
        self.assemble(phi, chi, modes, phi_prev, r_prev, r_out=out,
                      chi0=chi0)
        return out

    def newton_update(self, phi, chi, vd, tguess, p, modes=None,
                      phi_prev=None, r_prev=None, apply=True, active=None,
                      chi0=None, refresh_guess=False):
        """One Newton update v = -J^-1 R for every flamelet (steady, or
        Crank-Nicolson and/or with ``chi0`` per the bits of ``modes``).
        Applies it when ``apply`` (and then, with ``refresh_guess``, sets
        ``tguess`` to the temperature of the updated points); returns the
        per-flamelet sums of squares of v, (B, 2) on the device as
        [enthalpy row, species rows], and v (P, nv)."""
        self.source(phi, vd, tguess, p, jac=True)
        self.assemble(phi, chi, modes, phi_prev, r_prev, chi0=chi0)
        self.thomas()
        self.update(phi, active if apply else self.none_active,
                    tguess if (apply and refresh_guess) else None)
        return self.norm2, self.x
