"""Request server for chemistry on one GPU: continuous batching of
pointwise production-rate requests and steady-flamelet solves.

Requests of two kinds arrive over time, for any of the loaded mechanisms:

* ``WdotRequest`` -- net production rates of a device-resident state
  (ρ, T, Y) of any number of cells (``get_net_production_rates``);
* ``FlameletRequest`` -- one steady flamelet solve (the reference's
  ``FlameletSolver.solve``, ``flamelets/solver.py:476-600``).

The server keeps every state on the device and admits queued requests in
arrival order as long as the bytes they will hold fit in an HBM budget (and,
for flamelets, a slot limit).  Admitted production-rate requests are
launched at once on their own stream; admitted flamelets join the running
flamelet batch of their mechanism at the next iteration
(``FlameletStepper``), where each iteration advances every live flamelet by
one Newton update on a second stream, so the two kinds of work overlap on
the device.  Finished requests release their budget at once, which is what
lets newly arrived requests in.

Latency of a request is measured from its arrival time to the moment its
result is complete on the device.  ``serve_plain`` runs the same requests
one at a time with the plain drop-in (one kernel per production-rate
request; one ``FlameletSolver`` per flamelet) for comparison.
"""

from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .flamelets.stepper import FlameletStepper, SolveSettings


@dataclass
class WdotRequest:
    mechanism: str
    rho: torch.Tensor
    temperature: torch.Tensor
    mass_fractions: torch.Tensor

    @property
    def size(self) -> int:
        return int(self.temperature.numel())


@dataclass
class FlameletRequest:
    mechanism: str
    problem: object          # FlameletProblem
    domain: object = None    # Domain, for the plain drop-in

    @property
    def size(self) -> int:
        return len(self.problem.x)


@dataclass
class Ticket:
    id: int
    kind: str
    mechanism: str
    size: int
    t_arrival: float
    nbytes: int
    t_admit: Optional[float] = None
    t_done: Optional[float] = None
    result: object = None
    _event: object = None
    _request: object = None
    _done_event: object = None

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival

    @property
    def event(self):
        """Event recorded on the server's stream after the result was
        written: wait on it (``stream.wait_event``) before reading the
        result on another stream.  The ticket is finished once it has
        completed."""
        return self._done_event


class ChemistryServer:
    """Continuous-batching server; see the module docstring."""

    def __init__(self, gases: Dict[str, object], hbm_budget: int = 32 << 30,
                 max_flamelets: int = 1024,
                 settings: Optional[SolveSettings] = None, device=None,
                 sync_every: int = 4):
        self.gases = dict(gases)
        any_gas = next(iter(self.gases.values()))
        self.dev = torch.device(device) if device is not None \
            else any_gas.device
        self.budget = int(hbm_budget)
        self.max_flamelets = int(max_flamelets)
        # flamelet iterations queued per host sync of a batch (control flow
        # on the device in between; admissions happen at syncs)
        self.sync_every = int(sync_every)
        self.settings = settings or SolveSettings(equilibrate_steps=24,
                                                  newton_norm="scaled")
        # both streams at the default priority
        self.stream_w = torch.cuda.Stream(self.dev)
        # one stream per mechanism's flamelet batch (default priority), so
        # the batches of different mechanisms overlap on the device
        self.streams_f: Dict[str, torch.cuda.Stream] = {}
        self.steppers: Dict[str, FlameletStepper] = {}
        self.queue: collections.deque = collections.deque()
        self.inflight: List[Ticket] = []
        self.flamelets: Dict[Tuple[str, int], Ticket] = {}
        self.used = 0
        self.peak_used = 0
        self.done: List[Ticket] = []
        self._ids = 0
        self.t0 = time.perf_counter()

    # -- costs --------------------------------------------------------------
    def cost(self, req) -> int:
        """Device bytes a request holds while admitted."""
        gas = self.gases[req.mechanism]
        ns = gas.num_species
        if isinstance(req, WdotRequest):
            return 8 * ns * req.size  # the output; the inputs are resident
        nv = ns + 1
        return FlameletStepper.bytes_per_point(nv) * req.size

    def now(self) -> float:
        return time.perf_counter() - self.t0

    # -- interface ----------------------------------------------------------
    def submit(self, req, t_arrival: Optional[float] = None) -> Ticket:
        if req.mechanism not in self.gases:
            raise KeyError(f"mechanism {req.mechanism!r} is not loaded")
        if isinstance(req, FlameletRequest) and not self.gases[
                req.mechanism]._lib.has("pyro_flamelet_assemble"):
            raise ValueError(f"{req.mechanism}: library has no flamelet "
                             "kernels")
        kind = "wdot" if isinstance(req, WdotRequest) else "flamelet"
        nbytes = self.cost(req)
        if nbytes > self.budget:
            raise ValueError(f"request needs {nbytes} bytes, more than the "
                             f"budget of {self.budget}")
        t = Ticket(self._ids, kind, req.mechanism, req.size,
                   self.now() if t_arrival is None else t_arrival, nbytes,
                   _request=req)
        self._ids += 1
        if kind == "wdot":
            # the state was produced on the caller's stream
            self.stream_w.wait_stream(torch.cuda.current_stream(self.dev))
        self.queue.append(t)
        return t

    @property
    def busy(self) -> bool:
        # inflight: production-rate requests and finished flamelets whose
        # results are still being written on the device
        return bool(self.queue or self.inflight or self.flamelets)

    def _admit(self):
        while self.queue:
            t = self.queue[0]
            if self.used + t.nbytes > self.budget:
                break
            if t.kind == "flamelet" and len(self.flamelets) >= \
                    self.max_flamelets:
                break
            self.queue.popleft()
            self.used += t.nbytes
            self.peak_used = max(self.peak_used, self.used)
            t.t_admit = self.now()
            req = t._request
            if t.kind == "wdot":
                gas = self.gases[t.mechanism]
                with torch.cuda.stream(self.stream_w):
                    t.result = gas.get_net_production_rates(
                        req.rho, req.temperature, req.mass_fractions)
                    t._event = torch.cuda.Event()
                    t._event.record(self.stream_w)
                t._done_event = t._event
                self.inflight.append(t)
            else:
                st = self.steppers.get(t.mechanism)
                if st is None:
                    st = FlameletStepper(self.gases[t.mechanism],
                                         self.settings, device=self.dev,
                                         sync_every=self.sync_every)
                    self.steppers[t.mechanism] = st
                with torch.cuda.stream(self._stream_f(t.mechanism)):
                    (h,) = st.add([req.problem])
                self.flamelets[(t.mechanism, h)] = t

    def _finish(self, t: Ticket):
        t.t_done = self.now()
        self.used -= t.nbytes
        t._request = None
        self.done.append(t)

    def poll(self):
        keep = []
        for t in self.inflight:
            if t._event.query():
                t._event = None
                self._finish(t)
            else:
                keep.append(t)
        self.inflight = keep

    def _stream_f(self, mech: str) -> torch.cuda.Stream:
        st = self.streams_f.get(mech)
        if st is None:
            st = self.streams_f[mech] = torch.cuda.Stream(self.dev)
        return st

    def step(self):
        """Admit what fits, advance every flamelet batch by one iteration,
        retire finished work.  Every batch's iteration is launched (on its
        mechanism's stream) before any is waited for, so batches of
        different mechanisms run at the same time."""
        self._admit()
        launched = []
        for mech, st in self.steppers.items():
            if not st.n_live:
                continue
            with torch.cuda.stream(self._stream_f(mech)):
                if st.launch():
                    launched.append((mech, st))
        self.poll()
        for mech, st in launched:
            stream = self._stream_f(mech)
            with torch.cuda.stream(stream):
                finished = st.complete()
            for h, res in finished:
                t = self.flamelets.pop((mech, h))
                t.result = res
                # the result's final source evaluation and copies are still
                # queued on the mechanism's stream: the ticket finishes when
                # this event has completed
                t._event = torch.cuda.Event()
                t._event.record(stream)
                t._done_event = t._event
                self.inflight.append(t)
            self.poll()
        self.poll()

    def run(self, arrivals: Sequence[Tuple[float, object]],
            idle_sleep: float = 2e-4) -> List[Ticket]:
        """Serve ``(t_arrival, request)`` pairs (seconds from now, sorted),
        submitting each when its time comes; returns the tickets in arrival
        order once all are done."""
        self.t0 = time.perf_counter()
        tickets = []
        i = 0
        while i < len(arrivals) or self.busy:
            now = self.now()
            while i < len(arrivals) and arrivals[i][0] <= now:
                tickets.append(self.submit(arrivals[i][1], arrivals[i][0]))
                i += 1
            if self.busy:
                self.step()
            elif i < len(arrivals):
                time.sleep(min(idle_sleep, max(0.0, arrivals[i][0] - now)))
        torch.cuda.synchronize(self.dev)
        return tickets


def serve_plain(gases: Dict[str, object],
                arrivals: Sequence[Tuple[float, object]],
                settings: Optional[SolveSettings] = None):
    """The plain drop-in, one request at a time in arrival order (a FIFO
    server): every request is run alone and its service time measured; the
    latency of request i is then max(arrival_i, end_{i-1}) + service_i -
    arrival_i, exactly what a single plain server would give.  Returns
    ``(latencies, service_times, results)``."""
    from .flamelets.workload import solve_plain
    s = settings or SolveSettings(equilibrate_steps=24, newton_norm="scaled")
    service, results = [], []
    for _, req in arrivals:
        gas = gases[req.mechanism]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if isinstance(req, WdotRequest):
            out = gas.get_net_production_rates(req.rho, req.temperature,
                                               req.mass_fractions)
        else:
            kw = dict(newton_maxiter=s.newton_maxiter,
                      newton_tol=s.newton_tol,
                      bdf_newton_maxiter=s.bdf_newton_maxiter,
                      bdf_newton_tol=s.bdf_newton_tol,
                      bdf_time_step=s.bdf_time_step,
                      bdf_maxsteps=s.bdf_maxsteps, try_newton=s.try_newton,
                      max_attempts=s.max_attempts)
            out = solve_plain(gas, req.domain, req.problem,
                              equilibrate_steps=s.equilibrate_steps,
                              newton_norm=s.newton_norm, **kw)
        torch.cuda.synchronize()
        service.append(time.perf_counter() - t0)
        results.append(out)
    lat = fifo_latencies([ta for ta, _ in arrivals], service)
    return lat, np.array(service), results


def fifo_latencies(arrival_times, service_times) -> np.ndarray:
    """Latencies of a single first-in first-out server: request i starts at
    max(arrival_i, end_{i-1}) and ends service_i later."""
    lat, end = [], -np.inf
    for ta, sv in zip(arrival_times, service_times):
        start = max(ta, end)
        end = start + sv
        lat.append(end - ta)
    return np.array(lat, dtype=float)


def poisson_workload(gases: Dict[str, object], n_requests: int, rate: float,
                     seed: int = 5, wdot_fraction: float = 0.5,
                     wdot_log2_cells=(10, 20), flamelet_points=(64, 2048),
                     wdot_mechanisms: Optional[Sequence[str]] = None,
                     flamelet_mechanisms: Optional[Sequence[str]] = None):
    """``n_requests`` arrivals of a Poisson process of ``rate`` per second:
    production-rate requests of 2^U{lo..hi} cells (state drawn on the device:
    T ~ U[300, 2500] K, p log-uniform in [0.5, 20] atm, Y = 0.5
    Dirichlet(0.5) + 0.5 e_N2) and flamelet solves of U{lo..hi} points
    (Z-clustered grid, chi_st log-uniform in [1, 1000] 1/s, fuel H2:N2 at
    300 K, air at 800 K, 1 atm), mechanisms drawn uniformly from the lists.
    Returns a list of ``(t_arrival, request)``."""
    from .flamelets.workload import make_problem, streams
    from .mechanism import ONE_ATM
    rng = np.random.default_rng(seed)
    names = list(gases)
    wm = list(wdot_mechanisms or names)
    fm = list(flamelet_mechanisms or [n for n in names if gases[n]._lib.has(
        "pyro_flamelet_assemble")])
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_requests))
    t -= t[0]
    out = []
    st_cache = {}
    for k in range(n_requests):
        if rng.uniform() < wdot_fraction or not fm:
            name = wm[rng.integers(len(wm))]
            gas = gases[name]
            n = 1 << int(rng.integers(wdot_log2_cells[0],
                                      wdot_log2_cells[1] + 1))
            g = torch.Generator(device=gas.device).manual_seed(
                int(rng.integers(1 << 31)))
            T = 300.0 + 2200.0 * torch.rand(n, generator=g,
                                            dtype=torch.float64,
                                            device=gas.device)
            p = ONE_ATM * torch.exp(np.log(0.5) + np.log(40.0) * torch.rand(
                n, generator=g, dtype=torch.float64, device=gas.device))
            Y = torch._standard_gamma(torch.full(
                (gas.num_species, n), 0.5, dtype=torch.float64,
                device=gas.device), generator=g)
            Y = 0.5 * Y / Y.sum(0, keepdim=True)
            inert = "N2" if "N2" in gas.species_names else gas.species_names[-1]
            Y[gas.species_indices[inert]] += 0.5
            rho = gas.get_density(p, T, Y)
            out.append((float(t[k]), WdotRequest(name, rho, T, Y)))
        else:
            name = fm[rng.integers(len(fm))]
            gas = gases[name]
            if name not in st_cache:
                st_cache[name] = streams(gas)
            n = int(rng.integers(flamelet_points[0], flamelet_points[1] + 1))
            chi = float(np.exp(rng.uniform(0.0, np.log(1000.0))))
            dom, prob = make_problem(gas, st_cache[name], n, chi)
            out.append((float(t[k]), FlameletRequest(name, prob, dom)))
    torch.cuda.synchronize()
    return out


def summarize(latencies, t_first: float, t_last: float) -> dict:
    lat = np.asarray(latencies, dtype=float)
    span = max(t_last - t_first, 1e-12)
    return {"n": int(lat.size), "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "requests_per_s": lat.size / span}
