"""Device ``Thermochemistry``: the generated class's API over the HIP C ABI.

A class made by :func:`make_thermochem_class` is a drop-in for the class that
``PythonCodeGenerator.get_thermochem_class`` returns
(``pyrometheus/codegen/python.py:155-562``): same attributes, same method
names, same positional argument orders and the same species/reaction-leading
output layout, but every argument and result is a float64 torch tensor on a
ROCm device and every method runs one HIP kernel from the mechanism's
generated library.

Layout: per-species arrays are ``(ns, *cells)`` with the cell axes contiguous
(they are flattened to ``(ns, N)`` for the kernel; a leading dimension larger
than ``N`` is accepted through strided views).  Per-cell scalar arguments may
be Python numbers, 0-d tensors or tensors broadcastable to the cell shape.

Kernels run on the current torch stream; nothing synchronises except
``get_temperature``, which reads back the count of non-converged lanes to
raise the reference's ``RuntimeError`` (``codegen/python.py:396``).
"""

from __future__ import annotations

import ctypes
from numbers import Number
from typing import Optional, Tuple

import numpy as np
import torch

from .mechanism import GAS_CONSTANT, ONE_ATM, Mechanism
from .runtime import HipError, HipLibrary, load_mechanism_library

_SPECIES_THERMO = {"cp_r": 0, "h_rt": 1, "s_r": 2, "g_rt": 3, "dcp_r": 4,
                   "dh_rt": 5, "ds_r": 6}
_MIX_THERMO = {"cp": 0, "cv": 1, "h": 2, "e": 3}
_MIX_STATE = {"mmw": 0, "rspec": 1, "density": 2, "pressure": 3}


def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


class ThermochemistryBase:
    """Shared implementation; concrete classes are made per mechanism."""

    _lib: HipLibrary = None
    _mech: Mechanism = None

    def __init__(self, pyro_np=None, device=None):
        # ``pyro_np`` is accepted for signature compatibility with the
        # reference constructor; the array namespace here is always torch.
        mech = self._mech
        self.pyro_np = torch
        self.device = torch.device(device if device is not None else "cuda")
        self.model_name = mech.source
        self.num_elements = mech.n_elements
        self.num_species = mech.n_species
        self.num_reactions = mech.n_reactions
        self.num_falloff = mech.num_falloff
        self.one_atm = ONE_ATM
        self.gas_constant = GAS_CONSTANT
        self.big_number = 1.0e300
        self.species_names = mech.species_names
        self.species_indices = {n: i for i, n in enumerate(self.species_names)}
        self.molecular_weights = torch.tensor(mech.molecular_weights,
                                              dtype=torch.float64,
                                              device=self.device)
        self.inv_molecular_weights = 1 / self.molecular_weights

    # -- argument plumbing ----------------------------------------------------
    def _cells(self, y: torch.Tensor) -> Tuple[torch.Tensor, Tuple[int, ...]]:
        """Return a (ns, N) view/copy with unit cell stride, and the cell shape."""
        if not isinstance(y, torch.Tensor):
            y = torch.as_tensor(np.asarray(y), dtype=torch.float64,
                                device=self.device)
        if y.dtype != torch.float64:
            raise TypeError("mass fractions must be float64")
        if y.device.type != "cuda":
            raise TypeError("arrays must live on the ROCm device")
        if y.shape[0] != self.num_species:
            raise ValueError(f"leading axis must be num_species="
                             f"{self.num_species}, got {tuple(y.shape)}")
        cell_shape = tuple(y.shape[1:])
        y2 = y.reshape(self.num_species, -1)
        if y2.stride(1) != 1 or y2.stride(0) < y2.shape[1]:
            y2 = y2.contiguous()
        return y2, cell_shape

    def _scalar_field(self, x, n: int, cell_shape) -> Tuple[torch.Tensor, int]:
        """Per-cell input -> (tensor, stride) with stride 0 for broadcasts."""
        if isinstance(x, Number):
            return torch.tensor([float(x)], dtype=torch.float64,
                                device=self.device), 0
        if not isinstance(x, torch.Tensor):
            x = torch.as_tensor(np.asarray(x), dtype=torch.float64,
                                device=self.device)
        if x.dtype != torch.float64:
            raise TypeError("per-cell inputs must be float64")
        if x.numel() == 1:
            return x.reshape(1).contiguous(), 0
        x = torch.broadcast_to(x, cell_shape).reshape(-1)
        if x.numel() != n:
            raise ValueError("per-cell input does not match the cell count")
        return x.contiguous(), 1

    def _temperature_field(self, t, n, cell_shape) -> torch.Tensor:
        tt, s = self._scalar_field(t, n, cell_shape)
        if s == 0:
            tt = tt.expand(max(n, 1)).contiguous()
        return tt

    def _out(self, rows: Optional[int], n: int):
        shape = (n,) if rows is None else (rows, n)
        return torch.empty(shape, dtype=torch.float64, device=self.device)

    @staticmethod
    def _shape(out: torch.Tensor, rows: Optional[int], cell_shape):
        if rows is None:
            return out.reshape(cell_shape)
        return out.reshape((rows,) + tuple(cell_shape))

    # -- species / reaction naming --------------------------------------------
    def get_species_name(self, species_index):
        # the reference template reads self.species_name, which it never
        # defines (codegen/python.py:242-243); this returns the intended name
        return self.species_names[species_index]

    def get_species_index(self, species_name):
        return self.species_indices[species_name]

    # -- mixture state (codegen/python.py:248-285) -----------------------------
    def _mixture_state(self, which, a, temperature, mass_fractions):
        y, cs = self._cells(mass_fractions)
        n = y.shape[1]
        out = self._out(None, n)
        a_t, a_s = (None, 0) if a is None else self._scalar_field(a, n, cs)
        t_t, t_s = ((None, 0) if temperature is None
                    else self._scalar_field(temperature, n, cs))
        self._lib.call("pyro_mixture_state", _MIX_STATE[which], _ptr(a_t), a_s,
                       _ptr(t_t), t_s, y.data_ptr(), y.stride(0),
                       out.data_ptr(), n, _stream())
        return self._shape(out, None, cs)

    def get_specific_gas_constant(self, mass_fractions):
        return self._mixture_state("rspec", None, None, mass_fractions)

    def get_density(self, p, temperature, mass_fractions):
        return self._mixture_state("density", p, temperature, mass_fractions)

    def get_pressure(self, rho, temperature, mass_fractions):
        return self._mixture_state("pressure", rho, temperature, mass_fractions)

    def get_mixture_molecular_weight(self, mass_fractions):
        return self._mixture_state("mmw", None, None, mass_fractions)

    def _species_state(self, which, a, mass_fractions):
        y, cs = self._cells(mass_fractions)
        n = y.shape[1]
        out = self._out(self.num_species, n)
        a_t, a_s = self._scalar_field(a, n, cs)
        self._lib.call("pyro_species_state", which, a_t.data_ptr(), a_s,
                       y.data_ptr(), y.stride(0), out.data_ptr(), n, n,
                       _stream())
        return self._shape(out, self.num_species, cs)

    def get_concentrations(self, rho, mass_fractions):
        return self._species_state(0, rho, mass_fractions)

    def get_mole_fractions(self, mix_mol_weight, mass_fractions):
        return self._species_state(1, mix_mol_weight, mass_fractions)

    def get_mass_average_property(self, mass_fractions, spec_property):
        y, cs = self._cells(mass_fractions)
        p, _ = self._cells(spec_property)
        w = self.inv_molecular_weights.reshape(-1, 1)
        return (y * p * w).sum(dim=0).reshape(cs)

    # -- mixture thermo (codegen/python.py:293-311) ----------------------------
    def _mixture_thermo(self, which, temperature, mass_fractions):
        y, cs = self._cells(mass_fractions)
        n = y.shape[1]
        t, ts = self._scalar_field(temperature, n, cs)
        out = self._out(None, n)
        self._lib.call("pyro_mixture_thermo", _MIX_THERMO[which], t.data_ptr(),
                       ts, y.data_ptr(), y.stride(0), out.data_ptr(), n,
                       _stream())
        return self._shape(out, None, cs)

    def get_mixture_specific_heat_cp_mass(self, temperature, mass_fractions):
        return self._mixture_thermo("cp", temperature, mass_fractions)

    def get_mixture_specific_heat_cv_mass(self, temperature, mass_fractions):
        return self._mixture_thermo("cv", temperature, mass_fractions)

    def get_mixture_enthalpy_mass(self, temperature, mass_fractions):
        return self._mixture_thermo("h", temperature, mass_fractions)

    def get_mixture_internal_energy_mass(self, temperature, mass_fractions):
        return self._mixture_thermo("e", temperature, mass_fractions)

    # -- species thermo (codegen/python.py:313-358) ----------------------------
    def _species_thermo(self, which, temperature):
        if isinstance(temperature, Number):
            temperature = torch.tensor(float(temperature), dtype=torch.float64,
                                       device=self.device)
        t = temperature.reshape(-1).contiguous()
        n = t.numel()
        out = self._out(self.num_species, n)
        self._lib.call("pyro_species_thermo", _SPECIES_THERMO[which],
                       t.data_ptr(), 1, out.data_ptr(), n, n, _stream())
        return out.reshape((self.num_species,) + tuple(temperature.shape))

    def get_species_specific_heats_r(self, temperature):
        return self._species_thermo("cp_r", temperature)

    def get_species_enthalpies_rt(self, temperature):
        return self._species_thermo("h_rt", temperature)

    def get_species_entropies_r(self, temperature):
        return self._species_thermo("s_r", temperature)

    def get_species_gibbs_rt(self, temperature):
        return self._species_thermo("g_rt", temperature)

    def get_species_specific_heats_r_derivative(self, temperature):
        return self._species_thermo("dcp_r", temperature)

    def get_species_enthalpies_rt_derivative(self, temperature):
        return self._species_thermo("dh_rt", temperature)

    def get_species_entropies_rt_derivative(self, temperature):
        return self._species_thermo("ds_r", temperature)

    # -- reaction quantities (codegen/python.py:360-515) -----------------------
    def _reaction_quantities(self, temperature, concentrations, which):
        c, cs = self._cells(concentrations)
        n = c.shape[1]
        t = self._temperature_field(temperature, n, cs)
        outs = {k: self._out(self.num_reactions, n) for k in which}
        self._lib.call("pyro_reaction_quantities", t.data_ptr(), c.data_ptr(),
                       c.stride(0), _ptr(outs.get("kf")), _ptr(outs.get("nlk")),
                       _ptr(outs.get("rf")), _ptr(outs.get("rr")),
                       _ptr(outs.get("net")), n, n, _stream())
        return [self._shape(outs[k], self.num_reactions, cs) for k in which]

    def get_equilibrium_constants(self, temperature):
        """Returns -ln K_c per reaction (codegen/python.py:360-373)."""
        if isinstance(temperature, Number):
            temperature = torch.tensor(float(temperature), dtype=torch.float64,
                                       device=self.device)
        t = temperature.reshape(-1).contiguous()
        ones = torch.ones((self.num_species, t.numel()), dtype=torch.float64,
                          device=self.device)
        (nlk,) = self._reaction_quantities(t, ones, ["nlk"])
        return nlk.reshape((self.num_reactions,) + tuple(temperature.shape))

    def get_fwd_rate_coefficients(self, temperature, concentrations):
        return self._reaction_quantities(temperature, concentrations, ["kf"])[0]

    def get_net_rates_of_progress(self, temperature, concentrations):
        return self._reaction_quantities(temperature, concentrations, ["net"])[0]

    def get_fwd_rates_of_progress(self, temperature, concentrations):
        return self._reaction_quantities(temperature, concentrations, ["rf"])[0]

    def get_rev_rates_of_progress(self, temperature, concentrations):
        return self._reaction_quantities(temperature, concentrations, ["rr"])[0]

    # -- production rates (codegen/python.py:486-562) --------------------------
    def get_net_production_rates(self, rho, temperature, mass_fractions,
                                 out: Optional[torch.Tensor] = None):
        y, cs = self._cells(mass_fractions)
        n = y.shape[1]
        r = self._temperature_field(rho, n, cs)
        t = self._temperature_field(temperature, n, cs)
        if out is None:
            out = self._out(self.num_species, n)
            ldw = n
        else:
            ldw = out.stride(0)
        self._lib.call("pyro_wdot", r.data_ptr(), t.data_ptr(), y.data_ptr(),
                       y.stride(0), out.data_ptr(), ldw, n, _stream())
        return self._shape(out, self.num_species, cs) if ldw == n else out

    def get_creation_destruction_rates(self, rho, temperature, mass_fractions):
        y, cs = self._cells(mass_fractions)
        n = y.shape[1]
        r = self._temperature_field(rho, n, cs)
        t = self._temperature_field(temperature, n, cs)
        cd = self._out(self.num_species, n)
        dd = self._out(self.num_species, n)
        self._lib.call("pyro_creation_destruction", r.data_ptr(), t.data_ptr(),
                       y.data_ptr(), y.stride(0), cd.data_ptr(), dd.data_ptr(),
                       n, n, _stream())
        return (self._shape(cd, self.num_species, cs),
                self._shape(dd, self.num_species, cs))

    def get_creation_rates(self, rho, temperature, mass_fractions):
        return self.get_creation_destruction_rates(rho, temperature,
                                                   mass_fractions)[0]

    def get_destruction_rates(self, rho, temperature, mass_fractions):
        return self.get_creation_destruction_rates(rho, temperature,
                                                   mass_fractions)[1]

    # -- transport (codegen/python.py:564-641) ---------------------------------
    def _transport_lib(self):
        if not self._lib.has("pyro_species_transport"):
            raise HipError(f"mechanism {self._mech.name!r} has no transport "
                           "data (its library exports no transport kernels)")
        return self._lib

    def _species_transport(self, which, temperature, rows):
        lib = self._transport_lib()
        if isinstance(temperature, Number):
            temperature = torch.tensor(float(temperature), dtype=torch.float64,
                                       device=self.device)
        t = temperature.reshape(-1).contiguous()
        n = t.numel()
        out = torch.empty(rows + (n,), dtype=torch.float64, device=self.device)
        lib.call("pyro_species_transport", which, t.data_ptr(), 1,
                 out.data_ptr(), n, n, _stream())
        return out.reshape(rows + tuple(temperature.shape))

    def get_species_viscosities(self, temperature):
        """Species viscosities [Pa s], (ns, *cells)."""
        return self._species_transport(0, temperature, (self.num_species,))

    def get_species_thermal_conductivities(self, temperature):
        """Species thermal conductivities [W/(m K)], (ns, *cells)."""
        return self._species_transport(1, temperature, (self.num_species,))

    def get_species_binary_mass_diffusivities(self, temperature):
        """Binary diffusion coefficients times pressure [m^2/s Pa],
        (ns, ns, *cells)."""
        ns = self.num_species
        return self._species_transport(2, temperature, (ns, ns))

    def _mixture_transport(self, which, pressure, temperature,
                           mass_fractions):
        lib = self._transport_lib()
        y, cs = self._cells(mass_fractions)
        n = y.shape[1]
        t, ts = self._scalar_field(temperature, n, cs)
        p, ps = ((None, 0) if pressure is None
                 else self._scalar_field(pressure, n, cs))
        rows = self.num_species if which == 2 else None
        out = self._out(rows, n)
        lib.call("pyro_mixture_transport", which, _ptr(p), ps, t.data_ptr(),
                 ts, y.data_ptr(), y.stride(0), out.data_ptr(), n, n,
                 _stream())
        return self._shape(out, rows, cs)

    def get_mixture_viscosity_mixavg(self, temperature, mass_fractions):
        """Mixture viscosity by Wilke's rule [Pa s]."""
        return self._mixture_transport(0, None, temperature, mass_fractions)

    def get_mixture_thermal_conductivity_mixavg(self, temperature,
                                                mass_fractions):
        """Mixture conductivity, the mean of the arithmetic and harmonic
        mole-fraction averages [W/(m K)]."""
        return self._mixture_transport(1, None, temperature, mass_fractions)

    def get_species_mass_diffusivities_mixavg(self, pressure, temperature,
                                              mass_fractions):
        """Mixture-averaged species diffusivities [m^2/s], (ns, *cells)."""
        return self._mixture_transport(2, pressure, temperature,
                                       mass_fractions)

    def get_transport_fits(self):
        """The fit coefficients the library was built with: viscosity
        (ns, 5), conductivity (ns, 5), binary diffusivity (ns, ns, 5), in
        ascending powers of ln T (pyrometheus_amd.transport)."""
        lib = self._transport_lib()
        ns = self.num_species
        out = []
        for which, shape in ((0, (ns, 5)), (1, (ns, 5)), (2, (ns, ns, 5))):
            a = np.zeros(shape)
            lib.call("pyro_transport_fits", which,
                     a.ctypes.data_as(ctypes.c_void_p), a.size)
            out.append(a)
        return tuple(out)

    # -- temperature (codegen/python.py:375-396; make_pyro.py:255-299) ---------
    def _temperature(self, target, t_guess, y, do_energy, mode, tol, maxit):
        yy, cs = self._cells(y)
        n = yy.shape[1]
        e, es = self._scalar_field(target, n, cs)
        g, gs = self._scalar_field(t_guess, n, cs)
        out = self._out(None, n)
        fail = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._lib.call("pyro_temperature", e.data_ptr(), es, g.data_ptr(), gs,
                       yy.data_ptr(), yy.stride(0), out.data_ptr(), n,
                       int(do_energy), mode, float(tol), int(maxit),
                       fail.data_ptr(), None, _stream())
        return self._shape(out, None, cs), fail

    def get_temperature(self, enthalpy_or_energy, t_guess, y, do_energy=False):
        """Per-lane Newton; raises like the reference when a lane has not
        converged after 500 iterations.  As in the reference, only the literal
        ``False`` selects the enthalpy form (``do_energy is False``)."""
        energy = do_energy is not False
        t, fail = self._temperature(enthalpy_or_energy, t_guess, y, energy,
                                    0, 1.0e-6, 500)
        if int(fail.item()) > 0:
            raise RuntimeError("Temperature iteration failed to converge")
        return t

    def get_temperature_jump_bands(self, mass_fractions, do_energy=True):
        """Where ``get_temperature`` has no unique answer.

        The NASA-7 fits of a mechanism need not make h(T) continuous at a
        species' T_mid (the high set applies for ``T > T_mid``,
        ``chem_expr.py:51-57``).  For GRI-3.0 every mixture's e(T) drops by
        2e-5..5e-4 K c_v at T_mid = 1000 K, so targets there have two roots
        (one per side) and the Newton iteration, here as in the reference,
        returns whichever its guess leads to; where e(T) rises (GRI's HCNO at
        1382 K) targets inside the jump have no root and the iteration does
        not converge (``RuntimeError``, as in ``codegen/python.py:396``).

        Returns ``(t_mid, jump, half_width)``: the distinct T_mid values (G,),
        the jump e(T_mid+) - e(T_mid) (or of h) per cell in J/kg (G, *cells),
        and |jump| / c_v(T_mid) (c_p for h) in K: a target from a temperature
        within ``half_width`` of ``t_mid`` may have another root or none."""
        y, cs = self._cells(mass_fractions)
        mech = self._mech
        tm = np.array([s.t_mid for s in mech.species])
        groups = np.unique(tm)

        def h_rt(c, t):
            return (c[0] + c[1] * t / 2 + c[2] * t * t / 3 + c[3] * t ** 3 / 4
                    + c[4] * t ** 4 / 5 + c[5] / t)
        dmat = np.zeros((len(groups), mech.n_species))
        for k, s in enumerate(mech.species):
            g = int(np.searchsorted(groups, s.t_mid))
            dmat[g, k] = (GAS_CONSTANT * s.t_mid
                          * (h_rt(s.high, s.t_mid) - h_rt(s.low, s.t_mid))
                          / mech.molecular_weights[k])
        d = torch.tensor(dmat, dtype=torch.float64, device=self.device)
        jump = d @ y
        width = torch.empty_like(jump)
        which = "cv" if do_energy is not False else "cp"
        for g, t in enumerate(groups):
            pv = self._mixture_thermo(which, float(t), y)
            width[g] = jump[g].abs() / pv
        t_mid = torch.tensor(groups, dtype=torch.float64, device=self.device)
        return (t_mid, jump.reshape((len(groups),) + tuple(cs)),
                width.reshape((len(groups),) + tuple(cs)))

    def get_temperature_from_enthalpy(self, enthalpy, mass_fractions,
                                      temp_init, tol=1.0e-10, maxiter=500):
        """Flamelet hook (``flamelets/make_pyro.py:255-299``): Newton on
        ``h(T, Y) = h*`` stopping per lane once the step is ``<= tol``."""
        t, fail = self._temperature(enthalpy, temp_init, mass_fractions, False,
                                    1, tol, maxiter)
        if int(fail.item()) > 0:
            raise RuntimeError("Temperature iteration failed to converge")
        return t


def make_thermochem_class(name: str, lib: HipLibrary, mech: Mechanism):
    return type(name, (ThermochemistryBase,), {"_lib": lib, "_mech": mech})


def get_thermochem_class(mech: Mechanism):
    """Class bound to ``mech``'s HIP library (built on first use if missing).
    ``mech`` may also be a Cantera YAML path or a ``ct.Solution``."""
    from .mechanism import as_mechanism
    mech = as_mechanism(mech)
    return make_thermochem_class("Thermochemistry",
                                 load_mechanism_library(mech), mech)
