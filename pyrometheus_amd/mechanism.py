"""Mechanism intermediate representation (IR) and Cantera-YAML reader.

The reference obtains every mechanism constant from a Cantera ``Solution``
(``pyrometheus/cli.py:38``, ``test/backends.py:224``) and reads it through
``sol.species()``, ``sol.reactions()``, ``sol.reactant_stoich_coeff`` and
friends inside ``pyrometheus/chem_expr.py``.  Cantera is not available to this
package, so this module re-implements the part of Cantera that the generated
code depends on:

* YAML parsing of the ``phases`` / ``species`` / ``reactions`` sections, with
  the phase's species list deciding which species exist (so
  ``sandiego.yaml`` yields the 9 species its ``gas`` phase declares);
* Cantera's unit conversion to SI with kmol as quantity: the pre-exponential
  factor scales by ``(L^3/Q)^(order-1)`` where the order counts reactant
  stoichiometric coefficients (or explicit ``orders``), plus one for a
  mass-action third body, plus one for a falloff low-pressure rate;
  activation energies go to J/kmol;
* molecular weights from Cantera's atomic-weight table, summed in the phase's
  element order;
* reaction classification into the strings the reference tests
  (``reaction_type`` ``"Arrhenius"``, ``"three-body-Arrhenius"``,
  ``"falloff-Troe"``, ``"falloff-Lindemann"``; ``codegen/python.py:668-671``);
* Troe parameters as Cantera returns them from ``falloff_coeffs``
  (``T3`` and ``T1`` round-tripped through their reciprocals, ``T2`` dropped
  when zero), which ``chem_expr.troe_falloff_center_expr`` (``:348-373``)
  consumes.

Everything is immutable after construction and hashable, so a mechanism can
key a cache of generated kernels.
"""

from __future__ import annotations

import hashlib
import json
import math
import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import yaml

#: Universal gas constant in J/(kmol K), the value the generated code embeds
#: as ``gas_constant`` (``codegen/python.py:190``).
GAS_CONSTANT = 8314.46261815324
#: One standard atmosphere in Pa (``codegen/python.py:189``).
ONE_ATM = 101325.0

#: Atomic weights (kg/kmol) as tabulated by Cantera 3.x (IUPAC abridged).
ATOMIC_WEIGHTS = {
    "H": 1.008, "D": 2.014102, "Tr": 3.016049, "He": 4.002602,
    "Li": 6.94, "Be": 9.0121831, "B": 10.81, "C": 12.011, "N": 14.007,
    "O": 15.999, "F": 18.998403163, "Ne": 20.1797, "Na": 22.98976928,
    "Mg": 24.305, "Al": 26.9815384, "Si": 28.085, "P": 30.973761998,
    "S": 32.06, "Cl": 35.45, "Ar": 39.95, "K": 39.0983, "Ca": 40.078,
    "Fe": 55.845, "Kr": 83.798, "Xe": 131.293, "E": 5.48579909065e-4,
}

_QUANTITY_KMOL = {"kmol": 1.0, "mol": 1e-3, "molec": 1.0 / 6.02214076e26}
_LENGTH_M = {"m": 1.0, "cm": 1e-2, "mm": 1e-3}
_TIME_S = {"s": 1.0, "ms": 1e-3, "min": 60.0, "hr": 3600.0}
#: Activation-energy units to J/kmol; "K" is handled separately (multiply by R).
_ENERGY_J_PER_KMOL = {
    "J/kmol": 1.0, "J/mol": 1e3, "kJ/mol": 1e6, "cal/mol": 4184.0,
    "kcal/mol": 4.184e6, "eV": 9.648533212331001e7,
}


@dataclass(frozen=True)
class Species:
    """One species with its NASA-7 polynomial.

    ``low`` and ``high`` are the 7 NASA coefficients ``a0..a6`` of the two
    temperature ranges; ``t_mid`` splits them.  The reference selects the high
    set with the strict test ``T > t_mid`` (``chem_expr.py:51-57``).
    """

    name: str
    composition: Tuple[Tuple[str, float], ...]
    t_min: float
    t_mid: float
    t_max: float
    low: Tuple[float, ...]
    high: Tuple[float, ...]
    weight: float
    #: gas transport data of the YAML ``transport`` entry: (geometry,
    #: diameter [Angstrom], well-depth [K], dipole [Debye], polarizability
    #: [Angstrom^3], rotational-relaxation) or None
    transport: Optional[Tuple] = None

    def n_atoms(self, element: str) -> float:
        return dict(self.composition).get(element, 0.0)


@dataclass(frozen=True)
class Arrhenius:
    """Modified Arrhenius rate in SI/kmol units: ``A T^b exp(-Ea/(R T))``."""

    A: float
    b: float
    Ea: float  # J/kmol

    @property
    def t_a(self) -> float:
        """Activation temperature ``Ea/R`` (``chem_expr.py:317``)."""
        return self.Ea / GAS_CONSTANT


@dataclass(frozen=True)
class Reaction:
    """One reaction.  Stoichiometric dicts are stored sorted by species name,
    the iteration order of Cantera's ``Composition`` map."""

    equation: str
    reactants: Tuple[Tuple[str, float], ...]
    products: Tuple[Tuple[str, float], ...]
    reversible: bool
    kind: str  # "elementary" | "three-body" | "falloff"
    rate: Optional[Arrhenius] = None  # elementary and three-body
    low_rate: Optional[Arrhenius] = None  # falloff
    high_rate: Optional[Arrhenius] = None  # falloff
    falloff_type: Optional[str] = None  # "Troe" | "Lindemann"
    troe: Tuple[float, ...] = ()  # Cantera falloff_coeffs: A, T3, T1[, T2]
    efficiencies: Tuple[Tuple[str, float], ...] = ()
    default_efficiency: float = 1.0
    orders: Tuple[Tuple[str, float], ...] = ()
    duplicate: bool = False

    @property
    def reaction_type(self) -> str:
        if self.kind == "three-body":
            return "three-body-Arrhenius"
        if self.kind == "falloff":
            return f"falloff-{self.falloff_type}"
        return "Arrhenius"

    @property
    def is_falloff(self) -> bool:
        return self.kind == "falloff"

    @property
    def is_three_body(self) -> bool:
        return self.kind == "three-body"


@dataclass(frozen=True)
class Mechanism:
    """Immutable mechanism: the subset of a Cantera ``Solution`` the generated
    ``Thermochemistry`` class reads."""

    name: str
    source: str
    elements: Tuple[str, ...]
    species: Tuple[Species, ...]
    reactions: Tuple[Reaction, ...]
    _index: Dict[str, int] = field(default=None, compare=False, hash=False,
                                   repr=False)

    def __post_init__(self):
        object.__setattr__(self, "_index",
                           {s.name: i for i, s in enumerate(self.species)})

    # --- Cantera-like accessors used by the generators and the oracle ---
    @property
    def n_species(self) -> int:
        return len(self.species)

    @property
    def n_reactions(self) -> int:
        return len(self.reactions)

    @property
    def n_elements(self) -> int:
        return len(self.elements)

    @property
    def species_names(self) -> List[str]:
        return [s.name for s in self.species]

    def species_index(self, name: str) -> int:
        return self._index[name]

    @property
    def molecular_weights(self) -> np.ndarray:
        return np.array([s.weight for s in self.species])

    def reactant_stoich_coeff(self, k: int, i: int) -> float:
        return dict(self.reactions[i].reactants).get(self.species[k].name, 0.0)

    def product_stoich_coeff(self, k: int, i: int) -> float:
        return dict(self.reactions[i].products).get(self.species[k].name, 0.0)

    def n_atoms(self, k: int, e: int) -> float:
        return self.species[k].n_atoms(self.elements[e])

    def element_matrix(self) -> np.ndarray:
        """(n_elements, n_species) atom counts."""
        return np.array([[self.n_atoms(k, e) for k in range(self.n_species)]
                         for e in range(self.n_elements)])

    @property
    def falloff_reactions(self) -> List[Tuple[int, Reaction]]:
        return [(i, r) for i, r in enumerate(self.reactions)
                if r.reaction_type.startswith("falloff")]

    @property
    def three_body_reactions(self) -> List[Tuple[int, Reaction]]:
        return [(i, r) for i, r in enumerate(self.reactions)
                if r.reaction_type == "three-body-Arrhenius"]

    @property
    def num_falloff(self) -> int:
        return len(self.falloff_reactions)

    def digest(self) -> str:
        """Stable content hash (keys generated-code caches)."""
        payload = json.dumps(to_dict(self), sort_keys=True).encode()
        return hashlib.sha256(payload).hexdigest()[:16]


# ---------------------------------------------------------------------------
# YAML reading
# ---------------------------------------------------------------------------

_COLLIDER_RE = re.compile(r"\(\s*\+\s*([^)\s]+)\s*\)")


def _parse_side(side: str, species_names: Sequence[str]):
    """Parse one side of an equation into ({species: nu}, third_body).

    ``third_body`` is ``None``, ``"M"`` for ``+ M``, or ``("falloff", X)`` for
    ``(+ X)`` where X is ``M`` or a species name.
    """
    third = None
    m = _COLLIDER_RE.search(side)
    if m:
        third = ("falloff", m.group(1))
        side = side[:m.start()] + side[m.end():]
    stoich: Dict[str, float] = {}
    for term in side.split(" + "):
        term = term.strip()
        if not term:
            continue
        parts = term.split()
        if len(parts) == 2:
            nu, name = float(parts[0]), parts[1]
        elif len(parts) == 1:
            tok = parts[0]
            # "2OH"-style tokens are not used by the bundled mechanisms; a
            # leading number separated by a space is Cantera's canonical form.
            nu, name = 1.0, tok
        else:
            raise ValueError(f"cannot parse reaction term {term!r}")
        if name == "M":
            if third is not None:
                raise ValueError(f"two third bodies in {side!r}")
            third = "M"
            continue
        if name not in species_names:
            raise ValueError(f"unknown species {name!r} in {side!r}")
        stoich[name] = stoich.get(name, 0.0) + nu
    return stoich, third


def _split_equation(eq: str):
    for arrow, rev in (("<=>", True), ("=>", False), ("=", True)):
        if arrow in eq:
            lhs, rhs = eq.split(arrow, 1)
            return lhs, rhs, rev
    raise ValueError(f"no arrow in equation {eq!r}")


def _canonical_side(stoich: Dict[str, float], third) -> str:
    terms = []
    for name, nu in stoich.items():
        terms.append(name if nu == 1.0 else f"{nu:g} {name}")
    s = " + ".join(terms)
    if third == "M":
        s += " + M"
    elif isinstance(third, tuple):
        s += f" (+{third[1]})"
    return s


class _Units:
    def __init__(self, units: dict):
        units = units or {}
        self.length = _LENGTH_M[units.get("length", "m")]
        self.quantity = _QUANTITY_KMOL[units.get("quantity", "kmol")]
        self.time = _TIME_S[units.get("time", "s")]
        self.ea_unit = units.get("activation-energy",
                                 units.get("energy", "J/kmol"))
        if "energy" in units and "activation-energy" not in units:
            # energy is e.g. "cal" with quantity "mol"
            self.ea_unit = f"{units['energy']}/{units.get('quantity', 'kmol')}"

    def conc_factor(self) -> float:
        """(L^3/Q) expressed in m^3/kmol."""
        return self.length ** 3 / self.quantity

    def rate(self, d: dict, order: float) -> Arrhenius:
        if isinstance(d.get("A"), str) or isinstance(d.get("Ea"), str):
            raise ValueError("rate constants with explicit unit strings are "
                             "not supported")
        A = float(d["A"]) * self.conc_factor() ** (order - 1.0) / self.time
        b = float(d.get("b", 0.0))
        ea = float(d.get("Ea", 0.0))
        if self.ea_unit == "K":
            ea_si = ea * GAS_CONSTANT
        else:
            ea_si = ea * _ENERGY_J_PER_KMOL[self.ea_unit]
        return Arrhenius(A=A, b=b, Ea=ea_si)


def _troe_coeffs(troe: dict) -> Tuple[float, ...]:
    """Troe coefficients as Cantera's ``TroeRate.falloff_coeffs`` reports
    them: T3 and T1 are stored as reciprocals internally and inverted back;
    T2 appears only when nonzero."""
    a = float(troe["A"])
    t3 = float(troe["T3"])
    t1 = float(troe["T1"])
    t2 = float(troe.get("T2", 0.0))
    small = 1e-300
    rt3 = math.inf if abs(t3) < small else 1.0 / t3
    rt1 = math.inf if abs(t1) < small else 1.0 / t1
    c = [a, 1.0 / rt3, 1.0 / rt1]
    if abs(t2) >= small:
        c.append(t2)
    return tuple(c)


def _species_weight(comp: Dict[str, float], elements: Sequence[str]) -> float:
    w = 0.0
    for el in elements:
        n = comp.get(el, 0.0)
        if n:
            w += n * ATOMIC_WEIGHTS[el]
    return w


class _Yaml12Loader(yaml.SafeLoader):
    """SafeLoader with YAML 1.2 booleans (only true/false), as Cantera's
    yaml-cpp reads them: a species named ``NO`` must stay a string."""


_Yaml12Loader.yaml_implicit_resolvers = {
    k: [(tag, rx) for tag, rx in v if tag != "tag:yaml.org,2002:bool"]
    for k, v in yaml.SafeLoader.yaml_implicit_resolvers.items()
}
_Yaml12Loader.add_implicit_resolver(
    "tag:yaml.org,2002:bool", re.compile(r"^(?:true|True|TRUE|false|False|FALSE)$"),
    list("tTfF"))


def load_mechanism(path: str, phase: Optional[str] = None,
                   name: Optional[str] = None) -> Mechanism:
    """Read a Cantera YAML mechanism into the IR (``ct.Solution(path, phase)``
    equivalent for the fields the generated code uses)."""
    with open(path) as f:
        doc = yaml.load(f, Loader=_Yaml12Loader)
    return mechanism_from_dict(doc, phase=phase,
                               source=os.path.basename(path),
                               name=name or os.path.splitext(
                                   os.path.basename(path))[0])


def mechanism_from_dict(doc: dict, phase: Optional[str] = None,
                        source: str = "<dict>",
                        name: str = "mechanism") -> Mechanism:
    phases = doc["phases"]
    ph = phases[0] if phase is None else next(
        p for p in phases if p["name"] == phase)
    elements = tuple(ph["elements"])
    units = _Units(doc.get("units", {}))

    sp_section = ph.get("species", "all")
    all_sp = {s["name"]: s for s in doc["species"]}
    if sp_section == "all":
        sp_names = [s["name"] for s in doc["species"]]
    else:
        sp_names = list(sp_section)

    species = []
    for nm in sp_names:
        s = all_sp[nm]
        th = s["thermo"]
        if th.get("model") != "NASA7":
            raise ValueError(f"species {nm}: only NASA7 thermo is supported")
        tr = th["temperature-ranges"]
        data = th["data"]
        if len(tr) == 2:
            low = high = tuple(float(x) for x in data[0])
            t_min, t_mid, t_max = float(tr[0]), float(tr[1]), float(tr[1])
        else:
            t_min, t_mid, t_max = (float(x) for x in tr)
            low = tuple(float(x) for x in data[0])
            high = tuple(float(x) for x in data[1])
        comp = {k: float(v) for k, v in s["composition"].items()}
        tr_doc = s.get("transport")
        transport = None
        if tr_doc is not None and tr_doc.get("model", "gas") == "gas":
            transport = (str(tr_doc["geometry"]),
                         float(tr_doc["diameter"]),
                         float(tr_doc["well-depth"]),
                         float(tr_doc.get("dipole", 0.0)),
                         float(tr_doc.get("polarizability", 0.0)),
                         float(tr_doc.get("rotational-relaxation", 0.0)))
        species.append(Species(
            name=nm,
            composition=tuple(sorted(comp.items())),
            t_min=t_min, t_mid=t_mid, t_max=t_max, low=low, high=high,
            weight=_species_weight(comp, elements), transport=transport))

    rx_section = ph.get("reactions", "all")
    if rx_section in ("all", "declared-species"):
        rx_docs = doc.get("reactions", [])
    elif rx_section == "none":
        rx_docs = []
    else:
        raise ValueError(f"unsupported reactions spec {rx_section!r}")

    reactions = [_parse_reaction(r, sp_names, units) for r in rx_docs]
    return Mechanism(name=name, source=source, elements=elements,
                     species=tuple(species), reactions=tuple(reactions))


def _parse_reaction(r: dict, sp_names: Sequence[str], units: _Units
                    ) -> Reaction:
    eq = " ".join(r["equation"].split())
    lhs, rhs, reversible = _split_equation(eq)
    reac, third_l = _parse_side(lhs, sp_names)
    prod, third_r = _parse_side(rhs, sp_names)
    if third_l != third_r:
        raise ValueError(f"unbalanced third body in {eq!r}")
    rtype = r.get("type", "elementary")

    orders = {k: float(v) for k, v in (r.get("orders") or {}).items()}
    order = 0.0
    for name, nu in reac.items():
        order += orders.get(name, nu)
    # explicit orders for non-reactants also count toward the units
    for name, o in orders.items():
        if name not in reac:
            order += o

    common = dict(
        equation=eq,
        reactants=tuple(sorted(reac.items())),
        products=tuple(sorted(prod.items())),
        reversible=reversible,
        orders=tuple(sorted(orders.items())),
        duplicate=bool(r.get("duplicate", False)),
    )

    if rtype == "falloff" or isinstance(third_l, tuple):
        if rtype not in ("falloff",):
            raise ValueError(f"unsupported falloff type {rtype!r} in {eq!r}")
        collider = third_l[1] if isinstance(third_l, tuple) else "M"
        if "SRI" in r or "Tsang" in r:
            raise ValueError(f"only Troe/Lindemann falloff supported: {eq!r}")
        high = units.rate(r["high-P-rate-constant"], order)
        low = units.rate(r["low-P-rate-constant"], order + 1.0)
        if collider == "M":
            effs = {k: float(v) for k, v in (r.get("efficiencies") or {}).items()
                    if k in sp_names}
            default = float(r.get("default-efficiency", 1.0))
        else:
            effs = {collider: 1.0}
            default = 0.0
        if "Troe" in r:
            ftype, troe = "Troe", _troe_coeffs(r["Troe"])
        else:
            ftype, troe = "Lindemann", ()
        return Reaction(kind="falloff", low_rate=low, high_rate=high,
                        falloff_type=ftype, troe=troe,
                        efficiencies=tuple(sorted(effs.items())),
                        default_efficiency=default, **common)

    if rtype == "three-body" or third_l == "M":
        if third_l != "M":
            raise ValueError(
                f"three-body reaction without M is not supported: {eq!r}")
        rate = units.rate(r["rate-constant"], order + 1.0)
        effs = {k: float(v) for k, v in (r.get("efficiencies") or {}).items()
                if k in sp_names}
        default = float(r.get("default-efficiency", 1.0))
        return Reaction(kind="three-body", rate=rate,
                        efficiencies=tuple(sorted(effs.items())),
                        default_efficiency=default, **common)

    if rtype not in ("elementary", "reaction", "Arrhenius"):
        raise ValueError(f"unsupported reaction type {rtype!r} in {eq!r}")
    rate = units.rate(r["rate-constant"], order)
    return Reaction(kind="elementary", rate=rate, **common)


# ---------------------------------------------------------------------------
# Serialisation (also the schema the synthetic generator emits)
# ---------------------------------------------------------------------------

def to_dict(mech: Mechanism) -> dict:
    """Cantera-YAML-shaped dict in SI/kmol units (round-trips through
    :func:`mechanism_from_dict`)."""
    def rate(a: Arrhenius):
        return {"A": a.A, "b": a.b, "Ea": a.Ea}

    sp = []
    for s in mech.species:
        d = {
            "name": s.name,
            "composition": dict(s.composition),
            "thermo": {"model": "NASA7",
                       "temperature-ranges": [s.t_min, s.t_mid, s.t_max],
                       "data": [list(s.low), list(s.high)]},
        }
        if s.transport is not None:
            g, dia, eps, dip, pol, rot = s.transport
            d["transport"] = {"model": "gas", "geometry": g,
                              "diameter": dia, "well-depth": eps,
                              "dipole": dip, "polarizability": pol,
                              "rotational-relaxation": rot}
        sp.append(d)
    rx = []
    for r in mech.reactions:
        d = {"equation": r.equation}
        if r.kind == "three-body":
            d["type"] = "three-body"
        if r.kind == "falloff":
            d["type"] = "falloff"
            d["low-P-rate-constant"] = rate(r.low_rate)
            d["high-P-rate-constant"] = rate(r.high_rate)
            if r.falloff_type == "Troe":
                t = {"A": r.troe[0], "T3": r.troe[1], "T1": r.troe[2]}
                if len(r.troe) == 4:
                    t["T2"] = r.troe[3]
                d["Troe"] = t
        else:
            d["rate-constant"] = rate(r.rate)
        m = _COLLIDER_RE.search(r.equation)
        explicit_collider = m is not None and m.group(1) != "M"
        if r.kind != "elementary" and not explicit_collider:
            if r.efficiencies:
                d["efficiencies"] = dict(r.efficiencies)
            if r.default_efficiency != 1.0:
                d["default-efficiency"] = r.default_efficiency
        if r.orders:
            d["orders"] = dict(r.orders)
        if r.duplicate:
            d["duplicate"] = True
        rx.append(d)
    return {
        "units": {"length": "m", "quantity": "kmol",
                  "activation-energy": "J/kmol"},
        "phases": [{"name": mech.name, "thermo": "ideal-gas",
                    "elements": list(mech.elements),
                    "species": [s.name for s in mech.species],
                    "kinetics": "gas", "reactions": "all"}],
        "species": sp,
        "reactions": rx,
    }


def dump_yaml(mech: Mechanism, path: str) -> None:
    with open(path, "w") as f:
        yaml.safe_dump(to_dict(mech), f, sort_keys=False)


# ---------------------------------------------------------------------------
# Bundled mechanisms
# ---------------------------------------------------------------------------

MECH_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mechs")


def bundled(name: str) -> Mechanism:
    """Load one of the bundled mechanisms (``sandiego``, ``gri30``, ``hong``,
    ``uiuc``, ``uconn32``, ``bfer``) or a synthetic one (``synth9``,
    ``synth53``)."""
    if name == "synth9":
        from .synthetic import synthetic_h2_like
        return synthetic_h2_like()
    if name == "synth53":
        from .synthetic import synthetic_gri_like
        return synthetic_gri_like()
    return load_mechanism(os.path.join(MECH_DIR, f"{name}.yaml"), name=name)


def as_mechanism(sol, phase: Optional[str] = None) -> Mechanism:
    """The IR for what a code generator is handed: a ``Mechanism`` as is, a
    path to a Cantera YAML file, or an object with Cantera's
    ``Solution.write_yaml`` (a ``ct.Solution``, written to a temporary YAML
    file and read back; Cantera itself is optional)."""
    if isinstance(sol, Mechanism):
        return sol
    if isinstance(sol, (str, os.PathLike)):
        return load_mechanism(os.fspath(sol), phase=phase)
    if hasattr(sol, "write_yaml"):
        import tempfile
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "mechanism.yaml")
            sol.write_yaml(path)
            return load_mechanism(path, phase=phase,
                                  name=getattr(sol, "name", None)
                                  or "mechanism")
    raise TypeError(f"cannot read a mechanism from {type(sol).__name__}")
