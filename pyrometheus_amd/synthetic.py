"""Seeded synthetic mechanisms with prescribed reaction-type counts.

Two workloads name mechanisms by shape rather than by file: a 9-species,
21-reaction H2/O2/N2-sized one (3 Troe falloff, 4 three-body) and a
53-species, 325-reaction methane-sized one (26 Troe + 3 Lindemann falloff,
12 three-body).  The bundled files do not have the first shape
(``sandiego.yaml`` has 24 reactions, 2 Troe, 5 three-body), so this module
builds both from a seed:

* species compositions are taken from a real mechanism of the same size
  (sandiego's 9 species, GRI-3.0's 53), so every generated reaction can be
  element-balanced and element conservation stays a testable invariant;
* reactions are drawn from the element-balanced association
  (``A + B <=> C``) and exchange (``A + B <=> C + D``) candidates;
* ``ln A ~ U[ln 1e6, ln 1e18]`` in SI/kmol units, ``b ~ U[-2, 2]``,
  ``Ea ~ U[0, 2e5]`` J/mol (about 15% exactly zero so the ``A T^b`` branch of
  ``chem_expr.rate_coefficient_expr`` is exercised), all reversible;
* NASA-7 coefficients are least-squares quartic fits of a smooth
  ``cp/R = a + b tanh((T - Tc)/w)`` on [200, 1000] K and [1000, 3500] K, with
  ``a5``/``a6`` set from a drawn formation enthalpy/entropy and made
  continuous in h and s at 1000 K.

The output is an ordinary :class:`~pyrometheus_amd.mechanism.Mechanism`
(serialisable to the Cantera YAML schema with
:func:`~pyrometheus_amd.mechanism.dump_yaml`).
"""

from __future__ import annotations

import itertools
import math
from typing import Dict, List, Tuple

import numpy as np

from .mechanism import (Arrhenius, Mechanism, Reaction, Species,
                        _species_weight, bundled)

T_LO, T_MID, T_HI = 200.0, 1000.0, 3500.0


def _fit_nasa(rng: np.random.Generator) -> Tuple[Tuple[float, ...],
                                                  Tuple[float, ...]]:
    a = rng.uniform(2.5, 6.0)
    b = rng.uniform(0.3, 3.0)
    tc = rng.uniform(700.0, 2000.0)
    w = rng.uniform(300.0, 1200.0)

    def cp(t):
        return a + b * np.tanh((t - tc) / w)

    def fit(t0, t1):
        t = np.linspace(t0, t1, 200)
        # fit in a scaled variable for conditioning, then expand to monomials
        s = 1000.0
        c = np.polynomial.polynomial.polyfit(t / s, cp(t), 4)
        return [c[j] / s ** j for j in range(5)]

    lo = fit(T_LO, T_MID)
    hi = fit(T_MID, T_HI)

    def hpoly(c, t):  # h/RT without the a5/T term
        return c[0] + c[1] / 2 * t + c[2] / 3 * t**2 + c[3] / 4 * t**3 \
            + c[4] / 5 * t**4

    def spoly(c, t):  # s/R without a6
        return c[0] * math.log(t) + c[1] * t + c[2] / 2 * t**2 \
            + c[3] / 3 * t**3 + c[4] / 4 * t**4

    t0 = 298.15
    h298 = rng.uniform(-3.0e4, 6.0e4)  # h/R at 298.15 K, in K
    s298 = rng.uniform(15.0, 35.0)
    a5_lo = h298 - t0 * hpoly(lo, t0)
    a6_lo = s298 - spoly(lo, t0)
    a5_hi = a5_lo + T_MID * (hpoly(lo, T_MID) - hpoly(hi, T_MID))
    a6_hi = a6_lo + spoly(lo, T_MID) - spoly(hi, T_MID)
    low = tuple(float(x) for x in lo) + (float(a5_lo), float(a6_lo))
    high = tuple(float(x) for x in hi) + (float(a5_hi), float(a6_hi))
    return low, high


def _comp_key(comp: Dict[str, float]) -> Tuple:
    return tuple(sorted((k, v) for k, v in comp.items() if v))


def _add(c1, c2):
    out = dict(c1)
    for k, v in c2.items():
        out[k] = out.get(k, 0.0) + v
    return _comp_key(out)


def _candidates(species: List[Species], inert: set):
    comps = {s.name: dict(s.composition) for s in species}
    names = [s.name for s in species if s.name not in inert]
    by_comp: Dict[Tuple, List[str]] = {}
    for n in names:
        by_comp.setdefault(_comp_key(comps[n]), []).append(n)
    pair_sum: Dict[Tuple, List[Tuple[str, str]]] = {}
    for a, b in itertools.combinations_with_replacement(names, 2):
        pair_sum.setdefault(_add(comps[a], comps[b]), []).append((a, b))
    assoc = []
    for key, pairs in pair_sum.items():
        for c in by_comp.get(key, []):
            for p in pairs:
                if c not in p:
                    assoc.append((p, (c,)))
    exch = []
    for key, pairs in pair_sum.items():
        for p, q in itertools.combinations(pairs, 2):
            if set(p) & set(q):
                continue
            exch.append((p, q))
    return assoc, exch


def _stoich(names) -> Tuple[Tuple[str, float], ...]:
    d: Dict[str, float] = {}
    for n in names:
        d[n] = d.get(n, 0.0) + 1.0
    return tuple(sorted(d.items()))


def _side_str(names) -> str:
    d = dict(_stoich(names))
    return " + ".join(n if v == 1.0 else f"{v:g} {n}" for n, v in d.items())


def _arrhenius(rng, order: float, zero_ea_frac: float = 0.15) -> Arrhenius:
    ln_a = rng.uniform(math.log(1e6), math.log(1e18))
    b = rng.uniform(-2.0, 2.0)
    ea = 0.0 if rng.uniform() < zero_ea_frac else rng.uniform(0.0, 2.0e5)
    return Arrhenius(A=math.exp(ln_a), b=b, Ea=ea * 1e3)  # J/mol -> J/kmol


def _efficiencies(rng, species: List[Species]) -> Tuple[Tuple[str, float], ...]:
    n = int(rng.integers(2, 5))
    picks = rng.choice(len(species), size=n, replace=False)
    return tuple(sorted((species[i].name, float(round(rng.uniform(0.5, 12.0), 3)))
                        for i in picks))


def synthetic_mechanism(base: str, name: str, seed: int, n_reactions: int,
                        n_troe: int, n_lindemann: int, n_three_body: int,
                        inert=("N2",)) -> Mechanism:
    """Generate a mechanism over ``base``'s species compositions."""
    rng = np.random.default_rng(seed)
    ref = bundled(base)
    species = []
    for s in ref.species:
        low, high = _fit_nasa(rng)
        comp = dict(s.composition)
        species.append(Species(name=s.name, composition=s.composition,
                               t_min=T_LO, t_mid=T_MID, t_max=T_HI,
                               low=low, high=high,
                               weight=_species_weight(comp, ref.elements)))
    assoc, exch = _candidates(species, set(inert))
    rng.shuffle(assoc)
    rng.shuffle(exch)
    n_assoc = n_troe + n_lindemann + n_three_body
    n_exch = n_reactions - n_assoc
    if len(assoc) < n_assoc:
        raise ValueError("not enough balanced association candidates")

    def pick(pool, n):
        if n <= len(pool):
            return list(pool[:n])
        # allow duplicates once every distinct candidate is used
        idx = rng.integers(0, len(pool), size=n - len(pool))
        return list(pool) + [pool[i] for i in idx]

    kinds = (["troe"] * n_troe + ["lindemann"] * n_lindemann
             + ["three-body"] * n_three_body + ["elementary"] * n_exch)
    rng.shuffle(kinds)
    assoc_it = iter(pick(assoc, n_assoc))
    exch_it = iter(pick(exch, n_exch))
    seen = set()
    reactions = []
    for kind in kinds:
        if kind == "elementary":
            lhs, rhs = next(exch_it)
        else:
            lhs, rhs = next(assoc_it)
        order = float(len(lhs))
        reac, prod = _stoich(lhs), _stoich(rhs)
        key = (kind in ("troe", "lindemann"), kind == "three-body", reac, prod)
        dup = key in seen
        seen.add(key)
        if kind == "elementary":
            eq = f"{_side_str(lhs)} <=> {_side_str(rhs)}"
            reactions.append(Reaction(
                equation=eq, reactants=reac, products=prod, reversible=True,
                kind="elementary", rate=_arrhenius(rng, order), duplicate=dup))
        elif kind == "three-body":
            eq = f"{_side_str(lhs)} + M <=> {_side_str(rhs)} + M"
            reactions.append(Reaction(
                equation=eq, reactants=reac, products=prod, reversible=True,
                kind="three-body", rate=_arrhenius(rng, order + 1),
                efficiencies=_efficiencies(rng, species),
                default_efficiency=1.0, duplicate=dup))
        else:
            eq = f"{_side_str(lhs)} (+M) <=> {_side_str(rhs)} (+M)"
            if kind == "troe":
                t3 = rng.uniform(50.0, 500.0)
                t1 = rng.uniform(1000.0, 8000.0)
                troe = (float(rng.uniform(0.1, 0.9)), 1.0 / (1.0 / t3),
                        1.0 / (1.0 / t1))
                if rng.uniform() < 0.5:
                    troe = troe + (float(rng.uniform(1000.0, 8000.0)),)
                ftype = "Troe"
            else:
                troe, ftype = (), "Lindemann"
            reactions.append(Reaction(
                equation=eq, reactants=reac, products=prod, reversible=True,
                kind="falloff", high_rate=_arrhenius(rng, order),
                low_rate=_arrhenius(rng, order + 1), falloff_type=ftype,
                troe=troe, efficiencies=_efficiencies(rng, species),
                default_efficiency=1.0, duplicate=dup))
    return Mechanism(name=name, source=f"synthetic:{name}:seed{seed}",
                     elements=ref.elements, species=tuple(species),
                     reactions=tuple(reactions))


def synthetic_h2_like(seed: int = 101) -> Mechanism:
    """9 species / 21 reactions: 3 Troe falloff, 4 three-body, 14 elementary."""
    return synthetic_mechanism("sandiego", "synth9", seed, n_reactions=21,
                               n_troe=3, n_lindemann=0, n_three_body=4)


def synthetic_gri_like(seed: int = 102) -> Mechanism:
    """53 species / 325 reactions: 26 Troe + 3 Lindemann falloff,
    12 three-body, 284 elementary."""
    return synthetic_mechanism("gri30", "synth53", seed, n_reactions=325,
                               n_troe=26, n_lindemann=3, n_three_body=12,
                               inert=("N2", "AR"))
