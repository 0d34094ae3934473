"""Mechanism reader: counts, Cantera unit conversion, synthetic generator."""
import math

import numpy as np
import pytest

from pyrometheus_amd.mechanism import (GAS_CONSTANT, bundled, dump_yaml,
                                       load_mechanism, mechanism_from_dict,
                                       to_dict)


def _types(m):
    d = {}
    for r in m.reactions:
        d[r.reaction_type] = d.get(r.reaction_type, 0) + 1
    return d


def test_gri30_counts():
    m = bundled("gri30")
    assert (m.n_species, m.n_reactions) == (53, 325)
    assert sum(r.reversible for r in m.reactions) == 309
    assert _types(m) == {"Arrhenius": 284, "three-body-Arrhenius": 12,
                         "falloff-Troe": 26, "falloff-Lindemann": 3}
    assert sorted({s.t_mid for s in m.species}) == [1000.0, 1368.0, 1382.0,
                                                     1478.0]
    # YAML 1.2 booleans: the species "NO" stays a string
    assert "NO" in m.species_names


def test_sandiego_phase_species():
    m = bundled("sandiego")
    # the phase declares 9 of the file's 12 species
    assert m.species_names == ["H2", "H", "O2", "O", "OH", "HO2", "H2O2",
                               "H2O", "N2"]
    assert m.n_reactions == 24
    assert _types(m) == {"Arrhenius": 17, "three-body-Arrhenius": 5,
                         "falloff-Troe": 2}


def test_unit_conversion_sandiego():
    m = bundled("sandiego")
    r0 = m.reactions[0]  # H + O2 <=> OH + O, A 3.52e16 cm3/mol/s, Ea kJ/mol
    assert math.isclose(r0.rate.A, 3.52e16 * 1e-3, rel_tol=1e-15)
    assert math.isclose(r0.rate.Ea, 71.42e6, rel_tol=1e-15)
    r4 = m.reactions[4]  # H + H + M: third body adds an order
    assert r4.kind == "three-body"
    assert math.isclose(r4.rate.A, 1.3e18 * 1e-6, rel_tol=1e-15)
    assert dict(r4.efficiencies) == {"H2": 2.5, "H2O": 12.0}
    r9 = m.reactions[9]  # H + O2 (+M) <=> HO2 (+M)
    assert r9.reaction_type == "falloff-Troe"
    assert math.isclose(r9.high_rate.A, 4.65e12 * 1e-3, rel_tol=1e-15)
    assert math.isclose(r9.low_rate.A, 5.75e19 * 1e-6, rel_tol=1e-15)
    # Troe T3 and T1 are reported through their reciprocals, as Cantera does
    assert r9.troe[0] == 0.5
    assert r9.troe[1] == 1.0 / (1.0 / 1e-30)
    assert r9.troe[2] == 1.0 / (1.0 / 1e30)
    assert len(r9.troe) == 3


def test_unit_conversion_cal_and_colliders():
    g = bundled("gri30")
    # O + H2 <=> H + OH: A 3.87e4, b 2.7, Ea 6260 cal/mol
    r = next(r for r in g.reactions if r.equation.startswith("O + H2 <=>"))
    assert math.isclose(r.rate.A, 3.87e4 * 1e-3, rel_tol=1e-15)
    assert math.isclose(r.rate.Ea, 6260 * 4184.0, rel_tol=1e-15)
    assert math.isclose(r.rate.t_a, 6260 * 4184.0 / GAS_CONSTANT)
    h = bundled("hong")
    r3 = h.reactions[2]  # H + O2 (+ Ar) <=> HO2 (+ Ar)
    assert dict(r3.efficiencies) == {"Ar": 1.0}
    assert r3.default_efficiency == 0.0
    assert len(r3.troe) == 4
    u = bundled("uiuc")
    r0 = u.reactions[0]
    assert not r0.reversible and dict(r0.orders) == {"C2H4": 0.5, "O2": 0.65}
    # the rate order uses the explicit orders: 0.5 + 0.65
    assert math.isclose(r0.rate.A, 1e12 * (1e-3) ** (1.15 - 1), rel_tol=1e-14)


def test_molecular_weights():
    m = bundled("sandiego")
    w = dict(zip(m.species_names, m.molecular_weights))
    assert w["H2O"] == 2 * 1.008 + 15.999
    assert w["N2"] == 2 * 14.007
    g = bundled("gri30")
    wg = dict(zip(g.species_names, g.molecular_weights))
    assert math.isclose(wg["CH4"], 12.011 + 4 * 1.008, rel_tol=1e-15)
    assert wg["AR"] == 39.95


@pytest.mark.parametrize("name", ["sandiego", "gri30", "hong", "uiuc"])
def test_roundtrip(name, tmp_path):
    m = bundled(name)
    m2 = mechanism_from_dict(to_dict(m), name=name, source=m.source)
    assert m2.species == m.species and m2.reactions == m.reactions
    p = tmp_path / f"{name}.yaml"
    dump_yaml(m, str(p))
    m3 = load_mechanism(str(p), name=name)
    assert m3.reactions == m.reactions


@pytest.mark.parametrize("name,counts", [
    ("synth9", (9, 21, {"Arrhenius": 14, "three-body-Arrhenius": 4,
                        "falloff-Troe": 3})),
    ("synth53", (53, 325, {"Arrhenius": 284, "three-body-Arrhenius": 12,
                           "falloff-Troe": 26, "falloff-Lindemann": 3})),
])
def test_synthetic(name, counts):
    m = bundled(name)
    ns, nr, types = counts
    assert (m.n_species, m.n_reactions) == (ns, nr)
    assert _types(m) == types
    assert all(r.reversible for r in m.reactions)
    E = m.element_matrix()
    for r in m.reactions:
        lhs = sum(E[:, m.species_index(s)] * v for s, v in r.reactants)
        rhs = sum(E[:, m.species_index(s)] * v for s, v in r.products)
        assert np.array_equal(lhs, rhs), r.equation
    # deterministic in the seed
    assert bundled(name).digest() == m.digest()
    # h and s are continuous at T_mid by construction
    for s in m.species:
        t = s.t_mid

        def h(c):
            return (c[0] + c[1] / 2 * t + c[2] / 3 * t**2 + c[3] / 4 * t**3
                    + c[4] / 5 * t**4 + c[5] / t)
        assert math.isclose(h(s.low), h(s.high), rel_tol=1e-12, abs_tol=1e-9)


def test_as_mechanism_paths_and_solution_like(tmp_path):
    """A generator accepts the IR, a YAML path, or anything with Cantera's
    ``write_yaml`` (stand-in object; Cantera is not installed)."""
    import os
    from pyrometheus_amd.mechanism import MECH_DIR, as_mechanism, bundled
    m = bundled("sandiego")
    assert as_mechanism(m) is m
    p = os.path.join(MECH_DIR, "sandiego.yaml")
    assert as_mechanism(p).digest() == m.digest()

    class SolutionLike:
        name = "sandiego"

        def write_yaml(self, path):
            with open(p) as src, open(path, "w") as dst:
                dst.write(src.read())

    assert as_mechanism(SolutionLike()).digest() == m.digest()
    with pytest.raises(TypeError):
        as_mechanism(3.0)
