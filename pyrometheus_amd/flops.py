"""Algorithmic FLOP count of one net-production-rate evaluation per cell.

Convention (fixed per mechanism, independent of how a kernel computes):
add, sub, mul, div count 1; FMA counts 2; each exp, log, log10 and pow counts
``W_T`` = 20.  With ns species, nr reactions (nrev reversible), nf falloff
(nT Troe), n3 three-body, E explicit efficiency entries, R_i / P_i the
distinct reactant / product species of reaction i and nnz = sum(R_i + P_i):

  F_arith = 10 + 2 ns + 21 ns + sum_rev(2 (R_i + P_i) + 2) + 4 (nr + nf)
            + ns + 2 E + 23 nT + 4 (nf - nT) + (sum R_i + nr)
            + sum_rev(P_i + 3) + n3 + 2 nnz
  N_trans = 1 + (nr + nf) + 6 nT + nrev

The terms follow the generated ``get_net_production_rates`` call tree
(``pyrometheus/codegen/python.py:486-495``): concentrations, NASA-7 Gibbs
functions, -ln K_c, Arrhenius exponents, third-body sums, Troe blending,
rates of progress and the species sums.  For GRI-3.0 this gives
F_arith = 11,860 and N_trans = 820, i.e. 28,260 FLOP per cell.
"""

from __future__ import annotations

import json
import os

from .mechanism import Mechanism

W_TRANS = 20


def wdot_flops(mech: Mechanism, w_trans: int = W_TRANS) -> dict:
    ns, nr = mech.n_species, mech.n_reactions
    rev = [r for r in mech.reactions if r.reversible]
    nf = mech.num_falloff
    n_troe = sum(1 for r in mech.reactions
                 if r.is_falloff and r.falloff_type == "Troe")
    n3 = len(mech.three_body_reactions)
    n_eff = sum(len(r.efficiencies) for r in mech.reactions
                if r.kind in ("three-body", "falloff"))
    R = [len(r.reactants) for r in mech.reactions]
    P = [len(r.products) for r in mech.reactions]
    nnz = sum(R) + sum(P)
    sum_rev_rp = sum(2 * (len(r.reactants) + len(r.products)) + 2 for r in rev)
    sum_rev_p = sum(len(r.products) + 3 for r in rev)
    f_arith = (10 + 2 * ns + 21 * ns + sum_rev_rp + 4 * (nr + nf) + ns
               + 2 * n_eff + 23 * n_troe + 4 * (nf - n_troe) + (sum(R) + nr)
               + sum_rev_p + n3 + 2 * nnz)
    n_trans = 1 + (nr + nf) + 6 * n_troe + len(rev)
    return {"mechanism": mech.name, "ns": ns, "nr": nr, "nrev": len(rev),
            "nf": nf, "n_troe": n_troe, "n3": n3, "n_eff": n_eff,
            "sum_R": sum(R), "sum_P_rev": sum(len(r.products) for r in rev),
            "nnz": nnz, "f_arith": f_arith, "n_trans": n_trans,
            "w_trans": w_trans, "flop_per_cell": f_arith + w_trans * n_trans,
            "bytes_per_cell": 8 * (2 * ns + 2)}


FLOP_JSON = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mechs",
                         "flop_counts.json")


def write_flop_json(names=("sandiego", "synth9", "hong", "gri30", "synth53")):
    from .mechanism import bundled
    data = {n: wdot_flops(bundled(n)) for n in names}
    with open(FLOP_JSON, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    return data
