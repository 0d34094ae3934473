"""Transport tables of the ``hip`` target.

For a mechanism whose species all carry transport data, the generated
translation unit gets the fit coefficients of :mod:`pyrometheus_amd.transport`
(the forms of ``chem_expr.py:190-243``; the binary-diffusion fits as the
full symmetric (ns, ns) table) and the constant factors of Wilke's mixture
rule (``chem_expr.viscosity_mixture_rule_wilke_expr``, ``:246-258``), once for the device (``__constant__``, read with scalar
loads) and once for the host (``pyro_transport_fits`` reports them), and
includes ``pyro_transport_kernels.hpp`` for the kernels.
"""

from __future__ import annotations

import numpy as np

from ..mechanism import Mechanism
from ..transport import has_transport, transport_fits


def _rows(a, lit) -> str:
    return ", ".join("{" + ", ".join(lit(x) for x in row) + "}" for row in a)


def emit_transport_tables(w, mech: Mechanism, lit) -> bool:
    """Write the tables into namespace ``mech``; returns False (writing
    nothing) when a species lacks transport data."""
    if not has_transport(mech):
        return False
    visc, cond, diff = transport_fits(mech)
    ns = mech.n_species
    full = diff.reshape(ns * ns, 5)
    W = mech.molecular_weights
    # sqrt((mu_k / mu_j) sqrt(W_j / W_k)) = (|q_k| / |q_j|) (W_j / W_k)^(1/4)
    # with mu = sqrt(T) q^2, and the divisor as a factor
    wilke_s = (W[None, :] / W[:, None]) ** 0.25
    wilke_d = 1.0 / np.sqrt(8.0 * (1.0 + W[:, None] / W[None, :]))
    w("// transport fits in ln T (pyrometheus_amd.transport) and Wilke's")
    w("// constant factors")
    w("constexpr int kHasTransport = 1;")
    for nm, a in (("kTrVisc", visc), ("kTrCond", cond)):
        w(f"__constant__ double {nm}[NS][5] = {{{_rows(a, lit)}}};")
        w(f"constexpr double {nm}Host[NS][5] = {{{_rows(a, lit)}}};")
    w("__constant__ double kTrDiff[NS * NS][5] = {" + _rows(full, lit) + "};")
    w("constexpr double kTrDiffHost[NS * NS][5] = {"
      + _rows(full, lit) + "};")
    w("__constant__ double kTrW[NS] = {" + ", ".join(lit(x) for x in W)
      + "};")
    w(f"__constant__ double kTrWilkeS[NS][NS] = {{{_rows(wilke_s, lit)}}};")
    w(f"__constant__ double kTrWilkeD[NS][NS] = {{{_rows(wilke_d, lit)}}};")
    w("")
    return True
