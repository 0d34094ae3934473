"""The ``hip`` code-generation target.

``HipCodeGenerator.generate(name, mech)`` lowers the expression semantics of
``pyrometheus/chem_expr.py`` into one HIP C++ translation unit for gfx950:

* one cell per lane; every mechanism constant (NASA-7 coefficients, Arrhenius
  parameters, Troe coefficients, third-body efficiencies, stoichiometry) is a
  compile-time immediate, so nothing is loaded per reaction;
* species vectors are register arrays indexed only by constants;
* the kinetics are emitted reaction-major: each reaction's ``k_f``, ``-ln K_c``
  and rate of progress are computed and immediately folded into the
  per-species accumulators, so no ``(nr, N)`` intermediate exists;
* the mechanism-independent kernels and the C ABI come from the hand-written
  ``csrc/pyro_thermo_kernels.hpp``, which the generated unit includes.

Expression-level correspondences with the reference (all FP64):

=============================  =============================================
reference                      generated form
=============================  =============================================
``nasa7_conditional``          coefficient select on ``T > T_mid``, Horner
(``chem_expr.py:51-174``)      (strict ``>``: ``T == T_mid`` takes low)
``rate_coefficient_expr``      ``Ea == 0``: ``A*T^b`` (``b`` integer: exact
(``chem_expr.py:310-324``)     multiplies; else ``A*exp(b ln T)``);
                               else ``exp(ln A + b ln T - Ta/T)``
``third_body_efficiencies``    ``d*sum(C) + sum((eff-d)*C_k)``
(``chem_expr.py:327-345``)
Troe / Lindemann               ``Fcent``, ``f1`` with the ``Pr > 0`` guard,
(``chem_expr.py:348-421``)     ``F = 10^(Fcent/(1+f1^2))``
``equilibrium_constants_expr`` ``sum_p nu g - sum_r nu g - dnu*c0``;
(``chem_expr.py:274-302``)     irreversible placeholder ``-0.17364695002734*T``
``rate_of_progress_expr``      ``k_f*(prod C^nu' - exp(-ln K_c) prod C^nu'')``
(``chem_expr.py:428-460``)     with the ``orders`` zip quirk kept
``production_rate_expr``       accumulate ``+/- nu*q`` per species
=============================  =============================================

The arithmetic is reassociated (Horner, fused multiply-add) relative to the
numpy source, so results agree with the reference to rounding, not bitwise.
"""

from __future__ import annotations

import math
import os
from typing import Dict, List, Optional, Sequence, Tuple

from ..mechanism import (GAS_CONSTANT, ONE_ATM, Mechanism, Reaction,
                         as_mechanism)
from . import CodeGenerationOptions, CodeGenerator
from .hip_transport import emit_transport_tables

ABI_VERSION = 1
#: mechanisms up to this many species also get the dense-Jacobian kernels
#: (the NS x NS Jacobian lives in registers) and the flamelet kernels
DENSE_JAC_MAX_SPECIES = 16
CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "csrc")


def lit(x: float) -> str:
    """Exact C++ double literal."""
    x = float(x)
    if math.isinf(x):
        return "(__builtin_inf())" if x > 0 else "(-__builtin_inf())"
    if math.isnan(x):
        return "(__builtin_nan(\"\"))"
    r = repr(x)
    if "e" not in r and "." not in r and "inf" not in r:
        r += ".0"
    return r


def _is_small_int(x: float, bound: int = 4) -> bool:
    return float(x).is_integer() and abs(x) <= bound


class _Writer:
    def __init__(self):
        self.lines: List[str] = []
        self.ind = 0

    def __call__(self, s: str = ""):
        self.lines.append(("  " * self.ind + s) if s else "")

    def block(self, head: str):
        self(head + " {")
        self.ind += 1

    def end(self, tail: str = "}"):
        self.ind -= 1
        self(tail)

    def text(self) -> str:
        return "\n".join(self.lines) + "\n"


#: Code-shape choices of the fused kinetics kernels.
#:  pin   -- pin each species accumulator after every reaction (an empty asm
#:           with a read-write operand), so the compiler cannot sink the
#:           accumulation adds to the final stores and keep every rate live;
#:  kc    -- "species": K_c^-1 as a product of per-species exp(g'_k) factors
#:           (one exp per species instead of one per reaction); "exp": one
#:           exp of -ln K_c per reversible reaction;
#:  lazy  -- the omega-dot and JVP kernels open species (load C_k, compute
#:           exp(g'_k)) at their first use and store each rate after its last
#:           update, along a reaction order that keeps few species open
#:           (``reaction_order``); ``order_tries`` greedy restarts;
#:  wdot_waves / deriv_waves -- waves per SIMD the compiler must budget
#:           registers for in k_wdot / k_wdot_jvp (0: its own choice);
#:  lazy_fence -- an empty asm with a memory clobber every that many
#:           reactions of the lazy passes;
#:  prefetch -- the lazy omega-dot pass loads a species' mass fraction that
#:           many reactions before the species opens;
#:  wdot_block -- threads per block of k_wdot;
#:  recip1 -- the lazy omega-dot pass takes the reciprocal in K_c^-1 with
#:           one Newton step after v_rcp_f64 (relative error below 2.3e-15,
#:           measured) instead of two: 1.7 % faster (the same in the falloff
#:           terms changed the register allocation for the worse: 84 spills,
#:           11 % slower);
#:  ystash -- the lazy omega-dot pass keeps the mass fractions of that
#:           many species (the last to open) in LDS from its total-
#:           concentration pass instead of reading them from HBM again where
#:           they open: 39 of GRI-3.0's 53 (78 KB per 256-thread block, two
#:           blocks per CU) cut its HBM traffic from 26.8 to 22.6 GB per
#:           2^24 cells and its time by 1 % (53 leave one block per CU: 40 %
#:           slower);
#:  zform -- the lazy omega-dot pass keeps F_k = exp(-g'_k) and
#:           Z_k = C_k exp(g'_k) per species and writes a reversible rate as
#:           q = k_f prod_r F_r (prod_r Z_r - (RT/p0)^dnu prod_p Z_p): 53
#:           reciprocals instead of one per reversible reaction (309 for
#:           GRI-3.0), 4 % fewer VALU instructions, 6 % faster.
#:  kinv_f -- the lazy JVP / VJP passes keep F_k = 1 / E_k per open species
#:           and take K_c^-1 as a product of E (products) and F (reactants)
#:           without a reciprocal per reaction (off: the extra live values
#:           cost more than the reciprocals, JVP 9.9 -> 10.3 ms and VJP
#:           8.7 -> 9.0 ms on 2^22 GRI-3.0 cells);
#:  jac_groups -- groups of rows (balanced by work) in the source-Jacobian
#:           row pass: a large chunk runs every group in one thread (one
#:           wave per SIMD), a small one (a flamelet batch) gives each group
#:           its own wave for the same cells, so that more waves share the
#:           work (whole-chunk runs with split groups measured 14-25 %
#:           slower: the row accumulators, the concentrations and the
#:           hoisted cache loads outgrow 256 registers);
#:  jac_waves -- waves per SIMD the compiler must budget registers for in
#:           k_source_jacobian_rows (0: its own choice);
#:  jac_fence -- an empty asm with a memory clobber every that many
#:           reactions of a source-Jacobian row (0: none);
#:  jac_fused_groups -- row groups of the one-pass source Jacobian (the most
#:           waves that can share 64 cells);
#:  jac_fused_pin -- in the one-pass rows, the accumulators pass an empty asm
#:           every that many reactions (0: never; then the scheduler
#:           interleaves whole rows and spills: 2x slower);
#:  jac_fused_pin_ptrs -- the LDS table pointers pass the same asm, so later
#:           table reads stay below it (1, 2, 3 or 4 reactions per pin,
#:           with or without: within 1 % on 2^22 GRI-3.0 cells);
#:  jac_fused_ctable -- the one-pass rows read their constants from one
#:           deduplicated table through scalar loads (off: literals);
#:  nt_jac -- the one-pass source Jacobian writes its entries with streaming
#:           (non-temporal) stores when consecutive cells are adjacent in
#:           memory (entry-major): 36.9 -> 25.5 ms on 2^22 GRI-3.0 cells,
#:           the same values.  Per-cell blocks (the flamelet solver's layout)
#:           keep ordinary stores: there each lane's store lands on its own
#:           line, and streaming them made a GRI-3.0 flamelet solve about
#:           30 % slower;
#:  jac_stage -- per-cell blocks (each cell's row contiguous) pass each row
#:           through a per-wave LDS buffer, 8 entries of 64 cells at a time,
#:           so that a wave's store writes 8 cells x 64 contiguous bytes
#:           instead of touching 64 lines;
#:  jac_stage_pipe -- those rows pipelined: each chunk is read back from LDS
#:           while the next one is computed, and stored after that one's
#:           LDS writes;
#:  nt_wdot -- the same for omega-dot's rates (11.23 -> 11.12 ms on 2^24
#:           GRI-3.0 cells, timed side by side);
#:  nt_rq -- the same for the per-reaction outputs (k_f and the like: no
#:           change measured, off).
#: (GRI-3.0, 2^24 cells: 16.6 ms with the species arrays held in registers,
#: 12.1 ms lazy at two waves per SIMD, 11.35 ms with zform.)
DEFAULT_TUNING = {"pin": True, "pin_every": 3, "kc": "species", "ctable": True,
                  "lazy": True, "wdot_waves": 2, "deriv_waves": 0,
                  "lazy_fence": 1, "order_tries": 8, "prefetch": 4,
                  "zform": True, "jac_groups": 4, "ystash": 39}


def element_shift(mech: Mechanism):
    """Integer element potentials (mu_e, nu_e) that centre the shifted Gibbs
    functions g'_k = g_k - sum_e a_ek (mu_e / T + nu_e) near zero over
    300..3500 K (least squares, rounded to integers so every per-species
    shift is exact).  For an element-balanced reaction sum nu g' = sum nu g,
    so K_c is unchanged; exp(g'_k) stays far from overflow even at 300 K."""
    import numpy as np
    Ts = np.linspace(300.0, 3500.0, 33)
    rows, rhs = [], []
    A = mech.element_matrix().T  # (ns, ne)
    for s_i, sp in enumerate(mech.species):
        for t in Ts:
            c = sp.high if t > sp.t_mid else sp.low
            g = (c[0] * (1 - math.log(t)) - c[1] / 2 * t - c[2] / 6 * t ** 2
                 - c[3] / 12 * t ** 3 - c[4] / 20 * t ** 4 + c[5] / t - c[6])
            rows.append(np.concatenate([A[s_i] / t, A[s_i]]))
            rhs.append(g)
    sol = np.linalg.lstsq(np.array(rows), np.array(rhs), rcond=None)[0]
    ne = mech.n_elements
    mu = np.round(sol[:ne])
    nu = np.round(sol[ne:])
    return mu, nu


class _Lowering:
    """Lowers one mechanism; methods return C++ expressions or emit blocks."""

    def __init__(self, mech: Mechanism, tuning: Optional[dict] = None):
        self.m = mech
        self.ns = mech.n_species
        self.idx = {s.name: i for i, s in enumerate(mech.species)}
        self.tmids = sorted({s.t_mid for s in mech.species})
        self.tmid_flag = {t: f"hi{j}" for j, t in enumerate(self.tmids)}
        self.tuning = dict(DEFAULT_TUNING)
        self.tuning.update(tuning or {})
        mu, nu = element_shift(mech)
        A = mech.element_matrix().T
        self.shift_m = [float(A[k] @ mu) for k in range(self.ns)]
        self.shift_n = [float(A[k] @ nu) for k in range(self.ns)]
        E = mech.element_matrix()

        def balanced_integer(r):
            if r.orders:
                return False
            nus = [v for _, v in r.reactants] + [v for _, v in r.products]
            if not all(float(v).is_integer() and 0 < v <= 4 for v in nus):
                return False
            lhs = sum(E[:, self.idx[sp]] * v for sp, v in r.reactants)
            rhs = sum(E[:, self.idx[sp]] * v for sp, v in r.products)
            return bool((abs(lhs - rhs) < 1e-12).all())
        self.species_kc = (self.tuning["kc"] == "species" and all(
            balanced_integer(r) for r in mech.reactions if r.reversible))
        self.tab = None  # list of tabulated constants while emitting a table
        # value -> index into tab, when constants are shared while emitting
        self.tab_index = None
        # exp of the Arrhenius and Gibbs exponents (bounded arguments)
        self.fexp = "exp"
        # reciprocal in K_c^-1 (products of Gibbs factors)
        self.frecip = "pyro::recip"
        # K_c^-1 from E[] of the products and F[] = 1 / E[] of the
        # reactants (set while emitting the lazy derivative passes)
        self.kinv_f = False

    _INLINE = {0.0, 0.5, -0.5, 1.0, -1.0, 2.0, -2.0, 4.0, -4.0}

    def c(self, x: float) -> str:
        """A constant operand: a literal, or an entry of the current constant
        table (read through scalar loads) when one is being emitted."""
        x = float(x)
        if self.tab is None or x in self._INLINE or (
                x.is_integer() and -16 <= x <= 64):
            return lit(x)
        if self.tab_index is not None:
            j = self.tab_index.get(x)
            if j is None:
                j = self.tab_index[x] = len(self.tab)
                self.tab.append(x)
            return f"K[{j}]"
        self.tab.append(x)
        return f"K[{len(self.tab) - 1}]"

    # --- thermo ---------------------------------------------------------
    def _flags(self, w: _Writer, T: str = "T"):
        for t, f in self.tmid_flag.items():
            w(f"const bool {f} = {T} > {lit(t)};")

    def _sel(self, sp, a_hi: float, a_lo: float) -> str:
        if a_hi == a_lo:
            return lit(a_hi)
        return f"({self.tmid_flag[sp.t_mid]} ? {lit(a_hi)} : {lit(a_lo)})"

    @staticmethod
    def _horner(coeffs: Sequence[str], x: str) -> str:
        # coeffs[0] + x*(coeffs[1] + x*(...))
        expr = coeffs[-1]
        for c in reversed(coeffs[:-1]):
            expr = f"fma({x}, {expr}, {c})"
        return expr

    def emit_species_fn(self, w: _Writer, fname: str, args: str, gen):
        w(f"__device__ __forceinline__ void {fname}({args}, double (&o)[NS]) {{")
        w.ind += 1
        self._flags(w)
        for k, sp in enumerate(self.m.species):
            w(f"o[{k}] = {gen(sp)};  // {sp.name}")
        w.ind -= 1
        w("}")

    def cp_expr(self, sp) -> str:
        hi, lo = sp.high, sp.low
        cs = [self._sel(sp, hi[j], lo[j]) for j in range(5)]
        return self._horner(cs, "T")

    def h_expr(self, sp) -> str:
        hi, lo = sp.high, sp.low
        cs = [self._sel(sp, hi[0], lo[0])]
        for j, d in ((1, 2), (2, 3), (3, 4), (4, 5)):
            cs.append(self._sel(sp, hi[j] / d, lo[j] / d))
        poly = self._horner(cs, "T")
        return f"fma({self._sel(sp, hi[5], lo[5])}, invT, {poly})"

    def s_expr(self, sp) -> str:
        hi, lo = sp.high, sp.low
        cs = ["0.0", self._sel(sp, hi[1], lo[1])]
        for j, d in ((2, 2), (3, 3), (4, 4)):
            cs.append(self._sel(sp, hi[j] / d, lo[j] / d))
        # T*(a1 + T*(a2/2 + ...)) + a0*lnT + a6
        inner = self._horner(cs[1:], "T")
        return (f"fma({self._sel(sp, hi[0], lo[0])}, lnT, "
                f"fma(T, {inner}, {self._sel(sp, hi[6], lo[6])}))")

    def dcp_expr(self, sp) -> str:
        hi, lo = sp.high, sp.low
        cs = [self._sel(sp, hi[1], lo[1]), self._sel(sp, 2 * hi[2], 2 * lo[2]),
              self._sel(sp, 3 * hi[3], 3 * lo[3]),
              self._sel(sp, 4 * hi[4], 4 * lo[4])]
        return self._horner(cs, "T")

    def dh_expr(self, sp) -> str:
        hi, lo = sp.high, sp.low
        cs = [self._sel(sp, hi[1] / 2, lo[1] / 2),
              self._sel(sp, 2 * hi[2] / 3, 2 * lo[2] / 3),
              self._sel(sp, 3 * hi[3] / 4, 3 * lo[3] / 4),
              self._sel(sp, 4 * hi[4] / 5, 4 * lo[4] / 5)]
        return (f"fma({self._sel(sp, -hi[5], -lo[5])}, invT * invT, "
                f"{self._horner(cs, 'T')})")

    def ds_expr(self, sp) -> str:
        hi, lo = sp.high, sp.low
        cs = [self._sel(sp, hi[j], lo[j]) for j in (1, 2, 3, 4)]
        return f"fma({self._sel(sp, hi[0], lo[0])}, invT, {self._horner(cs, 'T')})"

    # --- shifted Gibbs exponentials --------------------------------------
    def emit_gibbs_exp(self, w: _Writer):
        """E[k] = exp(g'_k): both NASA ranges in the basis
        (1 - ln T, T, T^2, T^3, T^4, 1/T, 1), one select per species."""
        self.emit_gibbs_prologue(w)
        w("double E[NS];")
        for k in range(self.ns):
            self.emit_gibbs_one(w, k)

    def emit_gibbs_prologue(self, w: _Writer):
        w("const double omlnT = 1.0 - lnT;")
        w("const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;")
        self._flags(w)
        w("const double rtp = kGasConstant * T / kOneAtm;")
        w("const double prt = kOneAtm / (kGasConstant * T);")

    def emit_gibbs_one(self, w: _Writer, k: int):
        sp = self.m.species[k]

        def coeffs(c):
            return [c[0], -c[1] / 2, -c[2] / 6, -c[3] / 12, -c[4] / 20,
                    c[5] - self.shift_m[k], -(c[6] + self.shift_n[k])]
        basis = ["omlnT", "T", "T2", "T3", "T4", "invT"]

        def poly(cs):
            expr = self.c(cs[6])
            for cc, b in zip(reversed(cs[:6]), reversed(basis)):
                expr = f"fma({self.c(cc)}, {b}, {expr})"
            return expr
        hi, lo = coeffs(sp.high), coeffs(sp.low)
        if hi == lo:
            w(f"E[{k}] = {self.fexp}({poly(hi)});  // {sp.name}")
        else:
            # both ranges evaluated and pinned so the choice stays a
            # per-lane select rather than a divergent branch
            w(f"{{ double gh = {poly(hi)}, gl = {poly(lo)};")
            w('  asm volatile("" : "+v"(gh), "+v"(gl));')
            w(f"  E[{k}] = {self.fexp}({self.tmid_flag[sp.t_mid]} ? gh : gl);"
              f" }}  // {sp.name}")

    def kinv_expr(self, r: Reaction) -> str:
        """exp(-ln K_c) from the E[] factors (balanced, integer nu); with
        ``kinv_f`` the reactants' factors come from F[] = 1 / E[] and no
        reciprocal is taken per reaction."""
        def side(d, extra, arr="E"):
            ts = []
            for sp, nu in d:
                ts += [f"{arr}[{self.idx[sp]}]"] * int(nu)
            ts += extra
            return " * ".join(ts) if ts else "1.0"
        dnu = int(sum(nu for _, nu in r.products)
                  - sum(nu for _, nu in r.reactants))
        if self.kinv_f:
            num = side(r.products, ["rtp"] * dnu if dnu > 0 else [])
            den = side(r.reactants, ["prt"] * (-dnu) if dnu < 0 else [], "F")
            return f"({num}) * ({den})"
        num_extra = ["rtp"] * dnu if dnu > 0 else []
        den_extra = ["rtp"] * (-dnu) if dnu < 0 else []
        return (f"({side(r.products, num_extra)}) * "
                f"{self.frecip}({side(r.reactants, den_extra)})")

    def zform_terms(self, r: Reaction) -> str:
        """prod_r F_r * (prod_r Z_r - (RT/p0)^dnu prod_p Z_p), the rate of
        progress over k_f of a reversible reaction (balanced, integer nu)
        from the per-species F_k = exp(-g'_k), Z_k = C_k exp(g'_k)."""
        def side(d, arr, extra=()):
            ts = []
            for sp, nu in d:
                ts += [f"{arr}[{self.idx[sp]}]"] * int(nu)
            ts += list(extra)
            return " * ".join(ts) if ts else "1.0"
        dnu = int(sum(nu for _, nu in r.products)
                  - sum(nu for _, nu in r.reactants))
        ex = ["rtp"] * dnu if dnu > 0 else ["prt"] * (-dnu)
        return (f"({side(r.reactants, 'F')}) * ({side(r.reactants, 'Z')} - "
                f"{side(r.products, 'Z', ex)})")

    def pin(self, w: _Writer, names: List[str]):
        if self.tuning["pin"] and names:
            ops = ", ".join(f'"+v"({n})' for n in names)
            w(f'asm volatile("" : {ops});')

    # --- kinetics -------------------------------------------------------
    def arrhenius(self, rate) -> str:
        a, b, ta = rate.A, rate.b, rate.t_a
        c = self.c
        if ta == 0:
            if b == 0:
                return c(a)
            if _is_small_int(b):
                if b < 0:
                    return f"({c(a)} * pyro::ipow<{int(-b)}>(invT))"
                return f"({c(a)} * pyro::ipow<{int(b)}>(T))"
            return f"({c(a)} * {self.fexp}({c(b)} * lnT))"
        lna = math.log(a)
        if b == 0:
            return f"{self.fexp}(fma({c(-ta)}, invT, {c(lna)}))"
        return (f"{self.fexp}(fma({c(b)}, lnT, fma({c(-ta)}, invT, "
                f"{c(lna)})))")

    def rate_ratio(self, lo, hi) -> str:
        """k_lo / k_hi as one expression: a constant times an integer power
        of T when both are pure power laws, else one exp of the difference
        of the Arrhenius exponents."""
        db = lo.b - hi.b
        if lo.t_a == 0 and hi.t_a == 0 and _is_small_int(db):
            c = self.c(lo.A / hi.A)
            if db == 0:
                return c
            if db < 0:
                return f"({c} * pyro::ipow<{int(-db)}>(invT))"
            return f"({c} * pyro::ipow<{int(db)}>(T))"
        dlna = math.log(lo.A) - math.log(hi.A)
        dta = lo.t_a - hi.t_a
        return (f"{self.fexp}(fma({self.c(db)}, lnT, fma({self.c(-dta)}, "
                f"invT, {self.c(dlna)})))")

    def third_body(self, r: Reaction, C: str = "C", Ctot: str = "Ctot") -> str:
        """[M] = d Ctot + sum_k (eff_k - d) C_k (chem_expr.py:327-345); with
        ``C``/``Ctot`` naming tangent arrays it gives the tangent of [M]."""
        d = r.default_efficiency
        terms = []
        for sp, eff in r.efficiencies:
            coef = eff - d
            if coef != 0.0:
                terms.append((coef, self.idx[sp]))
        expr = "0.0" if d == 0 else (Ctot if d == 1.0 else f"({self.c(d)} * {Ctot})")
        for coef, k in terms:
            expr = f"fma({self.c(coef)}, {C}[{k}], {expr})"
        return expr

    def efficiency_terms(self, r: Reaction) -> List[Tuple[int, float]]:
        """(k, eff_k - d) for the explicit colliders of a three-body or
        falloff reaction."""
        d = r.default_efficiency
        return [(self.idx[sp], e - d) for sp, e in r.efficiencies
                if e - d != 0.0]

    def uses_ctot(self) -> bool:
        return any(r.default_efficiency != 0 for r in self.m.reactions
                   if r.kind in ("three-body", "falloff"))

    def kf_block(self, w: _Writer, r: Reaction, out: str = "kf"):
        """Emit statements defining ``out`` = forward rate coefficient with
        falloff and three-body factors (codegen/python.py:398-473)."""
        if r.kind == "falloff":
            w(f"const double khi = {self.arrhenius(r.high_rate)};")
            w(f"const double pr = ({self.third_body(r)}) * "
              f"{self.rate_ratio(r.low_rate, r.high_rate)};")
            if r.falloff_type == "Troe":
                tp = r.troe
                c = self.c
                t1 = f"({c(1.0 - tp[0])} * exp(T * {c(-1.0 / tp[1])}))"
                t2 = f"({c(tp[0])} * exp(T * {c(-1.0 / tp[2])}))"
                s = f"{t1} + {t2}"
                if len(tp) == 4:
                    s += f" + exp({c(-tp[3])} * invT)"
                w(f"const double fc = log10({s});")
                w("const double lpr = log10(pr);")
                w("const double cc = fma(-0.67, fc, -0.4);")
                w("const double nn = fma(-1.27, fc, 0.75);")
                w("const double f1 = pr > 0.0 ? (lpr + cc) * pyro::recip("
                  f"fma(-0.14, lpr + cc, nn)) : {lit(-1 / 0.14)};")
                w("const double fF = exp10(fc * pyro::recip(fma(f1, f1, 1.0)));")
                w(f"const double {out} = khi * (fF * pr * pyro::recip(1.0 + pr));")
            else:
                w(f"const double {out} = khi * (pr * pyro::recip(1.0 + pr));")
        elif r.kind == "three-body":
            w(f"const double {out} = {self.arrhenius(r.rate)} * ({self.third_body(r)});")
        else:
            w(f"const double {out} = {self.arrhenius(r.rate)};")

    # --- derivatives (forward mode, hand-written per reaction kind) ---------
    def dlnk(self, rate) -> str:
        """d ln k / dT = (b + Ta/T) / T."""
        b, ta = rate.b, rate.t_a
        if ta == 0:
            return "0.0" if b == 0 else f"({self.c(b)} * invT)"
        return f"(invT * fma({self.c(ta)}, invT, {self.c(b)}))"

    def kf_block_deriv(self, w: _Writer, r: Reaction):
        """Emit kf, dkfT = dkf/dT and (three-body, falloff) M and
        dkfM = dkf/d[M] (chem_expr.py:310-421 differentiated by hand)."""
        c = self.c
        if r.kind == "falloff":
            ln10 = math.log(10.0)
            w(f"const double khi = {self.arrhenius(r.high_rate)};")
            w(f"const double ahi = {self.dlnk(r.high_rate)};")
            w(f"const double alo = {self.dlnk(r.low_rate)};")
            w(f"const double ratio = {self.rate_ratio(r.low_rate, r.high_rate)};")
            w(f"const double M = {self.third_body(r)};")
            w("const double pr = M * ratio;")
            w("const double rpr1 = pyro::recip(1.0 + pr);")
            if r.falloff_type == "Troe":
                tp = r.troe
                w(f"const double e1 = exp(T * {c(-1.0 / tp[1])});")
                w(f"const double e2 = exp(T * {c(-1.0 / tp[2])});")
                S = f"fma({c(1.0 - tp[0])}, e1, {c(tp[0])} * e2)"
                dS = (f"fma({c((1.0 - tp[0]) * (-1.0 / tp[1]))}, e1, "
                      f"{c(tp[0] * (-1.0 / tp[2]))} * e2)")
                if len(tp) == 4:
                    w(f"const double e3 = exp({c(-tp[3])} * invT);")
                    S = f"({S} + e3)"
                    dS = f"fma({c(tp[3])} * invT * invT, e3, {dS})"
                w(f"const double sS = {S};")
                w(f"const double fc = log10(sS);")
                w(f"const double dfc = {dS} * pyro::recip(sS) * {c(1 / ln10)};")
                w("const double lpr = log10(pr);")
                w("const double cc = fma(-0.67, fc, -0.4);")
                w("const double nn = fma(-1.27, fc, 0.75);")
                w("const double u = lpr + cc;")
                w("const double rden = pyro::recip(fma(-0.14, u, nn));")
                w("const bool pos = pr > 0.0;")
                w(f"const double f1 = pos ? u * rden : {lit(-1 / 0.14)};")
                w("const double ropf = pyro::recip(fma(f1, f1, 1.0));")
                w("const double fF = exp10(fc * ropf);")
                w("const double dlF_df1 = -2.0 * fc * f1 * ropf * ropf;")
                w("const double df1_du = pos ? nn * rden * rden : 0.0;")
                w("const double df1_dn = pos ? -u * rden * rden : 0.0;")
                w("const double dlF_dL = dlF_df1 * df1_du;")
                w("const double dlF_dfc = fma(dlF_df1, fma(-0.67, df1_du, "
                  "-1.27 * df1_dn), ropf);")
                w(f"const double dlF_dT = fma(dlF_dL * {c(1 / ln10)}, alo - ahi, "
                  "dlF_dfc * dfc);")
                w("const double kf = khi * (fF * pr * rpr1);")
                w("const double dkfM = khi * fF * ratio * fma(dlF_dL, rpr1, "
                  "rpr1 * rpr1);")
                w(f"const double dkfT = kf * (ahi + fma({c(ln10)}, dlF_dT, "
                  "(alo - ahi) * rpr1));")
            else:
                w("const double kf = khi * (pr * rpr1);")
                w("const double dkfM = khi * ratio * rpr1 * rpr1;")
                w("const double dkfT = kf * fma(alo - ahi, rpr1, ahi);")
        elif r.kind == "three-body":
            w(f"const double k0 = {self.arrhenius(r.rate)};")
            w(f"const double M = {self.third_body(r)};")
            w("const double kf = k0 * M;")
            w("const double dkfM = k0;")
            w(f"const double dkfT = kf * {self.dlnk(r.rate)};")
        else:
            w(f"const double kf = {self.arrhenius(r.rate)};")
            w(f"const double dkfT = kf * {self.dlnk(r.rate)};")

    def prod_partials(self, pairs) -> List[Tuple[int, str]]:
        """(k, d prod / d C_k) for prod = prod_j C_j^nu_j, written without
        division so it stays finite when a concentration is zero."""
        out = []
        for i, (k, nu) in enumerate(pairs):
            others = [self._pow_term(kk, nn) for j, (kk, nn) in enumerate(pairs)
                      if j != i]
            if nu == 1.0:
                d = []
            elif _is_small_int(nu) and nu > 0:
                d = [lit(nu)] + [f"C[{k}]"] * (int(nu) - 1)
            else:
                d = [lit(nu), f"pow(C[{k}], {lit(nu - 1.0)})"]
            terms = d + others
            out.append((k, " * ".join(terms) if terms else "1.0"))
        return out

    def _pow_term(self, k: int, nu: float) -> str:
        if nu == 1.0:
            return f"C[{k}]"
        if _is_small_int(nu) and nu > 0:
            return "(" + " * ".join([f"C[{k}]"] * int(nu)) + ")"
        return f"pow(C[{k}], {lit(nu)})"

    def reactant_pairs(self, r: Reaction) -> List[Tuple[int, float]]:
        ind = [self.idx[sp] for sp, _ in r.reactants]
        if r.orders:
            nus = [o for _, o in r.orders]  # zip quirk, chem_expr.py:436-448
        else:
            nus = [nu for _, nu in r.reactants]
        return list(zip(ind, nus))

    def prod_expr(self, pairs) -> str:
        terms = [self._pow_term(k, nu) for k, nu in pairs]
        return " * ".join(terms) if terms else "1.0"

    def nlk_expr(self, r: Reaction) -> str:
        def side(d):
            ts = []
            for sp, nu in d:
                g = f"g[{self.idx[sp]}]"
                ts.append(g if nu == 1.0 else f"{lit(nu)} * {g}")
            return "(" + " + ".join(ts) + ")" if ts else "0.0"
        expr = f"{side(r.products)} - {side(r.reactants)}"
        dnu = sum(nu for _, nu in r.products) - sum(nu for _, nu in r.reactants)
        if dnu != 0:
            expr = f"fma({lit(-dnu)}, c0, {expr})"
        return expr

    def species_updates(self, r: Reaction, q: str) -> List[Tuple[int, float]]:
        """Net stoichiometric updates: production_rate_expr sums
        nu'' over products minus nu' over reactants."""
        upd: Dict[int, float] = {}
        for sp, nu in r.reactants:
            upd[self.idx[sp]] = upd.get(self.idx[sp], 0.0) - nu
        for sp, nu in r.products:
            upd[self.idx[sp]] = upd.get(self.idx[sp], 0.0) + nu
        return [(k, v) for k, v in upd.items() if v != 0.0]


def _emit_prelude(w: _Writer, L: _Lowering, name: str, digest: str,
                  wdot_block: int):
    m = L.m
    w(f"// Generated by pyrometheus_amd.codegen.hip for mechanism '{name}'")
    w(f"// source: {m.source}; digest {digest}")
    w(f"// {m.n_species} species, {m.n_reactions} reactions "
      f"({m.num_falloff} falloff, {len(m.three_body_reactions)} three-body)")
    w("// Do not edit: regenerate with HipCodeGenerator.generate().")
    w('#include "pyro_common.hpp"')
    w("")
    w("namespace mech {")
    w(f"constexpr int NS = {m.n_species};")
    w(f"constexpr int NR = {m.n_reactions};")
    w(f"constexpr int NF = {m.num_falloff};")
    w(f"constexpr int kAbiVersion = {ABI_VERSION};")
    w(f"constexpr int kWdotBlock = {wdot_block};")
    w(f"constexpr int kWdotWaves = {int(L.tuning.get('wdot_waves', 0))};")
    w(f"constexpr int kDerivWaves = {int(L.tuning.get('deriv_waves', 0))};")
    w(f"constexpr int kTempWaves = {int(L.tuning.get('temp_waves', 0))};")
    w(f"constexpr int kJacWaves = {int(L.tuning.get('jac_waves', 0))};")
    w(f"constexpr int kJacGroups = {int(L.tuning.get('jac_groups', 1))};")
    w("__device__ constexpr double kW[NS] = {"
      + ", ".join(lit(x) for x in m.molecular_weights) + "};")
    w("__device__ constexpr double kInvW[NS] = {"
      + ", ".join(lit(1.0 / x) for x in m.molecular_weights) + "};")
    # the same values where a kernel must read them from memory (scalar
    # loads) rather than have them folded into immediates
    w("__constant__ double kInvWMem[NS] = {"
      + ", ".join(lit(1.0 / x) for x in m.molecular_weights) + "};")
    w("using pyro::kGasConstant;")
    w("using pyro::kOneAtm;")
    w("")


def _emit_thermo(w: _Writer, L: _Lowering):
    L.emit_species_fn(w, "species_cp_r", "const double T", L.cp_expr)
    L.emit_species_fn(w, "species_h_rt", "const double T, const double invT",
                      L.h_expr)
    L.emit_species_fn(w, "species_s_r", "const double T, const double lnT",
                      L.s_expr)
    L.emit_species_fn(w, "species_dcp_r", "const double T", L.dcp_expr)
    L.emit_species_fn(w, "species_dh_rt", "const double T, const double invT",
                      L.dh_expr)
    L.emit_species_fn(w, "species_ds_r", "const double T, const double invT",
                      L.ds_expr)
    # mixture NASA-7 coefficients of a cell, for the temperature Newton:
    # sum_k yw_k h_k/RT and sum_k yw_k cp_k/R are polynomials in T whose
    # coefficients M_j = sum_k yw_k a_kj do not change between iterations.
    # Species are grouped by T_mid (the range flag is the same for all the
    # species of a group); single-range species join group 0 in both
    # ranges.  The pre-pass reads each species' coefficients from a constant
    # table through a pointer passed through an empty asm (scalar loads
    # where they are used); an iteration then costs one polynomial per group.
    tm = sorted({sp.t_mid for sp in L.m.species if sp.high != sp.low}) or [None]
    gi = {t: j for j, t in enumerate(tm)}
    ng = len(tm)
    table_at = len(w.lines)
    w(f"constexpr int kMixGroups = {ng};")
    w("__device__ __forceinline__ void mixture_nasa_coeffs("
      "const double (&yw)[NS], double (&M)[kMixGroups][2][6], double& syw) {")
    w.ind += 1
    L.tab = []
    w("pyro::cptr K = (pyro::cptr)kTabMixHCp;")
    w("#pragma unroll")
    w("for (int g = 0; g < kMixGroups; ++g)")
    w("#pragma unroll")
    w("  for (int r = 0; r < 2; ++r)")
    w("#pragma unroll")
    w("    for (int j = 0; j < 6; ++j) M[g][r][j] = 0.0;")
    w("syw = 0.0;")
    for k, sp in enumerate(L.m.species):
        g = 0 if sp.high == sp.low else gi[sp.t_mid]
        w(f"{{  // {sp.name}")
        w('  asm volatile("" : "+s"(K));')
        w(f"  const double y = yw[{k}];")
        w("  syw += y;")
        for r, c in ((0, sp.high), (1, sp.low)):
            for j in range(6):
                w(f"  M[{g}][{r}][{j}] = fma(y, {L.c(c[j])}, M[{g}][{r}][{j}]);")
        w("}")
    w.ind -= 1
    w("}")
    tab = L.tab or [0.0]
    w.lines.insert(table_at, "__constant__ double kTabMixHCp[] = {"
                   + ", ".join(lit(x) for x in tab) + "};")
    L.tab = None
    # sum_k yw_k (h_k/RT - shift) and sum_k yw_k (cp_k/R - shift) at T from
    # the mixture coefficients: h/RT = a0 + a1 T/2 + a2 T^2/3 + a3 T^3/4 +
    # a4 T^4/5 + a5/T, cp/R = a0 + a1 T + a2 T^2 + a3 T^3 + a4 T^4
    w("__device__ __forceinline__ void mixture_h_cp_eval(const double T, "
      "const double invT, const double (&M)[kMixGroups][2][6], "
      "const double syw, const double shift, double& sh, double& sc) {")
    w.ind += 1
    L._flags(w)
    w("double h = 0.0, c = 0.0;")
    for g, t in enumerate(tm):
        sel = L.tmid_flag[t] if t is not None else "true"
        w("{")
        w.ind += 1
        w(f"const bool hi = {sel};")
        w("double m[6];")
        w("#pragma unroll")
        w(f"for (int j = 0; j < 6; ++j) m[j] = hi ? M[{g}][0][j] : M[{g}][1][j];")
        w("h += fma(m[5], invT, fma(T, fma(T, fma(T, fma(T, 0.2 * m[4], "
          "0.25 * m[3]), (1.0 / 3.0) * m[2]), 0.5 * m[1]), m[0]));")
        w("c += fma(T, fma(T, fma(T, fma(T, m[4], m[3]), m[2]), m[1]), m[0]);")
        w.ind -= 1
        w("}")
    w("sh = fma(-shift, syw, h);")
    w("sc = fma(-shift, syw, c);")
    w.ind -= 1
    w("}")
    w("__device__ __forceinline__ void species_g_rt(const double T, "
      "const double lnT, const double invT, double (&o)[NS]) {")
    w.ind += 1
    w("double h[NS], s[NS];")
    w("species_h_rt(T, invT, h);")
    w("species_s_r(T, lnT, s);")
    w("#pragma unroll")
    w("for (int k = 0; k < NS; ++k) o[k] = h[k] - s[k];")
    w.ind -= 1
    w("}")
    w("")


def _kin_prologue(w: _Writer, L: _Lowering, need_g: bool):
    w("const double lnT = log(T);")
    w("const double invT = 1.0 / T;")
    if need_g:
        w("const double c0 = log(kOneAtm / (kGasConstant * T));")
        w("double g[NS];")
        w("species_g_rt(T, lnT, invT, g);")
    if L.uses_ctot():
        w("double Ctot = 0.0;")
        w("#pragma unroll")
        w("for (int k = 0; k < NS; ++k) Ctot += C[k];")


def _emit_wdot(w: _Writer, L: _Lowering):
    m = L.m
    w("// fused net production rates (codegen/python.py:475-495)")
    table_at = len(w.lines)
    w("__device__ __forceinline__ void net_production_rates(const double T, "
      "const double (&C)[NS], double (&w)[NS]) {")
    w.ind += 1
    if L.tuning.get("ctable", True):
        L.tab = []
        # constants are read from a table in constant memory through a
        # pointer the optimiser cannot see through: uniform scalar loads
        # replace per-use literal materialisation
        w("pyro::cptr K = (pyro::cptr)kTabWdot;")
        w('asm volatile("" : "+s"(K));')
    rev = any(r.reversible for r in m.reactions)
    if L.species_kc:
        _kin_prologue(w, L, need_g=False)
        if rev:
            L.emit_gibbs_exp(w)
    else:
        _kin_prologue(w, L, need_g=rev)
    w("#pragma unroll")
    w("for (int k = 0; k < NS; ++k) w[k] = 0.0;")
    pending = set()
    for i, r in enumerate(m.reactions):
        w(f"{{  // R{i}: {r.equation} [{r.reaction_type}]")
        w.ind += 1
        L.kf_block(w, r)
        rf = L.prod_expr(L.reactant_pairs(r))
        if r.reversible:
            rr = L.prod_expr([(L.idx[sp], nu) for sp, nu in r.products])
            kinv = L.kinv_expr(r) if L.species_kc else f"exp({L.nlk_expr(r)})"
            w(f"const double q = kf * ({rf} - {kinv} * {rr});")
        else:
            w(f"const double q = kf * ({rf});")
        upd = L.species_updates(r, "q")
        for k, nu in upd:
            if nu == 1.0:
                w(f"w[{k}] += q;")
            elif nu == -1.0:
                w(f"w[{k}] -= q;")
            else:
                w(f"w[{k}] = fma({lit(nu)}, q, w[{k}]);")
        pending.update(k for k, _ in upd)
        if (i + 1) % L.tuning.get("pin_every", 1) == 0 or i + 1 == len(m.reactions):
            L.pin(w, [f"w[{k}]" for k in sorted(pending)])
            pending.clear()
        w.ind -= 1
        w("}")
    w.ind -= 1
    w("}")
    w("")
    if L.tab is not None:
        tab = L.tab or [0.0]
        decl = ("__constant__ double kTabWdot[] = {"
                + ", ".join(lit(x) for x in tab) + "};")
        w.lines.insert(table_at, decl)
        L.tab = None


def species_uses(L: _Lowering, r: Reaction):
    """(concentrations read, Gibbs factors read, rates written) by reaction
    ``r`` in the fused kernel, as sets of species indices."""
    c = {L.idx[sp] for sp, _ in r.reactants}
    e, w = set(), {k for k, _ in L.species_updates(r, "q")}
    if r.reversible:
        c |= {L.idx[sp] for sp, _ in r.products}
        if L.species_kc:
            e = {L.idx[sp] for sp, _ in list(r.reactants) + list(r.products)}
    if r.kind in ("three-body", "falloff"):
        c |= {k for k, _ in L.efficiency_terms(r)}
    return c, e, w


def _peak_live(uses, order) -> int:
    """Largest number of species values (concentrations, Gibbs factors,
    rates) open at once along ``order``."""
    first, last = {}, {}
    for pos, ri in enumerate(order):
        for which, ks in enumerate(uses[ri]):
            for k in ks:
                first.setdefault((which, k), pos)
                last[(which, k)] = pos
    delta = [0] * (len(order) + 1)
    for key, a in first.items():
        delta[a] += 1
        delta[last[key] + 1] -= 1
    peak = run = 0
    for d in delta:
        run += d
        peak = max(peak, run)
    return peak


def _greedy_order(uses, rank) -> List[int]:
    left: Dict[Tuple[int, int], int] = {}
    for u in uses:
        for which, ks in enumerate(u):
            for k in ks:
                left[(which, k)] = left.get((which, k), 0) + 1
    opened = set()
    remaining = list(range(len(uses)))
    order = []
    while remaining:
        best, best_key = None, None
        for ri in remaining:
            score = 0
            for which, ks in enumerate(uses[ri]):
                for k in ks:
                    score += (which, k) not in opened
                    score -= left[(which, k)] == 1
            key = (score, rank[ri])
            if best_key is None or key < best_key:
                best, best_key = ri, key
        remaining.remove(best)
        order.append(best)
        for which, ks in enumerate(uses[best]):
            for k in ks:
                opened.add((which, k))
                left[(which, k)] -= 1
    return order


def reaction_order(L: _Lowering, tries: int = 8, uses=None) -> List[int]:
    """An order of the reactions that keeps few species live at once: a
    species' concentration and Gibbs factor are opened by the first reaction
    that reads them and its rate is closed (stored) by the last that writes
    it.  Greedy: take the reaction that opens the fewest new values net of
    the values it closes, ties broken by a fixed ranking of the reactions
    (their index, then seeded permutations); the order with the smallest
    peak is kept.  The net production rates are sums over reactions, so the
    order changes only their rounding."""
    import random
    if uses is None:
        uses = [species_uses(L, r) for r in L.m.reactions]
    n = len(uses)
    best, best_peak = None, None
    for t in range(tries):
        rank = list(range(n))
        if t:
            random.Random(t).shuffle(rank)
        order = _greedy_order(uses, rank)
        peak = _peak_live(uses, order)
        if best_peak is None or peak < best_peak:
            best, best_peak = order, peak
    return best


def _emit_wdot_lazy(w: _Writer, L: _Lowering):
    """Net production rates of one cell read from and written to memory
    directly: reactions in ``reaction_order``, each species' concentration
    loaded (and Gibbs factor computed) just before its first use and its
    rate stored right after its last update, so that only the species of a
    sliding window occupy registers."""
    m = L.m
    w("// net production rates, species opened and closed along the "
      "reactions (codegen/python.py:475-495)")
    sig = ("__device__ __forceinline__ void net_production_rates_lazy("
           "const double rho, const double T, const double* __restrict__ Y, "
           "const int64_t ldY, const int64_t i, double* __restrict__ W, "
           "const int64_t ldW) {")
    rev = any(r.reversible for r in m.reactions)
    if not L.tuning.get("lazy", True) or (rev and not L.species_kc):
        w(sig)
        w.ind += 1
        w("double C[NS], w[NS];")
        w("#pragma unroll")
        w("for (int k = 0; k < NS; ++k) C[k] = kInvW[k] * rho * Y[k * ldY + i];")
        w("net_production_rates(T, C, w);")
        w("#pragma unroll")
        w("for (int k = 0; k < NS; ++k) W[k * ldW + i] = w[k];")
        w.ind -= 1
        w("}")
        w("")
        return
    table_at = len(w.lines)
    w(sig)
    w.ind += 1
    if L.tuning.get("recip1", True):
        L.frecip = "pyro::recip1"
    if L.tuning.get("ctable", True):
        L.tab = []
        w("pyro::cptr K = (pyro::cptr)kTabWdotLazy;")
        w('asm volatile("" : "+s"(K));')
    w("const double lnT = log(T);")
    w("const double invT = 1.0 / T;")
    if rev:
        L.emit_gibbs_prologue(w)
    order = reaction_order(L, L.tuning.get("order_tries", 8))
    uses = [species_uses(L, r) for r in m.reactions]
    # ystash: the mass fractions of that many species, the last to open,
    # kept in LDS by the total-concentration pass for their opening (one
    # 8-byte slot per thread of the block per species)
    first_use = {}
    for pos, ri in enumerate(order):
        for k in uses[ri][0]:
            first_use.setdefault(k, pos)
    nst = int(L.tuning.get("ystash", 0)) if L.uses_ctot() else 0
    stash = {k: n for n, k in enumerate(sorted(
        first_use, key=lambda k: -first_use[k])[:nst])}
    if stash:
        w(f"__shared__ double ys[{len(stash)}][kWdotBlock];")
        w("double* ysw = &ys[0][threadIdx.x];")
    if L.uses_ctot():
        # the total concentration needs every species up front; the values
        # are loaded again where each species opens, from LDS for the
        # stashed ones (each stored beside its term of the sum: as a
        # separate pass after the sum it measured 2 % slower)
        w("double Ctot = 0.0;")
        if not stash:
            w("#pragma unroll")
            w("for (int k = 0; k < NS; ++k) "
              "Ctot += kInvW[k] * rho * Y[k * ldY + i];")
        for k in range(L.ns) if stash else ():
            if k in stash:
                w(f"{{ const double y = Y[{k} * ldY + i]; "
                  f"Ctot += kInvW[{k}] * rho * y; "
                  f"ysw[{stash[k]} * kWdotBlock] = y; }}")
            else:
                w(f"Ctot += kInvW[{k}] * rho * Y[{k} * ldY + i];")
    if stash:
        w("pyro::lptr ysr = (pyro::lptr)ysw;")
    # Yg: Y as a global-memory pointer that is passed through an empty asm
    # before each window opens species, so the optimiser cannot reuse the
    # values of the pass above or hoist the loads of later windows
    w("pyro::gptr Yg = (pyro::gptr)Y;")
    w("double C[NS], E[NS], w[NS];")
    # zform: each species opened with a Gibbs factor also gets
    # F_k = 1 / E_k and Z_k = C_k E_k, so that a reversible reaction reads
    #   q = k_f prod_r F_r (prod_r Z_r - (RT/p0)^dnu prod_p Z_p)
    # (= k_f (prod_r C_r - K_c^-1 prod_p C_p)): one reciprocal per species
    # instead of one per reaction
    zform = rev and L.tuning.get("zform", False)
    if zform:
        w("double F[NS], Z[NS];")
    last_w = {}
    for pos, ri in enumerate(order):
        for k in uses[ri][2]:
            last_w[k] = pos
    open_c, open_e, open_w = set(), set(), set()
    pending = set()
    fence = L.tuning.get("lazy_fence", 4)
    # mass fractions are loaded ``prefetch`` reactions before the species
    # opens, so the load's latency overlaps the reactions in between
    dist = int(L.tuning.get("prefetch", 0))
    opens_at = {}
    seen = set()
    for pos, ri in enumerate(order):
        for k in sorted(uses[ri][0] - seen):
            opens_at.setdefault(max(0, pos - dist), []).append(k)
        seen |= uses[ri][0]
    w("double yv[NS];")
    for pos, ri in enumerate(order):
        r = m.reactions[ri]
        c_u, e_u, w_u = uses[ri]
        if pos in opens_at:
            w('asm volatile("" : "+s"(Yg));')
            if stash:
                w('asm volatile("" : "+v"(ysr));')
            for k in opens_at[pos]:
                if k in stash:
                    w(f"yv[{k}] = ysr[{stash[k]} * kWdotBlock];")
                else:
                    w(f"yv[{k}] = Yg[{k} * ldY + i];")
        for k in sorted(c_u - open_c):
            w(f"C[{k}] = kInvW[{k}] * rho * yv[{k}];")
        for k in sorted(e_u - open_e):
            L.emit_gibbs_one(w, k)
            if zform:
                w(f"F[{k}] = pyro::recip(E[{k}]);")
                w(f"Z[{k}] = C[{k}] * E[{k}];")
        for k in sorted(w_u - open_w):
            w(f"w[{k}] = 0.0;")
        open_c |= c_u
        open_e |= e_u
        open_w |= w_u
        w(f"{{  // R{ri}: {r.equation} [{r.reaction_type}]")
        w.ind += 1
        L.kf_block(w, r)
        rf = L.prod_expr(L.reactant_pairs(r))
        if r.reversible and zform:
            w(f"const double q = kf * ({L.zform_terms(r)});")
        elif r.reversible:
            rr = L.prod_expr([(L.idx[sp], nu) for sp, nu in r.products])
            w(f"const double q = kf * ({rf} - {L.kinv_expr(r)} * {rr});")
        else:
            w(f"const double q = kf * ({rf});")
        upd = L.species_updates(r, "q")
        for k, nu in upd:
            if nu == 1.0:
                w(f"w[{k}] += q;")
            elif nu == -1.0:
                w(f"w[{k}] -= q;")
            else:
                w(f"w[{k}] = fma({lit(nu)}, q, w[{k}]);")
        pending.update(k for k, _ in upd)
        w.ind -= 1
        w("}")
        done = sorted(k for k in w_u if last_w[k] == pos)
        if (pos + 1) % L.tuning.get("pin_every", 1) == 0 or done \
                or pos + 1 == len(order):
            L.pin(w, [f"w[{k}]" for k in sorted(pending)])
            pending.clear()
        for k in done:
            if L.tuning.get("nt_wdot", True):
                w(f"__builtin_nontemporal_store(w[{k}], &W[{k} * ldW + i]);")
            else:
                w(f"W[{k} * ldW + i] = w[{k}];")
        if fence and (pos + 1) % fence == 0:
            # keep the loads and stores of each window in place
            w('asm volatile("" ::: "memory");')
    for k in range(L.ns):
        if k not in open_w:
            w(f"W[{k} * ldW + i] = 0.0;")
    w.ind -= 1
    w("}")
    w("")
    L.frecip = "pyro::recip"
    if L.tab is not None:
        tab = L.tab or [0.0]
        decl = ("__constant__ double kTabWdotLazy[] = {"
                + ", ".join(lit(x) for x in tab) + "};")
        w.lines.insert(table_at, decl)
        L.tab = None


def _emit_wdot_jac(w: _Writer, L: _Lowering):
    """omega-dot with dw/dC (dense NS x NS, registers) and dw/dT."""
    m = L.m
    w("// omega-dot and its derivatives: J[k][j] = d w_k / d C_j, "
      "JT[k] = d w_k / d T")
    w("__device__ __forceinline__ void net_production_rates_jac(const double T, "
      "const double (&C)[NS], double (&w)[NS], double (&J)[NS][NS], "
      "double (&JT)[NS]) {")
    w.ind += 1
    rev = any(r.reversible for r in m.reactions)
    if L.species_kc:
        _kin_prologue(w, L, need_g=False)
        if rev:
            L.emit_gibbs_exp(w)
    else:
        _kin_prologue(w, L, need_g=rev)
    if rev:
        w("double hrt[NS];")
        w("species_h_rt(T, invT, hrt);")
    w("double Dm[NS];")
    w("#pragma unroll")
    w("for (int k = 0; k < NS; ++k) {")
    w("  w[k] = 0.0; JT[k] = 0.0; Dm[k] = 0.0;")
    w("#pragma unroll")
    w("  for (int j = 0; j < NS; ++j) J[k][j] = 0.0;")
    w("}")
    for i, r in enumerate(m.reactions):
        w(f"{{  // R{i}: {r.equation} [{r.reaction_type}]")
        w.ind += 1
        L.kf_block_deriv(w, r)
        rpairs = L.reactant_pairs(r)
        w(f"const double pf = {L.prod_expr(rpairs)};")
        dpf = dict(L.prod_partials(rpairs))
        dpb = {}
        if r.reversible:
            ppairs = [(L.idx[sp], nu) for sp, nu in r.products]
            w(f"const double pb = {L.prod_expr(ppairs)};")
            dpb = dict(L.prod_partials(ppairs))
            kinv = L.kinv_expr(r) if L.species_kc else f"exp({L.nlk_expr(r)})"
            w(f"const double kinv = {kinv};")
            dnu = sum(nu for _, nu in r.products) - sum(nu for _, nu in r.reactants)
            hs = [f"{lit(-nu)} * hrt[{L.idx[sp]}]" for sp, nu in r.products]
            hs += [f"{lit(nu)} * hrt[{L.idx[sp]}]" for sp, nu in r.reactants]
            w(f"const double dkinvT = kinv * invT * ({lit(dnu)} + "
              + " + ".join(hs) + ");")
            w("const double delta = fma(-kinv, pb, pf);")
            w("const double dqT = fma(dkfT, delta, -kf * dkinvT * pb);")
        else:
            w("const double delta = pf;")
            w("const double dqT = dkfT * delta;")
        w("const double q = kf * delta;")
        mdep = r.kind in ("three-body", "falloff")
        if mdep:
            w("const double dqM = dkfM * delta;")
        cols = sorted(set(dpf) | set(dpb))
        for j in cols:
            a = dpf.get(j)
            b = dpb.get(j)
            if a is not None and b is not None:
                w(f"const double dq{j} = kf * fma(-kinv, {b}, {a});")
            elif a is not None:
                w(f"const double dq{j} = kf * ({a});")
            else:
                w(f"const double dq{j} = -kf * kinv * ({b});")
        eff = []
        if mdep:
            d = r.default_efficiency
            eff = [(L.idx[sp], e - d) for sp, e in r.efficiencies if e - d != 0.0]
        for k, nu in L.species_updates(r, "q"):
            nl = lit(nu)
            w(f"w[{k}] = fma({nl}, q, w[{k}]);")
            w(f"JT[{k}] = fma({nl}, dqT, JT[{k}]);")
            for j in cols:
                w(f"J[{k}][{j}] = fma({nl}, dq{j}, J[{k}][{j}]);")
            if mdep:
                for j, e in eff:
                    w(f"J[{k}][{j}] = fma({lit(nu * e)}, dqM, J[{k}][{j}]);")
                if r.default_efficiency != 0.0:
                    w(f"Dm[{k}] = fma({lit(nu * r.default_efficiency)}, dqM, Dm[{k}]);")
        w.ind -= 1
        w("}")
    w("#pragma unroll")
    w("for (int k = 0; k < NS; ++k)")
    w("#pragma unroll")
    w("  for (int j = 0; j < NS; ++j) J[k][j] += Dm[k];")
    w.ind -= 1
    w("}")
    w("")


def _reverse_parts(w: _Writer, L: _Lowering, r: Reaction):
    """Emit pf, pb, kinv, dkinvT (reversible) and the partial-derivative
    maps of the forward and reverse products; returns (dpf, dpb, cols)."""
    rpairs = L.reactant_pairs(r)
    w(f"const double pf = {L.prod_expr(rpairs)};")
    dpf = dict(L.prod_partials(rpairs))
    dpb = {}
    if r.reversible:
        ppairs = [(L.idx[sp], nu) for sp, nu in r.products]
        w(f"const double pb = {L.prod_expr(ppairs)};")
        dpb = dict(L.prod_partials(ppairs))
        kinv = L.kinv_expr(r) if L.species_kc else f"exp({L.nlk_expr(r)})"
        w(f"const double kinv = {kinv};")
        dnu = sum(nu for _, nu in r.products) - sum(nu for _, nu in r.reactants)
        hs = [f"{lit(-nu)} * hrt[{L.idx[sp]}]" for sp, nu in r.products]
        hs += [f"{lit(nu)} * hrt[{L.idx[sp]}]" for sp, nu in r.reactants]
        w(f"const double dkinvT = kinv * invT * ({lit(dnu)} + "
          + " + ".join(hs) + ");")
        w("const double delta = fma(-kinv, pb, pf);")
    else:
        w("const double delta = pf;")
    return dpf, dpb, sorted(set(dpf) | set(dpb))


def _deriv_prologue(w: _Writer, L: _Lowering):
    rev = any(r.reversible for r in L.m.reactions)
    if L.species_kc:
        _kin_prologue(w, L, need_g=False)
        if rev:
            L.emit_gibbs_exp(w)
    else:
        _kin_prologue(w, L, need_g=rev)
    if rev:
        w("double hrt[NS];")
        w("species_h_rt(T, invT, hrt);")


def _emit_wdot_jvp(w: _Writer, L: _Lowering):
    """Forward mode, reaction-major: w and dw = dw/dT dT + dw/dC dC in one
    pass, one cell per lane, no Jacobian held."""
    m = L.m
    w("// omega-dot and its directional derivative (forward mode)")
    w("__device__ __forceinline__ void net_production_rates_jvp("
      "const double T, const double (&C)[NS], const double dT, "
      "const double (&dC)[NS], double (&w)[NS], double (&dw)[NS]) {")
    w.ind += 1
    _deriv_prologue(w, L)
    if L.uses_ctot():
        w("double dCtot = 0.0;")
        w("#pragma unroll")
        w("for (int k = 0; k < NS; ++k) dCtot += dC[k];")
    w("#pragma unroll")
    w("for (int k = 0; k < NS; ++k) { w[k] = 0.0; dw[k] = 0.0; }")
    pending = set()
    for i, r in enumerate(m.reactions):
        w(f"{{  // R{i}: {r.equation} [{r.reaction_type}]")
        w.ind += 1
        L.kf_block_deriv(w, r)
        dpf, dpb, cols = _reverse_parts(w, L, r)
        fv = " + ".join(f"({e}) * dC[{j}]" for j, e in dpf.items()) or "0.0"
        if r.reversible:
            bv = " + ".join(f"({e}) * dC[{j}]" for j, e in dpb.items()) or "0.0"
            w(f"const double ddelta = fma(-kinv, {bv}, {fv}) - dkinvT * dT * pb;")
        else:
            w(f"const double ddelta = {fv};")
        dk = "dkfT * dT"
        if r.kind in ("three-body", "falloff"):
            w(f"const double dM = {L.third_body(r, 'dC', 'dCtot')};")
            dk = "fma(dkfM, dM, dkfT * dT)"
        w("const double q = kf * delta;")
        w(f"const double dq = fma({dk}, delta, kf * ddelta);")
        upd = L.species_updates(r, "q")
        for k, nu in upd:
            nl = lit(nu)
            w(f"w[{k}] = fma({nl}, q, w[{k}]);")
            w(f"dw[{k}] = fma({nl}, dq, dw[{k}]);")
        pending.update(k for k, _ in upd)
        if (i + 1) % L.tuning.get("pin_every", 1) == 0 or i + 1 == len(m.reactions):
            L.pin(w, [f"w[{k}]" for k in sorted(pending)]
                  + [f"dw[{k}]" for k in sorted(pending)])
            pending.clear()
        w.ind -= 1
        w("}")
    w.ind -= 1
    w("}")
    w("")


def _emit_wdot_jvp_lazy(w: _Writer, L: _Lowering):
    """The forward-mode pass with species opened and closed along
    ``reaction_order`` as in ``_emit_wdot_lazy``: C_k and dC_k loaded, E_k
    and h_k/RT computed at first use, w_k and dw_k stored after their last
    update."""
    m = L.m
    w("// omega-dot and its directional derivative, species opened and "
      "closed along the reactions")
    sig = ("__device__ __forceinline__ void net_production_rates_jvp_lazy("
           "const double rho, const double T, const double* __restrict__ Y, "
           "const int64_t ldY, const double a, const double dT, "
           "const double* __restrict__ vY, const int64_t ldV, "
           "const int64_t i, double* __restrict__ wout, "
           "double* __restrict__ dwout, const int64_t ldO) {")
    rev = any(r.reversible for r in m.reactions)
    dc = ("kInvW[{k}] * fma(Y[{k} * ldY + i], a, "
          "vY ? rho * vY[{k} * ldV + i] : 0.0)")
    # branch-free form for the lazy pass: without vY the tangent term reads
    # Y and multiplies it by zero
    dcg = ("kInvW[{k}] * fma(Yg[{k} * ldY + i], a, "
           "sv * Vg[{k} * ldVv + i])")
    if not L.tuning.get("lazy", True) or (rev and not L.species_kc):
        w(sig)
        w.ind += 1
        w("double C[NS], dC[NS], w[NS], dw[NS];")
        w("#pragma unroll")
        w("for (int k = 0; k < NS; ++k) {")
        w("  C[k] = kInvW[k] * rho * Y[k * ldY + i];")
        w("  dC[k] = " + dc.format(k="k") + ";")
        w("}")
        w("net_production_rates_jvp(T, C, dT, dC, w, dw);")
        w("#pragma unroll")
        w("for (int k = 0; k < NS; ++k) {")
        w("  dwout[k * ldO + i] = dw[k];")
        w("  if (wout) wout[k * ldO + i] = w[k];")
        w("}")
        w.ind -= 1
        w("}")
        w("")
        return
    w(sig)
    w.ind += 1
    w("const double lnT = log(T);")
    w("const double invT = 1.0 / T;")
    if rev:
        L.emit_gibbs_prologue(w)
    w("const double sv = vY ? rho : 0.0;")
    w("const int64_t ldVv = vY ? ldV : ldY;")
    w("pyro::gptr Yg = (pyro::gptr)Y;")
    w("pyro::gptr Vg = (pyro::gptr)(vY ? vY : Y);")
    if L.uses_ctot():
        w("double Ctot = 0.0, dCtot = 0.0;")
        w("#pragma unroll")
        w("for (int k = 0; k < NS; ++k) {")
        w("  Ctot += kInvW[k] * rho * Yg[k * ldY + i];")
        w("  dCtot += " + dcg.format(k="k") + ";")
        w("}")
    w("double C[NS], dC[NS], E[NS], hrt[NS], w[NS], dw[NS];")
    L.kinv_f = L.species_kc and L.tuning.get("kinv_f", False)
    if L.kinv_f:
        w("double F[NS];")
    order = reaction_order(L, L.tuning.get("order_tries", 8))
    uses = [species_uses(L, r) for r in m.reactions]
    last_w = {}
    for pos, ri in enumerate(order):
        for k in uses[ri][2]:
            last_w[k] = pos
    open_c, open_e, open_w = set(), set(), set()
    pending = set()
    for pos, ri in enumerate(order):
        r = m.reactions[ri]
        c_u, e_u, w_u = uses[ri]
        if c_u - open_c:
            w('asm volatile("" : "+s"(Yg), "+s"(Vg));')
        for k in sorted(c_u - open_c):
            w(f"C[{k}] = kInvW[{k}] * rho * Yg[{k} * ldY + i];")
            w(f"dC[{k}] = " + dcg.format(k=k) + ";")
        for k in sorted(e_u - open_e):
            L.emit_gibbs_one(w, k)
            if L.kinv_f:
                w(f"F[{k}] = pyro::recip(E[{k}]);")
            w(f"hrt[{k}] = {L.h_expr(m.species[k])};")
        for k in sorted(w_u - open_w):
            w(f"w[{k}] = 0.0; dw[{k}] = 0.0;")
        open_c |= c_u
        open_e |= e_u
        open_w |= w_u
        w(f"{{  // R{ri}: {r.equation} [{r.reaction_type}]")
        w.ind += 1
        L.kf_block_deriv(w, r)
        dpf, dpb, cols = _reverse_parts(w, L, r)
        fv = " + ".join(f"({e}) * dC[{j}]" for j, e in dpf.items()) or "0.0"
        if r.reversible:
            bv = " + ".join(f"({e}) * dC[{j}]" for j, e in dpb.items()) or "0.0"
            w(f"const double ddelta = fma(-kinv, {bv}, {fv}) - dkinvT * dT * pb;")
        else:
            w(f"const double ddelta = {fv};")
        dk = "dkfT * dT"
        if r.kind in ("three-body", "falloff"):
            w(f"const double dM = {L.third_body(r, 'dC', 'dCtot')};")
            dk = "fma(dkfM, dM, dkfT * dT)"
        w("const double q = kf * delta;")
        w(f"const double dq = fma({dk}, delta, kf * ddelta);")
        upd = L.species_updates(r, "q")
        for k, nu in upd:
            nl = lit(nu)
            w(f"w[{k}] = fma({nl}, q, w[{k}]);")
            w(f"dw[{k}] = fma({nl}, dq, dw[{k}]);")
        pending.update(k for k, _ in upd)
        w.ind -= 1
        w("}")
        done = sorted(k for k in w_u if last_w[k] == pos)
        if (pos + 1) % L.tuning.get("pin_every", 1) == 0 or done \
                or pos + 1 == len(order):
            L.pin(w, [f"w[{k}]" for k in sorted(pending)]
                  + [f"dw[{k}]" for k in sorted(pending)])
            pending.clear()
        for k in done:
            w(f"dwout[{k} * ldO + i] = dw[{k}];")
            w(f"if (wout) wout[{k} * ldO + i] = w[{k}];")
        fence = L.tuning.get("lazy_fence", 4)
        if fence and (pos + 1) % fence == 0:
            w('asm volatile("" ::: "memory");')
    for k in range(L.ns):
        if k not in open_w:
            w(f"dwout[{k} * ldO + i] = 0.0;")
            w(f"if (wout) wout[{k} * ldO + i] = 0.0;")
    w.ind -= 1
    w("}")
    w("")
    L.kinv_f = False


def _emit_wdot_vjp(w: _Writer, L: _Lowering):
    """Reverse mode, reaction-major: given the cotangent wb of w, accumulate
    Cb = wb^T dw/dC and Tb = wb^T dw/dT in one pass."""
    m = L.m
    w("// cotangent of omega-dot pulled back to (C, T) (reverse mode)")
    w("__device__ __forceinline__ void net_production_rates_vjp("
      "const double T, const double (&C)[NS], const double (&wb)[NS], "
      "double (&Cb)[NS], double& Tb) {")
    w.ind += 1
    _deriv_prologue(w, L)
    w("#pragma unroll")
    w("for (int k = 0; k < NS; ++k) Cb[k] = 0.0;")
    w("double tb = 0.0, mbd = 0.0;  // T cotangent; default-efficiency sum")
    pending = set()
    for i, r in enumerate(m.reactions):
        w(f"{{  // R{i}: {r.equation} [{r.reaction_type}]")
        w.ind += 1
        upd = L.species_updates(r, "q")
        rb = " + ".join(f"{lit(nu)} * wb[{k}]" for k, nu in upd) or "0.0"
        w(f"const double rb = {rb};")
        L.kf_block_deriv(w, r)
        dpf, dpb, cols = _reverse_parts(w, L, r)
        if r.reversible:
            w("tb = fma(rb, fma(dkfT, delta, -kf * dkinvT * pb), tb);")
        else:
            w("tb = fma(rb, dkfT * delta, tb);")
        w("const double g = rb * kf;")
        for j in cols:
            a, b = dpf.get(j), dpb.get(j)
            if a is not None and b is not None:
                w(f"Cb[{j}] = fma(g, fma(-kinv, {b}, {a}), Cb[{j}]);")
            elif a is not None:
                w(f"Cb[{j}] = fma(g, {a}, Cb[{j}]);")
            else:
                w(f"Cb[{j}] = fma(-g * kinv, {b}, Cb[{j}]);")
        touched = set(cols)
        if r.kind in ("three-body", "falloff"):
            w("const double mb = rb * dkfM * delta;")
            for j, e in L.efficiency_terms(r):
                w(f"Cb[{j}] = fma({L.c(e)}, mb, Cb[{j}]);")
                touched.add(j)
            if r.default_efficiency != 0.0:
                w(f"mbd = fma({L.c(r.default_efficiency)}, mb, mbd);")
        pending.update(touched)
        if (i + 1) % L.tuning.get("pin_every", 1) == 0 or i + 1 == len(m.reactions):
            L.pin(w, [f"Cb[{k}]" for k in sorted(pending)] + ["tb", "mbd"])
            pending.clear()
        w.ind -= 1
        w("}")
    w("#pragma unroll")
    w("for (int k = 0; k < NS; ++k) Cb[k] += mbd;")
    w("Tb = tb;")
    w.ind -= 1
    w("}")
    w("")


def _emit_wdot_vjp_lazy(w: _Writer, L: _Lowering):
    """The reverse-mode pass with species opened and closed along
    ``reaction_order``: Y_k, C_k and the cotangent wb_k loaded and E_k,
    h_k/RT computed at first use; the cotangent Cb_k of C_k is final (up to
    the default-efficiency term) after the last reaction that reads C_k, and
    its contributions g_Yk = Cb_k rho / W_k and sum_k Cb_k Y_k / W_k to g_rho
    are taken there.  The default-efficiency cotangent mbd (added to every
    Cb_k) is known only at the end: it is added to the stored g_Y in a last
    read-modify-write pass and to g_rho through Ctot."""
    m = L.m
    w("// cotangent of omega-dot pulled back to (rho, T, Y), species opened "
      "and closed along the reactions")
    sig = ("__device__ __forceinline__ void net_production_rates_vjp_lazy("
           "const double rho, const double T, const double* __restrict__ Y, "
           "const int64_t ldY, const double* __restrict__ wbp, "
           "const int64_t ldB, const int64_t i, double* __restrict__ g_rho, "
           "double* __restrict__ g_T, double* __restrict__ g_Y, "
           "const int64_t ldG) {")
    rev = any(r.reversible for r in m.reactions)
    if not L.tuning.get("lazy", True) or (rev and not L.species_kc):
        w(sig)
        w.ind += 1
        w("double C[NS], b[NS], Cb[NS], Tb;")
        w("#pragma unroll")
        w("for (int k = 0; k < NS; ++k) {")
        w("  C[k] = kInvW[k] * rho * Y[k * ldY + i];")
        w("  b[k] = wbp[k * ldB + i];")
        w("}")
        w("net_production_rates_vjp(T, C, b, Cb, Tb);")
        w("double gr = 0.0;")
        w("#pragma unroll")
        w("for (int k = 0; k < NS; ++k) {")
        w("  gr = fma(Cb[k] * kInvW[k], Y[k * ldY + i], gr);")
        w("  if (g_Y) g_Y[k * ldG + i] = Cb[k] * rho * kInvW[k];")
        w("}")
        w("if (g_rho) g_rho[i] = gr;")
        w("if (g_T) g_T[i] = Tb;")
        w.ind -= 1
        w("}")
        w("")
        return
    w(sig)
    w.ind += 1
    w("const double lnT = log(T);")
    w("const double invT = 1.0 / T;")
    if rev:
        L.emit_gibbs_prologue(w)
    w("pyro::gptr Yg = (pyro::gptr)Y;")
    w("pyro::gptr Bg = (pyro::gptr)wbp;")
    uses_mbd = any(r.kind in ("three-body", "falloff")
                   and r.default_efficiency != 0.0 for r in m.reactions)
    if L.uses_ctot():
        w("double Ctot = 0.0;")
        w("#pragma unroll")
        w("for (int k = 0; k < NS; ++k) Ctot += kInvW[k] * rho * Yg[k * ldY + i];")
    w("double y[NS], C[NS], E[NS], hrt[NS], wb[NS], Cb[NS];")
    L.kinv_f = L.species_kc and L.tuning.get("kinv_f", False)
    if L.kinv_f:
        w("double F[NS];")
    w("double tb = 0.0, mbd = 0.0, gr = 0.0;")
    order = reaction_order(L, L.tuning.get("order_tries", 8))
    uses = [species_uses(L, r) for r in m.reactions]
    last_c = {}
    for pos, ri in enumerate(order):
        for k in uses[ri][0]:
            last_c[k] = pos
    open_c, open_e, open_b = set(), set(), set()
    pending = set()
    fence = L.tuning.get("lazy_fence", 4)
    for pos, ri in enumerate(order):
        r = m.reactions[ri]
        c_u, e_u, w_u = uses[ri]
        if (c_u - open_c) or (w_u - open_b):
            w('asm volatile("" : "+s"(Yg), "+s"(Bg));')
        for k in sorted(c_u - open_c):
            w(f"y[{k}] = Yg[{k} * ldY + i];")
            w(f"C[{k}] = kInvW[{k}] * rho * y[{k}];")
            w(f"Cb[{k}] = 0.0;")
        for k in sorted(w_u - open_b):
            w(f"wb[{k}] = Bg[{k} * ldB + i];")
        for k in sorted(e_u - open_e):
            L.emit_gibbs_one(w, k)
            if L.kinv_f:
                w(f"F[{k}] = pyro::recip(E[{k}]);")
            w(f"hrt[{k}] = {L.h_expr(m.species[k])};")
        open_c |= c_u
        open_e |= e_u
        open_b |= w_u
        w(f"{{  // R{ri}: {r.equation} [{r.reaction_type}]")
        w.ind += 1
        upd = L.species_updates(r, "q")
        rb = " + ".join(f"{lit(nu)} * wb[{k}]" for k, nu in upd) or "0.0"
        w(f"const double rb = {rb};")
        L.kf_block_deriv(w, r)
        dpf, dpb, cols = _reverse_parts(w, L, r)
        if r.reversible:
            w("tb = fma(rb, fma(dkfT, delta, -kf * dkinvT * pb), tb);")
        else:
            w("tb = fma(rb, dkfT * delta, tb);")
        w("const double g = rb * kf;")
        for j in cols:
            a, b = dpf.get(j), dpb.get(j)
            if a is not None and b is not None:
                w(f"Cb[{j}] = fma(g, fma(-kinv, {b}, {a}), Cb[{j}]);")
            elif a is not None:
                w(f"Cb[{j}] = fma(g, {a}, Cb[{j}]);")
            else:
                w(f"Cb[{j}] = fma(-g * kinv, {b}, Cb[{j}]);")
        touched = set(cols)
        if r.kind in ("three-body", "falloff"):
            w("const double mb = rb * dkfM * delta;")
            for j, e in L.efficiency_terms(r):
                w(f"Cb[{j}] = fma({L.c(e)}, mb, Cb[{j}]);")
                touched.add(j)
            if r.default_efficiency != 0.0:
                w(f"mbd = fma({L.c(r.default_efficiency)}, mb, mbd);")
        pending.update(touched)
        w.ind -= 1
        w("}")
        done = sorted(k for k in c_u if last_c[k] == pos)
        if (pos + 1) % L.tuning.get("pin_every", 1) == 0 or done \
                or pos + 1 == len(order):
            L.pin(w, [f"Cb[{k}]" for k in sorted(pending)] + ["tb", "mbd"])
            pending.clear()
        for k in done:
            w(f"gr = fma(Cb[{k}] * kInvW[{k}], y[{k}], gr);")
            w(f"if (g_Y) g_Y[{k} * ldG + i] = Cb[{k}] * rho * kInvW[{k}];")
        if fence and (pos + 1) % fence == 0:
            w('asm volatile("" ::: "memory");')
    closed = set(last_c)
    if uses_mbd:
        # every Cb_k gets mbd: g_rho through sum_k Y_k / W_k = Ctot / rho,
        # g_Y by a last pass over the stored values (read back through a
        # pointer the optimiser cannot relate to the stores above)
        w("gr = fma(mbd, Ctot / rho, gr);")
        w("if (g_Y) {")
        w.ind += 1
        w("pyro::gmptr gy = (pyro::gmptr)g_Y;")
        w('asm volatile("" : "+s"(gy));')
        for k in range(L.ns):
            if k in closed:
                w(f"gy[{k} * ldG + i] = fma(mbd * rho, kInvW[{k}], "
                  f"gy[{k} * ldG + i]);")
            else:
                w(f"gy[{k} * ldG + i] = mbd * rho * kInvW[{k}];")
        w.ind -= 1
        w("}")
    else:
        for k in range(L.ns):
            if k not in closed:
                w(f"if (g_Y) g_Y[{k} * ldG + i] = 0.0;")
    w("if (g_rho) g_rho[i] = gr;")
    w("if (g_T) g_T[i] = tb;")
    w.ind -= 1
    w("}")
    w("")
    L.kinv_f = False


def cache_slots(L: _Lowering):
    """Slot of each cached per-reaction quantity in the source-Jacobian
    workspace: k_f, k_r = k_f/K_c (reversible), q, dq/dT, dq/d[M]
    (three-body and falloff).  Returns ({(i, name): slot}, n_slots)."""
    slots = {}
    for i, r in enumerate(L.m.reactions):
        names = ["kf"] + (["kr"] if r.reversible else []) + ["q", "dqT"]
        if r.kind in ("three-body", "falloff"):
            names.append("dqM")
        for nm in names:
            slots[(i, nm)] = len(slots)
    return slots, len(slots)


def _emit_reaction_cache(w: _Writer, L: _Lowering):
    """First pass of the source Jacobian: per reaction, the quantities the
    row pass needs, stored to the workspace column of this cell."""
    m = L.m
    slots, nslots = cache_slots(L)
    w(f"constexpr int kCacheSlots = {nslots};")
    w("// cells per workspace chunk: the leading dimension of every slot is a")
    w("// compile-time constant, so slot addresses are a uniform base plus a")
    w("// constant and one per-lane offset shared by all slots")
    w("constexpr int kCacheLd = 65536;")
    w("// per-reaction k_f, k_r, q, dq/dT, dq/d[M] to the workspace")
    w("__device__ __forceinline__ void reaction_cache(const double T, "
      "const double (&C)[NS], double* __restrict__ work, const unsigned il) {")
    w.ind += 1
    _deriv_prologue(w, L)
    for i, r in enumerate(m.reactions):
        w(f"{{  // R{i}: {r.equation} [{r.reaction_type}]")
        w.ind += 1
        L.kf_block_deriv(w, r)
        _reverse_parts(w, L, r)
        def at(nm):
            return f"(work + {slots[(i, nm)]} * kCacheLd)[il]"
        w(f"{at('kf')} = kf;")
        if r.reversible:
            w(f"{at('kr')} = kf * kinv;")
            w(f"{at('dqT')} = fma(dkfT, delta, -kf * dkinvT * pb);")
        else:
            w(f"{at('dqT')} = dkfT * delta;")
        w(f"{at('q')} = kf * delta;")
        if r.kind in ("three-body", "falloff"):
            w(f"{at('dqM')} = dkfM * delta;")
        if (i + 1) % L.tuning.get("pin_every", 1) == 0:
            w('asm volatile("" ::: "memory");')
        w.ind -= 1
        w("}")
    w.ind -= 1
    w("}")
    w("")


def _emit_source_jacobian_rows(w: _Writer, L: _Lowering):
    """Second pass: row k+1 of dS/d[h; Y] from the cached quantities of the
    reactions that involve species k.  G_kj = dw_k/dC_j is accumulated from
    k_f dpf/dC_j - k_r dpb/dC_j (+ the [M] term); then, with
    A_k = sum_j G_kj C_j / rho and B_k = dw_k/dT,
      J[1+k][c] = W_k/rho ((A_k - w_k/rho) drho_c + B_k dT_c)
                  + [c = 1+m] (W_k/W_m) G_km.
    dT_c comes from LDS (dTl[m*hs] = dT of column 1+m)."""
    m = L.m
    slots, _ = cache_slots(L)
    ns = L.ns
    w("// rows of the source Jacobian from the reaction cache: the row groups")
    w("// G0..G1-1 (compile-time), one after the other")
    w("template <int G0, int G1>")
    w("__device__ __forceinline__ void source_jacobian_rows("
      "double (&C)[NS], const double rho, const double T, "
      "const double cp, const double mmw, const double vd, "
      "const double* __restrict__ work, const unsigned il, "
      "pyro::lptr dTl, const int hs, "
      "double* __restrict__ o, const int64_t s_rc, "
      "double* __restrict__ S, const int64_t ldS) {")
    w.ind += 1
    w("pyro::cptr IW = (pyro::cptr)kInvWMem;")
    # the cache as a global-memory pointer: loads from a uniform base plus
    # the lane's 32-bit offset rather than a 64-bit address per load
    w("pyro::gptr wg = (pyro::gptr)work;")
    w("double irho = 1.0 / rho;")
    w("double rT = rho * (1.0 / T);")
    w("double dT0 = 1.0 / cp;")
    w("double dr0 = -rT * dT0;")
    w("double rm = -rho * mmw;  // drho_m = -rT dT_m + rm / W_m")
    w(f"constexpr int NV = {ns + 1};")
    # rows 1..ns split into kJacGroups groups of about equal work (longest
    # first onto the lightest group); row 0 goes with group 0.  Group rg is
    # one wave-uniform branch.
    inv = {k: [] for k in range(ns)}
    for i, r in enumerate(m.reactions):
        for k, nu in L.species_updates(r, "q"):
            inv[k].append((i, nu))
    ng = int(L.tuning.get("jac_groups", 1))
    rows = list(range(L.tuning.get("jac_rows", ns)))
    load = [0.0] * ng
    load[0] = ns
    group_of = {}
    for k in sorted(rows, key=lambda k: -(len(inv[k]) * 8 + ns)):
        g = min(range(ng), key=lambda g: load[g])
        group_of[k] = g
        load[g] += len(inv[k]) * 8 + ns
    def emit_row0():
        w("{  // row 0: d(vd/rho)")
        w.ind += 1
        w("const double v0 = -vd * irho * irho;")
        w("o[0] = v0 * dr0;")
        w("#pragma unroll")
        w("for (int q = 0; q < NS; ++q) {")
        w("  const double dT = dTl[q * hs];")
        w("  o[(1 + q) * s_rc] = v0 * fma(-rT, dT, rm * kInvW[q]);")
        w("}")
        w("if (S) S[0] = vd * irho;")
        w.ind -= 1
        w("}")

    def emit_row(k):
        w(f"{{  // row {k + 1}: {m.species[k].name}")
        w.ind += 1
        # fresh copies of C and of the per-cell scalars for this row:
        # products of them are then not shared with (and kept live for)
        # other rows
        for a in range(0, ns, 24):
            ops = ", ".join(f'"+v"(C[{j}])' for j in range(a, min(ns, a + 24)))
            w(f'asm volatile("" : {ops});')
        w('asm volatile("" : "+v"(rm), "+v"(rT), "+v"(dr0), "+v"(dT0), '
          '"+v"(irho));')
        # fresh copies of the cache and LDS pointers: this row's loads cannot
        # be hoisted above the previous row (and kept live meanwhile)
        w('asm volatile("" : "+s"(wg), "+v"(dTl), "+s"(IW));')
        w("double G[NS];")
        w("#pragma unroll")
        w("for (int j = 0; j < NS; ++j) G[j] = 0.0;")
        w("double wk = 0.0, bk = 0.0, dm = 0.0;")
        jf = int(L.tuning.get("jac_fence", 0))
        for n_r, (i, nu) in enumerate(inv[k]):
            if jf and n_r and n_r % jf == 0:
                # at most jac_fence reactions' loads in flight
                w('asm volatile("" ::: "memory");')
            r = m.reactions[i]
            nl = lit(nu)
            w(f"{{  // R{i}: {r.equation}")
            w.ind += 1
            def at(nm, i=i):
                return f"(wg + {slots[(i, nm)]} * kCacheLd)[il]"
            w(f"const double kf = {at('kf')};")
            w(f"wk = fma({nl}, {at('q')}, wk);")
            w(f"bk = fma({nl}, {at('dqT')}, bk);")
            rpairs = L.reactant_pairs(r)
            dpf = dict(L.prod_partials(rpairs))
            dpb = {}
            if r.reversible:
                w(f"const double kr = {at('kr')};")
                ppairs = [(L.idx[sp], nn) for sp, nn in r.products]
                dpb = dict(L.prod_partials(ppairs))
            w(f"const double a = {nl} * kf;")
            if dpb:
                w(f"const double b = {lit(-nu)} * kr;")
            for j in sorted(set(dpf) | set(dpb)):
                if j in dpf:
                    w(f"G[{j}] = fma(a, {dpf[j]}, G[{j}]);")
                if j in dpb:
                    w(f"G[{j}] = fma(b, {dpb[j]}, G[{j}]);")
            if r.kind in ("three-body", "falloff"):
                w(f"const double mq = {nl} * {at('dqM')};")
                for j, e in L.efficiency_terms(r):
                    w(f"G[{j}] = fma({lit(e)}, mq, G[{j}]);")
                if r.default_efficiency != 0.0:
                    w(f"dm = fma({lit(r.default_efficiency)}, mq, dm);")
            w.ind -= 1
            w("}")
        w("double A = 0.0;")
        w("#pragma unroll")
        w("for (int j = 0; j < NS; ++j) { G[j] += dm; A = fma(G[j], C[j], A); }")
        w(f"const double f = kW[{k}] * irho;")
        w("const double al = f * fma(-wk, irho, A * irho);")
        w("const double be = f * bk;")
        w(f"double* row = o + {(k + 1) * (ns + 1)} * s_rc;")
        w("row[0] = fma(al, dr0, be * dT0);")
        w("#pragma unroll")
        w("for (int q = 0; q < NS; ++q) {")
        w("  const double dT = dTl[q * hs];")
        w("  const double dr = fma(-rT, dT, rm * IW[q]);")
        w(f"  row[(1 + q) * s_rc] = fma(al, dr, fma(be, dT, "
          f"kW[{k}] * IW[q] * G[q]));")
        w("}")
        w(f"if (S) S[{k + 1} * ldS] = f * wk;")
        # keep the next row's loads below this row's stores: one row's
        # accumulators live at a time
        w('asm volatile("" ::: "memory");')
        w.ind -= 1
        w("}")

    # one thread per cell: every row, in natural order
    w(f"if constexpr (G0 == 0 && G1 == {ng}) {{")
    w.ind += 1
    emit_row0()
    for k in rows:
        emit_row(k)
    w.ind -= 1
    w("} else {")
    w.ind += 1
    # one wave per group of rows (the split form)
    for g in range(ng):
        w(f"if constexpr (G0 <= {g} && {g} < G1) {{")
        w.ind += 1
        if g == 0:
            emit_row0()
        for k in sorted(r for r in rows if group_of[r] == g):
            emit_row(k)
        w.ind -= 1
        w("}")
    w.ind -= 1
    w("}")
    w.ind -= 1
    w("}")
    # one group per wave: the group of wave rg (wave-uniform)
    w("__device__ __forceinline__ void source_jacobian_rows_group(int rg, "
      "double (&C)[NS], const double rho, const double T, "
      "const double cp, const double mmw, const double vd, "
      "const double* __restrict__ work, const unsigned il, "
      "pyro::lptr dTl, const int hs, "
      "double* __restrict__ o, const int64_t s_rc, "
      "double* __restrict__ S, const int64_t ldS) {")
    w.ind += 1
    w("switch (rg) {")
    for g in range(ng):
        w(f"  case {g}: source_jacobian_rows<{g}, {g + 1}>(C, rho, T, cp, mmw, "
          "vd, work, il, dTl, hs, o, s_rc, S, ldS); break;")
    w("  default: break;")
    w("}")
    w.ind -= 1
    w("}")
    w("")


def _kinv_linear(L: _Lowering, r: Reaction):
    """ln K_c^-1 of a reversible reaction as linear forms in the Gibbs basis
    (1 - ln T, T, T^2, T^3, T^4, 1/T, 1):
        ln K_c^-1 = sum_j s_j g_j/RT + dnu ln(R T / p0)
    (chem_expr.py:274-300; s_j the net stoichiometric coefficient, products
    positive), with g/RT = a0 (1 - ln T) - a1 T/2 - a2 T^2/6 - a3 T^3/12
    - a4 T^4/20 + a5/T - a6 of each species' range.  Returns (const,
    {T_mid: (hi, lo)}): the coefficients of the single-range species and of
    the dnu term, and per T_mid those of the two-range species of that T_mid
    in its high and low range."""
    def gcoef(a):
        return [a[0], -a[1] / 2, -a[2] / 6, -a[3] / 12, -a[4] / 20, a[5],
                -a[6]]
    net: Dict[int, float] = {}
    for sp, nu in r.products:
        net[L.idx[sp]] = net.get(L.idx[sp], 0.0) + nu
    for sp, nu in r.reactants:
        net[L.idx[sp]] = net.get(L.idx[sp], 0.0) - nu
    dnu = sum(nu for _, nu in r.products) - sum(nu for _, nu in r.reactants)
    const = [0.0] * 7
    const[0] -= dnu
    const[6] += dnu * (1.0 + math.log(GAS_CONSTANT / ONE_ATM))
    groups: Dict[float, Tuple[List[float], List[float]]] = {}
    for k in sorted(net):
        s, sp = net[k], L.m.species[k]
        if s == 0.0:
            continue
        if tuple(sp.high) == tuple(sp.low):
            for b, c in enumerate(gcoef(sp.high)):
                const[b] += s * c
        else:
            hi, lo = groups.setdefault(sp.t_mid, ([0.0] * 7, [0.0] * 7))
            for b, (ch, cl) in enumerate(zip(gcoef(sp.high), gcoef(sp.low))):
                hi[b] += s * ch
                lo[b] += s * cl
    return const, groups


def _emit_kinv_linear(w: _Writer, L: _Lowering, r: Reaction, out: str):
    """Emit ``out`` = ln K_c^-1 from ``_kinv_linear``: one form per T_mid of
    the reaction's two-range species, both ranges evaluated and pinned so
    the range choice stays a per-lane select."""
    basis = ["omlnT", "T", "T2", "T3", "T4", "invT"]

    def poly(cs):
        expr = L.c(cs[6])
        for cc, b in zip(reversed(cs[:6]), reversed(basis)):
            if cc != 0.0:
                expr = f"fma({L.c(cc)}, {b}, {expr})"
        return expr
    const, groups = _kinv_linear(L, r)
    if len(groups) == 1:
        (t, (hi, lo)), = groups.items()
        ch = [a + b for a, b in zip(const, hi)]
        cl = [a + b for a, b in zip(const, lo)]
        w(f"double lh = {poly(ch)}, ll = {poly(cl)};")
        w('asm volatile("" : "+v"(lh), "+v"(ll));')
        w(f"const double {out} = {L.tmid_flag[t]} ? lh : ll;")
        return
    w(f"double {out} = {poly(const)};")
    for t in sorted(groups):
        hi, lo = groups[t]
        w(f"{{ double gh = {poly(hi)}, gl = {poly(lo)};")
        w('  asm volatile("" : "+v"(gh), "+v"(gl));')
        w(f"  {out} += {L.tmid_flag[t]} ? gh : gl; }}")


def _fused_row_cost(L: _Lowering, k: int, inv) -> float:
    """Relative work of row k+1 in the one-pass form (exponentials and
    logarithms dominate): per reaction k_f and K_c^-1, the falloff blend,
    and the row's stores."""
    cost = 0.0
    for i, _ in inv[k]:
        r = L.m.reactions[i]
        cost += 2.0 if r.reversible else 1.0
        if r.kind == "falloff":
            cost += 6.0 if r.falloff_type == "Troe" else 2.0
    return cost + 0.1 * L.ns


def _emit_jac_fused_gibbs(w: _Writer, L: _Lowering, gibbs: bool):
    """The one-pass source Jacobian's per-cell Gibbs factors: for the
    species k = rg (mod G), E_k = exp(g'_k) and F_k = exp(-g'_k) to this
    cell's LDS columns Et, Ft (stride kJacFusedCells), g'_k = g_k/RT
    shifted by element potentials (the shifts cancel in a balanced
    reaction; the same exponents as ``emit_gibbs_exp``)."""
    w(f"constexpr bool kJacFusedGibbs = {'true' if gibbs else 'false'};")
    w("// cells per block of the one-pass kernel: the row stride of its LDS")
    w("// tables (species-major)")
    w("constexpr int kJacFusedCells = 64;")
    w("__device__ __forceinline__ void jac_fused_gibbs(const double T, "
      "const int rg, const int G, double* Et, double* Ft) {")
    w.ind += 1
    if gibbs:
        w("constexpr int hs = kJacFusedCells;")
        w("const double lnT = log(T);")
        w("const double invT = 1.0 / T;")
        L.emit_gibbs_prologue(w)
        basis = ["omlnT", "T", "T2", "T3", "T4", "invT"]
        for k, sp in enumerate(L.m.species):
            def coeffs(c, k=k):
                return [c[0], -c[1] / 2, -c[2] / 6, -c[3] / 12, -c[4] / 20,
                        c[5] - L.shift_m[k], -(c[6] + L.shift_n[k])]

            def poly(cs):
                expr = L.c(cs[6])
                for cc, b in zip(reversed(cs[:6]), reversed(basis)):
                    expr = f"fma({L.c(cc)}, {b}, {expr})"
                return expr
            hi, lo = coeffs(sp.high), coeffs(sp.low)
            w(f"if ({k} % G == rg) {{  // {sp.name}")
            w.ind += 1
            if hi == lo:
                w(f"const double g = {poly(hi)};")
            else:
                w(f"double gh = {poly(hi)}, gl = {poly(lo)};")
                w('asm volatile("" : "+v"(gh), "+v"(gl));')
                w(f"const double g = {L.tmid_flag[sp.t_mid]} ? gh : gl;")
            w(f"Et[{k} * hs] = exp(g);")
            w(f"Ft[{k} * hs] = exp(-g);")
            w.ind -= 1
            w("}")
    w.ind -= 1
    w("}")
    w("")


def _emit_source_jacobian_fused(w: _Writer, L: _Lowering):
    """One-pass source Jacobian: the same row assembly as
    ``_emit_source_jacobian_rows``, but each row evaluates k_f, K_c^-1 and
    their T- and [M]-derivatives for the reactions of its species from T
    and the concentrations instead of reading them from a per-reaction
    workspace.  K_c^-1 is a product of per-species factors E_k = exp(g'_k)
    (products) and F_k = exp(-g'_k) (reactants) that the kernel tabulates
    per cell in LDS (``jac_fused_gibbs``; mechanisms whose reversible
    reactions are balanced with integer coefficients), else one exponential
    of a linear form per reaction (``_kinv_linear``); its T-derivative takes
    the species' h/RT from the dT table (dT_m = -(R T/cp) (h_m/RT) / W_m,
    so h_m/RT = dT_m W_m rkT with rkT = -cp/(R T)).  Rows are split into
    kJacFusedGroups groups of about equal work; a wave takes a contiguous
    range of groups, chosen at run time, so one copy of the code serves any
    number of waves per cell."""
    m = L.m
    ns = L.ns
    ng = int(L.tuning.get("jac_fused_groups", 8))
    inv = {k: [] for k in range(ns)}
    for i, r in enumerate(m.reactions):
        for k, nu in L.species_updates(r, "q"):
            inv[k].append((i, nu))
    rows = list(range(L.tuning.get("jac_rows", ns)))
    cost = {k: _fused_row_cost(L, k, inv) for k in rows}
    load = [0.0] * ng
    load[0] = 0.1 * ns
    group_of = {}
    for k in sorted(rows, key=lambda k: -cost[k]):
        g = min(range(ng), key=lambda g: load[g])
        group_of[k] = g
        load[g] += cost[k]
    pin_n = int(L.tuning.get("jac_fused_pin", 2))
    pin_ptrs = bool(L.tuning.get("jac_fused_pin_ptrs", True))
    gibbs = L.species_kc
    _emit_jac_fused_gibbs(w, L, gibbs)
    table_at = len(w.lines)
    w(f"constexpr int kJacFusedGroups = {ng};")
    # per-wave staging buffer of per-cell rows: [8 entries][kJacStageLd]
    # doubles, cell c of entry j at j * kJacStageLd + c.  The wave writes
    # with lane = cell (consecutive addresses) and reads 8 cells x 8
    # entries per group of 32 lanes; a row length of 4 (mod 32) doubles
    # puts those 32 reads on 32 distinct banks.
    w("constexpr int kJacStageLd = 68;")
    w("// rows of the source Jacobian in one pass: the row groups g0..g1-1;")
    w("// Cp0, dTl0, Ep0 and Fp0 point at this lane's cell in the LDS tables")
    w("// of C, dT, E and F (the last two unused without kJacFusedGibbs)")
    w("__device__ __forceinline__ void source_jacobian_fused("
      "int g0, int g1, pyro::lptr Cp0, const double Ctot0, "
      "const double rho, const double T0, const double cp, "
      "const double mmw, const double vd, pyro::lptr dTl0, pyro::lptr Ep0, "
      "pyro::lptr Fp0, pyro::gmptr o0, const int64_t s_rc, pyro::gmptr S0, "
      "const int64_t ldS, const int mode, pyro::lmptr sw0, pyro::lptr sr0, "
      "pyro::gmptr ob0, const unsigned lob, const int64_t s8) {")
    w.ind += 1
    w("constexpr int hs = kJacFusedCells;")
    # one table for every row: a reaction met again in another row reuses
    # its constants (or, with jac_fused_ctable off, literals)
    if L.tuning.get("jac_fused_ctable", True):
        L.tab, L.tab_index = [], {}
    w("pyro::cptr K0 = (pyro::cptr)kTabJacFused;")
    w("pyro::cptr IW0 = (pyro::cptr)kInvWMem;")
    w("const double irho0 = 1.0 / rho;")
    w("const double rT0 = rho * (1.0 / T0);")
    w("const double dT00 = 1.0 / cp;")
    w("const double dr00 = -rT0 * dT00;")
    w("const double rm0 = -rho * mmw;  // drho_m = -rT dT_m + rm / W_m")
    w("const double lnT0 = log(T0);")
    w("const double invT0 = 1.0 / T0;")
    w("const double omlnT0 = 1.0 - lnT0;")
    w("const double T20 = T0 * T0, T30 = T20 * T0, T40 = T30 * T0;")
    w("const double rkT0 = -cp * invT0 * (1.0 / kGasConstant);")
    w("const double rtp0 = kGasConstant * T0 / kOneAtm;")
    w("const double prt0 = kOneAtm / (kGasConstant * T0);")
    L._flags(w, "T0")
    scal = ["T", "rm", "rT", "dr0", "dT0", "irho", "lnT", "invT", "omlnT",
            "T2", "T3", "T4", "rkT", "rtp", "prt"] + (
                ["Ctot"] if L.uses_ctot() else [])
    ptrs = ["Cp", "dTl"] + (["Ep", "Fp"] if gibbs else [])

    def pin_chunks(names, n=20):
        for a in range(0, len(names), n):
            ops = ", ".join(f'"+v"({x})' for x in names[a:a + n])
            w(f'asm volatile("" : {ops});')

    def row_locals():
        # row-local copies of the per-cell scalars and of the table, output
        # and constant pointers, passed through an empty asm: nothing of a
        # row is computed or loaded ahead of it (and kept live meanwhile),
        # and no value is carried from row to row
        w("double " + ", ".join(f"{x} = {x}0" for x in scal) + ";")
        w("pyro::lptr " + ", ".join(f"{x} = {x}0" for x in ptrs) + ";")
        w("pyro::gmptr o = o0, S = S0, ob = ob0;")
        w("pyro::lmptr sw = sw0;")
        w("pyro::lptr sr = sr0;")
        w("pyro::cptr K = K0, IW = IW0;")
        pin_chunks(scal + ptrs + ["o", "S", "sw", "sr"])
        w('asm volatile("" : "+s"(IW), "+s"(K), "+s"(ob));')
        w("const pyro::LdsColumn<hs> C{Cp};")
        if gibbs:
            w("const pyro::LdsColumn<hs> E{Ep}, F{Fp};")

    nvv = ns + 1
    nt_jac = bool(L.tuning.get("nt_jac", True))
    # rows shorter than one chunk of 8 entries keep direct stores
    stage = bool(L.tuning.get("jac_stage", True)) and nvv >= 8

    def emit_stores(r, v0, vq):
        """Store row r of this lane's cell: entry 0 = v0, entry 1 + q = vq
        with {q} replaced.  mode 1 (consecutive cells adjacent): streaming
        stores; mode 2 (each cell's row contiguous, a full block of 64
        cells): through this wave's LDS buffer, 8 entries of the 64 cells
        at a time, so that each store instruction writes 8 cells x 64
        contiguous bytes at the scalar base ob plus this lane's byte offset
        lob; mode 0: ordinary stores at the given strides."""
        def loop(streaming):
            w(f"pyro::gmptr row = o + {r * nvv} * s_rc;")
            # the store addresses are formed here, not ahead of the row
            w('asm volatile("" : "+v"(row));')
            w(f"row[0] = {v0};")
            w("#pragma unroll")
            w("for (int q = 0; q < NS; ++q) {")
            val = vq.format(q="q")
            if streaming:
                w(f"  __builtin_nontemporal_store({val}, "
                  "&row[(1 + q) * s_rc]);")
            else:
                w(f"  row[(1 + q) * s_rc] = {val};")
            w("}")

        def staged():
            # chunks of 8 entries; the last one ends at entry nvv - 1 and
            # may overlap the one before (the same values stored twice)
            starts = list(range(0, nvv - 7, 8))
            if nvv % 8:
                starts.append(nvv - 8)
            pipe = bool(L.tuning.get("jac_stage_pipe", True))

            def put(e0):
                # the chunk's values first, then its LDS writes: a table
                # read between two writes would wait for the earlier write
                vals = [v0 if e == 0 else vq.format(q=e - 1)
                        for e in range(e0, e0 + 8)]
                w("{")
                w("const double " + ", ".join(
                    f"x{j} = {v}" for j, v in enumerate(vals)) + ";")
                for j in range(8):
                    w(f"sw[{j} * kJacStageLd] = x{j};")
                w("}")

            def get():
                w("const double " + ", ".join(
                    f"y{pp} = sr[{pp * 8}]" for pp in range(8)) + ";")

            def out(e0):
                for pp in range(8):
                    w(f"pyro::store_at(ob + ({pp} * s8 + {r * nvv + e0}), "
                      f"lob, y{pp});")

            # LDS runs in order per wave: a chunk's writes land before its
            # reads, and its reads before the next chunk's writes; the
            # empty asm statements keep the compiler to that order
            bar = 'asm volatile("" ::: "memory");'
            if not pipe:
                for e0 in starts:
                    put(e0)
                    w(bar)
                    w("{")
                    get()
                    out(e0)
                    w("}")
                    w(bar)
                return
            # pipelined: read chunk c - 1 back, compute and write chunk c
            # while that read is in flight, then store chunk c - 1
            put(starts[0])
            w(bar)
            for c in range(1, len(starts) + 1):
                w("{")
                get()
                w(bar)
                if c < len(starts):
                    put(starts[c])
                    w(bar)
                out(starts[c - 1])
                w("}")
            w(bar)

        def streaming():
            # the empty asm statements open and close the streaming arm so
            # that it shares no leading or trailing instructions with the
            # others: merged stores would lose the non-temporal hint
            w('asm volatile("" ::: "memory");')
            loop(True)
            w('asm volatile("" ::: "memory");')

        arms = ([("mode == 2", staged)] if stage else []) + (
            [("mode == 1", streaming)] if nt_jac else [])
        for n_a, (cond, body) in enumerate(arms):
            w(("if" if n_a == 0 else "} else if") + f" ({cond}) {{")
            w.ind += 1
            body()
            w.ind -= 1
        if arms:
            w("} else {")
            w.ind += 1
        loop(False)
        if arms:
            w.ind -= 1
            w("}")

    def emit_row0():
        w("{  // row 0: d(vd/rho)")
        w.ind += 1
        row_locals()
        w("const double v0 = -vd * irho * irho;")
        emit_stores(0, "v0 * dr0",
                    "v0 * fma(-rT, dTl[{q} * hs], rm * IW[{q}])")
        w("if (S) S[0] = vd * irho;")
        w.ind -= 1
        w("}")

    def emit_reaction(i, nu):
        r = m.reactions[i]
        nl = lit(nu)
        w(f"{{  // R{i}: {r.equation}")
        w.ind += 1
        L.kf_block_deriv(w, r)
        rpairs = L.reactant_pairs(r)
        w(f"const double pf = {L.prod_expr(rpairs)};")
        dpf = dict(L.prod_partials(rpairs))
        dpb = {}
        if r.reversible:
            ppairs = [(L.idx[sp], n2) for sp, n2 in r.products]
            w(f"const double pb = {L.prod_expr(ppairs)};")
            dpb = dict(L.prod_partials(ppairs))
            if gibbs:
                L.kinv_f = True
                w(f"const double kinv = {L.kinv_expr(r)};")
                L.kinv_f = False
            else:
                _emit_kinv_linear(w, L, r, "lk")
                w("const double kinv = exp(lk);")
            # sum over reactants minus products of nu h/RT, from the dT table
            net: Dict[int, float] = {}
            for sp, n2 in r.reactants:
                net[L.idx[sp]] = net.get(L.idx[sp], 0.0) + n2
            for sp, n2 in r.products:
                net[L.idx[sp]] = net.get(L.idx[sp], 0.0) - n2
            expr = "0.0"
            for j in sorted(net):
                if net[j] != 0.0:
                    expr = (f"fma({L.c(net[j] * m.molecular_weights[j])}, "
                            f"dTl[{j} * hs], {expr})")
            dnu = (sum(n2 for _, n2 in r.products)
                   - sum(n2 for _, n2 in r.reactants))
            w(f"const double dkinvT = kinv * invT * fma(rkT, {expr}, "
              f"{lit(dnu)});")
            w("const double delta = fma(-kinv, pb, pf);")
            w("const double dqT = fma(dkfT, delta, -kf * dkinvT * pb);")
        else:
            w("const double delta = pf;")
            w("const double dqT = dkfT * delta;")
        w(f"wk = fma({nl}, kf * delta, wk);")
        w(f"bk = fma({nl}, dqT, bk);")
        w(f"const double a = {nl} * kf;")
        if dpb:
            w(f"const double b = {lit(-nu)} * (kf * kinv);")
        touched = []
        for j in sorted(set(dpf) | set(dpb)):
            if j in dpf:
                w(f"G[{j}] = fma(a, {dpf[j]}, G[{j}]);")
            if j in dpb:
                w(f"G[{j}] = fma(b, {dpb[j]}, G[{j}]);")
            touched.append(j)
        if r.kind in ("three-body", "falloff"):
            w(f"const double mq = {nl} * dkfM * delta;")
            for j, e in L.efficiency_terms(r):
                w(f"G[{j}] = fma({L.c(e)}, mq, G[{j}]);")
                touched.append(j)
            if r.default_efficiency != 0.0:
                w(f"dm = fma({L.c(r.default_efficiency)}, mq, dm);")
        w.ind -= 1
        w("}")
        return touched

    def emit_row(k):
        w(f"{{  // row {k + 1}: {m.species[k].name}")
        w.ind += 1
        row_locals()
        w("double G[NS];")
        w("#pragma unroll")
        w("for (int j = 0; j < NS; ++j) G[j] = 0.0;")
        w("double wk = 0.0, bk = 0.0, dm = 0.0;")
        pending = set()
        for n_r, (i, nu) in enumerate(inv[k]):
            pending.update(emit_reaction(i, nu))
            if pin_n and (n_r + 1) % pin_n == 0:
                # at most pin_n reactions of the row interleaved: the
                # accumulators settle and later table reads stay below
                pin_chunks([f"G[{j}]" for j in sorted(pending)]
                           + ["wk", "bk", "dm"]
                           + (ptrs if pin_ptrs else []))
                pending.clear()
        # A = sum_j (G_j + dm) C_j with the default-efficiency term dm
        # common to every column taken once
        w("double A = 0.0;")
        w("#pragma unroll")
        w("for (int j = 0; j < NS; ++j) A = fma(G[j], C[j], A);")
        if L.uses_ctot():
            w("A = fma(dm, Ctot, A);")
        w(f"const double f = kW[{k}] * irho;")
        w("const double al = f * fma(-wk, irho, A * irho);")
        w("const double be = f * bk;")
        # entry (1+k, 1+q) = al drho_q + be dT_q + (W_k/W_q) (G_q + dm) with
        # drho_q = -rT dT_q + rm/W_q:  ga dT_q + (cr + W_k G_q)/W_q
        w("const double ga = fma(-al, rT, be);")
        w(f"const double cr = fma({L.c(m.molecular_weights[k])}, dm, al * rm);")
        wk_c = L.c(m.molecular_weights[k])
        emit_stores(k + 1, "fma(al, dr0, be * dT0)",
                    f"fma(ga, dTl[{{q}} * hs], IW[{{q}}] * "
                    f"fma({wk_c}, G[{{q}}], cr))")
        w(f"if (S) S[{k + 1} * ldS] = f * wk;")
        w('asm volatile("" ::: "memory");')
        w.ind -= 1
        w("}")

    w("for (int g = g0; g < g1; ++g) {")
    w.ind += 1
    w("switch (g) {")
    for g in range(ng):
        w(f"case {g}: {{")
        w.ind += 1
        if g == 0:
            emit_row0()
        for k in sorted(r for r in rows if group_of[r] == g):
            emit_row(k)
        w("break;")
        w.ind -= 1
        w("}")
    w("default: break;")
    w("}")
    w.ind -= 1
    w("}")
    w.ind -= 1
    w("}")
    w("")
    tab = L.tab or [0.0]
    w.lines.insert(table_at, "__constant__ double kTabJacFused[] = {"
                   + ", ".join(lit(x) for x in tab) + "};")
    L.tab, L.tab_index = None, None


def _emit_reaction_quantities(w: _Writer, L: _Lowering):
    m = L.m
    w("// per-reaction k_f, -ln K_c, r_f, r_r, r_net to memory "
      "(codegen/python.py:360-515)")
    # MASK selects the outputs at compile time (1 k_f, 2 -ln K_c, 4 r_f,
    # 8 r_r, 16 r_net), so a call for k_f alone does not evaluate the Gibbs
    # functions or the reverse rates
    w("template <int MASK>")
    w("__device__ __forceinline__ void reaction_quantities(const double T, "
      "const double (&C)[NS], const int64_t i, const int64_t ld, "
      "double* __restrict__ okf, double* __restrict__ onlk, "
      "double* __restrict__ orf, double* __restrict__ orr, "
      "double* __restrict__ onet) {")
    w.ind += 1
    # nt_rq: streaming (non-temporal) stores of the outputs
    nt = bool(L.tuning.get("nt_rq", False))

    def st(ptr, idx, val):
        if nt:
            return f"__builtin_nontemporal_store({val}, &{ptr}[{idx}]);"
        return f"{ptr}[{idx}] = {val};"
    _kin_prologue(w, L, need_g=True)
    for j, r in enumerate(m.reactions):
        idx = f"{j} * ld + i"
        w(f"{{  // R{j}: {r.equation} [{r.reaction_type}]")
        w.ind += 1
        L.kf_block(w, r)
        rf = L.prod_expr(L.reactant_pairs(r))
        if r.reversible:
            rr = L.prod_expr([(L.idx[sp], nu) for sp, nu in r.products])
            w(f"const double nlk = {L.nlk_expr(r)};")
            w(f"const double pf = {rf};")
            w(f"const double pb = {rr};")
            w("if ((MASK & 1) && okf) " + st("okf", idx, "kf"))
            w("if ((MASK & 2) && onlk) " + st("onlk", idx, "nlk"))
            w("if ((MASK & 4) && orf) " + st("orf", idx, "kf * pf"))
            w("if ((MASK & 8) && orr) " + st("orr", idx, "kf * exp(nlk) * pb"))
            w("if ((MASK & 16) && onet) " + st("onet", idx, "kf * (pf - exp(nlk) * pb)"))
        else:
            w(f"const double pf = {rf};")
            w("if ((MASK & 1) && okf) " + st("okf", idx, "kf"))
            w("if ((MASK & 2) && onlk) " + st("onlk", idx, "-0.17364695002734 * T"))
            w("if ((MASK & 4) && orf) " + st("orf", idx, "kf * pf"))
            w("if ((MASK & 8) && orr) " + st("orr", idx, "0.0"))
            w("if ((MASK & 16) && onet) " + st("onet", idx, "kf * pf"))
        w.ind -= 1
        w("}")
    w.ind -= 1
    w("}")
    w("")


def _emit_creation_destruction(w: _Writer, L: _Lowering):
    m = L.m
    w("// creation/destruction split (chem_expr.py:542-567)")
    w("__device__ __forceinline__ void creation_destruction(const double T, "
      "const double (&C)[NS], double (&cd)[NS], double (&dd)[NS]) {")
    w.ind += 1
    _kin_prologue(w, L, need_g=any(r.reversible for r in m.reactions))
    w("#pragma unroll")
    w("for (int k = 0; k < NS; ++k) { cd[k] = 0.0; dd[k] = 0.0; }")
    for j, r in enumerate(m.reactions):
        w(f"{{  // R{j}: {r.equation}")
        w.ind += 1
        L.kf_block(w, r)
        w(f"const double qf = kf * ({L.prod_expr(L.reactant_pairs(r))});")
        if r.reversible:
            rr = L.prod_expr([(L.idx[sp], nu) for sp, nu in r.products])
            w(f"const double qr = kf * exp({L.nlk_expr(r)}) * ({rr});")
        for sp, nu in r.reactants:
            k = L.idx[sp]
            w(f"dd[{k}] = fma({lit(nu)}, qf, dd[{k}]);")
            if r.reversible:
                w(f"cd[{k}] = fma({lit(nu)}, qr, cd[{k}]);")
        for sp, nu in r.products:
            k = L.idx[sp]
            w(f"cd[{k}] = fma({lit(nu)}, qf, cd[{k}]);")
            if r.reversible:
                w(f"dd[{k}] = fma({lit(nu)}, qr, dd[{k}]);")
        w.ind -= 1
        w("}")
    w.ind -= 1
    w("}")
    w("")


def _emit_creation_destruction_lazy(w: _Writer, L: _Lowering):
    """Creation and destruction rates of one cell with species opened and
    closed along a reaction order as in ``_emit_wdot_lazy``: C_k loaded and
    g_k/RT computed at first use, both rates of a species stored after the
    last reaction it takes part in (chem_expr.py:542-567)."""
    m = L.m
    w("// creation/destruction split, species opened and closed along the "
      "reactions (chem_expr.py:542-567)")
    w("__device__ __forceinline__ void creation_destruction_lazy("
      "const double rho, const double T, const double* __restrict__ Y, "
      "const int64_t ldY, const int64_t i, double* __restrict__ cdot, "
      "double* __restrict__ ddot, const int64_t ldW) {")
    w.ind += 1
    if not L.tuning.get("lazy", True):
        w("double C[NS], cd[NS], dd[NS];")
        w("#pragma unroll")
        w("for (int k = 0; k < NS; ++k) C[k] = kInvW[k] * rho * Y[k * ldY + i];")
        w("creation_destruction(T, C, cd, dd);")
        w("#pragma unroll")
        w("for (int k = 0; k < NS; ++k) {")
        w("  cdot[k * ldW + i] = cd[k];")
        w("  ddot[k * ldW + i] = dd[k];")
        w("}")
        w.ind -= 1
        w("}")
        w("")
        return
    rev = any(r.reversible for r in m.reactions)
    w("const double lnT = log(T);")
    w("const double invT = 1.0 / T;")
    if rev:
        w("const double c0 = log(kOneAtm / (kGasConstant * T));")
        L._flags(w)
    if L.uses_ctot():
        w("double Ctot = 0.0;")
        w("#pragma unroll")
        w("for (int k = 0; k < NS; ++k) Ctot += kInvW[k] * rho * Y[k * ldY + i];")
    w("pyro::gptr Yg = (pyro::gptr)Y;")
    w("double C[NS], g[NS], cd[NS], dd[NS];")
    uses = []
    for r in m.reactions:
        c = {L.idx[sp] for sp, _ in r.reactants}
        e = set()
        spc = {L.idx[sp] for sp, _ in list(r.reactants) + list(r.products)}
        if r.reversible:
            c |= {L.idx[sp] for sp, _ in r.products}
            e = set(spc)
        if r.kind in ("three-body", "falloff"):
            c |= {k for k, _ in L.efficiency_terms(r)}
        uses.append((c, e, spc))
    order = reaction_order(L, L.tuning.get("order_tries", 8), uses)
    last_w = {}
    for pos, ri in enumerate(order):
        for k in uses[ri][2]:
            last_w[k] = pos
    open_c, open_e, open_w = set(), set(), set()
    pending = set()
    for pos, ri in enumerate(order):
        r = m.reactions[ri]
        c_u, e_u, w_u = uses[ri]
        if c_u - open_c:
            w('asm volatile("" : "+s"(Yg));')
        for k in sorted(c_u - open_c):
            w(f"C[{k}] = kInvW[{k}] * rho * Yg[{k} * ldY + i];")
        for k in sorted(e_u - open_e):
            sp = m.species[k]
            w(f"g[{k}] = {L.h_expr(sp)} - ({L.s_expr(sp)});")
        for k in sorted(w_u - open_w):
            w(f"cd[{k}] = 0.0; dd[{k}] = 0.0;")
        open_c |= c_u
        open_e |= e_u
        open_w |= w_u
        w(f"{{  // R{ri}: {r.equation}")
        w.ind += 1
        L.kf_block(w, r)
        w(f"const double qf = kf * ({L.prod_expr(L.reactant_pairs(r))});")
        if r.reversible:
            rr = L.prod_expr([(L.idx[sp], nu) for sp, nu in r.products])
            w(f"const double qr = kf * exp({L.nlk_expr(r)}) * ({rr});")
        for sp, nu in r.reactants:
            k = L.idx[sp]
            w(f"dd[{k}] = fma({lit(nu)}, qf, dd[{k}]);")
            if r.reversible:
                w(f"cd[{k}] = fma({lit(nu)}, qr, cd[{k}]);")
        for sp, nu in r.products:
            k = L.idx[sp]
            w(f"cd[{k}] = fma({lit(nu)}, qf, cd[{k}]);")
            if r.reversible:
                w(f"dd[{k}] = fma({lit(nu)}, qr, dd[{k}]);")
        pending |= w_u
        w.ind -= 1
        w("}")
        done = sorted(k for k in w_u if last_w[k] == pos)
        if (pos + 1) % L.tuning.get("pin_every", 1) == 0 or done \
                or pos + 1 == len(order):
            L.pin(w, [f"cd[{k}]" for k in sorted(pending)]
                  + [f"dd[{k}]" for k in sorted(pending)])
            pending.clear()
        for k in done:
            w(f"cdot[{k} * ldW + i] = cd[{k}];")
            w(f"ddot[{k} * ldW + i] = dd[{k}];")
    for k in range(L.ns):
        if k not in open_w:
            w(f"cdot[{k} * ldW + i] = 0.0;")
            w(f"ddot[{k} * ldW + i] = 0.0;")
    w.ind -= 1
    w("}")
    w("")


class HipCodeGenerator(CodeGenerator):
    """``hip`` target: mechanism IR -> HIP C++ translation unit."""

    #: overrides of DEFAULT_TUNING (class-level; set before generate())
    tuning: Optional[dict] = None

    @staticmethod
    def get_name() -> str:
        return "hip"

    @staticmethod
    def supports_overloading() -> bool:
        return True

    @staticmethod
    def generate(name: str, sol: Mechanism,
                 opts: CodeGenerationOptions = None) -> str:
        if opts is None:
            opts = CodeGenerationOptions()
        if opts.directive_offload is not None:
            raise TypeError(
                "OpenMP/ACC directive based offloading is not supported for "
                "HIP code generation")
        if opts.scalar_type not in (None, "double"):
            raise TypeError("the hip target computes in double precision only")
        mech = as_mechanism(sol)
        L = _Lowering(mech, HipCodeGenerator.tuning)
        w = _Writer()
        wdot_block = int(L.tuning.get("wdot_block", 256))
        _emit_prelude(w, L, name, mech.digest(), wdot_block)
        _emit_thermo(w, L)
        _emit_wdot(w, L)
        _emit_wdot_lazy(w, L)
        _emit_reaction_quantities(w, L)
        _emit_creation_destruction(w, L)
        _emit_creation_destruction_lazy(w, L)
        _emit_wdot_jvp(w, L)
        _emit_wdot_jvp_lazy(w, L)
        _emit_wdot_vjp(w, L)
        _emit_wdot_vjp_lazy(w, L)
        _emit_reaction_cache(w, L)
        _emit_source_jacobian_rows(w, L)
        _emit_source_jacobian_fused(w, L)
        dense = mech.n_species <= DENSE_JAC_MAX_SPECIES
        if dense:
            _emit_wdot_jac(w, L)
        transport = emit_transport_tables(w, mech, lit)
        w("}  // namespace mech")
        w("")
        w('#include "pyro_thermo_kernels.hpp"')
        w('#include "pyro_flamelet_kernels.hpp"')
        if dense:
            w('#include "pyro_flamelet_dense.hpp"')
        if transport:
            w('#include "pyro_transport_kernels.hpp"')
        return w.text()

    @staticmethod
    def compile_class(name: str, source: str, mech: Mechanism = None):
        """Compile generated HIP ``source`` into a shared library and return
        the device ``Thermochemistry`` class bound to it (the counterpart of
        ``PythonCodeGenerator.compile_class``, ``codegen/python.py:689-697``)."""
        from ..runtime import build_mechanism_library
        from ..thermochem import make_thermochem_class
        lib = build_mechanism_library(name, source)
        return make_thermochem_class(name, lib, mech)

    @staticmethod
    def get_thermochem_class(sol: Mechanism):
        """Generate, compile and load the class for ``sol``
        (``codegen/python.py:699-707``)."""
        from ..runtime import load_mechanism_library
        from ..thermochem import make_thermochem_class
        mech = as_mechanism(sol)
        lib = load_mechanism_library(mech)
        return make_thermochem_class("Thermochemistry", lib, mech)
