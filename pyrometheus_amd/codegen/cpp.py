"""The ``cpp`` target: a header-only C++20 class template of the mechanism,
callable from host code and from HIP device code.

The reference's C++ target (``pyrometheus/codegen/cpp.py:81-660``) renders
one ``struct NAME<DataTypeT, ContainerT>`` of static methods whose arguments
are ``ContainerT`` scalars and ``std::array`` species/reaction vectors, with
``DataTypeT`` constrained by the ``IntegralLike`` concept (``:96-109``) so
that user number types (automatic-differentiation duals, interval types)
pass through.  This generator keeps that interface -- the struct's name,
its constants, method names, argument orders and return types -- and adds
one thing for the GPU: every numerical method is ``__host__ __device__``
when the header is compiled by hipcc, so a user kernel can evaluate the
mechanism per thread (one cell per lane) without the package's libraries.

The expressions follow the reference's (``chem_expr.py:49-567``): NASA-7
polynomials selected at T_mid, Arrhenius as ``A T^b`` or
``exp(ln A + b ln T - T_a/T)``, third-body sums with default efficiencies,
Lindemann and Troe falloff with the reference's guard on the reduced
pressure, ``-ln K_c`` from the Gibbs functions with the ``Delta nu ln(p0/RT)``
term and the placeholder ``-0.17364695002734 T`` of irreversible reactions,
the reactant ``orders`` paired with the reactant list in order (as the
reference's ``zip`` does), and the reference's fixed-point loop for the
temperature (500 iterations, |dT| < 1e-6).  Transport methods are emitted
when every species has transport data, on the fits of
:mod:`pyrometheus_amd.transport`.
"""

from __future__ import annotations

from typing import List

import numpy as np

from ..mechanism import GAS_CONSTANT, ONE_ATM, Mechanism, Reaction, \
    as_mechanism
from ..transport import has_transport, transport_fits
from . import CodeGenerationOptions, CodeGenerator
from .hip import lit

file_extension = "hpp"


def _sum(terms: List[str], empty: str = "ContainerT(0.0)") -> str:
    return " + ".join(terms) if terms else empty


def _pow(base: str, nu: float) -> str:
    """base**nu as the reference's generated code computes it for an
    integer exponent (repeated product) and otherwise with pow."""
    if float(nu).is_integer() and 1 <= nu <= 4:
        return "*".join([base] * int(nu))
    return f"pow({base}, ContainerT({lit(nu)}))"


class _Emit:
    def __init__(self, mech: Mechanism):
        self.m = mech
        self.idx = {s.name: k for k, s in enumerate(mech.species)}
        self.falloff = [(i, r) for i, r in enumerate(mech.reactions)
                        if r.kind == "falloff"]
        self.three_body = [(i, r) for i, r in enumerate(mech.reactions)
                           if r.kind == "three-body"]

    # -- NASA-7 (chem_expr.py:49-174) ---------------------------------------
    @staticmethod
    def _cp(c, t):
        return (f"({lit(c[0])} + {lit(c[1])}*{t} + {lit(c[2])}*{t}*{t} "
                f"+ {lit(c[3])}*{t}*{t}*{t} + {lit(c[4])}*{t}*{t}*{t}*{t})")

    @staticmethod
    def _h(c, t):
        return (f"({lit(c[0])} + {lit(c[1] / 2)}*{t} + {lit(c[2] / 3)}*{t}*{t}"
                f" + {lit(c[3] / 4)}*{t}*{t}*{t} + {lit(c[4] / 5)}*{t}*{t}*{t}"
                f"*{t} + {lit(c[5])}/{t})")

    @staticmethod
    def _s(c, t):
        return (f"({lit(c[0])}*log({t}) + {lit(c[1])}*{t} + "
                f"{lit(c[2] / 2)}*{t}*{t} + {lit(c[3] / 3)}*{t}*{t}*{t} + "
                f"{lit(c[4] / 4)}*{t}*{t}*{t}*{t} + {lit(c[6])})")

    def nasa(self, sp, gen, t="temperature"):
        return (f"({t} > ContainerT({lit(sp.t_mid)}) ? "
                f"ContainerT{gen(sp.high, t)} : ContainerT{gen(sp.low, t)})")

    # -- kinetics (chem_expr.py:274-430) ------------------------------------
    @staticmethod
    def arrhenius(rate, t="temperature") -> str:
        if rate.t_a == 0:
            if rate.b == 0:
                return f"ContainerT({lit(rate.A)})"
            return f"{lit(rate.A)}*pow({t}, ContainerT({lit(rate.b)}))"
        return (f"exp({lit(np.log(rate.A))} + {lit(rate.b)}*log({t}) "
                f"- {lit(rate.t_a)}/{t})")

    def third_body(self, r: Reaction, c="concentrations") -> str:
        eff = dict(r.efficiencies)
        nondef = [f"{lit(e)}*{c}[{self.idx[sp]}]" for sp, e in
                  r.efficiencies]
        default = [f"{c}[{k}]" for k, s in enumerate(self.m.species)
                   if s.name not in eff]
        terms = list(nondef)
        if default:
            terms.append(f"{lit(r.default_efficiency)}*({' + '.join(default)})")
        return "(" + _sum(terms) + ")"

    def neg_log_kc(self, r: Reaction) -> str:
        sr = [f"{lit(nu)}*g0_rt[{self.idx[sp]}]" for sp, nu in r.reactants]
        sp_ = [f"{lit(nu)}*g0_rt[{self.idx[sp]}]" for sp, nu in r.products]
        dnu = sum(nu for _, nu in r.products) - sum(nu for _, nu in
                                                     r.reactants)
        out = f"({_sum(sp_)}) - ({_sum(sr)})"
        if dnu != 0:
            out += f" - {lit(dnu)}*c0"
        return out

    def reactant_orders(self, r: Reaction):
        ind = [self.idx[sp] for sp, _ in r.reactants]
        nus = [o for _, o in r.orders] if r.orders else \
            [nu for _, nu in r.reactants]
        return list(zip(ind, nus))

    def product(self, pairs, c="concentrations") -> str:
        if not pairs:
            return "ContainerT(1.0)"
        return "*".join(_pow(f"{c}[{k}]", nu) for k, nu in pairs)

    def stoich(self, k: int):
        name = self.m.species[k].name
        fwd = [(i, nu) for i, r in enumerate(self.m.reactions)
               for sp, nu in r.reactants if sp == name]
        rev = [(i, nu) for i, r in enumerate(self.m.reactions)
               for sp, nu in r.products if sp == name]
        return fwd, rev


def _header(name: str, mech: Mechanism) -> str:
    E = _Emit(mech)
    m = mech
    ns, nr, nf = m.n_species, m.n_reactions, len(E.falloff)
    W = m.molecular_weights
    o: List[str] = []
    w = o.append
    w("// Generated by pyrometheus_amd.codegen.cpp for mechanism "
      f"'{m.name}' ({ns} species, {nr} reactions).")
    w("// Do not edit: regenerate with CppCodeGenerator.generate().")
    w("#pragma once")
    w("")
    w("#if __cplusplus < 202002L")
    w('#error "Pyrometheus requires C++20 or later."')
    w("#endif")
    w("")
    w("#include <array>")
    w("#include <cmath>")
    w("#include <concepts>")
    w("#include <string>")
    w("#include <utility>")
    w("")
    w("#ifndef PYRO_HD")
    w("#if defined(__HIPCC__)")
    w("#include <hip/hip_runtime.h>")
    w("#define PYRO_HD __host__ __device__")
    w("#else")
    w("#define PYRO_HD")
    w("#endif")
    w("#endif")
    w("")
    w("namespace pyro {")
    w("")
    w("using std::exp;")
    w("using std::fabs;")
    w("using std::log;")
    w("using std::log10;")
    w("using std::pow;")
    w("using std::sqrt;")
    w("")
    w("#ifndef PYRO_INTEGRAL_LIKE_DEFINED")
    w("#define PYRO_INTEGRAL_LIKE_DEFINED")
    w("// the operations the generated methods apply to DataTypeT")
    w("template <typename T>")
    w("concept IntegralLike = requires(T a, T b) {")
    for req in ("{ T(1.0) } -> std::same_as<T>;",
                "{ a } -> std::convertible_to<double>;",
                "{ a + b } -> std::same_as<T>;", "{ a - b } -> std::same_as<T>;",
                "{ a * b } -> std::same_as<T>;", "{ a / b } -> std::same_as<T>;",
                "{ -a } -> std::same_as<T>;", "{ +a } -> std::same_as<T>;",
                "{ a += b } -> std::same_as<T&>;",
                "{ a -= b } -> std::same_as<T&>;",
                "{ a *= b } -> std::same_as<T&>;",
                "{ a /= b } -> std::same_as<T&>;",
                "{ a == b } -> std::same_as<bool>;",
                "{ a != b } -> std::same_as<bool>;",
                "{ a < b } -> std::same_as<bool>;",
                "{ a > b } -> std::same_as<bool>;",
                "{ a <= b } -> std::same_as<bool>;",
                "{ a >= b } -> std::same_as<bool>;",
                "{ log(a) } -> std::same_as<T>;", "{ exp(a) } -> std::same_as<T>;",
                "{ sqrt(a) } -> std::same_as<T>;",
                "{ pow(a, b) } -> std::same_as<T>;"):
        w("    " + req)
    w("};")
    w("#endif")
    w("")
    w("template <IntegralLike _DataTypeT = double, "
      "typename _ContainerT = _DataTypeT>")
    w(f"struct {name} {{")
    w(f"  constexpr static int num_species = {ns};")
    w(f"  constexpr static int num_reactions = {nr};")
    w(f"  constexpr static int num_falloff = {nf};")
    w("  constexpr static const char* species_names[] = {"
      + ", ".join(f'"{s.name}"' for s in m.species) + "};")
    w("  constexpr static const char* element_names[] = {"
      + ", ".join(f'"{e}"' for e in m.elements) + "};")
    w("")
    w("  using DataTypeT = _DataTypeT;")
    w("  using ContainerT = _ContainerT;")
    w("  using SpeciesT = std::array<ContainerT, num_species>;")
    w("  using Species2T = std::array<SpeciesT, num_species>;")
    w("  using ReactionsT = std::array<ContainerT, num_reactions>;")
    w("")
    w("  constexpr static double molecular_weights[] = {"
      + ", ".join(lit(x) for x in W) + "};")
    w("  constexpr static double inv_molecular_weights[] = {"
      + ", ".join(lit(1.0 / x) for x in W) + "};")
    w(f"  constexpr static double gas_constant = {lit(GAS_CONSTANT)};")
    w(f"  constexpr static double one_atm = {lit(ONE_ATM)};")
    w("")
    w("  static int nSpecies() { return num_species; }")
    w("  static std::string get_species_name(int index) "
      "{ return species_names[index]; }")
    w("  static std::string get_element_name(int index) "
      "{ return element_names[index]; }")
    w("  static int get_species_index(const std::string& name) {")
    for k, s in enumerate(m.species):
        w(f'    if (name == "{s.name}") return {k};')
    w("    return -1;")
    w("  }")
    w("  static int get_element_index(const std::string& name) {")
    for k, e in enumerate(m.elements):
        w(f'    if (name == "{e}") return {k};')
    w("    return -1;")
    w("  }")
    w("")

    def inv_w_sum(y="mass_fractions"):
        return " + ".join(f"inv_molecular_weights[{k}]*{y}[{k}]"
                          for k in range(ns))

    def fn(sig: str, body: List[str]):
        w(f"  PYRO_HD static {sig} {{")
        for line in body:
            w("    " + line)
        w("  }")
        w("")

    fn("ContainerT get_specific_gas_constant(SpeciesT const& mass_fractions)",
       [f"return gas_constant * ({inv_w_sum()});"])
    fn("ContainerT get_mix_molecular_weight(SpeciesT const& mass_fractions)",
       [f"return 1.0 / ({inv_w_sum()});"])
    fn("ContainerT get_mixture_molecular_weight(SpeciesT const& "
       "mass_fractions)", [f"return 1.0 / ({inv_w_sum()});"])
    fn("SpeciesT get_concentrations(ContainerT rho, SpeciesT const& "
       "mass_fractions)",
       ["return SpeciesT{" + ", ".join(
           f"inv_molecular_weights[{k}]*mass_fractions[{k}]*rho"
           for k in range(ns)) + "};"])
    fn("SpeciesT get_mole_fractions(ContainerT mix_mol_weight, SpeciesT const&"
       " mass_fractions)",
       ["return SpeciesT{" + ", ".join(
           f"inv_molecular_weights[{k}]*mass_fractions[{k}]*mix_mol_weight"
           for k in range(ns)) + "};"])
    fn("ContainerT get_mass_average_property(SpeciesT const& mass_fractions, "
       "SpeciesT const& spec_property)",
       ["return " + " + ".join(
           f"mass_fractions[{k}]*spec_property[{k}]*inv_molecular_weights[{k}]"
           for k in range(ns)) + ";"])
    fn("ContainerT get_mixture_specific_heat_cv_mass(ContainerT temperature, "
       "SpeciesT const& mass_fractions)",
       ["SpeciesT cp0_r = get_species_specific_heats_r(temperature);",
        "for (int i = 0; i < num_species; ++i) cp0_r[i] -= 1.0;",
        "return gas_constant * get_mass_average_property(mass_fractions, "
        "cp0_r);"])
    fn("ContainerT get_mixture_specific_heat_cp_mass(ContainerT temperature, "
       "SpeciesT const& mass_fractions)",
       ["const SpeciesT cp0_r = get_species_specific_heats_r(temperature);",
        "return gas_constant * get_mass_average_property(mass_fractions, "
        "cp0_r);"])
    fn("ContainerT get_mixture_enthalpy_mass(ContainerT temperature, "
       "SpeciesT const& mass_fractions)",
       ["const SpeciesT h0_rt = get_species_enthalpies_rt(temperature);",
        "return gas_constant * get_mass_average_property(mass_fractions, "
        "h0_rt) * temperature;"])
    fn("ContainerT get_mixture_internal_energy_mass(ContainerT temperature, "
       "SpeciesT const& mass_fractions)",
       ["SpeciesT e0_rt = get_species_enthalpies_rt(temperature);",
        "for (int i = 0; i < num_species; ++i) e0_rt[i] -= 1.0;",
        "return gas_constant * get_mass_average_property(mass_fractions, "
        "e0_rt) * temperature;"])
    fn("ContainerT get_density(ContainerT p, ContainerT temperature, "
       "SpeciesT const& mass_fractions)",
       ["const ContainerT mmw = get_mix_molecular_weight(mass_fractions);",
        "const ContainerT rt = gas_constant * temperature;",
        "return p * mmw / rt;"])
    fn("ContainerT get_pressure(ContainerT density, ContainerT temperature, "
       "SpeciesT const& mass_fractions)",
       ["const ContainerT mmw = get_mix_molecular_weight(mass_fractions);",
        "const ContainerT rt = gas_constant * temperature;",
        "return density * rt / mmw;"])
    for nm, gen in (("get_species_specific_heats_r", E._cp),
                    ("get_species_enthalpies_rt", E._h),
                    ("get_species_entropies_r", E._s)):
        fn(f"SpeciesT {nm}(ContainerT temperature)",
           ["return SpeciesT{"] + [f"  {E.nasa(s, gen)},"
                                   for s in m.species] + ["};"])
    fn("SpeciesT get_species_gibbs_rt(ContainerT temperature)",
       ["const SpeciesT h0_rt = get_species_enthalpies_rt(temperature);",
        "const SpeciesT s0_r = get_species_entropies_r(temperature);",
        "SpeciesT g0_rt;",
        "for (int i = 0; i < num_species; ++i) g0_rt[i] = h0_rt[i] - "
        "s0_r[i];",
        "return g0_rt;"])
    body = ["const ContainerT rt = gas_constant * temperature;",
            "const ContainerT c0 = log(one_atm / rt);",
            "const SpeciesT g0_rt = get_species_gibbs_rt(temperature);",
            "(void)c0;",
            "return ReactionsT{"]
    for r in m.reactions:
        body.append("  " + (E.neg_log_kc(r) if r.reversible
                            else "-0.17364695002734*temperature") + ",")
    body.append("};")
    fn("ReactionsT get_equilibrium_constants(ContainerT temperature)", body)
    fn("ContainerT get_temperature(ContainerT energy_or_enthalpy, ContainerT "
       "t_guess, SpeciesT const& mass_fractions, bool const do_energy = true)",
       ["const int num_iter = 500;",
        "const double tol = 1.0e-06;",
        "ContainerT iter_temp = t_guess;",
        "for (int iter = 0; iter < num_iter; ++iter) {",
        "  const ContainerT iter_rhs = energy_or_enthalpy - (do_energy",
        "      ? get_mixture_internal_energy_mass(iter_temp, mass_fractions)",
        "      : get_mixture_enthalpy_mass(iter_temp, mass_fractions));",
        "  const ContainerT iter_deriv = -(do_energy",
        "      ? get_mixture_specific_heat_cv_mass(iter_temp, mass_fractions)",
        "      : get_mixture_specific_heat_cp_mass(iter_temp, mass_fractions));",
        "  const ContainerT iter_dt = -iter_rhs / iter_deriv;",
        "  iter_temp += iter_dt;",
        "  if (fabs(iter_dt) < tol) break;",
        "}",
        "return iter_temp;"])
    if E.falloff:
        body = ["using FalloffT = std::array<ContainerT, num_falloff>;",
                "const FalloffT k_high = {"]
        body += [f"  {E.arrhenius(r.high_rate)}," for _, r in E.falloff]
        body += ["};", "const FalloffT k_low = {"]
        body += [f"  {E.arrhenius(r.low_rate)}," for _, r in E.falloff]
        body += ["};", "const FalloffT reduced_pressure = {"]
        body += [f"  {E.third_body(r)}*k_low[{j}]/k_high[{j}],"
                 for j, (_, r) in enumerate(E.falloff)]
        body += ["};", "const FalloffT falloff_center = {"]
        for _, r in E.falloff:
            if r.falloff_type == "Troe":
                tp = r.troe
                s = (f"{lit(1.0 - tp[0])}*exp(-temperature/{lit(tp[1])}) + "
                     f"{lit(tp[0])}*exp(-temperature/{lit(tp[2])})")
                if len(tp) == 4:
                    s += f" + exp(-{lit(tp[3])}/temperature)"
                body.append(f"  log10({s}),")
            else:
                body.append("  ContainerT(1.0),")
        body += ["};", "const FalloffT falloff_factor = {"]
        for j, (_, r) in enumerate(E.falloff):
            if r.falloff_type == "Troe":
                lr = f"log10(reduced_pressure[{j}])"
                c = f"(-0.4 - 0.67*falloff_center[{j}])"
                n = f"(0.75 - 1.27*falloff_center[{j}])"
                body.append(f"  (reduced_pressure[{j}] > 0.0 ? ContainerT(("
                            f"{lr} + {c})/({n} - 0.14*({lr} + {c}))) : "
                            f"ContainerT({lit(-1 / 0.14)})),")
            else:
                body.append("  ContainerT(0.0),")
        body += ["};"]
        for j, (i, r) in enumerate(E.falloff):
            f = (f"pow(ContainerT(10.0), falloff_center[{j}]/(1.0 + "
                 f"falloff_factor[{j}]*falloff_factor[{j}]))"
                 if r.falloff_type == "Troe" else "ContainerT(1.0)")
            body.append(f"k_fwd[{i}] = k_high[{j}]*{f}*reduced_pressure[{j}]"
                        f"/(1.0 + reduced_pressure[{j}]);")
        fn("void get_falloff_rates(ContainerT temperature, SpeciesT const& "
           "concentrations, ReactionsT& k_fwd)", body)
    body = ["ReactionsT k_fwd = {"]
    for r in m.reactions:
        body.append("  " + ("ContainerT(0.0)" if r.kind == "falloff"
                            else E.arrhenius(r.rate)) + ",")
    body.append("};")
    if E.falloff:
        body.append("get_falloff_rates(temperature, concentrations, k_fwd);")
    for i, r in E.three_body:
        body.append(f"k_fwd[{i}] *= {E.third_body(r)};")
    body.append("(void)concentrations;")
    body.append("return k_fwd;")
    fn("ReactionsT get_fwd_rate_coefficients(ContainerT temperature, "
       "SpeciesT const& concentrations)", body)

    def rev_prod(r):
        return E.product([(E.idx[sp], nu) for sp, nu in r.products])

    body = ["const ReactionsT k_fwd = get_fwd_rate_coefficients(temperature, "
            "concentrations);",
            "const ReactionsT log_k_eq = get_equilibrium_constants("
            "temperature);",
            "(void)log_k_eq;",
            "return ReactionsT{"]
    for i, r in enumerate(m.reactions):
        fw = E.product(E.reactant_orders(r))
        if r.reversible:
            body.append(f"  k_fwd[{i}]*({fw} - exp(log_k_eq[{i}])*"
                        f"{rev_prod(r)}),")
        else:
            body.append(f"  k_fwd[{i}]*{fw},")
    body.append("};")
    fn("ReactionsT get_net_rates_of_progress(ContainerT temperature, "
       "SpeciesT const& concentrations)", body)
    body = ["const ReactionsT k_fwd = get_fwd_rate_coefficients(temperature, "
            "concentrations);", "return ReactionsT{"]
    body += [f"  k_fwd[{i}]*{E.product(E.reactant_orders(r))},"
             for i, r in enumerate(m.reactions)]
    body.append("};")
    fn("ReactionsT get_fwd_rates_of_progress(ContainerT temperature, "
       "SpeciesT const& concentrations)", body)
    body = ["const ReactionsT k_fwd = get_fwd_rate_coefficients(temperature, "
            "concentrations);",
            "const ReactionsT log_k_eq = get_equilibrium_constants("
            "temperature);",
            "(void)log_k_eq;",
            "return ReactionsT{"]
    for i, r in enumerate(m.reactions):
        body.append(f"  k_fwd[{i}]*exp(log_k_eq[{i}])*{rev_prod(r)},"
                    if r.reversible else "  ContainerT(0.0),")
    body.append("};")
    fn("ReactionsT get_rev_rates_of_progress(ContainerT temperature, "
       "SpeciesT const& concentrations)", body)

    def species_sums(fwd_of, rev_of):
        out = []
        for k in range(ns):
            fwd, rev = E.stoich(k)
            out.append(fwd_of(fwd, rev) if rev_of is None
                       else (fwd_of(fwd, rev), rev_of(fwd, rev)))
        return out

    def terms(pairs, arr):
        return [f"{lit(nu)}*{arr}[{i}]" for i, nu in pairs]

    body = ["const SpeciesT concentrations = get_concentrations(rho, "
            "mass_fractions);",
            "const ReactionsT r_net = get_net_rates_of_progress(temperature, "
            "concentrations);",
            "return SpeciesT{"]
    for k in range(ns):
        fwd, rev = E.stoich(k)
        body.append(f"  ({_sum(terms(rev, 'r_net'))}) - "
                    f"({_sum(terms(fwd, 'r_net'))}),")
    body.append("};")
    fn("SpeciesT get_net_production_rates(ContainerT rho, ContainerT "
       "temperature, SpeciesT const& mass_fractions)", body)

    def made(k):
        fwd, rev = E.stoich(k)
        return (f"({_sum(terms(rev, 'r_fwd'))}) + "
                f"({_sum(terms(fwd, 'r_rev'))})")

    def lost(k):
        fwd, rev = E.stoich(k)
        return (f"({_sum(terms(fwd, 'r_fwd'))}) + "
                f"({_sum(terms(rev, 'r_rev'))})")

    pro = ["const SpeciesT concentrations = get_concentrations(rho, "
           "mass_fractions);",
           "const ReactionsT r_fwd = get_fwd_rates_of_progress(temperature, "
           "concentrations);",
           "const ReactionsT r_rev = get_rev_rates_of_progress(temperature, "
           "concentrations);"]
    sig = ("(ContainerT rho, ContainerT temperature, SpeciesT const& "
           "mass_fractions)")
    fn("SpeciesT get_creation_rates" + sig,
       pro + ["return SpeciesT{"] + [f"  {made(k)}," for k in range(ns)]
       + ["};"])
    fn("SpeciesT get_destruction_rates" + sig,
       pro + ["return SpeciesT{"] + [f"  {lost(k)}," for k in range(ns)]
       + ["};"])
    fn("std::pair<SpeciesT, SpeciesT> get_creation_destruction_rates" + sig,
       pro + ["const SpeciesT cdot = {"] + [f"  {made(k)}," for k in range(ns)]
       + ["};", "const SpeciesT ddot = {"]
       + [f"  {lost(k)}," for k in range(ns)]
       + ["};", "return std::make_pair(cdot, ddot);"])
    if has_transport(m):
        _transport(fn, m)
    w("};")
    w("")
    w("}  // namespace pyro")
    return "\n".join(o) + "\n"


def _transport(fn, m: Mechanism):
    """The transport methods (cpp.py:580-660 of the reference; the
    expressions of chem_expr.py:190-267) on this package's fits."""
    visc, cond, diff = transport_fits(m)
    ns = m.n_species
    W = m.molecular_weights

    def poly(c):
        lt = "log(temperature)"
        return (f"({lit(c[0])} + {lit(c[1])}*{lt} + {lit(c[2])}*{lt}*{lt} + "
                f"{lit(c[3])}*{lt}*{lt}*{lt} + {lit(c[4])}*{lt}*{lt}*{lt}*{lt})")

    fn("SpeciesT get_species_viscosities(ContainerT temperature)",
       ["return SpeciesT{"] + [f"  sqrt(temperature)*{poly(c)}*{poly(c)},"
                               for c in visc] + ["};"])
    fn("SpeciesT get_species_thermal_conductivities(ContainerT temperature)",
       ["return SpeciesT{"] + [f"  sqrt(temperature)*{poly(c)},"
                               for c in cond] + ["};"])
    body = ["return Species2T{"]
    for i in range(ns):
        body.append("  SpeciesT{")
        body += [f"    sqrt(temperature)*temperature*{poly(diff[i, j])},"
                 for j in range(ns)]
        body.append("  },")
    body.append("};")
    fn("Species2T get_species_binary_mass_diffusivities(ContainerT "
       "temperature)", body)
    pre = ["const ContainerT mmw = get_mixture_molecular_weight("
           "mass_fractions);",
           "const SpeciesT mole_fractions = get_mole_fractions(mmw, "
           "mass_fractions);"]
    body = pre + ["const SpeciesT viscosities = get_species_viscosities("
                  "temperature);", "SpeciesT mix_rule_f;"]
    for sp in range(ns):
        body.append(f"mix_rule_f[{sp}] = " + " + ".join(
            f"mole_fractions[{j}]*(1.0 + sqrt((viscosities[{sp}]/"
            f"viscosities[{j}])*{lit(np.sqrt(W[j] / W[sp]))}))*(1.0 + sqrt(("
            f"viscosities[{sp}]/viscosities[{j}])*"
            f"{lit(np.sqrt(W[j] / W[sp]))}))/{lit(np.sqrt(8 * (1 + W[sp] / W[j])))}"
            for j in range(ns)) + ";")
    body.append("return " + " + ".join(
        f"mole_fractions[{i}]*viscosities[{i}]/mix_rule_f[{i}]"
        for i in range(ns)) + ";")
    fn("ContainerT get_mixture_viscosity_mixavg(ContainerT temperature, "
       "SpeciesT const& mass_fractions)", body)
    body = pre + ["const SpeciesT conductivities = "
                  "get_species_thermal_conductivities(temperature);",
                  "const ContainerT lhs = " + " + ".join(
                      f"mole_fractions[{i}]*conductivities[{i}]"
                      for i in range(ns)) + ";",
                  "const ContainerT rhs = 1.0/(" + " + ".join(
                      f"mole_fractions[{i}]/conductivities[{i}]"
                      for i in range(ns)) + ");",
                  "return 0.5*(lhs + rhs);"]
    fn("ContainerT get_mixture_thermal_conductivity_mixavg(ContainerT "
       "temperature, SpeciesT const& mass_fractions)", body)
    body = pre + ["const Species2T bdiff_ij = "
                  "get_species_binary_mass_diffusivities(temperature);",
                  "SpeciesT out;"]
    for sp in range(ns):
        xs = " + ".join(f"mole_fractions[{i}]/bdiff_ij[{i}][{sp}]"
                        for i in range(ns))
        body.append(f"{{ const ContainerT denom = ({xs}) - mole_fractions"
                    f"[{sp}]/bdiff_ij[{sp}][{sp}];")
        body.append(f"  out[{sp}] = denom > 0.0 ? ContainerT((mmw - "
                    f"mole_fractions[{sp}]*{lit(W[sp])})/(pressure*mmw*denom))"
                    f" : ContainerT(bdiff_ij[{sp}][{sp}]/pressure); }}")
    body.append("return out;")
    fn("SpeciesT get_species_mass_diffusivities_mixavg(ContainerT pressure, "
       "ContainerT temperature, SpeciesT const& mass_fractions)", body)


class CppCodeGenerator(CodeGenerator):
    """The reference's ``cpp`` target (``codegen/cpp.py:663-700``) for this
    package's mechanism IR, with ``__host__ __device__`` methods under
    hipcc."""

    @staticmethod
    def get_name() -> str:
        return "cpp"

    @staticmethod
    def supports_overloading() -> bool:
        return False

    @staticmethod
    def generate(name: str, sol, opts: CodeGenerationOptions = None) -> str:
        if opts is None:
            opts = CodeGenerationOptions()
        if opts.directive_offload is not None:
            raise TypeError(
                "OpenMP/ACC directive based offloading is not supported for "
                "C++ code generation")
        return _header(name, as_mechanism(sol))
