"""Hand-written derivatives of the net production rates on the device.

The reference obtains derivatives by running the generated class under JAX
(``flamelets/equations.py:124-129``: ``vmap(jacfwd(local_source_terms))``).
Here the generator emits two dedicated passes per mechanism, one cell per
lane and reaction-major like the primal kernel:

- forward mode (``pyro_wdot_jvp``): ω̇ and dω̇ = ∂ω̇/∂ρ v_ρ + ∂ω̇/∂T v_T +
  ∂ω̇/∂Y v_Y;
- reverse mode (``pyro_wdot_vjp``): given w̄, the cotangents w̄ᵀ∂ω̇/∂ρ,
  w̄ᵀ∂ω̇/∂T, w̄ᵀ∂ω̇/∂Y;

and ``pyro_source_jacobian`` gives the full (ns+1) x (ns+1) Jacobian of the
flamelet source S = [vd/ρ; W_k ω̇_k/ρ] with respect to φ = [h; Y]
(``equations.py:361-389``) in two passes: one stores per reaction k_f, k_r,
q, ∂q/∂T and ∂q/∂[M] to a workspace, the other assembles each row from the
reactions of its species and the chain rule through T(h, Y) and ρ(T, Y).

``NetProductionRates`` wraps the pair in a ``torch.autograd.Function`` so
ω̇ takes part in torch's reverse mode (``backward`` is the VJP kernel) and
forward mode (``jvp`` is the JVP kernel).
"""

from __future__ import annotations

from typing import Optional

import torch
from torch.autograd.function import once_differentiable

from .thermochem import _ptr, _stream


def _flat_cells(gas, rho, temperature, mass_fractions):
    y, cs = gas._cells(mass_fractions)
    n = y.shape[1]
    r = gas._temperature_field(rho, n, cs)
    t = gas._temperature_field(temperature, n, cs)
    return r, t, y, n, cs


def _species_arg(gas, v, n, cs):
    if v is None:
        return None
    v2, vcs = gas._cells(v)
    if v2.shape[1] != n:
        raise ValueError("tangent/cotangent does not match the cell count")
    return v2


def wdot_jvp(gas, rho, temperature, mass_fractions, v_rho=None, v_T=None,
             v_Y=None, want_primal=True):
    """(ω̇, dω̇) along (v_rho, v_T, v_Y); missing tangents are zero."""
    r, t, y, n, cs = _flat_cells(gas, rho, temperature, mass_fractions)
    vr = None if v_rho is None else gas._temperature_field(v_rho, n, cs)
    vt = None if v_T is None else gas._temperature_field(v_T, n, cs)
    vy = _species_arg(gas, v_Y, n, cs)
    ns = gas.num_species
    dw = gas._out(ns, n)
    w = gas._out(ns, n) if want_primal else None
    gas._lib.call("pyro_wdot_jvp", r.data_ptr(), t.data_ptr(), y.data_ptr(),
                  y.stride(0), _ptr(vr), _ptr(vt), _ptr(vy),
                  vy.stride(0) if vy is not None else n, _ptr(w),
                  dw.data_ptr(), n, n, _stream())
    shape = (ns,) + tuple(cs)
    return (None if w is None else w.reshape(shape)), dw.reshape(shape)


def wdot_vjp(gas, rho, temperature, mass_fractions, wbar):
    """(w̄ᵀ∂ω̇/∂ρ, w̄ᵀ∂ω̇/∂T, w̄ᵀ∂ω̇/∂Y) for a cotangent w̄ of ω̇."""
    r, t, y, n, cs = _flat_cells(gas, rho, temperature, mass_fractions)
    b = _species_arg(gas, wbar, n, cs)
    g_r = gas._out(None, n)
    g_t = gas._out(None, n)
    g_y = gas._out(gas.num_species, n)
    gas._lib.call("pyro_wdot_vjp", r.data_ptr(), t.data_ptr(), y.data_ptr(),
                  y.stride(0), b.data_ptr(), b.stride(0), g_r.data_ptr(),
                  g_t.data_ptr(), g_y.data_ptr(), n, n, _stream())
    return (g_r.reshape(cs), g_t.reshape(cs),
            g_y.reshape((gas.num_species,) + tuple(cs)))


def source_jacobian(gas, temperature, pressure, mass_fractions, vd=None,
                    layout: str = "cell", out: Optional[torch.Tensor] = None,
                    source: Optional[torch.Tensor] = None,
                    work: Optional[torch.Tensor] = None):
    """∂S/∂φ per cell for S = [vd/ρ; W_k ω̇_k/ρ], φ = [h; Y], ρ = pW̄/(RT),
    at a temperature consistent with h.

    ``layout="cell"``: an (N, nv, nv) tensor, the reference's per-point
    block layout (``vmap(jacfwd(...))``, ``equations.py:124-129``).
    ``layout="entry"``: an (nv, nv, N) tensor, entry-major, which the kernel
    writes with fully coalesced stores.  ``source``, if given, an (nv, N)
    tensor, receives S.  The library's default form evaluates every row in
    one pass and needs no workspace; its two-pass form
    (``pyro_source_jacobian_form``) processes fixed chunks through a
    workspace of ``pyro_source_jacobian_work_doubles()`` doubles (the
    per-reaction cache of one chunk): pass ``work`` to reuse one."""
    y, cs = gas._cells(mass_fractions)
    n = y.shape[1]
    t = gas._temperature_field(temperature, n, cs)
    p, ps = gas._scalar_field(pressure, n, cs)
    v = None if vd is None else gas._temperature_field(vd, n, cs)
    nv = gas.num_species + 1
    if layout == "cell":
        shape, s_rc, s_i = (n, nv, nv), 1, nv * nv
    elif layout == "entry":
        shape, s_rc, s_i = (nv, nv, n), n, 1
    else:
        raise ValueError(f"unknown layout {layout!r}")
    if out is None:
        out = torch.empty(shape, dtype=torch.float64, device=gas.device)
    elif tuple(out.shape) != shape or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous {shape} tensor")
    if source is not None and (tuple(source.shape) != (nv, n)
                               or source.stride(1) != 1):
        raise ValueError(f"source must be an ({nv}, {n}) tensor")
    work = source_jacobian_workspace(gas, work)
    gas._lib.call("pyro_source_jacobian", t.data_ptr(), p.data_ptr(), ps,
                  y.data_ptr(), y.stride(0), _ptr(v), out.data_ptr(), s_rc,
                  s_i, _ptr(source),
                  source.stride(0) if source is not None else n,
                  _ptr(work), n, _stream())
    return out


def set_source_jacobian_form(gas, form: int, waves: int = 0) -> int:
    """Select how ``source_jacobian`` evaluates the Jacobian in this process
    (``pyro_source_jacobian_form``): 1 one pass, every row evaluating the
    rate coefficients of its species' reactions (the default); 0 two passes
    through a per-reaction cache.  ``waves`` (one pass): the waves sharing
    each 64 cells, 1, 2, 4 or 8, or 0 for the library default; the values
    do not depend on it.  Returns the previous form."""
    if form not in (0, 1) or waves not in (0, 1, 2, 4, 8):
        raise ValueError(f"unsupported source Jacobian form {form}, "
                         f"waves {waves}")
    return int(gas._lib.call("pyro_source_jacobian_form", int(form),
                             int(waves)))


def source_jacobian_workspace(gas, work: Optional[torch.Tensor] = None):
    """The workspace ``source_jacobian`` needs in the library's current form
    (checked, or allocated); None when it needs none."""
    need = int(gas._lib.call("pyro_source_jacobian_work_doubles"))
    if need == 0:
        return None
    if work is None:
        return torch.empty(need, dtype=torch.float64, device=gas.device)
    if work.dtype != torch.float64 or work.numel() < need or \
            not work.is_contiguous():
        raise ValueError(f"workspace must hold {need} contiguous doubles")
    return work


class NetProductionRates(torch.autograd.Function):
    """ω̇(ρ, T, Y) with the hand-written VJP as ``backward`` and JVP as
    ``jvp``.  Call as ``NetProductionRates.apply(gas, rho, T, Y)`` with
    float64 device tensors rho (cells), T (cells), Y (ns, cells)."""

    @staticmethod
    def forward(ctx, gas, rho, temperature, mass_fractions):
        ctx.gas = gas
        ctx.save_for_backward(rho, temperature, mass_fractions)
        ctx.save_for_forward(rho, temperature, mass_fractions)
        return gas.get_net_production_rates(rho, temperature, mass_fractions)

    @staticmethod
    @once_differentiable
    def backward(ctx, wbar):
        rho, temperature, y = ctx.saved_tensors
        g_r, g_t, g_y = wdot_vjp(ctx.gas, rho, temperature, y,
                                 wbar.contiguous())
        cells = tuple(y.shape[1:])

        def reduce(g, like):
            # per-cell gradient back to the input's shape: a scalar or a
            # broadcast input gets the sum over the cells it reached
            return g.reshape(cells).sum_to_size(like.shape)
        return (None,
                reduce(g_r, rho) if ctx.needs_input_grad[1] else None,
                reduce(g_t, temperature) if ctx.needs_input_grad[2] else None,
                g_y.reshape(y.shape) if ctx.needs_input_grad[3] else None)

    @staticmethod
    def jvp(ctx, _gas_t, v_rho, v_T, v_Y):
        rho, temperature, y = ctx.saved_tensors
        _, dw = wdot_jvp(ctx.gas, rho, temperature, y, v_rho, v_T,
                         None if v_Y is None else v_Y.contiguous(),
                         want_primal=False)
        return dw


def net_production_rates(gas, rho, temperature, mass_fractions):
    """Differentiable ``get_net_production_rates`` (reverse and forward
    mode through the hand-written kernels)."""
    return NetProductionRates.apply(gas, rho, temperature, mass_fractions)
