"""Derivatives of the net production rates: hand-written JVP / VJP kernels,
their torch.autograd.Function, and the per-cell flamelet source Jacobian.

References (the reference's JAX gradients cannot run here, JAX is absent):
* complex-step derivatives of the numpy oracle (error ~1e-16);
* central finite differences of the device primal, h = 1e-6 relative;
* the transpose identity <wb, J v> = <J^T wb, v>;
* the dense register Jacobian of the flamelet kernels (<= 16 species).

Tolerances (FP64): a directional derivative dw_k is compared to 1e-10 of
its gross scale sum_i |nu_ki| (|dr_f,i| + |dr_r,i|) (raw relative error is
meaningless where the derivative cancels); central differences to 1e-6 of
the same scale; the source Jacobian to 1e-9 of each row's largest entry.
"""
import numpy as np
import pytest
import torch

from oracle.flamelet_np import FlameletEquations as OracleFlamelet
from oracle.thermochem_np import Thermochemistry as Oracle
from pyrometheus_amd import bundled, get_thermochem_class
from tests.helpers import random_states

MECHS = ["sandiego", "hong", "uiuc", "gri30", "synth53"]
H = 1e-30


def _inert(m):
    return "N2" if "N2" in m.species_names else m.species_names[-1]


def cs_derivative(ref, rho, T, Y, v_rho, v_T, v_Y):
    """Complex-step dw along (v_rho, v_T, v_Y) and its gross scale
    sum_i |nu_ki| (|dr_f,i| + |dr_r,i|)."""
    r = rho + 1j * H * v_rho
    t = T + 1j * H * v_T
    y = Y + 1j * H * v_Y
    dw = np.imag(ref.get_net_production_rates(r, t, y)) / H
    c = ref.get_concentrations(r, y)
    drf = np.abs(np.imag(ref.get_fwd_rates_of_progress(t, c)) / H)
    drr = np.abs(np.imag(ref.get_rev_rates_of_progress(t, c)) / H)
    scale = np.zeros_like(dw)
    for k in range(ref.num_species):
        fwd, rev = ref._stoich_lists(k)
        for i, nu in fwd + rev:
            scale[k] += abs(nu) * (drf[i] + drr[i])
    return dw, scale


def test_oracle_complex_step_vs_central_differences():
    """The oracle's complex-step JVP (the reference used below) agrees with
    central differences of the oracle itself with second-order error."""
    m = bundled("gri30")
    ref = Oracle(m)
    T, p, Y = random_states(m, 6, seed=21)
    rho = ref.get_density(p, T, Y)
    rng = np.random.default_rng(5)
    vr, vt, vy = (rho * rng.normal(size=6), T * rng.normal(size=6),
                  Y * rng.normal(size=Y.shape))
    dw, scale = cs_derivative(ref, rho, T, Y, vr, vt, vy)
    errs = []
    for h in (1e-4, 1e-5):
        fd = (ref.get_net_production_rates(rho + h * vr, T + h * vt, Y + h * vy)
              - ref.get_net_production_rates(rho - h * vr, T - h * vt,
                                             Y - h * vy)) / (2 * h)
        errs.append(np.max(np.abs(fd - dw) / (scale + 1e-300)))
    assert errs[1] < 1e-6
    assert errs[0] / max(errs[1], 1e-300) > 50  # O(h^2): 100x per decade


_cache = {}


def setup(name, device):
    if name not in _cache:
        m = bundled(name)
        _cache[name] = (m, get_thermochem_class(m)(device=device), Oracle(m))
    return _cache[name]


def dev(a, device):
    return torch.tensor(np.asarray(a), dtype=torch.float64, device=device)


def host(t):
    return t.detach().cpu().numpy()


def _state(m, ref, n, seed):
    T, p, Y = random_states(m, n, seed=seed, inert=_inert(m))
    rho = ref.get_density(p, T, Y)
    return rho, T, p, Y


def _padded_cell_jacobian(gas, T, p, Y, vd, pad):
    """Per-cell blocks with `pad` unused doubles after each block, through
    the C ABI (s_rc = 1, s_i = nv^2 + pad); returns the (N, nv, nv) view."""
    from pyrometheus_amd.thermochem import _stream
    # the same input preparation as source_jacobian
    y, cs = gas._cells(Y)
    n = y.shape[1]
    t = gas._temperature_field(T, n, cs)
    pp, ps = gas._scalar_field(p, n, cs)
    v = gas._temperature_field(vd, n, cs)
    nv = gas.num_species + 1
    out = torch.full((n, nv * nv + pad), float("nan"), dtype=torch.float64,
                     device=y.device)
    st = gas._lib.call("pyro_source_jacobian", t.data_ptr(), pp.data_ptr(),
                       ps, y.data_ptr(), y.stride(0), v.data_ptr(),
                       out.data_ptr(), 1, nv * nv + pad, None, n, None, n,
                       _stream())
    assert st == 0, st
    assert bool(torch.isnan(out[:, nv * nv:]).all())
    return out[:, :nv * nv].reshape(n, nv, nv)


@pytest.mark.gpu
@pytest.mark.parametrize("name", MECHS)
def test_jvp_matches_complex_step(name, device):
    from pyrometheus_amd.adjoint import wdot_jvp
    m, gas, ref = setup(name, device)
    n = 515
    print(f"[jvp] {name} n={n}", flush=True)
    rho, T, p, Y = _state(m, ref, n, seed=31)
    rng = np.random.default_rng(32)
    vr, vt = rho * rng.normal(size=n), T * rng.normal(size=n)
    vy = rng.normal(size=Y.shape)
    dw_ref, scale = cs_derivative(ref, rho, T, Y, vr, vt, vy)
    w, dw = wdot_jvp(gas, dev(rho, device), dev(T, device), dev(Y, device),
                     dev(vr, device), dev(vt, device), dev(vy, device))
    err = np.max(np.abs(host(dw) - dw_ref) / (scale + 1e-300))
    assert err <= 1e-10, err
    w0 = gas.get_net_production_rates(dev(rho, device), dev(T, device),
                                      dev(Y, device))
    assert torch.equal(w, w0) or float((w - w0).abs().max()) <= 1e-12 * float(
        w0.abs().max())
    # each direction on its own, zero tangents passed as NULL
    for kw, (a, b, c) in (("T", (0 * vr, vt, 0 * vy)),
                          ("Y", (0 * vr, 0 * vt, vy))):
        d_ref, sc = cs_derivative(ref, rho, T, Y, a, b, c)
        args = dict(v_T=dev(vt, device)) if kw == "T" else dict(
            v_Y=dev(vy, device))
        _, d = wdot_jvp(gas, dev(rho, device), dev(T, device), dev(Y, device),
                        want_primal=False, **args)
        e = np.max(np.abs(host(d) - d_ref) / (sc + 1e-300))
        assert e <= 1e-10, (kw, e)


@pytest.mark.gpu
@pytest.mark.parametrize("name", MECHS)
def test_vjp_matches_complex_step(name, device):
    """g = wb^T J against the oracle Jacobian, one complex-step direction
    per input (rho, T, Y_j); tolerance 1e-10 of sum_k |wb_k| scale_k."""
    from pyrometheus_amd.adjoint import wdot_vjp
    m, gas, ref = setup(name, device)
    n = 257
    print(f"[vjp] {name} n={n}", flush=True)
    rho, T, p, Y = _state(m, ref, n, seed=41)
    rng = np.random.default_rng(42)
    wb = rng.normal(size=Y.shape)
    g_r, g_t, g_y = (host(x) for x in wdot_vjp(
        gas, dev(rho, device), dev(T, device), dev(Y, device), dev(wb, device)))
    z = np.zeros(n)
    zy = np.zeros_like(Y)
    # directions rho and T are taken relative (v = rho, v = T)
    dirs = [("rho", rho, z, zy, g_r * rho), ("T", z, T, zy, g_t * T)]
    for j in range(m.n_species):
        e = zy.copy()
        e[j] = 1.0
        dirs.append((f"Y{j}", z, z, e, g_y[j]))
    for nm, a, b, c, got in dirs:
        dw, sc = cs_derivative(ref, rho, T, Y, a, b, c)
        want = np.sum(wb * dw, axis=0)
        tol = 1e-10 * np.sum(np.abs(wb) * sc, axis=0) + 1e-300
        err = np.max(np.abs(got - want) / tol)
        assert err <= 1.0, (nm, err)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["sandiego", "gri30"])
def test_central_differences_and_transpose(name, device):
    """JVP against central differences of the device primal at h = 1e-6
    relative, and <wb, J v> = <J^T wb, v> per cell."""
    from pyrometheus_amd.adjoint import wdot_jvp, wdot_vjp
    m, gas, ref = setup(name, device)
    n = 4099
    print(f"[fd] {name} n={n}", flush=True)
    rho, T, p, Y = _state(m, ref, n, seed=51)
    rng = np.random.default_rng(52)
    vr, vt, vy = (rho * rng.normal(size=n), T * rng.normal(size=n),
                  Y * rng.normal(size=Y.shape))
    d = lambda a: dev(a, device)
    _, dw = wdot_jvp(gas, d(rho), d(T), d(Y), d(vr), d(vt), d(vy))
    h = 1e-6
    wp = gas.get_net_production_rates(d(rho + h * vr), d(T + h * vt),
                                      d(Y + h * vy))
    wm = gas.get_net_production_rates(d(rho - h * vr), d(T - h * vt),
                                      d(Y - h * vy))
    fd = host((wp - wm) / (2 * h))
    sub = slice(0, 300)  # gross scale from the oracle on a subset
    _, scale = cs_derivative(ref, rho[sub], T[sub], Y[:, sub], vr[sub],
                             vt[sub], vy[:, sub])
    err = np.max(np.abs(fd[:, sub] - host(dw)[:, sub]) / (scale + 1e-300))
    assert err <= 1e-6, err
    wb = rng.normal(size=Y.shape)
    g_r, g_t, g_y = wdot_vjp(gas, d(rho), d(T), d(Y), d(wb))
    lhs = (d(wb) * dw).sum(0)
    rhs = g_r * d(vr) + g_t * d(vt) + (g_y * d(vy)).sum(0)
    gross = (d(wb).abs() * dw.abs()).sum(0) + (g_r * d(vr)).abs() + \
        (g_t * d(vt)).abs() + (g_y * d(vy)).abs().sum(0)
    assert float(((lhs - rhs).abs() / gross).max()) <= 1e-11


@pytest.mark.gpu
def test_autograd_function(device):
    """torch reverse mode through NetProductionRates gives the VJP kernel's
    result; torch forward mode gives the JVP kernel's."""
    import torch.autograd.forward_ad as fwAD
    from pyrometheus_amd.adjoint import (net_production_rates, wdot_jvp,
                                         wdot_vjp)
    m, gas, ref = setup("gri30", device)
    n = 1000
    rho, T, p, Y = _state(m, ref, n, seed=61)
    d = lambda a: dev(a, device)
    r_t = d(rho).requires_grad_()
    t_t = d(T).requires_grad_()
    y_t = d(Y).requires_grad_()
    w = net_production_rates(gas, r_t, t_t, y_t)
    wb = torch.randn(w.shape, dtype=torch.float64, device=device,
                     generator=torch.Generator(device=device).manual_seed(7))
    (w * wb).sum().backward()
    g_r, g_t, g_y = wdot_vjp(gas, d(rho), d(T), d(Y), wb)
    assert torch.equal(r_t.grad, g_r)
    assert torch.equal(t_t.grad, g_t)
    assert torch.equal(y_t.grad, g_y)
    # composition: rho computed from T by torch ops, gradient flows through
    t2 = d(T).requires_grad_()
    rho2 = d(p) * gas.get_mixture_molecular_weight(d(Y)) / (gas.gas_constant * t2)
    (net_production_rates(gas, rho2, t2, d(Y)) * wb).sum().backward()
    gr2, gt2, _ = wdot_vjp(gas, rho2.detach(), d(T), d(Y), wb)
    want = gt2 + gr2 * (-rho2.detach() / d(T))
    assert float(((t2.grad - want).abs() / (want.abs() + gt2.abs()
                                             + 1e-300)).max()) <= 1e-12
    # forward mode
    vr, vt, vy = d(rho) * 0.1, d(T) * 0.01, d(Y) * 0.3
    with fwAD.dual_level():
        out = net_production_rates(gas, fwAD.make_dual(d(rho), vr),
                                   fwAD.make_dual(d(T), vt),
                                   fwAD.make_dual(d(Y), vy))
        tangent = fwAD.unpack_dual(out).tangent
    _, dw = wdot_jvp(gas, d(rho), d(T), d(Y), vr, vt, vy, want_primal=False)
    assert torch.equal(tangent, dw)
    # a scalar (0-d) density and a (1,) temperature that require grad get
    # the per-cell gradients summed over the cells they were broadcast to
    r0 = torch.tensor(float(rho[0]), dtype=torch.float64, device=device,
                      requires_grad=True)
    t0 = torch.full((1,), float(T[0]), dtype=torch.float64, device=device,
                    requires_grad=True)
    (net_production_rates(gas, r0, t0, d(Y)) * wb).sum().backward()
    g_r0, g_t0, _ = wdot_vjp(gas, r0.detach(), t0.detach(), d(Y), wb)
    assert r0.grad.shape == () and t0.grad.shape == (1,)
    assert torch.allclose(r0.grad, g_r0.sum(), rtol=1e-13, atol=0.0)
    assert torch.allclose(t0.grad, g_t0.sum().reshape(1), rtol=1e-13,
                          atol=0.0)


# uiuc (nv = 8, no per-species Gibbs factors) stages each per-cell row as
# one chunk of 8 entries, sandiego (nv = 10) as two overlapping chunks
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["sandiego", "gri30", "uiuc"])
def test_source_jacobian_forms_agree(name, device):
    """The one-pass form splits each 64 cells' rows over 1, 2, 4 or 8 waves
    (by default 4 above 32,768 cells, 8 below): every split gives the same
    values bit for bit, also a 67-cell call against the first cells of a
    40,000-cell call.  The two-pass form (per-reaction cache; one thread
    per cell above 32,768 cells, the waves of a block splitting the rows
    below) is bit for bit alike across its own forms too, and agrees with
    the one pass to rounding (K_c^-1 there is a product of per-species
    Gibbs factors, here one exponential per reaction)."""
    import torch
    from pyrometheus_amd.adjoint import (set_source_jacobian_form,
                                         source_jacobian)
    m, gas, ref = setup(name, device)
    n = 40000
    print(f"[source jac forms] {name} n={n}", flush=True)
    rho, T, p, Y = _state(m, ref, n, seed=73)
    d = lambda a: dev(a, device)
    vd = np.random.default_rng(74).uniform(0, 1e3, n)
    Tt, pt, Yt, vt = d(T), d(p), d(Y), d(vd)
    k = 67
    small_args = (Tt[:k].contiguous(), pt[:k].contiguous(),
                  Yt[:, :k].contiguous(), vt[:k].contiguous())
    try:
        big = source_jacobian(gas, Tt, pt, Yt, vt, layout="entry")
        assert bool(torch.isfinite(big).all())
        small = source_jacobian(gas, *small_args, layout="entry")
        assert torch.equal(big[:, :, :k], small)
        # per-cell blocks: full blocks of 64 cells pass their rows through
        # LDS, the partial last block stores directly; both give the
        # entry-major values, as does a padded block stride
        bigc = source_jacobian(gas, Tt, pt, Yt, vt, layout="cell")
        assert torch.equal(bigc, big.permute(2, 0, 1))
        assert torch.equal(_padded_cell_jacobian(gas, Tt, pt, Yt, vt, 5),
                           bigc)
        for waves in (1, 2, 4, 8):
            set_source_jacobian_form(gas, 1, waves)
            again = source_jacobian(gas, *small_args, layout="entry")
            assert torch.equal(again, small), waves
            again = source_jacobian(gas, *small_args, layout="cell")
            assert torch.equal(again, small.permute(2, 0, 1)), waves
        set_source_jacobian_form(gas, 0)
        big0 = source_jacobian(gas, Tt, pt, Yt, vt, layout="entry")
        small0 = source_jacobian(gas, *small_args, layout="entry")
        assert torch.equal(big0[:, :, :k], small0)
    finally:
        set_source_jacobian_form(gas, 1)
    # one pass against two: per row, relative to the row's largest entry
    scale = big0.abs().amax(dim=1, keepdim=True) + 1e-300
    err = float(((big - big0).abs() / scale).max())
    print(f"  one pass vs cached: {err:.2e}", flush=True)
    assert err <= 1e-11, err


@pytest.mark.gpu
@pytest.mark.parametrize("form", [1, 0])
@pytest.mark.parametrize("name", ["sandiego", "gri30", "uiuc"])
def test_source_jacobian(name, form, device):
    """(ns+1)^2 source Jacobian per cell against the oracle's complex step
    through T(h, Y) (both layouts; the one-pass form 1 and the two-pass
    form 0), and, for the 9-species mechanism, against the dense register
    Jacobian of the flamelet kernel."""
    from pyrometheus_amd.adjoint import (set_source_jacobian_form,
                                         source_jacobian)
    m, gas, ref = setup(name, device)
    n = 67
    print(f"[source jac] {name} form={form} n={n}", flush=True)
    set_source_jacobian_form(gas, form)
    try:
        _check_source_jacobian(m, gas, ref, n, device)
    finally:
        set_source_jacobian_form(gas, 1)


def _check_source_jacobian(m, gas, ref, n, device):
    from pyrometheus_amd.adjoint import source_jacobian
    rho, T, p, Y = _state(m, ref, n, seed=71)
    T = np.clip(T, 900.0, None)  # keep the oracle's T(h) Newton away from 300 K
    h = ref.get_mixture_enthalpy_mass(T, Y)
    phi = np.vstack([h[None], Y])
    vd = np.random.default_rng(72).uniform(0, 1e3, n)
    fl = OracleFlamelet(m, np.linspace(0, 1, n), Y[:, 0], Y[:, -1])
    J_ref = fl.source_jacobian(phi, vd, T, p)           # (n, nv, nv)
    S_ref, T_ref = fl.source_terms(phi, vd, T, p)
    d = lambda a: dev(a, device)
    S = torch.empty((m.n_species + 1, n), dtype=torch.float64, device=device)
    J = host(source_jacobian(gas, d(T_ref), d(p), d(Y), d(vd), source=S))
    Je = host(source_jacobian(gas, d(T_ref), d(p), d(Y), d(vd),
                              layout="entry"))
    assert np.array_equal(J, np.moveaxis(Je, 2, 0))
    rowmax = np.abs(J_ref).max(axis=2, keepdims=True) + 1e-300
    err = np.max(np.abs(J - J_ref) / rowmax)
    assert err <= 1e-9, err
    assert np.max(np.abs(host(S) - S_ref) / (np.abs(S_ref).max(1, keepdims=True)
                                            + 1e-300)) <= 1e-10
    if gas._lib.has("pyro_flamelet_source"):
        from pyrometheus_amd.flamelets.engine import FlameletBatch
        b = FlameletBatch(gas, [np.linspace(0, 1, n)], Y[:, 0], Y[:, -1])
        b.source(d(phi), d(vd), d(T_ref), d(p), jac=True)
        Jd = host(b.J)
        assert np.max(np.abs(Jd - J) / rowmax) <= 1e-11


@pytest.mark.gpu
def test_source_jacobian_conservation_large(device):
    """2^20 GRI cells, the full (ns+1)^2 source Jacobian: mass and element
    conservation of S_k = W_k w_k / rho hold for every state, so every
    column of the species rows sums to zero, unweighted (sum_k S_k = 0) and
    with weights a_ek / W_k (sum_k a_ek w_k = 0), to rounding of the column's
    gross sum; the one-pass and two-pass forms agree on the first 65,536
    cells to 1e-11 of each row's largest entry."""
    from pyrometheus_amd.adjoint import (set_source_jacobian_form,
                                         source_jacobian)
    m, gas, ref = setup("gri30", device)
    n = 1 << 20
    print(f"[source jac conservation] gri30 n={n}", flush=True)
    g = torch.Generator(device=device).manual_seed(83)
    T = 300.0 + 2200.0 * torch.rand(n, generator=g, dtype=torch.float64,
                                    device=device)
    Y = torch._standard_gamma(torch.full((m.n_species, n), 0.5,
                                         dtype=torch.float64, device=device),
                              generator=g)
    Y /= Y.sum(0, keepdim=True)
    J = source_jacobian(gas, T, 101325.0, Y, layout="entry")
    Js = J[1:]                                           # (ns, nv, n)
    mass = Js.sum(0).abs() / (Js.abs().sum(0) + 1e-300)
    E = torch.tensor(m.element_matrix(), dtype=torch.float64, device=device)
    Ew = E / gas.molecular_weights[None, :]              # a_ek / W_k
    flat = Js.reshape(m.n_species, -1)                    # contiguous
    el = (Ew @ flat).abs() / ((Ew.abs() @ flat.abs()) + 1e-300)
    print(f"  mass {float(mass.max()):.2e} elements {float(el.max()):.2e}",
          flush=True)
    assert bool(torch.isfinite(J).all())
    assert float(mass.max()) <= 1e-12
    assert float(el.max()) <= 1e-12
    k = 1 << 16
    head = J[:, :, :k].clone()
    del J, Js
    torch.cuda.empty_cache()
    set_source_jacobian_form(gas, 0)
    try:
        J0 = source_jacobian(gas, T[:k].contiguous(), 101325.0,
                             Y[:, :k].contiguous(), layout="entry")
    finally:
        set_source_jacobian_form(gas, 1)
    scale = J0.abs().amax(dim=1, keepdim=True) + 1e-300
    err = float(((head - J0).abs() / scale).max())
    print(f"  one pass vs two passes: {err:.2e}", flush=True)
    assert err <= 1e-11, err


@pytest.mark.gpu
def test_large_batch_invariants(device):
    """2^20 GRI cells: the JVP conserves mass and elements like w itself
    (sum_k W_k dw_k = 0, sum_k a_ek dw_k = 0), and the transpose identity
    holds everywhere."""
    from pyrometheus_amd.adjoint import wdot_jvp, wdot_vjp
    m, gas, ref = setup("gri30", device)
    n = 1 << 20
    g = torch.Generator(device=device).manual_seed(81)
    T = 300.0 + 2200.0 * torch.rand(n, generator=g, dtype=torch.float64,
                                    device=device)
    Y = torch._standard_gamma(torch.full((m.n_species, n), 0.5,
                                         dtype=torch.float64, device=device),
                              generator=g)
    Y /= Y.sum(0, keepdim=True)
    rho = gas.get_density(101325.0, T, Y)
    vr = rho * torch.randn(n, generator=g, dtype=torch.float64, device=device)
    vt = T * torch.randn(n, generator=g, dtype=torch.float64, device=device)
    vy = torch.randn(Y.shape, generator=g, dtype=torch.float64, device=device)
    w, dw = wdot_jvp(gas, rho, T, Y, vr, vt, vy)
    W = gas.molecular_weights[:, None]
    gross = (W * dw).abs().sum(0) + 1e-300
    assert float(((W * dw).sum(0).abs() / gross).max()) <= 1e-11
    E = torch.tensor(m.element_matrix(), dtype=torch.float64, device=device)
    el = E @ dw
    el_gross = E.abs() @ dw.abs() + 1e-300
    assert float((el.abs() / el_gross).max()) <= 1e-11
    wb = torch.randn(Y.shape, generator=g, dtype=torch.float64, device=device)
    g_r, g_t, g_y = wdot_vjp(gas, rho, T, Y, wb)
    lhs = (wb * dw).sum(0)
    rhs = g_r * vr + g_t * vt + (g_y * vy).sum(0)
    tot = (wb * dw).abs().sum(0) + (g_r * vr).abs() + (g_t * vt).abs() + \
        (g_y * vy).abs().sum(0)
    assert float(((lhs - rhs).abs() / tot).max()) <= 1e-11
    assert torch.isfinite(g_y).all()
