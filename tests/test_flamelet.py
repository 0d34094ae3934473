"""Flamelet path: oracle self-consistency (CPU) and device parity (GPU).

Tolerances (FP64; Newton |v| < 1e-8 and CN |v| < 1e-9 as in the reference
notebook -- the unscaled update norm is dominated by the enthalpy row and
bottoms out near 1e-9 in double precision): source terms 1e-12 of the gross production scale;
source Jacobian 1e-9 of each row's largest entry (oracle by complex step);
residual 1e-12 of |chi/2 Lap phi| + |S| + |BC|; converged flamelets agree
with the oracle's to 1e-8 relative; a flamelet solved
alone and inside a batch agrees bit for bit.
"""
import numpy as np
import pytest

from oracle import flamelet_np as F
from pyrometheus_amd.mechanism import ONE_ATM, bundled


def _setup(n=33, chi_st=100.0, t_ox=500.0, cluster=False):
    m = bundled("sandiego")
    y_ox, y_fu, h_ox, h_fu = F.stream_compositions(m, t_ox=t_ox, t_fu=300.0)
    z_st = F.stoichiometric_mixture_fraction(m, y_ox, y_fu)
    x = F.clustered_grid(n, z_st) if cluster else F.uniform_grid(n)
    phi0, _ = F.burke_schumann(m, x, y_ox, y_fu, h_ox, h_fu)
    chi = F.dissipation_profile(x, z_st, chi_st)
    gas = F.Thermochemistry(m)
    T0 = F.temperature_from_enthalpy(gas, phi0[0], phi0[1:], 1500.0)
    return m, x, y_ox, y_fu, h_ox, h_fu, z_st, phi0, chi, T0


def test_oracle_jacobian_fd_eoc():
    """Complex-step source Jacobian against central differences: the error
    falls at second order in the step (mirrors test/test.py:349-374)."""
    m, x, y_ox, y_fu, h_ox, h_fu, z_st, phi0, chi, T0 = _setup(n=9)
    eq = F.FlameletEquations(m, x, y_ox, y_fu)
    vd = np.zeros(len(x))
    J = eq.source_jacobian(phi0, vd, T0, ONE_ATM)
    errs = []
    for d in (1e-3, 1e-4):
        Jfd = np.empty_like(J)
        for v in range(eq.nv):
            scale = np.maximum(np.abs(phi0[v]), 1e-3) if v == 0 else 1e-2
            hstep = d * scale
            pp, pm = phi0.copy(), phi0.copy()
            pp[v] += hstep
            pm[v] -= hstep
            Sp, _ = eq.source_terms(pp, vd, T0, ONE_ATM)
            Sm, _ = eq.source_terms(pm, vd, T0, ONE_ATM)
            Jfd[:, :, v] = ((Sp - Sm) / (2 * hstep)).T
        errs.append(np.linalg.norm(Jfd[2:-2] - J[2:-2]) / np.linalg.norm(J[2:-2]))
    eoc = np.log(errs[0] / errs[1]) / np.log(10.0)
    assert errs[1] < 1e-6 and eoc > 1.8, (errs, eoc)


def test_oracle_solve_invariants():
    m, x, y_ox, y_fu, h_ox, h_fu, z_st, phi0, chi, T0 = _setup(n=25)
    s = F.FlameletSolver(m, x, y_ox, y_fu)
    phi, T, ok = s.solve(10, 1e-8, 20, 1e-9, 1e-5, 10, False, 10, chi,
                         np.zeros(len(x)), T0, ONE_ATM, h_ox, h_fu, phi0)
    assert ok
    # vd = 0: the enthalpy row decouples and is linear in Z
    lin = h_ox + (h_fu - h_ox) * x
    assert np.max(np.abs(phi[0] - lin)) < 1e-8 * abs(h_ox - h_fu)
    assert np.max(np.abs(phi[1:].sum(0) - 1)) < 1e-12
    assert T.max() > 1500.0


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def gas(device):
    from pyrometheus_amd import get_thermochem_class
    return get_thermochem_class(bundled("sandiego"))(device=device)


def _dev(a, device):
    import torch
    return torch.as_tensor(np.asarray(a), dtype=torch.float64, device=device)


def _state(phi, device):
    from pyrometheus_amd.flamelets.state import FlameletState
    return FlameletState(_dev(phi[0], device), _dev(phi[1:], device))


def _domain(x):
    from pyrometheus_amd.flamelets.domain import Domain, DomainConfig
    d = Domain(DomainConfig(num_x=len(x), x_l=0.0, x_r=1.0))
    d.x = np.asarray(x)
    from pyrometheus_amd.flamelets.domain import laplacian_coefficients
    d.laplacian = laplacian_coefficients(d.x)
    return d


@pytest.mark.gpu
@pytest.mark.parametrize("cluster", [False, True])
def test_gpu_source_jacobian_residual(gas, device, cluster):
    from pyrometheus_amd.flamelets.solver import FlameletEquations
    m, x, y_ox, y_fu, h_ox, h_fu, z_st, phi0, chi, T0 = _setup(n=41,
                                                               cluster=cluster)
    rng = np.random.default_rng(3)
    phi = phi0.copy()
    phi[1:, 1:-1] *= rng.uniform(0.8, 1.2, phi[1:, 1:-1].shape)
    phi[1:, 1:-1] += 1e-4 * rng.uniform(size=phi[1:, 1:-1].shape)
    phi[1:] /= phi[1:].sum(0)
    vd = rng.uniform(0, 1e3, len(x))
    ref = F.FlameletEquations(m, x, y_ox, y_fu)
    eq = FlameletEquations(gas, _domain(x), y_ox, y_fu)
    st = _state(phi, device)
    # source terms
    S_ref, _ = ref.source_terms(phi, vd, T0, ONE_ATM)
    S = eq.source_terms(st, _dev(vd, device), _dev(T0, device), ONE_ATM)
    rho = ref.gas.get_density(ONE_ATM, F.temperature_from_enthalpy(
        ref.gas, phi[0], phi[1:], T0), phi[1:])
    T = F.temperature_from_enthalpy(ref.gas, phi[0], phi[1:], T0)
    wscale = ref.gas.production_rate_scale(rho, T, phi[1:])
    sscale = m.molecular_weights[:, None] * wscale / rho
    got = np.vstack([S.enthalpy.cpu().numpy()[None], S.mass_fractions.cpu().numpy()])
    assert np.max(np.abs(got[0] - S_ref[0]) / np.abs(S_ref[0])) < 1e-13
    assert np.max(np.abs(got[1:] - S_ref[1:]) / (sscale + 1e-300)) < 1e-12
    # source Jacobian against the complex-step oracle
    J_ref = ref.source_jacobian(phi, vd, T0, ONE_ATM)
    J = eq.source_jacobian(st, _dev(vd, device), _dev(T0, device),
                           ONE_ATM).cpu().numpy()
    rowmax = np.max(np.abs(J_ref), axis=2, keepdims=True) + 1e-300
    err = np.max(np.abs(J - J_ref) / rowmax)
    assert err < 1e-9, err
    # residual and blocks
    R_ref = ref.rhs(phi, chi, vd, T0, ONE_ATM, h_ox, h_fu)
    R = eq.rhs(st, _dev(chi, device), _dev(vd, device), _dev(T0, device),
               ONE_ATM, h_ox, h_fu)
    Rg = np.vstack([R.enthalpy.cpu().numpy()[None], R.mass_fractions.cpu().numpy()])
    # rounding scale of the stencil: the terms, not their (cancelling) sum
    a, b, c = (np.abs(v) for v in ref.abc)
    ap = np.abs(phi)
    lap = b * ap
    lap[:, 1:] += a[1:] * ap[:, :-1]
    lap[:, :-1] += c[:-1] * ap[:, 1:]
    lap *= 0.5 * chi
    rscale = lap + np.abs(S_ref) + np.abs(ref.boundary_conditions(h_ox, h_fu))
    rscale[1:] += sscale
    assert np.max(np.abs(Rg - R_ref) / (rscale + 1e-300)) < 1e-12
    L_ref, C_ref, U_ref = ref.jac(phi, chi, vd, T0, ONE_ATM)
    L, C, U = (t.cpu().numpy() for t in eq.jac(
        st, _dev(chi, device), _dev(vd, device), _dev(T0, device), ONE_ATM))
    assert np.max(np.abs(L - L_ref)) <= 1e-14 * np.max(np.abs(L_ref))
    assert np.max(np.abs(U - U_ref)) <= 1e-14 * np.max(np.abs(U_ref))
    cmax = np.max(np.abs(C_ref), axis=2, keepdims=True)
    assert np.max(np.abs(C - C_ref) / cmax) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("semantics", ["exact", "reference"])
def test_gpu_newton_step_and_solve(gas, device, semantics):
    from pyrometheus_amd.flamelets.solver import FlameletSolver
    m, x, y_ox, y_fu, h_ox, h_fu, z_st, phi0, chi, T0 = _setup(n=33)
    vd = np.zeros(len(x))
    ref = F.FlameletSolver(m, x, y_ox, y_fu, semantics=semantics)
    solver = FlameletSolver(_domain(x), gas, (y_ox, y_fu),
                            semantics=semantics)
    st = _state(phi0, device)
    args = (_dev(chi, device), _dev(vd, device), _dev(T0, device), ONE_ATM,
            h_ox, h_fu)
    v = solver.flamelet_newton_step(st, *args)
    v_ref = ref.flamelet_newton_step(phi0, chi, vd, T0, ONE_ATM, h_ox, h_fu)
    vg = np.vstack([v.enthalpy.cpu().numpy()[None], v.mass_fractions.cpu().numpy()])
    assert np.max(np.abs(vg - v_ref), axis=1).max() <= 1e-8 * np.abs(v_ref).max()
    params = (10, 1e-8, 20, 1e-9, 1e-5, 10, False, 10)
    phi_r, T_r, ok_r = ref.solve(*params, chi, vd, T0, ONE_ATM, h_ox, h_fu, phi0)
    state, T = solver.solve(False, *params, *args[:3], ONE_ATM, h_ox, h_fu, st)
    assert ok_r and solver.converged
    pg = np.vstack([state.enthalpy.cpu().numpy()[None],
                    state.mass_fractions.cpu().numpy()])
    assert np.max(np.abs(pg[0] - phi_r[0])) <= 1e-8 * np.abs(phi_r[0]).max()
    assert np.max(np.abs(pg[1:] - phi_r[1:])) <= 1e-8
    assert np.max(np.abs(T.cpu().numpy() - T_r)) <= 1e-5
    lin = h_ox + (h_fu - h_ox) * x
    assert np.max(np.abs(pg[0] - lin)) < 1e-8 * abs(h_ox - h_fu)
    assert np.max(np.abs(pg[1:].sum(0) - 1)) < 1e-12


def _problems(n_flamelets, seed, sizes=None):
    from pyrometheus_amd.flamelets.solver import FlameletProblem
    rng = np.random.default_rng(seed)
    m = bundled("sandiego")
    y_ox, y_fu, h_ox, h_fu = F.stream_compositions(m, t_ox=800.0, t_fu=300.0)
    z_st = F.stoichiometric_mixture_fraction(m, y_ox, y_fu)
    gas = F.Thermochemistry(m)
    out = []
    for j in range(n_flamelets):
        n = sizes[j] if sizes else int(rng.integers(24, 80))
        x = F.clustered_grid(n, z_st) if j % 2 else F.uniform_grid(n)
        chi = F.dissipation_profile(x, z_st, float(np.exp(rng.uniform(0, np.log(1000)))))
        phi, _ = F.burke_schumann(m, x, y_ox, y_fu, h_ox, h_fu)
        T0 = F.temperature_from_enthalpy(gas, phi[0], phi[1:], 1500.0)
        out.append(FlameletProblem(x=x, chi=chi, phi=phi, t_guess=T0,
                                   y_ox=y_ox, y_fu=y_fu, h_ox=h_ox, h_fu=h_fu))
    return out


def _same(a, b):
    import torch
    return bool(((a == b) | (torch.isnan(a) & torch.isnan(b))).all())


@pytest.mark.gpu
def test_gpu_batch_alone_equals_in_mix(gas, device):
    from pyrometheus_amd.flamelets.solver import solve_batch
    probs = _problems(6, seed=4)
    mix = solve_batch(gas, probs, equilibrate_steps=24)
    for j in (0, 3, 5):
        alone = solve_batch(gas, [probs[j]], equilibrate_steps=24)[0]
        assert alone.converged == mix[j].converged
        assert alone.newton_iterations == mix[j].newton_iterations
        for nm, a, b in (("h", alone.state.enthalpy, mix[j].state.enthalpy),
                         ("Y", alone.state.mass_fractions,
                          mix[j].state.mass_fractions),
                         ("T", alone.temperature, mix[j].temperature)):
            assert _same(a, b), (j, nm, float((a - b).abs().max()))
    assert all(r.converged for r in mix), [r.converged for r in mix]
    # the hotter flame sits at the lower dissipation rate
    tmax = [float(r.temperature.max()) for r in mix]
    chis = [p.chi[len(p.chi) // 2] for p in probs]
    assert tmax[int(np.argmin(chis))] > tmax[int(np.argmax(chis))]


@pytest.mark.gpu
def test_gpu_graph_replay_matches_eager(gas, device):
    """Every iteration replayed from a captured HIP graph -- host work and a
    synchronisation between replays, a new capture after each re-pack --
    gives the update norms of every iteration and the final states of the
    eager launches bit for bit."""
    from pyrometheus_amd.flamelets.stepper import (FlameletStepper,
                                                   SolveSettings)
    probs = _problems(6, seed=4)
    out = {}
    for graphs in (False, True):
        st = FlameletStepper(gas, SolveSettings(equilibrate_steps=24),
                             graphs=graphs, device_control=False)
        st.norm_log = []
        handles = st.add(probs)
        res = st.run()
        out[graphs] = (st, [res[h] for h in handles])
    (st_e, eager), (st_g, graph) = out[False], out[True]
    assert st_e.graph_captures == 0 and st_g.graph_captures >= 1
    assert len(st_g.norm_log) == len(st_e.norm_log) > 20
    for it, (a, b) in enumerate(zip(st_e.norm_log, st_g.norm_log)):
        assert np.array_equal(a, b), it
    for j, (a, b) in enumerate(zip(eager, graph)):
        assert a.converged == b.converged
        for nm, u, v in (("h", a.state.enthalpy, b.state.enthalpy),
                         ("Y", a.state.mass_fractions,
                          b.state.mass_fractions),
                         ("T", a.temperature, b.temperature)):
            assert _same(u, v), (j, nm)


@pytest.mark.gpu
def test_gpu_batch_matches_oracle(gas, device):
    from pyrometheus_amd.flamelets.solver import solve_batch
    probs = _problems(3, seed=7, sizes=[21, 30, 17])
    res = solve_batch(gas, probs, equilibrate_steps=24)
    m = bundled("sandiego")
    for p, r in zip(probs, res):
        ref = F.FlameletSolver(m, p.x, p.y_ox, p.y_fu)
        phi_e, T_e = ref.equilibrate(p.phi, p.t_guess, p.pressure, p.h_ox,
                                     p.h_fu)
        phi_r, T_r, ok = ref.solve(10, 1e-8, 20, 1e-9, 1e-5, 10, False, 10,
                                   p.chi, np.zeros(len(p.x)), T_e,
                                   p.pressure, p.h_ox, p.h_fu, phi_e)
        assert ok and r.converged
        Y = r.state.mass_fractions.cpu().numpy()
        assert np.max(np.abs(Y - phi_r[1:])) <= 1e-8
        assert np.max(np.abs(r.temperature.cpu().numpy() - T_r)) <= 1e-5


@pytest.mark.gpu
def test_gpu_fine_grid_scaled_norm(gas, device):
    """Fine clustered grids (N in [500, 900]): the unscaled update norm of
    the enthalpy row stalls far above 1e-8 J/kg, so the solves use the scaled
    norm.  Batched and plain solves converge and agree to rounding, and the
    converged residual of every row is below 1e-10 of its gross scale
    |chi/2| (|a phi_-| + |b phi| + |c phi_+|) + |S|."""
    import torch
    from pyrometheus_amd.flamelets.solver import (FlameletSolver,
                                                  solve_batch)
    from pyrometheus_amd.flamelets.workload import flamelet_batch, solve_plain
    doms, probs, chis = flamelet_batch(gas, n_flamelets=3, seed=11,
                                       n_min=500, n_max=900)
    res = solve_batch(gas, probs, equilibrate_steps=24, newton_norm="scaled")
    assert all(r.converged for r in res)
    st, T, ok = solve_plain(gas, doms[0], probs[0], newton_norm="scaled")
    assert ok
    dY = float((st.mass_fractions - res[0].state.mass_fractions).abs().max())
    assert dY <= 1e-10, dY
    for d, p, r in zip(doms, probs, res):
        eq = FlameletSolver(d, gas, (p.y_ox, p.y_fu)).gov_eqns
        vd = torch.zeros(len(p.x), dtype=torch.float64, device=device)
        R = eq.rhs(r.state, p.chi, vd, r.temperature, p.pressure, p.h_ox,
                   p.h_fu)
        S = eq.source_terms(r.state, vd, r.temperature, p.pressure)
        phi = np.vstack([r.state.enthalpy.cpu().numpy()[None],
                         r.state.mass_fractions.cpu().numpy()])
        Rn = np.vstack([R.enthalpy.cpu().numpy()[None],
                        R.mass_fractions.cpu().numpy()])
        Sn = np.vstack([S.enthalpy.cpu().numpy()[None],
                        S.mass_fractions.cpu().numpy()])
        a, b, c = d.laplacian
        ap = np.abs(phi)
        gross = np.abs(b) * ap
        gross[:, 1:] += np.abs(a[1:]) * ap[:, :-1]
        gross[:, :-1] += np.abs(c[:-1]) * ap[:, 1:]
        gross = 0.5 * np.abs(p.chi) * gross + np.abs(Sn)
        rel = np.abs(Rn).max(axis=1) / gross.max(axis=1)
        assert rel.max() <= 1e-10, rel


@pytest.mark.gpu
def test_gpu_general_jacobian_matches_register_form(gas, device):
    """The general source Jacobian (pyro_source_jacobian, the form
    mechanisms with more than 16 species use) against the register form of
    pyro_flamelet_source: the same flamelets converge to the same solutions
    (1e-10 in Y, 1e-6 K)."""
    from pyrometheus_amd.flamelets.solver import solve_batch
    probs = _problems(3, seed=12, sizes=[30, 41, 25])
    kw = dict(equilibrate_steps=24, newton_norm="scaled")
    a = solve_batch(gas, probs, **kw)
    b = solve_batch(gas, probs, jacobian="general", **kw)
    for ra, rb in zip(a, b):
        assert ra.converged and rb.converged
        dy = float((ra.state.mass_fractions - rb.state.mass_fractions)
                   .abs().max())
        dt = float((ra.temperature - rb.temperature).abs().max())
        assert dy <= 1e-10 and dt <= 1e-6, (dy, dt)


@pytest.mark.gpu
def test_gpu_53_species_flamelets(device):
    """GRI-3.0 flamelets (nv = 54): every flamelet converges, one solved
    alone equals itself in the batch bit for bit, h is linear in Z, the
    mass fractions sum to one and the converged residual is below 1e-10 of
    its gross scale (enthalpy row; mass-fraction rows together)."""
    import torch
    from pyrometheus_amd import get_thermochem_class
    from pyrometheus_amd.flamelets.solver import FlameletSolver, solve_batch
    from pyrometheus_amd.flamelets.workload import flamelet_batch
    g53 = get_thermochem_class(bundled("gri30"))(device=device)
    doms, probs, chis = flamelet_batch(g53, n_flamelets=3, seed=13, n_min=24,
                                       n_max=40, chi_max=300.0)
    print("[gri flamelets] sizes", [len(p.x) for p in probs], flush=True)
    kw = dict(equilibrate_steps=24, newton_norm="scaled")
    res = solve_batch(g53, probs, **kw)
    assert all(r.converged for r in res), [r.converged for r in res]
    alone = solve_batch(g53, [probs[1]], **kw)[0]
    assert _same(alone.state.mass_fractions, res[1].state.mass_fractions)
    assert _same(alone.temperature, res[1].temperature)
    for d, p, r in zip(doms, probs, res):
        h = r.state.enthalpy.cpu().numpy()
        lin = p.h_ox + (p.h_fu - p.h_ox) * p.x
        assert np.max(np.abs(h - lin)) <= 1e-8 * abs(p.h_ox - p.h_fu)
        assert float((r.state.mass_fractions.sum(0) - 1).abs().max()) <= 1e-12
        assert float(r.temperature.max()) > 1500.0
        eq = FlameletSolver(d, g53, (p.y_ox, p.y_fu)).gov_eqns
        vd = torch.zeros(len(p.x), dtype=torch.float64, device=device)
        R = eq.rhs(r.state, p.chi, vd, r.temperature, p.pressure, p.h_ox,
                   p.h_fu)
        S = eq.source_terms(r.state, vd, r.temperature, p.pressure)
        phi = np.vstack([h[None], r.state.mass_fractions.cpu().numpy()])
        Rn = np.vstack([R.enthalpy.cpu().numpy()[None],
                        R.mass_fractions.cpu().numpy()])
        Sn = np.vstack([S.enthalpy.cpu().numpy()[None],
                        S.mass_fractions.cpu().numpy()])
        a, b, c = d.laplacian
        ap = np.abs(phi)
        gross = np.abs(b) * ap
        gross[:, 1:] += np.abs(a[1:]) * ap[:, :-1]
        gross[:, :-1] += np.abs(c[:-1]) * ap[:, 1:]
        gross = 0.5 * np.abs(p.chi) * gross + np.abs(Sn)
        # the enthalpy row against its own scale; the mass-fraction rows
        # (one unit, 1/s) against the largest of theirs: species that stay
        # at round-off level (carbon species in H2/air, 1e-47..1e-25) carry
        # no meaningful relative residual of their own
        rel_h = np.abs(Rn[0]).max() / gross[0].max()
        rel_y = np.abs(Rn[1:]).max() / gross[1:].max()
        assert max(rel_h, rel_y) <= 1e-10, (rel_h, rel_y)


@pytest.mark.gpu
def test_gpu_notebook_case(gas, device):
    """The reference notebook's run (examples/flamelet/flamelet.ipynb cells
    2-7: sandiego, N = 101, T_ox = 500 K, chi_st = 100 then a restart at
    1000, HP-equilibrium start) on the device FlameletSolver with its
    default (reference) semantics: every printed T_max of both time marches
    to 3 decimals, every printed Newton |v| above the rounding floor to 5
    digits, the last one (rounding) within a factor 3 and below the
    tolerance, the same iteration counts; the converged states match the
    oracle's to 1e-10 in Y and 1e-7 K."""
    from pyrometheus_amd.flamelets.solver import FlameletSolver
    from tests.test_flamelet_pins import (NB_TMAX_100, NB_TMAX_1000,
                                          NB_V_100, NB_V_1000,
                                          ROUNDING_FLOOR, notebook_case)
    m, x, y_ox, y_fu, h_ox, h_fu, z_st, phi, T = notebook_case()
    solver = FlameletSolver(_domain(x), gas, (y_ox, y_fu))
    assert solver.semantics == "reference"
    ref = F.FlameletSolver(m, x, y_ox, y_fu, semantics="reference")
    st, Td = _state(phi, device), _dev(T, device)
    vd = np.zeros(len(x))
    for chi_st, tmax, vn in ((100.0, NB_TMAX_100, NB_V_100),
                             (1000.0, NB_TMAX_1000, NB_V_1000)):
        print(f"[notebook] chi_st={chi_st}", flush=True)
        chi = F.dissipation_profile(x, z_st, chi_st)
        solver.log = []
        st, Td = solver.solve(False, 10, 1e-8, 20, 1e-9, 1e-5, 10, False, 10,
                              _dev(chi, device), _dev(vd, device), Td,
                              ONE_ATM, h_ox, h_fu, st)
        assert solver.converged
        cn = [e for e in solver.log if e[0] == "cn"]
        nw = [e for e in solver.log if e[0] == "newton"]
        assert [f"{e[3]:.3f}" for e in cn] == tmax
        assert len(nw) == len(vn)
        for (_, it, v), r in zip(nw, vn):
            if r > ROUNDING_FLOOR:
                assert f"{v:.4e}" == f"{r:.4e}", (chi_st, it, v, r)
            else:
                assert v < 1e-8 and r / 3 <= v <= 3 * r, (chi_st, it, v, r)
        phi, T, ok = ref.solve(10, 1e-8, 20, 1e-9, 1e-5, 10, False, 10, chi,
                               vd, T, ONE_ATM, h_ox, h_fu, phi)
        assert ok
        Y = st.mass_fractions.cpu().numpy()
        assert np.max(np.abs(Y - phi[1:])) <= 1e-10
        assert np.max(np.abs(Td.cpu().numpy() - T)) <= 1e-7


@pytest.mark.gpu
def test_gpu_failed_flamelet_stops_at_once(gas, device):
    """One flamelet of a batch starts with a NaN enthalpy at one point: that
    point's temperature Newton does not converge, its source is NaN, the
    flamelet's update is not finite and it ends failed after its first
    iteration (not after every attempt), with the reason given; the other
    flamelets of the batch give their results alone bit for bit.  The
    plain solver stops on the same flamelet at once too."""
    import copy
    from pyrometheus_amd.flamelets.solver import FlameletSolver, solve_batch
    probs = _problems(4, seed=4)
    bad = copy.copy(probs[1])
    bad.phi = probs[1].phi.copy()
    bad.phi[0, len(bad.x) // 2] = np.nan
    mix = solve_batch(gas, [probs[0], bad, probs[2], probs[3]],
                      equilibrate_steps=24)
    r = mix[1]
    assert not r.converged and r.failure and "not finite" in r.failure
    assert r.newton_iterations == 0 and r.time_steps == 0
    for j in (0, 2, 3):
        alone = solve_batch(gas, [probs[j]], equilibrate_steps=24)[0]
        assert mix[j].converged and mix[j].failure is None
        assert _same(alone.state.mass_fractions, mix[j].state.mass_fractions)
        assert _same(alone.temperature, mix[j].temperature)
    solver = FlameletSolver(_domain(bad.x), gas, (bad.y_ox, bad.y_fu))
    st, _ = solver.solve(False, 10, 1e-8, 20, 1e-9, 1e-5, 10, False, 10,
                         _dev(bad.chi, device),
                         _dev(np.zeros(len(bad.x)), device),
                         _dev(bad.t_guess, device), ONE_ATM, bad.h_ox,
                         bad.h_fu, _state(bad.phi, device))
    assert not solver.converged and solver.failure == "non-finite update"
    assert len(solver.log) <= 1


@pytest.mark.gpu
@pytest.mark.parametrize("norm,sync", [("reference", 1), ("reference", 8),
                                       ("scaled", 5)])
def test_gpu_device_control_matches_host(gas, device, norm, sync):
    """The control flow on the device (pyro_flamelet_control after every
    update, the host reading it back once per ``sync`` iterations) gives
    the host control flow's results bit for bit: states, temperatures,
    convergence, Newton and time-step counts, attempts; a flamelet alone
    equals itself in the batch."""
    from pyrometheus_amd.flamelets.stepper import (FlameletStepper,
                                                   SolveSettings)
    probs = _problems(7, seed=5)
    out = {}
    for dc in (False, True):
        st = FlameletStepper(gas, SolveSettings(equilibrate_steps=24,
                                                newton_norm=norm),
                             graphs=True, device_control=dc,
                             sync_every=sync)
        handles = st.add(probs)
        res = st.run()
        out[dc] = [res[h] for h in handles]
    for j, (a, b) in enumerate(zip(out[False], out[True])):
        assert (a.converged, a.newton_iterations, a.time_steps,
                a.attempts) == (b.converged, b.newton_iterations,
                                b.time_steps, b.attempts), j
        for nm, u, v in (("h", a.state.enthalpy, b.state.enthalpy),
                         ("Y", a.state.mass_fractions,
                          b.state.mass_fractions),
                         ("T", a.temperature, b.temperature)):
            assert _same(u, v), (j, nm)
    st = FlameletStepper(gas, SolveSettings(equilibrate_steps=24,
                                            newton_norm=norm),
                         graphs=True, device_control=True, sync_every=sync)
    (h,) = st.add([probs[3]])
    alone = st.run()[h]
    assert _same(alone.state.mass_fractions, out[True][3].state.mass_fractions)
