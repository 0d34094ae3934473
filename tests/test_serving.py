"""Request server: FIFO baseline arithmetic (CPU), and on the device that
every result is bit-identical alone and in the mix, that admission keeps
the held bytes within the HBM budget, and that all requests complete."""
import numpy as np
import pytest
import torch

from pyrometheus_amd.serving import fifo_latencies


def test_fifo_latencies():
    lat = fifo_latencies([0.0, 0.5, 0.6, 5.0], [1.0, 1.0, 0.2, 0.1])
    # ends: 1.0, 2.0, 2.2, 5.1
    assert np.allclose(lat, [1.0, 1.5, 1.6, 0.1])
    assert fifo_latencies([], []).size == 0


def _same(a, b):
    return bool(((a == b) | (torch.isnan(a) & torch.isnan(b))).all())


@pytest.fixture(scope="module")
def gases(device):
    from pyrometheus_amd import bundled, get_thermochem_class
    return {n: get_thermochem_class(bundled(n))(device=device)
            for n in ("sandiego", "gri30")}


@pytest.mark.gpu
def test_alone_equals_in_mix_and_budget(gases, device):
    from pyrometheus_amd.serving import (ChemistryServer, FlameletRequest,
                                         poisson_workload)
    from pyrometheus_amd.flamelets.solver import solve_batch
    arrivals = poisson_workload(gases, 14, rate=40.0, seed=9,
                                wdot_log2_cells=(10, 14),
                                flamelet_points=(40, 160),
                                flamelet_mechanisms=["sandiego"])
    kinds = [type(r).__name__ for _, r in arrivals]
    print("[serve] mix", kinds, flush=True)
    assert "WdotRequest" in kinds and "FlameletRequest" in kinds
    probe = ChemistryServer(gases)
    costs = [probe.cost(r) for _, r in arrivals]
    budget = 3 * max(costs)  # forces requests to wait for budget
    srv = ChemistryServer(gases, hbm_budget=budget, max_flamelets=3)
    mix = srv.run(arrivals)
    assert all(t.t_done is not None for t in mix)
    assert srv.peak_used <= budget and srv.used == 0
    assert any(t.t_admit > t.t_arrival + 1e-3 for t in mix)  # some waited
    for (ta, req), t in zip(arrivals, mix):
        alone = ChemistryServer(gases).run([(0.0, req)])[0]
        if t.kind == "wdot":
            assert torch.equal(alone.result, t.result)
            ref = gases[req.mechanism].get_net_production_rates(
                req.rho, req.temperature, req.mass_fractions)
            assert torch.equal(ref, t.result)
        else:
            a, b = alone.result, t.result
            assert a.converged and b.converged
            assert _same(a.state.enthalpy, b.state.enthalpy)
            assert _same(a.state.mass_fractions, b.state.mass_fractions)
            assert _same(a.temperature, b.temperature)
            c = solve_batch(gases[req.mechanism], [req.problem],
                            equilibrate_steps=24, newton_norm="scaled")[0]
            assert _same(c.state.mass_fractions, b.state.mass_fractions)


@pytest.mark.gpu
def test_alone_equals_in_mix_whole_mix(gases, device):
    """The whole serving mix: production-rate requests on both mechanisms
    and flamelets of 64..2040 points on both the 9- and the 53-species
    mechanism (one near the 2048-point top of the range), arriving while
    others run under a budget that makes some wait.  Every result equals
    the same request served alone and, for flamelets, solve_batch alone,
    bit for bit."""
    from pyrometheus_amd.serving import (ChemistryServer, FlameletRequest,
                                         poisson_workload)
    from pyrometheus_amd.flamelets.solver import solve_batch
    from pyrometheus_amd.flamelets.workload import make_problem, streams
    wd = [a for m, seed in (("gri30", 21), ("sandiego", 22))
          for a in poisson_workload(gases, 2, rate=100.0, seed=seed,
                                    wdot_log2_cells=(12, 18),
                                    wdot_fraction=1.0, wdot_mechanisms=[m])]
    flam = []
    for name, n, chi in (("gri30", 2040, 40.0), ("sandiego", 1500, 300.0),
                         ("gri30", 64, 900.0), ("sandiego", 64, 5.0),
                         ("gri30", 700, 150.0)):
        dom, prob = make_problem(gases[name], streams(gases[name]), n, chi)
        flam.append(FlameletRequest(name, prob, dom))
    reqs = [wd[0][1], flam[0], flam[1], wd[1][1], flam[2], flam[3],
            wd[2][1], flam[4], wd[3][1]]
    arrivals = [(0.02 * k, r) for k, r in enumerate(reqs)]
    mechs = {(type(r).__name__, r.mechanism) for r in reqs}
    assert len(mechs) == 4, mechs
    probe = ChemistryServer(gases)
    budget = 2 * max(probe.cost(r) for r in reqs)
    srv = ChemistryServer(gases, hbm_budget=budget)
    mix = srv.run(arrivals)
    print("[serve whole mix] latencies",
          [round(t.latency, 3) for t in mix], flush=True)
    assert srv.peak_used <= budget and srv.used == 0
    assert any(t.t_admit > t.t_arrival + 1e-3 for t in mix)
    for (ta, req), t in zip(arrivals, mix):
        alone = ChemistryServer(gases).run([(0.0, req)])[0]
        if t.kind == "wdot":
            assert torch.equal(alone.result, t.result)
            continue
        a, b = alone.result, t.result
        assert a.converged and b.converged, (req.mechanism, req.size)
        assert _same(a.state.enthalpy, b.state.enthalpy)
        assert _same(a.state.mass_fractions, b.state.mass_fractions)
        assert _same(a.temperature, b.temperature)
        c = solve_batch(gases[req.mechanism], [req.problem],
                        equilibrate_steps=24, newton_norm="scaled")[0]
        assert _same(c.state.mass_fractions, b.state.mass_fractions)
