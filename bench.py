"""Benchmark: net production rates of a 53-species FP64 mechanism, and
batched versus plain steady flamelet solves.

One step = one ``get_net_production_rates`` evaluation (one fused HIP kernel)
over 2^24 cells of GRI-3.0 (53 species, 325 reactions) with the state already
resident in HBM: T ~ U[300, 2500] K, p log-uniform in [0.5, 20] atm,
Y = 0.5 Dirichlet(0.5) + 0.5 e_N2 (synthetic, seeded).

Prints one JSON line with the throughput in Mcells/s, the FP64 roofline of the
kernel (algorithmic FLOPs per cell from ``pyrometheus_amd.flops`` divided by
the kernel's average launch time, measured with HIP events on the stream it
runs on), and the CPU baseline: the numpy oracle (the reference's op
structure) timed on a bounded sample on this host.

The ``pointwise`` object times the memory-bound methods (species h/RT,
mixture c_p, T from e, k_f) on GRI-3.0 against HBM bandwidth.

The ``derivatives`` object times the hand-written VJP and JVP of the net
production rates and the full (ns+1)^2 flamelet-source Jacobian on 2^22
GRI-3.0 cells (seed 3, cotangent seed 33), with a central-difference check
of the VJP at h = 1e-6 T.

The ``transport`` object times the mixture-averaged transport properties
(Wilke viscosity, conductivity, species diffusivities) on 2^22 GRI-3.0 cells.

The ``flamelets`` object of the same line solves 256 steady flamelets of the
9-species H2/air mechanism (grid sizes uniform in [64, 1024] clustered about
Z_st, chi_st log-uniform in [1, 1000] 1/s, fuel H2:N2 at 300 K, air at 800 K,
1 atm) with ``solve_batch`` (one launch sequence per iteration for every
unfinished flamelet) and times the plain drop-in (one ``FlameletSolver`` per
flamelet) on a sample of the same problems; both reach Newton |v| < 1e-8
under the scaled update norm.  Latency of a batched flamelet is the time
from submission of the batch to that flamelet's convergence.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells N]
                       [--mech gri30] [--cpu-seconds S] [--no-cpu]
                       [--flamelets B] [--plain-sample K] [--no-flamelets]
"""

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0
METRIC = ("Mcells/s \u03c9\u0307 eval (53-sp FP64); flamelet solves/s + p99 "
          "latency vs plain")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=1 << 24)
    ap.add_argument("--mech", default="gri30")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES "
                         "(3 launches); calibrates HBM byte counters")
    ap.add_argument("--flamelets", type=int, default=256)
    ap.add_argument("--plain-sample", type=int, default=8)
    ap.add_argument("--newton-norm", default="reference",
                    choices=("reference", "scaled"))
    ap.add_argument("--cpu-flamelets", type=int, default=2,
                    help="flamelets of the CPU (numpy) flamelet baseline")
    ap.add_argument("--flamelet-reps", type=int, default=3)
    ap.add_argument("--flamelet-seed", type=int, default=4)
    ap.add_argument("--no-flamelets", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=96)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-flamelet-mechs", default="sandiego,gri30")
    ap.add_argument("--deriv-cells", type=int, default=1 << 22)
    ap.add_argument("--no-derivatives", action="store_true")
    ap.add_argument("--no-pointwise", action="store_true")
    ap.add_argument("--transport-cells", type=int, default=1 << 22)
    ap.add_argument("--no-transport", action="store_true")
    ap.add_argument("--no-wdot", action="store_true",
                    help="skip the omega-dot leg (profiling the flamelets)")
    ap.add_argument("--traffic-json",
                    default=os.path.join(ROOT, "profiles", "wdot_gri30_traffic.json"))
    return ap.parse_args()


def device_inputs(mech, n, seed, device):
    import torch
    from pyrometheus_amd.mechanism import ONE_ATM
    g = torch.Generator(device=device).manual_seed(seed)
    T = 300.0 + 2200.0 * torch.rand(n, generator=g, dtype=torch.float64,
                                    device=device)
    lnp = torch.log(torch.tensor(0.5)) + torch.log(torch.tensor(40.0)) * \
        torch.rand(n, generator=g, dtype=torch.float64, device=device)
    p = ONE_ATM * torch.exp(lnp)
    ns = mech.n_species
    # Dirichlet(0.5) by normalised Gamma(0.5) draws, species-major (ns, n)
    alpha = torch.full((ns, n), 0.5, dtype=torch.float64, device=device)
    Y = torch._standard_gamma(alpha, generator=g)
    Y /= Y.sum(0, keepdim=True)
    Y *= 0.5
    Y[mech.species_index("N2")] += 0.5
    del alpha
    return T, p, Y


def cpu_model():
    """The host CPU's model name from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def host_cores():
    """CPUs this process may run on (the box's share, not the machine's)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _cpu_wdot_chunks(task):
    """Worker: the numpy oracle's net production rates on ``nchunks`` chunks
    of 2^16 cells drawn from ``seed``; returns (cells, seconds in ω̇)."""
    import numpy as np
    from oracle.thermochem_np import Thermochemistry
    from pyrometheus_amd import bundled
    from pyrometheus_amd.mechanism import ONE_ATM
    name, seed, nchunks, chunk = task
    mech = bundled(name)
    gas = Thermochemistry(mech)
    rng = np.random.default_rng(seed)
    done, elapsed = 0, 0.0
    for _ in range(nchunks):
        T = rng.uniform(300.0, 2500.0, chunk)
        p = ONE_ATM * np.exp(rng.uniform(np.log(0.5), np.log(20.0), chunk))
        Y = 0.5 * rng.dirichlet(0.5 * np.ones(mech.n_species), chunk).T
        Y[mech.species_index("N2")] += 0.5
        rho = gas.get_density(p, T, Y)
        t1 = time.perf_counter()
        gas.get_net_production_rates(rho, T, Y)
        elapsed += time.perf_counter() - t1
        done += chunk
    return done, elapsed


def cpu_baseline(mech, seconds, seed, flamelet_probs=None):
    """The numpy oracle (the reference's op structure: one whole-array ufunc
    per generated term) on this host: 1 process on chunks of 2^16 cells
    until ~``seconds`` of work, then a process pool over every available
    core on as many chunks per process; optionally the oracle's flamelet
    solve on a sample of the benchmark's flamelets (1 core, and a pool)."""
    import multiprocessing as mp
    chunk = 1 << 16
    cores = host_cores()
    cells, el = 0, 0.0
    nch = 0
    while el < seconds:
        c, e = _cpu_wdot_chunks((mech.name, seed + nch, 1, chunk))
        cells, el, nch = cells + c, el + e, nch + 1
    one = cells / el / 1e6
    per = max(1, nch // 2)
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_wdot_chunks, [(mech.name, 1000 + seed + j, per,
                                           chunk) for j in range(cores)])
    wall = time.perf_counter() - t0
    pool_cells = sum(r[0] for r in res)
    out = {"value": one, "unit": "Mcells/s", "cores": 1, "kind": "port",
           "sample": f"{cells} cells of {mech.name} in chunks of {chunk} "
                     "(seeded like the workload), numpy restatement of the "
                     "generated code (whole-array ufunc per term), 1 "
                     f"process; CPU {cpu_model()}; JAX absent",
           "all_cores": {"value": pool_cells / wall / 1e6,
                         "unit": "Mcells/s", "cores": cores,
                         "sample": f"{pool_cells} cells, {per} chunks per "
                                   f"process, {cores} processes (wall time "
                                   "including process start)"},
           "cpu_model": cpu_model()}
    if flamelet_probs:
        out["flamelets"] = cpu_flamelet_baseline(flamelet_probs, cores)
    return out


def _cpu_source_jacobian(task):
    """Worker: the numpy oracle's (ns+1)^2 source Jacobian (complex step
    through T(h, Y), one perturbed evaluation per column) of ``n`` cells of
    ``name`` drawn like the workload; returns (cells, seconds)."""
    import numpy as np
    from oracle.flamelet_np import FlameletEquations
    from oracle.thermochem_np import Thermochemistry
    from pyrometheus_amd import bundled
    from pyrometheus_amd.mechanism import ONE_ATM
    name, seed, n = task
    mech = bundled(name)
    gas = Thermochemistry(mech)
    rng = np.random.default_rng(seed)
    T = rng.uniform(300.0, 2500.0, n)
    p = ONE_ATM * np.exp(rng.uniform(np.log(0.5), np.log(20.0), n))
    Y = 0.5 * rng.dirichlet(0.5 * np.ones(mech.n_species), n).T
    Y[mech.species_index("N2")] += 0.5
    phi = np.vstack([gas.get_mixture_enthalpy_mass(T, Y)[None], Y])
    eq = FlameletEquations(mech, np.linspace(0.0, 1.0, n), Y[:, 0], Y[:, -1])
    t0 = time.perf_counter()
    eq.source_jacobian(phi, np.zeros(n), T, p)
    return n, time.perf_counter() - t0


def cpu_small_baselines(cores):
    """The numpy oracle on config 1 (synth9 ω̇, 1 process and a pool over
    ``cores``) and config 3 (the GRI-3.0 source Jacobian, 1 process)."""
    import multiprocessing as mp
    c1, e1 = _cpu_wdot_chunks(("synth9", 1, 4, 1 << 16))
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_wdot_chunks, [("synth9", 2000 + j, 4, 1 << 16)
                                          for j in range(cores)])
    wall = time.perf_counter() - t0
    pool_cells = sum(r[0] for r in res)
    nj, ej = _cpu_source_jacobian(("gri30", 3, 512))
    return {
        "config1": {
            "value": c1 / e1 / 1e6, "unit": "Mcells/s", "cores": 1,
            "kind": "port",
            "sample": "4 chunks of 65,536 synth9 cells, numpy oracle",
            "all_cores": {"value": pool_cells / wall / 1e6,
                          "unit": "Mcells/s", "cores": cores,
                          "sample": f"{pool_cells} cells, 4 chunks per "
                                    f"process, {cores} processes (wall time "
                                    "including process start)"}},
        "source_jacobian": {
            "value": nj / ej / 1e6, "unit": "Mcells/s", "cores": 1,
            "kind": "port",
            "sample": f"{nj} gri30 cells drawn like the workload, the "
                      "oracle's complex-step source Jacobian (one complex "
                      "evaluation of the source terms through T(h, Y) per "
                      "column), 1 process; the reference differentiates "
                      "with JAX, which is absent"}}


def _cpu_flamelet(prob):
    """Worker: the numpy oracle's solve of one benchmark flamelet (the same
    equilibration phase and hybrid solve as the device); ``prob`` is the
    tuple (x, chi, phi, t_guess, y_ox, y_fu, h_ox, h_fu, pressure); returns
    (seconds, converged, N)."""
    import numpy as np
    from oracle import flamelet_np as F
    from pyrometheus_amd.mechanism import bundled
    x, chi, phi, tg, y_ox, y_fu, h_ox, h_fu, p = prob
    mech = bundled("sandiego")
    t0 = time.perf_counter()
    s = F.FlameletSolver(mech, x, y_ox, y_fu)
    phi, T = s.equilibrate(phi, tg, p, h_ox, h_fu)
    _, _, ok = s.solve(10, 1e-8, 20, 1e-9, 1e-5, 10, False, 10, chi,
                       np.zeros(len(x)), T, p, h_ox, h_fu, phi)
    return time.perf_counter() - t0, bool(ok), len(x)


def cpu_flamelet_baseline(probs, cores):
    """Flamelet solves/s of the numpy oracle: the given problems one after
    another in one process, and ``cores`` copies of them spread over a
    process pool."""
    import multiprocessing as mp
    import numpy as np
    probs = [(p.x, p.chi, p.phi, p.t_guess, p.y_ox, p.y_fu, p.h_ox, p.h_fu,
              p.pressure) for p in probs]
    one = [_cpu_flamelet(p) for p in probs]
    t_one = sum(r[0] for r in one)
    ctx = mp.get_context("spawn")
    batch = [probs[j % len(probs)] for j in range(cores)]
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_flamelet, batch)
    wall = time.perf_counter() - t0
    return {"value": len(one) / t_one, "unit": "solves/s", "cores": 1,
            "kind": "port",
            "mean_n": float(np.mean([r[2] for r in one])),
            "sample": f"{len(one)} of the benchmark's flamelets (N = "
                      f"{[r[2] for r in one]}), size-stratified, numpy "
                      "restatement of the flamelet solver (complex-step "
                      "source Jacobian, LU block Thomas), equilibration "
                      "and hybrid solve as on the device",
            "converged": f"{sum(r[1] for r in one)}/{len(one)}",
            "all_cores": {"value": len(batch) / wall, "unit": "solves/s",
                          "cores": cores,
                          "sample": f"{len(batch)} solves over {cores} "
                                    "processes (wall time including process "
                                    "start)"}}


def pointwise_leg(args, device):
    """The HBM-bound pointwise methods on GRI-3.0 over 2^24 cells (2^22 for
    the per-reaction k_f): species h/RT, mixture c_p, T from the internal
    energy (the reference's energy form with T_guess = T + U[-300, 300] K)
    and k_f, each timed with HIP events on the launch stream, with the
    compulsory bytes per cell and the resulting fraction of HBM bandwidth."""
    import torch
    from pyrometheus_amd import bundled, get_thermochem_class
    mech = bundled("gri30")
    gas = get_thermochem_class(mech)(device=device)
    n = args.cells
    ns, nr = mech.n_species, mech.n_reactions
    T, p, Y = device_inputs(mech, n, args.seed, device)
    g = torch.Generator(device=device).manual_seed(args.seed + 1)
    tg = (T + 600.0 * (torch.rand(n, generator=g, dtype=torch.float64,
                                  device=device) - 0.5)).clamp(200.0, 3500.0)
    e = gas.get_mixture_internal_energy_mass(T, Y)
    del p

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        ms = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            torch.cuda.synchronize()
            ms.append(a.elapsed_time(b))
        return sum(ms) / len(ms)

    def entry(ms, cells, bytes_cell):
        gbs = bytes_cell * cells / (ms * 1e-3) / 1e9
        return {"ms": ms, "cells": cells, "mcells_s": cells / ms / 1e3,
                "bytes_per_cell": bytes_cell, "hbm_gbs": gbs,
                "hbm_frac": gbs / HBM_PEAK_GBS}
    print("[bench] pointwise", file=sys.stderr, flush=True)
    res = {"workload": f"gri30, {n} cells (seed {args.seed})"}
    res["species_enthalpies_rt"] = entry(
        timed(lambda: gas.get_species_enthalpies_rt(T)), n, 8 * (ns + 1))
    res["mixture_cp_mass"] = entry(
        timed(lambda: gas.get_mixture_specific_heat_cp_mass(T, Y)), n,
        8 * (ns + 2))
    ms_t = timed(lambda: gas.get_temperature(e, tg, Y, True))
    t_out = gas.get_temperature(e, tg, Y, True)
    # GRI-3.0's e(T) is not continuous at T_mid: a target from within a
    # band around T_mid may have a second root, which Newton may reach
    t_mid, _, width = gas.get_temperature_jump_bands(Y)
    near = ((T[None, :] - t_mid[:, None]).abs() <= 2.0 * width + 1e-12).any(0)
    err = (t_out - T).abs()
    resid = ((gas.get_mixture_internal_energy_mass(t_out, Y) - e).abs()
             / gas.get_mixture_specific_heat_cv_mass(t_out, Y))
    res["temperature_from_energy"] = dict(
        entry(ms_t, n, 8 * (ns + 3)),
        max_abs_err_K_outside_bands=float(err[~near].max().item()),
        cells_in_bands=int(near.sum().item()),
        max_abs_err_K_in_bands=(float(err[near].max().item())
                                if bool(near.any()) else 0.0),
        max_root_residual_K=float(resid.max().item()),
        max_band_half_width_K=float(width.max().item()),
        bands="cells within twice |e(T_mid+) - e(T_mid)|/c_v of a T_mid "
              "(GRI-3.0: e drops by 2e-5..5e-4 K c_v at 1000 K, two roots)")
    del e, tg, t_out, near, err, resid, width
    torch.cuda.empty_cache()
    m = n // 4
    Tm, Ym = T[:m].contiguous(), Y[:, :m].contiguous()
    rho = gas.get_density(torch.full((m,), 101325.0, dtype=torch.float64,
                                     device=device), Tm, Ym)
    C = gas.get_concentrations(rho, Ym)
    res["fwd_rate_coefficients"] = entry(
        timed(lambda: gas.get_fwd_rate_coefficients(Tm, C)), m,
        8 * (ns + 1 + nr))
    del C, Tm, Ym, rho, T, Y
    torch.cuda.empty_cache()
    return res


def derivatives_leg(args, device):
    """Config-3 derivatives on GRI-3.0: the VJP of the net production rates
    with respect to (rho, T, Y) for a cotangent w ~ N(0, 1), the JVP, and the
    full (ns+1)^2 Jacobian of the flamelet source S = [vd/rho; W_k w_k/rho]
    with respect to (h, Y) (vd = 0, entry-major layout), each timed with HIP
    events on the launch stream; plus a central-difference check of the
    VJP's temperature cotangent (h = 1e-6 T) on a sample of cells."""
    import torch
    from pyrometheus_amd import bundled, get_thermochem_class
    from pyrometheus_amd.adjoint import (wdot_jvp, wdot_vjp, source_jacobian,
                                         source_jacobian_workspace,
                                         set_source_jacobian_form)
    mech = bundled("gri30")
    gas = get_thermochem_class(mech)(device=device)
    n = args.deriv_cells
    ns, nv = mech.n_species, mech.n_species + 1
    T, p, Y = device_inputs(mech, n, 3, device)
    rho = gas.get_density(p, T, Y)
    g = torch.Generator(device=device).manual_seed(33)
    wbar = torch.randn((ns, n), generator=g, dtype=torch.float64, device=device)
    vT = torch.randn(n, generator=g, dtype=torch.float64, device=device)
    vY = 1e-3 * torch.randn((ns, n), generator=g, dtype=torch.float64,
                            device=device)
    out = torch.empty((ns, n), dtype=torch.float64, device=device)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        ms = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            torch.cuda.synchronize()
            ms.append(a.elapsed_time(b))
        return sum(ms) / len(ms)

    res = {"workload": f"gri30, {n} cells (seed 3), cotangent N(0,1) "
                       "(seed 33), vd = 0"}
    print("[bench] derivatives", file=sys.stderr, flush=True)
    ms_w = timed(lambda: gas.get_net_production_rates(rho, T, Y, out=out), 5)
    ms_v = timed(lambda: wdot_vjp(gas, rho, T, Y, wbar), 5)
    ms_j = timed(lambda: wdot_jvp(gas, rho, T, Y, None, vT, vY), 5)
    res.update({"wdot_ms": ms_w, "vjp_ms": ms_v, "jvp_ms": ms_j,
                "vjp_over_wdot": ms_v / ms_w, "jvp_over_wdot": ms_j / ms_w,
                "vjp_mcells_s": n / ms_v / 1e3})
    # central difference of <wbar, w(T)> in T against the VJP's g_T
    sel = torch.arange(0, n, max(1, n // 4096), device=device)[:4096]
    rs, ts, ys, ws = rho[sel], T[sel], Y[:, sel].contiguous(), wbar[:, sel]
    _, g_t, _ = wdot_vjp(gas, rs, ts, ys, ws.contiguous())
    h = 1e-6 * ts
    fp = (ws * gas.get_net_production_rates(rs, ts + h, ys)).sum(0)
    fm = (ws * gas.get_net_production_rates(rs, ts - h, ys)).sum(0)
    fd = (fp - fm) / (2 * h)
    scale = (ws.abs() * gas.get_net_production_rates(rs, ts, ys).abs()
             ).sum(0) / ts
    res["fd_check"] = {"quantity": "g_T vs central difference, h = 1e-6 T, "
                                   "4096 cells",
                       "max_err_over_scale": float(
                           ((g_t - fd).abs() / scale).max().item())}
    del wbar, vT, vY, out
    torch.cuda.empty_cache()
    jac = torch.empty((nv, nv, n), dtype=torch.float64, device=device)
    # the two-pass form (per-reaction cache) first, for comparison
    set_source_jacobian_form(gas, 0)
    try:
        work = source_jacobian_workspace(gas)
        ms_J0 = timed(lambda: source_jacobian(gas, T, p, Y, None,
                                              layout="entry", out=jac,
                                              work=work), 3)
    finally:
        set_source_jacobian_form(gas, 1)
    del work
    work = source_jacobian_workspace(gas)
    ms_J = timed(lambda: source_jacobian(gas, T, p, Y, None, layout="entry",
                                         out=jac, work=work), 3)
    # the flamelet solver's per-point block layout, in the same memory
    ms_Jc = timed(lambda: source_jacobian(gas, T, p, Y, None, layout="cell",
                                          out=jac.view(n, nv, nv),
                                          work=work), 3)
    bytes_cell = 8 * (nv * nv + ns + 2)
    res["source_jacobian"] = {
        "ms": ms_J, "mcells_s": n / ms_J / 1e3, "layout": "entry (nv, nv, N)",
        "form": "one pass (k_source_jacobian_fused: each row evaluates its "
                "reactions; no workspace)",
        "two_pass_ms": ms_J0,
        "cell_layout_ms": ms_Jc,
        "cell_layout_hbm_frac": bytes_cell * n / (ms_Jc * 1e-3) / 1e9
                                / HBM_PEAK_GBS,
        "output_gb": 8 * nv * nv * n / 1e9,
        "roofline": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                     "bytes_per_cell": bytes_cell,
                     "achieved": bytes_cell * n / (ms_J * 1e-3) / 1e9,
                     "frac": bytes_cell * n / (ms_J * 1e-3) / 1e9
                     / HBM_PEAK_GBS}}
    assert torch.isfinite(jac[:, :, :4096]).all().item()
    del jac, work, T, p, Y, rho
    torch.cuda.empty_cache()
    return res


def flamelet_leg(args, device):
    """Batched (``solve_batch``) against plain (``FlameletSolver``) solves of
    the same flamelets; kernel times of the batched run from HIP events."""
    import numpy as np
    import torch
    from pyrometheus_amd import bundled, get_thermochem_class
    from pyrometheus_amd.flamelets.solver import solve_batch
    from pyrometheus_amd.flamelets.workload import (flamelet_batch,
                                                    solve_plain)
    mech = bundled("sandiego")
    gas = get_thermochem_class(mech)(device=device)
    doms, probs, chis = flamelet_batch(gas, n_flamelets=args.flamelets,
                                       seed=args.flamelet_seed)
    kw = dict(equilibrate_steps=24, newton_norm=args.newton_norm)
    sizes = np.array([len(p.x) for p in probs])
    # block-Thomas schedule for nv = 10 on this device, tuned on this
    # workload's flamelet sizes (synthetic blocks)
    from pyrometheus_amd.flamelets.autotune import tune_thomas
    tuned = tune_thomas(gas, sizes=sizes, seed=args.flamelet_seed)
    print(f"[bench] block-Thomas schedule: {tuned['kind']} "
          f"alpha={tuned['alpha']:.3g}", file=sys.stderr, flush=True)
    print(f"[bench] flamelets: {len(probs)} batched, points {sizes.sum()}",
          file=sys.stderr, flush=True)
    solve_batch(gas, probs[:2], **kw)  # warm-up
    torch.cuda.synchronize()
    def batched(events=None):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = solve_batch(gas, probs, kernel_events=events, **kw)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, res

    runs = []
    for rep in range(args.flamelet_reps):
        wall, res = batched()
        runs.append((wall, res))
        print(f"[bench] batched rep {rep}: {wall:.3f} s, converged "
              f"{sum(r.converged for r in res)}", file=sys.stderr, flush=True)
    walls = [r[0] for r in runs]
    wall, res = runs[int(np.argsort(walls)[len(walls) // 2])]
    # one more run with HIP events around every kernel for the breakdown
    events, stats = {}, {}
    torch.cuda.synchronize()
    solve_batch(gas, probs, kernel_events=events, stats=stats, **kw)
    torch.cuda.synchronize()
    kms = {k: sum(a.elapsed_time(e) for a, e in v) for k, v in events.items()}
    tkey = ("pyro_flamelet_thomas_partitioned"
            if "pyro_flamelet_thomas_partitioned" in events
            else "pyro_flamelet_thomas")
    iters = len(events[tkey])
    lat_b = np.array([r.finished_at for r in res])
    resid = max_scaled_residual(gas, probs, res)
    plain_t, plain_ok = [], 0
    # size-stratified sample: the flamelets at evenly spaced ranks of N
    order = np.argsort(sizes, kind="stable")
    k = min(args.plain_sample, len(probs))
    sample = order[np.round(np.linspace(0, len(probs) - 1, k)).astype(int)]
    for j in sample:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, _, ok = solve_plain(gas, doms[j], probs[j], semantics="exact",
                               **kw)
        torch.cuda.synchronize()
        plain_t.append(time.perf_counter() - t0)
        plain_ok += bool(ok)
        print(f"[bench] plain {j}: N={len(probs[j].x)} {plain_t[-1]:.3f} s "
              f"ok={ok}", file=sys.stderr, flush=True)
    plain_t = np.array(plain_t)
    batched_sps = len(probs) / wall
    plain_sps = len(plain_t) / plain_t.sum()
    nv = mech.n_species + 1
    rows = stats["point_iterations"]  # block rows through block Thomas
    thomas_ms = kms[tkey]
    # block Thomas on (nv x nv) central blocks with scalar coupling: per
    # block row it reads C (nv^2), rhs (nv), L and U, writes x (nv)
    bytes_row = 8 * (nv * nv + 2 * nv + 2)
    flop_row = 4.67 * nv ** 3 + 8 * nv ** 2
    # the CPU flamelet baseline's sample (bounded CPU time): flamelets at
    # the 5 % .. 25 % ranks of N; its result is also scaled to the batch's
    # mean N (CPU time per solve grows about linearly with N)
    kc = min(args.cpu_flamelets, len(probs))
    cpu_sample = [probs[j] for j in order[np.round(
        np.linspace(0.05, 0.25, kc) * (len(probs) - 1)).astype(int)]]
    args.batch_mean_n = float(sizes.mean())
    return cpu_sample, {
        "workload": f"{len(probs)} steady flamelets, {mech.name} "
                    f"({mech.n_species} species), N in [64, 1024] "
                    "Z-clustered, chi_st log-uniform [1, 1000] 1/s, "
                    "T_ox 800 K, T_fu 300 K, 1 atm, equilibrium start",
        "batched_launch": "each iteration replayed from a captured HIP "
                          "graph (re-captured when the batch is re-packed)",
        "thomas_schedule": tuned,
        "batched_solves_per_s": batched_sps,
        "plain_solves_per_s": plain_sps,
        "speedup_vs_plain": batched_sps / plain_sps,
        "batched_wall_s": wall,
        "batched_wall_s_reps": walls,
        "batched_latency_p50_s": float(np.percentile(lat_b, 50)),
        "batched_latency_p99_s": float(np.percentile(lat_b, 99)),
        "plain_latency_p50_s": float(np.percentile(plain_t, 50)),
        "plain_latency_p99_s": float(np.percentile(plain_t, 99)),
        "plain_sample": f"{len(plain_t)} of the same problems at evenly "
                        "spaced ranks of N (mean N "
                        f"{sizes[sample].mean():.0f}, batch mean "
                        f"{sizes.mean():.0f}), one FlameletSolver each "
                        "(same iteration as the batch: exact source "
                        "Jacobian), sequential",
        "residual": resid,
        "converged_batched": f"{sum(r.converged for r in res)}/{len(res)}",
        "converged_plain": f"{plain_ok}/{len(plain_t)}",
        "newton_norm": ("reference (|v|_2 over every row, "
                        "solver.py:136-139) < 1e-8"
                        if args.newton_norm == "reference" else
                        "scaled (enthalpy row / max(|h_ox|, |h_fu|)) < 1e-8"),
        "batch_iterations": iters,
        "kernel_ms": kms,
        "thomas_roofline": {
            "bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
            "achieved": bytes_row * rows / (thomas_ms * 1e-3) / 1e9,
            "frac": bytes_row * rows / (thomas_ms * 1e-3) / 1e9
            / HBM_PEAK_GBS,
            "bytes_per_block_row": bytes_row,
            "flop_per_block_row": flop_row,
            "fp64_tflops": flop_row * rows / (thomas_ms * 1e-3) / 1e12,
            "kernel": tkey,
            "note": "partitions of ~sqrt(2N) rows swept from both ends by "
                    "two wavefronts each, interface system, recovery: "
                    "latency-bound chains"},
    }


def max_scaled_residual(gas, probs, results):
    """Converged residual of every flamelet against its gross scale: for the
    enthalpy row and for the mass-fraction rows together, max |R| over
    max (|chi/2| (|a phi_-| + |b phi| + |c phi_+|) + |S|) of that flamelet,
    R and S evaluated by the device kernels at the returned state and
    temperature; returns the worst over the batch and the count above
    1e-10."""
    import numpy as np
    import torch
    from pyrometheus_amd.flamelets.engine import FlameletBatch
    b = FlameletBatch(gas, [p.x for p in probs],
                      np.stack([p.y_ox for p in probs]),
                      np.stack([p.y_fu for p in probs]))
    b.set_boundary_enthalpy([p.h_ox for p in probs],
                            [p.h_fu for p in probs])
    dev = b.dev
    phi = torch.cat([torch.cat([r.state.enthalpy[None],
                                r.state.mass_fractions]) for r in results],
                    dim=1).contiguous()
    T = torch.cat([r.temperature for r in results]).contiguous()
    chi = torch.as_tensor(np.concatenate([p.chi for p in probs]),
                          dtype=torch.float64, device=dev)
    pres = b.per_point([p.pressure for p in probs])
    vd = torch.zeros(b.P, dtype=torch.float64, device=dev)
    R = torch.empty_like(phi)
    b.residual(phi, chi, vd, T, pres, R)
    S = b.S.clone()
    ap = phi.abs()
    gross = b.lb.abs() * ap
    nxt = torch.zeros_like(ap)
    prv = torch.zeros_like(ap)
    prv[:, 1:] = ap[:, :-1]
    nxt[:, :-1] = ap[:, 1:]
    gross = 0.5 * chi.abs() * (gross + b.la.abs() * prv + b.lc.abs() * nxt)
    gross += S.abs()
    worst = []
    for a, e in zip(b.offsets_np[:-1], b.offsets_np[1:]):
        r, g = R[:, a:e].abs(), gross[:, a:e]
        worst.append(max(float(r[0].max() / g[0].max()),
                         float(r[1:].max() / g[1:].max())))
    worst = np.array(worst)
    return {"max_scaled": float(worst.max()),
            "above_1e-10": int((worst > 1e-10).sum()),
            "flamelets": len(worst),
            "definition": "per flamelet, enthalpy row and mass-fraction rows "
                          "separately: max|R| / max(|chi/2|(|a phi_-| + "
                          "|b phi| + |c phi_+|) + |S|)"}


def transport_leg(args, device):
    """Mixture-averaged transport on GRI-3.0: Wilke's viscosity, the
    conductivity mean and the species diffusivities over 2^22 cells, each
    timed with HIP events on the launch stream.  The double sums over
    species pairs make the first and last compute-bound: the FP64 rate
    counts the operations of the pair loop per cell (7 per pair for Wilke,
    19 per pair for the diffusivities: the ln-T polynomial, the
    reciprocal's estimate refinement and the accumulation).  The CPU line
    times the numpy oracle on 4,096 of the cells, one core."""
    import numpy as np
    import torch
    from oracle.transport_np import Transport
    from pyrometheus_amd import bundled, get_thermochem_class
    mech = bundled("gri30")
    gas = get_thermochem_class(mech)(device=device)
    n, ns = args.transport_cells, mech.n_species
    T, p, Y = device_inputs(mech, n, args.seed + 5, device)
    print("[bench] transport", file=sys.stderr, flush=True)
    stream = torch.cuda.current_stream()

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            torch.cuda.synchronize()
            ms.append(a.elapsed_time(b))
        return sum(ms) / len(ms)

    res = {"workload": f"gri30, {n} cells (seed {args.seed + 5}), "
                       "fits of pyrometheus_amd.transport"}
    legs = (("mixture_viscosity_wilke",
             lambda: gas.get_mixture_viscosity_mixavg(T, Y), 7 * ns * ns,
             8 * (ns + 2)),
            ("mixture_conductivity",
             lambda: gas.get_mixture_thermal_conductivity_mixavg(T, Y),
             None, 8 * (ns + 2)),
            ("species_diffusivities_mixavg",
             lambda: gas.get_species_mass_diffusivities_mixavg(p, T, Y),
             19 * ns * ns, 8 * (2 * ns + 2)))
    for name, fn, flop, nbytes in legs:
        ms = timed(fn)
        e = {"ms": ms, "mcells_s": n / ms / 1e3, "bytes_per_cell": nbytes,
             "hbm_frac": nbytes * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        if flop:
            tf = flop * n / (ms * 1e-3) / 1e12
            e.update(flop_per_cell=flop, fp64_tflops=tf,
                     fp64_frac=tf / FP64_PEAK_TFLOPS)
        res[name] = e
    # CPU: the oracle on a slice, evaluated on the library's own fits
    m = 4096
    Tc, pc, Yc = (T[:m].cpu().numpy(), p[:m].cpu().numpy(),
                  Y[:, :m].cpu().numpy())
    ref = Transport(mech, *gas.get_transport_fits())
    cpu = {}
    for name, fn in (("mixture_viscosity_wilke",
                      lambda: ref.get_mixture_viscosity_mixavg(Tc, Yc)),
                     ("species_diffusivities_mixavg",
                      lambda: ref.get_species_mass_diffusivities_mixavg(
                          pc, Tc, Yc))):
        t0 = time.perf_counter()
        fn()
        cpu[name] = m / (time.perf_counter() - t0) / 1e6
    res["cpu_baseline"] = {"unit": "Mcells/s", "cores": 1, "kind": "port",
                           "sample": f"{m} of the cells, numpy oracle",
                           **cpu}
    del T, p, Y
    torch.cuda.empty_cache()
    return res


def config1_leg(args, device):
    """The CPU-runnable configuration on the device: net production rates
    and T from e (guesses T +- 300 K) of the 9-species synthetic H2-like
    mechanism (21 reactions: 3 Troe, 4 three-body) over 65,536 cells,
    seed 1."""
    import torch
    from pyrometheus_amd import bundled, get_thermochem_class
    from pyrometheus_amd.flops import wdot_flops
    mech = bundled("synth9")
    gas = get_thermochem_class(mech)(device=device)
    n = 1 << 16
    T, p, Y = device_inputs(mech, n, 1, device)
    rho = gas.get_density(p, T, Y)
    g = torch.Generator(device=device).manual_seed(11)
    tg = (T + 600.0 * (torch.rand(n, generator=g, dtype=torch.float64,
                                  device=device) - 0.5)).clamp(200.0, 3500.0)
    e = gas.get_mixture_internal_energy_mass(T, Y)
    out = torch.empty((mech.n_species, n), dtype=torch.float64, device=device)
    stream = torch.cuda.current_stream()

    def timed(fn, reps=20):
        fn()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True),
               torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for a, b in ev:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        return sum(a.elapsed_time(b) for a, b in ev) / reps
    ms_w = timed(lambda: gas.get_net_production_rates(rho, T, Y, out=out))
    ms_t = timed(lambda: gas.get_temperature(e, tg, Y, True))
    t_out = gas.get_temperature(e, tg, Y, True)
    fl = wdot_flops(mech)
    return {"workload": f"synth9 ({mech.n_species} species, "
                        f"{mech.n_reactions} reactions), {n} cells (seed 1)",
            "wdot_ms": ms_w, "wdot_mcells_s": n / ms_w / 1e3,
            "wdot_fp64_frac": fl["flop_per_cell"] * n / (ms_w * 1e-3)
            / 1e12 / FP64_PEAK_TFLOPS,
            "wdot_hbm_frac": fl["bytes_per_cell"] * n / (ms_w * 1e-3) / 1e9
            / HBM_PEAK_GBS,
            "temperature_ms": ms_t,
            "temperature_mcells_s": n / ms_t / 1e3,
            "temperature_max_abs_err_K": float((t_out - T).abs().max()),
            "note": "65,536 cells fill 256 of the chip's 1,024 SIMDs "
                    "with one wave each: launch- and latency-bound"}


def serving_leg(args, device):
    """Poisson arrivals of production-rate requests (2^10..2^20 cells) and
    flamelet solves (64..2048 points), each on the 9- or the 53-species
    mechanism with probability 1/2 (``--serve-flamelet-mechs`` restricts
    the flamelets), served by the continuous-batching server and by the
    plain drop-in one request at a time (FIFO), on the same arrival
    trace."""
    import numpy as np
    import torch
    from pyrometheus_amd import bundled, get_thermochem_class
    from pyrometheus_amd.serving import (ChemistryServer, poisson_workload,
                                         serve_plain, summarize)
    gases = {n: get_thermochem_class(bundled(n))(device=device)
             for n in ("sandiego", "gri30")}
    fmechs = [m for m in args.serve_flamelet_mechs.split(",") if m]
    arrivals = poisson_workload(gases, args.serve_requests, args.serve_rate,
                                seed=5, flamelet_mechanisms=fmechs)
    warm = poisson_workload(gases, 4, 100.0, seed=6, wdot_log2_cells=(10, 12),
                            flamelet_points=(64, 96),
                            flamelet_mechanisms=["sandiego"])
    ChemistryServer(gases).run(warm)
    print(f"[bench] serving: {len(arrivals)} requests at "
          f"{args.serve_rate}/s", file=sys.stderr, flush=True)
    srv = ChemistryServer(gases)
    tickets = srv.run(arrivals)
    lat = np.array([t.latency for t in tickets])
    t_first = min(t.t_arrival for t in tickets)
    t_last = max(t.t_done for t in tickets)
    kinds = np.array([t.kind for t in tickets])
    out = {"workload": f"{len(arrivals)} Poisson arrivals at "
                       f"{args.serve_rate}/s (seed 5): 50% omega-dot of "
                       "2^U{10..20} cells on sandiego or gri30, 50% flamelet "
                       "solves of U{64..2048} points on "
                       + " or ".join(fmechs),
           "batched": summarize(lat, t_first, t_last),
           "batched_by_kind": {k: summarize(lat[kinds == k], t_first, t_last)
                               for k in ("wdot", "flamelet")},
           "peak_hbm_bytes": srv.peak_used,
           "flamelets_converged": int(sum(
               t.result.converged for t in tickets if t.kind == "flamelet"))}
    print("[bench] serving: plain", file=sys.stderr, flush=True)
    lat_p, service, _ = serve_plain(gases, arrivals)
    t_end = max(ta + l for (ta, _), l in zip(arrivals, lat_p))
    out["plain"] = summarize(lat_p, 0.0, t_end)
    out["plain_by_kind"] = {k: summarize(lat_p[kinds == k], 0.0, t_end)
                            for k in ("wdot", "flamelet")}
    out["plain_service_s_total"] = float(service.sum())
    out["throughput_ratio"] = (out["batched"]["requests_per_s"]
                               / out["plain"]["requests_per_s"])
    out["p99_ratio_plain_over_batched"] = (out["plain"]["p99_s"]
                                           / out["batched"]["p99_s"])
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not "
              "supported (cells shard per GPU with no exchange; see the "
              "design notes)", file=sys.stderr)
        sys.exit(2)
    import torch
    from pyrometheus_amd import bundled, get_thermochem_class
    from pyrometheus_amd.flops import wdot_flops

    assert torch.cuda.is_available(), "bench.py needs a ROCm device"
    device = torch.device("cuda:0")
    if args.copy_probe:
        src = torch.ones(args.copy_probe // 8, dtype=torch.float64,
                         device=device)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 8}), flush=True)
        return
    if args.no_wdot:
        line = {"metric": METRIC, "n_gpus": 1, "data": "synthetic"}
        if not args.no_pointwise:
            line["pointwise"] = pointwise_leg(args, device)
        if not args.no_derivatives:
            line["derivatives"] = derivatives_leg(args, device)
        if not args.no_transport:
            line["transport"] = transport_leg(args, device)
        if not args.no_flamelets:
            _, line["flamelets"] = flamelet_leg(args, device)
        if not args.no_serving:
            line["serving"] = serving_leg(args, device)
        print(json.dumps(line), flush=True)
        return
    mech = bundled(args.mech)
    gas = get_thermochem_class(mech)(device=device)
    n = args.cells
    T, p, Y = device_inputs(mech, n, args.seed, device)
    rho = gas.get_density(p, T, Y)
    del p
    out = torch.empty((mech.n_species, n), dtype=torch.float64, device=device)
    print(f"[bench] {mech.name} cells={n} warmup={args.warmup} "
          f"steps={args.steps}", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        gas.get_net_production_rates(rho, T, Y, out=out)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        gas.get_net_production_rates(rho, T, Y, out=out)
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [a.elapsed_time(b) for a, b in ev]
    ms_step = wall * 1e3 / args.steps
    avg_launch = sum(launch_ms) / len(launch_ms)
    assert torch.isfinite(out).all().item()

    fl = wdot_flops(mech)
    achieved = fl["flop_per_cell"] * n / (avg_launch * 1e-3) / 1e12
    traffic = None
    if os.path.exists(args.traffic_json):
        with open(args.traffic_json) as f:
            tj = json.load(f)
        if tj.get("mechanism") == mech.name and tj.get("cells") == n:
            traffic = tj.get("bytes_per_launch")
    line = {
        "metric": METRIC,
        "value": n / (ms_step * 1e-3) / 1e6,
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"net production rates, {mech.name} "
                               f"({mech.n_species} species, "
                               f"{mech.n_reactions} reactions), {n} cells, "
                               "SoA FP64 state resident in HBM",
                   "cells": n, "mechanism": mech.name},
        "roofline": {"bound": "fp64-valu", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": traffic,
                     "pipe": "FP64 VALU (MI355X FP64 vector and matrix peaks "
                             "are both 78.6 TFLOP/s)",
                     "flop_per_cell": fl["flop_per_cell"],
                     "f_arith": fl["f_arith"], "n_trans": fl["n_trans"],
                     "w_trans": fl["w_trans"],
                     "kernel_ms": avg_launch,
                     "algorithmic_bytes_per_launch": fl["bytes_per_cell"] * n,
                     "hbm_gbs_algorithmic": fl["bytes_per_cell"] * n
                     / (avg_launch * 1e-3) / 1e9},
    }
    del rho, T, Y, out
    torch.cuda.empty_cache()
    if not args.no_pointwise:
        line["pointwise"] = pointwise_leg(args, device)
    if not args.no_derivatives:
        line["derivatives"] = derivatives_leg(args, device)
    if not args.no_transport:
        line["transport"] = transport_leg(args, device)
    line["config1"] = config1_leg(args, device)
    cpu_probs = None
    if not args.no_flamelets:
        cpu_probs, line["flamelets"] = flamelet_leg(args, device)
    if not args.no_serving:
        line["serving"] = serving_leg(args, device)
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(mech, args.cpu_seconds, args.seed,
                                            cpu_probs)
        fl = line["cpu_baseline"].get("flamelets")
        if fl and getattr(args, "batch_mean_n", None):
            f = fl["mean_n"] / args.batch_mean_n
            fl["value_at_batch_mean_n"] = fl["value"] * f
            fl["all_cores"]["value_at_batch_mean_n"] = (
                fl["all_cores"]["value"] * f)
        line["cpu_baseline"].update(cpu_small_baselines(host_cores()))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
