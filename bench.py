"""Benchmark: net production rates of a 53-species FP64 mechanism.

One step = one ``get_net_production_rates`` evaluation (one fused HIP kernel)
over 2^24 cells of GRI-3.0 (53 species, 325 reactions) with the state already
resident in HBM: T ~ U[300, 2500] K, p log-uniform in [0.5, 20] atm,
Y = 0.5 Dirichlet(0.5) + 0.5 e_N2 (synthetic, seeded).

Prints one JSON line with the throughput in Mcells/s, the FP64 roofline of the
kernel (algorithmic FLOPs per cell from ``pyrometheus_amd.flops`` divided by
the kernel's average launch time, measured with HIP events on the stream it
runs on), and the CPU baseline: the numpy oracle (the reference's op
structure) timed on a bounded sample on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells N]
                       [--mech gri30] [--cpu-seconds S] [--no-cpu]
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


def cpu_baseline(mech, seconds, seed):
    """numpy oracle on chunks of 2^16 cells until ~``seconds`` of work."""
    import numpy as np
    from oracle.thermochem_np import Thermochemistry
    from pyrometheus_amd.mechanism import ONE_ATM
    gas = Thermochemistry(mech)
    rng = np.random.default_rng(seed)
    chunk = 1 << 16
    done, t0 = 0, time.perf_counter()
    elapsed = 0.0
    while elapsed < seconds:
        T = rng.uniform(300.0, 2500.0, chunk)
        p = ONE_ATM * np.exp(rng.uniform(np.log(0.5), np.log(20.0), chunk))
        Y = 0.5 * rng.dirichlet(0.5 * np.ones(mech.n_species), chunk).T
        Y[mech.species_index("N2")] += 0.5
        rho = gas.get_density(p, T, Y)
        t1 = time.perf_counter()
        gas.get_net_production_rates(rho, T, Y)
        elapsed += time.perf_counter() - t1
        done += chunk
    return {"value": done / elapsed / 1e6, "unit": "Mcells/s", "cores": 1,
            "kind": "port",
            "sample": f"{done} cells of {mech.name} in chunks of {chunk}, "
                      f"numpy restatement of the generated code (whole-array "
                      f"ufunc per term), 1 process, CPU "
                      f"{platform.processor() or platform.machine()}, "
                      f"{os.cpu_count()} host threads visible; JAX absent"}


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
        "metric": "Mcells/s omega-dot eval (53-sp FP64)",
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
        "roofline": {"bound": "mfma", "achieved": achieved,
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
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(mech, args.cpu_seconds, args.seed)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
