"""Autotuner for the block-Thomas schedule of flamelet batches.

A schedule is either ``"sequential"`` -- one wavefront walks each flamelet's
block rows (``k_block_thomas``) -- or ``"partitioned"`` with a flamelet of N
points cut into partitions of max(8, round(alpha sqrt(N))) rows, each swept
from both ends by two sweeps, then the nv-sized interface recursion and the
recovery (``pyro_flamelet_thomas_partitioned``).  Which is fastest, and with
which alpha, depends on the block size, on the device and on the spread of
flamelet sizes: short partitions mean short sweeps but a longer interface
recursion.

For 16 < nv <= 63 the partitioned solve has two forms: ``"mfma"``, a
blocked Gauss-Jordan sweep whose trailing updates, and an interface
recursion whose nv x nv products, run on the FP64 matrix cores, and
``"rowlane"``, one logical row per lane with the pivot row broadcast by
``v_readlane`` and the products as LDS broadcasts; both are candidates.

The schedule is kept per (device, nv), never per batch: a flamelet's
partitioning then depends on its own size only, so its arithmetic -- and its
result, bit for bit -- is the same alone and in any batch.  ``tune_thomas``
times every candidate on a synthetic batch with HIP events, checks that all
candidates solve it (they differ only in the order of elimination), and
stores the fastest; ``FlameletBatch(thomas="auto")`` uses the stored
schedule, or the default (partitioned, alpha = sqrt 2) before any tuning.
"""

from __future__ import annotations

import math
from typing import Dict, Optional, Sequence, Tuple

import numpy as np
import torch

DEFAULT = {"kind": "partitioned", "alpha": math.sqrt(2.0), "sweep": "mfma"}
SWEEPS = {"rowlane": 0, "mfma": 1}
ALPHAS = (0.5, 0.75, 1.0, math.sqrt(2.0), 2.0, 3.0)

_SCHEDULES: Dict[Tuple[str, int], dict] = {}


def _key(nv: int, device) -> Tuple[str, int]:
    device = torch.device(device)
    name = (torch.cuda.get_device_name(device) if device.type == "cuda"
            else device.type)
    return name, int(nv)


def schedule(nv: int, device) -> dict:
    """The stored schedule for blocks of size ``nv`` on ``device``, or the
    default."""
    return dict(_SCHEDULES.get(_key(nv, device), DEFAULT))


def set_schedule(nv: int, device, sched: dict):
    if sched.get("kind") not in ("sequential", "partitioned") or \
            sched.get("sweep", DEFAULT["sweep"]) not in SWEEPS:
        raise ValueError(f"unknown block-Thomas schedule {sched!r}")
    _SCHEDULES[_key(nv, device)] = {
        "kind": sched["kind"],
        "alpha": float(sched.get("alpha", DEFAULT["alpha"])),
        "sweep": sched.get("sweep", DEFAULT["sweep"])}


def reset():
    """Forget every tuned schedule."""
    _SCHEDULES.clear()


def _synthetic_batch(gas, sizes, seed, device):
    from .engine import FlameletBatch
    grids = [np.linspace(0.0, 1.0, int(n)) for n in sizes]
    zeros = np.zeros(gas.num_species)
    b = FlameletBatch(gas, grids, zeros, zeros, device=device,
                      thomas="sequential")
    g = torch.Generator(device=b.dev).manual_seed(seed)
    dd = dict(generator=g, dtype=torch.float64, device=b.dev)
    nv = b.nv
    # well-conditioned blocks with scalar couplings, the flamelet structure
    b.J.copy_(torch.randn((b.P, nv, nv), **dd)
              + (nv + 4.0) * torch.eye(nv, dtype=torch.float64,
                                       device=b.dev))
    b.L.copy_(torch.rand(b.P, **dd) - 0.5)
    b.U.copy_(torch.rand(b.P, **dd) - 0.5)
    b.rhs.copy_(torch.randn((b.P, nv), **dd))
    return b


def tune_thomas(gas, sizes: Optional[Sequence[int]] = None,
                alphas: Sequence[float] = ALPHAS, reps: int = 5,
                seed: int = 0, device=None, store: bool = True) -> dict:
    """Time the sequential and partitioned schedules (one per alpha) on a
    synthetic batch of flamelets with ``sizes`` points (default: 256 sizes
    uniform in [64, 1024], seed ``seed``) and keep the fastest for this
    mechanism's block size on this device.  Returns the chosen schedule with
    the median milliseconds of every candidate."""
    device = torch.device(device) if device is not None else gas.device
    if sizes is None:
        sizes = np.random.default_rng(seed).integers(64, 1025, 256)
    b = _synthetic_batch(gas, sizes, seed, device)
    candidates = [("sequential", None, DEFAULT["sweep"])]
    if b.supports_partitioned():
        sweeps = ("mfma", "rowlane") if b.has_sweep_forms() else ("mfma",)
        candidates += [("partitioned", float(a), sw) for a in alphas
                       for sw in sweeps]
    stream = torch.cuda.current_stream(b.dev)
    times, ref = {}, None
    for kind, alpha, sweep in candidates:
        b.set_thomas(kind, alpha, sweep=sweep)
        b.thomas()
        torch.cuda.synchronize(b.dev)
        x = b.x.clone()
        if ref is None:
            ref = x
        else:
            err = float(((x - ref).abs().max() / ref.abs().max()).item())
            if not err <= 1e-8:
                raise RuntimeError(f"block-Thomas schedule {kind} "
                                   f"alpha={alpha} {sweep} disagrees "
                                   f"({err:.2e})")
        ms = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            b.thomas()
            e1.record(stream)
            torch.cuda.synchronize(b.dev)
            ms.append(e0.elapsed_time(e1))
        name = kind if alpha is None else f"{kind} alpha={alpha:.3g}"
        if alpha is not None and b.has_sweep_forms():
            name += f" {sweep}"
        times[name] = (float(np.median(ms)), kind, alpha, sweep)
    best = min(times.values(), key=lambda t: t[0])
    chosen = {"kind": best[1],
              "alpha": best[2] if best[2] is not None else DEFAULT["alpha"],
              "sweep": best[3]}
    if store:
        set_schedule(b.nv, device, chosen)
    return dict(chosen, ms={k: v[0] for k, v in times.items()},
                nv=b.nv, systems=len(sizes), points=int(np.sum(sizes)))
