"""Block-tridiagonal solves on the device.

``block_thomas(lower, central, upper, rhs)`` keeps the reference signature
and block layout (``pyrometheus/flamelets/linear_solver.py:6-103``): lower
and upper are ``(N-1, nv, nv)``, central ``(N, nv, nv)``, rhs ``(N, nv)``;
the result is ``(N, nv)``.  ``block_thomas_batched`` solves many systems of
uneven length in one launch (one wave per system) from concatenated block
rows and an offsets vector.  Both run the HIP kernel of ``libpyro_core``;
each block row is reduced by Gauss-Jordan elimination with partial pivoting
(the reference calls LU with partial pivoting, so results agree to rounding).
"""

from __future__ import annotations

from typing import Optional

import torch

from ..runtime import HipError, load_core_library


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _check(t: torch.Tensor, name: str):
    if t.dtype != torch.float64 or t.device.type != "cuda":
        raise TypeError(f"{name} must be a float64 tensor on the ROCm device")


def block_thomas_batched(lower: torch.Tensor, central: torch.Tensor,
                         upper: torch.Tensor, rhs: torch.Tensor,
                         offsets: torch.Tensor, diag: bool = False,
                         out: Optional[torch.Tensor] = None,
                         work: Optional[torch.Tensor] = None,
                         check_singular: bool = False) -> torch.Tensor:
    """Solve the systems whose block rows are ``offsets[s]:offsets[s+1]``.

    ``central`` is ``(P, nv, nv)`` and ``rhs`` ``(P, nv)``.  With ``diag``
    false, ``lower[g]``/``upper[g]`` are ``(nv, nv)`` blocks (row ``g+1`` to
    ``g`` / row ``g`` to ``g+1``); with ``diag`` true they are ``(P,)``
    scalars multiplying the identity.  The last entry of each system in
    ``lower``/``upper`` is not read."""
    lib = load_core_library()
    for t, nm in ((lower, "lower"), (central, "central"), (upper, "upper"),
                  (rhs, "rhs")):
        _check(t, nm)
    P, nv = rhs.shape
    if central.shape != (P, nv, nv):
        raise ValueError("central must be (P, nv, nv)")
    want = (P,) if diag else (P, nv, nv)
    if lower.shape != want or upper.shape != want:
        raise ValueError(f"lower/upper must be {want}")
    if not lib.call("pyro_block_thomas_supported", nv):
        raise HipError(f"no block-Thomas kernel for nv={nv}")
    offsets = offsets.to(device=rhs.device, dtype=torch.int64).contiguous()
    nsys = offsets.numel() - 1
    lower, central, upper, rhs = (t.contiguous() for t in
                                  (lower, central, upper, rhs))
    if out is None:
        out = torch.empty_like(rhs)
    if work is None or work.numel() < P * nv * (nv + 1):
        work = torch.empty(P * nv * (nv + 1), dtype=torch.float64,
                           device=rhs.device)
    flag = (torch.zeros(1, dtype=torch.int32, device=rhs.device)
            if check_singular else None)
    lib.call("pyro_block_thomas_batched", lower.data_ptr(), central.data_ptr(),
             upper.data_ptr(), rhs.data_ptr(), out.data_ptr(),
             offsets.data_ptr(), nsys, nv, int(diag), work.data_ptr(),
             None if flag is None else flag.data_ptr(), _stream())
    if flag is not None and int(flag.item()) > 0:
        raise RuntimeError(f"{int(flag.item())} block-tridiagonal system(s) "
                           "met a singular pivot block")
    return out


def block_thomas_batched_noalloc(lower: torch.Tensor, central: torch.Tensor,
                                 upper: torch.Tensor, rhs: torch.Tensor,
                                 offsets: torch.Tensor, out: torch.Tensor,
                                 diag: bool = False,
                                 n_singular: Optional[torch.Tensor] = None
                                 ) -> torch.Tensor:
    """``block_thomas_batched`` through the allocation-free entry point: the
    library's fixed per-device workspace replaces ``work``, so nothing is
    allocated (safe under graph capture).  Every argument, ``out`` included,
    must already be a contiguous device tensor (``offsets`` int64,
    ``n_singular`` a zeroed int32 counter or None); the batch may hold at
    most ``scratch_doubles() // (nv * (nv + 1))`` block rows.  Calls that
    overlap on one device race on the shared workspace."""
    lib = load_core_library()
    P, nv = rhs.shape
    for t in (lower, central, upper, rhs, out, offsets):
        if not t.is_contiguous() or t.device.type != "cuda":
            raise TypeError("arguments must be contiguous device tensors")
    if offsets.dtype != torch.int64 or out.shape != rhs.shape:
        raise TypeError("offsets must be int64 and out shaped like rhs")
    cap = scratch_doubles() // (nv * (nv + 1))
    if P > cap:
        raise ValueError(f"{P} block rows of size {nv} exceed the fixed "
                         f"workspace ({cap} rows)")
    lib.call("pyro_block_thomas_batched_noalloc", lower.data_ptr(),
             central.data_ptr(), upper.data_ptr(), rhs.data_ptr(),
             out.data_ptr(), offsets.data_ptr(), offsets.numel() - 1, P, nv,
             int(diag), None if n_singular is None else n_singular.data_ptr(),
             _stream())
    return out


def scratch_doubles() -> int:
    """Size in doubles of the allocation-free entry point's workspace."""
    return int(load_core_library().functions[
        "pyro_block_thomas_scratch_doubles"]())


def block_thomas(lower: torch.Tensor, central: torch.Tensor,
                 upper: torch.Tensor, rhs: torch.Tensor) -> torch.Tensor:
    """Drop-in for the reference ``block_thomas`` on one system."""
    n, nv = rhs.shape
    pad = torch.zeros((1, nv, nv), dtype=torch.float64, device=rhs.device)
    lo = torch.cat([lower.reshape(-1, nv, nv), pad]) if n > 1 else pad
    up = torch.cat([upper.reshape(-1, nv, nv), pad]) if n > 1 else pad
    offs = torch.tensor([0, n], dtype=torch.int64, device=rhs.device)
    return block_thomas_batched(lo, central, up, rhs, offs, diag=False)
