"""The allocation-free block-Thomas entry point against the allocating one.

``pyro_block_thomas_batched_noalloc`` runs the kernels of
``pyro_block_thomas_batched`` with the library's fixed per-device workspace
instead of a caller-provided one, so both must give the same solution bit
for bit.  Two calls in a row (the second with fewer, shorter systems, so it
runs over the first call's leftover workspace), for the flamelet form
(scalar couplings, nv = 10) and general blocks (nv = 54), both through the
C ABI; plus the host-side rejection of a batch beyond the workspace.
"""
import numpy as np
import pytest
import torch

from pyrometheus_amd.runtime import HipError, load_core_library

pytestmark = pytest.mark.gpu


def _batch(rng, sizes, nv, diag, device):
    P = int(sum(sizes))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    C = rng.normal(size=(P, nv, nv)) + 4.0 * nv * np.eye(nv)
    if diag:
        L, U = rng.uniform(-1, 1, P), rng.uniform(-1, 1, P)
    else:
        L, U = rng.normal(size=(P, nv, nv)), rng.normal(size=(P, nv, nv))
    R = rng.normal(size=(P, nv))
    d = lambda a: torch.tensor(a, dtype=torch.float64, device=device)
    return (d(L), d(C), d(U), d(R),
            torch.tensor(off, dtype=torch.int64, device=device), P)


def _both(lib, L, C, U, R, off, P, nv, diag, device):
    nsys = off.numel() - 1
    x_ref = torch.full((P, nv), -1.0, dtype=torch.float64, device=device)
    x_na = torch.full((P, nv), -2.0, dtype=torch.float64, device=device)
    work = torch.empty(P * nv * (nv + 1), dtype=torch.float64, device=device)
    f_ref = torch.zeros(1, dtype=torch.int32, device=device)
    f_na = torch.zeros(1, dtype=torch.int32, device=device)
    stream = torch.cuda.current_stream().cuda_stream
    lib.call("pyro_block_thomas_batched", L.data_ptr(), C.data_ptr(),
             U.data_ptr(), R.data_ptr(), x_ref.data_ptr(), off.data_ptr(),
             nsys, nv, int(diag), work.data_ptr(), f_ref.data_ptr(), stream)
    lib.call("pyro_block_thomas_batched_noalloc", L.data_ptr(), C.data_ptr(),
             U.data_ptr(), R.data_ptr(), x_na.data_ptr(), off.data_ptr(),
             nsys, P, nv, int(diag), f_na.data_ptr(), stream)
    torch.cuda.synchronize()
    return x_ref, x_na, int(f_ref.item()), int(f_na.item())


@pytest.mark.parametrize("nv,diag,first,second", [
    (10, True, (64, 1024, 40), (64, 300, 9)),
    (54, False, (30, 200, 12), (8, 90, 5)),
])
def test_noalloc_matches_allocating_entry(nv, diag, first, second, device):
    lib = load_core_library()
    cap = int(lib.functions["pyro_block_thomas_scratch_doubles"]())
    rng = np.random.default_rng(nv)
    for lo, hi, nsys in (first, second):
        sizes = rng.integers(lo, hi + 1, nsys)
        print(f"[noalloc] nv={nv} diag={diag} systems={nsys} "
              f"rows={int(sizes.sum())}", flush=True)
        L, C, U, R, off, P = _batch(rng, sizes, nv, diag, device)
        assert P * nv * (nv + 1) <= cap
        x_ref, x_na, f_ref, f_na = _both(lib, L, C, U, R, off, P, nv, diag,
                                         device)
        assert f_ref == 0 and f_na == 0
        assert torch.isfinite(x_ref).all()
        assert torch.equal(x_ref, x_na)


def test_noalloc_rejects_oversized_batch(device):
    lib = load_core_library()
    cap = int(lib.functions["pyro_block_thomas_scratch_doubles"]())
    nv = 10
    rows = cap // (nv * (nv + 1)) + 1
    off = torch.tensor([0, rows], dtype=torch.int64, device=device)
    dummy = torch.zeros(1, dtype=torch.float64, device=device)
    with pytest.raises(HipError, match="status -4"):
        lib.call("pyro_block_thomas_batched_noalloc", dummy.data_ptr(),
                 dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(),
                 dummy.data_ptr(), off.data_ptr(), 1, rows, nv, 1, None,
                 torch.cuda.current_stream().cuda_stream)
