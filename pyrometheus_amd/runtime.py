"""Building and loading the HIP shared libraries behind the C ABI.

Two kinds of library exist:

* one per mechanism, ``_gen/libpyro_<name>_<digest>.so``, compiled from the
  ``hip`` target's generated source plus ``csrc/pyro_thermo_kernels.hpp``;
* the mechanism-independent ``_gen/libpyro_core.so`` (batched block-Thomas and
  flamelet assembly) from ``csrc/pyro_core.hip``.

Libraries are built in-tree with ``hipcc --offload-arch=gfx950`` and loaded
with :mod:`ctypes`.  If a library is missing and ``hipcc`` is not on the
machine, loading raises: there is no CPU fallback for any product path.
"""

from __future__ import annotations

import ctypes
import os
import shutil
import subprocess
import threading
from typing import Dict, Optional

from .mechanism import Mechanism

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC_DIR = os.path.join(PKG_DIR, "csrc")
GEN_DIR = os.path.join(PKG_DIR, "_gen")
ARCH = os.environ.get("PYRO_OFFLOAD_ARCH", "gfx950")

# Every HIP translation unit is compiled with packed-FP32 instructions
# disabled; the FP64 paths never need them.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-rdc", "-ffp-contract=fast",
]

_lock = threading.Lock()
_loaded: Dict[str, "HipLibrary"] = {}


class HipError(RuntimeError):
    """A C-ABI call returned a negative status."""


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise HipError("hipcc not found; the HIP libraries must be built "
                       "before use (run __graft_entry__.build())")
    return exe


def compile_hip(src_path: str, out_path: str, extra=()) -> None:
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{CSRC_DIR}", *extra, src_path,
           "-o", out_path + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise HipError(f"hipcc failed for {src_path}:\n{res.stderr[-4000:]}")
    os.replace(out_path + ".tmp", out_path)


_SRC_DIGEST = None


def source_digest() -> str:
    """Hash of everything a generated library is compiled from besides the
    mechanism: the hand-written headers, the generator and the flags."""
    global _SRC_DIGEST
    if _SRC_DIGEST is None:
        import hashlib
        h = hashlib.sha256(" ".join(HIPCC_FLAGS).encode())
        files = sorted(f for f in os.listdir(CSRC_DIR) if f.endswith(".hpp"))
        for f in files:
            with open(os.path.join(CSRC_DIR, f), "rb") as fh:
                h.update(fh.read())
        for f in (("codegen", "hip.py"), ("codegen", "hip_transport.py"),
                  ("transport.py",)):
            with open(os.path.join(PKG_DIR, *f), "rb") as fh:
                h.update(fh.read())
        _SRC_DIGEST = h.hexdigest()[:8]
    return _SRC_DIGEST


def mechanism_lib_path(mech: Mechanism) -> str:
    return os.path.join(
        GEN_DIR, f"libpyro_{mech.name}_{mech.digest()}_{source_digest()}.so")


def generate_source(mech: Mechanism) -> str:
    from .codegen.hip import HipCodeGenerator
    return HipCodeGenerator.generate(mech.name, mech)


def build_mechanism_library(name_or_mech, source: Optional[str] = None,
                            force: bool = False) -> str:
    """Generate (if needed) and compile the library for a mechanism.
    Returns the ``.so`` path."""
    os.makedirs(GEN_DIR, exist_ok=True)
    if isinstance(name_or_mech, Mechanism):
        mech = name_or_mech
        so = mechanism_lib_path(mech)
        src = so[:-3] + ".hip"
        if source is None:
            source = generate_source(mech)
    else:
        import hashlib
        digest = hashlib.sha256(source.encode()).hexdigest()[:16]
        so = os.path.join(GEN_DIR, f"libpyro_{name_or_mech}_{digest}.so")
        src = so[:-3] + ".hip"
    if os.path.exists(so) and not force:
        return so
    with open(src, "w") as f:
        f.write(source)
    compile_hip(src, so)
    return so


def core_lib_path() -> str:
    import hashlib
    with open(os.path.join(CSRC_DIR, "pyro_core.hip"), "rb") as f:
        d = hashlib.sha256(f.read() + source_digest().encode()).hexdigest()[:8]
    return os.path.join(GEN_DIR, f"libpyro_core_{d}.so")


def prune_stale_libraries(keep) -> list:
    """Delete generated sources and libraries under ``_gen/`` built from
    older sources (their file names carry other digests) and not in
    ``keep``; returns the removed paths."""
    if not os.path.isdir(GEN_DIR):
        return []
    src = source_digest()
    keep = {os.path.basename(k)[:-3] for k in keep}
    removed = []
    for f in sorted(os.listdir(GEN_DIR)):
        stem, ext = os.path.splitext(f)
        if ext not in (".so", ".hip", ".hpp") or \
                not stem.startswith("libpyro_"):
            continue
        if stem in keep or stem.endswith("_" + src):
            continue
        os.remove(os.path.join(GEN_DIR, f))
        removed.append(f)
    return removed


def build_core_library(force: bool = False) -> str:
    os.makedirs(GEN_DIR, exist_ok=True)
    so = core_lib_path()
    src = os.path.join(CSRC_DIR, "pyro_core.hip")
    if os.path.exists(so) and not force:
        return so
    compile_hip(src, so)
    return so


# ---------------------------------------------------------------------------
# ctypes binding
# ---------------------------------------------------------------------------
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I = ctypes.c_int
_D = ctypes.c_double

#: argument signatures of the per-mechanism C ABI (include/pyro_hip.h)
MECH_SIGNATURES = {
    "pyro_mech_info": (_P, _I),
    "pyro_species_thermo": (_I, _P, _I64, _P, _I64, _I64, _P),
    "pyro_mixture_thermo": (_I, _P, _I64, _P, _I64, _P, _I64, _P),
    "pyro_mixture_state": (_I, _P, _I64, _P, _I64, _P, _I64, _P, _I64, _P),
    "pyro_species_state": (_I, _P, _I64, _P, _I64, _P, _I64, _I64, _P),
    "pyro_reaction_quantities": (_P, _P, _I64, _P, _P, _P, _P, _P, _I64,
                                 _I64, _P),
    "pyro_wdot": (_P, _P, _P, _I64, _P, _I64, _I64, _P),
    "pyro_creation_destruction": (_P, _P, _P, _I64, _P, _P, _I64, _I64, _P),
    "pyro_temperature": (_P, _I64, _P, _I64, _P, _I64, _P, _I64, _I, _I, _D,
                         _I, _P, _P, _P),
    "pyro_wdot_jvp": (_P, _P, _P, _I64, _P, _P, _P, _I64, _P, _P, _I64,
                      _I64, _P),
    "pyro_wdot_vjp": (_P, _P, _P, _I64, _P, _I64, _P, _P, _P, _I64, _I64, _P),
    "pyro_source_jacobian": (_P, _P, _I64, _P, _I64, _P, _P, _I64, _I64, _P,
                             _I64, _P, _I64, _P),
    "pyro_source_jacobian_work_doubles": (),
    "pyro_source_jacobian_form": (_I, _I),
    "pyro_flamelet_source": (_P, _I64, _P, _P, _I64, _P, _I64, _P, _I64, _P,
                             _P, _I64, _D, _I, _P, _I, _P),
    "pyro_flamelet_assemble": (_P, _I64, _P, _I64, _P, _P, _P, _P, _P, _P,
                               _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                               _I64, _P),
    "pyro_flamelet_thomas": (_P, _P, _P, _P, _P, _P, _I64, _P, _P, _P),
    "pyro_flamelet_thomas_sweep_form": (_I,),
    "pyro_flamelet_thomas_partitioned": (_P, _P, _P, _P, _P, _P, _I64, _P,
                                         _P, _P, _I64, _P, _P, _P, _I64, _P,
                                         _P, _P, _P, _P, _P),
    "pyro_flamelet_update": (_P, _I64, _P, _P, _P, _P, _P, _P, _I64, _P,
                             _I64, _P),
    "pyro_flamelet_eos_rt": (_P, _I64, _P, _I64, _P, _P, _I64, _I64, _D, _I,
                             _P, _I, _P),
    "pyro_species_transport": (_I, _P, _I64, _P, _I64, _I64, _P),
    "pyro_mixture_transport": (_I, _P, _I64, _P, _I64, _P, _I64, _P, _I64,
                               _I64, _P),
    "pyro_transport_fits": (_I, _P, _I64),
    "pyro_flamelet_adjoint_blocks": (_P, _P, _P, _P, _P, _P, _P, _P, _P,
                                     _I64, _P),
}


class HipLibrary:
    """A loaded ``.so`` with typed entry points; negative status raises."""

    def __init__(self, path: str, signatures: Dict[str, tuple]):
        self.path = path
        self._lib = ctypes.CDLL(path, mode=os.RTLD_LOCAL | os.RTLD_NOW)
        self.functions = {}
        for name, argtypes in signatures.items():
            fn = getattr(self._lib, name, None)
            if fn is None:
                continue
            fn.argtypes = list(argtypes)
            # size queries return int64; everything else an int status
            fn.restype = (ctypes.c_int64 if name.endswith("_doubles")
                          else ctypes.c_int)
            self.functions[name] = fn

    def has(self, name: str) -> bool:
        return name in self.functions

    def call(self, name: str, *args) -> int:
        fn = self.functions.get(name)
        if fn is None:
            raise HipError(f"{os.path.basename(self.path)} does not export "
                           f"{name}")
        st = fn(*args)
        if st < 0:
            raise HipError(f"{name} failed with status {st}")
        return st


def load_mechanism_library(mech: Mechanism, build: bool = True) -> HipLibrary:
    """Load the library for ``mech``; build it first if it is missing and
    ``build`` is true."""
    so = mechanism_lib_path(mech)
    with _lock:
        if so in _loaded:
            return _loaded[so]
        if not os.path.exists(so):
            if not build:
                raise HipError(f"HIP library for mechanism {mech.name!r} is "
                               f"not built ({so})")
            build_mechanism_library(mech)
        lib = HipLibrary(so, MECH_SIGNATURES)
        info = (ctypes.c_int64 * 4)()
        lib.call("pyro_mech_info", ctypes.cast(info, ctypes.c_void_p), 4)
        if info[0] != mech.n_species or info[1] != mech.n_reactions:
            raise HipError(f"{so} does not match mechanism {mech.name}")
        lib.info = tuple(info)
        _loaded[so] = lib
        return lib


CORE_SIGNATURES = {
    "pyro_block_thomas_supported": (_I,),
    "pyro_block_thomas_batched": (_P, _P, _P, _P, _P, _P, _I64, _I, _I, _P,
                                  _P, _P),
    "pyro_block_thomas_scratch_doubles": (),
    "pyro_block_thomas_batched_noalloc": (_P, _P, _P, _P, _P, _P, _I64, _I64,
                                          _I, _I, _P, _P),
    "pyro_flamelet_control": (_P, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _P),
    "pyro_flamelet_state_copy": (_P, _P, _P, _I64, _I, _P, _P, _I64, _P),
}


def load_core_library(build: bool = True) -> HipLibrary:
    so = core_lib_path()
    with _lock:
        if so in _loaded:
            return _loaded[so]
        if not os.path.exists(so):
            if not build:
                raise HipError(f"HIP core library is not built ({so})")
            build_core_library()
        lib = HipLibrary(so, CORE_SIGNATURES)
        _loaded[so] = lib
        return lib
