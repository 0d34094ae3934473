"""Build of the C++ header kernel check: the generated header of a
mechanism (pyrometheus_amd.codegen.cpp) and tests/cpp_target_shim.cpp
compiled by hipcc for gfx950 into ``pyrometheus_amd/_gen/`` (called by
``__graft_entry__.build()``; the GPU test loads the result)."""

import hashlib
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SHIM = os.path.join(HERE, "cpp_target_shim.cpp")


def _digest(name: str) -> str:
    from pyrometheus_amd import bundled
    from pyrometheus_amd.codegen import cpp
    from pyrometheus_amd.runtime import source_digest
    h = hashlib.sha256((bundled(name).digest() + source_digest()).encode())
    for path in (SHIM, cpp.__file__):
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:8]


def check_library_path(name: str) -> str:
    from pyrometheus_amd.runtime import GEN_DIR
    return os.path.join(GEN_DIR, f"libpyro_cppcheck_{name}_{_digest(name)}.so")


def build_check_library(name: str) -> str:
    from pyrometheus_amd import bundled
    from pyrometheus_amd.codegen.cpp import CppCodeGenerator
    from pyrometheus_amd.runtime import GEN_DIR, compile_hip
    os.makedirs(GEN_DIR, exist_ok=True)
    so = check_library_path(name)
    hdr = so[:-3] + ".hpp"
    with open(hdr, "w") as f:
        f.write(CppCodeGenerator.generate(name, bundled(name)))
    compile_hip(SHIM, so, extra=("-x", "hip", "-std=c++20",
                                 f'-DPYRO_HEADER="{hdr}"',
                                 f"-DPYRO_NAME={name}"))
    return so
