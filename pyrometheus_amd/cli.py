"""Command line: generate the source of a mechanism for a target.

``python -m pyrometheus_amd -l hip|cpp -m MECH.yaml -o OUT -n NAME [-p PHASE]``
takes the reference's options (``pyrometheus/cli.py:11-42``: ``-l/--lang``
from the generator registry, ``-m/--mech``, ``-o/--output``,
``-n/--name/--namespace``, ``-p/--phase``) and writes the generated source.
The mechanism is read by this package's Cantera-YAML reader (Cantera itself
is not needed).  ``--build`` additionally compiles the source for gfx950
into a shared library next to the output (``OUT.so``) with the package's
compiler flags (``hip`` only: the ``cpp`` target is a header).
"""

from __future__ import annotations

import argparse
import os
import sys
from typing import Optional, Sequence


def _version() -> str:
    try:
        from importlib.metadata import version
        return version("pyrometheus_amd")
    except Exception:  # not installed as a distribution: the source tree
        return "source tree"


def main(argv: Optional[Sequence[str]] = None) -> int:
    from . import get_code_generators
    from .mechanism import load_mechanism

    generators = get_code_generators()
    parser = argparse.ArgumentParser(
        prog="pyrometheus_amd",
        description="Code generation for combustion thermochemistry "
                    "(MI355X HIP libraries, C++ header).")
    parser.add_argument("--version", action="version",
                        version=f"%(prog)s {_version()}")
    parser.add_argument("-l", "--lang", "--language",
                        help="Language to generate code for.",
                        choices=generators.keys(), required=True)
    parser.add_argument("-m", "--mech", "--mechanism",
                        help="Path to the mechanism file.", required=True)
    parser.add_argument("-o", "--output",
                        help="Path to the output file.", required=True)
    parser.add_argument("-n", "--name", "--namespace",
                        help="Namespace to use for the generated code.",
                        required=True)
    parser.add_argument("-p", "--phase",
                        help="Phase name to use for the generated code.")
    parser.add_argument("--build", action="store_true",
                        help="Also compile the source into OUTPUT's stem "
                             "with .so (hipcc, gfx950).")
    args = parser.parse_args(argv)

    mech = load_mechanism(args.mech, args.phase)
    source = generators[args.lang].generate(args.name, mech)
    with open(args.output, "w") as f:
        f.write(source)
    if args.build:
        if args.lang != "hip":
            parser.error("--build compiles the hip target only")
        from .runtime import compile_hip
        so = os.path.splitext(args.output)[0] + ".so"
        compile_hip(os.path.abspath(args.output), so)
        print(so)
    return 0


if __name__ == "__main__":
    sys.exit(main())
