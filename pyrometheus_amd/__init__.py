"""pyrometheus_amd: MI355X-native thermochemistry and flamelet hot path.

The generator registry mirrors ``pyrometheus/__init__.py:40-45``; this package
registers the ``hip`` target (device libraries) and the ``cpp`` target (the
reference's C++ header, host and HIP device callable).  ``get_thermochem_class(mech)`` returns the
device class for a mechanism (the counterpart of
``PythonCodeGenerator.get_thermochem_class``).
"""

import typing

from .codegen import CodeGenerationOptions, CodeGenerator  # noqa: F401
from .codegen.cpp import CppCodeGenerator
from .codegen.hip import HipCodeGenerator
from .mechanism import Mechanism, bundled, load_mechanism  # noqa: F401


def get_code_generators() -> typing.Dict[str, CodeGenerator]:
    return {HipCodeGenerator.get_name(): HipCodeGenerator,
            CppCodeGenerator.get_name(): CppCodeGenerator}


def get_thermochem_class(mech: Mechanism):
    """Device ``Thermochemistry`` class for ``mech`` (builds its HIP library
    on first use when it is missing)."""
    from .thermochem import get_thermochem_class as _get
    return _get(mech)


__all__ = ["CodeGenerationOptions", "CodeGenerator", "HipCodeGenerator",
           "Mechanism", "bundled", "load_mechanism", "get_code_generators",
           "get_thermochem_class"]
