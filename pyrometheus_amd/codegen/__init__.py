"""Code-generator plugin interface.

Mirrors the reference's generator ABC (``pyrometheus/codegen/__init__.py:8-44``):
a generator is a class of static methods ``get_name()``,
``supports_overloading()`` and ``generate(name, sol, opts)``, configured by a
frozen :class:`CodeGenerationOptions`.  ``sol`` is this package's
:class:`~pyrometheus_amd.mechanism.Mechanism` (the reference passes a Cantera
``Solution``; the IR carries the same information).
"""

from abc import abstractmethod
from dataclasses import dataclass
from typing import Optional


@dataclass(frozen=True)
class CodeGenerationOptions:
    scalar_type: Optional[str] = None
    directive_offload: Optional[str] = None


class CodeGenerator:
    @staticmethod
    @abstractmethod
    def get_name() -> str:
        """Returns the name (slug) of the code generator."""

    @staticmethod
    @abstractmethod
    def supports_overloading() -> bool:
        """Returns whether the code generator supports operator overloading."""

    @staticmethod
    @abstractmethod
    def generate(name: str, sol, opts: CodeGenerationOptions = None) -> str:
        """Generate thermochemistry source for the mechanism ``sol``."""
