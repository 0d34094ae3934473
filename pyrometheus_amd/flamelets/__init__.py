"""Steady mixture-fraction flamelets on the device
(``pyrometheus/flamelets/__init__.py:32-48``)."""

from .domain import Domain, DomainConfig
from .linear_solver import block_thomas, block_thomas_batched
from .solver import (FlameletEquations, FlameletProblem, FlameletResult,
                     FlameletSolver, solve_batch)
from .state import FlameletState, _array_to_state, _state_to_array

__all__ = ["Domain", "DomainConfig", "FlameletEquations", "FlameletSolver",
           "FlameletState", "FlameletProblem", "FlameletResult",
           "solve_batch", "block_thomas", "block_thomas_batched",
           "_state_to_array", "_array_to_state"]
