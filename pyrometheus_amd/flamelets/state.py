"""Flamelet state container (``pyrometheus/flamelets/state.py:191-316``)."""

from dataclasses import dataclass

import torch


@dataclass
class FlameletState:
    """``enthalpy`` (N,) and ``mass_fractions`` (ns, N), float64 tensors."""

    enthalpy: torch.Tensor
    mass_fractions: torch.Tensor

    @property
    def num_vars(self):
        return 1 + self.mass_fractions.shape[0]

    @property
    def spatial_shape(self):
        return tuple(self.enthalpy.shape)

    def _binary(self, other, op):
        if isinstance(other, FlameletState):
            return FlameletState(op(self.enthalpy, other.enthalpy),
                                 op(self.mass_fractions, other.mass_fractions))
        return FlameletState(op(self.enthalpy, other),
                             op(self.mass_fractions, other))

    def __add__(self, o):
        return self._binary(o, torch.add)

    def __sub__(self, o):
        return self._binary(o, torch.sub)

    def __mul__(self, o):
        return self._binary(o, torch.mul)

    __rmul__ = __mul__

    def zeros_like(self):
        return FlameletState(torch.zeros_like(self.enthalpy),
                             torch.zeros_like(self.mass_fractions))


def _state_to_array(state: FlameletState) -> torch.Tensor:
    """(nv, N): row 0 enthalpy, rows 1.. mass fractions."""
    return torch.vstack([state.enthalpy.reshape(1, -1), state.mass_fractions])


def _array_to_state(array: torch.Tensor) -> FlameletState:
    return FlameletState(array[0], array[1:])
