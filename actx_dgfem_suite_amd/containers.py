"""Array containers of the three operators and their flattening to the field
order the C ABI uses.

* ``WaveState(u, v)`` mirrors equations/wave.py:63-77 (``v`` an object array
  of 3 leaves).
* ``ConservedEulerField(mass, energy, momentum)`` mirrors upstream
  ``grudge.models.euler.ConservedEulerField`` as used in
  equations/euler.py:37,91-93.
* Maxwell state is a flat object array ``[E_x, E_y, E_z, H_x, H_y, H_z]``
  (equations/maxwell.py:61-65, 73-74).

Leaves are ``(K, Np)`` arrays (torch tensors on the device, numpy on the host).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Callable

import numpy as np


def obj_array(items) -> np.ndarray:
    out = np.empty(len(items), dtype=object)
    for k, x in enumerate(items):
        out[k] = x
    return out


@dataclass(frozen=True)
class WaveState:
    u: Any
    v: np.ndarray  # object array of 3 leaves

    def __post_init__(self):
        if not (isinstance(self.v, np.ndarray) and self.v.dtype.char == "O"):
            raise TypeError("WaveState.v must be an object array")


@dataclass(frozen=True)
class ConservedEulerField:
    mass: Any
    energy: Any
    momentum: np.ndarray  # object array of 3 leaves

    def __post_init__(self):
        if not (isinstance(self.momentum, np.ndarray) and self.momentum.dtype.char == "O"):
            raise TypeError("ConservedEulerField.momentum must be an object array")


def flatten_fields(state) -> list:
    """Leaves in C-ABI field order."""
    if isinstance(state, WaveState):
        return [state.u, *state.v]
    if isinstance(state, ConservedEulerField):
        return [state.mass, state.energy, *state.momentum]
    if isinstance(state, np.ndarray) and state.dtype.char == "O":
        return list(state)
    if isinstance(state, (list, tuple)):
        return list(state)
    raise TypeError(f"unsupported state container {type(state)}")


def unflatten_like(template, leaves: list):
    if isinstance(template, WaveState):
        return WaveState(u=leaves[0], v=obj_array(leaves[1:4]))
    if isinstance(template, ConservedEulerField):
        return ConservedEulerField(mass=leaves[0], energy=leaves[1],
                                   momentum=obj_array(leaves[2:5]))
    if isinstance(template, np.ndarray) and template.dtype.char == "O":
        return obj_array(leaves)
    if isinstance(template, tuple):
        return tuple(leaves)
    if isinstance(template, list):
        return list(leaves)
    raise TypeError(f"unsupported state container {type(template)}")


def map_leaves(f: Callable, state):
    if isinstance(state, (WaveState, ConservedEulerField)) or (
            isinstance(state, np.ndarray) and state.dtype.char == "O"):
        return unflatten_like(state, [f(x) for x in flatten_fields(state)])
    if isinstance(state, (list, tuple)):
        return type(state)(map_leaves(f, x) for x in state)
    return f(state)
