"""Particles: SoA device tensors (mirrors the reference's particles.py:10-122).

Each particle field is one zero-initialised 1-D tensor in device memory,
accessible as ``particles.name`` or ``particles["name"]``.  ``ParticlesView``
hands out read-only access (a copy-free view; torch has no read-only flag, so
the view returns tensors that callers must not modify).
"""

from __future__ import annotations

import types

import numpy as np
import torch

from .spatial_arrays import default_device, to_torch_dtype


class _FrozenArrayMapping:
    __slots__ = ("_shape", "_dtypes", "_arrays")

    def __init__(self, **arrays: torch.Tensor) -> None:
        shapes = {tuple(a.shape) for a in arrays.values()}
        if len(shapes) != 1:
            raise ValueError("All arrays must have the same shape. Got shapes: " + ", ".join(str(s) for s in shapes))
        object.__setattr__(self, "_shape", shapes.pop())
        object.__setattr__(self, "_dtypes", types.MappingProxyType({n: a.dtype for n, a in arrays.items()}))
        object.__setattr__(self, "_arrays", types.MappingProxyType(arrays))

    def __setattr__(self, name, value):
        raise AttributeError(f"{self.__class__.__name__} is immutable")

    def __getattr__(self, name: str) -> torch.Tensor:
        try:
            return object.__getattribute__(self, "_arrays")[name]
        except KeyError:
            raise AttributeError(f"'{self.__class__.__name__}' object has no attribute '{name}'") from None

    def __getitem__(self, name: str) -> torch.Tensor:
        return object.__getattribute__(self, "_arrays")[name]

    def __contains__(self, name: str) -> bool:
        return name in object.__getattribute__(self, "_arrays")

    @property
    def shape(self) -> tuple:
        return object.__getattribute__(self, "_shape")

    @property
    def arrays(self):
        return object.__getattribute__(self, "_arrays")

    @property
    def dtypes(self):
        return object.__getattribute__(self, "_dtypes")

    def to_numpy(self, name: str) -> np.ndarray:
        """Host copy of one field."""
        return self[name].detach().cpu().numpy()

    def __repr__(self) -> str:
        fields = ", ".join(f"{n}:{d}" for n, d in self.dtypes.items())
        return f"{self.__class__.__name__}(shape={self.shape}, fields={{ {fields} }})"


class Particles(_FrozenArrayMapping):
    __slots__ = ("_length", "_device")

    def __init__(self, nparticles: int, *, device=None, **fields) -> None:
        device = torch.device(device) if device is not None else default_device()
        object.__setattr__(self, "_length", int(nparticles))
        object.__setattr__(self, "_device", device)
        arrays = {name: torch.zeros((int(nparticles),), dtype=to_torch_dtype(dt), device=device)
                  for name, dt in fields.items()}
        super().__init__(**arrays)

    def __len__(self) -> int:
        return object.__getattribute__(self, "_length")

    @property
    def device(self) -> torch.device:
        return object.__getattribute__(self, "_device")


class ParticlesView(_FrozenArrayMapping):
    """Read-only access to a Particles object's arrays.

    With ``order`` (a callable returning None or an int32 device tensor whose
    entry j is the particle held in slot j), the arrays are kept in a
    spatial order and each field is returned in the particles' own order: a
    new tensor, out[order[j]] = slot j (without an order, the arrays
    themselves)."""

    __slots__ = ("_length", "_order")

    def __init__(self, parent: Particles, order=None) -> None:
        object.__setattr__(self, "_length", len(parent))
        object.__setattr__(self, "_order", order)
        super().__init__(**dict(parent.arrays))

    def __len__(self) -> int:
        return object.__getattribute__(self, "_length")

    def __getitem__(self, name: str) -> torch.Tensor:
        a = super().__getitem__(name)
        order_fn = object.__getattribute__(self, "_order")
        order = order_fn() if order_fn is not None else None
        if order is None:
            return a
        out = torch.empty_like(a)
        out[order.long()] = a
        return out

    def __getattr__(self, name: str) -> torch.Tensor:
        if name in object.__getattribute__(self, "_arrays"):
            return self[name]
        raise AttributeError(f"'{self.__class__.__name__}' object has no attribute '{name}'")

    @property
    def arrays(self):
        """The fields in the particles' order (new tensors while the arrays are
        kept in a spatial order, else the arrays themselves)."""
        raw = object.__getattribute__(self, "_arrays")
        order_fn = object.__getattribute__(self, "_order")
        if order_fn is None or order_fn() is None:
            return raw
        return types.MappingProxyType({name: self[name] for name in raw})
