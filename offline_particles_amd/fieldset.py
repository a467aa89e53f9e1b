"""Fieldset: the field and constant registry (mirrors the reference's fieldset.py:11-241)."""

from __future__ import annotations

import types
from typing import Any, Mapping

import numpy as np

from .fields import Field


class Fieldset:
    """Fields of a simulation plus named constants and index bounds.

    Default index bounds are ``[0, size-1]`` of the centred grid in each
    dimension; they are registered as the constants ``zidx_min`` ...
    ``xidx_max`` that the validation kernel reads (fieldset.py:45-70).
    """

    def __init__(self, t_size: int, z_size: int, y_size: int, x_size: int, *,
                 constants: Mapping[str, Any] | None = None, zidx_bounds=None, yidx_bounds=None,
                 xidx_bounds=None, **fields: Field) -> None:
        self._t_size = t_size
        self._z_size = z_size
        self._y_size = y_size
        self._x_size = x_size
        if zidx_bounds is None:
            zidx_bounds = (0, z_size - 1)
        if yidx_bounds is None:
            yidx_bounds = (0, y_size - 1)
        if xidx_bounds is None:
            xidx_bounds = (0, x_size - 1)
        self._zidx_bounds = (np.float64(zidx_bounds[0]), np.float64(zidx_bounds[1]))
        self._yidx_bounds = (np.float64(yidx_bounds[0]), np.float64(yidx_bounds[1]))
        self._xidx_bounds = (np.float64(xidx_bounds[0]), np.float64(xidx_bounds[1]))
        self._fields: dict[str, Field] = {}
        self._constants: dict[str, Any] = {}
        if constants is not None:
            for name, value in constants.items():
                self.add_constant(name, value)
        self.add_constant("zidx_min", self._zidx_bounds[0])
        self.add_constant("zidx_max", self._zidx_bounds[1])
        self.add_constant("yidx_min", self._yidx_bounds[0])
        self.add_constant("yidx_max", self._yidx_bounds[1])
        self.add_constant("xidx_min", self._xidx_bounds[0])
        self.add_constant("xidx_max", self._xidx_bounds[1])
        for name, field in fields.items():
            self.add_field(name, field)

    t_size = property(lambda self: self._t_size)
    z_size = property(lambda self: self._z_size)
    y_size = property(lambda self: self._y_size)
    x_size = property(lambda self: self._x_size)
    zidx_bounds = property(lambda self: self._zidx_bounds)
    yidx_bounds = property(lambda self: self._yidx_bounds)
    xidx_bounds = property(lambda self: self._xidx_bounds)
    zidx_min = property(lambda self: self._zidx_bounds[0])
    zidx_max = property(lambda self: self._zidx_bounds[1])
    yidx_min = property(lambda self: self._yidx_bounds[0])
    yidx_max = property(lambda self: self._yidx_bounds[1])
    xidx_min = property(lambda self: self._xidx_bounds[0])
    xidx_max = property(lambda self: self._xidx_bounds[1])

    @property
    def simulation_shape(self) -> tuple:
        return (self._t_size, self._z_size, self._y_size, self._x_size)

    @property
    def fields(self) -> Mapping[str, Field]:
        return types.MappingProxyType(self._fields)

    @property
    def constants(self) -> Mapping[str, Any]:
        return types.MappingProxyType(self._constants)

    def add_field(self, name: str, field: Field) -> None:
        if name in self:
            raise KeyError(f"Field '{name}' already exists in Fieldset. First remove it before adding a new one.")
        try:
            field.validate_shape(self.simulation_shape)
        except ValueError as e:
            raise ValueError(f"Error validating shape of Field '{name}'.") from e
        self._fields[name] = field

    def add_constant(self, name: str, value: Any) -> None:
        if name in self._constants or name in self:
            raise KeyError(f"'{name}' already exists in Fieldset. First remove it before adding a new one.")
        self._constants[name] = _numpyify_constant(value)

    def remove(self, name: str) -> None:
        if name in self._constants:
            del self._constants[name]
            return
        if name in self._fields:
            del self._fields[name]
            return
        raise KeyError(f"Field '{name}' does not exist in Fieldset. Cannot remove.")

    def __getitem__(self, name: str) -> Field:
        if name in self._fields:
            return self._fields[name]
        raise KeyError(f"Field '{name}' does not exist in Fieldset.")

    def __contains__(self, name: str) -> bool:
        return name in self._fields

    def keys(self):
        return self._fields.keys()

    def values(self):
        return self._fields.values()

    def items(self):
        return self._fields.items()


def _numpyify_constant(value: Any):
    """A Python scalar from a single-valued input (fieldset.py:233-241)."""
    try:
        arr = np.asarray(value)
        if arr.size != 1:
            raise ValueError(f"Expected a single value, got array of size {arr.size}.")
        return arr.item()
    except (ValueError, TypeError) as e:
        raise ValueError(f"Cannot convert value '{value}' to a numpy scalar.") from e
