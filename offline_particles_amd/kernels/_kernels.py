"""Particle kernel plugin API (mirrors the reference's kernels/_kernels.py:12-149).

A ``KernelFunction`` is ``fn(particles, scalars, fielddata) -> None``.  A
``ParticleKernel`` wraps one or more of them with the particle fields (and
dtypes), scalars and simulation fields they need; ``merge_*`` combine the
declarations of several kernels and raise ``TypeError`` on conflicting
dtypes.  The declarations of the built-in kernels are identical to the
reference's, so merged particle layouts are the same.
"""

from __future__ import annotations

import types
from typing import Callable, Iterable, Mapping

import numpy as np

KernelFunction = Callable[..., None]

DEFAULT_PARTICLE_FIELDS: dict = {
    "status": np.dtype(np.uint8),
    "zidx": np.dtype(np.float64),
    "yidx": np.dtype(np.float64),
    "xidx": np.dtype(np.float64),
}


class ParticleKernel:
    """A kernel to be executed on particles."""

    def __init__(self, fn, particle_fields: Mapping, scalars: Mapping, simulation_fields: Iterable[str]):
        funcs = (fn,) if callable(fn) else tuple(fn)
        if not all(callable(f) for f in funcs):
            raise TypeError("All kernel functions must be callable")
        self._funcs = funcs
        self._particle_fields = {k: np.dtype(v) for k, v in particle_fields.items()}
        self._scalars = {k: np.dtype(v) for k, v in scalars.items()}
        self._simulation_fields = frozenset(simulation_fields)

    @property
    def particle_fields(self) -> Mapping:
        return types.MappingProxyType(self._particle_fields)

    @property
    def scalars(self) -> Mapping:
        return types.MappingProxyType(self._scalars)

    @property
    def simulation_fields(self) -> frozenset:
        return self._simulation_fields

    @property
    def functions(self) -> tuple:
        return self._funcs

    @staticmethod
    def func_name(fn) -> str:
        return getattr(fn, "__qualname__", getattr(fn, "__name__", repr(fn)))

    def __repr__(self) -> str:
        funcs = ", ".join(self.func_name(fn) for fn in self._funcs)
        return (f"{self.__class__.__name__}(funcs=[{funcs}], particle_fields={list(self._particle_fields)}, "
                f"scalars={list(self._scalars)}, simulation_fields={sorted(self._simulation_fields)})")

    def __str__(self) -> str:
        return "Particle Kernel: " + " → ".join(self.func_name(fn) for fn in self._funcs)

    def __call__(self, particles, scalars, fielddata) -> None:
        for fn in self._funcs:
            fn(particles, scalars, fielddata)

    @classmethod
    def chain(cls, *kernels: "ParticleKernel") -> "ParticleKernel":
        funcs = tuple(fn for k in kernels for fn in k._funcs)
        return cls(funcs, merge_particle_fields(kernels), merge_scalars(kernels), merge_simulation_fields(kernels))

    def chain_with(self, *others: "ParticleKernel") -> "ParticleKernel":
        return self.__class__.chain(self, *others)


def merge_particle_fields(kernels: Iterable[ParticleKernel]) -> dict:
    merged = DEFAULT_PARTICLE_FIELDS.copy()
    for kernel in kernels:
        for name, dtype in kernel.particle_fields.items():
            if name in merged:
                if merged[name] != dtype:
                    raise TypeError(f"Conflicting dtypes for particle field '{name}': {merged[name]} vs {dtype}")
            else:
                merged[name] = dtype
    return merged


def merge_scalars(kernels: Iterable[ParticleKernel]) -> dict:
    merged: dict = {}
    for kernel in kernels:
        for name, dtype in kernel.scalars.items():
            if name in merged:
                if merged[name] != dtype:
                    raise TypeError(f"Conflicting dtypes for scalar '{name}': {merged[name]} vs {dtype}")
            else:
                merged[name] = dtype
    return merged


def merge_simulation_fields(kernels: Iterable[ParticleKernel]) -> frozenset:
    merged: set = set()
    for kernel in kernels:
        merged.update(kernel.simulation_fields)
    return frozenset(merged)
