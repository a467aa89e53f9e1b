"""Output declarations and writer interfaces (mirrors output/_output.py:14-171).

An ``Output`` names a particle field that one or more kernels fill; a writer
turns outputs into events (time, one per output, finalise) that the
simulation runs on its schedule, the output kernels launching before each
output event.
"""

from __future__ import annotations

import abc
import dataclasses
import functools
from typing import Mapping

import numpy as np

from ..kernels import ParticleKernel, merge_particle_fields
from ..simulation import Event, SimulationState


@dataclasses.dataclass(frozen=True)
class Output:
    """Class defining a single output."""

    name: str
    particle_field: str
    dtype: np.dtype = dataclasses.field(init=False)
    kernels: tuple

    def __init__(self, name: str, *kernels: ParticleKernel, particle_field: str | None = None) -> None:
        if particle_field is None:
            particle_field = name
        kernel_fields = merge_particle_fields(kernels)
        if particle_field not in kernel_fields:
            raise ValueError(f"Particle field '{particle_field}' not found in provided kernels.")
        object.__setattr__(self, "name", name)
        object.__setattr__(self, "particle_field", particle_field)
        object.__setattr__(self, "dtype", np.dtype(kernel_fields[particle_field]))
        object.__setattr__(self, "kernels", kernels)


class AbstractOutputWriter(abc.ABC):
    """Interface for output writers."""

    @property
    @abc.abstractmethod
    def name(self) -> str: ...

    @property
    @abc.abstractmethod
    def outputs(self) -> Mapping[str, Output]: ...

    @abc.abstractmethod
    def write_time(self, state: SimulationState) -> None: ...

    @abc.abstractmethod
    def write_output(self, name: str, state: SimulationState) -> None: ...

    @abc.abstractmethod
    def finalise_write_round(self, state: SimulationState) -> None: ...

    def event_name(self, output_name: str) -> str:
        return f"{self.name}:{output_name}"

    def create_events(self) -> list:
        """Events: write time, one per output (its kernels launch first), finalise."""
        events = [Event(self.event_name("time"), self.write_time)]
        for name, output in self.outputs.items():
            events.append(Event(self.event_name(name), functools.partial(self.write_output, name), *output.kernels))
        events.append(Event(self.event_name("finalise"), self.finalise_write_round))
        return events


class AbstractOutputWriterBuilder(abc.ABC):
    """Abstract base class for output writer builders."""

    @property
    @abc.abstractmethod
    def name(self) -> str: ...

    @property
    @abc.abstractmethod
    def outputs(self) -> Mapping[str, Output]: ...

    @abc.abstractmethod
    def add_output(self, *outputs: Output, **kwargs) -> None: ...

    @abc.abstractmethod
    def remove_output(self, name: str) -> None: ...

    @abc.abstractmethod
    def build(self, nparticles: int, time_type: np.dtype) -> AbstractOutputWriter: ...
