"""An output writer that keeps (time, particle) arrays in host memory.

Same events and checks as the reference's ZarrOutputWriter
(output/_zarr_writer.py:16-98): every round appends one time value and one
row per output, and ``finalise_write_round`` checks that every array has
grown by exactly one row.  Rows are copied device -> pinned host memory
asynchronously on the current stream; ``arrays()`` waits for the copies.
"""

from __future__ import annotations

import types
from typing import Mapping

import numpy as np
import torch

from ._output import AbstractOutputWriter, AbstractOutputWriterBuilder, Output


class MemoryOutputWriter(AbstractOutputWriter):
    def __init__(self, name: str, nparticles: int, time_type, outputs: dict) -> None:
        self._name = name
        self._nparticles = int(nparticles)
        self._time_type = np.dtype(time_type)
        self._outputs = types.MappingProxyType(dict(outputs))
        self._times: list = []
        self._rows: dict = {k: [] for k in outputs}
        self._count = 0
        self._device_copies = False

    @property
    def name(self) -> str:
        return self._name

    @property
    def outputs(self) -> Mapping[str, Output]:
        return self._outputs

    def write_time(self, state) -> None:
        self._times.append(state.time)

    def write_output(self, name: str, state) -> None:
        if name not in self._outputs:
            raise KeyError(f"Output variable '{name}' not found.")
        src = state.particles[self._outputs[name].particle_field]
        row = torch.empty(src.shape, dtype=src.dtype, pin_memory=src.is_cuda)
        row.copy_(src, non_blocking=src.is_cuda)
        self._device_copies |= src.is_cuda
        self._rows[name].append(row)

    def finalise_write_round(self, state) -> None:
        expected = self._count + 1
        if len(self._times) != expected:
            raise RuntimeError(f"Time output has {len(self._times)} entries, expected {expected}.")
        for name, rows in self._rows.items():
            if len(rows) != expected:
                raise RuntimeError(f"Output '{name}' has {len(rows)} time entries, expected {expected}.")
        self._count += 1

    def arrays(self) -> dict:
        """{"time": (T,), name: (T, nparticles)} numpy arrays."""
        if self._device_copies:
            torch.cuda.synchronize()
        out = {"time": np.array(self._times, dtype=self._time_type)}
        for name, rows in self._rows.items():
            dt = self._outputs[name].dtype
            out[name] = (np.stack([r.numpy() for r in rows]).astype(dt, copy=False) if rows
                         else np.zeros((0, self._nparticles), dt))
        return out


class MemoryOutputBuilder(AbstractOutputWriterBuilder):
    def __init__(self, name: str) -> None:
        self._name = name
        self._outputs: dict = {}

    @property
    def name(self) -> str:
        return self._name

    @property
    def outputs(self) -> Mapping[str, Output]:
        return types.MappingProxyType(dict(self._outputs))

    def add_output(self, *outputs: Output, **kwargs) -> None:
        for output in outputs:
            if output.name in self._outputs:
                raise KeyError(f"Output variable with name '{output.name}' already exists.")
            self._outputs[output.name] = output

    def remove_output(self, name: str) -> None:
        if name not in self._outputs:
            raise KeyError(f"Output variable '{name}' does not exist.")
        del self._outputs[name]

    def build(self, nparticles: int, time_type=np.dtype(np.float64)) -> MemoryOutputWriter:
        return MemoryOutputWriter(self._name, nparticles, time_type, self._outputs)
