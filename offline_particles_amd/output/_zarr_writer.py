"""Trajectory output to Zarr v3 stores (mirrors output/_zarr_writer.py:16-223).

Same builder/writer API, events and checks as the reference's
ZarrOutputBuilder / ZarrOutputWriter: a ``time`` array of shape (T,) chunked
(1,), and one (T, nparticles) array per output chunked (1, min(chunksize,
nparticles)) with dimension names (time, particle), each grown by one row per
write round.  The store is a local directory (the reference's ``store``
accepts any zarr StoreLike; a path is the common one) written in the layout
zarr.create_array gives (``_zarr.py``), without the zarr package.

A row leaves the device through pinned host memory.  With
``async_writes=True`` the copy and the chunk files of a round are written by
host threads while the simulation keeps stepping; a round's rows become
visible (the arrays' shape grows) once all its chunks are on disk, rounds in
order, and ``flush()`` waits for every pending round.  By default each round
is on disk when ``finalise_write_round`` returns, as with the reference.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import types
from typing import Any, Mapping

import numpy as np
import torch

from ._output import AbstractOutputWriter, AbstractOutputWriterBuilder, Output
from .._zarr import ZarrArray

DEFAULT_CHUNKSIZE = 10_000

# array_kwargs understood here (zarr.create_array keywords with a local meaning)
_ARRAY_KWARGS = ("fill_value", "compressors", "attributes")


def _create_kwargs(kw: dict) -> dict:
    bad = [k for k in kw if k not in _ARRAY_KWARGS]
    if bad:
        raise TypeError(f"unsupported array keyword(s) {bad}; supported: {list(_ARRAY_KWARGS)}")
    out = {}
    if "fill_value" in kw:
        out["fill_value"] = kw["fill_value"]
    if "attributes" in kw:
        out["attributes"] = kw["attributes"]
    # zarr.create_array's default compressors="auto" gives numeric v3 arrays
    # the zstd codec (level 0), so the default here is zstd as well
    c = kw.get("compressors", "auto")
    if c is None or (isinstance(c, (list, tuple)) and not c):
        out["compressor"] = None
    elif c in ("auto", "zstd") or (isinstance(c, (list, tuple)) and list(c) == ["zstd"]):
        out["compressor"] = "zstd"
    else:
        raise ValueError(f"compressors must be None, 'auto' or 'zstd', got {c!r}")
    return out


class ZarrOutputWriter(AbstractOutputWriter):
    """Writes each round's time and outputs as new rows of Zarr v3 arrays."""

    def __init__(self, name: str, store, time_array: ZarrArray, outputs: dict, *, async_writes: bool = False,
                 workers: int = 8) -> None:
        self._name = name
        self._store = store
        self._time_array = time_array
        self._outputs = types.MappingProxyType(dict(outputs))
        self._output_count = 0
        self._time_count = 0
        self._row_counts = {k: 0 for k in outputs}
        self._async = bool(async_writes)
        self._pool = cf.ThreadPoolExecutor(max_workers=max(1, int(workers)))
        self._committer = cf.ThreadPoolExecutor(max_workers=1)
        self._round: list = []
        self._pending: list = []

    @property
    def name(self) -> str:
        return self._name

    @property
    def store(self):
        return self._store

    @property
    def outputs(self) -> Mapping[str, Output]:
        return types.MappingProxyType({key: output for key, (output, _) in self._outputs.items()})

    def write_time(self, state) -> None:
        t = self._time_count
        value = np.asarray([state.time], dtype=self._time_array.dtype)
        self._round.append(self._pool.submit(self._time_array.write_row, t, value))
        self._time_count += 1

    def write_output(self, name: str, state) -> None:
        if name not in self._outputs:
            raise KeyError(f"Output variable '{name}' not found.")
        output, array = self._outputs[name]
        src = state.particles[output.particle_field]
        t = self._row_counts[name]
        if torch.is_tensor(src) and src.is_cuda:
            row = torch.empty(src.shape, dtype=src.dtype, pin_memory=True)
            row.copy_(src, non_blocking=True)
            ready = torch.cuda.Event()
            ready.record()
        else:
            row = src.detach().clone() if torch.is_tensor(src) else np.array(src)
            ready = None
        # the particle axis in column blocks of whole chunks, written in parallel
        n = array.shape[1]
        c = array.chunks[1]
        step = max(c, -(-n // (4 * c)) * c)
        for j in range(0, max(n, 1), step):
            self._round.append(self._pool.submit(self._write_block, array, t, j, min(n, j + step), row, ready))
        self._row_counts[name] += 1

    @staticmethod
    def _write_block(array: ZarrArray, t: int, j0: int, j1: int, row, ready) -> None:
        if ready is not None:
            ready.synchronize()
        values = row.numpy() if torch.is_tensor(row) else row
        values = np.asarray(values, dtype=array.dtype)
        c = array.chunks[1]
        for j in range(j0, j1, c):
            array.write_chunk((t, j // c), values[j:min(j + c, j1)].reshape(1, -1))

    def finalise_write_round(self, state) -> None:
        """Confirm that every output was written this round, then commit the round (zarr_writer.py:80-97)."""
        expected = self._output_count + 1
        problem = None
        if self._time_count != expected:
            problem = f"Time output has {self._time_count} entries, expected {expected}."
        for name, count in self._row_counts.items():
            if problem is None and count != expected:
                problem = f"Output '{name}' has {count} time entries, expected {expected}."
        if problem is not None:
            # the chunk writes already submitted this round finish before the
            # error reaches the caller, so no host thread still writes the store
            cf.wait(self._round)
            raise RuntimeError(problem)
        futures, self._round = self._round, []
        arrays = [self._time_array] + [a for _, a in self._outputs.values()]
        commit = self._committer.submit(self._commit, futures, arrays, expected)
        self._output_count += 1
        if self._async:
            self._pending.append(commit)
            done = [f for f in self._pending if f.done()]
            self._pending = [f for f in self._pending if not f.done()]
            for f in done:
                f.result()  # a failed earlier round raises here
        else:
            commit.result()

    @staticmethod
    def _commit(futures, arrays, rows: int) -> None:
        for f in futures:
            f.result()
        for a in arrays:
            a.resize((rows,) + tuple(a.shape[1:]))

    def flush(self) -> None:
        """Wait until every committed round is on disk (raises a failed write)."""
        pending, self._pending = self._pending, []
        for f in pending:
            f.result()

    def close(self) -> None:
        self.flush()
        self._pool.shutdown(wait=True)
        self._committer.shutdown(wait=True)


class ZarrOutputBuilder(AbstractOutputWriterBuilder):
    """Builder for Zarr output (zarr_writer.py:100-223)."""

    def __init__(self, name: str, store, *, chunksize: int = DEFAULT_CHUNKSIZE,
                 particle_dimension_name: str = "particle", time_name: str = "time", overwrite: bool = False,
                 array_kwargs: dict[str, Any] | None = None, time_array_kwargs: dict[str, Any] | None = None,
                 async_writes: bool = False) -> None:
        self._name = name
        self._store = os.fspath(store)
        self._outputs: dict = {}
        self._chunksize = int(chunksize)
        if self._chunksize < 1:
            raise ValueError("chunksize must be >= 1")
        self._particle_dimension_name = particle_dimension_name
        self._time_name = time_name
        self._overwrite = overwrite
        self._array_kwargs = dict(array_kwargs or {})
        _create_kwargs(self._array_kwargs)
        self._time_array_kwargs = self._array_kwargs.copy()
        if time_array_kwargs is not None:
            self._time_array_kwargs.update(time_array_kwargs)
        _create_kwargs(self._time_array_kwargs)
        self._async = async_writes

    @property
    def name(self) -> str:
        return self._name

    @property
    def outputs(self) -> Mapping[str, Output]:
        return types.MappingProxyType({key: output for key, (output, _) in self._outputs.items()})

    def add_output(self, *outputs: Output, **kwargs) -> None:
        array_kwargs = self._array_kwargs.copy()
        array_kwargs.update(kwargs)
        _create_kwargs(array_kwargs)
        for output in outputs:
            if output.name in self._outputs:
                raise KeyError(f"Output variable with name '{output.name}' already exists.")
            self._outputs[output.name] = (output, array_kwargs)

    def remove_output(self, name: str) -> None:
        if name not in self._outputs:
            raise KeyError(f"Output variable '{name}' does not exist.")
        del self._outputs[name]

    def build(self, nparticles: int, time_type=np.dtype(np.float64)) -> ZarrOutputWriter:
        time_array = ZarrArray.create(self._store, self._time_name, shape=(0,), dtype=time_type, chunks=(1,),
                                      dimension_names=(self._time_name,), overwrite=self._overwrite,
                                      **_create_kwargs(self._time_array_kwargs))
        outputs = {
            name: (output, ZarrArray.create(self._store, name, shape=(0, int(nparticles)), dtype=output.dtype,
                                            chunks=(1, max(1, min(self._chunksize, int(nparticles)))),
                                            dimension_names=(self._time_name, self._particle_dimension_name),
                                            overwrite=self._overwrite, **_create_kwargs(kw)))
            for name, (output, kw) in self._outputs.items()
        }
        return ZarrOutputWriter(self._name, self._store, time_array, outputs, async_writes=self._async)
