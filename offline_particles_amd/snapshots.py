"""Snapshot sources read on demand: the on-disk side of the time window.

The reference reads large snapshot sets lazily through dask arrays
(``TimeDependentField.from_dask``, fields.py:443-458; ``ChunkedDaskArray``,
spatial_arrays.py:228-292, which loads the chunks a launch needs).  Here a
``SnapshotSource`` hands out whole snapshots as host arrays, decoded on host
threads, and prefetches the next one; ``ZarrSnapshots`` reads them from a
Zarr v3 array of shape (T, ...) (``_zarr``, no zarr package).

Consumers:
- ``TimeDependentField(source, ...)`` keeps the two or three snapshots in use
  on the device (the drop-in path);
- ``AB3Engine(..., stream_slots=k)`` streams them through its ring of device
  slots: host decode on a thread, pinned staging, copy on a side stream.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import threading

import numpy as np

from ._zarr import ZarrArray


class SnapshotSource:
    """Base of lazily read (T, ...) snapshot sets.  Subclasses implement ``_read``."""

    def __init__(self, shape, dtype, *, workers: int = 4) -> None:
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self._pool = cf.ThreadPoolExecutor(max_workers=max(1, int(workers)))
        self._lock = threading.Lock()
        self._pending: dict = {}
        self.reads = 0

    @property
    def ndim(self) -> int:
        return len(self.shape)

    def __len__(self) -> int:
        return self.shape[0]

    def _read(self, i: int) -> np.ndarray:
        out = np.empty(self.shape[1:], dtype=self.dtype)
        self._read_into(i, out)
        return out

    def _read_into(self, i: int, out: np.ndarray) -> None:  # pragma: no cover - abstract
        raise NotImplementedError

    def submit(self, fn, *args) -> cf.Future:
        """Run fn(*args) on the source's host threads."""
        return self._pool.submit(fn, *args)

    def read_into(self, i: int, out: np.ndarray) -> None:
        """Snapshot i (mod T) into ``out`` (e.g. pinned staging memory), on the calling thread."""
        self._read_into(i % len(self), out)
        self.reads += 1

    def prefetch(self, i: int) -> None:
        """Start decoding snapshot i (mod T) on a host thread if it is not already."""
        i %= len(self)
        with self._lock:
            if i not in self._pending:
                self._pending[i] = self._pool.submit(self._read, i)

    def read(self, i: int) -> np.ndarray:
        """Snapshot i (mod T) as a host array (waits for a prefetch in flight)."""
        i %= len(self)
        with self._lock:
            fut = self._pending.pop(i, None)
        arr = fut.result() if fut is not None else self._read(i)
        self.reads += 1
        return arr

    def read_all(self) -> np.ndarray:
        return np.stack([self.read(i) for i in range(len(self))])


class ArraySnapshots(SnapshotSource):
    """Snapshots of any (T, ...) array with basic slicing (a dask array, a
    ``ZarrArray``, a memory map): snapshot i is ``data[i]``, computed when the
    slice has a ``compute()`` method (dask)."""

    def __init__(self, data, *, workers: int = 2) -> None:
        super().__init__(data.shape, data.dtype, workers=workers)
        if len(self.shape) < 2:
            raise ValueError("a snapshot array needs a time axis and a spatial one")
        self._data = data

    def _read_into(self, i: int, out: np.ndarray) -> None:
        block = self._data[i]
        if hasattr(block, "compute"):
            block = block.compute()
        out[...] = np.asarray(block)


class ZarrSnapshots(SnapshotSource):
    """Snapshots of a Zarr v3 array (T, ...) in a local directory store; each
    snapshot is decoded chunk by chunk on ``workers`` threads."""

    def __init__(self, store, name: str, *, workers: int = 8) -> None:
        self._array = ZarrArray(os.fspath(store), name)
        super().__init__(self._array.shape, self._array.dtype, workers=4)
        if len(self._array.shape) < 2:
            raise ValueError(f"{name}: a snapshot array needs a time axis and a spatial one")
        self._chunk_pool = cf.ThreadPoolExecutor(max_workers=max(1, int(workers)))

    def _read_into(self, i: int, out: np.ndarray) -> None:
        a = self._array
        c0 = a.chunks[0]
        if c0 == 1 and tuple(a.chunks[1:]) == tuple(a.shape[1:]) and out.dtype == a.dtype:
            # one chunk per snapshot: the file goes straight into `out`
            if a.read_chunk_into((i,) + (0,) * (len(a.shape) - 1), out.reshape((1,) + out.shape)):
                return
        counts = [-(-n // c) for n, c in zip(a.shape[1:], a.chunks[1:])]

        def one(idx):
            chunk = a.read_chunk((i // c0,) + idx)[i % c0]
            dst = tuple(slice(k * c, min((k + 1) * c, n)) for k, c, n in zip(idx, a.chunks[1:], a.shape[1:]))
            out[dst] = chunk[tuple(slice(0, s.stop - s.start) for s in dst)]

        list(self._chunk_pool.map(one, np.ndindex(*counts)))


def write_zarr_snapshots(store, name: str, data: np.ndarray, *, chunks=None, compressor: str | None = None,
                         dimension_names=None, overwrite: bool = False) -> ZarrArray:
    """Write a (T, ...) array of snapshots as a Zarr v3 array (one time step per chunk
    along axis 0 by default), e.g. to stage model output for ``ZarrSnapshots``."""
    data = np.asarray(data)
    chunks = tuple(chunks) if chunks is not None else (1,) + data.shape[1:]
    arr = ZarrArray.create(os.fspath(store), name, shape=data.shape, dtype=data.dtype, chunks=chunks,
                           dimension_names=dimension_names, compressor=compressor, overwrite=overwrite)
    for idx in np.ndindex(*[-(-n // c) for n, c in zip(data.shape, chunks)]):
        sl = tuple(slice(k * c, min((k + 1) * c, n)) for k, c, n in zip(idx, chunks, data.shape))
        arr.write_chunk(idx, data[sl])
    return arr
