"""Zarr v3 arrays on a local directory store, without the zarr package.

The reference writes its trajectories with zarr-python 3 (uv.lock: zarr
3.1.5), whose ``zarr.create_array`` lays an array out as (Zarr v3 core
specification):

- ``<root>/zarr.json``: group metadata ``{"zarr_format": 3, "node_type": "group"}``;
- ``<root>/<name>/zarr.json``: array metadata (shape, data_type, regular chunk
  grid, ``default`` chunk key encoding with separator "/", fill_value, codecs,
  dimension_names);
- ``<root>/<name>/c/<i>/<j>``: one file per chunk, the full chunk shape in C
  order (edge chunks padded with the fill value), passed through the codecs:
  ``bytes`` (little endian) then optionally ``zstd`` (zarr-python's default
  compressor; available here through pyarrow).

This module writes and reads exactly that layout, so a store written here
opens with zarr-python and vice versa for these codecs.
"""

from __future__ import annotations

import json
import os
import shutil
from typing import Sequence

import numpy as np

_V3_TYPES = {
    np.dtype(np.bool_): "bool", np.dtype(np.int8): "int8", np.dtype(np.int16): "int16",
    np.dtype(np.int32): "int32", np.dtype(np.int64): "int64", np.dtype(np.uint8): "uint8",
    np.dtype(np.uint16): "uint16", np.dtype(np.uint32): "uint32", np.dtype(np.uint64): "uint64",
    np.dtype(np.float16): "float16", np.dtype(np.float32): "float32", np.dtype(np.float64): "float64",
}
_FROM_V3 = {v: k for k, v in _V3_TYPES.items()}


def _zstd():
    import pyarrow

    if not pyarrow.Codec.is_available("zstd"):  # pragma: no cover
        raise RuntimeError("zstd is not available (pyarrow without zstd)")
    return pyarrow.Codec("zstd")


def _write_json(path: str, obj) -> None:
    """Replace ``path`` with the JSON of ``obj`` atomically (a reader never sees half a file)."""
    tmp = f"{path}.tmp{os.getpid()}"
    with open(tmp, "w") as f:
        json.dump(obj, f, indent=2)
    os.replace(tmp, path)


def _fill_json(dtype: np.dtype, value):
    if dtype.kind == "f":
        v = float(value)
        if np.isnan(v):
            return "NaN"
        if np.isinf(v):
            return "Infinity" if v > 0 else "-Infinity"
        return v
    if dtype.kind == "b":
        return bool(value)
    return int(value)


def _fill_value(dtype: np.dtype, js):
    if isinstance(js, str):
        return dtype.type({"NaN": np.nan, "Infinity": np.inf, "-Infinity": -np.inf}[js])
    return dtype.type(js)


def ensure_group(root: str) -> None:
    """Create the store directory and its root group metadata if missing."""
    os.makedirs(root, exist_ok=True)
    meta = os.path.join(root, "zarr.json")
    if not os.path.exists(meta):
        _write_json(meta, {"zarr_format": 3, "node_type": "group", "attributes": {}})


class ZarrArray:
    """One Zarr v3 array of a local directory store (read, append along axis 0)."""

    def __init__(self, root: str, name: str) -> None:
        self.path = os.path.join(root, name)
        with open(os.path.join(self.path, "zarr.json")) as f:
            meta = json.load(f)
        if meta.get("zarr_format") != 3 or meta.get("node_type") != "array":
            raise ValueError(f"{self.path} is not a Zarr v3 array")
        grid = meta["chunk_grid"]
        if grid["name"] != "regular":
            raise ValueError(f"unsupported chunk grid {grid['name']!r}")
        enc = meta["chunk_key_encoding"]
        self._sep = enc.get("configuration", {}).get("separator", "/" if enc["name"] == "default" else ".")
        self._prefix = "c" + self._sep if enc["name"] == "default" else ""
        self.meta = meta
        self.shape = tuple(meta["shape"])
        self.dtype = _FROM_V3[meta["data_type"]]
        self.chunks = tuple(grid["configuration"]["chunk_shape"])
        self.fill_value = _fill_value(self.dtype, meta["fill_value"])
        names = [c["name"] for c in meta["codecs"]]
        if not names or names[0] != "bytes" or any(n not in ("bytes", "zstd") for n in names):
            raise ValueError(f"unsupported codecs {names}")
        endian = meta["codecs"][0].get("configuration", {}).get("endian", "little")
        self._dtype_disk = self.dtype.newbyteorder("<" if endian == "little" else ">")
        self._zstd = "zstd" in names

    @classmethod
    def create(cls, root: str, name: str, *, shape: Sequence[int], dtype, chunks: Sequence[int],
               dimension_names: Sequence[str] | None = None, fill_value=0, compressor: str | None = None,
               attributes: dict | None = None, overwrite: bool = False) -> "ZarrArray":
        """Create an empty array (zarr.create_array's layout)."""
        dtype = np.dtype(dtype)
        if dtype not in _V3_TYPES:
            raise TypeError(f"unsupported dtype {dtype}")
        if compressor not in (None, "zstd"):
            raise ValueError(f"unsupported compressor {compressor!r} (None or 'zstd')")
        ensure_group(root)
        path = os.path.join(root, name)
        if os.path.exists(path):
            if not overwrite:
                raise FileExistsError(f"{path} exists (overwrite=False)")
            shutil.rmtree(path)
        os.makedirs(path)
        codecs = [{"name": "bytes", "configuration": {"endian": "little"}}]
        if compressor == "zstd":
            codecs.append({"name": "zstd", "configuration": {"level": 0, "checksum": False}})
        meta = {
            "zarr_format": 3,
            "node_type": "array",
            "shape": [int(s) for s in shape],
            "data_type": _V3_TYPES[dtype],
            "chunk_grid": {"name": "regular", "configuration": {"chunk_shape": [int(c) for c in chunks]}},
            "chunk_key_encoding": {"name": "default", "configuration": {"separator": "/"}},
            "fill_value": _fill_json(dtype, fill_value),
            "codecs": codecs,
            "attributes": dict(attributes or {}),
        }
        if dimension_names is not None:
            meta["dimension_names"] = list(dimension_names)
        _write_json(os.path.join(path, "zarr.json"), meta)
        return cls(root, name)

    # ---------------------------------------------------------------- chunks

    @property
    def ndim(self) -> int:
        return len(self.shape)

    def chunk_path(self, idx: Sequence[int]) -> str:
        return os.path.join(self.path, self._prefix + self._sep.join(str(int(i)) for i in idx))

    def encode_chunk(self, block: np.ndarray) -> bytes:
        """A full chunk (shape == chunks) through the codecs."""
        raw = np.ascontiguousarray(block, dtype=self._dtype_disk).tobytes()
        return _zstd().compress(raw, asbytes=True) if self._zstd else raw

    def decode_chunk(self, data: bytes) -> np.ndarray:
        n = int(np.prod(self.chunks)) * self.dtype.itemsize
        if self._zstd:
            data = _zstd().decompress(data, decompressed_size=n, asbytes=True)
        if len(data) != n:
            raise ValueError(f"chunk of {len(data)} bytes, expected {n}")
        return np.frombuffer(data, dtype=self._dtype_disk).astype(self.dtype, copy=False).reshape(self.chunks)

    def read_chunk_into(self, idx: Sequence[int], out: np.ndarray) -> bool:
        """Read an uncompressed, native-order chunk file straight into ``out`` (C-contiguous,
        the chunk's shape and dtype); False if the chunk needs decoding (then nothing is read)."""
        if self._zstd or self._dtype_disk != self.dtype or not out.flags.c_contiguous:
            return False
        path = self.chunk_path(idx)
        if not os.path.exists(path):
            out[...] = self.fill_value
            return True
        buf = memoryview(out.reshape(-1).view(np.uint8))
        with open(path, "rb", buffering=0) as f:
            got = f.readinto(buf)
        if got != out.nbytes:
            raise ValueError(f"chunk {path}: {got} bytes, expected {out.nbytes}")
        return True

    def write_chunk(self, idx: Sequence[int], block: np.ndarray) -> None:
        """Write chunk ``idx`` from ``block`` (its part inside the array; the rest is fill)."""
        if block.ndim != len(self.chunks) or any(b > c for b, c in zip(block.shape, self.chunks)):
            raise ValueError(f"block of shape {block.shape} does not fit chunks {self.chunks}")
        full = np.full(self.chunks, self.fill_value, dtype=self.dtype)
        full[tuple(slice(0, s) for s in block.shape)] = block
        path = self.chunk_path(idx)
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = f"{path}.tmp{os.getpid()}"
        with open(tmp, "wb") as f:
            f.write(self.encode_chunk(full))
        os.replace(tmp, path)

    def read_chunk(self, idx: Sequence[int]) -> np.ndarray:
        path = self.chunk_path(idx)
        if not os.path.exists(path):
            return np.full(self.chunks, self.fill_value, dtype=self.dtype)
        with open(path, "rb") as f:
            return self.decode_chunk(f.read())

    # ---------------------------------------------------------------- metadata

    def resize(self, shape: Sequence[int]) -> None:
        self.meta["shape"] = [int(s) for s in shape]
        self.shape = tuple(self.meta["shape"])
        _write_json(os.path.join(self.path, "zarr.json"), self.meta)

    # ---------------------------------------------------------------- whole rows / reads

    def write_row(self, t: int, values: np.ndarray) -> None:
        """Write row ``t`` along axis 0 (chunk extent 1 on axis 0) of a 1-D or 2-D array."""
        if self.chunks[0] != 1:
            raise ValueError("write_row needs chunks of extent 1 along axis 0")
        values = np.asarray(values)
        if len(self.shape) == 1:
            self.write_chunk((t,), values.reshape(1))
            return
        c = self.chunks[1]
        for j in range(0, self.shape[1], c):
            self.write_chunk((t, j // c), values[j:j + c].reshape(1, -1))

    def __getitem__(self, key) -> np.ndarray:
        """Basic indexing (integers and slices with step 1) through the chunks."""
        if not isinstance(key, tuple):
            key = (key,)
        key = key + (slice(None),) * (len(self.shape) - len(key))
        sl, squeeze = [], []
        for k, n in zip(key, self.shape):
            if isinstance(k, (int, np.integer)):
                k = int(k) + (n if k < 0 else 0)
                if not 0 <= k < n:
                    raise IndexError(k)
                sl.append((k, k + 1))
                squeeze.append(True)
            else:
                start, stop, step = k.indices(n)
                if step != 1:
                    raise IndexError("slices with a step are not supported")
                sl.append((start, max(start, stop)))
                squeeze.append(False)
        out = np.empty([b - a for a, b in sl], dtype=self.dtype)
        ranges = [range(a // c, -(-b // c)) if b > a else range(0) for (a, b), c in zip(sl, self.chunks)]
        for idx in np.ndindex(*[len(r) for r in ranges]):
            ci = [r[i] for r, i in zip(ranges, idx)]
            chunk = self.read_chunk(ci)
            src, dst = [], []
            for (a, b), c, k in zip(sl, self.chunks, ci):
                lo, hi = max(a, k * c), min(b, (k + 1) * c)
                src.append(slice(lo - k * c, hi - k * c))
                dst.append(slice(lo - a, hi - a))
            out[tuple(dst)] = chunk[tuple(src)]
        return out.reshape([s for s, q in zip(out.shape, squeeze) if not q])


def open_array(root: str, name: str) -> ZarrArray:
    return ZarrArray(os.fspath(root), name)
