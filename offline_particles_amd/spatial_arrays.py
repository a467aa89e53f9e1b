"""Spatial arrays resident in HBM.

Mirrors the reference's spatial_arrays.py: ``Stagger`` (offsets and staggered
sizes of the ROMS C-grid, spatial_arrays.py:18-68), ``BBox``
(spatial_arrays.py:15) and ``NumpyArray`` (spatial_arrays.py:181-225), which
holds the whole array and returns it unchanged for any bounding box.  Here the
whole array lives in device memory (``DeviceArray``).  ``ChunkedArray`` is the
counterpart of the reference's ``ChunkedDaskArray`` (spatial_arrays.py:228-350):
it reads from a chunked host array (a dask array, a Zarr array of this
package, a numpy array with a chunk layout) only the chunks that cover the
particles' bounding box, into device memory.
"""

from __future__ import annotations

import collections
import enum

import numpy as np
import torch

BBox = collections.namedtuple("BBox", ("zmin", "zmax", "ymin", "ymax", "xmin", "xmax"))


class Stagger(str, enum.Enum):
    """Staggering of one dimension (spatial_arrays.py:18-68)."""

    CENTER = "center"
    LEFT = "left"
    RIGHT = "right"
    INNER = "inner"
    OUTER = "outer"
    INVARIANT = "invariant"

    def __str__(self) -> str:
        return self.value

    @property
    def offset(self) -> float | None:
        """Offset between centred and staggered indices."""
        if self is Stagger.CENTER:
            return 0.0
        if self in (Stagger.LEFT, Stagger.OUTER):
            return 0.5
        if self in (Stagger.RIGHT, Stagger.INNER):
            return -0.5
        return None

    def expected_size(self, N: int) -> int | None:
        """Size of this dimension given the centred size N."""
        if self in (Stagger.CENTER, Stagger.LEFT, Stagger.RIGHT):
            return N
        if self is Stagger.OUTER:
            return N + 1
        if self is Stagger.INNER:
            return N - 1
        return None

    @property
    def is_invariant(self) -> bool:
        return self is Stagger.INVARIANT

    @property
    def is_active(self) -> bool:
        return self is not Stagger.INVARIANT

    @property
    def on_face(self) -> bool:
        return self in (Stagger.LEFT, Stagger.RIGHT, Stagger.INNER, Stagger.OUTER)

    @property
    def at_center(self) -> bool:
        return self is Stagger.CENTER


_TORCH_DTYPES = {
    np.dtype(np.float64): torch.float64,
    np.dtype(np.float32): torch.float32,
    np.dtype(np.uint8): torch.uint8,
    np.dtype(np.int32): torch.int32,
    np.dtype(np.int64): torch.int64,
    np.dtype(np.uint16): torch.uint16 if hasattr(torch, "uint16") else torch.int16,
}


def to_torch_dtype(dtype) -> torch.dtype:
    if isinstance(dtype, torch.dtype):
        return dtype
    return _TORCH_DTYPES[np.dtype(dtype)]


def default_device() -> torch.device:
    return torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() else torch.device("cpu")


def as_device_tensor(data, device=None, dtype=None) -> torch.Tensor:
    """Copy host data (numpy / torch) into a contiguous device tensor."""
    device = torch.device(device) if device is not None else default_device()
    if isinstance(data, torch.Tensor):
        t = data
    else:
        arr = np.ascontiguousarray(data)
        t = torch.from_numpy(arr)
    if dtype is not None:
        t = t.to(to_torch_dtype(dtype))
    return t.to(device).contiguous()


class SpatialArray:
    """Base of spatial arrays: staggers and offsets (spatial_arrays.py:71-178)."""

    def __init__(self, z_stagger: Stagger, y_stagger: Stagger, x_stagger: Stagger) -> None:
        self._z_stagger = Stagger(z_stagger)
        self._y_stagger = Stagger(y_stagger)
        self._x_stagger = Stagger(x_stagger)

    @property
    def z_stagger(self) -> Stagger:
        return self._z_stagger

    @property
    def y_stagger(self) -> Stagger:
        return self._y_stagger

    @property
    def x_stagger(self) -> Stagger:
        return self._x_stagger

    @property
    def stagger(self) -> tuple:
        return (self._z_stagger, self._y_stagger, self._x_stagger)

    @property
    def offsets(self) -> tuple:
        return tuple(s.offset for s in self.stagger)

    @property
    def dmask(self) -> tuple:
        return tuple(s.is_active for s in self.stagger)

    @property
    def active_offsets(self) -> tuple:
        return tuple(s.offset for s in self.stagger if s.is_active)


class DeviceArray(SpatialArray):
    """A whole spatial array resident in device memory (the reference's NumpyArray)."""

    def __init__(self, data, z_stagger, y_stagger, x_stagger, device=None) -> None:
        super().__init__(z_stagger, y_stagger, x_stagger)
        self._data = as_device_tensor(data, device)

    @property
    def data(self) -> torch.Tensor:
        return self._data

    @property
    def dtype(self):
        return self._data.dtype

    @property
    def shape(self) -> tuple:
        return tuple(self._data.shape)

    def get_data_subset(self, bounding_box: BBox):
        """Whole array and unchanged offsets for any bounding box (spatial_arrays.py:204-225)."""
        return self._data, self.active_offsets


# The reference's name for the in-memory array type.
NumpyArray = DeviceArray


def chunk_bounds(chunks, shape) -> tuple:
    """Chunk boundaries 0, c0, c0 + c1, ..., n of every dimension
    (spatial_arrays.py:243): ``chunks`` per dimension is either the sizes of
    its chunks (dask) or one regular chunk size (Zarr)."""
    out = []
    for c, n in zip(chunks, shape):
        if isinstance(c, (int, np.integer)):
            b = np.append(np.arange(0, int(n), max(1, int(c))), int(n))
        else:
            b = np.concatenate([[0], np.cumsum(np.asarray(c, dtype=np.int64))])
            if b[-1] != n:
                raise ValueError(f"chunks {tuple(c)} do not cover a dimension of size {n}")
        out.append(b.astype(np.int64))
    return tuple(out)


def compute_new_lower_bound(dim_min: float, offset: float, bounds) -> int:
    """Start of the chunk holding dim_min + offset, clamped to 0
    (spatial_arrays.py:308-327)."""
    idx = int(np.searchsorted(bounds, dim_min + offset, side="right")) - 1
    return int(bounds[max(0, idx)])


def compute_new_upper_bound(dim_max: float, offset: float, bounds) -> int:
    """End of the chunk holding dim_max + offset + 1 (the exclusive upper
    index), clamped to the array's size (spatial_arrays.py:330-350)."""
    idx = int(np.searchsorted(bounds, dim_max + offset + 1, side="left"))
    return int(bounds[min(len(bounds) - 1, idx)])


class ChunkedArray(SpatialArray):
    """A spatial array read chunk by chunk from a chunked host array
    (the reference's ChunkedDaskArray, spatial_arrays.py:228-295).

    ``data`` needs ``shape``, ``dtype``, ``chunks`` (per dimension, chunk sizes
    or one regular size) and basic slicing; a slice with a ``compute()`` method
    (dask) is computed.  With ``index`` the array is ``data[index]`` (e.g. one
    snapshot of a (T, ...) array) and is sliced as ``data[index + slices]``, so
    a lazily read array is never read whole.  ``get_data_subset`` returns the
    chunks covering the bounding box as a device tensor and the offsets into
    it, loading them only when the covering chunks change; ``None`` asks for
    the whole array."""

    needs_bbox = True

    def __init__(self, data, z_stagger, y_stagger, x_stagger, *, index: tuple = (), device=None,
                 storage_dtype=None) -> None:
        super().__init__(z_stagger, y_stagger, x_stagger)
        self._data = data
        self._index = tuple(index)
        k = len(self._index)
        self._shape = tuple(int(n) for n in data.shape[k:])
        self._bounds = chunk_bounds(tuple(data.chunks)[k:], self._shape)
        self._device = torch.device(device) if device is not None else default_device()
        self._storage = to_torch_dtype(storage_dtype if storage_dtype is not None else np.dtype(data.dtype))
        if len(self._shape) != sum(self.dmask):
            raise ValueError(f"array of {len(self._shape)} dimensions for {sum(self.dmask)} active staggers")
        self._subset = None
        self._subset_bounds = None
        # bumped at every load: names the subset's contents (its storage may be reused)
        self.version = 0
        self.loads = 0

    @property
    def dtype(self):
        return self._storage

    @property
    def shape(self) -> tuple:
        return self._shape

    @property
    def data(self) -> torch.Tensor:
        """The whole array on the device (read now if other chunks are loaded)."""
        return self.get_data_subset(None)[0]

    def _new_bounds(self, bounding_box) -> tuple:
        if bounding_box is None:
            return tuple((0, n) for n in self._shape)
        pairs = [(bounding_box[2 * d], bounding_box[2 * d + 1]) for d in range(3)]
        active = [p for p, on in zip(pairs, self.dmask) if on]
        return tuple((compute_new_lower_bound(lo, off, b), compute_new_upper_bound(hi, off, b))
                     for (lo, hi), off, b in zip(active, self.active_offsets, self._bounds))

    def get_data_subset(self, bounding_box):
        """(device tensor of the covering chunks, offsets into it) (spatial_arrays.py:258-295)."""
        new_bounds = self._new_bounds(bounding_box)
        offsets = tuple(off - lo for off, (lo, _) in zip(self.active_offsets, new_bounds))
        if new_bounds != self._subset_bounds:
            block = self._data[self._index + tuple(slice(lo, hi) for lo, hi in new_bounds)]
            if hasattr(block, "compute"):
                block = block.compute()
            self._subset = as_device_tensor(np.asarray(block), self._device, self._storage)
            self._subset_bounds = new_bounds
            self.version += 1
            self.loads += 1
        return self._subset, offsets


# The reference's name for the chunked array type.
ChunkedDaskArray = ChunkedArray
