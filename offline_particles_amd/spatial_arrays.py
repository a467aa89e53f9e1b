"""Spatial arrays resident in HBM.

Mirrors the reference's spatial_arrays.py: ``Stagger`` (offsets and staggered
sizes of the ROMS C-grid, spatial_arrays.py:18-68), ``BBox``
(spatial_arrays.py:15) and ``NumpyArray`` (spatial_arrays.py:181-225), which
holds the whole array and returns it unchanged for any bounding box.  Here the
whole array lives in device memory (``DeviceArray``).  The reference's lazy
``ChunkedDaskArray`` (disk-backed chunk loading) is not part of this package.
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
