"""Fields resident in HBM (mirrors the reference's fields.py).

``StaticField`` holds one device array.  ``TimeDependentField`` holds its
snapshots in device memory and keeps the reference's two-snapshot window
(``set_time_index``/``increment_time``/``decrement_time``, fields.py:340-392).

The one deliberate difference is where the time blend happens.  The
reference's ``get_field_data`` (fields.py:394-430) blends the *whole* window
into an FP64 array on every launch.  Here ``get_field_data`` returns a
``BlendWindow`` (the two snapshots plus the rounded weights) and the kernels
blend each stencil value they read, ``prev*gt + next*ft`` in the storage
type, which is the same arithmetic on the same operands: the values a kernel
interpolates are bit-identical to the reference's blended array, without the
O(grid) pass.  ``BlendWindow.materialize()`` produces the full blended array
when a caller needs it.
"""

from __future__ import annotations

import collections
import os

import numpy as np
import torch

from ._lib import OPX_BF16, OPX_F32, OPX_F64, storage_code
from ._zarr import ZarrArray
from .snapshots import ArraySnapshots, SnapshotSource, ZarrSnapshots
from .spatial_arrays import (BBox, ChunkedArray, DeviceArray, SpatialArray, Stagger, as_device_tensor, default_device,
                             to_torch_dtype)

_FieldDataTuple = collections.namedtuple("FieldData", ["array", "offsets"])


class FieldData(_FieldDataTuple):
    """The reference's FieldData (array, offsets) (fields.py:12).

    ``packed`` (an attribute, not a tuple item, so ``array, offsets = fd``
    still works) is None or ``(tensor, kind)``: a copy of the same values in
    which one load brings a node's stencil corners (``opx_pack_field_as``;
    OPX_PACK_QUAD for a static field: the (y, x) corners; OPX_PACK_HEX for a
    3-D snapshot pair: all 8 corners of both snapshots; OPX_PACK_OCT for a 2-D
    one: the (y, x) corners of both) or, inside ``sorted_stencils``, a stencil
    row's x-neighbours of both snapshots (OPX_PACK_TPAIR).  The kernels read
    it when present; the values they interpolate are unchanged."""

    packed = None


# Packed stencil copies for the drop-in kernels (F32 and BF16 device fields of
# 2 or 3 dimensions): a 3-D stencil of a snapshot pair then touches 1 cache
# line (HEX) instead of 8, a 2-D one 1 instead of 2.  They cost 16 (3-D
# pair), 8 (2-D pair) or 4 (static) values per node of device memory; False
# turns them off.
PACK_STENCILS = True


class unpacked:
    """Context in which get_field_data makes no packed copies (for kernels whose
    particles are sorted by cell, where neighbouring threads already share cache
    lines and the copies only add footprint)."""

    def __enter__(self):
        global PACK_STENCILS
        self._saved, PACK_STENCILS = PACK_STENCILS, False
        return self

    def __exit__(self, *exc):
        global PACK_STENCILS
        PACK_STENCILS = self._saved
        return False


# Inside sorted_stencils, snapshot pairs are packed as OPX_PACK_TPAIR.
_SORTED = False


class sorted_stencils:
    """Context for kernels whose particles are sorted by cell (a Simulation with
    spatial ordering): snapshot pairs get the TPAIR copy, the two snapshots
    interleaved node by node, instead of the OCT copy.  Neighbouring threads
    then share stencil rows, so the OCT copy's four copies of every node only
    add footprint; TPAIR stores each value once and still reads a row of both
    snapshots with one load."""

    def __enter__(self):
        global _SORTED
        self._saved, _SORTED = _SORTED, True
        return self

    def __exit__(self, *exc):
        global _SORTED
        _SORTED = self._saved
        return False


def _packable(t) -> bool:
    return (PACK_STENCILS and isinstance(t, torch.Tensor) and t.is_cuda and t.dim() in (2, 3)
            and t.dtype in (torch.float32, torch.bfloat16) and t.is_contiguous() and min(t.shape[-2:]) >= 2)


# The plain kernels read dx, dy and h's QUAD copies from one buffer holding the
# three side by side (opx_field.packed_stride), when they share a dtype and shape,
# and zeta's OCT copy with h's QUAD from another.
MERGE_STATIC_STENCILS = True

# Largest HEX copy made of one snapshot pair (16 values per node); larger
# 3-D fields get the OCT copy (8 values per node).
HEX_MAX_BYTES = 16 << 30


def _pair_kind(t: torch.Tensor) -> int:
    """The packed kind of a snapshot pair like `t`: TPAIR for sorted particles,
    else HEX for a 3-D field (a trilinear stencil in one line) within
    HEX_MAX_BYTES, else OCT."""
    from . import _lib

    if _SORTED:
        return _lib.OPX_PACK_TPAIR
    if t.dim() == 3 and t.shape[0] >= 2 and t.numel() * 16 * t.element_size() <= HEX_MAX_BYTES:
        return _lib.OPX_PACK_HEX
    return _lib.OPX_PACK_OCT


def _pack(prev: torch.Tensor, next_: torch.Tensor | None, kind: int | None = None) -> tuple:
    """(packed tensor, kind) of prev (QUAD) or of the pair (prev, next_) (OCT,
    or the kind given)."""
    from . import _lib

    dims = (1,) * (3 - prev.dim()) + tuple(prev.shape)
    if kind is None:
        kind = _lib.OPX_PACK_QUAD if next_ is None else _lib.OPX_PACK_OCT
    code = storage_code(prev.dtype)
    nbytes = _lib.ctypes.c_uint64()
    _lib.call("opx_packed_field_bytes", code, kind, *dims, _lib.ctypes.byref(nbytes))
    out = torch.empty(int(nbytes.value), dtype=torch.uint8, device=prev.device)
    _lib.call("opx_pack_field_as", prev.data_ptr(), None if next_ is None else next_.data_ptr(), code, kind, *dims,
              out.data_ptr(), torch.cuda.current_stream(prev.device).cuda_stream)
    return out, kind


class BlendWindow:
    """Snapshots It and It+1 of a time-dependent field plus blend weights.

    ``ft``/``gt`` are the weights rounded to the storage type as
    fields.py:423-424 does (``dtype.type(ft)``, ``dtype.type(1 - ft)``);
    bfloat16 storage blends in float32 with float32 weights.
    """

    __slots__ = ("prev", "next", "ft", "gt", "It")

    def __init__(self, prev: torch.Tensor, next: torch.Tensor, ft: float, gt: float, It: int = 0):
        self.prev = prev
        self.next = next
        self.ft = float(ft)
        self.gt = float(gt)
        self.It = It

    @property
    def shape(self) -> tuple:
        return tuple(self.prev.shape)

    @property
    def dtype(self):
        return self.prev.dtype

    @property
    def ndim(self) -> int:
        return self.prev.dim()

    def materialize(self) -> torch.Tensor:
        """The blended FP64 array the reference would hand to a kernel."""
        if self.prev.dtype == torch.float64:
            a = self.prev * self.gt
            b = self.next * self.ft
            return a + b
        work = torch.float32
        a = self.prev.to(work) * torch.tensor(self.gt, dtype=work, device=self.prev.device)
        b = self.next.to(work) * torch.tensor(self.ft, dtype=work, device=self.prev.device)
        return (a + b).to(torch.float64)


def blend_weights(time_index: float, dtype) -> tuple[int, float, float]:
    """(It, ft, gt) exactly as fields.py:409-424 computes them."""
    It, ft = divmod(time_index, 1)
    It = int(It)
    code = storage_code(dtype)
    if code == OPX_F64:
        ft = np.float64(ft)
        gt = np.float64(1 - ft)
    else:  # float32 and bfloat16 storage blend in float32
        ft = np.float32(ft)
        gt = np.float32(1 - ft)
    return It, float(ft), float(gt)


class Field:
    """Abstract base of fields (fields.py:15-119)."""

    def __init__(self, z_stagger, y_stagger, x_stagger) -> None:
        self._z_stagger = Stagger(z_stagger)
        self._y_stagger = Stagger(y_stagger)
        self._x_stagger = Stagger(x_stagger)
        self._z_offset = self._z_stagger.offset
        self._y_offset = self._y_stagger.offset
        self._x_offset = self._x_stagger.offset

    z_stagger = property(lambda self: self._z_stagger)
    y_stagger = property(lambda self: self._y_stagger)
    x_stagger = property(lambda self: self._x_stagger)
    z_offset = property(lambda self: self._z_offset)
    y_offset = property(lambda self: self._y_offset)
    x_offset = property(lambda self: self._x_offset)

    @property
    def stagger(self):
        return (self._z_stagger, self._y_stagger, self._x_stagger)

    @property
    def all_offsets(self):
        return (self._z_offset, self._y_offset, self._x_offset)

    @property
    def active_offsets(self) -> tuple:
        return tuple(o for o, s in zip(self.all_offsets, self.stagger) if s.is_active)

    @property
    def dmask(self):
        return tuple(s.is_active for s in self.stagger)

    @property
    def nspatial_dims(self) -> int:
        return len(self.spatial_shape)

    # True when get_field_data reads only what the particles' bounding box
    # covers (chunked fields), so the launcher computes the box for it
    needs_bbox = False

    def _expected_spatial(self, simulation_shape):
        staggered = (
            self.z_stagger.expected_size(simulation_shape[1]),
            self.y_stagger.expected_size(simulation_shape[2]),
            self.x_stagger.expected_size(simulation_shape[3]),
        )
        return tuple(s for s in staggered if s is not None)


class StaticField(Field):
    """A field that does not change in time (fields.py:122-231)."""

    def __init__(self, data: SpatialArray):
        super().__init__(data.z_stagger, data.y_stagger, data.x_stagger)
        self._data = data

    @property
    def data(self) -> SpatialArray:
        return self._data

    @property
    def needs_bbox(self) -> bool:
        return bool(getattr(self._data, "needs_bbox", False))

    @property
    def dtype(self):
        return self._data.dtype

    @property
    def spatial_shape(self) -> tuple:
        return self._data.shape

    def __repr__(self) -> str:
        return (f"StaticField(shape={self._data.shape}, z_stagger={self.z_stagger}, "
                f"y_stagger={self.y_stagger}, x_stagger={self.x_stagger})")

    def validate_shape(self, simulation_shape) -> None:
        expected = self._expected_spatial(simulation_shape)
        if tuple(self._data.shape) != expected:
            raise ValueError(f"Expected shape {expected} but data has shape {tuple(self._data.shape)}")

    def get_field_data(self, time_index: float, bbox: BBox) -> FieldData:
        array, offsets = self._data.get_data_subset(bbox)
        fd = FieldData(array, offsets)
        if _packable(array):
            # a chunked array's version names the loaded chunks (storage may be reused)
            key = (array.data_ptr(), tuple(array.shape), array.dtype, getattr(self._data, "version", 0))
            cached = getattr(self, "_packed", None)
            if cached is None or cached[0] != key:
                cached = self._packed = (key, _pack(array, None))
            fd.packed = cached[1]
        return fd

    @classmethod
    def from_numpy(cls, data, z_stagger, y_stagger, x_stagger, *, device=None) -> "StaticField":
        return cls(DeviceArray(data, Stagger(z_stagger), Stagger(y_stagger), Stagger(x_stagger), device))

    from_array = from_numpy

    @classmethod
    def from_dask(cls, data, z_stagger, y_stagger, x_stagger, *, device=None, storage_dtype=None) -> "StaticField":
        """A field read chunk by chunk from a chunked array (fields.py:216-231):
        each launch holds on the device only the chunks that cover the
        particles' bounding box (``ChunkedArray``).  ``data`` is a dask array or
        any array with ``chunks`` and basic slicing (e.g. a ``ZarrArray``)."""
        return cls(ChunkedArray(data, Stagger(z_stagger), Stagger(y_stagger), Stagger(x_stagger), device=device,
                                storage_dtype=storage_dtype))


class TimeDependentField(Field):
    """A field given as snapshots on a time axis, resident in HBM (fields.py:237-458).

    ``data`` is a (T, *spatial) host or device array, or a ``SnapshotSource``
    (e.g. ``ZarrSnapshots``) read lazily: then only the snapshots in use stay
    on the device (the last few, the next one decoded ahead on a host thread),
    as the reference's dask-backed fields load what a launch needs
    (fields.py:443-458).  A chunked host array (a dask array, a ``ZarrArray``)
    is read lazily the same way, a snapshot at a time; with
    ``spatial_array_factory=ChunkedArray`` (``from_dask``) each launch reads
    only the chunks of its two snapshots that cover the particles' bounding
    box (the reference's ChunkedDaskArray slices).  With ``cyclic=True`` snapshot ``It`` is
    ``data[It % T]`` so a short set of periodic snapshots serves an arbitrarily
    long time axis; otherwise ``It`` must lie in ``[0, T-2]`` as in the
    reference.
    """

    # snapshots of a lazily read field kept on the device
    LAZY_CACHE = 3

    def __init__(self, data, z_stagger, y_stagger, x_stagger, spatial_array_factory=None,
                 output_dtype=np.float64, *, cyclic: bool = False, device=None, storage_dtype=None):
        super().__init__(z_stagger, y_stagger, x_stagger)
        if data.ndim < 2:
            raise ValueError("TimeDependentField requires at least 2 dimensions (time + spatial).")
        if data.shape[0] < 2 and not cyclic:
            raise ValueError("TimeDependentField requires at least 2 time steps.")
        # chunked: the host array, sliced per launch by the bounding box
        self._chunked = None
        if spatial_array_factory is ChunkedArray:
            if isinstance(data, (SnapshotSource, torch.Tensor)):
                raise TypeError("a chunked field needs a host array with chunks (a dask array, a ZarrArray)")
            self._chunked = data
            self._chunk_slices = collections.OrderedDict()
            data = ArraySnapshots(data)
        elif (hasattr(data, "chunks") and not isinstance(data, (np.ndarray, torch.Tensor, SnapshotSource))):
            # a lazily read chunked array: a snapshot at a time, as the reference
            # slices data[It] (fields.py:340-392)
            data = ArraySnapshots(data)
        if isinstance(data, SnapshotSource):
            self._source = data
            self._snapshots = None
            self._device = torch.device(device) if device is not None else default_device()
            self._storage = to_torch_dtype(storage_dtype if storage_dtype is not None else data.dtype)
            self._cache = collections.OrderedDict()
            self._shape = tuple(data.shape)
        else:
            self._source = None
            self._snapshots = as_device_tensor(data, device, storage_dtype)
            self._shape = tuple(self._snapshots.shape)
        self._num_timesteps = int(self._shape[0])
        self._cyclic = bool(cyclic)
        self._output_dtype = output_dtype
        self._It = 0

    @classmethod
    def from_zarr(cls, store, name: str, z_stagger, y_stagger, x_stagger, *, preload: bool = False,
                  cyclic: bool = False, device=None, storage_dtype=None, workers: int = 8) -> "TimeDependentField":
        """A field from a (T, ...) Zarr v3 array of a local store: every snapshot
        loaded now (``preload``) or read lazily through ``ZarrSnapshots``."""
        src = ZarrSnapshots(store, name, workers=workers)
        data = src.read_all() if preload else src
        return cls(data, z_stagger, y_stagger, x_stagger, cyclic=cyclic, device=device, storage_dtype=storage_dtype)

    @classmethod
    def from_zarr_chunked(cls, store, name: str, z_stagger, y_stagger, x_stagger, *, cyclic: bool = False,
                          device=None, storage_dtype=None) -> "TimeDependentField":
        """A field over a (T, ...) Zarr v3 array of a local store whose launches
        read only the chunks covering the particles' bounding box (the Zarr
        form of ``from_dask``)."""
        return cls.from_dask(ZarrArray(os.fspath(store), name), z_stagger, y_stagger, x_stagger, cyclic=cyclic,
                             device=device, storage_dtype=storage_dtype)

    @classmethod
    def from_dask(cls, data, z_stagger, y_stagger, x_stagger, *, preload_space: bool = False, cyclic: bool = False,
                  device=None, storage_dtype=None) -> "TimeDependentField":
        """A field over a chunked (T, ...) array (fields.py:443-458): with
        ``preload_space`` each snapshot in use is read whole, otherwise each
        launch reads the chunks covering the particles' bounding box
        (``ChunkedArray``).  ``data`` is a dask array or any array with
        ``chunks`` and basic slicing (e.g. a ``ZarrArray``)."""
        factory = DeviceArray if preload_space else ChunkedArray
        return cls(data, z_stagger, y_stagger, x_stagger, factory, cyclic=cyclic, device=device,
                   storage_dtype=storage_dtype)

    @property
    def needs_bbox(self) -> bool:
        return self._chunked is not None

    def _chunk_slice(self, It: int) -> ChunkedArray:
        """Snapshot It as a ChunkedArray (the last three kept, so the window's
        snapshots keep their loaded chunks as it moves, fields.py:340-366)."""
        idx = It % self._num_timesteps if self._cyclic else It
        a = self._chunk_slices.get(idx)
        if a is None:
            a = self._chunk_slices[idx] = ChunkedArray(self._chunked, self.z_stagger, self.y_stagger, self.x_stagger,
                                                       index=(idx,), device=self._device, storage_dtype=self._storage)
            while len(self._chunk_slices) > 3:
                self._chunk_slices.popitem(last=False)
        else:
            self._chunk_slices.move_to_end(idx)
        return a

    @property
    def source(self):
        """The SnapshotSource of a lazily read field (None when every snapshot is resident)."""
        return self._source

    @property
    def data(self) -> torch.Tensor:
        if self._source is not None:
            raise TypeError("this field reads its snapshots lazily: use snapshot(It) or .source")
        return self._snapshots

    @property
    def dtype(self):
        return self._storage if self._source is not None else self._snapshots.dtype

    @property
    def cyclic(self) -> bool:
        return self._cyclic

    @property
    def num_snapshots(self) -> int:
        return self._num_timesteps

    @property
    def spatial_shape(self) -> tuple:
        return tuple(self._shape[1:])

    def __repr__(self) -> str:
        return (f"TimeDependentField(shape={self._shape}, z_stagger={self.z_stagger}, "
                f"y_stagger={self.y_stagger}, x_stagger={self.x_stagger}, cyclic={self._cyclic})")

    def validate_shape(self, simulation_shape) -> None:
        expected = self._expected_spatial(simulation_shape)
        if self.spatial_shape != expected:
            raise ValueError(f"Expected spatial shape {expected} but data has shape {self.spatial_shape}")
        if not self._cyclic and self._num_timesteps != simulation_shape[0]:
            raise ValueError(f"Expected {simulation_shape[0]} snapshots, data has {self._num_timesteps}")

    def snapshot(self, It: int) -> torch.Tensor:
        """Device view of snapshot It."""
        idx = It % self._num_timesteps if self._cyclic else It
        if self._source is None:
            return self._snapshots[idx]
        if not 0 <= idx < self._num_timesteps:
            raise IndexError(f"snapshot {It} out of range")
        t = self._cache.get(idx)
        if t is None:
            host = self._source.read(idx)
            t = torch.from_numpy(np.ascontiguousarray(host)).to(self._device, self._storage)
            self._cache[idx] = t
            while len(self._cache) > self.LAZY_CACHE:
                self._cache.popitem(last=False)
            nxt = idx + 1
            if self._cyclic or nxt < self._num_timesteps:
                self._source.prefetch(nxt)
        else:
            self._cache.move_to_end(idx)
        return t

    @property
    def previous_time_slice(self) -> torch.Tensor:
        return self.snapshot(self._It)

    @property
    def next_time_slice(self) -> torch.Tensor:
        return self.snapshot(self._It + 1)

    def increment_time(self) -> None:
        if not self._cyclic and self._It == self._num_timesteps - 2:
            raise IndexError("Cannot increment past the penultimate timestep.")
        self._It += 1

    def decrement_time(self) -> None:
        if not self._cyclic and self._It == 0:
            raise IndexError("Cannot decrement past the first timestep.")
        self._It -= 1

    def set_time_index(self, It: int) -> None:
        """Move the window to [It, It+1] (fields.py:368-392)."""
        if It == self._It:
            return
        if It == self._It + 1:
            return self.increment_time()
        if It == self._It - 1:
            return self.decrement_time()
        if not self._cyclic and (It < 0 or It > self._num_timesteps - 2):
            raise IndexError(f"Valid range of time indices is 0,...,{self._num_timesteps - 2}, got {It}.")
        self._It = It

    # packed snapshot pairs kept on the device (an RK2 step near a snapshot
    # boundary reads two pairs)
    PACKED_PAIRS = 2

    def get_field_data(self, time_index: float, bbox: BBox) -> FieldData:
        It, ft, gt = blend_weights(time_index, self.dtype)
        self.set_time_index(It)
        versions = ()
        if self._chunked is not None:
            # the covering chunks of both snapshots (fields.py:414-416); the
            # versions name what was loaded
            pa, na = self._chunk_slice(It), self._chunk_slice(It + 1)
            prev, offsets = pa.get_data_subset(bbox)
            nxt, _ = na.get_data_subset(bbox)
            versions = (pa.version, na.version)
        else:
            prev, nxt = self.snapshot(It), self.snapshot(It + 1)
            offsets = self.active_offsets
        fd = FieldData(BlendWindow(prev, nxt, ft, gt, It), offsets)
        if _packable(prev):
            # the snapshot indices name the contents (a lazily read field may hand
            # out a freed snapshot's address for another one), the pointers the storage
            K = self._num_timesteps
            idx = (It % K, (It + 1) % K) if self._cyclic else (It, It + 1)
            kind = _pair_kind(prev)
            key = idx + (prev.data_ptr(), nxt.data_ptr(), tuple(prev.shape), prev.dtype, kind) + versions
            cache = self.__dict__.setdefault("_packed_pairs", collections.OrderedDict())
            packed = cache.get(key)
            if packed is None:
                packed = cache[key] = _pack(prev, nxt, kind)
                while len(cache) > self.PACKED_PAIRS:
                    cache.popitem(last=False)
            else:
                cache.move_to_end(key)
            fd.packed = packed
        return fd

    @classmethod
    def from_numpy(cls, data, z_stagger, y_stagger, x_stagger, **kw) -> "TimeDependentField":
        return cls(data, z_stagger, y_stagger, x_stagger, **kw)

    from_array = from_numpy


__all__ = ["FieldData", "BlendWindow", "Field", "StaticField", "TimeDependentField", "blend_weights",
           "OPX_F32", "OPX_F64", "OPX_BF16"]
