"""Chunked fields on the CPU: ``spatial_arrays.ChunkedArray``,
``StaticField.from_dask``, ``TimeDependentField.from_dask`` and
``from_zarr_chunked`` (the reference's ChunkedDaskArray,
spatial_arrays.py:228-350, and its factories, fields.py:216-231, 443-458).

dask is not installed here, so a small stand-in with dask's interface (shape,
dtype, per-dimension chunk sizes, lazy slices with ``compute()``) plays its
part and counts what is read.  The bounds rule is pinned by brute force
against what it must give: the smallest chunk-aligned range holding every node
that a stencil of a position in the box weights (floor(g) and ceil(g) for g
in [min + offset, max + offset], clipped to the array)."""

import math
import types

import numpy as np
import pytest
import torch

from offline_particles_amd.fields import BlendWindow, StaticField, TimeDependentField
from offline_particles_amd.launcher import _NO_BBOX, Launcher
from offline_particles_amd.snapshots import write_zarr_snapshots
from offline_particles_amd.spatial_arrays import (BBox, ChunkedArray, Stagger, chunk_bounds,
                                                  compute_new_lower_bound, compute_new_upper_bound)


class FakeDask:
    """dask.array's interface over a numpy array: slicing is lazy, compute() reads."""

    def __init__(self, a, chunks):
        self._a = np.asarray(a)
        self.shape, self.dtype = self._a.shape, self._a.dtype
        self.chunks = tuple(tuple(c) for c in chunks)
        self.ndim = self._a.ndim
        self.reads = []

    def __getitem__(self, key):
        parent = self

        class Lazy:
            def compute(self):
                parent.reads.append(key)
                return parent._a[key].copy()

        return Lazy()


def _regular(n, c):
    return tuple([c] * (n // c) + ([n % c] if n % c else []))


@pytest.mark.parametrize("seed", range(4))
def test_bounds_rule_is_the_smallest_covering_chunk_range(seed):
    rng = np.random.default_rng(seed)
    for _ in range(300):
        n = int(rng.integers(2, 40))
        sizes, left = [], n
        while left:
            sizes.append(int(rng.integers(1, min(left, 9) + 1)))
            left -= sizes[-1]
        b = chunk_bounds((tuple(sizes),), (n,))[0]
        off = float(rng.choice([-0.5, 0.0, 0.5]))
        lo_pos = float(rng.uniform(-3, n + 2))
        hi_pos = lo_pos + float(rng.uniform(0, n))
        if rng.random() < 0.3:  # integral box edges
            lo_pos, hi_pos = float(math.floor(lo_pos)), float(math.floor(hi_pos))
        gl, gu = lo_pos + off, hi_pos + off
        want_lo = max(int(x) for x in b if x <= max(0, math.floor(gl)))
        want_hi = min(int(x) for x in b if x >= min(n, math.ceil(gu) + 1))
        assert compute_new_lower_bound(lo_pos, off, b) == want_lo, (sizes, off, lo_pos)
        assert compute_new_upper_bound(hi_pos, off, b) == want_hi, (sizes, off, hi_pos)


def test_chunk_bounds_of_both_chunk_conventions():
    assert [list(b) for b in chunk_bounds(((4, 4, 2), (3, 3)), (10, 6))] == [[0, 4, 8, 10], [0, 3, 6]]
    assert [list(b) for b in chunk_bounds((4, 3), (10, 6))] == [[0, 4, 8, 10], [0, 3, 6]]
    with pytest.raises(ValueError):
        chunk_bounds(((4, 4),), (10,))


def test_chunked_array_loads_only_the_covering_chunks():
    a = np.arange(6 * 20 * 30, dtype=np.float32).reshape(6, 20, 30)
    d = FakeDask(a, ((3, 3), _regular(20, 8), _regular(30, 8)))
    arr = ChunkedArray(d, Stagger.CENTER, Stagger.LEFT, Stagger.CENTER, device="cpu")
    assert arr.needs_bbox and arr.shape == (6, 20, 30)
    box = BBox(0.2, 1.7, 9.0, 11.5, 3.0, 12.25)
    sub, offs = arr.get_data_subset(box)
    yoff = Stagger.LEFT.offset
    # z: chunk 0; y: rows 9 + yoff .. 12.5 + yoff, chunks [8, 16); x: chunks [0, 16)
    lo = (0, compute_new_lower_bound(9.0, yoff, chunk_bounds((_regular(20, 8),), (20,))[0]), 0)
    assert lo[1] == 8
    assert torch.equal(sub, torch.from_numpy(a[0:3, 8:16, 0:16]))
    assert offs == (0.0 - 0, yoff - 8, 0.0 - 0)
    assert len(d.reads) == 1 and arr.version == 1
    # a box covered by the same chunks reads nothing; one covered by others reads
    # again (as the reference, on any change of the covering range)
    arr.get_data_subset(BBox(0.5, 1.0, 9.5, 10.0, 2.0, 9.0))
    assert len(d.reads) == 1 and arr.version == 1
    sub2, offs2 = arr.get_data_subset(BBox(0.5, 3.5, 9.5, 10.0, 4.0, 20.0))
    assert len(d.reads) == 2 and arr.version == 2
    assert torch.equal(sub2, torch.from_numpy(a[0:6, 8:16, 0:24]))
    # None asks for the whole array
    whole, offs3 = arr.get_data_subset(None)
    assert torch.equal(whole, torch.from_numpy(a)) and offs3 == arr.active_offsets
    assert torch.equal(arr.data, torch.from_numpy(a))


def test_chunked_array_of_invariant_dimensions_takes_the_active_ones():
    a = np.arange(20 * 30, dtype=np.float64).reshape(20, 30)
    arr = ChunkedArray(FakeDask(a, (_regular(20, 8), _regular(30, 8))), Stagger.INVARIANT, Stagger.CENTER,
                       Stagger.RIGHT, device="cpu")
    sub, offs = arr.get_data_subset(BBox(-100.0, 100.0, 17.0, 18.0, 0.0, 3.0))
    assert torch.equal(sub, torch.from_numpy(a[16:20, 0:8]))
    assert offs == (0.0 - 16, Stagger.RIGHT.offset)
    with pytest.raises(ValueError):
        ChunkedArray(FakeDask(a, (_regular(20, 8), _regular(30, 8))), Stagger.CENTER, Stagger.CENTER,
                     Stagger.CENTER, device="cpu")


def test_static_field_from_dask():
    a = np.random.default_rng(0).standard_normal((20, 30))
    f = StaticField.from_dask(FakeDask(a, (_regular(20, 8), _regular(30, 8))), "invariant", "center", "center",
                              device="cpu")
    assert f.needs_bbox and f.spatial_shape == (20, 30)
    fd = f.get_field_data(0.0, BBox(0.0, 0.0, 1.0, 2.0, 9.0, 10.0))
    assert torch.equal(fd.array, torch.from_numpy(a[0:8, 8:16])) and fd.offsets == (0.0, -8.0)
    assert not StaticField.from_numpy(a, "invariant", "center", "center", device="cpu").needs_bbox


@pytest.mark.parametrize("cyclic", [False, True])
def test_time_dependent_field_from_dask(cyclic):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((4, 6, 20, 30)).astype(np.float32)
    d = FakeDask(a, ((1,) * 4, (6,), _regular(20, 8), _regular(30, 8)))
    f = TimeDependentField.from_dask(d, "center", "center", "left", cyclic=cyclic, device="cpu")
    assert f.needs_bbox
    box = BBox(0.0, 5.0, 1.0, 6.0, 20.0, 21.0)
    fd = f.get_field_data(1.25, box)
    assert isinstance(fd.array, BlendWindow)
    xoff = Stagger.LEFT.offset
    assert torch.equal(fd.array.prev, torch.from_numpy(a[1, :, 0:8, 16:24]))
    assert torch.equal(fd.array.next, torch.from_numpy(a[2, :, 0:8, 16:24]))
    assert fd.offsets == (0.0, 0.0, xoff - 16)
    assert fd.array.ft == pytest.approx(0.25) and fd.array.gt == pytest.approx(0.75)
    # the window moves on: snapshot 2's chunks are reused, only snapshot 3's are read
    n = len(d.reads)
    f.get_field_data(2.5, box)
    assert len(d.reads) == n + 1 and d.reads[-1][0] == 3
    if cyclic:
        fd = f.get_field_data(3.5, box)  # snapshots 3 and 0
        assert torch.equal(fd.array.next, torch.from_numpy(a[0, :, 0:8, 16:24]))
    # whole snapshots for callers that take the whole field (the fused engine)
    assert torch.equal(f.snapshot(1), torch.from_numpy(a[1]))
    # preload_space: whole snapshots, offsets unchanged, no bounding box needed
    g = TimeDependentField.from_dask(FakeDask(a, d.chunks), "center", "center", "left", preload_space=True,
                                     device="cpu")
    assert not g.needs_bbox
    fd = g.get_field_data(1.25, box)
    assert torch.equal(fd.array.prev, torch.from_numpy(a[1])) and fd.offsets == (0.0, 0.0, xoff)


def test_time_dependent_field_from_zarr_chunked(tmp_path):
    a = np.random.default_rng(2).standard_normal((3, 5, 17, 23)).astype(np.float32)
    write_zarr_snapshots(tmp_path / "s.zarr", "u", a, chunks=(1, 5, 8, 8))
    f = TimeDependentField.from_zarr_chunked(tmp_path / "s.zarr", "u", "center", "center", "center", device="cpu")
    fd = f.get_field_data(0.5, BBox(1.0, 2.0, 9.0, 9.5, 17.0, 21.0))
    assert torch.equal(fd.array.prev, torch.from_numpy(a[0, :, 8:16, 16:23]))
    assert torch.equal(fd.array.next, torch.from_numpy(a[1, :, 8:16, 16:23]))
    assert fd.offsets == (0.0, -8.0, -16.0)


def test_launcher_computes_the_box_only_for_chunked_fields():
    a = np.zeros((2, 4, 10, 12), dtype=np.float32)
    fs = types.SimpleNamespace(constants={}, fields={})
    resident = TimeDependentField.from_numpy(a, "center", "center", "center", device="cpu")
    chunked = TimeDependentField.from_dask(FakeDask(a, ((1, 1), (4,), (5, 5), (6, 6))), "center", "center",
                                           "center", device="cpu")
    fs.fields = {"r": resident, "c": chunked}
    launcher = Launcher(types.SimpleNamespace(constants={}, fields=fs.fields, __getitem__=None), index_padding=1)
    launcher._fieldset = fs.fields
    parts = {"zidx": torch.tensor([1.0, 2.0]), "yidx": torch.tensor([3.0, 7.5]), "xidx": torch.tensor([0.5, 4.0])}
    assert launcher._bbox_for(types.SimpleNamespace(simulation_fields=("r",)), parts) == _NO_BBOX
    assert launcher._bbox_for(types.SimpleNamespace(simulation_fields=("r", "c")), parts) == \
        BBox(0.0, 3.0, 2.0, 8.5, -0.5, 5.0)
