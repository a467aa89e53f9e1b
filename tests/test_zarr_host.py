"""The Zarr v3 trajectory store (output/_zarr.py, output/_zarr_writer.py; the
reference's output/_zarr_writer.py:16-223 with zarr-python 3.1.5), checked
on the CPU against the Zarr v3 core specification's layout: group and array
metadata documents, the regular chunk grid with the "default" chunk key
encoding ("c/<i>/<j>"), full-size edge chunks padded with the fill value, the
bytes codec (little endian) and zstd frames.  No zarr-python is available to
write golden stores, so the layout is pinned to the specification."""

import json
import os
import types

import numpy as np
import pytest
import torch

from offline_particles_amd.kernels.roms import compute_z_kernel
from offline_particles_amd.output import Output, ZarrArray, ZarrOutputBuilder, open_array


def _state(t, **fields):
    return types.SimpleNamespace(time=t, particles={k: torch.as_tensor(v) for k, v in fields.items()})


def _outputs():
    return Output("depth", compute_z_kernel("_depth"), particle_field="_depth")


def test_store_layout_follows_the_v3_spec(tmp_path):
    root = str(tmp_path / "traj.zarr")
    b = ZarrOutputBuilder("zarr", root, chunksize=4, array_kwargs={"compressors": None})
    b.add_output(_outputs())
    w = b.build(10)
    rows = [np.arange(10, dtype=np.float64) * (r + 1) for r in range(3)]
    for r, row in enumerate(rows):
        w.write_time(_state(100.0 * r, _depth=row))
        w.write_output("depth", _state(100.0 * r, _depth=row))
        w.finalise_write_round(None)
    assert json.load(open(os.path.join(root, "zarr.json"))) == {"zarr_format": 3, "node_type": "group",
                                                                "attributes": {}}
    meta = json.load(open(os.path.join(root, "depth", "zarr.json")))
    assert meta == {
        "zarr_format": 3, "node_type": "array", "shape": [3, 10], "data_type": "float64",
        "chunk_grid": {"name": "regular", "configuration": {"chunk_shape": [1, 4]}},
        "chunk_key_encoding": {"name": "default", "configuration": {"separator": "/"}},
        "fill_value": 0.0, "codecs": [{"name": "bytes", "configuration": {"endian": "little"}}],
        "attributes": {}, "dimension_names": ["time", "particle"]}
    tmeta = json.load(open(os.path.join(root, "time", "zarr.json")))
    assert tmeta["shape"] == [3] and tmeta["chunk_grid"]["configuration"]["chunk_shape"] == [1]
    assert tmeta["dimension_names"] == ["time"]
    # chunk (r, j) holds particles 4j .. 4j+3 of row r as raw little-endian FP64; the edge chunk is
    # full size, padded with the fill value
    for r, row in enumerate(rows):
        for j in range(3):
            raw = open(os.path.join(root, "depth", "c", str(r), str(j)), "rb").read()
            assert len(raw) == 4 * 8
            got = np.frombuffer(raw, dtype="<f8")
            exp = np.zeros(4)
            part = row[4 * j:4 * j + 4]
            exp[:len(part)] = part
            assert np.array_equal(got, exp)
        t = np.frombuffer(open(os.path.join(root, "time", "c", str(r)), "rb").read(), dtype="<f8")
        assert np.array_equal(t, [100.0 * r])
    a = open_array(root, "depth")
    assert a.shape == (3, 10) and np.array_equal(a[:, :], np.stack(rows))
    assert np.array_equal(a[1, 3:9], rows[1][3:9]) and a[2, 9] == rows[2][9]
    assert np.array_equal(open_array(root, "time")[:], [0.0, 100.0, 200.0])


def test_default_codecs_are_bytes_then_zstd(tmp_path):
    """zarr.create_array's default compressors="auto" gives a numeric v3 array
    the codecs [bytes (little endian), zstd (level 0)]; the writer does the same
    when array_kwargs names no compressors."""
    root = str(tmp_path / "d.zarr")
    b = ZarrOutputBuilder("zarr", root, chunksize=4)
    b.add_output(_outputs())
    w = b.build(6)
    row = np.arange(6, dtype=np.float64)
    w.write_time(_state(1.0, _depth=row))
    w.write_output("depth", _state(1.0, _depth=row))
    w.finalise_write_round(None)
    for name in ("depth", "time"):
        meta = json.load(open(os.path.join(root, name, "zarr.json")))
        assert meta["codecs"] == [{"name": "bytes", "configuration": {"endian": "little"}},
                                  {"name": "zstd", "configuration": {"level": 0, "checksum": False}}]
    assert open(os.path.join(root, "depth", "c", "0", "0"), "rb").read()[:4] == bytes.fromhex("28b52ffd")
    assert np.array_equal(open_array(root, "depth")[0], row)


def test_zstd_chunks_and_fill_value(tmp_path):
    root = str(tmp_path / "z.zarr")
    b = ZarrOutputBuilder("zarr", root, chunksize=6, array_kwargs={"compressors": "zstd", "fill_value": np.nan})
    b.add_output(_outputs())
    w = b.build(8)
    row = np.linspace(-1, 1, 8)
    w.write_time(_state(5.0, _depth=row))
    w.write_output("depth", _state(5.0, _depth=row))
    w.finalise_write_round(None)
    meta = json.load(open(os.path.join(root, "depth", "zarr.json")))
    assert meta["codecs"][1] == {"name": "zstd", "configuration": {"level": 0, "checksum": False}}
    assert meta["fill_value"] == "NaN"
    raw = open(os.path.join(root, "depth", "c", "0", "1"), "rb").read()
    assert raw[:4] == bytes.fromhex("28b52ffd")  # a zstd frame
    a = open_array(root, "depth")
    assert np.array_equal(a[0], row)
    edge = a.read_chunk((0, 1))
    assert np.array_equal(edge[0, :2], row[6:]) and np.all(np.isnan(edge[0, 2:]))


def test_round_checks_and_errors(tmp_path):
    root = str(tmp_path / "e.zarr")
    b = ZarrOutputBuilder("zarr", root)
    b.add_output(_outputs())
    with pytest.raises(KeyError):
        b.add_output(_outputs())
    with pytest.raises(TypeError):
        ZarrOutputBuilder("zarr", root, array_kwargs={"shards": (1, 10)})
    w = b.build(5)
    with pytest.raises(KeyError):
        w.write_output("nope", _state(0.0, _depth=np.zeros(5)))
    w.write_time(_state(0.0, _depth=np.zeros(5)))
    with pytest.raises(RuntimeError, match="depth"):
        w.finalise_write_round(None)  # the depth row of this round is missing
    with pytest.raises(FileExistsError):
        ZarrOutputBuilder("zarr", root).build(5)
    ZarrOutputBuilder("zarr", root, overwrite=True).build(5)
    b.remove_output("depth")
    with pytest.raises(KeyError):
        b.remove_output("depth")


def test_async_rounds_in_order(tmp_path):
    root = str(tmp_path / "a.zarr")
    b = ZarrOutputBuilder("zarr", root, chunksize=1000, async_writes=True)
    b.add_output(_outputs())
    w = b.build(2500)
    rng = np.random.default_rng(0)
    rows = [rng.normal(size=2500) for _ in range(12)]
    for r, row in enumerate(rows):
        w.write_time(_state(float(r), _depth=row))
        w.write_output("depth", _state(float(r), _depth=row))
        w.finalise_write_round(None)
    w.flush()
    a = open_array(root, "depth")
    assert a.shape == (12, 2500)
    assert np.array_equal(a[:, :], np.stack(rows))
    w.close()


def test_array_reads_other_chunkings(tmp_path):
    """Reads of a 3-D array across chunk boundaries, as a snapshot source would do."""
    root = str(tmp_path / "s.zarr")
    data = np.arange(4 * 5 * 7, dtype=np.float32).reshape(4, 5, 7)
    a = ZarrArray.create(root, "u", shape=data.shape, dtype=np.float32, chunks=(1, 3, 4))
    for i in range(4):
        for j in range(2):
            for k in range(2):
                a.write_chunk((i, j, k), data[i:i + 1, 3 * j:3 * j + 3, 4 * k:4 * k + 4])
    b = open_array(root, "u")
    assert b.dtype == np.float32 and b.chunks == (1, 3, 4)
    assert np.array_equal(b[:, :, :], data)
    assert np.array_equal(b[2], data[2])
    assert np.array_equal(b[1:3, 2:5, 3:6], data[1:3, 2:5, 3:6])


def test_zarr_snapshots_read_back_and_prefetch(tmp_path):
    from offline_particles_amd.snapshots import ZarrSnapshots, write_zarr_snapshots

    rng = np.random.default_rng(1)
    data = rng.normal(size=(5, 4, 6, 9)).astype(np.float32)
    write_zarr_snapshots(tmp_path / "s.zarr", "u", data, chunks=(1, 2, 4, 5))
    write_zarr_snapshots(tmp_path / "s.zarr", "v", data, compressor="zstd")
    for name in ("u", "v"):
        src = ZarrSnapshots(tmp_path / "s.zarr", name)
        assert src.shape == data.shape and src.dtype == np.float32 and len(src) == 5
        src.prefetch(3)
        assert np.array_equal(src.read(3), data[3])
        assert np.array_equal(src.read(7), data[2])  # index mod T
        assert np.array_equal(src.read_all(), data)


def test_lazy_time_dependent_field_equals_resident(tmp_path):
    from offline_particles_amd.fields import TimeDependentField
    from offline_particles_amd.launcher import _NO_BBOX
    from offline_particles_amd.snapshots import ZarrSnapshots, write_zarr_snapshots

    rng = np.random.default_rng(2)
    data = rng.normal(size=(6, 3, 5, 7)).astype(np.float32)
    write_zarr_snapshots(tmp_path / "f.zarr", "u", data)
    eager = TimeDependentField(data, "center", "center", "inner", cyclic=True, device="cpu")
    lazy = TimeDependentField(ZarrSnapshots(tmp_path / "f.zarr", "u"), "center", "center", "inner", cyclic=True,
                              device="cpu")
    pre = TimeDependentField.from_zarr(tmp_path / "f.zarr", "u", "center", "center", "inner", preload=True,
                                       cyclic=True, device="cpu")
    assert lazy.source is not None and pre.source is None
    with pytest.raises(TypeError):
        lazy.data
    assert lazy.dtype == eager.dtype and lazy.spatial_shape == eager.spatial_shape
    for tidx in (0.25, 1.5, 4.75, 9.125, 2.0):
        a, b, c = (f.get_field_data(np.float64(tidx), _NO_BBOX).array for f in (eager, lazy, pre))
        for w in (b, c):
            assert (a.ft, a.gt, a.It) == (w.ft, w.gt, w.It)
            assert torch.equal(a.prev, w.prev) and torch.equal(a.next, w.next)
    assert len(lazy._cache) <= TimeDependentField.LAZY_CACHE
    strict = TimeDependentField(ZarrSnapshots(tmp_path / "f.zarr", "u"), "center", "center", "inner",
                                device="cpu")
    with pytest.raises(IndexError):
        strict.snapshot(6)
