"""Sparse (tiled) snapshot streaming: a slot holds only the tiles the particles can reach.

Gate: with every tile outside the loaded ones filled with NaN (so a read of a
tile that was not loaded would make the particle non-finite), the tiled
streamed engine gives the same bits as the engine with every snapshot
resident, for every packed layout, AB3 and RK2, FP32 and BF16, with releases
added later elsewhere on the grid (so regions grow and slots are topped up).
The entry points behind it are checked against numpy: the occupancy map, the
box copies and the row-band pack.
"""

import ctypes

import numpy as np
import pytest
import torch

from offline_particles_amd import _lib
from offline_particles_amd.engine import AB3Engine
from offline_particles_amd.serving import ReleaseServer, Request
from offline_particles_amd.workloads import build_fieldset, turbulent_workload

from ._helpers import POS

pytestmark = pytest.mark.gpu


def _wl():
    # hourly snapshots, dt = 120 s: a snapshot boundary every 30 steps
    return turbulent_workload(NX=96, NY=80, NZ=12, nparticles=1, steps=400, nsnap_unique=4)


def _box(wl, rng, m, y0, x0, size=6.0):
    """m particles in a size x size cell box at (y0, x0), over the depth."""
    return (rng.uniform(1, wl.NZ - 2, m), y0 + rng.uniform(0, size, m), x0 + rng.uniform(0, size, m))


def _state(eng):
    z, y, x = (t.cpu().numpy() for t in eng.positions())
    return {"zidx": z, "yidx": y, "xidx": x, "status": eng.status().cpu().numpy()}


def _same(a, b, what):
    for k in ("status",) + POS:
        assert np.array_equal(a[k], b[k], equal_nan=True), (what, k)


def _run(wl, device, *, storage, scheme, layout, tiles, rels, steps, graph_steps=6, margin=4):
    fs = build_fieldset(wl, device=device if tiles is None else "cpu")
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), storage_dtype=storage, device=device, capacity=6000,
                    rebin_every=17, stream_slots=None if tiles is None else 3, stream_tiles=tiles, scheme=scheme)
    if layout is not None:
        eng.set_layout(layout)
    eng.graph_steps = graph_steps
    if tiles is not None:
        eng._poison_unloaded = True
        eng.tile_margin_steps = margin
        # the host runs ahead of the device here: bound how stale a map may be
        eng.max_occupancy_lag = 12
    done = 0
    for at, pos, t0, n, start in rels:
        if at > done:
            eng.step(at - done)
            done = at
        eng.add_release(*pos, start_time=t0, nsteps=n, start_step=start)
    eng.step(steps - done)
    eng.check_clock()
    return eng, _state(eng)


def _releases(wl, seed):
    rng = np.random.default_rng(seed)
    return [
        # (added at step, positions, start time, steps, start step)
        (0, _box(wl, rng, 1500, 10.0, 12.0), 100.0, 150, 0),
        (0, _box(wl, rng, 800, 55.0, 70.0), 300.0, 120, 5),
        # added mid-run on an empty part of the grid: its tiles must be loaded
        # into slots already resident
        (40, _box(wl, rng, 600, 60.0, 15.0), 100.0 + 40 * 120.0, 100, 40),
    ]


@pytest.mark.parametrize("layout", [_lib.OPX_LAYOUT_PAIR, _lib.OPX_LAYOUT_YPAIR, _lib.OPX_LAYOUT_QUAD,
                                    _lib.OPX_LAYOUT_OCT, _lib.OPX_LAYOUT_RING])
def test_tiled_stream_matches_resident_layouts(device, layout):
    wl = _wl()
    rels = _releases(wl, 11)
    _, ref = _run(wl, device, storage=torch.float32, scheme="ab3", layout=layout, tiles=None, rels=rels, steps=155)
    eng, got = _run(wl, device, storage=torch.float32, scheme="ab3", layout=layout, tiles=4, rels=rels, steps=155)
    _same(ref, got, f"tiled vs resident, layout {layout}")
    assert not np.isnan(got["xidx"]).any()
    # sparse: fewer bytes than whole snapshots; the late release forced top-ups
    assert eng.snapshot_bytes < 0.9 * eng.snapshot_full_bytes, (eng.snapshot_bytes, eng.snapshot_full_bytes)
    assert eng.tile_topups >= 1


@pytest.mark.parametrize("storage,scheme", [(torch.bfloat16, "ab3"), (torch.float32, "rk2"),
                                            (torch.bfloat16, "rk2")])
def test_tiled_stream_matches_resident_schemes(device, storage, scheme):
    wl = _wl()
    rels = _releases(wl, 12)
    _, ref = _run(wl, device, storage=storage, scheme=scheme, layout=None, tiles=None, rels=rels, steps=150)
    eng, got = _run(wl, device, storage=storage, scheme=scheme, layout=None, tiles=4, rels=rels, steps=150)
    _same(ref, got, f"tiled vs resident, {storage} {scheme}")
    assert eng.snapshot_bytes < 0.9 * eng.snapshot_full_bytes, (eng.snapshot_bytes, eng.snapshot_full_bytes)
    assert eng.tile_topups >= 1


def test_tiled_stream_lagging_maps_and_wide_tiles(device):
    """Maps forced to be waited for (lag bound 1 step) and tiles wider than the grid's
    clouds (16 cells) give the same bits."""
    wl = _wl()
    rels = _releases(wl, 13)
    _, ref = _run(wl, device, storage=torch.float32, scheme="ab3", layout=None, tiles=None, rels=rels, steps=120)
    fs = build_fieldset(wl, device="cpu")
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=6000, rebin_every=17,
                    stream_slots=3, stream_tiles=16)
    eng._poison_unloaded = True
    eng.max_occupancy_lag = 1
    eng.launch_steps = 8
    done = 0
    for at, pos, t0, n, start in rels:
        if at > done:
            eng.step(at - done)
            done = at
        eng.add_release(*pos, start_time=t0, nsteps=n, start_step=start)
    eng.step(120 - done)
    _same(ref, _state(eng), "lagging maps, 16-cell tiles")


def test_tiled_serving_alone_equals_mix(device):
    """Served requests with tiled streaming equal each request run alone with every
    snapshot resident."""
    wl = _wl()
    rng = np.random.default_rng(14)
    reqs = [Request(*_box(wl, rng, m, y0, x0), nsteps=n, arrival_step=a)
            for m, y0, x0, n, a in ((700, 8.0, 10.0, 60, 0), (500, 60.0, 70.0, 45, 7), (900, 30.0, 40.0, 80, 13),
                                    (300, 65.0, 12.0, 50, 31))]
    fs = build_fieldset(wl, device="cpu")
    srv = ReleaseServer(fs, wl.time_array, np.float64(wl.dt), storage_dtype=torch.float32, device=device,
                        stream_slots=3, max_particles=4000, max_releases=4, rebin_every=13, graph_steps=6,
                        stream_tiles=8)
    srv.engine._poison_unloaded = True
    srv.engine.tile_margin_steps = 4
    done = srv.run(reqs, realtime=False)
    assert len(done) == len(reqs)
    assert srv.stats["snapshot_bytes"] < srv.stats["snapshot_full_bytes"]
    fs_dev = build_fieldset(wl, device=device)
    for r in done:
        eng = AB3Engine(fs_dev, wl.time_array, np.float64(wl.dt), device=device, capacity=r.size, rebin_every=0)
        eng.add_release(r.zidx, r.yidx, r.xidx, start_time=r.start_time, nsteps=r.nsteps)
        eng.step(r.nsteps)
        alone = _state(eng)
        got = {k: r.result[k].numpy() for k in ("status",) + POS}
        _same(alone, got, f"request of {r.size} particles")


# ---------------------------------------------------------------- the entry points

def test_tile_occupancy_matches_numpy(device):
    wl = _wl()
    rng = np.random.default_rng(15)
    fs = build_fieldset(wl, device=device)
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=5000)
    z = rng.uniform(0, wl.NZ - 1, 5000)
    y = rng.uniform(-1.5, wl.NY + 0.5, 5000)  # some outside the grid: edge tiles
    x = rng.uniform(-1.5, wl.NX + 0.5, 5000)
    x[::97] = np.nan  # non-finite: not counted
    eng.add_release(z, y, x, nsteps=1)
    st = eng.state
    st.status[: 4000: 5] = 128  # INACTIVE: not counted
    T = 7
    nty, ntx = -(-wl.NY // T), -(-wl.NX // T)
    out = torch.full((nty * ntx,), 9, dtype=torch.uint8, device=device)
    _lib.call("opx_tile_occupancy", ctypes.byref(st.struct(eng.n)), ctypes.byref(eng.grid), T, out.data_ptr(), nty,
              ntx, torch.cuda.current_stream().cuda_stream)
    got = out.cpu().numpy().reshape(nty, ntx)
    keep = np.ones(5000, bool)
    keep[: 4000: 5] = False
    keep &= np.isfinite(x)
    # the stored cells: floor of the index, clamped to [-2, N + 1], then to the grid
    cy = np.clip(np.clip(np.floor(y[keep]), -2, wl.NY + 1), 0, wl.NY - 1).astype(int)
    cx = np.clip(np.clip(np.floor(x[keep]), -2, wl.NX + 1), 0, wl.NX - 1).astype(int)
    want = np.zeros((nty, ntx), np.uint8)
    want[cy // T, cx // T] = 1
    assert np.array_equal(got, want)


@pytest.mark.parametrize("fn", ["opx_copy_boxes_h2d", "opx_gather_boxes_h2d"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16, torch.float64])
def test_copy_boxes_matches_numpy(device, dtype, fn):
    rng = np.random.default_rng(16)
    nz, ny, nx = 5, 37, 53
    host = torch.from_numpy(rng.standard_normal((nz, ny, nx)).astype(np.float32)).to(dtype).pin_memory()
    dev = torch.zeros((nz, ny, nx), dtype=dtype, device=device)
    # 200 boxes: more than one launch of the gather kernel (128 boxes each)
    boxes = [[0, 4, 0, 53], [10, 37, 5, 6], [20, 30, 40, 53], [36, 37, 0, 1]]
    boxes += [[5 + i % 5, 6 + i % 5, i % 50, i % 50 + 3] for i in range(196)]
    boxes = np.array(boxes, dtype=np.int32)
    _lib.call(fn, dev.data_ptr(), host.data_ptr(), nz, ny, nx, host.element_size(), boxes.ctypes.data, len(boxes),
              torch.cuda.current_stream().cuda_stream)
    got = dev.cpu()
    want = torch.zeros_like(host)
    for y0, y1, x0, x1 in boxes:
        want[:, y0:y1, x0:x1] = host[:, y0:y1, x0:x1]
    assert torch.equal(got, want)
    with pytest.raises(_lib.OpxInvalidArgument):
        bad = np.array([[0, 38, 0, 1]], dtype=np.int32)
        _lib.call(fn, dev.data_ptr(), host.data_ptr(), nz, ny, nx, host.element_size(), bad.ctypes.data, 1,
                  torch.cuda.current_stream().cuda_stream)


def test_gather_boxes_refuses_unmapped_source(device):
    """Pageable host memory has no device mapping: refused before any launch."""
    host = torch.zeros((2, 8, 8), dtype=torch.bfloat16)  # not pinned
    dev = torch.zeros((2, 8, 8), dtype=torch.bfloat16, device=device)
    boxes = np.array([[0, 8, 0, 8]], dtype=np.int32)
    with pytest.raises(_lib.OpxInvalidArgument, match="not mapped"):
        _lib.call("opx_gather_boxes_h2d", dev.data_ptr(), host.data_ptr(), 2, 8, 8, 2, boxes.ctypes.data, 1,
                  torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("storage", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("layout", [_lib.OPX_LAYOUT_PAIR, _lib.OPX_LAYOUT_QUAD, _lib.OPX_LAYOUT_OCT,
                                    _lib.OPX_LAYOUT_YPAIR])
def test_pack_window_rows_equals_full_pack(device, storage, layout):
    """Packing rows [y0, y1) writes exactly those rows of the full pack (each field's
    rows in the range, every level) and nothing else."""
    wl = _wl()
    fs = build_fieldset(wl, device=device)
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), storage_dtype=storage, device=device, capacity=10)
    w, g = eng.window, eng.grid
    sizes = [-(-b // 256) * 256 for b in eng.packed_window_bytes(layout)]
    stream = torch.cuda.current_stream().cuda_stream
    full = [torch.zeros(n, dtype=torch.uint8, device=device) for n in sizes]
    part = [torch.zeros(n, dtype=torch.uint8, device=device) for n in sizes]
    _lib.call("opx_pack_window", ctypes.byref(w), ctypes.byref(g), layout, 1, 2, *(b.data_ptr() for b in full),
              stream)
    y0, y1 = 23, 61
    _lib.call("opx_pack_window_rows", ctypes.byref(w), ctypes.byref(g), layout, 1, 2, y0, y1,
              *(b.data_ptr() for b in part), stream)
    NZ, NY, NX = wl.NZ, wl.NY, wl.NX
    shapes = [(NZ, NY, NX - 1), (NZ, NY - 1, NX), (NZ + 1, NY, NX), (1, NY, NX)]
    for f, p, shp, nbytes in zip(full, part, shapes, eng.packed_window_bytes(layout)):
        nodes = shp[0] * shp[1] * shp[2]
        per = nbytes // nodes
        F = f[: nbytes].cpu().numpy().reshape(shp[0], shp[1], shp[2], per)
        P = p[: nbytes].cpu().numpy().reshape(shp[0], shp[1], shp[2], per)
        r1 = min(y1, shp[1])
        assert np.array_equal(P[:, y0:r1], F[:, y0:r1])
        assert not P[:, :y0].any() and not P[:, r1:].any()
    with pytest.raises(_lib.OpxInvalidArgument):
        _lib.call("opx_pack_window_rows", ctypes.byref(w), ctypes.byref(g), layout, 1, 2, 5, 5,
                  *(b.data_ptr() for b in part), stream)
