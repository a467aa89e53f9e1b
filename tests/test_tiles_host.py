"""Host logic of the tiled snapshot streaming: tile dilation and copy boxes (CPU only)."""

import numpy as np
import pytest

from offline_particles_amd.engine import dilate_tiles, tile_boxes


def _brute_dilate(mask, d):
    out = np.zeros_like(mask)
    for i, j in np.argwhere(mask):
        out[max(0, i - d): i + d + 1, max(0, j - d): j + d + 1] = True
    return out


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("d", [0, 1, 2, 5, 40])
def test_dilate_tiles_matches_brute_force(seed, d):
    rng = np.random.default_rng(seed)
    mask = rng.random((rng.integers(1, 30), rng.integers(1, 40))) < 0.05
    assert np.array_equal(dilate_tiles(mask, d), _brute_dilate(mask, d))


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("gap", [0, 1, 3])
def test_tile_boxes_cover_exactly(seed, gap):
    """The boxes cover every set tile, nothing outside the set tiles and the bridged
    gaps (runs within a tile row at most `gap` tiles apart), do not overlap, stay in
    the grid, and `covered` is the union of the boxes."""
    rng = np.random.default_rng(seed)
    T = int(rng.integers(3, 40))
    ny, nx = int(rng.integers(1, 300)), int(rng.integers(1, 300))
    nty, ntx = -(-ny // T), -(-nx // T)
    mask = rng.random((nty, ntx)) < rng.uniform(0.02, 0.6)
    boxes, covered = tile_boxes(mask, T, ny, nx, gap)
    assert boxes.dtype == np.int32 and boxes.shape[1] == 4
    cells = np.zeros((ny, nx), np.int32)
    for y0, y1, x0, x1 in boxes:
        assert 0 <= y0 < y1 <= ny and 0 <= x0 < x1 <= nx
        cells[y0:y1, x0:x1] += 1
    assert cells.max(initial=0) <= 1  # no overlap
    # the covered tiles, expanded to cells (clipped at the grid edge), are exactly the boxes
    want = np.kron(covered, np.ones((T, T), np.int32))[:ny, :nx]
    assert np.array_equal(cells, want)
    assert (covered | ~mask).all()  # every set tile covered
    # outside the set tiles only gaps of at most `gap` tiles between two runs of a row
    extra = covered & ~mask
    for ty, tx in np.argwhere(extra):
        row = mask[ty]
        left = np.flatnonzero(row[:tx])
        right = np.flatnonzero(row[tx + 1:])
        assert left.size and right.size
        assert (tx + 1 + right[0]) - left[-1] - 1 <= gap


def _loop_tile_boxes(mask, tile, ny, nx, gap):
    """Row by row: the runs of each tile row (bridging gaps of at most `gap`), then
    each run stacked over the consecutive rows that hold the same run."""
    runs_of = []
    covered = np.zeros_like(mask)
    for row in mask:
        runs, j = [], 0
        while j < row.size:
            if not row[j]:
                j += 1
                continue
            k = j
            while k < row.size and row[k]:
                k += 1
            if runs and j - runs[-1][1] <= gap:
                runs[-1] = (runs[-1][0], k)
            else:
                runs.append((j, k))
            j = k
        runs_of.append(runs)
    boxes = []
    for ty, runs in enumerate(runs_of):
        for a, b in runs:
            covered[ty, a:b] = True
            if ty and (a, b) in runs_of[ty - 1]:
                continue  # part of the box opened in an earlier row
            end = ty + 1
            while end < len(runs_of) and (a, b) in runs_of[end]:
                end += 1
            boxes.append((ty * tile, min(end * tile, ny), a * tile, min(b * tile, nx)))
    return np.array(sorted(boxes), dtype=np.int32).reshape(-1, 4), covered


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("gap", [0, 1, 2, 4])
def test_tile_boxes_equal_row_by_row_construction(seed, gap):
    """The array form gives the row-by-row construction's boxes, in the same order,
    and the same covered mask, on random masks and on clustered ones (patches of
    tiles, as the particles of a request occupy them)."""
    rng = np.random.default_rng(100 + seed)
    nty, ntx = int(rng.integers(1, 60)), int(rng.integers(1, 70))
    if seed % 2:
        mask = rng.random((nty, ntx)) < rng.uniform(0.02, 0.7)
    else:
        mask = np.zeros((nty, ntx), bool)
        for _ in range(int(rng.integers(0, 20))):
            y, x = int(rng.integers(0, nty)), int(rng.integers(0, ntx))
            mask[y:y + int(rng.integers(1, 5)), x:x + int(rng.integers(1, 5))] = True
    T = int(rng.integers(2, 33))
    ny = int(rng.integers((nty - 1) * T + 1, nty * T + 1))
    nx = int(rng.integers((ntx - 1) * T + 1, ntx * T + 1))
    boxes, covered = tile_boxes(mask, T, ny, nx, gap)
    want_boxes, want_covered = _loop_tile_boxes(mask, T, ny, nx, gap)
    assert boxes.dtype == np.int32
    assert np.array_equal(boxes, want_boxes)
    assert np.array_equal(covered, want_covered)
    assert covered is not mask


def test_tile_boxes_stack_rows_and_clip():
    m = np.zeros((5, 6), bool)
    m[1, 1] = m[2, 1] = True  # the same run in two rows: one box
    m[1, 4] = True
    m[4, 5] = True  # the last tile, clipped to the grid
    boxes, covered = tile_boxes(m, 32, 150, 190, gap=1)
    assert boxes.tolist() == [[32, 64, 128, 160], [32, 96, 32, 64], [128, 150, 160, 190]]
    boxes, _ = tile_boxes(m, 32, 150, 190, gap=2)  # tiles 1 and 4 of row 1 bridged
    assert [32, 64, 32, 160] in boxes.tolist()
    assert tile_boxes(np.zeros((3, 3), bool), 8, 24, 24)[0].shape == (0, 4)


@pytest.mark.parametrize("scheme,reach", [("ab3", 44.0 / 12.0), ("rk2", 1.0)])
def test_step_displacement_bound_holds_on_oracle_trajectories(oracle_lib, scheme, reach):
    """The bound the tiled streaming dilates by: one step moves a particle by at most
    dt x reach x R cells horizontally, R the largest |u| / dx, |v| / dy of the
    snapshots (reach 44/12 for AB3 with its start-up, max(alpha, |b1| + |b2|) = 1
    for RK2 with alpha = 2/3).  Checked step by step on the FP64 oracle."""
    from oracle import OracleSimulation

    from offline_particles_amd.workloads import turbulent_workload

    from ._helpers import oracle_fields

    wl = turbulent_workload(NX=48, NY=40, NZ=8, nparticles=3000, steps=80, nsnap_unique=3, dtype=np.float64)
    f = wl.fields
    R = max(np.abs(f["u"]).max() / f["dx"].min(), np.abs(f["v"]).max() / f["dy"].min())
    sim = OracleSimulation(oracle_fields(wl), wl.time_array, wl.dt, scheme=scheme, nparticles=3000, hc=wl.hc,
                           NZ=wl.NZ, bounds=wl.bounds)
    for k, a in zip(("zidx", "yidx", "xidx"), wl.positions):
        sim.p[k][:] = a
    if scheme == "ab3":
        sim.p["status"][:] = 10  # MULTISTEP_1: Euler, AB2, then AB3
    worst = 0.0
    for _ in range(60):
        y0, x0 = sim.p["yidx"].copy(), sim.p["xidx"].copy()
        sim.step()
        live = (sim.p["status"] & 128) == 0
        d = np.maximum(np.abs(sim.p["yidx"] - y0), np.abs(sim.p["xidx"] - x0))[live]
        worst = max(worst, float(d.max(initial=0.0)))
    assert worst > 0.05 * wl.dt * R  # the flow moves the particles
    assert worst <= wl.dt * reach * R, (worst, wl.dt * reach * R)


@pytest.mark.parametrize("scheme", ["ab3", "rk2"])
def test_run_displacement_bound_holds_on_oracle_trajectories(oracle_lib, scheme):
    """The bound the tiled streaming widens the tile map by: m steps move a particle
    at most dt x R x (m + 4) cells horizontally from where it was for AB3 (with its
    Euler and AB2 start-up) and dt x R x (m s + alpha + s) for RK2 (s = |b1| + |b2| =
    1 for alpha = 2/3), R the largest |u| / dx, |v| / dy of the snapshots.  Checked
    over every run of m = 1..40 steps of FP64 oracle trajectories, and the engine
    widens by exactly these constants."""
    from oracle import OracleSimulation

    from offline_particles_amd.workloads import turbulent_workload

    from ._helpers import oracle_fields

    wl = turbulent_workload(NX=48, NY=40, NZ=8, nparticles=2000, steps=80, nsnap_unique=3, dtype=np.float64)
    f = wl.fields
    R = max(np.abs(f["u"]).max() / f["dx"].min(), np.abs(f["v"]).max() / f["dy"].min())
    sim = OracleSimulation(oracle_fields(wl), wl.time_array, wl.dt, scheme=scheme, nparticles=2000, hc=wl.hc,
                           NZ=wl.NZ, bounds=wl.bounds)
    for k, a in zip(("zidx", "yidx", "xidx"), wl.positions):
        sim.p[k][:] = a
    if scheme == "ab3":
        sim.p["status"][:] = 10  # MULTISTEP_1: Euler, AB2, then AB3
    ys, xs = [sim.p["yidx"].copy()], [sim.p["xidx"].copy()]
    for _ in range(40):
        sim.step()
        ys.append(sim.p["yidx"].copy())
        xs.append(sim.p["xidx"].copy())
    live = (sim.p["status"] & 128) == 0
    ys, xs = np.array(ys)[:, live], np.array(xs)[:, live]
    s, c = (1.0, 4.0) if scheme == "ab3" else (1.0, 2.0 / 3.0 + 1.0)
    worst = 0.0
    for m in range(1, 41):
        d = np.maximum(np.abs(ys[m:] - ys[:-m]), np.abs(xs[m:] - xs[:-m])).max()
        assert d <= wl.dt * R * (m * s + c), (m, d, wl.dt * R * (m * s + c))
        worst = max(worst, d / (wl.dt * R * (m * s + c)))
    assert worst > 0.05  # the flow moves the particles

    from offline_particles_amd.engine import run_reach

    assert np.allclose(run_reach(scheme, 2.0 / 3.0), (s, c))
