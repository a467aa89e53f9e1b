"""Fused AB3 engine (split positions, FP32/BF16 fields) against the oracle.

Gate for the FP32 form: max position deviation <= 1e-5 grid-index units after
100 steps with status codes bit-exact.  BF16 storage is a precision form with
no reference counterpart; its deviation is reported and bounded loosely here
(the at-scale 1000-step figure is measured separately).
"""

import numpy as np
import pytest
import torch

from offline_particles_amd.engine import AB3Engine
from offline_particles_amd.workloads import build_fieldset, eddy_workload, turbulent_workload

from ._helpers import POS, max_dev, run_oracle

pytestmark = pytest.mark.gpu


def _engine(wl, device, *, storage=torch.float32, rebin_every=25, xlow=2, nsteps=100, positions=None,
            capacity=None, scheme="ab3"):
    fs = build_fieldset(wl, device=device)
    pos = wl.positions if positions is None else positions
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), storage_dtype=storage, device=device,
                    capacity=capacity or max(1, len(pos[0])), rebin_every=rebin_every, xlow=xlow, scheme=scheme)
    eng.add_release(*pos, nsteps=nsteps)
    return eng


def _state(eng):
    z, y, x = (t.cpu().numpy() for t in eng.positions())
    return {"zidx": z, "yidx": y, "xidx": x, "status": eng.status().cpu().numpy()}


def _turb():
    return turbulent_workload(NX=96, NY=80, NZ=12, nparticles=20_000, steps=100, nsnap_unique=4)


def test_fused_fp32_vs_oracle_100_steps(device):
    wl = _turb()
    sim_o = run_oracle(wl, "ab3", 100)
    eng = _engine(wl, device)
    eng.step(100)
    eng.check_clock()
    g = _state(eng)
    assert np.array_equal(g["status"], sim_o.p["status"]), "status differs"
    devs = {k: max_dev(g[k], sim_o.p[k]) for k in POS}
    print("fused fp32 max deviation vs oracle after 100 steps:", devs, flush=True)
    assert max(devs.values()) <= 1e-5


def test_fused_fp32_eddy_vs_oracle(device):
    wl = eddy_workload(steps=100, dtype=np.float32)
    sim_o = run_oracle(wl, "ab3", 100)
    eng = _engine(wl, device)
    eng.step(100)
    g = _state(eng)
    assert np.array_equal(g["status"], sim_o.p["status"])
    devs = {k: max_dev(g[k], sim_o.p[k]) for k in POS}
    print("fused fp32 eddy max deviation:", devs, flush=True)
    assert max(devs.values()) <= 1e-5


def test_rebin_does_not_change_results(device):
    wl = _turb()
    a = _engine(wl, device, rebin_every=0)
    b = _engine(wl, device, rebin_every=7, xlow=3)
    a.step(40)
    b.step(40)
    sa, sb = _state(a), _state(b)
    for k in ("status",) + POS:
        assert np.array_equal(sa[k], sb[k], equal_nan=True), k


def test_bf16_storage_close(device):
    wl = _turb()
    sim_o = run_oracle(wl, "ab3", 100)
    eng = _engine(wl, device, storage=torch.bfloat16)
    eng.step(100)
    g = _state(eng)
    both = (g["status"] < 128) & (sim_o.p["status"] < 128)
    devs = {k: float(np.max(np.abs(g[k][both] - sim_o.p[k][both]))) for k in POS}
    print("bf16 max deviation after 100 steps:", devs, flush=True)
    assert max(devs.values()) < 0.05


def test_releases_alone_equal_mix(device):
    """Each release gives the same result alone and stepped together with others."""
    wl = _turb()
    rng = np.random.default_rng(5)
    n = [3000, 500, 7000]
    starts = [0.0, 1234.5, 400.0]  # off the dt grid
    steps = [60, 30, 45]
    first = [0, 10, 3]
    pos = [tuple(rng.uniform(1, m - 2, k) for m in (wl.NZ, wl.NY, wl.NX)) for k in n]
    fs = build_fieldset(wl, device=device)
    mix = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=sum(n), rebin_every=9)
    for p, t0, ns, s0 in zip(pos, starts, steps, first):
        mix.add_release(*p, start_time=t0, nsteps=ns, start_step=s0)
    mix.step(70)
    mix.check_clock()
    got = _state(mix)
    off = 0
    for p, t0, ns, s0, k in zip(pos, starts, steps, first, n):
        alone = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=k, rebin_every=0)
        alone.add_release(*p, start_time=t0, nsteps=ns, start_step=0)
        alone.step(ns)
        ref = _state(alone)
        for key in ("status",) + POS:
            assert np.array_equal(got[key][off:off + k], ref[key], equal_nan=True), key
        off += k


@pytest.mark.parametrize("scheme", ["ab3", "rk2"])
def test_off_grid_start_vs_oracle(device, scheme):
    """A release starting off the dt grid (its own FP64 clock: blend weights and
    snapshot pairs at t0 + n dt, RK2's stage-2 time t + alpha dt), alone and
    stepped in a mix with a release on the grid, against the oracle started at
    the same time."""
    wl = _turb()
    rng = np.random.default_rng(21)
    pos = tuple(rng.uniform(1, m - 2, 6000) for m in (wl.NZ, wl.NY, wl.NX))
    other = tuple(rng.uniform(1, m - 2, 2000) for m in (wl.NZ, wl.NY, wl.NX))
    t0, steps = 1234.5, 70  # crosses the snapshot boundaries at 3600 and 7200 s
    status = 0 if scheme == "rk2" else 10
    sim_o = run_oracle(wl, scheme, steps, positions=pos, status=np.full(6000, status, np.uint8), time=t0)
    fs = build_fieldset(wl, device=device)
    for mixed in (False, True):
        eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=8000, rebin_every=11,
                        scheme=scheme)
        eng.launch_steps = 8
        if mixed:
            eng.add_release(*other, start_time=0.0, nsteps=steps + 5)
        r = eng.add_release(*pos, start_time=t0, nsteps=steps, start_step=3 if mixed else 0)
        eng.step(steps + (3 if mixed else 0))
        eng.check_clock()
        sl = eng.release_slice(r)
        g = dict(zip(POS, (t[sl].cpu().numpy() for t in eng.positions())))
        assert np.array_equal(eng.status()[sl].cpu().numpy(), sim_o.p["status"]), (scheme, mixed)
        devs = {k: max_dev(g[k], sim_o.p[k]) for k in POS}
        print(f"{scheme} release at t0 = {t0} s ({'mixed' if mixed else 'alone'}): max deviation {devs}", flush=True)
        assert max(devs.values()) <= 1e-5


def test_clock_leaves_time_array(device):
    wl = eddy_workload(steps=4, nparticles=100, dtype=np.float32)
    eng = _engine(wl, device, nsteps=1000)
    eng.step(len(wl.time_array) * 40)
    with pytest.raises(ValueError):
        eng.check_clock()


@pytest.mark.parametrize("storage", [torch.float32, torch.bfloat16])
def test_stencil_layouts_identical(device, storage):
    """RING, PAIR, QUAD, OCT, YPAIR and LDS-staged PAIR stencils blend the same values in the same order:
    identical bits."""
    from offline_particles_amd import _lib

    wl = _turb()
    res = []
    for layout in (_lib.OPX_LAYOUT_RING, _lib.OPX_LAYOUT_PAIR, _lib.OPX_LAYOUT_QUAD, _lib.OPX_LAYOUT_PAIR_LDS,
                   _lib.OPX_LAYOUT_OCT, _lib.OPX_LAYOUT_YPAIR):
        eng = _engine(wl, device, storage=storage, rebin_every=11)
        eng.set_layout(layout)
        eng.step(100)  # crosses snapshot boundaries (hourly snapshots, dt = 120 s)
        res.append(_state(eng))
    for other in res[1:]:
        for k in ("status",) + POS:
            assert np.array_equal(res[0][k], other[k], equal_nan=True), k


def test_graph_replays_match_eager(device):
    """Steps replayed from captured HIP graphs give the eager result after every replay.

    Releases are added between replays and the runs cross snapshot boundaries
    (the packed windows change between replays)."""
    wl = _turb()
    rng = np.random.default_rng(11)
    pos = [tuple(rng.uniform(1, m - 2, k) for m in (wl.NZ, wl.NY, wl.NX)) for k in (4000, 2500)]
    fs = build_fieldset(wl, device=device)
    engines = []
    for graph_steps in (0, 6):
        eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=6500, rebin_every=13,
                        max_releases=4)
        eng.graph_steps = graph_steps
        eng.add_release(*pos[0], start_time=10.0, nsteps=90)
        engines.append(eng)
    for chunk in range(8):
        if chunk == 2:
            for eng in engines:
                eng.add_release(*pos[1], start_time=3000.0, nsteps=40, start_step=eng.global_step + 1)
        for eng in engines:
            eng.step(12)
        torch.cuda.synchronize()
        a, b = _state(engines[0]), _state(engines[1])
        for k in ("status",) + POS:
            assert np.array_equal(a[k], b[k], equal_nan=True), (chunk, k)
    assert len(engines[1]._graphs) > 0


# ---------------------------------------------------------------- fused RK2

def test_fused_rk2_fp32_vs_oracle_100_steps(device):
    wl = _turb()
    sim_o = run_oracle(wl, "rk2", 100)
    eng = _engine(wl, device, scheme="rk2")
    eng.step(100)
    eng.check_clock()
    g = _state(eng)
    assert np.array_equal(g["status"], sim_o.p["status"]), "status differs"
    devs = {k: max_dev(g[k], sim_o.p[k]) for k in POS}
    print("fused rk2 fp32 max deviation vs oracle after 100 steps:", devs, flush=True)
    assert max(devs.values()) <= 1e-5


def test_fused_rk2_eddy_vs_oracle(device):
    """RK2 on the eddy case (dt = 300 s, snapshots 3600 s apart: stage 2 and
    the update cross snapshot boundaries)."""
    wl = eddy_workload(steps=100, dtype=np.float32)
    sim_o = run_oracle(wl, "rk2", 100)
    eng = _engine(wl, device, scheme="rk2")
    eng.step(100)
    g = _state(eng)
    assert np.array_equal(g["status"], sim_o.p["status"])
    devs = {k: max_dev(g[k], sim_o.p[k]) for k in POS}
    print("fused rk2 eddy max deviation:", devs, flush=True)
    assert max(devs.values()) <= 1e-5


@pytest.mark.parametrize("storage", [torch.float32, torch.bfloat16])
def test_fused_rk2_layouts_and_graphs_identical(device, storage):
    """RK2: every stencil layout, re-binning and graph replays give identical bits."""
    from offline_particles_amd import _lib

    wl = _turb()
    res = []
    for layout, graph_steps, rebin in ((_lib.OPX_LAYOUT_RING, 0, 0), (_lib.OPX_LAYOUT_PAIR, 5, 11),
                                       (_lib.OPX_LAYOUT_QUAD, 8, 7), (_lib.OPX_LAYOUT_OCT, 3, 5),
                                       (_lib.OPX_LAYOUT_YPAIR, 4, 9)):
        eng = _engine(wl, device, storage=storage, rebin_every=rebin, scheme="rk2")
        eng.set_layout(layout)
        eng.graph_steps = graph_steps
        eng.step(70)
        res.append(_state(eng))
    for other in res[1:]:
        for k in ("status",) + POS:
            assert np.array_equal(res[0][k], other[k], equal_nan=True), k


@pytest.mark.parametrize("storage", [torch.float32, torch.bfloat16])
def test_lds_tiles_staged_and_fallback(device, storage):
    """The LDS-tiled step matches the global-read step bit for bit both on waves
    whose stencil boxes fit the LDS pool (sorted particles) and on waves that
    fall back to global reads (unsorted particles spread over the grid), with
    particles leaving the domain and a snapshot boundary in the run."""
    from offline_particles_amd import _lib

    wl = _turb()
    for rebin in (9, 0):
        res = []
        for layout in (_lib.OPX_LAYOUT_PAIR, _lib.OPX_LAYOUT_PAIR_LDS):
            eng = _engine(wl, device, storage=storage, rebin_every=rebin, xlow=1)
            eng.set_layout(layout)
            eng.graph_steps = 4
            eng.step(45)
            res.append(_state(eng))
        for k in ("status",) + POS:
            assert np.array_equal(res[0][k], res[1][k], equal_nan=True), (rebin, k)


@pytest.mark.parametrize("scheme", ["ab3", "rk2"])
@pytest.mark.parametrize("storage", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("layout", ["ring", "pair", "quad", "oct", "ypair"])
def test_several_steps_per_launch_identical(device, storage, layout, scheme):
    """k AB3 or RK2 steps per launch (state in registers, unchanged stencils blended again) give the bits of
    one step per launch: across snapshot boundaries, re-bins and graph replays, with releases that start,
    end and cross a time-array edge inside a launch, and particles that leave the domain."""
    from offline_particles_amd.engine import LAYOUT_NAMES

    code = {v: k for k, v in LAYOUT_NAMES.items()}[layout]
    wl = _turb()
    rng = np.random.default_rng(11)
    n = [6000, 900, 3000]
    starts = [0.0, 1234.5, 400.0]  # off the dt grid
    steps = [95, 30, 41]
    first = [0, 13, 5]
    pos = [tuple(rng.uniform(0.5, m - 1.5, k) for m in (wl.NZ, wl.NY, wl.NX)) for k in n]
    # a release at the lateral edges (some drift out mid-launch) with one non-finite position
    m_edge = 800
    ze = rng.uniform(0.5, wl.NZ - 1.5, m_edge)
    ye = np.where(rng.random(m_edge) < 0.5, 0.01, wl.NY - 1.01)
    xe = np.where(rng.random(m_edge) < 0.5, 0.01, wl.NX - 1.01)
    ye[::2] = rng.uniform(0.5, wl.NY - 1.5, m_edge // 2)
    xe[0] = np.nan
    pos.append((ze, ye, xe))
    n, starts, steps, first = n + [m_edge], starts + [0.0], steps + [90], first + [2]
    fs = build_fieldset(wl, device=device)
    res = []
    for launch_steps, graph_steps in ((1, 0), (3, 0), (8, 0), (8, 8), (16, 5), (64, 0)):
        eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), storage_dtype=storage, device=device,
                        capacity=sum(n), rebin_every=23, scheme=scheme)
        eng.set_layout(code)
        for p, t0, ns, s0 in zip(pos, starts, steps, first):
            eng.add_release(*p, start_time=t0, nsteps=ns, start_step=s0)
        eng.launch_steps = launch_steps
        eng.graph_steps = graph_steps
        eng.step(100)
        eng.check_clock()
        res.append(_state(eng))
    assert (res[0]["status"] == 130).sum() > 10, "too few particles left the domain"
    assert (res[0]["status"] == 129).any(), "the non-finite particle was not flagged"
    for other in res[1:]:
        for k in ("status",) + POS:
            assert np.array_equal(res[0][k], other[k], equal_nan=True), k
