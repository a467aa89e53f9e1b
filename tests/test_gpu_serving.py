"""Streamed snapshot window, release retirement/extraction and the serving loop.

Gate: every result is bit-identical to the same release stepped alone with
every snapshot resident (the fused step is per particle and per release
clock, so streaming, re-binning, compaction and the mix change nothing).
"""

import numpy as np
import pytest
import torch

from offline_particles_amd.engine import AB3Engine
from offline_particles_amd.serving import ReleaseServer, Request
from offline_particles_amd.workloads import build_fieldset, turbulent_workload

from ._helpers import POS

pytestmark = pytest.mark.gpu


def _wl():
    # hourly snapshots, dt = 120 s: a snapshot boundary every 30 steps
    return turbulent_workload(NX=96, NY=80, NZ=12, nparticles=1, steps=400, nsnap_unique=4)


def _release(wl, rng, m):
    return tuple(rng.uniform(1, n - 2, m) for n in (wl.NZ, wl.NY, wl.NX))


def _alone(fs, wl, pos, *, start_time, nsteps, storage=torch.float32):
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), storage_dtype=storage, capacity=len(pos[0]),
                    rebin_every=0)
    eng.add_release(*pos, start_time=start_time, nsteps=nsteps)
    eng.step(nsteps)
    z, y, x = (t.cpu().numpy() for t in eng.positions())
    return {"zidx": z, "yidx": y, "xidx": x, "status": eng.status().cpu().numpy()}


def _same(a, b, what):
    for k in ("status",) + POS:
        assert np.array_equal(a[k], b[k], equal_nan=True), (what, k)


@pytest.mark.parametrize("storage", [torch.float32, torch.bfloat16])
def test_streamed_window_matches_resident(device, storage):
    wl = _wl()
    rng = np.random.default_rng(3)
    pos = [_release(wl, rng, 3000), _release(wl, rng, 1500)]
    res = []
    for slots in (None, 3):
        fs = build_fieldset(wl, device=device if slots is None else "cpu")
        eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), storage_dtype=storage, device=device, capacity=4500,
                        rebin_every=17, stream_slots=slots)
        eng.graph_steps = 6
        eng.add_release(*pos[0], start_time=100.0, nsteps=150)
        eng.add_release(*pos[1], start_time=300.0, nsteps=120, start_step=5)  # 400 s behind the first
        eng.step(155)
        eng.check_clock()
        z, y, x = (t.cpu().numpy() for t in eng.positions())
        res.append({"zidx": z, "yidx": y, "xidx": x, "status": eng.status().cpu().numpy()})
        if slots:
            # 150 steps x 120 s crosses 5 snapshot boundaries: each snapshot loaded once
            assert 5 <= eng.snapshot_loads <= 9, eng.snapshot_loads
    _same(res[0], res[1], "streamed vs resident")


def test_streamed_window_too_narrow_raises(device):
    wl = _wl()
    rng = np.random.default_rng(4)
    fs = build_fieldset(wl, device="cpu")
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=200, stream_slots=3)
    eng.add_release(*_release(wl, rng, 100), start_time=0.0, nsteps=10)
    eng.add_release(*_release(wl, rng, 100), start_time=4 * 3600.0, nsteps=10)  # 4 snapshots apart
    with pytest.raises(ValueError, match="streamed window"):
        eng.step(1)


def test_extract_retire_compact_reuse(device):
    wl = _wl()
    rng = np.random.default_rng(5)
    fs = build_fieldset(wl, device=device)
    pos = [_release(wl, rng, m) for m in (2000, 700, 1200, 900)]
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=4000, rebin_every=9,
                    max_releases=4)
    r0 = eng.add_release(*pos[0], start_time=0.0, nsteps=40)
    r1 = eng.add_release(*pos[1], start_time=0.0, nsteps=20)
    r2 = eng.add_release(*pos[2], start_time=0.0, nsteps=60)
    eng.step(20)
    assert eng.steps_left(r1) == 0
    z, y, x, st = (t.cpu().numpy() for t in eng.extract(r1))
    _same({"zidx": z, "yidx": y, "xidx": x, "status": st},
          _alone(fs, wl, pos[1], start_time=0.0, nsteps=20), "extract r1")
    eng.retire(r1)
    assert eng.retired_particles == 700
    eng.rebin()  # drops r1's particles, frees its id
    assert eng.n == 3200 and eng.retired_particles == 0
    t_now = 0.0
    for _ in range(20):
        t_now = t_now + wl.dt
    r3 = eng.add_release(*pos[3], start_time=t_now, nsteps=30, start_step=eng.global_step)
    assert r3 == r1  # id recycled
    eng.step(40)
    for r, p, t0, ns in ((r0, pos[0], 0.0, 40), (r2, pos[2], 0.0, 60), (r3, pos[3], t_now, 30)):
        z, y, x, st = (t.cpu().numpy() for t in eng.extract(r))
        _same({"zidx": z, "yidx": y, "xidx": x, "status": st}, _alone(fs, wl, p, start_time=t0, nsteps=ns),
              f"release {r}")
    with pytest.raises(RuntimeError):
        eng.positions()


def test_serving_results_alone_equal_mix(device):
    """Requests admitted while others run (deterministic arrival steps, a
    particle budget small enough to make requests wait for compaction) give
    each request the result it gets served alone."""
    wl = _wl()
    rng = np.random.default_rng(6)
    sizes = [1500, 300, 2200, 800, 1200, 2500, 100]
    steps = [60, 25, 90, 40, 70, 30, 55]
    arrive = [0, 3, 10, 10, 31, 45, 47]
    reqs = [Request(*_release(wl, rng, m), nsteps=ns, arrival_step=a) for m, ns, a in zip(sizes, steps, arrive)]
    fs = build_fieldset(wl, device="cpu")
    srv = ReleaseServer(fs, wl.time_array, np.float64(wl.dt), storage_dtype=torch.float32, device=device,
                        stream_slots=3, max_particles=4000, max_releases=4, rebin_every=13, graph_steps=6)
    done = srv.run(reqs, realtime=False)
    assert len(done) == len(reqs)
    st = srv.stats
    print("serving stats:", st, flush=True)
    assert st["compactions"] >= 1 and st["peak_live_particles"] <= 4000
    # some requests had to wait for slots
    assert any(r.start_step > r.arrival_step for r in reqs)
    fs_dev = build_fieldset(wl, device=device)
    for req in reqs:
        got = {k: v.numpy() for k, v in req.result.items()}
        ref = _alone(fs_dev, wl, (req.zidx, req.yidx, req.xidx), start_time=req.start_time, nsteps=req.nsteps)
        _same(got, ref, f"request of {req.size} admitted at step {req.start_step}")
        assert req.latency >= 0


def test_served_requests_vs_oracle_off_grid_clock(device):
    """A service clock started off the dt grid: every served request (FP32
    storage, streamed window) against the oracle run of its particles from the
    request's start time."""
    from ._helpers import max_dev, run_oracle

    wl = _wl()
    rng = np.random.default_rng(16)
    sizes, steps, arrive = [1200, 500, 900], [50, 35, 40], [0, 7, 19]
    reqs = [Request(*_release(wl, rng, m), nsteps=ns, arrival_step=a) for m, ns, a in zip(sizes, steps, arrive)]
    fs = build_fieldset(wl, device="cpu")
    srv = ReleaseServer(fs, wl.time_array, np.float64(wl.dt), storage_dtype=torch.float32, device=device,
                        stream_slots=3, max_particles=2000, max_releases=4, rebin_every=9, graph_steps=6,
                        start_time=777.25)
    srv.run(reqs, realtime=False)
    for req in reqs:
        assert req.start_time != round(req.start_time / wl.dt) * wl.dt
        sim_o = run_oracle(wl, "ab3", req.nsteps, positions=(req.zidx, req.yidx, req.xidx), time=req.start_time)
        got = {k: v.numpy() for k, v in req.result.items()}
        assert np.array_equal(got["status"], sim_o.p["status"]), req.start_step
        dev = max(max_dev(got[k], sim_o.p[k]) for k in POS)
        print(f"served request of {req.size} from t = {req.start_time} s: max deviation {dev:.3g}", flush=True)
        assert dev <= 1e-5


def test_serving_realtime_latency(device):
    # requests may run back to back: a time axis for 12 x 80 steps
    wl = turbulent_workload(NX=96, NY=80, NZ=12, nparticles=1, steps=1200, nsnap_unique=4)
    rng = np.random.default_rng(8)
    reqs = [Request(*_release(wl, rng, int(m)), nsteps=int(ns), arrival=float(t))
            for m, ns, t in zip(rng.integers(100, 3000, 12), rng.integers(20, 80, 12), np.sort(rng.uniform(0, 0.3, 12)))]
    fs = build_fieldset(wl, device="cpu")
    srv = ReleaseServer(fs, wl.time_array, np.float64(wl.dt), device=device, stream_slots=3,
                        max_particles=1 << 16, graph_steps=8)
    done = srv.run(reqs)
    st = srv.stats
    print("realtime serving stats:", st, flush=True)
    assert st["requests"] == 12
    assert 0 < st["latency_p50_s"] <= st["latency_p99_s"] < 60
    for r in done:
        assert r.completed_at >= r.arrival


def test_finished_release_is_parked_by_rebin(device):
    """A release whose steps are done stops moving; the next re-bin moves its
    particles behind the active ones (not launched) without changing them."""
    wl = _wl()
    rng = np.random.default_rng(9)
    fs = build_fieldset(wl, device=device)
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=3000, rebin_every=0)
    eng.add_release(*_release(wl, rng, 1000), start_time=0.0, nsteps=10)
    eng.add_release(*_release(wl, rng, 2000), start_time=0.0, nsteps=40)
    eng.step(10)
    before = [t.cpu().numpy().copy() for t in eng.extract(0)]
    eng.rebin()
    n_active = int(eng._n_active_dev.item())
    live = int((eng.status()[1000:] < 128).sum().item())
    assert n_active == live  # only the running release's active particles are launched
    eng.step(30)
    after = [t.cpu().numpy() for t in eng.extract(0)]
    for a, b in zip(before, after):
        assert np.array_equal(a, b, equal_nan=True)
