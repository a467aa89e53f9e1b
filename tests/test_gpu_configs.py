"""The workload configurations at their full sizes, on the GPU.

1024 x 768 x 50 grid, 8 hourly FP32 snapshots cycled (turbulent k^-3 flow,
2% land), AB3 from MULTISTEP_1, dt = 120 s:

- plain FP32 drop-in, 4M particles x 100 steps, AB3 from MULTISTEP_1 and AB3
  with an RK2 start-up (the configuration's "AB3 with RK2 startup"): a seeded
  100k subsample equals the oracle's run of those 100k bit for bit, status
  included;
- fused FP32, 16M particles x 100 steps (32 steps per launch, re-binning):
  the subsample is within 1e-5 grid-index units of the oracle with status
  bit-exact, and equal bit for bit to the fused run of those 100k alone
  (particles are independent, so the 16M run must give each the same bits);
- fused BF16, 16M: the subsample equals the 100k-alone BF16 run bit for bit
  (its deviation from the oracle is printed);
- the 256-release batching mix (sizes 1k-1M, starts over 24 h, 6-72 h,
  10 km boxes), BF16, stepped together from graphs: the smallest, the
  largest, the one starting farthest off the dt grid and one ending inside a
  launch each equal the release run alone.

2048 x 1536 x 60 grid, BF16 fields streamed through a 3-snapshot window:
requests with deterministic arrival steps served with a particle budget that
makes some wait; each equals the request stepped alone with every snapshot
resident.

Workloads are generated once per module (about 11 s and 36 s of host time).
"""

import time

import numpy as np
import pytest
import torch

from offline_particles_amd.engine import AB3Engine
from offline_particles_amd.workloads import batched_releases, build_fieldset, build_simulation, turbulent_workload

from ._helpers import POS, max_dev, run_oracle

pytestmark = pytest.mark.gpu

N_PLAIN = 4_000_000
N_FUSED = 16_000_000
N_SUB = 100_000
STEPS = 100


def _log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def grid1024():
    t = time.perf_counter()
    # a time axis long enough for the batching mix (24 h of starts + up to 72 h)
    wl = turbulent_workload(nparticles=N_FUSED, steps=int(96 * 3600 / 120) + 40, dtype=np.float32)
    rng = np.random.default_rng(2024)
    sub = np.sort(rng.choice(N_PLAIN, N_SUB, replace=False))
    _log(f"1024x768x50 workload with {N_FUSED} particles in {time.perf_counter() - t:.1f}s")
    return wl, sub


_ORACLE_RUNS = {}


def _oracle_sub(grid1024, scheme="ab3"):
    if scheme not in _ORACLE_RUNS:
        wl, sub = grid1024
        t = time.perf_counter()
        sim = run_oracle(wl, scheme, STEPS, positions=tuple(a[sub] for a in wl.positions))
        _log(f"oracle: {N_SUB} particles x {STEPS} steps ({scheme}) in {time.perf_counter() - t:.1f}s")
        _ORACLE_RUNS[scheme] = {k: sim.p[k].copy() for k in ("status",) + POS}
    return _ORACLE_RUNS[scheme]


@pytest.fixture(scope="module")
def oracle_sub(grid1024):
    yield _oracle_sub(grid1024)
    _ORACLE_RUNS.clear()


def _fused(wl, device, pos, storage, *, launch_steps=32, rebin_every=32):
    fs = build_fieldset(wl, device=device)
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), storage_dtype=storage, device=device,
                    capacity=len(pos[0]), rebin_every=rebin_every, xlow=4)
    eng.launch_steps = launch_steps
    eng.add_release(*pos, nsteps=STEPS)
    eng.step(STEPS)
    eng.check_clock()
    z, y, x = (t.cpu().numpy() for t in eng.positions())
    out = {"zidx": z, "yidx": y, "xidx": x, "status": eng.status().cpu().numpy()}
    del eng, fs
    torch.cuda.empty_cache()
    return out


def _same(a, b, what, idx=None):
    for k in ("status",) + POS:
        got = a[k] if idx is None else a[k][idx]
        assert np.array_equal(got, b[k], equal_nan=True), (what, k)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("scheme", ["ab3", "ab3_rk2"])
def test_plain_fp32_4M_subsample_bit_exact(device, grid1024, oracle_sub, scheme):
    """AB3 from MULTISTEP_1, and AB3 whose first two steps are RK2 steps."""
    wl, sub = grid1024
    oracle_sub = _oracle_sub(grid1024, scheme)
    pos = tuple(a[:N_PLAIN] for a in wl.positions)
    sim = build_simulation(wl, scheme, device=device, positions=pos)
    t = time.perf_counter()
    for _ in range(STEPS):
        sim.step()
    torch.cuda.synchronize()
    _log(f"plain FP32 drop-in ({scheme}): {N_PLAIN} particles x {STEPS} steps in {time.perf_counter() - t:.2f}s")
    g = {k: sim.particles.to_numpy(k) for k in ("status",) + POS}
    del sim
    torch.cuda.empty_cache()
    assert np.array_equal(g["status"][sub], oracle_sub["status"]), "status differs"
    devs = {k: max_dev(g[k][sub], oracle_sub[k]) for k in POS}
    _log(f"plain FP32 4M ({scheme}): max deviation of the 100k subsample from the oracle:", devs)
    assert max(devs.values()) == 0.0
    # the run moved the particles and some left the domain
    assert np.mean(np.abs(g["xidx"][sub] - pos[2][sub])) > 0.1


@pytest.mark.timeout(300)
def test_fused_fp32_16M_subsample(device, grid1024, oracle_sub):
    wl, sub = grid1024
    full = _fused(wl, device, wl.positions, torch.float32)
    assert np.array_equal(full["status"][sub], oracle_sub["status"]), "status differs"
    devs = {k: max_dev(full[k][sub], oracle_sub[k]) for k in POS}
    _log("fused FP32 16M: max deviation of the 100k subsample from the oracle:", devs)
    assert max(devs.values()) <= 1e-5
    alone = _fused(wl, device, tuple(a[sub] for a in wl.positions), torch.float32, launch_steps=8, rebin_every=0)
    _same(full, alone, "16M run vs the subsample alone", sub)


@pytest.mark.timeout(300)
def test_fused_bf16_16M_subsample(device, grid1024, oracle_sub):
    wl, sub = grid1024
    full = _fused(wl, device, wl.positions, torch.bfloat16)
    alone = _fused(wl, device, tuple(a[sub] for a in wl.positions), torch.bfloat16, launch_steps=1, rebin_every=0)
    _same(full, alone, "16M BF16 run vs the subsample alone", sub)
    both = (full["status"][sub] < 128) & (oracle_sub["status"] < 128)
    d = max(float(np.max(np.abs(full[k][sub][both] - oracle_sub[k][both]))) for k in POS)
    _log(f"fused BF16 16M: max deviation of the 100k subsample from the oracle after {STEPS} steps: {d:.3g}")
    assert d < 0.05


@pytest.mark.timeout(300)
def test_batching_256_releases_mix_equals_alone(device, grid1024):
    wl, _ = grid1024
    rels = batched_releases(wl, nrel=256, size_range=(1_000, 1_000_000))
    dt = wl.dt
    start_step = [int(np.floor(r.start_time / dt)) for r in rels]
    ends = [s + r.nsteps for s, r in zip(start_step, rels)]
    total = sum(r.size for r in rels)
    launch_steps = 16
    fs = build_fieldset(wl, device=device)
    mix = AB3Engine(fs, wl.time_array, np.float64(dt), storage_dtype=torch.bfloat16, device=device, capacity=total,
                    max_releases=len(rels), rebin_every=64, xlow=4)
    mix.launch_steps = launch_steps
    mix.graph_steps = 8
    order = np.argsort(start_step, kind="stable")
    last = max(ends)
    t = time.perf_counter()
    nxt, g, ids = 0, 0, {}
    while g < last:
        while nxt < len(rels) and start_step[order[nxt]] <= g:
            i = int(order[nxt])
            r = rels[i]
            ids[i] = mix.add_release(*r.positions, start_time=r.start_time, nsteps=r.nsteps, start_step=start_step[i])
            nxt += 1
        upcoming = start_step[order[nxt]] if nxt < len(rels) else last
        k = max(1, min(upcoming, last) - g)
        mix.step(k)
        g += k
    mix.check_clock()
    torch.cuda.synchronize()
    _log(f"batching mix: {len(rels)} releases, {total} particles, {last} global steps in "
         f"{time.perf_counter() - t:.1f}s")
    sizes = [r.size for r in rels]
    off_grid = [abs(r.start_time / dt - round(r.start_time / dt)) for r in rels]
    boundaries = set(start_step)
    inside = [i for i in range(len(rels)) if ends[i] % launch_steps and ends[i] not in boundaries]
    picks = {"smallest": int(np.argmin(sizes)), "largest": int(np.argmax(sizes)),
             "off-grid start": int(np.argmax(off_grid)), "ends inside a launch": inside[len(inside) // 2]}
    for what, i in picks.items():
        r = rels[i]
        sl = mix.release_slice(ids[i])
        z, y, x = (t_[sl].cpu().numpy() for t_ in mix.positions())
        got = {"zidx": z, "yidx": y, "xidx": x, "status": mix.status()[sl].cpu().numpy()}
        alone = AB3Engine(fs, wl.time_array, np.float64(dt), storage_dtype=torch.bfloat16, device=device,
                          capacity=r.size, rebin_every=0)
        alone.launch_steps = 8
        alone.add_release(*r.positions, start_time=r.start_time, nsteps=r.nsteps)
        alone.step(r.nsteps)
        alone.check_clock()
        za, ya, xa = (t_.cpu().numpy() for t_ in alone.positions())
        ref = {"zidx": za, "yidx": ya, "xidx": xa, "status": alone.status().cpu().numpy()}
        _same(got, ref, f"{what} release ({r.size} particles, start {r.start_time:.1f} s, {r.nsteps} steps)")
        _log(f"batching: {what} release {i} ({r.size} particles, start {r.start_time:.1f} s, {r.nsteps} steps, "
             f"ends at global step {ends[i]}) equals its run alone")
        del alone
    del mix, fs
    torch.cuda.empty_cache()


@pytest.fixture(scope="module")
def grid2048():
    t = time.perf_counter()
    wl = turbulent_workload(NX=2048, NY=1536, NZ=60, nparticles=1, steps=400, nsnap_unique=4, dtype=np.float32)
    _log(f"2048x1536x60 workload in {time.perf_counter() - t:.1f}s")
    return wl


@pytest.mark.timeout(300)
def test_serving_2048_grid_alone_equals_mix(device, grid2048):
    from offline_particles_amd.serving import ReleaseServer, Request

    wl = grid2048
    rng = np.random.default_rng(12)
    sizes = [2_000_000, 1_000, 350_000, 60_000, 1_200_000, 5_000]
    steps = [90, 40, 120, 75, 60, 100]
    arrive = [0, 0, 8, 20, 33, 61]

    def seed(m):
        y0, x0 = rng.uniform(1, wl.NY - 12), rng.uniform(1, wl.NX - 12)
        return (rng.uniform(0.5, wl.NZ - 1.5, m), y0 + rng.uniform(0, 10, m), x0 + rng.uniform(0, 10, m))

    reqs = [Request(*seed(m), nsteps=ns, arrival_step=a) for m, ns, a in zip(sizes, steps, arrive)]
    t = time.perf_counter()
    fs_host = build_fieldset(wl, device="cpu", storage_dtype=torch.bfloat16)
    srv = ReleaseServer(fs_host, wl.time_array, np.float64(wl.dt), storage_dtype=torch.bfloat16, device=device,
                        stream_slots=3, max_particles=3_000_000, max_releases=4, rebin_every=24, graph_steps=8,
                        launch_steps=8)
    done = srv.run(reqs, realtime=False)
    _log(f"served {len(done)} requests on 2048x1536x60 in {time.perf_counter() - t:.1f}s:", srv.stats)
    assert len(done) == len(reqs)
    assert any(r.start_step > r.arrival_step for r in reqs), "no request waited for particle slots"
    del srv, fs_host
    torch.cuda.empty_cache()
    fs = build_fieldset(wl, device=device, storage_dtype=torch.bfloat16)
    for req in reqs:
        alone = AB3Engine(fs, wl.time_array, np.float64(wl.dt), storage_dtype=torch.bfloat16, device=device,
                          capacity=req.size, rebin_every=0)
        alone.launch_steps = 8
        alone.add_release(req.zidx, req.yidx, req.xidx, start_time=req.start_time, nsteps=req.nsteps)
        alone.step(req.nsteps)
        alone.check_clock()
        z, y, x = (t_.cpu().numpy() for t_ in alone.positions())
        ref = {"zidx": z, "yidx": y, "xidx": x, "status": alone.status().cpu().numpy()}
        got = {k: v.numpy() for k, v in req.result.items()}
        _same(got, ref, f"request of {req.size} admitted at step {req.start_step}")
        del alone
    _log("serving 2048x1536x60: every request equals its run alone")
    del fs
    torch.cuda.empty_cache()
