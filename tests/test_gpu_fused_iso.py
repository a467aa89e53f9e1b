"""Fused isopycnal-following AB3 (AB3Engine(scheme="ab3_iso"),
opx_ab3_iso_fused_steps) against the oracle's ab3_isopycnal_following.

Gate (the fused FP32 gate): max position deviation <= 1e-5 grid-index units
after 100 steps with status codes bit-exact.  Layouts, several steps per
launch, re-binning and graph replays must give identical bits.
"""

import numpy as np
import pytest
import torch

from offline_particles_amd import _lib
from offline_particles_amd.engine import AB3Engine
from offline_particles_amd.workloads import add_density, build_fieldset, eddy_workload, turbulent_workload

from ._helpers import POS, max_dev, run_oracle

pytestmark = pytest.mark.gpu

CONSTS = {"w_max": 5e-4, "rho_scale": 0.2}


def _wl(kind):
    if kind == "eddy":
        wl = eddy_workload(steps=100, nparticles=5000, dtype=np.float32)
    else:
        wl = turbulent_workload(NX=96, NY=80, NZ=12, nparticles=20_000, steps=100, nsnap_unique=4)
    wl = add_density(wl)
    rho_t = np.random.default_rng(2).uniform(1026.2, 1027.8, wl.nparticles)
    return wl, rho_t


def _engine(wl, rho_t, device, *, layout=None, launch_steps=1, rebin_every=25, xlow=2, nsteps=100, graph_steps=0):
    fs = build_fieldset(wl, device=device, constants=CONSTS)
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=wl.nparticles,
                    rebin_every=rebin_every, xlow=xlow, scheme="ab3_iso")
    if layout is not None:
        eng.set_layout(layout)
    eng.launch_steps = launch_steps
    eng.graph_steps = graph_steps
    eng.add_release(*wl.positions, nsteps=nsteps, rho_t=rho_t)
    return eng


def _state(eng):
    z, y, x = (t.cpu().numpy() for t in eng.positions())
    return {"zidx": z, "yidx": y, "xidx": x, "status": eng.status().cpu().numpy()}


@pytest.mark.parametrize("kind", ["eddy", "turbulent"])
def test_fused_iso_vs_oracle_100_steps(device, kind):
    wl, rho_t = _wl(kind)
    print(f"fused iso vs oracle: {kind}, {wl.nparticles} particles, 100 steps", flush=True)
    sim_o = run_oracle(wl, "ab3_iso", 100, rho_t=rho_t, **CONSTS)
    eng = _engine(wl, rho_t, device)
    eng.step(100)
    eng.check_clock()
    g = _state(eng)
    assert np.array_equal(g["status"], sim_o.p["status"]), "status differs"
    devs = {k: max_dev(g[k], sim_o.p[k]) for k in POS}
    print("fused iso max deviation vs oracle after 100 steps:", devs, flush=True)
    assert max(devs.values()) <= 1e-5
    # the density drive moved the particles through the column
    act = g["status"] < 128
    assert np.max(np.abs(g["zidx"][act] - wl.positions[0][act])) > 0.1


def test_fused_iso_layouts_launch_steps_graphs_identical(device):
    wl, rho_t = _wl("turbulent")
    cases = [(_lib.OPX_LAYOUT_RING, 1, 0), (_lib.OPX_LAYOUT_PAIR, 1, 0), (_lib.OPX_LAYOUT_QUAD, 8, 0),
             (_lib.OPX_LAYOUT_OCT, 16, 8), (_lib.OPX_LAYOUT_YPAIR, 16, 0), (_lib.OPX_LAYOUT_PAIR, 64, 0),
             (_lib.OPX_LAYOUT_PAIR_LDS, 1, 0)]
    ref = None
    for layout, m, graph in cases:
        print(f"fused iso: layout {layout}, {m} steps per launch, graph steps {graph}", flush=True)
        eng = _engine(wl, rho_t, device, layout=layout, launch_steps=m, rebin_every=11, graph_steps=graph)
        eng.step(100)
        torch.cuda.synchronize()
        st = _state(eng)
        if ref is None:
            ref = st
            continue
        for k in ("status",) + POS:
            assert np.array_equal(ref[k], st[k], equal_nan=True), (layout, m, graph, k)


def test_fused_iso_argument_errors(device):
    wl, rho_t = _wl("eddy")
    fs = build_fieldset(wl, device=device, constants=CONSTS)
    with pytest.raises(TypeError):
        AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, storage_dtype=torch.bfloat16,
                  scheme="ab3_iso")
    fs_no_consts = build_fieldset(wl, device=device)
    with pytest.raises(KeyError):
        AB3Engine(fs_no_consts, wl.time_array, np.float64(wl.dt), device=device, scheme="ab3_iso")
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=10, scheme="ab3_iso")
    pos = tuple(a[:10] for a in wl.positions)
    with pytest.raises(ValueError):
        eng.add_release(*pos, nsteps=5)  # rho_t missing
    with pytest.raises(ValueError):
        eng.add_release(*pos, nsteps=5, rho_t=rho_t[:3])
    w = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=10)
    with pytest.raises(ValueError):
        w.add_release(*pos, nsteps=5, rho_t=rho_t[:10])  # w advection has no target density
    # one scalar target for a whole release
    eng.add_release(*pos, nsteps=5, rho_t=1027.0)
    eng.step(5)
    assert eng.active_count() > 0
