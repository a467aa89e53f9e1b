"""compute_z and the output interpolation kernels, run as output events of a
Simulation, against the oracle at the same positions and time index (FP64:
bit-exact)."""

import numpy as np
import pytest

from oracle import driver as od
from offline_particles_amd import SimulationBuilder
from offline_particles_amd.kernels.roms import compute_z_kernel
from offline_particles_amd.models import roms
from offline_particles_amd.output import MemoryOutputBuilder, Output, linearly_interpolate_fields
from offline_particles_amd.workloads import OFFSETS, TIME_DEPENDENT, build_fieldset, eddy_workload

pytestmark = pytest.mark.gpu


def _oracle_field(wl, name):
    return od.OracleField(wl.fields[name], OFFSETS[name], time_dependent=name in TIME_DEPENDENT)


def _expected(wl, name, tidx, z, y, x):
    L = od.load()
    a = _oracle_field(wl, name).field_data(tidx, L).copy()
    o = OFFSETS[name]
    if a.ndim == 3:
        return od.trilinear(a, z + o[0], y + o[1], x + o[2])
    if a.ndim == 2:
        return od.bilinear(a, y + o[0], x + o[1])
    return od.linear(a, z + o[0])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_output_events_vs_oracle(device, dtype):
    wl = eddy_workload(steps=30, nparticles=1200, dtype=dtype)
    fs = build_fieldset(wl, device=device)
    ts = roms.rk2_w_advection_timestepper(wl.time_array, np.float64(wl.dt))
    b = SimulationBuilder(ts, fs)
    mem = MemoryOutputBuilder("mem")
    outs = linearly_interpolate_fields(fs, "u", "w", "zeta", "h", "C")
    mem.add_output(*outs)
    mem.add_output(Output("depth", compute_z_kernel("_depth"), particle_field="_depth"))
    b.add_output_writer(mem, n=5)
    # outputs of one field dtype share one particle field ("_output_<dtype>"),
    # filled just before each is written
    assert {o.name: o.particle_field for o in outs} == {
        n: f"_output_{wl.fields[n].dtype}" for n in ("u", "w", "zeta", "h", "C")}
    sim = b.build_simulation(len(wl.positions[0]), device=device)
    # the first write round ran at construction, with every index 0
    snaps = [(float(sim.tidx), np.zeros(len(wl.positions[0])), np.zeros(len(wl.positions[0])),
              np.zeros(len(wl.positions[0])), np.zeros(len(wl.positions[0]), np.uint8))]
    sim.set_indices(*wl.positions)
    for _ in range(20):
        sim.step()
        if sim.iteration % 5 == 0:
            snaps.append((float(sim.tidx),) + tuple(sim.particles.to_numpy(k) for k in ("zidx", "yidx", "xidx",
                                                                                     "status")))
    arrs = sim.output_writers["mem"].arrays()
    assert arrs["time"].shape == (len(snaps),)
    for r, (tidx, z, y, x, st) in enumerate(snaps):
        act = st < 128
        for name in ("u", "w", "zeta", "h", "C"):
            exp = _expected(wl, name, tidx, z[act], y[act], x[act])
            got = arrs[name][r][act]
            assert np.array_equal(got, exp), (name, r, float(np.max(np.abs(got - exp))))
        h = _expected(wl, "h", tidx, z[act], y[act], x[act])
        zeta = _expected(wl, "zeta", tidx, z[act], y[act], x[act])
        C = _expected(wl, "C", tidx, z[act], y[act], x[act])
        exp = od.compute_z(z[act], wl.NZ, wl.hc, h, C, zeta)
        assert np.array_equal(arrs["depth"][r][act], exp), ("depth", r)


def test_zarr_writer_matches_memory_writer(device, tmp_path):
    """The Zarr v3 trajectory store written during a run holds the rows the memory
    writer holds (same rounds, same values), synchronous and asynchronous."""
    from offline_particles_amd.output import ZarrOutputBuilder, open_array

    wl = eddy_workload(steps=30, nparticles=2500)
    fs = build_fieldset(wl, device=device)
    ts = roms.rk2_w_advection_timestepper(wl.time_array, np.float64(wl.dt))
    b = SimulationBuilder(ts, fs)
    mem = MemoryOutputBuilder("mem")
    outs = linearly_interpolate_fields(fs, "u", "zeta")
    depth = Output("depth", compute_z_kernel("_depth"), particle_field="_depth")
    mem.add_output(*outs, depth)
    b.add_output_writer(mem, n=4)
    roots = {}
    for mode in (False, True):
        roots[mode] = str(tmp_path / f"traj_{int(mode)}.zarr")
        z = ZarrOutputBuilder(f"zarr{int(mode)}", roots[mode], chunksize=1000, async_writes=mode)
        z.add_output(*outs, depth)
        b.add_output_writer(z, n=4)
    n = len(wl.positions[0])
    sim = b.build_simulation(n, device=device)
    # the oracle's own run of the same particles: its positions and time index
    # at each write round give the rows the store must hold
    from ._helpers import run_oracle

    sim_o = run_oracle(wl, "rk2", 0)
    rounds = [(float(sim_o.tidx), np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n, np.uint8))]
    sim.set_indices(*wl.positions)
    for _ in range(20):
        sim.step()
        sim_o.step()
        if sim.iteration % 4 == 0:
            rounds.append((float(sim_o.tidx),) + tuple(sim_o.p[k].copy() for k in ("zidx", "yidx", "xidx",
                                                                                     "status")))
    sim.output_writers["zarr1"].flush()
    exp = sim.output_writers["mem"].arrays()
    for mode, root in roots.items():
        assert np.array_equal(open_array(root, "time")[:], exp["time"]), mode
        for name in ("u", "zeta", "depth"):
            got = open_array(root, name)
            assert got.shape == exp[name].shape, (mode, name)
            assert np.array_equal(got[:, :], exp[name]), (mode, name)
    # the Zarr rows against the oracle's outputs at the oracle's positions
    got = {name: open_array(roots[False], name)[:, :] for name in ("u", "zeta", "depth")}
    assert got["u"].shape[0] == len(rounds)
    for r, (tidx, z, y, x, st) in enumerate(rounds):
        act = st < 128
        for name in ("u", "zeta"):
            assert np.array_equal(got[name][r][act], _expected(wl, name, tidx, z[act], y[act], x[act])), (name, r)
        h, zeta, C = (_expected(wl, k, tidx, z[act], y[act], x[act]) for k in ("h", "zeta", "C"))
        assert np.array_equal(got["depth"][r][act], od.compute_z(z[act], wl.NZ, wl.hc, h, C, zeta)), ("depth", r)
