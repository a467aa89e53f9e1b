"""Parity of the HIP drop-in kernels with the CPU oracle (same seeded inputs).

Gate: FP64 fields within 1e-9 grid-index units of max position deviation after
100 steps, FP32 fields within 1e-5, status codes bit-exact.  Schemes: RK2, AB3
(MULTISTEP start-up) and AB3 whose first two steps are RK2 steps ("ab3_rk2",
against the oracle's composition of the same kernels).  The drop-in keeps
the reference's FP64 operation order and is built without FMA contraction, so
the observed deviation is expected to be exactly zero; the test reports it.
"""

import numpy as np
import pytest

from offline_particles_amd.workloads import eddy_workload, turbulent_workload

from ._helpers import POS, build_gpu_sim, gpu_state, max_dev, run_oracle

pytestmark = pytest.mark.gpu


def _compare(sim_o, sim_g, tol):
    g = gpu_state(sim_g)
    assert np.array_equal(g["status"], sim_o.p["status"]), "status differs"
    devs = {k: max_dev(g[k], sim_o.p[k]) for k in POS + ("z",)}
    print("max deviation vs oracle:", devs, flush=True)
    for k in POS:
        assert devs[k] <= tol, (k, devs[k])
    return devs


@pytest.mark.parametrize("scheme", ["rk2", "ab3", "ab3_rk2"])
def test_eddy_fp64_100_steps(device, scheme):
    wl = eddy_workload(steps=100)
    sim_o = run_oracle(wl, scheme, 100)
    sim_g = build_gpu_sim(wl, scheme, device)
    for _ in range(100):
        sim_g.step()
    devs = _compare(sim_o, sim_g, 1e-9)
    assert (sim_o.p["status"] == 130).any()  # some particles leave through the surface/bed


@pytest.mark.parametrize("scheme", ["rk2", "ab3", "ab3_rk2"])
def test_turbulent_fp32_100_steps(device, scheme):
    wl = turbulent_workload(NX=96, NY=80, NZ=12, nparticles=20_000, steps=100, nsnap_unique=4)
    sim_o = run_oracle(wl, scheme, 100)
    sim_g = build_gpu_sim(wl, scheme, device)
    for _ in range(100):
        sim_g.step()
    _compare(sim_o, sim_g, 1e-5)


def test_edge_cases_status(device):
    """NaN/inf indices, inclusive bounds, inactive particles, start-up codes."""
    wl = eddy_workload(steps=10, nparticles=16)
    NZ, NY, NX = wl.NZ, wl.NY, wl.NX
    z = np.array([0.0, NZ - 1, -1e-12, NZ - 1 + 1e-9, np.nan, 3.0, 3.0, 3.0, 5.0, 5.0, 5.0, 5.0, 1.0, 2.0, 0.5, 7.5])
    y = np.array([0.0, NY - 1, 10.0, 10.0, 10.0, np.inf, 10.0, 10.0, 40.0, 40.0, 40.0, -0.0, NY - 1, 30.0, 47.5, 47.5])
    x = np.array([0.0, NX - 1, 10.0, 10.0, 10.0, 10.0, -np.inf, NX - 1 + 1e-6, 60.0, 60.0, 60.0, 60.0, 64.0, 64.0, 63.5, 63.5])
    status = np.array([10, 10, 10, 10, 10, 10, 10, 10, 128, 255, 11, 0, 10, 10, 10, 10], np.uint8)
    for scheme in ("rk2", "ab3"):
        st = status if scheme == "ab3" else np.where(status >= 128, status, 0).astype(np.uint8)
        sim_o = run_oracle(wl, scheme, 5, positions=(z, y, x), status=st)
        sim_g = build_gpu_sim(wl, scheme, device, positions=(z, y, x), status=st)
        for _ in range(5):
            sim_g.step()
        _compare(sim_o, sim_g, 1e-12)


def test_zero_particles(device):
    wl = eddy_workload(steps=4, nparticles=0)
    sim_g = build_gpu_sim(wl, "ab3", device)
    for _ in range(3):
        sim_g.step()
    assert len(sim_g.particles) == 0


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-9), (np.float32, 1e-5)])
def test_isopycnal_following_100_steps(device, dtype, tol):
    """AB3 isopycnal following (models/roms/__init__.py:109-140) against the oracle."""
    from offline_particles_amd import SimulationBuilder
    from offline_particles_amd.models import roms
    from offline_particles_amd.workloads import add_density, build_fieldset

    wl = add_density(eddy_workload(steps=100, nparticles=5000, dtype=dtype))
    n = wl.nparticles
    rho_t = np.random.default_rng(2).uniform(1026.2, 1027.8, n)
    consts = {"w_max": 5e-4, "rho_scale": 0.2}
    sim_o = run_oracle(wl, "ab3_iso", 100, rho_t=rho_t, **consts)
    fs = build_fieldset(wl, device=device, constants=consts)
    ts = roms.ab3_isopycnal_following_timestepper(wl.time_array, np.float64(wl.dt))
    sim_g = SimulationBuilder(ts, fs).build_simulation(n, device=device)
    sim_g.set_indices(*wl.positions)
    sim_g.set_particle_field("status", 10)
    sim_g.set_particle_field("rho_t", rho_t)
    for _ in range(100):
        sim_g.step()
    _compare(sim_o, sim_g, tol)


@pytest.mark.parametrize("scheme", ["ab3", "rk2"])
def test_spatial_order_changes_nothing_the_user_sees(device, scheme):
    """Simulation(spatial_order_every=N) permutes every particle array into the
    Morton order of the cells every N steps: positions, status, every particle
    field and the output writer's rows, read in the particles' own order, are
    the bits of the run in the caller's order; writes after a reorder land on
    the right particles."""
    from offline_particles_amd import SimulationBuilder
    from offline_particles_amd.kernels.roms import compute_z_kernel
    from offline_particles_amd.models import roms
    from offline_particles_amd.output import MemoryOutputBuilder, Output, linearly_interpolate_fields
    from offline_particles_amd.workloads import build_fieldset

    wl = turbulent_workload(NX=96, NY=80, NZ=12, nparticles=20_000, steps=100, nsnap_unique=4)
    fs = build_fieldset(wl, device=device)
    sims = []
    for every in (0, 7):
        ts = (roms.ab3_w_advection_timestepper if scheme == "ab3" else roms.rk2_w_advection_timestepper)(
            wl.time_array, np.float64(wl.dt))
        b = SimulationBuilder(ts, fs)
        mem = MemoryOutputBuilder("mem")
        mem.add_output(*linearly_interpolate_fields(fs, "u", "zeta"),
                       Output("depth", compute_z_kernel("_depth"), particle_field="_depth"))
        b.add_output_writer(mem, n=5)
        sim = b.build_simulation(wl.nparticles, device=device, spatial_order_every=every)
        sim.set_indices(*wl.positions)
        if scheme == "ab3":
            sim.set_particle_field("status", 10)
        sims.append(sim)
    for k in range(60):
        for sim in sims:
            sim.step()
        if k == 40:
            # a write after several reorders: particle 123 made inactive in both runs
            for sim in sims:
                st = sim.particles.to_numpy("status")
                st[123] = 128
                sim.set_particle_field("status", st)
    plain, ordered = sims
    order = ordered.spatial_order.cpu().numpy()
    assert np.array_equal(np.sort(order), np.arange(wl.nparticles)) and not np.array_equal(order,
                                                                                        np.arange(wl.nparticles))
    for name in plain.particles.arrays:
        a, b = plain.particles.to_numpy(name), ordered.particles.to_numpy(name)
        assert np.array_equal(a, b, equal_nan=True), name
    assert plain.particles.to_numpy("status")[123] == 128
    ra, rb = plain.output_writers["mem"].arrays(), ordered.output_writers["mem"].arrays()
    for name in ra:
        assert np.array_equal(ra[name], rb[name], equal_nan=True), name
