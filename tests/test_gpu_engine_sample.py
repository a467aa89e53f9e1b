"""Output events of the fused form (AB3Engine.sample, opx_sample_split) against
the drop-in output kernels and compute_z at the engine's positions (FP64:
bit-exact), each release at its own time, and against the oracle."""

import numpy as np
import pytest
import torch

from oracle import driver as od
from offline_particles_amd.engine import AB3Engine
from offline_particles_amd.kernels import merge_particle_fields
from offline_particles_amd.kernels.output import (bilinear_interpolation_kernel, linear_interpolation_kernel,
                                                  trilinear_interpolation_kernel)
from offline_particles_amd.kernels.roms import compute_z_kernel
from offline_particles_amd.launcher import Launcher
from offline_particles_amd.particles import Particles
from offline_particles_amd.workloads import OFFSETS, TIME_DEPENDENT, build_fieldset, turbulent_workload

pytestmark = pytest.mark.gpu

NAMES = ("u", "v", "w", "zeta", "h", "C")


def _drop_in(fs, tidx, z, y, x, status, device):
    """The drop-in output kernels (one launch per field) at FP64 indices."""
    kernels = {"u": trilinear_interpolation_kernel("u", "_u"), "v": trilinear_interpolation_kernel("v", "_v"),
               "w": trilinear_interpolation_kernel("w", "_w"), "zeta": bilinear_interpolation_kernel("zeta", "_zeta"),
               "h": bilinear_interpolation_kernel("h", "_h"), "C": linear_interpolation_kernel("C", "_C"),
               "depth": compute_z_kernel("_depth")}
    P = Particles(len(z), device=device, **merge_particle_fields(kernels.values()))
    for k, a in (("zidx", z), ("yidx", y), ("xidx", x)):
        P[k].copy_(a)
    P["status"].copy_(status)
    launcher = Launcher(fs)
    for kern in kernels.values():
        launcher.launch_kernel(kern, P, tidx)
    return {k: P["_" + k].cpu().numpy() for k in kernels}


def _tidx(ta, t):
    idx = np.searchsorted(ta, t, side="right") - 1
    return idx + (t - ta[idx]) / (ta[idx + 1] - ta[idx])


def test_sample_equals_drop_in_outputs_per_release(device):
    wl = turbulent_workload(NX=96, NY=80, NZ=12, nparticles=12_000, steps=80, nsnap_unique=4)
    fs = build_fieldset(wl, device=device)
    rng = np.random.default_rng(4)
    sizes, starts = (7000, 5000), (0.0, 1234.5)
    pos = [tuple(rng.uniform(1, m - 2, k) for m in (wl.NZ, wl.NY, wl.NX)) for k in sizes]
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=sum(sizes), rebin_every=9)
    eng.launch_steps = 4
    for p, t0 in zip(pos, starts):
        eng.add_release(*p, start_time=t0, nsteps=60)
    eng.step(37)
    out = eng.sample(fs, NAMES, depth="depth")
    torch.cuda.synchronize()
    z, y, x = eng.positions()
    st = eng.status()
    act_all = (st < 128).cpu().numpy()
    assert act_all.sum() > 0.9 * sum(sizes)
    off = 0
    for r, k in enumerate(sizes):
        sl = slice(off, off + k)
        tidx = _tidx(wl.time_array, eng._t[r])
        print(f"release {r}: {k} particles at tidx {tidx:.4f}", flush=True)
        exp = _drop_in(fs, tidx, z[sl], y[sl], x[sl], st[sl], device)
        act = act_all[sl]
        for name in NAMES + ("depth",):
            got = out[name][sl].cpu().numpy()
            assert np.array_equal(got[act], exp[name][act]), (r, name)
            assert np.all(np.isnan(got[~act])), (r, name)
        off += k
    # the two releases are sampled at different times
    assert _tidx(wl.time_array, eng._t[0]) != _tidx(wl.time_array, eng._t[1])

    # and the oracle (the reference's full-grid blend + trilinear) at a few particles of release 0
    L = od.load()
    tidx = _tidx(wl.time_array, eng._t[0])
    zz, yy, xx = (a[:200].cpu().numpy() for a in (z, y, x))
    sel = act_all[:200]
    a = od.OracleField(wl.fields["u"], OFFSETS["u"], time_dependent="u" in TIME_DEPENDENT).field_data(tidx, L)
    o = OFFSETS["u"]
    ref = od.trilinear(a.copy(), zz[sel] + o[0], yy[sel] + o[1], xx[sel] + o[2])
    assert np.array_equal(out["u"][:200].cpu().numpy()[sel], ref)


def test_sample_argument_errors(device):
    wl = turbulent_workload(NX=32, NY=24, NZ=6, nparticles=100, steps=10, nsnap_unique=2)
    fs = build_fieldset(wl, device=device)
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=100)
    eng.add_release(*wl.positions, nsteps=5)
    with pytest.raises(ValueError):
        eng.sample(fs)
    with pytest.raises(KeyError):
        eng.sample(fs, ("no_such_field",))
    eng.step(3)
    out = eng.sample(fs, ("u",))
    assert out["u"].shape == (100,) and out["u"].dtype == torch.float64
