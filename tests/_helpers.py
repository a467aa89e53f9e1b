"""Shared helpers: run a workload on the oracle and on the HIP drop-in."""

from __future__ import annotations

import numpy as np

from offline_particles_amd.workloads import OFFSETS, TIME_DEPENDENT

POS = ("zidx", "yidx", "xidx")


def oracle_fields(wl):
    from oracle import OracleField

    return {k: OracleField(v, OFFSETS[k], k in TIME_DEPENDENT) for k, v in wl.fields.items()}


def run_oracle(wl, scheme, steps, *, positions=None, status=None, threads=None, **kw):
    from oracle import OracleSimulation

    pos = wl.positions if positions is None else positions
    sim = OracleSimulation(oracle_fields(wl), wl.time_array, wl.dt, scheme=scheme, nparticles=len(pos[0]),
                           hc=wl.hc, NZ=wl.NZ, bounds=wl.bounds, threads=threads, **kw)
    for k, a in zip(POS, pos):
        sim.p[k][:] = a
    if status is not None:
        sim.p["status"][:] = status
    elif scheme in ("ab3", "ab3_iso"):
        sim.p["status"][:] = 10
    sim.run(steps)
    return sim


def build_gpu_sim(wl, scheme, device, *, positions=None, status=None, storage_dtype=None):
    from offline_particles_amd.workloads import build_simulation

    return build_simulation(wl, scheme, device=device, storage_dtype=storage_dtype, positions=positions,
                            status=status)


def gpu_state(sim):
    return {k: sim.particles.to_numpy(k) for k in ("status",) + POS + ("z",)}


def max_dev(a, b):
    """Max |a - b| over entries finite in both; NaN placement must agree."""
    fa, fb = np.isfinite(a), np.isfinite(b)
    assert np.array_equal(fa, fb), "non-finite entries differ"
    if not fa.any():
        return 0.0
    return float(np.max(np.abs(a[fa] - b[fa])))
