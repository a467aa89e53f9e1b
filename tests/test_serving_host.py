"""Host logic of the serving form (no GPU): admission queue, request generator, slot size."""

import numpy as np
import pytest
import torch

from offline_particles_amd.engine import SplitState
from offline_particles_amd.serving import SLOT_BYTES, AdmissionQueue, Request, poisson_requests
from offline_particles_amd.workloads import turbulent_workload


def _req(m, arrival=0.0):
    z = np.zeros(m)
    return Request(z, z, z, nsteps=10, arrival=arrival)


def test_admission_fifo_by_slots_and_ids():
    q = AdmissionQueue(capacity=1000, max_releases=2)
    for m in (600, 300, 50):
        q.submit(_req(m))
    assert q.head_fits(used_slots=0, used_ids=0)
    assert q.pop().size == 600
    # 300 fits beside 600; ids: 1 of 2 in use
    assert q.head_fits(600, 1)
    assert q.pop().size == 300
    # the 50-particle request fits the slots but no release id is free
    assert not q.head_fits(900, 2)
    assert q.head_fits(900, 1)
    # strict FIFO: a blocked head blocks the smaller requests behind it
    q2 = AdmissionQueue(capacity=1000, max_releases=8)
    q2.submit(_req(900))
    q2.submit(_req(10))
    assert not q2.head_fits(200, 1)


def test_admission_rejects_oversized():
    q = AdmissionQueue(capacity=100, max_releases=4)
    with pytest.raises(ValueError, match="exceeds the capacity"):
        q.submit(_req(101))


def test_slot_bytes_match_state_arrays():
    s = SplitState(10, "cpu")
    per = sum(t.element_size() for t in (s.status, s.cell, s.fz, s.fy, s.fx, s.hz1, s.hy1, s.hx1, s.hz2, s.hy2,
                                         s.hx2, s.pid, s.rel, s.z))
    assert per == SplitState.BYTES_PER_PARTICLE == 55
    assert SLOT_BYTES == 2 * per + 16 + 8
    assert s.status.dtype == torch.uint8 and bool((s.status == 128).all())  # empty slots are INACTIVE


def test_poisson_requests_statistics():
    wl = turbulent_workload(NX=64, NY=48, NZ=8, nparticles=1, steps=10, nsnap_unique=2)
    reqs = poisson_requests(wl, rate=20.0, duration=50.0, size_range=(1_000, 20_000), steps_range=(720, 7200),
                            seed=1)
    n = len(reqs)
    assert abs(n - 1000) < 4 * np.sqrt(1000)  # Poisson count, mean 1000
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) >= 0) and arr[-1] < 50.0
    sizes = np.array([r.size for r in reqs])
    assert sizes.min() >= 1_000 and sizes.max() <= 20_000
    # log-uniform: the median is near the geometric mean of the range
    assert abs(np.log(np.median(sizes)) - 0.5 * (np.log(1e3) + np.log(2e4))) < 0.25
    steps = np.array([r.nsteps for r in reqs])
    assert steps.min() >= 720 and steps.max() <= 7200
    for r in reqs[:20]:
        assert r.yidx.min() >= 1 and r.yidx.max() <= wl.NY - 2
        assert r.xidx.min() >= 1 and r.xidx.max() <= wl.NX - 2
        assert np.ptp(r.xidx) <= 10 and np.ptp(r.yidx) <= 10
