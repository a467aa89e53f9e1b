"""The device release clocks (opx_clock_advance_steps) against the
reference's time index, Timestepper.get_time_index (timesteppers.py:64-85):
idx = searchsorted(time_array, t, side="right") - 1, tidx = idx + (t - t0) /
(t1 - t0), then divmod(tidx, 1) with the FP32 weight (fields.py:409-424),
and the RK2 intermediate time t + alpha dt rounded twice
(timesteppers.py:250-251).  numpy's searchsorted is the checker of the
device's search, which starts from where a uniformly spaced array would hold
t: uniform arrays, arrays of uneven spacing (the search's fallback), repeated
entries, times on array entries and times leaving the array (the release
stops and the error flag is set, where the reference raises)."""

import ctypes

import numpy as np
import pytest
import torch

from offline_particles_amd import _lib

pytestmark = pytest.mark.gpu


def _time_array(kind, rng, n=400):
    if kind == "uniform":
        return np.arange(n, dtype=np.float64) * 3600.0
    gaps = rng.uniform(0.3, 3.0, n - 1) * 3600.0
    if kind == "repeats":
        gaps[rng.choice(n - 1, 12, replace=False)] = 0.0
    return np.concatenate([[0.0], np.cumsum(gaps)])


def _host_entry(ta, t, nslots, alpha, dt):
    """(ok, ft, ft1, ftm, s0, s0p, s0m) of one step at time t, as the reference."""
    def tindex(x):
        if not (ta[0] <= x <= ta[-1]):
            return None
        idx = int(np.searchsorted(ta, x, side="right")) - 1
        if idx + 1 >= len(ta):
            return None
        return idx + (x - ta[idx]) / (ta[idx + 1] - ta[idx])

    def split(tidx):
        It, f = divmod(tidx, 1.0)
        return np.float32(f), int(It) % nslots

    a, b = tindex(t), tindex(t + dt)
    m = tindex(t + np.float64(alpha * dt)) if alpha > 0 else 0.0
    if a is None or b is None or m is None:
        return None
    (ft, s0), (ft1, s0p) = split(a), split(b)
    ftm, s0m = split(m) if alpha > 0 else (np.float32(0.0), 0)
    return ft, ft1, ftm, s0, s0p, s0m


@pytest.mark.parametrize("kind", ["uniform", "uneven", "repeats"])
@pytest.mark.parametrize("alpha", [0.0, 2.0 / 3.0])
def test_clock_matches_the_reference_time_index(device, kind, alpha):
    rng = np.random.default_rng({"uniform": 1, "uneven": 2, "repeats": 3}[kind])
    ta = _time_array(kind, rng)
    R, K, nslots, dt = 257, 40, 3, 120.0
    t0 = rng.uniform(ta[0], ta[-1], R)
    t0[:20] = ta[rng.choice(len(ta), 20)]       # on array entries (and the last one)
    t0[20:30] = ta[-1] - rng.uniform(0, 30 * dt, 10)  # leave the array within the steps
    clocks = (_lib.opx_release_clock * R)()
    for r in range(R):
        clocks[r].t, clocks[r].start_step, clocks[r].steps_left = float(t0[r]), 0, K
    cbytes = torch.frombuffer(bytearray(clocks), dtype=torch.uint8).to(device)
    table = torch.zeros(K * R * 16, dtype=torch.uint8, device=device)
    ta_d = torch.from_numpy(ta).to(device)
    err = torch.zeros(1, dtype=torch.int32, device=device)
    _lib.call("opx_clock_advance_steps", cbytes.data_ptr(), table.data_ptr(), R, K, R, None, ta_d.data_ptr(),
              len(ta), nslots, 0, None, dt, alpha, err.data_ptr(), torch.cuda.current_stream().cuda_stream)
    got = (_lib.opx_release_step * (K * R)).from_buffer_copy(table.cpu().numpy().tobytes())
    stopped_any = False
    for r in range(R):
        t, left = np.float64(t0[r]), K
        for k in range(K):
            e = got[k * R + r]
            active = (e.slots >> 30) & 1
            want = _host_entry(ta, t, nslots, alpha, dt) if left > 0 else None
            if want is None:
                assert active == 0, (kind, r, k)
                left = 0
                stopped_any |= k < K
                continue
            ft, ft1, ftm, s0, s0p, s0m = want
            assert active == 1, (kind, r, k, t)
            assert (np.float32(e.ft), np.float32(e.ft1)) == (ft, ft1), (kind, r, k, t)
            assert (e.slots & 1023, (e.slots >> 10) & 1023) == (s0, s0p), (kind, r, k, t)
            if alpha > 0:
                assert np.float32(e.ftm) == ftm and (e.slots >> 20) & 1023 == s0m, (kind, r, k, t)
            t = t + np.float64(dt)
            left -= 1
    assert stopped_any and int(err.item()) == 1
