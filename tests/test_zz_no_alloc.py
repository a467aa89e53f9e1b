"""opx_rebin_noalloc (the re-bin with the library's static scratch buffer)
against opx_rebin (the caller's workspace), through the C-ABI: every array of
the permuted state and the active / live counts bit for bit, on two calls in a
row (the second over fewer particles, so stale scratch contents would show),
and the size limit refused before any launch."""

import ctypes

import numpy as np
import pytest
import torch

from offline_particles_amd import _lib
from offline_particles_amd.engine import _STATE_F32, AB3Engine
from offline_particles_amd.workloads import build_fieldset, turbulent_workload

pytestmark = pytest.mark.gpu

_ARRAYS = ("status", "cell", "pid", "rel", "z") + _STATE_F32


def _state_arrays(st, n):
    return {k: getattr(st, k)[:n].cpu().numpy() for k in _ARRAYS}


def test_rebin_noalloc_equals_rebin(device):
    wl = turbulent_workload(NX=96, NY=80, NZ=12, nparticles=1, steps=60, nsnap_unique=4)
    fs = build_fieldset(wl, device=device)
    rng = np.random.default_rng(31)
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=40_000, rebin_every=0)
    for m, t0, ns in ((25_000, 0.0, 40), (15_000, 700.0, 10)):
        eng.add_release(*(rng.uniform(-0.5, n - 0.5, m) for n in (wl.NZ, wl.NY, wl.NX)), start_time=t0, nsteps=ns)
    eng.step(20)  # histories filled, some particles out of the domain, the second release finished
    torch.cuda.synchronize()
    src = eng.state
    stream = torch.cuda.current_stream(device).cuda_stream
    for n in (40_000, 9_000):  # the second call is smaller
        print(f"rebin_noalloc vs rebin: {n} particles", flush=True)
        outs = []
        for noalloc in (False, True):
            dst = eng._new_state(src.capacity)
            counts = torch.full((2,), -7, dtype=torch.int32, device=device)
            s_src, s_dst = src.struct(n), dst.struct(n)
            args = (ctypes.byref(s_src), ctypes.byref(s_dst), ctypes.byref(eng.grid), 3, eng._clocks.data_ptr(),
                    len(eng._clocks_host))
            if noalloc:
                _lib.call("opx_rebin_noalloc", *args, counts.data_ptr(), counts.data_ptr() + 4, stream)
            else:
                need = ctypes.c_uint64()
                _lib.call("opx_rebin_workspace_bytes", n, ctypes.byref(need))
                work = torch.empty(int(need.value), dtype=torch.uint8, device=device)
                _lib.call("opx_rebin", *args, work.data_ptr(), work.numel(), counts.data_ptr(),
                          counts.data_ptr() + 4, stream)
            torch.cuda.synchronize()
            outs.append((_state_arrays(dst, n), counts.cpu().numpy()))
        (a, ca), (b, cb) = outs
        assert np.array_equal(ca, cb) and 0 < ca[0] <= ca[1] <= n, (ca, cb)
        for k in _ARRAYS:
            assert np.array_equal(a[k], b[k], equal_nan=True), (n, k)
        # a permutation of the input: every particle once
        assert np.array_equal(np.sort(a["pid"]), np.sort(src.pid[:n].cpu().numpy()))

    # beyond the static buffer: refused before any launch
    big = src.struct(8_000_000)
    with pytest.raises(_lib.OpxInvalidArgument, match="static"):
        _lib.call("opx_rebin_noalloc", ctypes.byref(big), ctypes.byref(big), ctypes.byref(eng.grid), 3, None, 0,
                  eng._n_active_dev.data_ptr(), None, stream)
