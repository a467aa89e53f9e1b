"""The drop-in over chunked fields (``TimeDependentField.from_dask``,
``StaticField.from_dask``; the reference's ChunkedDaskArray slices,
spatial_arrays.py:228-350, fields.py:216-231, 443-458).

Gate: each launch holds on the device only the chunks covering the particles'
bounding box plus the timestepper's index padding, and the results are the
bits of the same run over resident fields.  The subset's offsets are the
stagger offsets minus the subset's start (integers), so x + offset stays exact
in FP64 and every stencil base and weight is unchanged.  dask is not installed:
a stand-in with its interface (chunk sizes, lazy slices with ``compute()``)
holds the host arrays; a Zarr store read chunk by chunk is the second source.
"""

import numpy as np
import pytest
import torch

from offline_particles_amd.fields import StaticField, TimeDependentField
from offline_particles_amd.fieldset import Fieldset
from offline_particles_amd.snapshots import write_zarr_snapshots
from offline_particles_amd.workloads import STAGGERS, TIME_DEPENDENT, build_simulation, turbulent_workload

from ._helpers import gpu_state
from .test_chunked_host import FakeDask, _regular

pytestmark = pytest.mark.gpu


def _wl():
    wl = turbulent_workload(NX=96, NY=80, NZ=12, nparticles=4_000, steps=60, nsnap_unique=4)
    rng = np.random.default_rng(5)
    n = len(wl.positions[0])
    # particles in a sub-box, so the covering chunks are fewer than all
    z = rng.uniform(1.0, wl.NZ - 2.0, n)
    y = rng.uniform(30.0, 50.0, n)
    x = rng.uniform(40.0, 70.0, n)
    wl.positions = (z, y, x)
    return wl


def _chunked_fieldset(wl, device, cy, cx, zarr_dir=None):
    fields = {}
    for name, arr in wl.fields.items():
        zs, ys, xs = STAGGERS[name]
        dims = arr.shape[1:] if name in TIME_DEPENDENT else arr.shape
        # chunk y and x (the last two dims), keep z whole
        chunks = [(n,) for n in dims]
        if len(dims) >= 2:
            chunks[-2] = _regular(dims[-2], cy)
            chunks[-1] = _regular(dims[-1], cx)
        if name in TIME_DEPENDENT:
            if zarr_dir is not None:
                write_zarr_snapshots(zarr_dir, name, arr, chunks=(1,) + tuple(c[0] for c in chunks))
                fields[name] = TimeDependentField.from_zarr_chunked(zarr_dir, name, zs, ys, xs, cyclic=True,
                                                                    device=device)
            else:
                fields[name] = TimeDependentField.from_dask(FakeDask(arr, ((1,) * len(arr),) + tuple(chunks)),
                                                            zs, ys, xs, cyclic=True, device=device)
        elif len(dims) >= 2:
            fields[name] = StaticField.from_dask(FakeDask(arr, tuple(chunks)), zs, ys, xs, device=device)
        else:
            fields[name] = StaticField.from_numpy(arr, zs, ys, xs, device=device)
    return Fieldset(len(wl.time_array), wl.NZ, wl.NY, wl.NX, constants={"hc": wl.hc, "NZ": wl.NZ}, **fields)


@pytest.mark.parametrize("scheme", ["ab3", "rk2"])
@pytest.mark.parametrize("chunks", [(16, 16), (8, 24)])
@pytest.mark.parametrize("source", ["dask", "zarr"])
def test_chunked_fields_give_the_resident_bits(device, tmp_path, scheme, chunks, source):
    wl = _wl()
    resident = build_simulation(wl, scheme, device=device)
    fs = _chunked_fieldset(wl, device, *chunks, zarr_dir=tmp_path / "f.zarr" if source == "zarr" else None)
    chunked = build_simulation(wl, scheme, device=device, fieldset=fs)
    for _ in range(50):  # across snapshot boundaries (hourly snapshots, 120 s steps)
        resident.step()
        chunked.step()
    torch.cuda.synchronize()
    a, b = gpu_state(resident), gpu_state(chunked)
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=True), (scheme, chunks, source, k)
    # the launches held the covering chunks only
    u = fs["u"]
    held = [s for s in u._chunk_slices.values() if s._subset is not None]
    assert held and all(np.prod(s._subset.shape) < np.prod(u.spatial_shape) for s in held)
    assert fs["h"].data.loads >= 1 and np.prod(fs["h"].data._subset.shape) < np.prod(fs["h"].spatial_shape)
