"""Host logic of the output layer (no GPU): Output declarations, kernel choice
by field dims, writer events and round checks (output/_output.py,
output/_linear_interpolation.py, output/_zarr_writer.py:61-98)."""

import types

import numpy as np
import pytest
import torch

from offline_particles_amd.kernels.output import bilinear_interpolation_kernel, linear_interpolation_kernel
from offline_particles_amd.kernels.roms import compute_z_kernel
from offline_particles_amd.output import MemoryOutputBuilder, Output, linearly_interpolate_fields
from offline_particles_amd.workloads import build_fieldset, eddy_workload


def test_output_needs_declared_particle_field():
    k = compute_z_kernel("_depth")
    assert Output("depth", k, particle_field="_depth").dtype == np.float64
    with pytest.raises(ValueError, match="not found"):
        Output("depth", k)


def test_interpolation_kernel_choice_and_declarations():
    wl = eddy_workload(NX=16, NY=12, NZ=5, nparticles=10, steps=2)
    fs = build_fieldset(wl, device="cpu")
    outs = {o.name: o for o in linearly_interpolate_fields(fs, "u", "zeta", "C")}
    names = {n: o.kernels[0].functions[0].func.__name__ for n, o in outs.items()}
    assert names == {"u": "trilinear_interpolation_kernel_function",
                     "zeta": "bilinear_interpolation_kernel_function",
                     "C": "linear_interpolation_kernel_function"}
    assert outs["C"].kernels[0].functions[0].keywords["dimension_idx"] == "zidx"
    assert outs["zeta"].kernels[0].functions[0].keywords["dimension_idx0"] == "yidx"
    assert outs["u"].particle_field == "_output_float64"
    assert set(outs["u"].kernels[0].simulation_fields) == {"u"}
    with pytest.raises(KeyError):
        linearly_interpolate_fields(fs, "salt")
    with pytest.raises(ValueError):
        linear_interpolation_kernel("C", dimension="t")
    with pytest.raises(ValueError):
        bilinear_interpolation_kernel("zeta", dimensions=("x", "y"))


def test_memory_writer_rounds():
    k = compute_z_kernel("_depth")
    b = MemoryOutputBuilder("mem")
    b.add_output(Output("depth", k, particle_field="_depth"))
    with pytest.raises(KeyError):
        b.add_output(Output("depth", k, particle_field="_depth"))
    w = b.build(4, np.dtype(np.float64))
    ev = w.create_events()
    assert [e.name for e in ev] == ["mem:time", "mem:depth", "mem:finalise"]
    assert ev[1].kernels == (k,)
    for t in range(3):
        st = types.SimpleNamespace(time=np.float64(10.0 * t),
                                   particles={"_depth": torch.arange(4, dtype=torch.float64) + t})
        for e in ev:
            e(st)
    a = w.arrays()
    assert np.array_equal(a["time"], [0.0, 10.0, 20.0])
    assert np.array_equal(a["depth"], np.arange(4)[None, :] + np.arange(3)[:, None])
    # a round with an output missing fails the finalise check
    ev[0](st)
    with pytest.raises(RuntimeError, match="expected 4"):
        ev[2](st)
