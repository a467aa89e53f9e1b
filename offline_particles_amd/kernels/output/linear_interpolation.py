"""Output by linear interpolation of field data (mirrors kernels/output/linear_interpolation.pyx).

Each kernel function writes, for every active particle, the (blended) field
value at the particle's indices into a particle field: linear over one index,
bilinear over two, trilinear over (zidx, yidx, xidx) -- one HIP kernel each
(``opx_interp_linear`` / ``_bilinear`` / ``_trilinear``), FP64 with the
reference's operation order.
"""

from __future__ import annotations

import ctypes
from functools import partial
from typing import Literal

import numpy as np
import torch

from ... import _lib
from .._kernels import ParticleKernel
from .._marshal import _check_tensor, field_struct, stream_handle

__all__ = [
    "linear_interpolation_kernel_function",
    "bilinear_interpolation_kernel_function",
    "trilinear_interpolation_kernel_function",
    "linear_interpolation_kernel",
    "bilinear_interpolation_kernel",
    "trilinear_interpolation_kernel",
    "vertical_interpolation_kernel",
    "horizontal_interpolation_kernel",
]


def _f64(particles, name):
    return _check_tensor(particles[name], name, torch.float64)


def _status(particles):
    return _check_tensor(particles["status"], "status", torch.uint8)


def linear_interpolation_kernel_function(particles, scalars, fielddata, dimension_idx, field_name,
                                         particle_name) -> None:
    """Perform linear interpolation on a 1D array (linear_interpolation.pyx:34-71)."""
    f = field_struct(fielddata[field_name], field_name)
    _lib.call("opx_interp_linear", len(particles), _status(particles), _f64(particles, dimension_idx),
              ctypes.byref(f), _f64(particles, particle_name), stream_handle())


def bilinear_interpolation_kernel_function(particles, scalars, fielddata, dimension_idx0, dimension_idx1,
                                           field_name, particle_name) -> None:
    """Perform bilinear interpolation on a 2D array (linear_interpolation.pyx:73-110)."""
    f = field_struct(fielddata[field_name], field_name)
    _lib.call("opx_interp_bilinear", len(particles), _status(particles), _f64(particles, dimension_idx0),
              _f64(particles, dimension_idx1), ctypes.byref(f), _f64(particles, particle_name), stream_handle())


def trilinear_interpolation_kernel_function(particles, scalars, fielddata, field_name, particle_name) -> None:
    """Perform trilinear interpolation on a 3D array (linear_interpolation.pyx:112-145)."""
    f = field_struct(fielddata[field_name], field_name)
    _lib.call("opx_interp_trilinear", len(particles), _status(particles), _f64(particles, "zidx"),
              _f64(particles, "yidx"), _f64(particles, "xidx"), ctypes.byref(f), _f64(particles, particle_name),
              stream_handle())


def linear_interpolation_kernel(field: str, output_name: str | None = None,
                                dimension: Literal["z", "y", "x"] = "z") -> ParticleKernel:
    """Return a ParticleKernel that performs linear interpolation of field data."""
    if dimension not in {"z", "y", "x"}:
        raise ValueError(f"Invalid dimension: {dimension}. Valid options are 'z', 'y', or 'x'.")
    dimension_idx = dimension + "idx"
    if output_name is None:
        output_name = field
    fn = partial(linear_interpolation_kernel_function, dimension_idx=dimension_idx, field_name=field,
                 particle_name=output_name)
    return ParticleKernel(fn, particle_fields={"status": np.uint8, dimension_idx: np.float64,
                                               output_name: np.float64},
                          scalars=dict(), simulation_fields=[field])


def bilinear_interpolation_kernel(field: str, output_name: str | None = None,
                                  dimensions: tuple = ("y", "x")) -> ParticleKernel:
    """Return a ParticleKernel that performs bilinear interpolation of field data."""
    valid_dimensions = {("z", "y"), ("z", "x"), ("y", "x")}
    if dimensions not in valid_dimensions:
        raise ValueError(f"Invalid dimensions: {dimensions}. Valid options are {valid_dimensions}.")
    d0, d1 = dimensions[0] + "idx", dimensions[1] + "idx"
    if output_name is None:
        output_name = field
    fn = partial(bilinear_interpolation_kernel_function, dimension_idx0=d0, dimension_idx1=d1, field_name=field,
                 particle_name=output_name)
    return ParticleKernel(fn, particle_fields={"status": np.uint8, d0: np.float64, d1: np.float64,
                                               output_name: np.float64},
                          scalars=dict(), simulation_fields=[field])


def trilinear_interpolation_kernel(field: str, output_name: str | None = None) -> ParticleKernel:
    """Return a ParticleKernel that performs trilinear interpolation of field data."""
    if output_name is None:
        output_name = field
    fn = partial(trilinear_interpolation_kernel_function, field_name=field, particle_name=output_name)
    return ParticleKernel(fn, particle_fields={"status": np.uint8, "zidx": np.float64, "yidx": np.float64,
                                               "xidx": np.float64, output_name: np.float64},
                          scalars=dict(), simulation_fields=[field])


def vertical_interpolation_kernel(field: str, output_name: str | None = None) -> ParticleKernel:
    """Return a ParticleKernel that performs vertical linear interpolation of field data."""
    return linear_interpolation_kernel(field, output_name, dimension="z")


def horizontal_interpolation_kernel(field: str, output_name: str | None = None) -> ParticleKernel:
    """Return a ParticleKernel that performs horizontal bilinear interpolation of field data."""
    return bilinear_interpolation_kernel(field, output_name, dimensions=("y", "x"))
