"""Validation kernels (mirrors the reference's kernels/validation.pyx:1-124).

``finite_indices`` flags non-finite indices NONFINITE (129);
``domain_bounds`` flags indices outside the inclusive bounds OUT_OF_DOMAIN
(130); ``validation_kernel`` chains them.  Both skip INACTIVE particles.
Each launches one HIP kernel through libopx.
"""

from __future__ import annotations

import numpy as np

from .. import _lib
from ._kernels import ParticleKernel
from ._marshal import byref, particles_struct, scalars_struct, stream_handle

__all__ = ["finite_indices", "domain_bounds", "finite_indices_kernel", "domain_bounds_kernel",
           "validation_kernel"]

_BOUNDS = ("zidx_min", "zidx_max", "yidx_min", "yidx_max", "xidx_min", "xidx_max")


def finite_indices(particles, scalars, fielddata) -> None:
    """validation.pyx:21-43"""
    p = particles_struct(particles)
    _lib.call("opx_finite_indices", byref(p), stream_handle())


def domain_bounds(particles, scalars, fielddata) -> None:
    """validation.pyx:45-77"""
    p = particles_struct(particles)
    s = scalars_struct(scalars, _BOUNDS)
    _lib.call("opx_domain_bounds", byref(p), byref(s), stream_handle())


_POS = {"status": np.uint8, "zidx": np.float64, "yidx": np.float64, "xidx": np.float64}

finite_indices_kernel = ParticleKernel(finite_indices, particle_fields=_POS, scalars={}, simulation_fields=[])
domain_bounds_kernel = ParticleKernel(domain_bounds, particle_fields=_POS,
                                      scalars={k: np.float64 for k in _BOUNDS}, simulation_fields=[])
validation_kernel = ParticleKernel.chain(finite_indices_kernel, domain_bounds_kernel)
