"""compute_z kernel (mirrors kernels/roms/_roms_core.pyx:17-112).

``compute_z_kernel(particle_field)`` writes the depth z of each active
particle's (zidx, yidx, xidx) into ``particle_field`` with h and zeta
(bilinear) and C (linear), one HIP kernel (``opx_compute_z``).
"""

from __future__ import annotations

import ctypes
import functools

import numpy as np
import torch

from ... import _lib
from .._kernels import ParticleKernel
from .._marshal import _check_tensor, field_struct, stream_handle

__all__ = ["compute_z_kernel", "compute_z_kernel_function"]


def compute_z_kernel_function(particles, scalars, fielddata, particle_field) -> None:
    """Compute the physical vertical coordinate for particles."""
    for k in ("hc", "NZ"):
        if k not in scalars:
            raise KeyError(k)
    st = _check_tensor(particles["status"], "status", torch.uint8)
    z, y, x, out = (_check_tensor(particles[k], k, torch.float64)
                    for k in ("zidx", "yidx", "xidx", particle_field))
    h, zeta, C = (field_struct(fielddata[k], k) for k in ("h", "zeta", "C"))
    _lib.call("opx_compute_z", len(particles), st, z, y, x, ctypes.byref(h), ctypes.byref(zeta), ctypes.byref(C),
              float(scalars["hc"]), int(scalars["NZ"]), out, stream_handle())


def compute_z_kernel(particle_field: str = "z") -> ParticleKernel:
    """Return a kernel that computes the physical vertical coordinate for particles."""
    fn = functools.partial(compute_z_kernel_function, particle_field=particle_field)
    return ParticleKernel(
        fn,
        particle_fields={"status": np.uint8, "zidx": np.float64, "yidx": np.float64, "xidx": np.float64,
                         particle_field: np.float64},
        scalars={"hc": np.float64, "NZ": np.int32},
        simulation_fields=["h", "C", "zeta"],
    )
