"""RK2 with ROMS w advection (mirrors kernels/roms/rk2_w_advection.pyx:1-455).

Three stages, each one HIP kernel:
- step 1 at (x_n, t_n): z from zidx, then u/dx, v/dy, w (rk2_w_advection.pyx:31-153);
- step 2 at the intermediate point x_n + alpha dt k1 and t_n + alpha dt, with
  zidx from z by the inverse stretching (:155-275);
- update with b1 = 1/(2 alpha) on stage 1 and b2 = 1 - b1 on stage 2, as the
  reference code does, then zidx from z at t_{n+1} (:278-354).
"""

from __future__ import annotations

import numpy as np

from ... import _lib
from .._kernels import ParticleKernel
from .._marshal import byref, fields_struct, particles_struct, scalars_struct, stream_handle

__all__ = [
    "rk2_w_advection_step_1", "rk2_w_advection_step_2", "rk2_w_advection_update",
    "rk2_w_advection_step_1_kernel", "rk2_w_advection_step_2_kernel", "rk2_w_advection_update_kernel",
]

_FULL = ("u", "v", "w", "dx", "dy", "h", "C", "zeta")
_VERT = ("h", "C", "zeta")


def _launch(entry, particles, scalars, fielddata, fields, need, required):
    p = particles_struct(particles, required)
    s = scalars_struct(scalars, need)
    f = fields_struct(fielddata, fields)
    _lib.call(entry, byref(p), byref(s), byref(f), stream_handle())


def rk2_w_advection_step_1(particles, scalars, fielddata) -> None:
    """First step of RK2 with ROMS w advection."""
    _launch("opx_rk2_step_1", particles, scalars, fielddata, _FULL, ("hc", "NZ"),
            ("z", "_dxidx1", "_dyidx1", "_dz1"))


def rk2_w_advection_step_2(particles, scalars, fielddata) -> None:
    """Second step of RK2 with ROMS w advection."""
    _launch("opx_rk2_step_2", particles, scalars, fielddata, _FULL, ("_dt", "_RK2_alpha", "hc", "NZ"),
            ("z", "_dxidx1", "_dyidx1", "_dz1", "_dxidx2", "_dyidx2", "_dz2"))


def rk2_w_advection_update(particles, scalars, fielddata) -> None:
    """Update step of RK2 with ROMS w advection."""
    _launch("opx_rk2_update", particles, scalars, fielddata, _VERT, ("_dt", "_RK2_alpha", "hc", "NZ"),
            ("z", "_dxidx1", "_dyidx1", "_dz1", "_dxidx2", "_dyidx2", "_dz2"))


_F8 = np.float64
rk2_w_advection_step_1_kernel = ParticleKernel(
    rk2_w_advection_step_1,
    particle_fields={"status": np.uint8, "zidx": _F8, "yidx": _F8, "xidx": _F8, "z": _F8,
                     "_dxidx1": _F8, "_dyidx1": _F8, "_dz1": _F8},
    scalars={"hc": np.float64, "NZ": np.int32},
    simulation_fields=list(_FULL),
)
rk2_w_advection_step_2_kernel = ParticleKernel(
    rk2_w_advection_step_2,
    particle_fields={"status": np.uint8, "zidx": _F8, "yidx": _F8, "xidx": _F8, "z": _F8,
                     "_dxidx1": _F8, "_dyidx1": _F8, "_dz1": _F8, "_dxidx2": _F8, "_dyidx2": _F8, "_dz2": _F8},
    scalars={"_dt": np.float64, "_RK2_alpha": np.float64, "hc": np.float64, "NZ": np.int32},
    simulation_fields=list(_FULL),
)
rk2_w_advection_update_kernel = ParticleKernel(
    rk2_w_advection_update,
    particle_fields={"status": np.uint8, "zidx": _F8, "yidx": _F8, "xidx": _F8, "z": _F8,
                     "_dxidx1": _F8, "_dyidx1": _F8, "_dz1": _F8, "_dxidx2": _F8, "_dyidx2": _F8, "_dz2": _F8},
    scalars={"_dt": np.float64, "_RK2_alpha": np.float64, "hc": np.float64, "NZ": np.int32},
    simulation_fields=list(_VERT),
)
