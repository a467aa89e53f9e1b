"""AB3 with ROMS w advection (mirrors kernels/roms/ab3_w_advection.pyx:1-295).

``ab3_w_advection`` (one HIP kernel): f0 = (u/dx, v/dy, w) at (x_n, t_n),
Euler -> AB2 -> AB3 start-up from the MULTISTEP_1/MULTISTEP_2 status codes,
z from zidx, the AB3 update of (z, y, x) and the history shift
(ab3_w_advection.pyx:20-178).  ``ab3_post_w_advection`` recomputes zidx from z
with h and zeta at t_{n+1} (:180-229).
"""

from __future__ import annotations

import numpy as np

from ... import _lib
from .._kernels import ParticleKernel
from .._marshal import byref, fields_struct, particles_struct, scalars_struct, stream_handle

__all__ = ["ab3_w_advection", "ab3_post_w_advection", "ab3_w_advection_kernel", "ab3_post_w_advection_kernel",
           "rk2_startup_history_kernel"]

_FULL = ("u", "v", "w", "dx", "dy", "h", "C", "zeta")
_VERT = ("h", "C", "zeta")
_HIST = ("_dz0", "_dz1", "_dz2", "_dyidx0", "_dyidx1", "_dyidx2", "_dxidx0", "_dxidx1", "_dxidx2")


def ab3_w_advection(particles, scalars, fielddata) -> None:
    """Advect particles with the 3rd-order Adams-Bashforth scheme using w."""
    p = particles_struct(particles, ("z",) + _HIST)
    s = scalars_struct(scalars, ("_dt", "hc", "NZ"))
    f = fields_struct(fielddata, _FULL)
    _lib.call("opx_ab3_w_advection", byref(p), byref(s), byref(f), stream_handle())


def ab3_post_w_advection(particles, scalars, fielddata) -> None:
    """Post-advection step: zidx from the updated z."""
    p = particles_struct(particles, ("z",))
    s = scalars_struct(scalars, ("hc", "NZ"))
    f = fields_struct(fielddata, _VERT)
    _lib.call("opx_ab3_post_w_advection", byref(p), byref(s), byref(f), stream_handle())


_F8 = np.float64
ab3_w_advection_kernel = ParticleKernel(
    ab3_w_advection,
    particle_fields={"status": np.uint8, "zidx": _F8, "yidx": _F8, "xidx": _F8, "z": _F8,
                     **{k: _F8 for k in _HIST}},
    scalars={"_dt": np.float64, "hc": np.float64, "NZ": np.int32},
    simulation_fields=list(_FULL),
)
ab3_post_w_advection_kernel = ParticleKernel(
    ab3_post_w_advection,
    particle_fields={"status": np.uint8, "zidx": _F8, "yidx": _F8, "xidx": _F8, "z": _F8},
    scalars={"hc": np.float64, "NZ": np.int32},
    simulation_fields=list(_VERT),
)


def _rk2_startup_history(particles, scalars, fielddata, stage) -> None:
    """History bookkeeping of the RK2 start-up (timesteppers.ABRK2StartupTimestepper):
    "save" keeps the first step's stage-1 derivatives f(x0, t0) (left in _d*1) in
    _d*0; "restore" moves them to _d*2 after the second step, whose stage 1 left
    f(x1, t1) in _d*1.  Device-to-device copies of whole particle arrays."""
    for axis in ("z", "yidx", "xidx"):
        d0, d1, d2 = (particles[f"_d{axis}{k}"] for k in (0, 1, 2))
        if stage == "save":
            d0.copy_(d1)
        else:
            d2.copy_(d0)


def rk2_startup_history_kernel(stage: str) -> ParticleKernel:
    """The history step of the RK2 start-up of AB3 ("save" after the first RK2 step,
    "restore" after the second); no counterpart in the reference."""
    if stage not in ("save", "restore"):
        raise ValueError(f"stage must be 'save' or 'restore', not {stage!r}")
    import functools

    fn = functools.partial(_rk2_startup_history, stage=stage)
    fn.__qualname__ = f"rk2_startup_history_{stage}"
    return ParticleKernel(fn, particle_fields={k: _F8 for k in _HIST}, scalars={}, simulation_fields=[])
