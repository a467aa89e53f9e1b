"""AB3 that tracks an isopycnal (mirrors kernels/roms/ab3_isopycnal_following.pyx:1-330).

Horizontal advection as ab3_w_advection; the vertical rate is
``w = w_max * drho / (rho_scale + |drho|)`` with ``drho = rho - rho_t`` (the
particle's target density), which drives each particle towards its target
isopycnal (:126-134).  One HIP kernel (``opx_ab3_isopycnal_following``); the
post step recomputes zidx from z exactly as ab3_post_w_advection does.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from ... import _lib
from .._kernels import ParticleKernel
from .._marshal import _check_tensor, field_struct, fields_struct, particles_struct, scalars_struct, stream_handle

__all__ = ["ab3_isopycnal_following", "ab3_post_isopycnal_following", "ab3_isopycnal_following_kernel",
           "ab3_post_isopycnal_following_kernel"]

_FIELDS = ("u", "v", "dx", "dy", "h", "C", "zeta")
_VERT = ("h", "C", "zeta")
_HIST = ("_dz0", "_dz1", "_dz2", "_dyidx0", "_dyidx1", "_dyidx2", "_dxidx0", "_dxidx1", "_dxidx2")


def ab3_isopycnal_following(particles, scalars, fielddata) -> None:
    """Advect particles using the 3rd-order Adams-Bashforth scheme towards their target isopycnal."""
    p = particles_struct(particles, ("z",) + _HIST)
    rho_t = _check_tensor(particles["rho_t"], "rho_t", torch.float64)
    for k in ("w_max", "rho_scale"):
        if k not in scalars:
            raise KeyError(k)
    s = scalars_struct(scalars, ("_dt", "hc", "NZ"))
    f = fields_struct(fielddata, _FIELDS)
    rho = field_struct(fielddata["rho"], "rho")
    _lib.call("opx_ab3_isopycnal_following", ctypes.byref(p), rho_t, ctypes.byref(s), ctypes.byref(f),
              ctypes.byref(rho), float(scalars["w_max"]), float(scalars["rho_scale"]), stream_handle())


def ab3_post_isopycnal_following(particles, scalars, fielddata) -> None:
    """Post-advection step to update particle zidx from z (ab3_isopycnal_following.pyx:193-242)."""
    p = particles_struct(particles, ("z",))
    s = scalars_struct(scalars, ("hc", "NZ"))
    f = fields_struct(fielddata, _VERT)
    _lib.call("opx_ab3_post_w_advection", ctypes.byref(p), ctypes.byref(s), ctypes.byref(f), stream_handle())


_F8 = np.float64
ab3_isopycnal_following_kernel = ParticleKernel(
    ab3_isopycnal_following,
    particle_fields={"status": np.uint8, "zidx": _F8, "yidx": _F8, "xidx": _F8, "z": _F8, "rho_t": _F8,
                     **{k: _F8 for k in _HIST}},
    scalars={"_dt": np.float64, "hc": np.float64, "NZ": np.int32, "w_max": np.float64, "rho_scale": np.float64},
    simulation_fields=["u", "v", "rho", "dx", "dy", "h", "C", "zeta"],
)
ab3_post_isopycnal_following_kernel = ParticleKernel(
    ab3_post_isopycnal_following,
    particle_fields={"status": np.uint8, "zidx": _F8, "yidx": _F8, "xidx": _F8, "z": _F8},
    scalars={"hc": np.float64, "NZ": np.int32},
    simulation_fields=list(_VERT),
)
