"""ROMS timestepper factories (mirrors the reference's models/roms/__init__.py:37-106).

``rk2_w_advection_timestepper``, ``ab3_w_advection_timestepper`` and
``ab3_isopycnal_following_timestepper`` (:109-140) compose
the same kernels in the same order as the reference: the validation
pre-step, the RK2 stages (alpha = 2/3 by default) or the AB3 kernel and its
post-step, and an index padding of 5.
"""

from __future__ import annotations

from ...kernels.roms import (
    ab3_isopycnal_following_kernel,
    ab3_post_isopycnal_following_kernel,
    ab3_post_w_advection_kernel,
    ab3_w_advection_kernel,
    compute_z_kernel,
    rk2_w_advection_step_1_kernel,
    rk2_w_advection_step_2_kernel,
    rk2_w_advection_update_kernel,
)
from ...kernels.validation import validation_kernel
from ...timesteppers import ABRK2StartupTimestepper, ABTimestepper, RK2Timestepper

__all__ = [
    "compute_z_kernel",
    "ab3_w_advection_kernel", "ab3_post_w_advection_kernel", "ab3_w_advection_timestepper",
    "rk2_w_advection_step_1_kernel", "rk2_w_advection_step_2_kernel", "rk2_w_advection_update_kernel",
    "rk2_w_advection_timestepper",
    "ab3_isopycnal_following_kernel", "ab3_post_isopycnal_following_kernel", "ab3_isopycnal_following_timestepper",
]


def rk2_w_advection_timestepper(time_array, dt, *, time=None, time_unit=None, iteration: int = 0,
                                index_padding: int = 5, alpha: float = 2 / 3) -> RK2Timestepper:
    """RK2 timestepper with the ROMS w-advection kernels."""
    return RK2Timestepper(time_array, dt, rk_step_1_kernel=rk2_w_advection_step_1_kernel,
                          rk_step_2_kernel=rk2_w_advection_step_2_kernel,
                          rk_update_kernel=rk2_w_advection_update_kernel, time=time, time_unit=time_unit,
                          iteration=iteration, index_padding=index_padding, alpha=alpha,
                          pre_step_kernel=validation_kernel)


def ab3_w_advection_timestepper(time_array, dt, *, time=None, time_unit=None, iteration: int = 0,
                                index_padding: int = 5, startup: str = "multistep",
                                alpha: float = 2 / 3) -> ABTimestepper:
    """AB3 timestepper with the ROMS w-advection kernels.

    ``startup="multistep"`` (the reference's only form) starts each particle
    through its MULTISTEP status codes (Euler, AB2); ``startup="rk2"`` takes the
    first two steps with the RK2 w-advection kernels (parameter ``alpha``) and
    starts AB3 from their stage-1 derivatives (ABRK2StartupTimestepper)."""
    if startup == "multistep":
        return ABTimestepper(time_array, dt, ab_kernel=ab3_w_advection_kernel, time=time, time_unit=time_unit,
                             iteration=iteration, index_padding=index_padding, pre_step_kernel=validation_kernel,
                             post_step_kernel=ab3_post_w_advection_kernel)
    if startup == "rk2":
        return ABRK2StartupTimestepper(time_array, dt, ab3_w_advection_kernel, rk2_w_advection_step_1_kernel,
                                       rk2_w_advection_step_2_kernel, rk2_w_advection_update_kernel, alpha=alpha,
                                       time=time, time_unit=time_unit, iteration=iteration,
                                       index_padding=index_padding, pre_step_kernel=validation_kernel,
                                       post_step_kernel=ab3_post_w_advection_kernel)
    raise ValueError(f"startup must be 'multistep' or 'rk2', not {startup!r}")


def ab3_isopycnal_following_timestepper(time_array, dt, *, time=None, time_unit=None, iteration: int = 0,
                                        index_padding: int = 5) -> ABTimestepper:
    """AB3 timestepper with the ROMS isopycnal following kernels."""
    return ABTimestepper(time_array, dt, ab_kernel=ab3_isopycnal_following_kernel, time=time, time_unit=time_unit,
                         iteration=iteration, index_padding=index_padding, pre_step_kernel=validation_kernel,
                         post_step_kernel=ab3_post_isopycnal_following_kernel)
