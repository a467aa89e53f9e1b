"""Timesteppers (mirrors the reference's timesteppers.py:16-305).

The clock is FP64 (or datetime64) on the host; ``get_time_index`` and
``advance_time`` restate timesteppers.py:64-85 and :157-161 exactly, so each
launch sees the same fractional snapshot index as in the reference.  The
launch order of ``RK2Timestepper`` (timesteppers.py:241-257) and
``ABTimestepper`` (:296-305) is unchanged.
"""

from __future__ import annotations

import abc

import numpy as np

from .launcher import ScalarSource


class Timestepper(abc.ABC):
    """Handles particle advection time stepping."""

    _dt_scalar = ScalarSource("_dt", lambda self, tidx: self._normalised_dt)
    _time_scalar = ScalarSource("_time", lambda self, tidx: self.time)
    _tidx_scalar = ScalarSource("_tidx", lambda self, tidx: self._tidx)

    def __init__(self, time_array, dt, *, time_unit=None, time=None, iteration: int = 0,
                 index_padding: int = 0) -> None:
        time_array = np.asarray(time_array)
        if np.any(time_array[1:] <= time_array[:-1]):
            raise ValueError("time_array must be strictly increasing.")
        self._time_array = time_array
        if time_unit is None:
            if isinstance(dt, np.floating):
                time_unit = np.float64(1.0)
            else:
                raise ValueError("time_unit must be specified for dimensional time.")
        self._time_unit = time_unit
        self.set_dt(dt)
        if time is None:
            time = self._time_array[0]
        self.set_time(time)
        self.set_iteration(iteration)
        self._index_padding = 0
        self.set_index_padding(index_padding)

    @property
    def time_array(self) -> np.ndarray:
        return self._time_array

    def get_time_index(self, time) -> np.float64:
        """Fractional snapshot index of ``time`` (timesteppers.py:64-85)."""
        time_array = self._time_array
        if time < time_array[0] or time > time_array[-1]:
            raise ValueError("Time is out of bounds of the time array.")
        idx = np.searchsorted(time_array, time, side="right") - 1
        t0 = time_array[idx]
        t1 = time_array[idx + 1]
        fraction = (time - t0) / (t1 - t0)
        return idx + fraction

    def set_dt(self, dt) -> None:
        try:
            self._normalised_dt = np.float64(dt / self._time_unit)
        except Exception as e:
            raise TypeError(f"dt must be of the same type as time_unit={self._time_unit!r}") from e

    def set_time(self, time) -> None:
        try:
            _ = time + self.dt
        except Exception as e:
            raise TypeError(f"time must be compatible with dt={self.dt!r}") from e
        self._tidx = self.get_time_index(time)
        self._time = time

    def set_iteration(self, iteration: int) -> None:
        if iteration < 0:
            raise ValueError("Iteration must be non-negative.")
        self._iteration = iteration

    def set_index_padding(self, index_padding: int, force: bool = False) -> None:
        if index_padding < 0:
            raise ValueError("Index padding must be non-negative.")
        if force or index_padding > self._index_padding:
            self._index_padding = index_padding

    time_unit = property(lambda self: self._time_unit)
    dt = property(lambda self: self._normalised_dt * self._time_unit)
    time = property(lambda self: self._time)
    time_array = property(lambda self: self._time_array)
    iteration = property(lambda self: self._iteration)
    tidx = property(lambda self: self._tidx)
    index_padding = property(lambda self: self._index_padding)

    def advance_time(self) -> None:
        """time += dt, then the new time index (timesteppers.py:157-161)."""
        self._time += self.dt
        self._tidx = self.get_time_index(self._time)
        self._iteration += 1

    @property
    @abc.abstractmethod
    def kernels(self) -> tuple:
        ...

    @abc.abstractmethod
    def timestep_particles(self, particles, launcher) -> None:
        ...


class RK2Timestepper(Timestepper):
    """Two-stage explicit RK2 with parameter alpha (timesteppers.py:175-257)."""

    _alpha_scalar = ScalarSource("_RK2_alpha", lambda self, tidx: self._alpha)

    def __init__(self, time_array, dt, rk_step_1_kernel, rk_step_2_kernel, rk_update_kernel, *,
                 alpha: float = 2 / 3, time_unit=None, time=None, iteration: int = 0, index_padding: int = 0,
                 pre_step_kernel=None, post_step_kernel=None) -> None:
        super().__init__(time_array, dt, time_unit=time_unit, time=time, index_padding=index_padding,
                         iteration=iteration)
        self._rk_step_1_kernel = rk_step_1_kernel
        self._rk_step_2_kernel = rk_step_2_kernel
        self._rk_update_kernel = rk_update_kernel
        self._pre_step_kernel = pre_step_kernel
        self._post_step_kernel = post_step_kernel
        self._alpha = alpha

    @property
    def alpha(self) -> float:
        return self._alpha

    @property
    def kernels(self) -> tuple:
        pre = (self._pre_step_kernel,) if self._pre_step_kernel is not None else ()
        post = (self._post_step_kernel,) if self._post_step_kernel is not None else ()
        return pre + (self._rk_step_1_kernel, self._rk_step_2_kernel, self._rk_update_kernel) + post

    def timestep_particles(self, particles, launcher) -> None:
        if self._pre_step_kernel is not None:
            launcher.launch_kernel(self._pre_step_kernel, particles, self._tidx)
        launcher.launch_kernel(self._rk_step_1_kernel, particles, self._tidx)
        intermediate_time = self.time + self._alpha * self.dt
        intermediate_tidx = self.get_time_index(intermediate_time)
        launcher.launch_kernel(self._rk_step_2_kernel, particles, intermediate_tidx)
        self.advance_time()
        launcher.launch_kernel(self._rk_update_kernel, particles, self._tidx)
        if self._post_step_kernel is not None:
            launcher.launch_kernel(self._post_step_kernel, particles, self._tidx)


class ABTimestepper(Timestepper):
    """Adams-Bashforth timestepper (timesteppers.py:260-305)."""

    def __init__(self, time_array, dt, ab_kernel, *, time_unit=None, time=None, iteration: int = 0,
                 index_padding: int = 0, pre_step_kernel=None, post_step_kernel=None) -> None:
        super().__init__(time_array, dt, time_unit=time_unit, time=time, index_padding=index_padding,
                         iteration=iteration)
        self._ab_kernel = ab_kernel
        self._pre_step_kernel = pre_step_kernel
        self._post_step_kernel = post_step_kernel

    @property
    def kernels(self) -> tuple:
        pre = (self._pre_step_kernel,) if self._pre_step_kernel is not None else ()
        post = (self._post_step_kernel,) if self._post_step_kernel is not None else ()
        return pre + (self._ab_kernel,) + post

    def timestep_particles(self, particles, launcher) -> None:
        if self._pre_step_kernel is not None:
            launcher.launch_kernel(self._pre_step_kernel, particles, self._tidx)
        launcher.launch_kernel(self._ab_kernel, particles, self._tidx)
        self.advance_time()
        if self._post_step_kernel is not None:
            launcher.launch_kernel(self._post_step_kernel, particles, self._tidx)


class ABRK2StartupTimestepper(ABTimestepper):
    """AB3 whose first two steps are RK2 steps: a start-up variant with no
    counterpart in the reference, whose ABTimestepper starts through the
    MULTISTEP status codes (Euler, then AB2, ab3_w_advection.pyx:121-138).

    The first two steps run the RK2 sequence of RK2Timestepper
    (timesteppers.py:241-257: pre-step, stage 1 at t, stage 2 at
    t + alpha dt, advance, update at t + dt).  The RK2 and AB3 kernels share
    the derivative fields ``_d*1`` and ``_d*2`` (rk2_w_advection.pyx:378-380,
    ab3_w_advection.pyx:250-258), so the history is kept in them: after the
    first step stage 1's f(x0, t0) is saved in ``_d*0``; after the second,
    ``_d*2`` = f(x0, t0) while stage 1 left ``_d*1`` = f(x1, t1), the history
    AB3 needs at its first step from (x2, t2).  Particles start with status
    NORMAL (a MULTISTEP status would restart their history with Euler).
    """

    _alpha_scalar = ScalarSource("_RK2_alpha", lambda self, tidx: self._alpha)

    def __init__(self, time_array, dt, ab_kernel, rk_step_1_kernel, rk_step_2_kernel, rk_update_kernel, *,
                 alpha: float = 2 / 3, time_unit=None, time=None, iteration: int = 0, index_padding: int = 0,
                 pre_step_kernel=None, post_step_kernel=None) -> None:
        super().__init__(time_array, dt, ab_kernel, time_unit=time_unit, time=time, iteration=iteration,
                         index_padding=index_padding, pre_step_kernel=pre_step_kernel,
                         post_step_kernel=post_step_kernel)
        self._rk = (rk_step_1_kernel, rk_step_2_kernel, rk_update_kernel)
        self._alpha = alpha
        self._rk_steps_left = 2
        from .kernels.roms.ab3_w_advection import rk2_startup_history_kernel

        self._history = (rk2_startup_history_kernel("save"), rk2_startup_history_kernel("restore"))

    @property
    def alpha(self) -> float:
        return self._alpha

    @property
    def kernels(self) -> tuple:
        return super().kernels + self._rk + self._history

    def timestep_particles(self, particles, launcher) -> None:
        if self._rk_steps_left == 0:
            super().timestep_particles(particles, launcher)
            return
        rk1, rk2, update = self._rk
        if self._pre_step_kernel is not None:
            launcher.launch_kernel(self._pre_step_kernel, particles, self._tidx)
        launcher.launch_kernel(rk1, particles, self._tidx)
        intermediate_tidx = self.get_time_index(self.time + self._alpha * self.dt)
        launcher.launch_kernel(rk2, particles, intermediate_tidx)
        self.advance_time()
        launcher.launch_kernel(update, particles, self._tidx)
        self._rk_steps_left -= 1
        launcher.launch_kernel(self._history[0 if self._rk_steps_left else 1], particles, self._tidx)
