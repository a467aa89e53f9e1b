"""Simulation (mirrors the reference's simulation.py:27-367 for the stepping path).

``Simulation.step`` runs one timestepper step and then any scheduled events;
``set_indices`` / ``set_particle_field`` write the device particle arrays.
Events and schedulers are the reference's minimal host-side plumbing
(events/_events.py, events/_schedulers.py) so that event kernels launch at the
same iterations/times.  Output writers (``add_output_writer``) turn their
outputs into events as the reference's builder does (simulation.py:316-367).
"""

from __future__ import annotations

import dataclasses
import itertools
import types
from typing import Callable

import numpy as np
import torch

from .fields import sorted_stencils
from .kernels import merge_particle_fields
from .launcher import Launcher
from .particles import Particles, ParticlesView


@dataclasses.dataclass(frozen=True)
class SimulationState:
    time: object
    dt: object
    tidx: np.float64
    iteration: int
    particles: ParticlesView


class Event:
    """A function on the simulation state plus kernels launched before it."""

    def __init__(self, name: str, func: Callable, *kernels) -> None:
        self._name = name
        self._func = func
        self._kernels = kernels

    def __call__(self, state: SimulationState) -> None:
        self._func(state)

    name = property(lambda self: self._name)
    kernels = property(lambda self: self._kernels)


class IterationScheduler:
    """Triggers events every N iterations (events/_schedulers.py:13-71)."""

    def __init__(self) -> None:
        self._next = None
        self._events: dict = {}

    def _schedule_event(self, iteration, N, event) -> None:
        self._events.setdefault(iteration, []).append((N, event))

    @property
    def events(self):
        for lst in self._events.values():
            for _, e in lst:
                yield e

    def register_event(self, first: int, n: int, event: Event) -> None:
        self._schedule_event(first, n, event)
        self.set_next()

    def set_next(self) -> None:
        self._next = min(self._events.keys()) if self._events else None

    def __call__(self, iteration: int) -> list:
        out = []
        while self._next is not None and self._next <= iteration:
            for N, event in self._events.pop(self._next, []):
                out.append(event)
                self._schedule_event(self._next + N, N, event)
            self.set_next()
        return out


class TimeScheduler:
    """Triggers events every dt of simulated time (events/_schedulers.py:74-140)."""

    def __init__(self, *, forward: bool = True) -> None:
        self._forward = forward
        self._next_time = None
        self._events: dict = {}

    def _schedule_event(self, time, dt, event) -> None:
        self._events.setdefault(time, []).append((dt, event))

    @property
    def events(self):
        for lst in self._events.values():
            for _, e in lst:
                yield e

    def register_event(self, first, dt, event: Event) -> None:
        self._schedule_event(first, dt, event)
        self.set_next()

    def set_next(self) -> None:
        if not self._events:
            self._next_time = None
        else:
            self._next_time = min(self._events) if self._forward else max(self._events)

    def __call__(self, time) -> list:
        out = []
        while (nt := self._next_time) is not None and (nt <= time if self._forward else nt >= time):
            for dt, event in self._events.pop(nt, []):
                out.append(event)
                self._schedule_event(nt + dt, dt, event)
            self.set_next()
        return out


class Simulation:
    """A particle simulation on one device."""

    def __init__(self, nparticles: int, timestepper, fieldset, iteration_scheduler=None, time_scheduler=None,
                 output_writers=None, *, device=None, spatial_order_every: int = 0, spatial_order_xlow: int = 3) -> None:
        self._timestepper = timestepper
        self._fieldset = fieldset
        self._iteration_scheduler = iteration_scheduler or IterationScheduler()
        self._time_scheduler = time_scheduler or TimeScheduler()
        self._output_writers = types.MappingProxyType(dict(output_writers or {}))
        self._launcher = Launcher(fieldset)
        self._launcher.maybe_increase_index_padding(timestepper.index_padding)
        self._launcher.register_scalar_data_sources_from_object(self._timestepper)
        for event in itertools.chain(self._iteration_scheduler.events, self._time_scheduler.events):
            self._launcher.register_scalar_data_sources_from_object(event)
        kernels = list(self._timestepper.kernels)
        for event in itertools.chain(self._iteration_scheduler.events, self._time_scheduler.events):
            kernels.extend(event.kernels)
        particle_fields = merge_particle_fields(kernels)
        self._particles = Particles(nparticles, device=device, **particle_fields)
        # spatial ordering (opt-in): every `spatial_order_every` steps the
        # particle arrays are permuted into the Morton order of the particles'
        # cells, so the stencil gathers of neighbouring threads share cache
        # lines; self._order[j] is the particle held in slot j (None: identity)
        if spatial_order_every < 0:
            raise ValueError("spatial_order_every must be non-negative")
        self._order_every = int(spatial_order_every)
        self._order_xlow = int(spatial_order_xlow)
        self._order = None
        self._since_order = 0
        self._order_work = None
        self._particles_view = ParticlesView(self._particles, lambda: self._order)
        self._invoke_events()

    timestepper = property(lambda self: self._timestepper)
    fieldset = property(lambda self: self._fieldset)
    launcher = property(lambda self: self._launcher)
    time = property(lambda self: self._timestepper.time)
    iteration = property(lambda self: self._timestepper.iteration)
    dt = property(lambda self: self._timestepper.dt)
    tidx = property(lambda self: self._timestepper.tidx)
    particles = property(lambda self: self._particles_view)
    iteration_scheduler = property(lambda self: self._iteration_scheduler)
    time_scheduler = property(lambda self: self._time_scheduler)
    output_writers = property(lambda self: self._output_writers)

    @property
    def state(self) -> SimulationState:
        return SimulationState(time=self.time, dt=self.dt, tidx=self.tidx, iteration=self.iteration,
                               particles=self._particles_view)

    def _invoke_events(self) -> None:
        for event in itertools.chain(self._iteration_scheduler(self.iteration), self._time_scheduler(self.time)):
            for kernel in event.kernels:
                self._launcher.launch_kernel(kernel, self._particles, self.tidx)
            event(self.state)

    def _as_field(self, name, values):
        target = self._particles[name]
        if isinstance(values, torch.Tensor):
            t = values.to(device=target.device, dtype=target.dtype)
        else:
            np_dtype = torch.empty((), dtype=target.dtype).numpy().dtype
            arr = np.asarray(values, dtype=np_dtype)
            t = torch.from_numpy(np.ascontiguousarray(arr)).to(target.device)
        return torch.broadcast_to(t, target.shape)

    def _to_slots(self, t):
        """Values given in the particles' order, arranged in slot order."""
        return t if self._order is None else t[self._order.long()]

    def set_indices(self, zidx=None, yidx=None, xidx=None) -> None:
        """Set particle indices (simulation.py:178-207); all inputs are checked before any write."""
        vals = {}
        for name, v in (("zidx", zidx), ("yidx", yidx), ("xidx", xidx)):
            if v is not None:
                vals[name] = self._as_field(name, v)
        for name, t in vals.items():
            self._particles[name].copy_(self._to_slots(t))

    def set_particle_field(self, field_name: str, values) -> None:
        """Set one particle field (simulation.py:209-223)."""
        self._particles[field_name].copy_(self._to_slots(self._as_field(field_name, values)))

    @property
    def spatial_order(self):
        """None, or the int32 device tensor whose entry j is the particle held in slot j."""
        return self._order

    def reorder(self) -> None:
        """Permute every particle array into the Morton order of the particles'
        cells now (opx_spatial_order, opx_gather); what the particles hold, and
        what ``particles``, ``set_indices``, ``set_particle_field`` and the output
        writers see, does not change."""
        import ctypes

        from . import _lib
        from .kernels._marshal import stream_handle

        P = self._particles
        n = len(P)
        self._since_order = 0
        if n == 0:
            return
        dev = P["zidx"].device
        if self._order_work is None:
            need = ctypes.c_uint64()
            _lib.call("opx_spatial_order_workspace_bytes", n, ctypes.byref(need))
            self._order_work = torch.empty(int(need.value), dtype=torch.uint8, device=dev)
        fs = self._fieldset
        perm = torch.empty(n, dtype=torch.int32, device=dev)
        st = stream_handle()
        _lib.call("opx_spatial_order", n, P["status"].data_ptr(), P["zidx"].data_ptr(), P["yidx"].data_ptr(),
                  P["xidx"].data_ptr(), fs.z_size, fs.y_size, fs.x_size, self._order_xlow,
                  self._order_work.data_ptr(), self._order_work.numel(), perm.data_ptr(), st)
        tmp = {}
        for name, a in P.arrays.items():
            b = tmp.get(a.dtype)
            if b is None:
                b = tmp[a.dtype] = torch.empty_like(a)
            _lib.call("opx_gather", n, perm.data_ptr(), a.data_ptr(), b.data_ptr(), a.element_size(), st)
            a.copy_(b)
        order = self._order if self._order is not None else torch.arange(n, dtype=torch.int32, device=dev)
        new = torch.empty_like(order)
        _lib.call("opx_gather", n, perm.data_ptr(), order.data_ptr(), new.data_ptr(), 4, st)
        self._order = new

    def step(self) -> None:
        """Advance the particles by one timestep, then run scheduled events."""
        if self._order_every:
            if self._since_order % self._order_every == 0:
                self.reorder()
            # the particles are sorted by cell: snapshot pairs as TPAIR copies
            with sorted_stencils():
                self._timestepper.timestep_particles(self._particles, self._launcher)
                self._since_order += 1
                self._invoke_events()
            return
        self._timestepper.timestep_particles(self._particles, self._launcher)
        self._since_order += 1
        self._invoke_events()


class SimulationBuilder:
    """Builds a Simulation with events and output writers (simulation.py:232-367)."""

    def __init__(self, timestepper, fieldset) -> None:
        self._timestepper = timestepper
        self._fieldset = fieldset
        self._iteration_scheduler = IterationScheduler()
        self._time_scheduler = TimeScheduler()
        self._output_writers: dict = {}

    def every_n(self, n: int, event: Event, *, first: int | None = None) -> None:
        if n <= 0:
            raise ValueError("n must be a positive integer.")
        self._iteration_scheduler.register_event(0 if first is None else first, n, event)

    def every_dt(self, dt, event: Event, *, first=None) -> None:
        t = self._timestepper.time
        if first is None:
            first = t
        try:
            _ = t + dt
        except TypeError as e:
            raise TypeError(f"Incompatible dt type {type(dt)} for timestepper time type {type(t)}") from e
        self._time_scheduler.register_event(first, dt, event)

    def add_event(self, event: Event, *, n=None, dt=None, first=None) -> None:
        if n is not None and dt is not None:
            raise ValueError("Cannot specify both n and dt.")
        if n is not None:
            self.every_n(n, event, first=first)
        elif dt is not None:
            self.every_dt(dt, event, first=first)
        else:
            raise ValueError("Either n or dt must be specified.")

    def add_output_writer(self, builder, *, n=None, dt=None, first=None) -> None:
        """Add an output writer; its events run every n iterations or every dt."""
        name = builder.name
        if name in self._output_writers:
            raise ValueError(f"Output writer '{name}' already exists.")
        self._output_writers[name] = (builder, {"n": n, "dt": dt, "first": first})

    def build_simulation(self, nparticles: int, *, device=None, spatial_order_every: int = 0,
                         spatial_order_xlow: int = 3) -> Simulation:
        writers = {}
        time_type = self._timestepper.time_array.dtype
        for name, (builder, kwargs) in self._output_writers.items():
            writers[name] = builder.build(nparticles, time_type)
            for event in writers[name].create_events():
                self.add_event(event, **kwargs)
        return Simulation(nparticles, self._timestepper, self._fieldset, self._iteration_scheduler,
                          self._time_scheduler, writers, device=device, spatial_order_every=spatial_order_every,
                          spatial_order_xlow=spatial_order_xlow)
