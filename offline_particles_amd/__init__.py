"""MI355X (gfx950) implementation of offline_particles' advection hot path.

The public names mirror the reference package ``offline_particles``; kernels
run as hand-written HIP kernels in ``libopx.so`` (include/opx.h) on particle
and field data resident in device memory.
"""

from . import events, kernels, output
from .fields import BlendWindow, FieldData, StaticField, TimeDependentField
from .fieldset import Fieldset
from .models import roms
from .particles import Particles, ParticlesView
from .simulation import Event, Simulation, SimulationBuilder, SimulationState
from .spatial_arrays import BBox, ChunkedArray, DeviceArray, Stagger
from .timesteppers import ABTimestepper, RK2Timestepper, Timestepper

__all__ = [
    "TimeDependentField", "StaticField", "FieldData", "BlendWindow", "Fieldset", "Simulation",
    "SimulationBuilder", "Event", "SimulationState", "RK2Timestepper", "ABTimestepper", "Timestepper",
    "Particles", "ParticlesView", "Stagger", "BBox", "DeviceArray", "ChunkedArray", "events", "kernels", "output",
    "roms",
]
