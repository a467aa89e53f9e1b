"""Particle kernels (mirrors the reference's offline_particles.kernels)."""

from ._kernels import (
    DEFAULT_PARTICLE_FIELDS,
    KernelFunction,
    ParticleKernel,
    merge_particle_fields,
    merge_scalars,
    merge_simulation_fields,
)
from .status import ParticleStatus

__all__ = [
    "ParticleKernel",
    "KernelFunction",
    "ParticleStatus",
    "DEFAULT_PARTICLE_FIELDS",
    "merge_particle_fields",
    "merge_scalars",
    "merge_simulation_fields",
]
