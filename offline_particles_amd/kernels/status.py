"""Particle status codes (values of the reference's kernels/status.pxd:1-14)."""

from enum import IntEnum


class ParticleStatus(IntEnum):
    INACTIVE = 1 << 7  # bit 7 marks a particle the kernels skip
    NORMAL = 0
    NONFINITE = 1 | (1 << 7)
    OUT_OF_DOMAIN = 2 | (1 << 7)
    MULTISTEP_1 = 10
    MULTISTEP_2 = 11
