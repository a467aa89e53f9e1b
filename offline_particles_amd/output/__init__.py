"""Simulation output (mirrors the reference's offline_particles.output).

The reference's ZarrOutputWriter needs the zarr package, which this
environment does not have; ``MemoryOutputWriter`` provides the same events
and checks with host-memory arrays.
"""

from ._linear_interpolation import linearly_interpolate_fields
from ._memory_writer import MemoryOutputBuilder, MemoryOutputWriter
from ._output import AbstractOutputWriter, AbstractOutputWriterBuilder, Output

__all__ = [
    "AbstractOutputWriter",
    "AbstractOutputWriterBuilder",
    "Output",
    "MemoryOutputWriter",
    "MemoryOutputBuilder",
    "linearly_interpolate_fields",
]
