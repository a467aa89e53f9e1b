"""Simulation output (mirrors the reference's offline_particles.output).

``ZarrOutputBuilder`` / ``ZarrOutputWriter`` write the reference's Zarr v3
trajectory stores (time, particle) through this package's own Zarr v3 layer
(``_zarr``; the zarr package is not needed); ``MemoryOutputWriter`` keeps the
same rows in host memory.
"""

from ._linear_interpolation import linearly_interpolate_fields
from ._memory_writer import MemoryOutputBuilder, MemoryOutputWriter
from ._output import AbstractOutputWriter, AbstractOutputWriterBuilder, Output
from .._zarr import ZarrArray, open_array
from ._zarr_writer import DEFAULT_CHUNKSIZE, ZarrOutputBuilder, ZarrOutputWriter

__all__ = [
    "AbstractOutputWriter",
    "AbstractOutputWriterBuilder",
    "Output",
    "MemoryOutputWriter",
    "MemoryOutputBuilder",
    "ZarrOutputWriter",
    "ZarrOutputBuilder",
    "ZarrArray",
    "open_array",
    "DEFAULT_CHUNKSIZE",
    "linearly_interpolate_fields",
]
