"""CPU oracle for the advection hot path.  TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker (or the timed CPU
baseline).  The product package ``offline_particles_amd`` never imports it.

Contents
- ``opx_oracle.c`` / ``liboracle.so``: FP64 C restatement of the reference
  Cython kernels (validation, RK2 stages, AB3 + post) and of the full-grid time
  blend, OpenMP ``schedule(static)`` like the reference's ``prange``.
- ``driver.py``: a restatement of the reference's Python driver layer
  (``Timestepper.get_time_index``/``advance_time``, ``RK2Timestepper`` and
  ``ABTimestepper`` launch order, ``TimeDependentField.get_field_data`` blend
  semantics) that sequences the C kernels exactly as the reference does.
- ``ref_numerics.pyx``: a driver that compiles the reference's own
  ``linear.pxd`` and ``_vertical_coordinate.pxd`` into ``oracle/_ref`` (built
  only where the reference tree is present) to pin the oracle's L0 numerics.

Pinning status per oracle function is listed in ``driver.PINNING``.
"""

from .driver import (  # noqa: F401
    OracleField,
    OracleSimulation,
    lib,
    load,
    PINNING,
)
