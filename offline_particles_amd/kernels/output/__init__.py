"""Kernels for computing output quantities (mirrors kernels/output/__init__.py)."""

from .linear_interpolation import (
    bilinear_interpolation_kernel,
    horizontal_interpolation_kernel,
    linear_interpolation_kernel,
    trilinear_interpolation_kernel,
    vertical_interpolation_kernel,
)

__all__ = [
    "linear_interpolation_kernel",
    "bilinear_interpolation_kernel",
    "trilinear_interpolation_kernel",
    "vertical_interpolation_kernel",
    "horizontal_interpolation_kernel",
]
