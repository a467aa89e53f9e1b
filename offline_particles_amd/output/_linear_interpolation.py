"""Outputs that linearly interpolate field data (mirrors output/_linear_interpolation.py:18-56)."""

from __future__ import annotations

import numpy as np
import torch

from ..kernels.output import bilinear_interpolation_kernel, linear_interpolation_kernel, trilinear_interpolation_kernel
from ._output import Output

DMASK_DIM_MAPPING_2D = {
    (True, True, False): ("z", "y"),
    (True, False, True): ("z", "x"),
    (False, True, True): ("y", "x"),
}


def _dtype_name(dtype) -> str:
    """numpy name of a field's storage dtype (the reference's fields hold numpy arrays)."""
    if isinstance(dtype, torch.dtype):
        return "bfloat16" if dtype == torch.bfloat16 else str(torch.empty((), dtype=dtype).numpy().dtype)
    return str(np.dtype(dtype))


def linearly_interpolate_fields(fieldset, *variables: str, particle_field_prefix: str = "_output") -> list:
    """Output variables that linearly interpolate field data.

    The particle field of each output is ``f"{particle_field_prefix}_{dtype}"``
    (outputs of one dtype share it; each output event fills it just before
    it is written).
    """
    dims = ("z", "y", "x")
    outputs = []
    for var in variables:
        if var not in fieldset:
            raise KeyError(f"Field '{var}' not found in fieldset.")
        field = fieldset[var]
        dmask = tuple(field.dmask)
        ndim = field.nspatial_dims
        particle_field = f"{particle_field_prefix}_{_dtype_name(field.dtype)}"
        if ndim == 1:
            kernel = linear_interpolation_kernel(var, particle_field, dims[dmask.index(True)])
        elif ndim == 2:
            kernel = bilinear_interpolation_kernel(var, particle_field, DMASK_DIM_MAPPING_2D[dmask])
        elif ndim == 3:
            kernel = trilinear_interpolation_kernel(var, particle_field)
        else:
            raise ValueError(f"Field '{var}' has unsupported number of dimensions: {ndim}")
        outputs.append(Output(var, kernel, particle_field=particle_field))
    return outputs
