"""Kernel launcher (mirrors the reference's launcher.py:16-174).

``launch_kernel`` evaluates the kernel's declared scalars at the launch's time
index and fetches the declared fields' ``FieldData``, then calls the kernel.
The reference also reduces a particle bounding box on every launch
(launcher.py:122-148), which only its disk-backed chunked arrays use;
HBM-resident arrays ignore it (spatial_arrays.py:204-225), so here the box is
computed only when a field asks for it (``Field.needs_bbox``).
"""

from __future__ import annotations

from typing import Callable

from .spatial_arrays import BBox

_NO_BBOX = BBox(0.0, 0.0, 0.0, 0.0, 0.0, 0.0)


class ScalarSource:
    """Descriptor declaring a scalar data source; getter(self, tidx) -> scalar."""

    def __init__(self, name: str, getter: Callable) -> None:
        self.name = name
        self._getter = getter

    def __get__(self, obj, owner=None):
        if obj is None:
            return self

        def scalar_func(tidx):
            return self._getter(obj, tidx)

        return scalar_func


class Launcher:
    """Launches particle kernels with their scalars and field data."""

    def __init__(self, fieldset, *, index_padding: int = 0) -> None:
        self._scalar_data_sources: dict = {}
        self._fieldset = fieldset
        if index_padding < 0:
            raise ValueError("index_padding must be non-negative")
        self._index_padding = index_padding
        for name, value in self._fieldset.constants.items():
            self.register_scalar_data_source(name, self.create_value_scalar_source(value))

    @staticmethod
    def create_value_scalar_source(value):
        def value_func(tidx):
            return value

        return value_func

    def register_scalar_data_source(self, name: str, source) -> None:
        if name in self._scalar_data_sources:
            raise ValueError(
                f"Scalar data source '{name}' is already registered. Deregister it before registering a new one.")
        if name in self._fieldset.fields:
            raise ValueError(f"Scalar data source '{name}' conflicts with a field in the fieldset.")
        self._scalar_data_sources[name] = source

    def deregister_scalar_data_source(self, name: str) -> None:
        if name not in self._scalar_data_sources:
            raise ValueError(f"Scalar data source '{name}' is not registered.")
        del self._scalar_data_sources[name]

    def register_scalar_data_sources_from_object(self, obj) -> None:
        for attr_name in dir(type(obj)):
            attr = getattr(type(obj), attr_name)
            if isinstance(attr, ScalarSource):
                self.register_scalar_data_source(attr.name, getattr(obj, attr_name))

    @property
    def index_padding(self) -> int:
        return self._index_padding

    def maybe_increase_index_padding(self, padding: int) -> None:
        if padding > self._index_padding:
            self._index_padding = padding

    def construct_bbox(self, particles) -> BBox:
        """Bounding box of the particle indices plus padding (launcher.py:122-148)."""
        pad = self._index_padding
        z, y, x = particles["zidx"], particles["yidx"], particles["xidx"]
        return BBox(float(z.min()) - pad, float(z.max()) + pad, float(y.min()) - pad, float(y.max()) + pad,
                    float(x.min()) - pad, float(x.max()) + pad)

    def get_field_data(self, name: str, time_index: float, bbox: BBox):
        return self._fieldset[name].get_field_data(time_index, bbox)

    def _bbox_for(self, kernel, particles) -> BBox:
        if any(getattr(self._fieldset[n], "needs_bbox", False) for n in kernel.simulation_fields):
            return self.construct_bbox(particles)
        return _NO_BBOX

    def scalars_for(self, kernel, time_index: float) -> dict:
        return {name: self._scalar_data_sources[name](time_index) for name in kernel.scalars}

    def launch_kernel(self, kernel, particles, time_index: float) -> None:
        bbox = self._bbox_for(kernel, particles)
        scalars = self.scalars_for(kernel, time_index)
        fielddata = {name: self.get_field_data(name, time_index, bbox) for name in kernel.simulation_fields}
        kernel(particles, scalars, fielddata)
