"""Reference-structured CPU driver over the C oracle.  TEST INFRASTRUCTURE ONLY.

Restates the reference's Python driver layer (paths relative to
``src/offline_particles/`` in the reference tree):

- ``Timestepper.get_time_index`` (timesteppers.py:64-85) and
  ``advance_time`` (timesteppers.py:157-161);
- the launch order of ``RK2Timestepper.timestep_particles``
  (timesteppers.py:241-257) and ``ABTimestepper.timestep_particles``
  (timesteppers.py:296-305) with the validation pre-step and AB3 post-step of
  the ROMS factories (models/roms/__init__.py:37-106);
- ``TimeDependentField.get_field_data`` (fields.py:394-430): ``divmod`` of the
  time index, field-dtype blend weights and a full-grid blend into an FP64
  buffer on every launch; static non-FP64 fields are converted to FP64 on every
  launch like ``unpack_fielddata_*`` (kernels/_core.pyx:5-24);
- ``Launcher.launch_kernel``'s per-launch bounding box (launcher.py:122-148).

The snapshot axis may be cyclic: snapshot ``It`` reads ``data[It % K]``, so a
long run can use a short set of unique snapshots (the reference reads
``data[It]`` of an explicit array; a cyclic array gives the same values).
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

# Pinning of each oracle function: what checks it (see tests/test_oracle.py).
PINNING = {
    "orc_truncate_index/linear/bilinear/trilinear": "compiled reference linear.pxd (oracle/_ref) + "
    "tests/golden vectors + affine-exactness and extrapolation closed forms",
    "orc_compute_z/compute_zidx": "compiled reference _vertical_coordinate.pxd (oracle/_ref) + golden "
    "vectors + closed form zidx = NZ(z/h+1)-1/2 for C = sigma, zeta = 0 + round trip",
    "orc_blend_f64/f32": "numpy multiply/multiply/add in the field dtype (bitwise, same seeds)",
    "orc_finite_indices/domain_bounds": "exact comparisons on hand-built NaN/inf/boundary cases",
    "orc_rk2_step_1/step_2/update": "closed form for u = a + b t: a dt + b t dt + (alpha - 1/2) b dt^2 "
    "(distinguishes the code's b1 = 1/(2 alpha) from the tableau) + dx/dt = lambda x amplification + "
    "uniform-w vertical closed form",
    "orc_ab3/orc_ab3_post": "closed form for u = a + b t across Euler -> AB2 -> AB3 start-up "
    "(x_N - x_0 = a N dt + b (t_N^2 - t_0^2)/2 - b dt^2/2) + uniform-w vertical closed form",
    "ab3_rk2 (RK2 start-up of AB3, composed)": "closed forms: u = a + b t gives a N dt + b dt^2 N (N - 1) / 2 + "
    "b dt^2 (2 (alpha - 1/2) + (N - 2) / 2) after N steps; dx/dt = lambda x gives the RK2 amplification for two "
    "steps, then AB3 from f(x0), f(x1); uniform w over a flat bed",
    "driver (get_time_index, launch order, FP32 blend weights)": "restated from timesteppers.py / fields.py; "
    "numpy semantics checked directly; composition pinned by the RK2/AB3 closed forms with "
    "time-dependent fields",
}

INACTIVE = 128
NORMAL = 0
NONFINITE = 129
OUT_OF_DOMAIN = 130
MULTISTEP_1 = 10
MULTISTEP_2 = 11

_dp = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class _Field(ctypes.Structure):
    _fields_ = [
        ("a", _dp),
        ("n0", ctypes.c_int64), ("n1", ctypes.c_int64), ("n2", ctypes.c_int64),
        ("o0", ctypes.c_double), ("o1", ctypes.c_double), ("o2", ctypes.c_double),
    ]


class _Fields(ctypes.Structure):
    _fields_ = [(n, _Field) for n in ("u", "v", "w", "dx", "dy", "h", "C", "zeta")]


class _Particles(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("status", _u8p)] + [
        (n, _dp)
        for n in ("zidx", "yidx", "xidx", "z", "dz0", "dz1", "dz2", "dyidx0", "dyidx1", "dyidx2",
                  "dxidx0", "dxidx1", "dxidx2")
    ]


class _Scalars(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double), ("alpha", ctypes.c_double), ("hc", ctypes.c_double),
        ("NZ", ctypes.c_int32), ("pad_", ctypes.c_int32),
        ("zmin", ctypes.c_double), ("zmax", ctypes.c_double),
        ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
        ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
    ]


_lib = None


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference/src/offline_particles"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def load():
    """Load liboracle.so, building it first if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    L = ctypes.CDLL(_LIB_PATH)
    P, S, F = ctypes.POINTER(_Particles), ctypes.POINTER(_Scalars), ctypes.POINTER(_Fields)
    for name, args in {
        "orc_finite_indices": [P],
        "orc_domain_bounds": [P, S],
        "orc_rk2_step_1": [P, S, F],
        "orc_rk2_step_2": [P, S, F],
        "orc_rk2_update": [P, S, F],
        "orc_ab3": [P, S, F],
        "orc_ab3_post": [P, S, F],
    }.items():
        getattr(L, name).argtypes = args
        getattr(L, name).restype = None
    L.orc_ab3_iso.argtypes = [P, S, F, ctypes.POINTER(_Field), _dp, ctypes.c_double, ctypes.c_double]
    L.orc_ab3_iso.restype = None
    L.orc_output_trilinear.argtypes = [P, ctypes.POINTER(_Field), _dp]
    L.orc_output_trilinear.restype = None
    L.orc_output_compute_z.argtypes = [P, S, F, _dp]
    L.orc_output_compute_z.restype = None
    L.orc_blend_f64.argtypes = [_dp, _dp, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _dp]
    L.orc_blend_f32.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_float,
                                ctypes.c_float, _dp]
    L.orc_f32_to_f64.argtypes = [ctypes.c_void_p, ctypes.c_int64, _dp]
    L.orc_bbox.argtypes = [P, _dp]
    L.orc_linear.argtypes = [_dp, ctypes.c_int64, ctypes.c_double]
    L.orc_linear.restype = ctypes.c_double
    L.orc_bilinear.argtypes = [_dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_double]
    L.orc_bilinear.restype = ctypes.c_double
    L.orc_trilinear.argtypes = [_dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double]
    L.orc_trilinear.restype = ctypes.c_double
    L.orc_compute_z.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double]
    L.orc_compute_z.restype = ctypes.c_double
    L.orc_compute_zidx.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, _dp, ctypes.c_int64, ctypes.c_double]
    L.orc_compute_zidx.restype = ctypes.c_double
    L.orc_num_threads.restype = ctypes.c_int
    L.orc_set_num_threads.argtypes = [ctypes.c_int]
    _lib = L
    return L


def lib():
    return load()


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


# ---------------------------------------------------------------------------
# L0 vector helpers (used by the pinning tests)
# ---------------------------------------------------------------------------

def linear(a, idx):
    L = load()
    a = np.ascontiguousarray(a, np.float64)
    return np.array([L.orc_linear(_ptr(a), a.shape[0], float(i)) for i in np.atleast_1d(idx)])


def bilinear(a, i0, i1):
    L = load()
    a = np.ascontiguousarray(a, np.float64)
    return np.array([L.orc_bilinear(_ptr(a), a.shape[0], a.shape[1], float(x), float(y))
                     for x, y in zip(np.atleast_1d(i0), np.atleast_1d(i1))])


def trilinear(a, i0, i1, i2):
    L = load()
    a = np.ascontiguousarray(a, np.float64)
    return np.array([L.orc_trilinear(_ptr(a), *a.shape, float(x), float(y), float(z))
                     for x, y, z in zip(np.atleast_1d(i0), np.atleast_1d(i1), np.atleast_1d(i2))])


def compute_z(zidx, NZ, hc, h, C, zeta):
    L = load()
    return np.array([L.orc_compute_z(float(a), int(NZ), float(hc), float(b), float(c), float(d))
                     for a, b, c, d in np.broadcast(zidx, h, C, zeta)])


def compute_zidx(z, h, zeta, hc, NZ, Carr, C_offset):
    L = load()
    Carr = np.ascontiguousarray(Carr, np.float64)
    return np.array([L.orc_compute_zidx(float(a), float(b), float(c), float(hc), int(NZ), _ptr(Carr),
                                        Carr.shape[0], float(C_offset))
                     for a, b, c in np.broadcast(z, h, zeta)])


# ---------------------------------------------------------------------------
# Field + simulation driver
# ---------------------------------------------------------------------------

@dataclass
class OracleField:
    """A field as the reference's Fieldset holds it.

    ``data`` is the spatial array (static) or a (K, *spatial) array of
    snapshots (time dependent; snapshot It is ``data[It % K]`` when ``cyclic``,
    else ``data[It]``).  ``offsets`` are the active-dimension stagger offsets
    (spatial_arrays.py:29-40, 136-143).
    """

    data: np.ndarray
    offsets: tuple
    time_dependent: bool = False
    cyclic: bool = True
    _out: np.ndarray | None = field(default=None, repr=False)

    def snapshot(self, It: int) -> np.ndarray:
        K = self.data.shape[0]
        if self.cyclic:
            return self.data[It % K]
        if It < 0 or It > K - 1:
            raise IndexError(f"snapshot {It} out of range")
        return self.data[It]

    def field_data(self, tidx: float, L) -> np.ndarray:
        """Blended (or converted) FP64 array at time index ``tidx`` (fields.py:394-430)."""
        if not self.time_dependent:
            if self.data.dtype == np.float64:
                return np.ascontiguousarray(self.data)
            if self._out is None:
                self._out = np.empty(self.data.shape, np.float64)
            src = np.ascontiguousarray(self.data, np.float32)
            L.orc_f32_to_f64(src.ctypes.data, src.size, _ptr(self._out))
            return self._out
        It, ft = divmod(tidx, 1)
        It = int(It)
        prev = np.ascontiguousarray(self.snapshot(It))
        nxt = np.ascontiguousarray(self.snapshot(It + 1))
        if self._out is None:
            self._out = np.empty(prev.shape, np.float64)
        dtype = self.data.dtype
        ft = dtype.type(ft)
        gt = dtype.type(1 - ft)
        if dtype == np.float64:
            L.orc_blend_f64(_ptr(prev), _ptr(nxt), prev.size, float(ft), float(gt), _ptr(self._out))
        elif dtype == np.float32:
            L.orc_blend_f32(prev.ctypes.data, nxt.ctypes.data, prev.size, float(ft), float(gt),
                            _ptr(self._out))
        else:
            raise TypeError(f"oracle supports float32/float64 fields, got {dtype}")
        return self._out


_KERNEL_FIELDS = {
    "full": ("u", "v", "w", "dx", "dy", "h", "C", "zeta"),
    "iso": ("u", "v", "dx", "dy", "h", "C", "zeta"),
    "vertical": ("h", "C", "zeta"),
}

_PARTICLE_FIELDS = ("zidx", "yidx", "xidx", "z", "dz0", "dz1", "dz2", "dyidx0", "dyidx1", "dyidx2",
                    "dxidx0", "dxidx1", "dxidx2")


class OracleSimulation:
    """Reference-structured simulation on the CPU oracle.

    ``scheme`` is ``"rk2"`` (rk2_w_advection_timestepper), ``"ab3"``
    (ab3_w_advection_timestepper), ``"ab3_rk2"`` (the same with its first two
    steps taken by the RK2 kernels) or ``"ab3_iso"``
    (ab3_isopycnal_following_timestepper: ``fields`` has "rho", and
    ``rho_t``, ``w_max``, ``rho_scale`` are given); all use the validation
    pre-step.
    Particle state lives in ``self.p`` (numpy arrays named as the reference's
    particle fields, with RK2's ``_d*1/_d*2`` sharing the AB3 names).
    """

    def __init__(self, fields: dict, time_array, dt, *, scheme: str, nparticles: int, hc: float,
                 NZ: int, bounds, alpha: float = 2 / 3, time=None, threads: int | None = None,
                 bbox: bool = True, rho_t=None, w_max: float = 0.0, rho_scale: float = 1.0):
        if scheme not in ("rk2", "ab3", "ab3_iso", "ab3_rk2"):
            raise ValueError(scheme)
        # "ab3_rk2": AB3 whose first two steps are RK2 steps (a start-up variant
        # composed from the reference's kernels; no reference mode): the RK2 and
        # AB3 kernels share the derivative fields _d*1/_d*2
        # (rk2_w_advection.pyx:378-380, ab3_w_advection.pyx:250-258), so f(x0, t0)
        # from step 1's stage 1 is saved and put in _d*2 after step 2, whose stage
        # 1 leaves f(x1, t1) in _d*1
        self._rk_steps_left = 2 if scheme == "ab3_rk2" else 0
        self.L = load()
        if threads:
            self.L.orc_set_num_threads(int(threads))
        self.fields = fields
        self.scheme = scheme
        self.time_array = np.asarray(time_array)
        if np.any(self.time_array[1:] <= self.time_array[:-1]):
            raise ValueError("time_array must be strictly increasing.")
        self.dt = np.float64(dt)
        self.alpha = alpha
        self.time = self.time_array[0] if time is None else time
        self.tidx = self.get_time_index(self.time)
        self.iteration = 0
        self.do_bbox = bbox
        n = int(nparticles)
        self.p = {"status": np.zeros(n, np.uint8)}
        for name in _PARTICLE_FIELDS:
            self.p[name] = np.zeros(n, np.float64)
        self._cp = _Particles(n, self.p["status"].ctypes.data_as(_u8p),
                              *[_ptr(self.p[k]) for k in _PARTICLE_FIELDS])
        zmin, zmax, ymin, ymax, xmin, xmax = (float(b) for b in bounds)
        self._scalars = _Scalars(float(self.dt), float(alpha), float(hc), int(NZ), 0,
                                 zmin, zmax, ymin, ymax, xmin, xmax)
        self._bbox = np.zeros(6)
        if scheme == "ab3_iso":
            self.p["rho_t"] = np.ascontiguousarray(np.broadcast_to(np.asarray(rho_t, np.float64), (n,)))
            self.w_max = float(w_max)
            self.rho_scale = float(rho_scale)

    # timesteppers.py:64-85
    def get_time_index(self, time):
        t = self.time_array
        if time < t[0] or time > t[-1]:
            raise ValueError("Time is out of bounds of the time array.")
        idx = np.searchsorted(t, time, side="right") - 1
        t0 = t[idx]
        t1 = t[idx + 1]
        fraction = (time - t0) / (t1 - t0)
        return idx + fraction

    # timesteppers.py:157-161
    def advance_time(self):
        self.time += self.dt
        self.tidx = self.get_time_index(self.time)
        self.iteration += 1

    def _fields(self, tidx, which):
        F = _Fields()
        for name in _KERNEL_FIELDS[which]:
            fld = self.fields[name]
            arr = fld.field_data(tidx, self.L)
            shp = list(arr.shape) + [1] * (3 - arr.ndim)
            offs = list(fld.offsets) + [0.0] * (3 - len(fld.offsets))
            setattr(F, name, _Field(_ptr(arr), shp[0], shp[1], shp[2], *offs))
            # keep the array alive for the call
            setattr(self, f"_keep_{name}", arr)
        return F

    def _launch(self, fn, tidx, which=None):
        if self.do_bbox:
            self.L.orc_bbox(ctypes.byref(self._cp), _ptr(self._bbox))
        if which is None:
            fn(ctypes.byref(self._cp), ctypes.byref(self._scalars))
        else:
            F = self._fields(tidx, which)
            fn(ctypes.byref(self._cp), ctypes.byref(self._scalars), ctypes.byref(F))

    def _validate(self):
        if self.do_bbox:
            self.L.orc_bbox(ctypes.byref(self._cp), _ptr(self._bbox))
        self.L.orc_finite_indices(ctypes.byref(self._cp))
        self.L.orc_domain_bounds(ctypes.byref(self._cp), ctypes.byref(self._scalars))

    def _rk2_step(self):
        # timesteppers.py:241-257
        L = self.L
        self._validate()
        self._launch(L.orc_rk2_step_1, self.tidx, "full")
        intermediate_time = self.time + self.alpha * self.dt
        intermediate_tidx = self.get_time_index(intermediate_time)
        self._launch(L.orc_rk2_step_2, intermediate_tidx, "full")
        self.advance_time()
        self._launch(L.orc_rk2_update, self.tidx, "vertical")

    def step(self):
        L = self.L
        if self._rk_steps_left:
            self._rk2_step()
            self._rk_steps_left -= 1
            for axis in ("z", "yidx", "xidx"):
                if self._rk_steps_left:
                    self.p[f"d{axis}0"][:] = self.p[f"d{axis}1"]
                else:
                    self.p[f"d{axis}2"][:] = self.p[f"d{axis}0"]
        elif self.scheme == "rk2":
            self._rk2_step()
        elif self.scheme == "ab3_iso":
            # ABTimestepper (timesteppers.py:296-305) with the isopycnal kernels
            # (models/roms/__init__.py:109-140)
            self._validate()
            if self.do_bbox:
                self.L.orc_bbox(ctypes.byref(self._cp), _ptr(self._bbox))
            F = self._fields(self.tidx, "iso")
            rf = self.fields["rho"]
            ra = rf.field_data(self.tidx, self.L)
            self._keep_rho = ra
            R = _Field(_ptr(ra), *ra.shape, *rf.offsets)
            L.orc_ab3_iso(ctypes.byref(self._cp), ctypes.byref(self._scalars), ctypes.byref(F), ctypes.byref(R),
                          _ptr(self.p["rho_t"]), self.w_max, self.rho_scale)
            self.advance_time()
            self._launch(L.orc_ab3_post, self.tidx, "vertical")
        else:
            # timesteppers.py:296-305
            self._validate()
            self._launch(L.orc_ab3, self.tidx, "full")
            self.advance_time()
            self._launch(L.orc_ab3_post, self.tidx, "vertical")

    def run(self, nsteps: int):
        for _ in range(nsteps):
            self.step()

    # output events at the current time index: the launcher blends the field
    # over the whole grid (fields.py:394-430), then the particle loop runs
    def output_trilinear(self, name: str, out: np.ndarray) -> np.ndarray:
        """trilinear_interpolation_kernel(name) (linear_interpolation.pyx:107-136) into ``out``."""
        fld = self.fields[name]
        arr = fld.field_data(self.tidx, self.L)
        F = _Field(_ptr(arr), *arr.shape, *fld.offsets)
        self.L.orc_output_trilinear(ctypes.byref(self._cp), ctypes.byref(F), _ptr(out))
        return out

    def output_compute_z(self, out: np.ndarray) -> np.ndarray:
        """compute_z_kernel (_roms_core.pyx:17-73) into ``out``."""
        F = self._fields(self.tidx, "vertical")
        self.L.orc_output_compute_z(ctypes.byref(self._cp), ctypes.byref(self._scalars), ctypes.byref(F), _ptr(out))
        return out
