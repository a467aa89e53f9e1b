"""ctypes binding of libopx.so (declared in include/opx.h).

The library is required: if it is missing or fails to load, every kernel call
raises ``RuntimeError`` (there is no CPU fallback).  Build it with
``python -m offline_particles_amd._build`` or ``__graft_entry__.build()``.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _build

LIB_PATH = _build.LIB_PATH

# ---- enums (opx.h) ----
OPX_F64, OPX_F32, OPX_BF16 = 0, 1, 2
OPX_OK, OPX_ERR_INVALID, OPX_ERR_HIP, OPX_ERR_UNSUPPORTED = 0, -1, -2, -3


class OpxError(RuntimeError):
    """An error returned by libopx."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


class OpxInvalidArgument(OpxError, ValueError):
    """libopx rejected an argument (OPX_ERR_INVALID)."""


_vp = ctypes.c_void_p


OPX_PACK_NONE, OPX_PACK_QUAD, OPX_PACK_OCT, OPX_PACK_TPAIR, OPX_PACK_HEX = 0, 1, 2, 3, 4


class opx_field(ctypes.Structure):
    _fields_ = [
        ("prev", _vp), ("next", _vp),
        ("dtype", ctypes.c_int32), ("ndim", ctypes.c_int32),
        ("shape", ctypes.c_int64 * 3),
        ("offsets", ctypes.c_double * 3),
        ("ft", ctypes.c_double), ("gt", ctypes.c_double),
        ("packed", _vp), ("packed_kind", ctypes.c_int32), ("packed_stride", ctypes.c_int32),
    ]


ROMS_FIELD_NAMES = ("u", "v", "w", "dx", "dy", "h", "C", "zeta")


class opx_roms_fields(ctypes.Structure):
    _fields_ = [(n, opx_field) for n in ROMS_FIELD_NAMES]


class opx_scalars(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double), ("alpha", ctypes.c_double), ("hc", ctypes.c_double),
        ("NZ", ctypes.c_int32), ("pad_", ctypes.c_int32),
        ("zidx_min", ctypes.c_double), ("zidx_max", ctypes.c_double),
        ("yidx_min", ctypes.c_double), ("yidx_max", ctypes.c_double),
        ("xidx_min", ctypes.c_double), ("xidx_max", ctypes.c_double),
    ]


PARTICLE_PTR_NAMES = ("zidx", "yidx", "xidx", "z", "dz0", "dz1", "dz2", "dyidx0", "dyidx1", "dyidx2",
                      "dxidx0", "dxidx1", "dxidx2")


class opx_particles(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("status", _vp)] + [(n, _vp) for n in PARTICLE_PTR_NAMES]


# name -> argtypes; every entry point returns int
_SIGNATURES = {
    "opx_finite_indices": [ctypes.POINTER(opx_particles), _vp],
    "opx_domain_bounds": [ctypes.POINTER(opx_particles), ctypes.POINTER(opx_scalars), _vp],
    "opx_validate": [ctypes.POINTER(opx_particles), ctypes.POINTER(opx_scalars), _vp],
    "opx_rk2_step_1": [ctypes.POINTER(opx_particles), ctypes.POINTER(opx_scalars),
                       ctypes.POINTER(opx_roms_fields), _vp],
    "opx_rk2_step_2": [ctypes.POINTER(opx_particles), ctypes.POINTER(opx_scalars),
                       ctypes.POINTER(opx_roms_fields), _vp],
    "opx_rk2_update": [ctypes.POINTER(opx_particles), ctypes.POINTER(opx_scalars),
                       ctypes.POINTER(opx_roms_fields), _vp],
    "opx_ab3_w_advection": [ctypes.POINTER(opx_particles), ctypes.POINTER(opx_scalars),
                            ctypes.POINTER(opx_roms_fields), _vp],
    "opx_ab3_post_w_advection": [ctypes.POINTER(opx_particles), ctypes.POINTER(opx_scalars),
                                 ctypes.POINTER(opx_roms_fields), _vp],
}


class opx_split_state(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("status", _vp), ("cell", _vp), ("fz", _vp), ("fy", _vp), ("fx", _vp),
                ("hz1", _vp), ("hy1", _vp), ("hx1", _vp), ("hz2", _vp), ("hy2", _vp), ("hx2", _vp),
                ("pid", _vp), ("rel", _vp), ("z", _vp), ("rho_t", _vp)]


class opx_grid(ctypes.Structure):
    _fields_ = [("NZ", ctypes.c_int32), ("NY", ctypes.c_int32), ("NX", ctypes.c_int32),
                ("xbits", ctypes.c_int32), ("ybits", ctypes.c_int32), ("zbits", ctypes.c_int32),
                ("hc", ctypes.c_double),
                ("zidx_min", ctypes.c_double), ("zidx_max", ctypes.c_double),
                ("yidx_min", ctypes.c_double), ("yidx_max", ctypes.c_double),
                ("xidx_min", ctypes.c_double), ("xidx_max", ctypes.c_double)]


OPX_LAYOUT_RING, OPX_LAYOUT_PAIR, OPX_LAYOUT_QUAD, OPX_LAYOUT_PAIR_LDS, OPX_LAYOUT_OCT, OPX_LAYOUT_YPAIR = 0, 1, 2, 3, 4, 5
OPX_FLAG_TIME_OUT_OF_RANGE, OPX_FLAG_UNPACKED_PAIR = 1, 2


class opx_window(ctypes.Structure):
    _fields_ = [("u", _vp), ("v", _vp), ("w", _vp), ("zeta", _vp),
                ("u_stride", ctypes.c_int64), ("v_stride", ctypes.c_int64), ("w_stride", ctypes.c_int64),
                ("zeta_stride", ctypes.c_int64), ("dtype3", ctypes.c_int32), ("nslots", ctypes.c_int32),
                ("hdd", _vp), ("C", _vp), ("layout", ctypes.c_int32), ("npacked", ctypes.c_int32),
                ("pu", _vp), ("pv", _vp), ("pw", _vp), ("p2", _vp),
                ("pu_stride", ctypes.c_uint64), ("pv_stride", ctypes.c_uint64), ("pw_stride", ctypes.c_uint64),
                ("p2_stride", ctypes.c_uint64), ("dy", _vp), ("wlevels", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


OPX_SAMPLE_LINEAR, OPX_SAMPLE_BILINEAR, OPX_SAMPLE_TRILINEAR, OPX_SAMPLE_COMPUTE_Z = 1, 2, 3, 4


class opx_sample(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("field", ctypes.c_int32), ("field2", ctypes.c_int32),
                ("field3", ctypes.c_int32)]


class opx_release_clock(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("start_step", ctypes.c_int32), ("steps_left", ctypes.c_int32)]


class opx_release_step(ctypes.Structure):
    _fields_ = [("ft", ctypes.c_float), ("ft1", ctypes.c_float), ("ftm", ctypes.c_float), ("slots", ctypes.c_uint32)]


_P = ctypes.POINTER
_SIGNATURES.update({
    "opx_clock_advance": [_vp, _vp, ctypes.c_int32, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp,
                          ctypes.c_double, ctypes.c_double, _vp, _vp],
    "opx_clock_advance_steps": [_vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp, _vp, ctypes.c_int32,
                                ctypes.c_int32, ctypes.c_int32, _vp, ctypes.c_double, ctypes.c_double, _vp, _vp],
    "opx_rk2_fused_dev": [_P(opx_split_state), _P(opx_grid), _vp, ctypes.c_int32, ctypes.c_int32, _vp,
                          ctypes.c_double, ctypes.c_double, _vp, ctypes.c_int32, _vp],
    "opx_ab3_isopycnal_following": [_P(opx_particles), _vp, _P(opx_scalars), _P(opx_roms_fields), _P(opx_field),
                                    ctypes.c_double, ctypes.c_double, _vp],
    "opx_compute_z": [ctypes.c_int64, _vp, _vp, _vp, _vp, _P(opx_field), _P(opx_field), _P(opx_field),
                      ctypes.c_double, ctypes.c_int32, _vp, _vp],
    "opx_interp_linear": [ctypes.c_int64, _vp, _vp, _P(opx_field), _vp, _vp],
    "opx_interp_bilinear": [ctypes.c_int64, _vp, _vp, _vp, _P(opx_field), _vp, _vp],
    "opx_interp_trilinear": [ctypes.c_int64, _vp, _vp, _vp, _vp, _P(opx_field), _vp, _vp],
    "opx_rk2_fused": [_P(opx_split_state), _P(opx_grid), _P(opx_window), _vp, ctypes.c_double, ctypes.c_double, _vp,
                      ctypes.c_int32, _vp],
    "opx_rk2_fused_steps": [_P(opx_split_state), _P(opx_grid), _P(opx_window), _vp, ctypes.c_int32, ctypes.c_int32,
                            ctypes.c_double, ctypes.c_double, _vp, ctypes.c_int32, _vp, _vp],
    "opx_rk2_fused_steps_dev": [_P(opx_split_state), _P(opx_grid), _vp, ctypes.c_int32, ctypes.c_int32, _vp,
                                ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double, _vp, ctypes.c_int32,
                                _vp, _vp],
    "opx_ab3_fused_dev": [_P(opx_split_state), _P(opx_grid), _vp, ctypes.c_int32, ctypes.c_int32, _vp,
                          ctypes.c_double, _vp, ctypes.c_int32, _vp],
    "opx_ab3_fused": [_P(opx_split_state), _P(opx_grid), _P(opx_window), _vp, ctypes.c_double, _vp, ctypes.c_int32,
                      _vp],
    "opx_ab3_fused_steps": [_P(opx_split_state), _P(opx_grid), _P(opx_window), _vp, ctypes.c_int32, ctypes.c_int32,
                            ctypes.c_double, _vp, ctypes.c_int32, _vp, _vp],
    "opx_ab3_fused_steps_dev": [_P(opx_split_state), _P(opx_grid), _vp, ctypes.c_int32, ctypes.c_int32, _vp,
                                ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _vp, ctypes.c_int32, _vp, _vp],
    "opx_ab3_iso_fused_steps": [_P(opx_split_state), _P(opx_grid), _P(opx_window), _vp, ctypes.c_int32,
                                ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double, _vp, ctypes.c_int32,
                                _vp, _vp],
    "opx_ab3_iso_fused_steps_dev": [_P(opx_split_state), _P(opx_grid), _vp, ctypes.c_int32, _vp, ctypes.c_int32,
                                    ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double, _vp,
                                    ctypes.c_int32, _vp, _vp],
    "opx_sample_split": [_P(opx_split_state), _P(opx_grid), _vp, ctypes.c_int32, ctypes.c_int32, _vp, ctypes.c_int32,
                         ctypes.c_double, _vp, ctypes.c_int32, _vp],
    "opx_split_from_f64": [_P(opx_split_state), _P(opx_grid), _vp, _vp, _vp, _vp],
    "opx_split_to_f64": [_P(opx_split_state), _P(opx_grid), _vp, _vp, _vp, _vp],
    "opx_rebin_workspace_bytes": [ctypes.c_int64, _P(ctypes.c_uint64)],
    "opx_packed_bytes": [_P(opx_grid), ctypes.c_int32, ctypes.c_int32, _P(ctypes.c_uint64), _P(ctypes.c_uint64),
                         _P(ctypes.c_uint64), _P(ctypes.c_uint64)],
    "opx_pack_window": [_P(opx_window), _P(opx_grid), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp,
                        _vp, _vp],
    "opx_packed_field_bytes": [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                               _P(ctypes.c_uint64)],
    "opx_pack_field": [_vp, _vp, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _vp, _vp],
    "opx_pack_field_as": [_vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                          _vp, _vp],
    "opx_pack_window_boxes": [_P(opx_window), _P(opx_grid), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp,
                              ctypes.c_int32, _vp, _vp, _vp, _vp, _vp],
    "opx_pack_window_rows": [_P(opx_window), _P(opx_grid), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                             ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp, _vp, _vp],
    "opx_tile_occupancy": [_P(opx_split_state), _P(opx_grid), ctypes.c_int32, _vp, ctypes.c_int32, ctypes.c_int32,
                           _vp],
    "opx_copy_boxes_h2d": [_vp, _vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, _vp,
                           ctypes.c_int32, _vp],
    "opx_gather_boxes_h2d": [_vp, _vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, _vp,
                             ctypes.c_int32, _vp],
    "opx_rebin": [_P(opx_split_state), _P(opx_split_state), _P(opx_grid), ctypes.c_int32, _vp, ctypes.c_int32, _vp,
                  ctypes.c_uint64, _vp, _vp, _vp],
    "opx_spatial_order_workspace_bytes": [ctypes.c_int64, _P(ctypes.c_uint64)],
    "opx_spatial_order": [ctypes.c_int64, _vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                          ctypes.c_int32, _vp, ctypes.c_uint64, _vp, _vp],
    "opx_gather": [ctypes.c_int64, _vp, _vp, _vp, ctypes.c_int32, _vp],
    "opx_rebin_noalloc": [_P(opx_split_state), _P(opx_split_state), _P(opx_grid), ctypes.c_int32, _vp,
                          ctypes.c_int32, _vp, _vp, _vp],
    "opx_extract_release": [_P(opx_split_state), _P(opx_grid), ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64, _vp,
                            _vp, _vp, _vp, _vp],
})

# extra signatures registered by other modules
EXTRA_SIGNATURES: dict = {}

_lib = None
_load_error: Exception | None = None


def load():
    """Load libopx.so once; raise RuntimeError if it is missing."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libopx.so not found at {LIB_PATH}; build it with `python -m offline_particles_amd._build` "
            "(there is no CPU fallback)")
    try:
        L = ctypes.CDLL(LIB_PATH)
    except OSError as e:  # pragma: no cover - depends on the machine
        _load_error = e
        raise RuntimeError(f"failed to load {LIB_PATH}: {e}") from e
    L.opx_last_error.restype = ctypes.c_char_p
    L.opx_last_error.argtypes = []
    L.opx_abi_version.restype = ctypes.c_int
    for name, args in {**_SIGNATURES, **EXTRA_SIGNATURES}.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _lib = L
    return L


def register(name: str, argtypes, restype=ctypes.c_int):
    """Declare an additional entry point's signature (applied at load)."""
    EXTRA_SIGNATURES[name] = argtypes
    if _lib is not None:
        fn = getattr(_lib, name)
        fn.argtypes = argtypes
        fn.restype = restype


def call(name: str, *args) -> None:
    """Call an entry point; raise OpxError on a non-zero return code."""
    L = load()
    rc = getattr(L, name)(*args)
    if rc != OPX_OK:
        msg = L.opx_last_error().decode(errors="replace")
        if rc == OPX_ERR_INVALID:
            raise OpxInvalidArgument(name, rc, msg)
        raise OpxError(name, rc, msg)


def declared_symbols(header: str | None = None) -> list[str]:
    """Function names declared in include/opx.h (used by the export test)."""
    import re

    header = header or os.path.join(_build.INCLUDE, "opx.h")
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(opx_[a-z0-9_]+)\s*\(", text)))


def storage_code(dtype) -> int:
    """opx_dtype for a torch or numpy dtype."""
    import torch

    if dtype in (torch.float64, np.float64) or dtype == np.dtype(np.float64):
        return OPX_F64
    if dtype in (torch.float32, np.float32) or dtype == np.dtype(np.float32):
        return OPX_F32
    if dtype == torch.bfloat16:
        return OPX_BF16
    raise TypeError(f"unsupported field dtype {dtype}")
