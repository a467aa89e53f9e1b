"""Turn particles / scalars / FieldData into the C structs of include/opx.h.

Checks mirror what the reference's typed memoryviews enforce at kernel entry
(dtype, C-contiguity; kernels/_core.pyx:5-24 and e.g. ab3_w_advection.pyx:22-40)
plus the device check this path needs.  A missing particle field, scalar or
field raises ``KeyError`` like the reference's dictionary lookups.
"""

from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ..fields import BlendWindow

# opx_particles member -> reference particle field name
_PARTICLE_NAMES = {
    "zidx": "zidx", "yidx": "yidx", "xidx": "xidx", "z": "z",
    "dz0": "_dz0", "dz1": "_dz1", "dz2": "_dz2",
    "dyidx0": "_dyidx0", "dyidx1": "_dyidx1", "dyidx2": "_dyidx2",
    "dxidx0": "_dxidx0", "dxidx1": "_dxidx1", "dxidx2": "_dxidx2",
}


def _check_tensor(t: torch.Tensor, name: str, dtype: torch.dtype) -> int:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"particle field '{name}' must be a torch tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise ValueError(f"Buffer dtype mismatch for '{name}': expected {dtype}, got {t.dtype}")
    if not t.is_cuda:
        raise ValueError(f"particle field '{name}' must be in device memory (got {t.device})")
    if not t.is_contiguous():
        raise ValueError(f"particle field '{name}' must be C-contiguous")
    return t.data_ptr()


def particles_struct(particles, required=()) -> _lib.opx_particles:
    s = _lib.opx_particles()
    s.n = len(particles)
    s.status = _check_tensor(particles["status"], "status", torch.uint8)
    for member, name in _PARTICLE_NAMES.items():
        if name in particles:
            setattr(s, member, _check_tensor(particles[name], name, torch.float64))
        elif name in required:
            raise KeyError(name)
    return s


def scalars_struct(scalars: dict, need=()) -> _lib.opx_scalars:
    for k in need:
        if k not in scalars:
            raise KeyError(k)
    g = scalars.get
    return _lib.opx_scalars(
        float(g("_dt", 0.0)), float(g("_RK2_alpha", 0.0)), float(g("hc", 0.0)), int(g("NZ", 0)), 0,
        float(g("zidx_min", 0.0)), float(g("zidx_max", 0.0)),
        float(g("yidx_min", 0.0)), float(g("yidx_max", 0.0)),
        float(g("xidx_min", 0.0)), float(g("xidx_max", 0.0)),
    )


def _device_data(t: torch.Tensor, name: str) -> int:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"field '{name}' data must be a device tensor, got {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError(f"field '{name}' must be in device memory (got {t.device})")
    if not t.is_contiguous():
        raise ValueError(f"field '{name}' must be C-contiguous")
    return t.data_ptr()


def field_struct(fd, name: str) -> _lib.opx_field:
    arr, offsets = fd.array, fd.offsets
    f = _lib.opx_field()
    if isinstance(arr, BlendWindow):
        f.prev = _device_data(arr.prev, name)
        f.next = _device_data(arr.next, name)
        if arr.next.shape != arr.prev.shape or arr.next.dtype != arr.prev.dtype:
            raise ValueError(f"field '{name}': snapshot shapes/dtypes differ")
        f.ft, f.gt = arr.ft, arr.gt
        t = arr.prev
    else:
        f.prev = _device_data(arr, name)
        f.next = None
        t = arr
    f.dtype = _lib.storage_code(t.dtype)
    f.ndim = t.dim()
    if len(offsets) != t.dim():
        raise ValueError(f"field '{name}': {len(offsets)} offsets for {t.dim()} dims")
    for d in range(t.dim()):
        f.shape[d] = int(t.shape[d])
        f.offsets[d] = float(offsets[d])
    packed = getattr(fd, "packed", None)
    if packed is not None:
        f.packed = _device_data(packed[0], name)
        f.packed_kind = int(packed[1])
    return f


def merge_quads(quads, elem: int) -> torch.Tensor:
    """One buffer of 4 x `elem` bytes per node holding the QUAD copies `quads`
    (uint8 tensors of n x `elem` bytes, one per field) side by side, the rest
    zero: read with opx_field.packed_stride = 4 x elem, field j at j x elem."""
    n = quads[0].numel() // elem
    if len(quads) > 4 or any(q.numel() != n * elem for q in quads):
        raise ValueError("QUAD copies of different sizes")
    out = torch.zeros((n, 4, elem), dtype=torch.uint8, device=quads[0].device)
    for j, q in enumerate(quads):
        out[:, j] = q.view(n, elem)
    return out


def merge_elements(parts, sizes, stride: int) -> torch.Tensor:
    """One buffer of `stride` bytes per node holding the packed copies `parts`
    (uint8 tensors of n x sizes[j] bytes) side by side, part j at
    sum(sizes[:j]), the rest zero."""
    n = parts[0].numel() // sizes[0]
    if sum(sizes) > stride or any(p.numel() != n * e for p, e in zip(parts, sizes)):
        raise ValueError("packed copies of different node counts, or wider than the stride")
    out = torch.zeros((n, stride), dtype=torch.uint8, device=parts[0].device)
    at = 0
    for p, e in zip(parts, sizes):
        out[:, at:at + e] = p.view(n, e)
        at += e
    return out


def _merge_zeta_h(s, fielddata) -> None:
    """Point the zeta and h descriptors of `s` at one buffer holding, per node,
    zeta's OCT element (its snapshot pair) and then h's QUAD (64 bytes for FP32),
    when zeta has an OCT copy and h a QUAD copy of the same shape: the post step
    and compute_z then read h and zeta in one cache line.  The buffer is kept
    with zeta's copy of the pair (made again for each pair)."""
    fz, fh = s.zeta, s.h
    pz, ph = getattr(fielddata["zeta"], "packed", None), getattr(fielddata["h"], "packed", None)
    if pz is None or ph is None or pz[1] != _lib.OPX_PACK_OCT or ph[1] != _lib.OPX_PACK_QUAD:
        return
    if fz.ndim != 2 or fh.ndim != 2 or tuple(fz.shape[:2]) != tuple(fh.shape[:2]):
        return
    oct_ = 32 if fz.dtype == _lib.OPX_F32 else 16
    quad = 16 if fh.dtype == _lib.OPX_F32 else 8
    stride = 1 << (oct_ + quad - 1).bit_length()
    cached = getattr(pz[0], "_opx_zeta_h_merge", None)
    if cached is None or cached[0] is not ph[0]:
        # the cache holds h's copy, so its identity cannot be reused
        cached = (ph[0], merge_elements([pz[0], ph[0]], [oct_, quad], stride))
        pz[0]._opx_zeta_h_merge = cached
    base = cached[1].data_ptr()
    fz.packed, fz.packed_stride = base, stride
    fh.packed, fh.packed_stride = base + oct_, stride


_STATIC = ("dx", "dy", "h")


def _merge_static(s, fielddata) -> None:
    """Point the dx, dy and h descriptors of `s` at one buffer holding their QUAD
    copies side by side (64 bytes per node for FP32), when all three have QUAD
    copies of one dtype and shape: the three static stencils of a step are then
    one cache line instead of three.  The buffer is kept with dx's copy and made
    again when any of the three copies changes."""
    fs = [getattr(s, n) for n in _STATIC]
    packs = [getattr(fielddata[n], "packed", None) for n in _STATIC]
    if any(p is None or p[1] != _lib.OPX_PACK_QUAD for p in packs):
        return
    if len({(f.dtype, f.ndim, tuple(f.shape[:f.ndim])) for f in fs}) != 1:
        return
    tensors = [p[0] for p in packs]
    elem = 16 if fs[0].dtype == _lib.OPX_F32 else 8
    cached = getattr(tensors[0], "_opx_static_merge", None)
    if cached is None or cached[0] is not tensors[1] or cached[1] is not tensors[2]:
        # the cache holds dy's and h's copies, so their identity cannot be reused
        cached = (tensors[1], tensors[2], merge_quads(tensors, elem))
        tensors[0]._opx_static_merge = cached
    base = cached[2].data_ptr()
    for j, f in enumerate(fs):
        f.packed = base + j * elem
        f.packed_stride = 4 * elem


def fields_struct(fielddata: dict, names) -> _lib.opx_roms_fields:
    from ..fields import MERGE_STATIC_STENCILS

    s = _lib.opx_roms_fields()
    for name in names:
        setattr(s, name, field_struct(fielddata[name], name))
    if MERGE_STATIC_STENCILS and all(n in names for n in _STATIC):
        _merge_static(s, fielddata)
    if MERGE_STATIC_STENCILS and "zeta" in names and "h" in names:
        _merge_zeta_h(s, fielddata)  # h then reads from here, dx and dy from the static buffer
    return s


def stream_handle(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def byref(x):
    return ctypes.byref(x)
