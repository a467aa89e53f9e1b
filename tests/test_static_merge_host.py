"""The shared buffer of the static fields' QUAD copies (dx, dy, h side by side,
read through opx_field.packed_stride): layout and descriptor wiring, on host
tensors (no kernel runs)."""

import types

import numpy as np
import torch

from offline_particles_amd import _lib
from offline_particles_amd.kernels import _marshal


def _quad(n, elem, seed):
    return torch.from_numpy(np.random.default_rng(seed).integers(0, 256, n * elem, dtype=np.uint8))


def test_merge_quads_layout():
    n, elem = 37, 16
    qs = [_quad(n, elem, s) for s in range(3)]
    out = _marshal.merge_quads(qs, elem)
    assert out.shape == (n, 4, elem) and out.dtype == torch.uint8
    flat = out.reshape(-1).numpy()
    for j, q in enumerate(qs):
        for node in (0, 1, n // 2, n - 1):
            # field j's QUAD of `node` sits at node * 4 elem + j elem
            got = flat[node * 4 * elem + j * elem: node * 4 * elem + (j + 1) * elem]
            assert np.array_equal(got, q.numpy()[node * elem:(node + 1) * elem])
    assert not flat.reshape(n, 4, elem)[:, 3].any()  # the fourth slot is zero


def _fielddata(n, elem, kind=_lib.OPX_PACK_QUAD, seeds=(0, 1, 2)):
    return {name: types.SimpleNamespace(packed=(_quad(n, elem, s), kind)) for name, s in zip(("dx", "dy", "h"), seeds)}


def _struct(shape=(6, 8), dtype=_lib.OPX_F32):
    s = _lib.opx_roms_fields()
    for name in ("dx", "dy", "h"):
        f = getattr(s, name)
        f.dtype, f.ndim = dtype, len(shape)
        for d, v in enumerate(shape):
            f.shape[d] = v
    return s


def test_merge_static_points_descriptors_into_one_buffer():
    n, elem = 48, 16
    fd = _fielddata(n, elem)
    s = _struct()
    _marshal._merge_static(s, fd)
    merged = fd["dx"].packed[0]._opx_static_merge[2]
    for j, name in enumerate(("dx", "dy", "h")):
        f = getattr(s, name)
        assert f.packed == merged.data_ptr() + j * elem and f.packed_stride == 4 * elem
    # the same copies: the buffer is reused; a new copy of h: made again
    s2 = _struct()
    _marshal._merge_static(s2, fd)
    assert s2.dx.packed == s.dx.packed
    fd["h"].packed = (_quad(n, elem, 9), _lib.OPX_PACK_QUAD)
    s3 = _struct()
    _marshal._merge_static(s3, fd)
    new = fd["dx"].packed[0]._opx_static_merge[2]
    assert new is not merged
    assert np.array_equal(new[:, 2].reshape(-1).numpy(), fd["h"].packed[0].numpy())


def test_merge_static_leaves_mismatched_fields_alone():
    n, elem = 48, 16
    # a different shape for h
    s = _struct()
    s.h.shape[1] = 9
    fd = _fielddata(n, elem)
    _marshal._merge_static(s, fd)
    assert all(getattr(s, k).packed_stride == 0 for k in ("dx", "dy", "h"))
    # one field without a QUAD copy
    s = _struct()
    fd = _fielddata(n, elem)
    fd["dy"].packed = None
    _marshal._merge_static(s, fd)
    assert all(getattr(s, k).packed_stride == 0 for k in ("dx", "dy", "h"))


def test_merge_zeta_h_layout_and_descriptors():
    """zeta's OCT element and h's QUAD side by side per node (64 bytes for FP32,
    32 for BF16 zeta with FP32 h), both descriptors pointing into the buffer."""
    n = 30
    for zdt, hdt, oct_, quad, stride in ((_lib.OPX_F32, _lib.OPX_F32, 32, 16, 64),
                                         (_lib.OPX_BF16, _lib.OPX_F32, 16, 16, 32),
                                         (_lib.OPX_BF16, _lib.OPX_BF16, 16, 8, 32)):
        s = _lib.opx_roms_fields()
        for name, dt in (("zeta", zdt), ("h", hdt)):
            f = getattr(s, name)
            f.dtype, f.ndim = dt, 2
            f.shape[0], f.shape[1] = 5, 6
        z, h = _quad(n, oct_, 3), _quad(n, quad, 4)
        fd = {"zeta": types.SimpleNamespace(packed=(z, _lib.OPX_PACK_OCT)),
              "h": types.SimpleNamespace(packed=(h, _lib.OPX_PACK_QUAD))}
        _marshal._merge_zeta_h(s, fd)
        buf = z._opx_zeta_h_merge[1]
        assert buf.shape == (n, stride)
        assert s.zeta.packed == buf.data_ptr() and s.zeta.packed_stride == stride
        assert s.h.packed == buf.data_ptr() + oct_ and s.h.packed_stride == stride
        b = buf.numpy()
        assert np.array_equal(b[:, :oct_].reshape(-1), z.numpy())
        assert np.array_equal(b[:, oct_:oct_ + quad].reshape(-1), h.numpy())
        assert not b[:, oct_ + quad:].any()


def test_merge_zeta_h_skips_tpair_and_other_shapes():
    n = 30
    s = _lib.opx_roms_fields()
    for name in ("zeta", "h"):
        f = getattr(s, name)
        f.dtype, f.ndim = _lib.OPX_F32, 2
        f.shape[0], f.shape[1] = 5, 6
    fd = {"zeta": types.SimpleNamespace(packed=(_quad(n, 8, 1), _lib.OPX_PACK_TPAIR)),
          "h": types.SimpleNamespace(packed=(_quad(n, 16, 2), _lib.OPX_PACK_QUAD))}
    _marshal._merge_zeta_h(s, fd)  # sorted particles: the TPAIR copy stays alone
    assert s.zeta.packed_stride == 0 and s.h.packed_stride == 0
    s.h.shape[1] = 7
    fd["zeta"].packed = (_quad(n, 32, 1), _lib.OPX_PACK_OCT)
    _marshal._merge_zeta_h(s, fd)
    assert s.zeta.packed_stride == 0 and s.h.packed_stride == 0
