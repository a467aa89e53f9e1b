"""Packed stencil copies of the drop-in's fields (opx_field.packed).

Gate: the drop-in gives the same bits with the packed copies as without
(they hold the same values; the interpolants blend them with the same
arithmetic), for AB3, RK2 and AB3 with an RK2 start-up, FP32 and BF16
storage, spatially ordered (TPAIR copies of the snapshot pairs) or not (HEX
copies of 3-D pairs, OCT of 2-D ones); and opx_pack_field_as's layouts equal
the values numpy picks.  The
parity tests against the oracle run with the packed copies on (the default).
"""

import ctypes

import numpy as np
import pytest
import torch

from offline_particles_amd import _lib, fields
from offline_particles_amd.workloads import build_fieldset, build_simulation, turbulent_workload

from ._helpers import gpu_state

pytestmark = pytest.mark.gpu


def _wl():
    # hourly snapshots, dt = 120 s: 60 steps cross two snapshot boundaries
    return turbulent_workload(NX=96, NY=80, NZ=12, nparticles=20_000, steps=80, nsnap_unique=4)


@pytest.mark.parametrize("order", [0, 9])
@pytest.mark.parametrize("storage", [None, torch.bfloat16])
@pytest.mark.parametrize("scheme", ["ab3", "rk2", "ab3_rk2"])
def test_packed_stencils_change_nothing(device, scheme, storage, order):
    wl = _wl()
    res = {}
    try:
        for pack in (False, True, "separate"):
            # "separate": packed copies, but dx, dy and h each in its own buffer
            fields.PACK_STENCILS = bool(pack)
            fields.MERGE_STATIC_STENCILS = pack is True
            sim = build_simulation(wl, scheme, device=device, storage_dtype=storage, spatial_order_every=order)
            for _ in range(60):
                sim.step()
            torch.cuda.synchronize()
            res[pack] = gpu_state(sim)
            if pack is True:
                # dx, dy and h read from one buffer of their QUADs side by side
                merged = getattr(sim.fieldset["dx"]._packed[1][0], "_opx_static_merge", None)
                assert merged is not None and merged[2].numel() == 4 * sim.fieldset["dx"]._packed[1][0].numel()
                # and h with zeta's OCT copy of the pair (caller's order only)
                zh = [getattr(v[0], "_opx_zeta_h_merge", None) for v in sim.fieldset["zeta"]._packed_pairs.values()]
                assert any(x is not None for x in zh) == (order == 0), zh
                # the kinds the steps packed: TPAIR when the particles are sorted
                kinds = {k[-1] for k in sim.fieldset["u"]._packed_pairs}
                assert kinds == {_lib.OPX_PACK_TPAIR if order else _lib.OPX_PACK_HEX}, kinds
                kinds = {k[-1] for k in sim.fieldset["zeta"]._packed_pairs}
                assert kinds == {_lib.OPX_PACK_TPAIR if order else _lib.OPX_PACK_OCT}, kinds
                fd = sim.fieldset["u"].get_field_data(0.25, None)
                assert fd.packed is not None and fd.packed[1] == _lib.OPX_PACK_HEX
                assert sim.fieldset["h"].get_field_data(0.25, None).packed[1] == _lib.OPX_PACK_QUAD
    finally:
        fields.PACK_STENCILS = True
        fields.MERGE_STATIC_STENCILS = True
    for k, a in res[False].items():
        assert np.array_equal(a, res[True][k], equal_nan=True), (scheme, storage, k)
        assert np.array_equal(a, res["separate"][k], equal_nan=True), (scheme, storage, k)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("shape", [(5, 7, 9), (2, 3, 2)])
def test_pack_hex_matches_numpy(device, dtype, shape):
    rng = np.random.default_rng(5)
    nz, ny, nx = shape
    p = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(device, dtype)
    q = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(device, dtype)
    code = _lib.storage_code(dtype)
    nbytes = ctypes.c_uint64()
    _lib.call("opx_packed_field_bytes", code, _lib.OPX_PACK_HEX, nz, ny, nx, ctypes.byref(nbytes))
    assert nbytes.value == nz * ny * nx * 16 * p.element_size()
    out = torch.zeros(nbytes.value, dtype=torch.uint8, device=device)
    _lib.call("opx_pack_field_as", p.data_ptr(), q.data_ptr(), code, _lib.OPX_PACK_HEX, nz, ny, nx, out.data_ptr(),
              torch.cuda.current_stream().cuda_stream)
    raw = {2: np.uint16, 4: np.uint32}[p.element_size()]
    view = torch.int16 if dtype == torch.bfloat16 else torch.int32
    P, Q = (t.cpu().view(view).numpy().view(raw) for t in (p, q))
    got = out.cpu().numpy().view(raw).reshape(nz, ny, nx, 16)
    # corner (dz, dy, dx) at (dz * 2 + dy) * 2 + dx, prev's then next's; the last
    # level / row / column repeats the node
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    z1, y1, x1 = np.minimum(zz + 1, nz - 1), np.minimum(yy + 1, ny - 1), np.minimum(xx + 1, nx - 1)
    corners = [(a, b, c) for a in (zz, z1) for b in (yy, y1) for c in (xx, x1)]
    want = np.stack([A[a, b, c] for A in (P, Q) for a, b, c in corners], axis=-1)
    assert np.array_equal(got, want)
    # a 2-D field has no HEX copy
    with pytest.raises(RuntimeError):
        _lib.call("opx_packed_field_bytes", code, _lib.OPX_PACK_HEX, 1, ny, nx, ctypes.byref(nbytes))


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("shape", [(5, 7, 9), (1, 6, 11)])
def test_pack_tpair_matches_numpy(device, dtype, shape):
    rng = np.random.default_rng(4)
    nz, ny, nx = shape
    p = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(device, dtype)
    q = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(device, dtype)
    code = _lib.storage_code(dtype)
    nbytes = ctypes.c_uint64()
    _lib.call("opx_packed_field_bytes", code, _lib.OPX_PACK_TPAIR, nz, ny, nx, ctypes.byref(nbytes))
    assert nbytes.value == nz * ny * nx * 2 * p.element_size()
    out = torch.zeros(nbytes.value, dtype=torch.uint8, device=device)
    _lib.call("opx_pack_field_as", p.data_ptr(), q.data_ptr(), code, _lib.OPX_PACK_TPAIR, nz, ny, nx, out.data_ptr(),
              torch.cuda.current_stream().cuda_stream)
    raw = {2: np.uint16, 4: np.uint32}[p.element_size()]
    view = torch.int16 if dtype == torch.bfloat16 else torch.int32
    P, Q = (t.cpu().view(view).numpy().view(raw) for t in (p, q))
    got = out.cpu().numpy().view(raw).reshape(nz, ny, nx, 2)
    assert np.array_equal(got, np.stack([P, Q], axis=-1))
    # a static field has no TPAIR copy
    with pytest.raises(RuntimeError):
        _lib.call("opx_pack_field_as", p.data_ptr(), None, code, _lib.OPX_PACK_TPAIR, nz, ny, nx, out.data_ptr(),
                  torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("shape", [(5, 7, 9), (1, 6, 11)])
@pytest.mark.parametrize("pair", [False, True])
def test_pack_field_matches_numpy(device, dtype, shape, pair):
    rng = np.random.default_rng(3)
    nz, ny, nx = shape
    p = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(device, dtype)
    q = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(device, dtype)
    kind = _lib.OPX_PACK_OCT if pair else _lib.OPX_PACK_QUAD
    code = _lib.storage_code(dtype)
    nbytes = ctypes.c_uint64()
    _lib.call("opx_packed_field_bytes", code, kind, nz, ny, nx, ctypes.byref(nbytes))
    per = (8 if pair else 4) * p.element_size()
    assert nbytes.value == nz * ny * nx * per
    out = torch.zeros(nbytes.value, dtype=torch.uint8, device=device)
    _lib.call("opx_pack_field", p.data_ptr(), q.data_ptr() if pair else None, code, nz, ny, nx, out.data_ptr(),
              torch.cuda.current_stream().cuda_stream)
    raw = {2: np.uint16, 4: np.uint32}[p.element_size()]
    P = p.cpu().view(torch.int16 if dtype == torch.bfloat16 else torch.int32).numpy().view(raw)
    Q = q.cpu().view(torch.int16 if dtype == torch.bfloat16 else torch.int32).numpy().view(raw)
    got = out.cpu().numpy().view(raw).reshape(nz, ny, nx, -1)
    # corners (y, x), (y, x+1), (y+1, x), (y+1, x+1); the last row / column repeats the node
    yy, xx = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    y1, x1 = np.minimum(yy + 1, ny - 1), np.minimum(xx + 1, nx - 1)
    corners = [(yy, xx), (yy, x1), (y1, xx), (y1, x1)]
    # (two BF16 values per 32-bit word, the first in the low half: the same order as uint16)
    want = np.stack([A[:, a, b] for A in ((P, Q) if pair else (P,)) for a, b in corners], axis=-1)
    assert np.array_equal(got, want)
