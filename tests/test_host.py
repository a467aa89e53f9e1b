"""Host-side checks that need no GPU.

- libopx.so loads and exports every function include/opx.h declares;
- the ctypes structs have the C layout of include/opx.h (compiled with gcc);
- the reference-API mirror: kernel declarations, merge rules, time index and
  blend-weight semantics, Stagger, Fieldset validation, argument errors.
"""

import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

from offline_particles_amd import _build, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        _build.build()
    return _lib.load()


def test_library_exports_every_declared_symbol(lib):
    declared = _lib.declared_symbols()
    assert len(declared) >= 16
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    for s in declared:
        assert hasattr(lib, s)
    assert lib.opx_abi_version() == 1


_STRUCTS = {
    "opx_field": _lib.opx_field,
    "opx_roms_fields": _lib.opx_roms_fields,
    "opx_scalars": _lib.opx_scalars,
    "opx_particles": _lib.opx_particles,
    "opx_split_state": _lib.opx_split_state,
    "opx_grid": _lib.opx_grid,
    "opx_window": _lib.opx_window,
    "opx_release_clock": _lib.opx_release_clock,
    "opx_release_step": _lib.opx_release_step,
    "opx_sample": _lib.opx_sample,
}


def test_ctypes_layout_matches_header():
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "opx.h"', "int main(void) {"]
    for name, cls in _STRUCTS.items():
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{name}.{fname} %zu\\n", offsetof({name}, {fname}));')
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "abi.c")
        exe = os.path.join(d, "abi")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    c = dict(line.rsplit(" ", 1) for line in out.splitlines())
    for name, cls in _STRUCTS.items():
        assert int(c[name]) == ctypes.sizeof(cls), name
        for fname, _ in cls._fields_:
            assert int(c[f"{name}.{fname}"]) == getattr(cls, fname).offset, f"{name}.{fname}"


def test_kernel_declarations_match_reference():
    from offline_particles_amd.kernels.roms import (
        ab3_post_w_advection_kernel,
        ab3_w_advection_kernel,
        rk2_w_advection_step_1_kernel,
        rk2_w_advection_step_2_kernel,
        rk2_w_advection_update_kernel,
    )
    from offline_particles_amd.kernels.validation import validation_kernel

    f8, u1, i4 = np.dtype(np.float64), np.dtype(np.uint8), np.dtype(np.int32)
    hist = ["_dz0", "_dz1", "_dz2", "_dyidx0", "_dyidx1", "_dyidx2", "_dxidx0", "_dxidx1", "_dxidx2"]
    assert dict(ab3_w_advection_kernel.particle_fields) == {
        "status": u1, "zidx": f8, "yidx": f8, "xidx": f8, "z": f8, **{k: f8 for k in hist}}
    assert dict(ab3_w_advection_kernel.scalars) == {"_dt": f8, "hc": f8, "NZ": i4}
    assert ab3_w_advection_kernel.simulation_fields == frozenset("u v w dx dy h C zeta".split())
    assert ab3_post_w_advection_kernel.simulation_fields == frozenset({"h", "C", "zeta"})
    assert dict(rk2_w_advection_step_2_kernel.scalars) == {"_dt": f8, "_RK2_alpha": f8, "hc": f8, "NZ": i4}
    assert set(rk2_w_advection_step_1_kernel.particle_fields) == {
        "status", "zidx", "yidx", "xidx", "z", "_dxidx1", "_dyidx1", "_dz1"}
    assert rk2_w_advection_update_kernel.simulation_fields == frozenset({"h", "C", "zeta"})
    assert set(validation_kernel.scalars) == {"zidx_min", "zidx_max", "yidx_min", "yidx_max", "xidx_min",
                                              "xidx_max"}
    assert len(validation_kernel.functions) == 2


def test_merge_rules():
    from offline_particles_amd.kernels import ParticleKernel, merge_particle_fields, merge_scalars

    k1 = ParticleKernel(lambda *a: None, {"a": np.float64}, {"s": np.float64}, ["u"])
    k2 = ParticleKernel(lambda *a: None, {"a": np.float32}, {"s": np.int32}, ["v"])
    with pytest.raises(TypeError):
        merge_particle_fields([k1, k2])
    with pytest.raises(TypeError):
        merge_scalars([k1, k2])
    merged = merge_particle_fields([k1])
    assert list(merged)[:4] == ["status", "zidx", "yidx", "xidx"]
    chained = ParticleKernel.chain(k1, ParticleKernel(lambda *a: None, {}, {}, ["w"]))
    assert chained.simulation_fields == frozenset({"u", "w"})
    with pytest.raises(TypeError):
        ParticleKernel([1], {}, {}, [])


def test_time_index_and_blend_weights():
    from offline_particles_amd import RK2Timestepper
    from offline_particles_amd.fields import blend_weights
    from offline_particles_amd.kernels import ParticleKernel

    k = ParticleKernel(lambda *a: None, {}, {}, [])
    ta = np.array([0.0, 3600.0, 7200.0, 10800.0])
    ts = RK2Timestepper(ta, np.float64(300.0), k, k, k)
    assert ts.tidx == 0.0
    for _ in range(13):
        ts.advance_time()
    assert ts.time == 3900.0 and ts.iteration == 13
    assert ts.tidx == 1 + (3900.0 - 3600.0) / 3600.0
    with pytest.raises(ValueError):
        ts.get_time_index(10800.5)
    with pytest.raises(IndexError):
        ts.get_time_index(10800.0)
    with pytest.raises(ValueError):
        RK2Timestepper(ta, 300.0, k, k, k)  # a Python float dt needs time_unit
    with pytest.raises(ValueError):
        RK2Timestepper(ta[::-1], np.float64(1.0), k, k, k)
    import torch

    It, ft, gt = blend_weights(np.float64(2.3), torch.float32)
    assert It == 2 and ft == float(np.float32(np.float64(2.3) - 2))
    assert gt == float(np.float32(1) - np.float32(np.float64(2.3) - 2))
    It, ft, gt = blend_weights(np.float64(0.75), torch.float64)
    assert (It, ft, gt) == (0, 0.75, 0.25)


def test_stagger_and_fieldset_validation():
    import torch

    from offline_particles_amd import Fieldset, StaticField, Stagger, TimeDependentField

    assert Stagger("inner").offset == -0.5 and Stagger("outer").expected_size(10) == 11
    assert Stagger("invariant").offset is None and Stagger.LEFT.offset == 0.5
    h = StaticField.from_numpy(np.ones((4, 5)), "invariant", "center", "center", device="cpu")
    fs = Fieldset(3, 2, 4, 5, constants={"hc": 10.0}, h=h)
    assert fs.constants["zidx_max"] == 1.0 and fs.constants["hc"] == 10.0
    with pytest.raises(ValueError):
        fs.add_field("bad", StaticField.from_numpy(np.ones((4, 6)), "invariant", "center", "center", device="cpu"))
    with pytest.raises(KeyError):
        fs.add_field("h", h)
    u = TimeDependentField(np.zeros((3, 2, 4, 4)), "center", "center", "inner", device="cpu")
    fs.add_field("u", u)
    fd = u.get_field_data(1.25, None)
    assert fd.offsets == (0.0, 0.0, -0.5)
    assert fd.array.ft == 0.25 and fd.array.gt == 0.75 and fd.array.It == 1
    with pytest.raises(IndexError):
        u.get_field_data(2.5, None)  # window [2, 3] needs snapshot 3 of 3
    cyc = TimeDependentField(np.zeros((2, 2, 4, 4)), "center", "center", "inner", cyclic=True, device="cpu")
    assert cyc.get_field_data(7.5, None).array.It == 7
    assert torch.equal(cyc.snapshot(7), cyc.snapshot(1))


def test_kernel_rejects_host_tensors():
    from offline_particles_amd.kernels.validation import finite_indices
    from offline_particles_amd.particles import Particles

    p = Particles(4, device="cpu", status=np.uint8, zidx=np.float64, yidx=np.float64, xidx=np.float64)
    with pytest.raises(ValueError):
        finite_indices(p, {}, {})


def test_iso_fused_entry_point_rejects_bad_arguments(lib):
    """opx_ab3_iso_fused_steps checks its window and arguments on the host and
    returns an error before any launch: a w window (NZ + 1 levels), BF16
    storage, a missing rho_t, rho_scale <= 0.  (Every case fails a check, so
    nothing reaches the device.)"""
    NZ, NY, NX = 10, 8, 9
    g = _lib.opx_grid()
    g.NZ, g.NY, g.NX = NZ, NY, NX
    g.xbits, g.ybits, g.zbits = 4, 4, 4
    g.hc = 10.0
    buf = ctypes.create_string_buffer(64)
    p = ctypes.addressof(buf)
    s = _lib.opx_split_state()
    s.n = 1
    for name in ("status", "cell", "fz", "fy", "fx", "hz1", "hy1", "hx1", "hz2", "hy2", "hx2", "z", "rho_t"):
        setattr(s, name, p)
    w = _lib.opx_window()
    w.u = w.v = w.w = w.zeta = w.hdd = w.C = p
    w.u_stride = w.v_stride = w.w_stride = w.zeta_stride = (NZ + 1) * NY * NX
    w.dtype3, w.nslots = _lib.OPX_F32, 2
    w.wlevels = NZ + 1

    def rc(rho_scale=0.2):
        return lib.opx_ab3_iso_fused_steps(ctypes.byref(s), ctypes.byref(g), ctypes.byref(w), p, 0, 1, 120.0, 5e-4,
                                           rho_scale, None, 128, None, None)

    assert rc() == _lib.OPX_ERR_INVALID and b"levels" in lib.opx_last_error()
    w.wlevels = NZ
    w.dtype3 = _lib.OPX_BF16
    assert rc() == _lib.OPX_ERR_INVALID and b"FP32" in lib.opx_last_error()
    w.dtype3 = _lib.OPX_F32
    s.rho_t = None
    assert rc() == _lib.OPX_ERR_INVALID and b"rho_t" in lib.opx_last_error()
    s.rho_t = p
    assert rc(0.0) == _lib.OPX_ERR_INVALID and b"rho_scale" in lib.opx_last_error()
    # the w-advection step refuses a density window
    assert lib.opx_ab3_fused_steps(ctypes.byref(s), ctypes.byref(g), ctypes.byref(w), p, 0, 1, 120.0, None, 128,
                                   None, None) == _lib.OPX_ERR_INVALID
    assert b"levels" in lib.opx_last_error()


def test_fused_engine_rejects_bounds_beyond_split_range(lib):
    """The split state stores cells -2 .. N+1 per axis and clamps positions beyond
    them, so bounds outside (-2, N+1] would let a clamped particle pass the
    domain check: the engine raises ValueError and the C-ABI refuses the grid."""
    from offline_particles_amd import Fieldset, StaticField, TimeDependentField
    from offline_particles_amd.engine import AB3Engine

    NZ, NY, NX = 4, 5, 6
    T = 3
    td = {"u": ((NZ, NY, NX - 1), ("center", "center", "inner")),
          "v": ((NZ, NY - 1, NX), ("center", "inner", "center")),
          "w": ((NZ + 1, NY, NX), ("outer", "center", "center")),
          "zeta": ((NY, NX), ("invariant", "center", "center"))}
    st = {"h": ((NY, NX), ("invariant", "center", "center")), "dx": ((NY, NX), ("invariant", "center", "center")),
          "dy": ((NY, NX), ("invariant", "center", "center")), "C": ((NZ + 1,), ("outer", "invariant", "invariant"))}

    def fieldset(**bounds):
        f = {k: TimeDependentField(np.zeros((T,) + s, np.float32), *g, device="cpu") for k, (s, g) in td.items()}
        f.update({k: StaticField.from_numpy(np.ones(s, np.float32), *g, device="cpu") for k, (s, g) in st.items()})
        return Fieldset(T, NZ, NY, NX, constants={"hc": 10.0}, **bounds, **f)

    ta = np.arange(T, dtype=np.float64)
    for bounds in ({"xidx_bounds": (-5, NX + 4)}, {"xidx_bounds": (0, NX + 2)}, {"yidx_bounds": (-2, NY - 1)},
                   {"zidx_bounds": (0, NZ + 1.5)}):
        with pytest.raises(ValueError, match="bounds"):
            AB3Engine(fieldset(**bounds), ta, 0.1, device="cpu")

    # the C-ABI checks the same range before any launch
    g = _lib.opx_grid()
    g.NZ, g.NY, g.NX = NZ, NY, NX
    g.xbits, g.ybits, g.zbits = 4, 4, 4
    g.hc = 10.0
    g.zidx_max, g.yidx_max, g.xidx_max = NZ - 1, NY - 1, NX + 4
    g.xidx_min = -5
    s = _lib.opx_split_state()
    assert lib.opx_split_from_f64(ctypes.byref(s), ctypes.byref(g), None, None, None, None) == _lib.OPX_ERR_INVALID
    assert b"bounds" in lib.opx_last_error()
    g.xidx_min, g.xidx_max = -1.5, NX + 1
    assert lib.opx_split_from_f64(ctypes.byref(s), ctypes.byref(g), None, None, None, None) == _lib.OPX_OK


def test_rk2_startup_timestepper_declarations():
    """AB3 with an RK2 start-up declares the union of the AB3, RK2 and history
    kernels' particle fields (the RK2 and AB3 derivative names coincide, so the
    merge is conflict-free), the RK2 parameter as a scalar, and only the
    reference's two start-up names are accepted."""
    from offline_particles_amd.kernels import merge_particle_fields, merge_scalars
    from offline_particles_amd.models import roms
    from offline_particles_amd.timesteppers import ABRK2StartupTimestepper, ABTimestepper

    ta = np.arange(4, dtype=np.float64) * 3600.0
    ts = roms.ab3_w_advection_timestepper(ta, np.float64(120.0), startup="rk2")
    assert isinstance(ts, ABRK2StartupTimestepper) and ts.alpha == 2 / 3
    fields = merge_particle_fields(list(ts.kernels))
    for name in ("status", "zidx", "yidx", "xidx", "z", "_dz0", "_dz1", "_dz2", "_dyidx0", "_dyidx1", "_dyidx2",
                 "_dxidx0", "_dxidx1", "_dxidx2"):
        assert name in fields, name
    assert "_RK2_alpha" in merge_scalars(list(ts.kernels))
    assert type(roms.ab3_w_advection_timestepper(ta, np.float64(120.0))) is ABTimestepper
    with pytest.raises(ValueError):
        roms.ab3_w_advection_timestepper(ta, np.float64(120.0), startup="euler")


def test_packed_window_limited_to_4GiB(lib):
    """The step kernels address a packed stencil window with 32-bit byte
    offsets, so opx_packed_bytes refuses a layout whose window would exceed
    4 GiB (OCT with FP32 nodes on a 2048 x 1536 x 60 grid: 32 B x 61 levels x
    2048 x 1536 = 6.1 GB) and sizes the others (PAIR FP32: 8 B per node)."""
    g = _lib.opx_grid()
    g.NZ, g.NY, g.NX = 60, 1536, 2048
    g.xbits, g.ybits, g.zbits = 12, 11, 7
    g.hc = 10.0
    g.zidx_max, g.yidx_max, g.xidx_max = g.NZ - 1, g.NY - 1, g.NX - 1
    out = [ctypes.c_uint64() for _ in range(4)]
    args = [ctypes.byref(o) for o in out]
    assert lib.opx_packed_bytes(ctypes.byref(g), _lib.OPX_F32, _lib.OPX_LAYOUT_OCT, *args) == _lib.OPX_ERR_INVALID
    assert b"4 GiB" in lib.opx_last_error()
    assert lib.opx_packed_bytes(ctypes.byref(g), _lib.OPX_BF16, _lib.OPX_LAYOUT_OCT, *args) == _lib.OPX_OK
    assert out[2].value == 16 * 61 * 1536 * 2048 <= 1 << 32
    assert lib.opx_packed_bytes(ctypes.byref(g), _lib.OPX_F32, _lib.OPX_LAYOUT_PAIR, *args) == _lib.OPX_OK
    assert out[2].value == 8 * 61 * 1536 * 2048


# The reference's public names by module (its __init__.py files: __init__.py:1-25,
# events/__init__.py:1-11, kernels/__init__.py, kernels/output/__init__.py,
# kernels/roms/__init__.py, models/roms/__init__.py:8-20, output/__init__.py,
# fields.py, spatial_arrays.py, timesteppers.py): a script's imports resolve after
# the package name is switched.
REFERENCE_PUBLIC = {
    "": ["TimeDependentField", "StaticField", "Fieldset", "Simulation", "SimulationBuilder", "Event",
         "SimulationState", "RK2Timestepper", "Timestepper", "kernels", "output", "roms"],
    "events": ["Event", "SimulationState", "IterationScheduler", "TimeScheduler"],
    "kernels": ["ParticleKernel", "KernelFunction", "ParticleStatus", "merge_particle_fields", "merge_scalars",
                "merge_simulation_fields"],
    "kernels.output": ["linear_interpolation_kernel", "bilinear_interpolation_kernel",
                       "trilinear_interpolation_kernel", "vertical_interpolation_kernel",
                       "horizontal_interpolation_kernel"],
    "kernels.roms": ["compute_z_kernel", "ab3_w_advection_kernel", "ab3_post_w_advection_kernel",
                     "ab3_isopycnal_following_kernel", "ab3_post_isopycnal_following_kernel",
                     "rk2_w_advection_step_1_kernel", "rk2_w_advection_step_2_kernel",
                     "rk2_w_advection_update_kernel"],
    "models.roms": ["compute_z_kernel", "ab3_w_advection_kernel", "ab3_post_w_advection_kernel",
                    "ab3_w_advection_timestepper", "rk2_w_advection_step_1_kernel", "rk2_w_advection_step_2_kernel",
                    "rk2_w_advection_update_kernel", "rk2_w_advection_timestepper"],
    "output": ["AbstractOutputWriter", "AbstractOutputWriterBuilder", "Output", "ZarrOutputWriter",
               "ZarrOutputBuilder", "linearly_interpolate_fields"],
    "fields": ["Field", "StaticField", "TimeDependentField", "FieldData"],
    "spatial_arrays": ["BBox", "Stagger", "SpatialArray", "NumpyArray", "ChunkedDaskArray",
                       "compute_new_lower_bound", "compute_new_upper_bound"],
    "timesteppers": ["Timestepper", "RK2Timestepper", "ABTimestepper"],
    "launcher": ["Launcher", "ScalarSource"],
    "particles": ["Particles", "ParticlesView"],
}


@pytest.mark.parametrize("module", sorted(REFERENCE_PUBLIC))
def test_reference_public_names_resolve(module):
    import importlib

    mod = importlib.import_module("offline_particles_amd" + ("." + module if module else ""))
    missing = [n for n in REFERENCE_PUBLIC[module] if not hasattr(mod, n)]
    assert not missing, (module, missing)


@pytest.mark.parametrize("scheme", ["ab3", "rk2"])
@pytest.mark.parametrize("k", [1, 7, 128])
@pytest.mark.parametrize("uneven", [False, True])
def test_host_clock_run_equals_stepwise(scheme, k, uneven):
    """The engine's host clock mirror over a run of k steps at once
    (_advance_host_clocks_run) against k calls of the per-step mirror: the
    same snapshot pairs, times (bit for bit) and steps left."""
    import types

    from offline_particles_amd.engine import AB3Engine

    rng = np.random.default_rng(k + 10 * uneven + (scheme == "rk2"))
    n = 200
    ta = (np.concatenate([[0.0], np.cumsum(rng.uniform(0.4, 2.5, n - 1) * 3600.0)]) if uneven
          else np.arange(n, dtype=np.float64) * 3600.0)

    def engine():
        R = 300
        e = types.SimpleNamespace(time_array=ta, dt=np.float64(120.0), alpha=2.0 / 3.0, scheme=scheme,
                                  global_step=1000)
        e._t = rng.uniform(ta[0], ta[-1], R)
        e._t[:30] = ta[-1] - rng.uniform(0, 200 * 120.0, 30)  # leave the array within the run
        e._t[30:40] = ta[rng.integers(0, n, 10)]
        e._start = 1000 + rng.integers(-5, k + 5, R)
        e._left = rng.integers(-1, 2 * k + 2, R)
        e._time_index = lambda t: AB3Engine._time_index(e, t)
        return e

    a = engine()
    b = types.SimpleNamespace(**{kk: (v.copy() if isinstance(v, np.ndarray) else v) for kk, v in vars(a).items()})
    b._time_index = lambda t: AB3Engine._time_index(b, t)
    want = set()
    for _ in range(k):
        want |= AB3Engine._advance_host_clocks(a)
        a.global_step += 1
    got = AB3Engine._advance_host_clocks_run(b, k)
    assert got == want
    assert np.array_equal(a._t, b._t) and np.array_equal(a._left, b._left)
