#!/usr/bin/env python3
"""Throughput benchmark of the advection hot path on one MI355X.

Workload (default): 16M particles, AB3 (Euler -> AB2 -> AB3 start-up), dt =
120 s, on a synthetic 1024 x 768 x 50 ROMS grid with 8 hourly snapshots cycled
(turbulent k^-3 flow, 2% land).  A "step" is one AB3 step of every particle:
validation, the AB3 stage and the post-step zidx update.

Forms measured in the same run:
- plain_fp32: the drop-in behind the reference API (Simulation.step ->
  validation, ab3_w_advection, ab3_post_w_advection HIP kernels), FP32 fields,
  FP64 positions and FP64 interpolation, particles in the caller's order.
- plain_fp32_ordered: the same with Simulation(spatial_order_every=20): the
  particle arrays permuted into the Morton order of their cells every 20 steps
  (inside the timed loop); the user sees the same bits in the same order.
- fused_fp32 / fused_bf16: AB3Engine, one fused HIP kernel per step over
  split positions (packed int32 cell + 32-bit fixed-point offsets) with u, v,
  w stored in FP32 or BF16; bin shape, workgroup size and re-sort cadence are
  picked by the engine's autotuner during warm-up, and the periodic re-binning
  is inside the timed region.
- batching (``--batched NREL``, default 256): uneven releases (sizes log-uniform
  1k-1M, start over 24 h, 6-72 h long, 10 km boxes) on the same grid with FP32
  and with BF16 storage (``--mix-storage``), stepped together from captured HIP
  graphs with the autotuned bin shape, layout and re-sort cadence; one release
  of each mix (the one nearest 1000 steps, 5000 of its particles) is checked
  against the oracle from its own start time; reported under ``batching``.
- serving (``--serve SECONDS``, default 5): Poisson arrivals at 20/s for that
  many seconds, sizes 1k-2M, 1-10 simulated days, on a 2048 x 1536 x 60 grid
  (FP32 and BF16 storage, the same requests) streamed through a 3-snapshot
  window of which each slot receives only the 16 x 16-column tiles the
  particles can reach (``--serve-tiles``), admission by free HBM; every
  request runs to completion; p50/p99 latency under ``serving``.
- isopycnal (``--extras iso``): the isopycnal-following AB3 on the same
  workload with a stratified, drifting density field, as the drop-in
  (validation, ab3_isopycnal_following, post step) and as the fused engine
  form (AB3Engine(scheme="ab3_iso"), FP32 storage, tuned like the fused
  forms); reported under ``isopycnal``.
- output events (``--extras output``): compute_z plus the trilinear (u, v, w)
  and bilinear (zeta, h) output kernels of the drop-in API over every particle
  at a time index between snapshots, and the same six outputs from the fused
  engine's state (AB3Engine.sample, one kernel over the re-binned particles);
  reported under ``output_events``.
- io (``--extras io``, not in the default run): the workload's snapshots staged
  to a Zarr v3 store on local disk, the fused engine streaming them from that
  store beside streaming from pinned host memory, and the Zarr trajectory
  writer; reported under ``io``.
The headline ``value`` is the fastest 16M-particle AB3 form whose positions
after 1000 steps stay within 1e-3 grid-index units of the FP64 oracle (the
accuracy run; without one, FP32-storage forms only); every form is listed in
``forms`` with its deviation.

Prints ONE JSON line.  ``value`` is particle-steps/s of the headline form with
all inputs resident in HBM; particle-steps count the particles active during
the timed steps (average of the active counts before and after).  The fused
forms' timed window is the requested steps rounded up to whole re-sort cycles
of whole launches (``config.timed_steps``), starting with a re-bin.
``roofline`` is computed for the dominant kernel from algorithmic bytes
(compulsory particle state + distinct field nodes touched) over its average
duration measured with HIP events on its stream; ``roofline_valu`` gives the
several-step kernels' VALU issue rate (the instruction count per wave-step
from the committed counter summary, over the live duration), which is what
bounds them.  The CPU leg times the
reference-structured FP64 CPU oracle on the host cores over a bounded sample
(``cpu_baseline``) and runs it as the checker of the GPU forms' positions
after 1000 steps on a 100k-particle subsample (``accuracy``).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
import types

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# VALU issue ceiling: 256 CUs x 4 SIMDs at the 2.4 GHz peak engine clock, one
# wave64 VALU instruction per SIMD every 4 cycles (16 lanes per SIMD; the
# counter SQ_ACTIVE_INST_VALU counts these 4-cycle slots)
VALU_PEAK_INSTS = 256 * 4 * 2.4e9 / 4
# accuracy bar of the headline: max position deviation from the FP64 oracle
# after the accuracy run's steps (1000), in grid-index units
ACCURACY_BAR = 1e-3


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# algorithmic bytes
# ---------------------------------------------------------------------------

def touched_nodes(shape, offsets, pos, corners_3d=True):
    """Distinct stencil nodes of a field touched by the given positions (device tensors)."""
    import torch

    dev = pos[0].device
    if len(shape) == 3:
        n0, n1, n2 = shape
        idx = [torch.clamp(torch.trunc(p + o).to(torch.int64), 0, n - 2) for p, o, n in zip(pos, offsets, shape)]
        base = (idx[0] * n1 + idx[1]) * n2 + idx[2]
        deltas = [0, 1, n2, n2 + 1, n1 * n2, n1 * n2 + 1, n1 * n2 + n2, n1 * n2 + n2 + 1]
        total = n0 * n1 * n2
    else:
        n0, n1 = shape
        idx = [torch.clamp(torch.trunc(p + o).to(torch.int64), 0, n - 2) for p, o, n in zip(pos[1:], offsets, shape)]
        base = idx[0] * n1 + idx[1]
        deltas = [0, 1, n1, n1 + 1]
        total = n0 * n1
    mark = torch.zeros(total, dtype=torch.uint8, device=dev)
    for d in deltas:
        mark[base + d] = 1
    return int(mark.sum().item())


def field_bytes_touched(wl, sim, storage_bytes_3d, fields_3d=("u", "v", "w")):
    """Bytes of distinct field values one AB3 launch must read (2 snapshots of
    each 3-D field and of zeta)."""
    from offline_particles_amd.workloads import OFFSETS

    P = sim.particles
    act = P["status"] < 128
    pos = [P[k][act] for k in ("zidx", "yidx", "xidx")]
    b2 = wl.fields["h"].dtype.itemsize
    total = 0
    for name in fields_3d:
        total += 2 * storage_bytes_3d * touched_nodes(wl.fields[name].shape[1:], OFFSETS[name], pos)
    for name, nsnap in (("h", 1), ("dx", 1), ("dy", 1), ("zeta", 2)):
        shp = wl.fields[name].shape[-2:]
        total += nsnap * b2 * touched_nodes(shp, OFFSETS[name], pos)
    total += wl.fields["C"].nbytes
    return total


# plain drop-in AB3 kernel (ab3_w_advection): per active particle it reads
# status, 3 positions, 6 history values (1 + 24 + 48 B) and writes z,
# y, x, the 3 current derivatives and 6 history values (8 + 16 + 24 + 48 B).
PLAIN_AB3_STATE_BYTES = 73 + 96
# plain isopycnal-following AB3 kernel (ab3_isopycnal_following): the same
# plus the particle's target density rho_t (8 B read)
PLAIN_ISO_STATE_BYTES = PLAIN_AB3_STATE_BYTES + 8
# output kernels: status and the indices read (1 + 8 per index), one FP64 value written
OUTPUT_TRILINEAR_BYTES = 1 + 24 + 8
OUTPUT_BILINEAR_BYTES = 1 + 16 + 8

# isopycnal-following constants and targets of the isopycnal form: the density
# of add_density spans 1026-1028 kg/m^3 over the column
ISO_CONSTANTS = {"w_max": 5e-4, "rho_scale": 0.2}


def iso_targets(n):
    return np.random.default_rng(2).uniform(1026.2, 1027.8, n)


def with_density(wl):
    """A shallow copy of the workload with the time-dependent density field "rho" added."""
    import copy

    from offline_particles_amd.workloads import add_density

    w2 = copy.copy(wl)
    w2.fields = dict(wl.fields)
    return add_density(w2)


# ---------------------------------------------------------------------------
# forms
# ---------------------------------------------------------------------------

class _TimedKernel:
    """Wraps a ParticleKernel's functions with HIP events on the current stream."""

    def __init__(self, kernel):
        import torch

        from offline_particles_amd.kernels import ParticleKernel

        self.events = []
        self.active = False
        inner = kernel

        def timed(particles, scalars, fielddata):
            if self.active:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                inner(particles, scalars, fielddata)
                e1.record()
                self.events.append((e0, e1))
            else:
                inner(particles, scalars, fielddata)

        timed.__qualname__ = "timed_" + "_".join(ParticleKernel.func_name(f) for f in kernel.functions)
        self.kernel = ParticleKernel(timed, kernel.particle_fields, kernel.scalars, kernel.simulation_fields)

    def mean_ms(self):
        import torch

        torch.cuda.synchronize()
        if not self.events:
            return float("nan")
        return float(np.mean([a.elapsed_time(b) for a, b in self.events]))


def active_count(sim):
    return int((sim.particles["status"] < 128).sum().item())


def _time_plain(sim, tk, steps, warmup):
    """Wall time of ``steps`` Simulation steps after ``warmup``, and the timed
    kernel's mean duration; particle-steps count the active particles."""
    import torch

    for _ in range(warmup):
        sim.step()
    torch.cuda.synchronize()
    a0 = active_count(sim)
    tk.active = True
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    a1 = active_count(sim)
    tk.active = False
    return t1 - t0, 0.5 * (a0 + a1), tk.mean_ms()


def run_plain(wl, device, steps, warmup, order_every=0):
    """The drop-in through Simulation.step; with ``order_every`` > 0 the
    simulation keeps its particle arrays in the Morton order of the cells,
    refreshed every that many steps inside the timed loop (form
    plain_fp32_ordered)."""
    import torch

    from offline_particles_amd.kernels.roms import ab3_w_advection_kernel
    from offline_particles_amd.workloads import build_simulation

    tk = _TimedKernel(ab3_w_advection_kernel)
    sim = build_simulation(wl, "ab3", device=device, ab_kernel=tk.kernel, spatial_order_every=order_every)
    fbytes = field_bytes_touched(wl, sim, wl.fields["u"].dtype.itemsize)
    elapsed, act, kms = _time_plain(sim, tk, steps, warmup)
    algo = act * PLAIN_AB3_STATE_BYTES + fbytes
    res = {
        "form": "plain_fp32_ordered" if order_every else "plain_fp32",
        "spatial_order_every": order_every,
        "particle_steps_per_s": act * steps / elapsed,
        "ms_per_step": 1e3 * elapsed / steps,
        "active_particles": act,
        "kernel": "k_ab3 (opx_ab3_w_advection)",
        "kernel_ms": kms,
        "algo_bytes": algo,
        "field_bytes_touched": fbytes,
        "achieved_GBs": algo / (kms * 1e-3) / 1e9,
    }
    del sim
    torch.cuda.empty_cache()
    return res


def run_plain_iso(wl, device, steps, warmup):
    """The isopycnal-following AB3 drop-in (ab3_isopycnal_following_timestepper's
    kernels: validation, ab3_isopycnal_following, its post step) on the workload
    with a density field; FP32 fields, FP64 positions."""
    import torch

    from offline_particles_amd.kernels.roms import (ab3_isopycnal_following_kernel,
                                                    ab3_post_isopycnal_following_kernel)
    from offline_particles_amd.kernels.validation import validation_kernel
    from offline_particles_amd.simulation import SimulationBuilder
    from offline_particles_amd.timesteppers import ABTimestepper
    from offline_particles_amd.workloads import build_fieldset

    tk = _TimedKernel(ab3_isopycnal_following_kernel)
    fs = build_fieldset(wl, device=device, constants=ISO_CONSTANTS)
    ts = ABTimestepper(wl.time_array, np.float64(wl.dt), ab_kernel=tk.kernel, index_padding=5,
                       pre_step_kernel=validation_kernel, post_step_kernel=ab3_post_isopycnal_following_kernel)
    sim = SimulationBuilder(ts, fs).build_simulation(wl.nparticles, device=device)
    sim.set_indices(*wl.positions)
    sim.set_particle_field("status", 10)
    sim.set_particle_field("rho_t", iso_targets(wl.nparticles))
    fbytes = field_bytes_touched(wl, sim, wl.fields["u"].dtype.itemsize, fields_3d=("u", "v", "rho"))
    elapsed, act, kms = _time_plain(sim, tk, steps, warmup)
    algo = act * PLAIN_ISO_STATE_BYTES + fbytes
    res = {
        "form": "plain_iso_fp32",
        "particle_steps_per_s": act * steps / elapsed,
        "ms_per_step": 1e3 * elapsed / steps,
        "active_particles": act,
        "kernel": "k_ab3_iso (opx_ab3_isopycnal_following)",
        "kernel_ms": kms,
        "algo_bytes": algo,
        "field_bytes_touched": fbytes,
        "achieved_GBs": algo / (kms * 1e-3) / 1e9,
        "roofline_frac": algo / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "constants": dict(ISO_CONSTANTS),
    }
    del sim, fs
    torch.cuda.empty_cache()
    return res


OUTPUT_TIDX = 0.5  # output events between the first two snapshots


def run_output(wl, device, rounds, warmup):
    """Output events of the drop-in API: one round is compute_z plus trilinear u,
    v, w and bilinear zeta, h for every particle (compute_z_kernel and the
    linear-interpolation output kernels, one HIP kernel each), launched through
    the Launcher at a time index between two snapshots."""
    import torch

    from offline_particles_amd.kernels import merge_particle_fields
    from offline_particles_amd.kernels.output import (bilinear_interpolation_kernel,
                                                      trilinear_interpolation_kernel)
    from offline_particles_amd.kernels.roms import compute_z_kernel
    from offline_particles_amd.launcher import Launcher
    from offline_particles_amd.particles import Particles
    from offline_particles_amd.workloads import OFFSETS, build_fieldset

    kernels = {"z": compute_z_kernel("_z_out")}
    for name in ("u", "v", "w"):
        kernels[name] = trilinear_interpolation_kernel(name, f"_{name}_out")
    for name in ("zeta", "h"):
        kernels[name] = bilinear_interpolation_kernel(name, f"_{name}_out")
    timed = {k: _TimedKernel(v) for k, v in kernels.items()}
    fs = build_fieldset(wl, device=device)
    launcher = Launcher(fs)
    n = wl.nparticles
    P = Particles(n, device=device, **merge_particle_fields([t.kernel for t in timed.values()]))
    for k, a in zip(("zidx", "yidx", "xidx"), wl.positions):
        P[k].copy_(torch.from_numpy(a))

    def one_round():
        for t in timed.values():
            launcher.launch_kernel(t.kernel, P, OUTPUT_TIDX)

    for _ in range(warmup):
        one_round()
    torch.cuda.synchronize()
    for t in timed.values():
        t.active = True
    t0 = time.perf_counter()
    for _ in range(rounds):
        one_round()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kms = {k: t.mean_ms() for k, t in timed.items()}
    pos = [P[k] for k in ("zidx", "yidx", "xidx")]
    ub = wl.fields["u"].dtype.itemsize
    # the trilinear kernel of u: its particle bytes plus the distinct u nodes of both snapshots
    algo_u = n * OUTPUT_TRILINEAR_BYTES + 2 * ub * touched_nodes(wl.fields["u"].shape[1:], OFFSETS["u"], pos)
    res = {
        "form": "output_fp32",
        "outputs_per_round": len(kernels),
        "particle_outputs_per_s": n * len(kernels) * rounds / elapsed,
        "ms_per_round": 1e3 * elapsed / rounds,
        "particles": n,
        "kernel_ms": kms,
        "kernel": "k_interp_trilinear (opx_interp_trilinear), field u",
        "algo_bytes": algo_u,
        "achieved_GBs": algo_u / (kms["u"] * 1e-3) / 1e9,
        "roofline_frac": algo_u / (kms["u"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "tidx": OUTPUT_TIDX,
    }
    del P, fs
    torch.cuda.empty_cache()
    return res


def run_output_fused(wl, device, rounds, warmup):
    """Output events of the fused form: AB3Engine.sample of u, v, w (trilinear),
    zeta, h (bilinear) and z (compute_z) for every particle in one kernel
    (opx_sample_split) over the engine's re-binned state, at the engine's time
    after 16 steps."""
    import torch

    from offline_particles_amd.engine import AB3Engine
    from offline_particles_amd.workloads import OFFSETS, build_fieldset

    fs = build_fieldset(wl, device=device)
    n = wl.nparticles
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=n, rebin_every=64, xlow=3)
    eng.add_release(*wl.positions, nsteps=32)
    eng.launch_steps = 16
    eng.rebin()
    eng.step(16)
    names = ("u", "v", "w", "zeta", "h")
    for _ in range(warmup):
        eng.sample(fs, names, depth="z")
    torch.cuda.synchronize()
    eng.timer = []
    t0 = time.perf_counter()
    for _ in range(rounds):
        eng.sample(fs, names, depth="z")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kms = float(np.mean([a.elapsed_time(b) for a, b, _ in eng.timer]))
    eng.timer = None
    act_mask = eng.status() < 128
    act = int(act_mask.sum().item())
    z, y, x = eng.positions()
    pos = [z[act_mask], y[act_mask], x[act_mask]]
    # state read (status, cell, 3 offsets, release id, pid = 23 B) and six FP64
    # outputs written per active particle, plus the distinct field nodes read
    fb = sum(2 * 4 * touched_nodes(wl.fields[k].shape[1:], OFFSETS[k], pos) for k in ("u", "v", "w"))
    n2 = touched_nodes(wl.fields["h"].shape, OFFSETS["h"], pos)
    fb += (2 * 4 + 4) * n2 + wl.fields["C"].nbytes
    algo = act * (23 + 6 * 8) + fb
    res = {
        "form": "output_fused_fp32",
        "outputs_per_round": 6,
        "particle_outputs_per_s": act * 6 * rounds / elapsed,
        "ms_per_round": 1e3 * elapsed / rounds,
        "particles": n,
        "active_particles": act,
        "kernel": "k_sample (opx_sample_split)",
        "kernel_ms": kms,
        "algo_bytes": algo,
        "achieved_GBs": algo / (kms * 1e-3) / 1e9,
        "roofline_frac": algo / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }
    del eng, fs
    torch.cuda.empty_cache()
    return res


def run_io(wl, device, steps=60, rounds=4):
    """The data formats either side of the path, on local disk (a temporary
    directory): the workload's FP32 snapshots staged to a Zarr v3 store, the
    fused engine (FP32, 16M particles, 3-slot streamed window) fed from that
    store (host-thread decode, pinned staging, side-stream copies) beside the
    same engine fed from pinned host memory, and the Zarr trajectory writer
    writing rounds of three FP64 outputs of every particle."""
    import shutil
    import tempfile

    import torch

    from offline_particles_amd.engine import AB3Engine
    from offline_particles_amd.fields import StaticField, TimeDependentField
    from offline_particles_amd.fieldset import Fieldset
    from offline_particles_amd.kernels.roms import compute_z_kernel
    from offline_particles_amd.output import Output, ZarrOutputBuilder
    from offline_particles_amd.snapshots import write_zarr_snapshots
    from offline_particles_amd.workloads import STAGGERS, TIME_DEPENDENT

    tmp = tempfile.mkdtemp(prefix="opx_io_")
    try:
        root = os.path.join(tmp, "snapshots.zarr")
        t0 = time.perf_counter()
        staged = 0
        lazy, host = {}, {}
        for name, arr in wl.fields.items():
            zs, ys, xs = STAGGERS[name]
            if name in TIME_DEPENDENT:
                write_zarr_snapshots(root, name, arr)
                staged += arr.nbytes
                lazy[name] = TimeDependentField.from_zarr(root, name, zs, ys, xs, cyclic=True, device=device)
                host[name] = TimeDependentField(arr, zs, ys, xs, cyclic=True, device="cpu")
            else:
                lazy[name] = host[name] = StaticField.from_numpy(arr, zs, ys, xs, device=device)
        stage_s = time.perf_counter() - t0
        consts = {"hc": wl.hc, "NZ": wl.NZ}
        res = {"staged_GB": staged / 1e9, "stage_write_GBs": staged / stage_s / 1e9}
        for label, flds in (("zarr", lazy), ("pinned_host", host)):
            fs = Fieldset(len(wl.time_array), wl.NZ, wl.NY, wl.NX, constants=consts, **flds)
            eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), device=device, capacity=wl.nparticles,
                            rebin_every=64, xlow=3, stream_slots=3)
            eng.launch_steps = 16
            eng.add_release(*wl.positions, nsteps=steps + 16)
            eng.rebin()
            eng.step(16)
            torch.cuda.synchronize()
            loads0, bytes0 = eng.snapshot_loads, eng.snapshot_bytes
            t0 = time.perf_counter()
            eng.step(steps)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            res[f"engine_from_{label}"] = {
                "particle_steps_per_s": eng.active_count() * steps / el, "steps": steps,
                "snapshot_loads": eng.snapshot_loads - loads0,
                "snapshot_GBs": (eng.snapshot_bytes - bytes0) / el / 1e9, "seconds": el}
            del eng, fs
            torch.cuda.empty_cache()
        # trajectory writer: rounds of three FP64 outputs of every particle
        n = wl.nparticles
        b = ZarrOutputBuilder("zarr", os.path.join(tmp, "traj.zarr"))
        for k in ("z", "y", "x"):
            b.add_output(Output(k, compute_z_kernel(f"_{k}"), particle_field=f"_{k}"))
        w = b.build(n)
        parts = {f"_{k}": torch.from_numpy(np.ascontiguousarray(a)).to(device) for k, a in zip("zyx", wl.positions)}
        state = types.SimpleNamespace(time=0.0, particles=parts)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r in range(rounds):
            state.time = float(r)
            w.write_time(state)
            for k in ("z", "y", "x"):
                w.write_output(k, state)
            w.finalise_write_round(state)
        el = time.perf_counter() - t0
        w.close()
        res["trajectory_writer"] = {"rounds": rounds, "outputs_per_round": 3, "particles": n,
                                    "GB_written": rounds * 3 * n * 8 / 1e9, "GBs": rounds * 3 * n * 8 / el / 1e9,
                                    "seconds_per_round": el / rounds, "chunk_files_per_round": 3 * -(-n // 10_000)}
        return res
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def field_bytes_touched_fused(wl, pos, storage_bytes_3d, fields_3d=("u", "v", "w")):
    """Distinct field bytes one fused launch reads: u, v, w (or rho; 2 snapshots),
    packed h/dx/dy, zeta (2), C."""
    from offline_particles_amd.workloads import OFFSETS

    total = 0
    for name in fields_3d:
        total += 2 * storage_bytes_3d * touched_nodes(wl.fields[name].shape[1:], OFFSETS[name], pos)
    n2 = touched_nodes(wl.fields["h"].shape, OFFSETS["h"], pos)
    total += 16 * n2 + 2 * 4 * n2
    total += wl.fields["C"].nbytes
    return total


# fused AB3 kernel: per active particle it reads status (1), release id (2),
# cell (4), 3 offsets (12), 6 history values (24) and z (8) and writes cell,
# offsets, history and z (48 B).
FUSED_STATE_BYTES = 1 + 2 + 4 + 12 + 24 + 8 + 48
# fused RK2 kernel: reads status, release id, cell, 3 offsets and z; writes
# cell, offsets and z (no history).  Its field term is counted at the step's
# start position only (a lower bound: stage 2 reads around x*).
FUSED_RK2_STATE_BYTES = 1 + 2 + 4 + 12 + 8 + 24
# fused isopycnal-following AB3 kernel: the AB3 state plus the FP64 target density
FUSED_ISO_STATE_BYTES = FUSED_STATE_BYTES + 8

# stencil layout names -> OPX_LAYOUT_* values of the C-ABI
LAYOUTS = {"ring": 0, "pair": 1, "quad": 2, "pair_lds": 3, "oct": 4, "ypair": 5}


def run_fused(wl, device, steps, warmup, storage, rebin_every, xlow, tune=False, layout=None, block=128,
              scheme="ab3", launch_steps=1):
    import torch

    from offline_particles_amd.engine import AB3Engine
    from offline_particles_amd.workloads import build_fieldset

    iso = scheme == "ab3_iso"
    fs = build_fieldset(wl, device=device, constants=ISO_CONSTANTS if iso else None)
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), storage_dtype=storage, device=device,
                    capacity=wl.nparticles, rebin_every=rebin_every, xlow=xlow, scheme=scheme)
    del fs
    # the release runs as long as the time array allows; the timed steps must
    # all be real steps of every particle (checked below)
    nsteps_max = int((wl.time_array[-1] - wl.time_array[0]) // wl.dt) - 2
    eng.add_release(*wl.positions, nsteps=nsteps_max, **({"rho_t": iso_targets(wl.nparticles)} if iso else {}))
    if layout is not None:
        eng.set_layout(LAYOUTS[layout])
    eng.block = block
    eng.launch_steps = launch_steps
    eng.rebin()
    eng.step(warmup)
    if tune:
        # steps per launch from 1, 8, 16, 32 and 64 (1, 32 and 64 are the profiled launch sizes)
        log("autotune:", eng.autotune(launch_steps=(1, 8, 16, 32, 64, 128)))
    torch.cuda.synchronize()
    a0 = eng.active_count()
    z, y, x = eng.positions()
    act = eng.status() < 128
    sb = 2 if storage == torch.bfloat16 else 4
    fbytes = field_bytes_touched_fused(wl, [z[act], y[act], x[act]], sb,
                                       ("u", "v", "rho") if iso else ("u", "v", "w"))
    del z, y, x, act
    # time whole re-sort cycles of whole launches: the requested steps rounded
    # up to a multiple of the re-bin cadence (itself whole launches), starting
    # with a re-bin, so the window holds the re-bins at their tuned rate
    m_tuned = max(1, eng.launch_steps)
    cycle = eng.rebin_every if eng.rebin_every > 0 else m_tuned
    timed = cycle * max(1, -(-steps // cycle))
    if eng.rebin_every > 0:
        eng._since_rebin = eng.rebin_every
    eng.timer = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.step(timed)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # per launch: the launches that took the full step count (a re-bin can cut one short)
    m = max(n for _, _, n in eng.timer)
    kms = float(np.mean([a.elapsed_time(b) for a, b, n in eng.timer if n == m]))
    eng.timer = None
    a1 = eng.active_count()
    eng.check_clock()
    if eng.steps_left(0) <= 0:
        raise RuntimeError("the release ran out of steps inside the timed region; lengthen the time array")
    elapsed = t1 - t0
    actn = 0.5 * (a0 + a1)
    # algorithmic bytes per particle-step times the particle-steps of one launch
    state_bytes = {"ab3": FUSED_STATE_BYTES, "rk2": FUSED_RK2_STATE_BYTES, "ab3_iso": FUSED_ISO_STATE_BYTES}[scheme]
    algo = actn * state_bytes + fbytes
    algo_launch = m * algo
    name = {"ab3": "fused_", "rk2": "fused_rk2_", "ab3_iso": "fused_iso_"}[scheme] + ("bf16" if sb == 2 else "fp32")
    kname = {"ab3": "k_ab3_fused", "rk2": "k_rk2_fused", "ab3_iso": "k_ab3_iso_fused"}[scheme]
    entry = {"ab3": "opx_ab3_fused", "rk2": "opx_rk2_fused", "ab3_iso": "opx_ab3_iso_fused"}[scheme]
    if m > 1 or iso:
        entry += "_steps"
    res = {
        "form": name,
        "particle_steps_per_s": actn * timed / elapsed,
        "ms_per_step": 1e3 * elapsed / timed,
        "timed_steps": timed,
        "rebins_timed": timed // eng.rebin_every if eng.rebin_every > 0 else 0,
        "active_particles": actn,
        "kernel": f"{kname}<{'bf16' if sb == 2 else 'f32'}> ({entry})",
        "kernel_ms": kms,
        "steps_per_launch": m,
        "algo_bytes": algo_launch,
        "algo_bytes_per_step": algo,
        # logical gather bytes: what the particles' stencils name, before any cache reuse (u, v, w:
        # 8 nodes x 2 snapshots each; h, zeta x 2, dx, dy corners and the two C levels ~ 32 x 4 B)
        "gather_bytes_per_particle_step": 48 * sb + 32 * 4,
        "field_bytes_touched": fbytes,
        "achieved_GBs": algo_launch / (kms * 1e-3) / 1e9,
        "rebin_every": eng.rebin_every,
        "xlow": eng.xlow,
        "block": eng.block,
        "layout": {v: k for k, v in LAYOUTS.items()}[eng.layout],
        "tuning": eng.tuning,
    }
    del eng
    torch.cuda.empty_cache()
    return res


def run_serving(device, seconds, rate=20.0, grid=(2048, 1536, 60), nsnap=4, max_size=2_000_000, graph_steps=8,
                tune=True, max_particles=256_000_000, storage=None, wl=None, reqs=None, tiles=16, tile_copy="auto"):
    """Serving form: Poisson arrivals (``rate``/s for ``seconds`` s), sizes log-uniform on
    [1k, max_size], 1-10 simulated days (720-7200 steps of 120 s), BF16 fields streamed
    through a 3-slot window, admission by free HBM; every request runs to completion.
    ``tiles``: each snapshot slot receives only the tiles of that many rho columns
    square that the particles can reach (0: whole snapshots)."""
    import torch

    from offline_particles_amd.engine import AB3Engine
    from offline_particles_amd.serving import ReleaseServer, poisson_requests
    from offline_particles_amd.workloads import build_fieldset, turbulent_workload

    NX, NY, NZ = grid
    storage = torch.bfloat16 if storage is None else storage
    t = time.perf_counter()
    days = 10
    if wl is None:
        # snapshots cycle, so the time array can cover any service clock: room for
        # the longest request after arrivals over `seconds` at up to 40,000 steps/s
        # (the service clock runs fastest while few particles are live)
        wl = turbulent_workload(NX=NX, NY=NY, NZ=NZ, nparticles=1,
                                steps=int(days * 86400 / 120) + int(40_000 * seconds) + 40,
                                nsnap_unique=nsnap, dtype=np.float32)
    if reqs is None:
        reqs = poisson_requests(wl, rate=rate, duration=seconds, size_range=(1_000, max_size),
                                steps_range=(720, 7200))
    else:
        for r in reqs:  # fresh copies of the same requests (the server fills in results)
            r.release, r.start_step, r.result = -1, -1, None
    log(f"serving: {len(reqs)} requests, {sum(r.size for r in reqs)} particles, "
        f"{sum(r.size * r.nsteps for r in reqs):.3e} particle-steps; workload in {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    fs = build_fieldset(wl, device="cpu", storage_dtype=storage)
    tuning = None
    layout, xlow, block, rebin_every, launch_steps = None, 4, 128, 40, 1
    if tune:
        # autotune on a probe: the 40 largest requests started together
        probe = sorted(reqs, key=lambda r: -r.size)[:40]
        eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), storage_dtype=storage, device=device,
                        capacity=sum(r.size for r in probe), max_releases=64, stream_slots=3,
                        stream_tiles=tiles or None)
        for r in probe:
            eng.add_release(r.zidx, r.yidx, r.xidx, nsteps=400)
        eng.rebin()
        # launches stay within a server chunk (graph_steps) and the 3-slot window
        tuning = eng.autotune(xlows=(4, 5), blocks=(128,), cadence_probe=30,
                              launch_steps=tuple(m for m in (1, 4, 8) if m <= max(1, graph_steps)))
        tuning = {k: v for k, v in tuning.items() if k != "trials"}
        log("serving autotune:", tuning)
        layout = LAYOUTS[tuning["layout"]]
        xlow, block, rebin_every = tuning["xlow"], tuning["block"], tuning["rebin_every"]
        launch_steps = tuning["launch_steps"]
        del eng
        torch.cuda.empty_cache()
    srv = ReleaseServer(fs, wl.time_array, np.float64(wl.dt), storage_dtype=storage, device=device,
                        stream_slots=3, max_releases=4096, rebin_every=rebin_every, xlow=xlow, layout=layout,
                        block=block, graph_steps=graph_steps, max_particles=max_particles,
                        launch_steps=launch_steps, stream_tiles=tiles or None)
    if tiles:
        srv.engine.tile_copy = tile_copy
    log(f"serving: server ready in {time.perf_counter() - t:.1f}s, capacity {srv.capacity} particles")
    srv.run(reqs)
    res = dict(srv.stats)
    res.update({"form": "served_" + ("bf16" if storage == torch.bfloat16 else "fp32"), "grid": [NX, NY, NZ],
                "arrival_rate_per_s": rate, "arrival_seconds": seconds, "tuning": tuning})
    del srv, fs
    torch.cuda.empty_cache()
    return res, wl, reqs


def run_batched(device, nrel, max_size, storage, graph_steps, grid=(1024, 768, 50)):
    """Batching form: uneven releases (own start time, duration, 10 km boxes) stepped by shared launches.

    The bin shape, layout and workgroup size come from the autotuner run on a
    probe engine holding a sample of the releases; the main run then steps
    every release to completion in runs of one tuned launch length, adding the
    releases that start within a run before it (each still starts at its own
    step), with runs of steps replayed from captured HIP graphs.  Particle-steps are the scheduled
    ones (size x steps of each release); the fraction of particles that left
    the domain (and stopped early) is reported beside it.
    """
    import torch

    from offline_particles_amd.engine import AB3Engine
    from offline_particles_amd.workloads import batched_releases, build_fieldset, turbulent_workload

    NX, NY, NZ = grid
    t = time.perf_counter()
    wl = turbulent_workload(NX=NX, NY=NY, NZ=NZ, nparticles=1, steps=int(96 * 3600 / 120) + 40, dtype=np.float32)
    rels = batched_releases(wl, nrel=nrel, size_range=(1_000, max_size))
    total = sum(r.size for r in rels)
    log(f"batched: {nrel} releases, {total} particles, workload in {time.perf_counter() - t:.1f}s")
    fs = build_fieldset(wl, device=device)
    # autotune on a probe: the 40 largest releases started together
    probe = sorted(rels, key=lambda r: -r.size)[:40]
    eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), storage_dtype=storage, device=device,
                    capacity=sum(r.size for r in probe), max_releases=64)
    for r in probe:
        eng.add_release(*r.positions, nsteps=1200)
    eng.rebin()
    tuning = eng.autotune(xlows=(3, 4, 5), blocks=(128,), cadence_probe=30, launch_steps=(1, 8, 16, 32, 64, 128))
    log("batched autotune:", {k: v for k, v in tuning.items() if k != "trials"})
    del eng
    torch.cuda.empty_cache()

    dt = wl.dt
    start_step = [int(np.floor(r.start_time / dt)) for r in rels]
    order = np.argsort(start_step, kind="stable")
    last = max(s + r.nsteps for s, r in zip(start_step, rels))
    eng = AB3Engine(fs, wl.time_array, np.float64(dt), storage_dtype=storage, device=device, capacity=total,
                    max_releases=nrel, rebin_every=tuning["rebin_every"], xlow=tuning["xlow"])
    eng.set_layout(LAYOUTS[tuning["layout"]])
    eng.block = tuning["block"]
    eng.launch_steps = tuning["launch_steps"]
    eng.graph_steps = graph_steps
    scheduled = sum(r.size * r.nsteps for r in rels)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nxt = 0
    g = 0
    rel_ids = {}
    # releases starting within the next run of steps are added ahead of their
    # start step (their particles wait, skipped, until it), so each run is one
    # tuned launch length rather than the few steps between two starts
    run = max(eng.launch_steps, graph_steps, 1)
    while g < last:
        k = min(run, last - g)
        while nxt < nrel and start_step[order[nxt]] < g + k:
            r = rels[order[nxt]]
            rel_ids[int(order[nxt])] = eng.add_release(*r.positions, start_time=r.start_time, nsteps=r.nsteps,
                                                       start_step=start_step[order[nxt]])
            nxt += 1
        eng.step(k)
        g += k
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    eng.check_clock()
    exited = float((eng.status() >= 128).float().mean().item())
    # the release whose run is nearest 1000 steps, read back for the accuracy check
    # against the oracle (its first 5000 particles)
    pick = int(np.argmin([abs(r.nsteps - 1000) for r in rels]))
    sl = eng.release_slice(rel_ids[pick])
    sample = {"release": pick, "particles": min(5000, rels[pick].size),
              "start_time": rels[pick].start_time, "nsteps": rels[pick].nsteps}
    n_s = sample["particles"]
    pos = [t_[sl][:n_s].cpu().numpy() for t_ in eng.positions()]
    sample["result"] = dict(zip(("zidx", "yidx", "xidx"), pos), status=eng.status()[sl][:n_s].cpu().numpy())
    sample["positions"] = tuple(a[:n_s] for a in rels[pick].positions)
    res = {
        "form": "batched_" + ("bf16" if storage == torch.bfloat16 else "fp32"),
        "releases": nrel,
        "particles": total,
        "global_steps": last,
        "particle_steps_scheduled": scheduled,
        "seconds": elapsed,
        "particle_steps_per_s": scheduled / elapsed,
        "exited_fraction": exited,
        "graph_steps": graph_steps,
        "tuning": {k: v for k, v in tuning.items() if k != "trials"},
    }
    del eng, fs
    torch.cuda.empty_cache()
    return res, sample, wl


def batched_accuracy(wl, samples):
    """Max / RMS deviation of one release of the batching mix (read back from the mix)
    from the FP64 oracle run of the same particles from the release's start time
    over its steps; one oracle run serves every storage form."""
    from oracle import OracleField, OracleSimulation
    from offline_particles_amd.workloads import OFFSETS, TIME_DEPENDENT

    s0 = next(iter(samples.values()))
    t = time.perf_counter()
    fields = {k: OracleField(v, OFFSETS[k], k in TIME_DEPENDENT) for k, v in wl.fields.items()}
    n = s0["particles"]
    sim = OracleSimulation(fields, wl.time_array, wl.dt, scheme="ab3", nparticles=n, hc=wl.hc, NZ=wl.NZ,
                           bounds=wl.bounds, threads=_cpu_threads(), time=s0["start_time"])
    for k, a in zip(("zidx", "yidx", "xidx"), s0["positions"]):
        sim.p[k][:] = a
    sim.p["status"][:] = 10
    sim.run(s0["nsteps"])
    log(f"batched accuracy: oracle {n} x {s0['nsteps']} steps in {time.perf_counter() - t:.1f}s")
    out = {}
    for name, smp in samples.items():
        g = smp["result"]
        both = (g["status"] < 128) & (sim.p["status"] < 128)
        d = np.stack([np.abs(g[k] - sim.p[k]) for k in ("zidx", "yidx", "xidx")])[:, both].max(axis=0)
        out[name] = {"release": smp["release"], "particles": n, "steps": smp["nsteps"],
                     "start_time": smp["start_time"], "max": float(d.max()) if d.size else 0.0,
                     "rms": float(np.sqrt((d**2).mean())) if d.size else 0.0,
                     "status_mismatch": int((g["status"] != sim.p["status"]).sum())}
    return out


# ---------------------------------------------------------------------------
# HBM traffic measured by rocprofv3 (FETCH_SIZE x2 + WRITE_SIZE x1, see profiles/)
# ---------------------------------------------------------------------------

# (form, steps per launch) -> summary
_PROFILE_FOR_FORM = {
    ("plain_fp32", 1): "k_ab3_plain_turbulent16M.md",
    ("plain_fp32_ordered", 1): "k_ab3_plain_ordered_turbulent16M.md",
    ("fused_fp32", 1): "k_ab3_fused_f32_turbulent16M.md",
    ("fused_bf16", 1): "k_ab3_fused_bf16_turbulent16M.md",
    ("fused_rk2_fp32", 1): "k_rk2_fused_f32_turbulent16M.md",
    ("fused_fp32", 32): "k_ab3_fused_steps32_f32_turbulent16M.md",
    ("fused_bf16", 32): "k_ab3_fused_steps32_bf16_turbulent16M.md",
    ("fused_rk2_fp32", 32): "k_rk2_fused_steps32_f32_turbulent16M.md",
    ("fused_fp32", 64): "k_ab3_fused_steps64_f32_turbulent16M.md",
    ("fused_bf16", 64): "k_ab3_fused_steps64_bf16_turbulent16M.md",
    ("fused_rk2_fp32", 64): "k_rk2_fused_steps64_f32_turbulent16M.md",
    ("plain_iso_fp32", 1): "k_ab3_iso_plain_turbulent16M.md",
    ("output_fp32", 1): "k_interp_trilinear_turbulent16M.md",
    ("fused_iso_fp32", 32): "k_ab3_iso_fused_steps32_turbulent16M.md",
    ("fused_iso_fp32", 64): "k_ab3_iso_fused_steps64_turbulent16M.md",
    ("output_fused_fp32", 1): "k_sample_turbulent16M.md",
}


def _profile_file(table, form, particles, grid, steps_per_launch):
    """The committed summary for (form, steps per launch) at the profiled workload
    (16M particles, 1024x768x50): the exact launch size, else the nearest
    profiled one of the form.  Returns (path, profiled steps per launch)."""
    if particles != 16_000_000 or tuple(grid) != (1024, 768, 50):
        return None, None
    sizes = [m for f, m in table if f == form]
    if not sizes:
        return None, None
    m = min(sizes, key=lambda k: (abs(np.log(k) - np.log(max(1, steps_per_launch))), -k))
    return os.path.join(ROOT, "profiles", table[(form, m)]), m


def _profile_value(path, key):
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def profiled_traffic(form, particles, grid, steps_per_launch=1):
    """HBM bytes per launch of the form's dominant kernel from its committed rocprofv3
    summary at this workload (16M particles, 1024x768x50).  A launch size with no
    summary of its own takes the per-step traffic of the nearest profiled size of
    the form times its steps (the state is read and written once per launch, so
    this slightly overstates longer launches and understates shorter ones)."""
    path, m = _profile_file(_PROFILE_FOR_FORM, form, particles, grid, steps_per_launch)
    if path is None:
        return None
    v = _profile_value(path, "traffic_bytes_per_launch")
    return None if v is None else v * steps_per_launch / m


def profiled_traffic_source(form, particles, grid, steps_per_launch=1):
    path, m = _profile_file(_PROFILE_FOR_FORM, form, particles, grid, steps_per_launch)
    if path is None:
        return None
    src = os.path.relpath(path, ROOT)
    return src if m == steps_per_launch else f"{src} (per step, {m}-step launches, scaled to {steps_per_launch})"


# (form, steps per launch) -> shader-counter summary holding valu_insts_per_wave_step
_COUNTERS_FOR_FORM = {
    ("fused_bf16", 32): "k_ab3_fused_steps32_bf16_turbulent16M.md",
    ("fused_fp32", 32): "k_ab3_fused_steps32_f32_turbulent16M.md",
    ("fused_bf16", 64): "k_ab3_fused_steps64_bf16_turbulent16M.md",
    ("fused_fp32", 64): "k_ab3_fused_steps64_f32_turbulent16M.md",
    ("fused_rk2_fp32", 32): "k_rk2_fused_steps32_f32_turbulent16M.md",
    ("fused_rk2_fp32", 64): "k_rk2_fused_steps64_f32_turbulent16M.md",
    ("fused_iso_fp32", 32): "k_ab3_iso_fused_steps32_turbulent16M.md",
    ("fused_iso_fp32", 64): "k_ab3_iso_fused_steps64_turbulent16M.md",
}


def valu_roofline(form, particles, grid, steps_per_launch, active, kernel_ms):
    """VALU issue rate of the several-step kernel: the wave64 VALU instructions per
    wave-step that rocprofv3 counted (SQ_INSTS_VALU / (SQ_WAVES x steps)) for this
    form, times the wave-steps of the live launch, over its HIP-event duration;
    peak = 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction."""
    path, m = _profile_file(_COUNTERS_FOR_FORM, form, particles, grid, steps_per_launch)
    if path is None or not kernel_ms or kernel_ms != kernel_ms:
        return None
    x = _profile_value(path, "valu_insts_per_wave_step")
    if x is None:
        return None
    achieved = x * (active / 64.0) * steps_per_launch / (kernel_ms * 1e-3)
    return {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_INSTS,
            "unit": "wave64 VALU instructions/s", "frac": achieved / VALU_PEAK_INSTS,
            "valu_insts_per_wave_step": x, "source": os.path.relpath(path, ROOT)
            + ("" if m == steps_per_launch else f" ({m}-step launches)")}


# ---------------------------------------------------------------------------
# CPU baseline: reference-structured FP64 oracle on the host cores
# ---------------------------------------------------------------------------

def _cpu_model():
    """Model name plus family / model / stepping and the logical CPUs the host has."""
    info = {}
    ncpu = 0
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k == "processor":
                ncpu += 1
            elif k in ("model name", "cpu family", "model", "stepping") and k not in info:
                info[k] = v.strip()
    except OSError:  # pragma: no cover
        return None
    if not info:
        return None
    return (f"{info.get('model name', '?')} (family {info.get('cpu family', '?')}, model {info.get('model', '?')}, "
            f"stepping {info.get('stepping', '?')}; {ncpu} logical CPUs on the host, "
            f"{len(os.sched_getaffinity(0))} in this process's affinity)")


def _cpu_threads():
    """Host threads for the CPU legs: the cores this process may run on
    (sched_getaffinity), capped by OMP_NUM_THREADS when the machine sets it (the
    GPU host gives one GPU's job a share of 16 cores and sets it to 16)."""
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        ncpu = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        ncpu = min(ncpu, int(omp))
    return max(1, ncpu)


def _oracle_sim(wl, n, scheme, threads):
    from oracle import OracleField, OracleSimulation
    from offline_particles_amd.workloads import OFFSETS, TIME_DEPENDENT

    fields = {k: OracleField(v, OFFSETS[k], k in TIME_DEPENDENT) for k, v in wl.fields.items()}
    extra = dict(rho_t=iso_targets(wl.nparticles)[:n], **ISO_CONSTANTS) if scheme == "ab3_iso" else {}
    sim = OracleSimulation(fields, wl.time_array, wl.dt, scheme=scheme, nparticles=n, hc=wl.hc, NZ=wl.NZ,
                           bounds=wl.bounds, threads=threads, **extra)
    for k, a in zip(("zidx", "yidx", "xidx"), wl.positions):
        sim.p[k][:] = a[:n]
    return sim


def cpu_baseline(wl, seconds, max_particles, scheme="ab3", max_steps=20):
    threads = _cpu_threads()
    n = min(max_particles, wl.nparticles)
    sim = _oracle_sim(wl, n, scheme, threads)
    sim.p["status"][:] = 10
    sim.step()  # warm-up (start-up step)
    steps, t0 = 0, time.perf_counter()
    act = 0
    while True:
        act += int((sim.p["status"] < 128).sum())
        sim.step()
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= max_steps:
            break
    what = "AB3" if scheme == "ab3" else "isopycnal-following AB3"
    return {
        "value": act / el,
        "unit": "particle-steps/s",
        "cores": threads,
        "cpu_model": _cpu_model(),
        "kind": "port",
        "sample": f"{n} particles x {steps} {what} steps of the same workload on the FP64 C/OpenMP oracle "
                  f"(full-grid time blend per launch, validation + AB3 + post per step), {el:.1f} s",
    }


def cpu_output_baseline(wl, seconds, max_particles):
    """Output rounds of compute_z and trilinear u, v, w on the oracle, each the
    launcher's full-grid blend of the field followed by the particle loop."""
    threads = _cpu_threads()
    n = min(max_particles, wl.nparticles)
    sim = _oracle_sim(wl, n, "ab3", threads)
    sim.tidx = OUTPUT_TIDX
    out = np.zeros(n)
    rounds, t0 = 0, time.perf_counter()
    while True:
        sim.output_compute_z(out)
        for name in ("u", "v", "w"):
            sim.output_trilinear(name, out)
        rounds += 1
        el = time.perf_counter() - t0
        if el >= seconds or rounds >= 5:
            break
    return {
        "value": n * 4 * rounds / el,
        "unit": "particle-outputs/s",
        "cores": threads,
        "cpu_model": _cpu_model(),
        "kind": "port",
        "sample": f"{n} particles x {rounds} rounds of compute_z + trilinear u, v, w on the FP64 C/OpenMP "
                  f"oracle (full-grid blend per output, then the particle loop), {el:.1f} s",
    }


def accuracy(wl, device, nsub, steps):
    """Max / RMS position deviation of each form from the FP64 oracle after ``steps`` steps."""
    import torch

    from oracle import OracleField, OracleSimulation
    from offline_particles_amd.engine import AB3Engine
    from offline_particles_amd.workloads import OFFSETS, TIME_DEPENDENT, build_fieldset, build_simulation

    pos = tuple(a[:nsub] for a in wl.positions)
    t = time.perf_counter()
    fields = {k: OracleField(v, OFFSETS[k], k in TIME_DEPENDENT) for k, v in wl.fields.items()}
    threads = _cpu_threads()
    sim = OracleSimulation(fields, wl.time_array, wl.dt, scheme="ab3", nparticles=nsub, hc=wl.hc, NZ=wl.NZ,
                           bounds=wl.bounds, threads=threads)
    for k, a in zip(("zidx", "yidx", "xidx"), pos):
        sim.p[k][:] = a
    sim.p["status"][:] = 10
    sim.run(steps)
    log(f"accuracy: oracle {nsub} x {steps} steps in {time.perf_counter() - t:.1f}s")
    ref = {k: sim.p[k] for k in ("zidx", "yidx", "xidx", "status")}

    def report(name, z, y, x, st):
        both = (st < 128) & (ref["status"] < 128)
        d = np.stack([np.abs(z - ref["zidx"]), np.abs(y - ref["yidx"]), np.abs(x - ref["xidx"])])[:, both]
        dm = d.max(axis=0)
        return {"max": float(dm.max()), "rms": float(np.sqrt((dm**2).mean())),
                "p99": float(np.quantile(dm, 0.99)),
                "max_zyx": [float(v) for v in d.max(axis=1)],
                "status_mismatch": int((st != ref["status"]).sum()), "compared": int(both.sum())}

    out = {"particles": nsub, "steps": steps}
    for name, every in (("plain_fp32", 0), ("plain_fp32_ordered", 20)):
        g = build_simulation(wl, "ab3", device=device, positions=pos, spatial_order_every=every)
        for _ in range(steps):
            g.step()
        out[name] = report(name, *(g.particles.to_numpy(k) for k in ("zidx", "yidx", "xidx", "status")))
        del g
    fs = build_fieldset(wl, device=device)
    res_bf16 = None
    for name, dt in (("fused_fp32", torch.float32), ("fused_bf16", torch.bfloat16)):
        eng = AB3Engine(fs, wl.time_array, np.float64(wl.dt), storage_dtype=dt, device=device, capacity=nsub)
        eng.add_release(*pos, nsteps=steps)
        eng.step(steps)
        eng.check_clock()
        z, y, x = (t_.cpu().numpy() for t_ in eng.positions())
        st = eng.status().cpu().numpy()
        out[name] = report(name, z, y, x, st)
        if dt == torch.bfloat16:
            res_bf16 = (z, y, x, st)
        del eng
    # BF16 storage error vs kernel arithmetic: the oracle again on u, v, w rounded
    # to BF16 (held in FP32), against which the BF16 form differs only by its
    # arithmetic
    t = time.perf_counter()
    rounded = dict(wl.fields)
    for k in ("u", "v", "w"):
        rounded[k] = torch.from_numpy(wl.fields[k]).to(torch.bfloat16).to(torch.float32).numpy()
    fields_r = {k: OracleField(v, OFFSETS[k], k in TIME_DEPENDENT) for k, v in rounded.items()}
    sim = OracleSimulation(fields_r, wl.time_array, wl.dt, scheme="ab3", nparticles=nsub, hc=wl.hc, NZ=wl.NZ,
                           bounds=wl.bounds, threads=threads)
    for k, a in zip(("zidx", "yidx", "xidx"), pos):
        sim.p[k][:] = a
    sim.p["status"][:] = 10
    sim.run(steps)
    del fields_r, rounded
    log(f"accuracy: oracle on BF16-rounded fields {nsub} x {steps} steps in {time.perf_counter() - t:.1f}s")
    ref = {k: sim.p[k] for k in ("zidx", "yidx", "xidx", "status")}
    out["fused_bf16_vs_bf16_rounded_fields"] = report("fused_bf16_rounded", *res_bf16)
    log("accuracy:", out)
    return out


_DTYPE = {
    "plain_fp32": ("f64", "FP32 storage of every field; FP64 positions, blends, interpolation and vertical "
                          "transform (the drop-in's arithmetic)"),
    "plain_fp32_ordered": ("f64", "as plain_fp32, particle arrays in the Morton order of their cells"),
    "fused_fp32": ("f32", "FP32 storage of u, v, w; FP32 interpolation of u and v; FP64 blends of w, h, zeta, "
                          "FP64 vertical transform and clocks; positions as packed cells plus 32-bit fixed-point "
                          "offsets"),
    "fused_bf16": ("bf16", "BF16 storage of u, v, w, widened to FP32 for interpolation; otherwise as fused_fp32"),
}


def _meets_bar(name, acc):
    """Whether a form's accuracy run met the bar; without one, only FP32-storage forms
    qualify (BF16 storage misses the bar on this flow, see the accuracy report)."""
    if acc is None:
        return "bf16" not in name
    a = acc.get(name)
    return a is not None and a["max"] <= ACCURACY_BAR


def build_output(args, forms, grid, cpu, acc, batching, serving, isopycnal=None, output_events=None) -> dict:
    """The JSON line: headline = the fastest AB3 form at the 16M-particle workload
    among those whose accuracy run meets the bar."""
    NX, NY, NZ = grid
    for k, f in forms.items():
        if acc is not None and k in acc:
            f["max_dev_vs_oracle"] = acc[k]["max"]
            f["meets_accuracy_bar"] = acc[k]["max"] <= ACCURACY_BAR
    ab3_all = [f for k, f in forms.items() if "rk2" not in k]
    ab3 = [f for f in ab3_all if _meets_bar(f["form"], acc)]
    bar_met = bool(ab3)
    if not ab3:
        ab3 = ab3_all or list(forms.values())
    if not ab3:
        # no 16M form was asked for: the batching form heads the line
        if not batching:
            raise SystemExit("bench.py: no form was measured (give --forms or --batched)")
        b = next(iter(batching.values()))
        ab3 = [dict(b, kernel="k_ab3_fused (batched releases)", ms_per_step=float("nan"), achieved_GBs=None,
                    algo_bytes=None)]
    head = max(ab3, key=lambda f: f["particle_steps_per_s"])
    m = head.get("steps_per_launch", 1)
    dtype, detail = _DTYPE.get(head["form"], ("f32", None))

    out = {
        "metric": "particle-steps/s (16M, AB3) plain vs batched/BF16; p50/p99 release latency",
        "value": head["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "dtype_detail": detail,
        "data": "synthetic (seeded turbulent ROMS fields, uniform random particles)",
        "config": {
            "workload": f"{'RK2' if 'rk2' in head['form'] else 'AB3'} dt=120s, {args.particles} particles, "
                        f"{NX}x{NY}x{NZ} grid, 8 hourly snapshots (FP32 data; the {head['form']} form)",
            "form": head["form"],
            "particles": args.particles,
            "grid": [NX, NY, NZ],
            "timed_steps": head.get("timed_steps", args.steps),
            "headline_rule": (f"fastest AB3 form whose max deviation from the FP64 oracle after "
                              f"{acc['steps'] if acc else '-'} steps is <= {ACCURACY_BAR:g} grid-index units"
                              + ("" if acc else " (no accuracy run: FP32-storage forms only)")
                              + ("" if bar_met else "; NO form met it, fastest form shown")),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": head["kernel"],
            "achieved": head["achieved_GBs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": head["achieved_GBs"] / HBM_PEAK_GBS if head["achieved_GBs"] is not None else None,
            "traffic": profiled_traffic(head["form"], args.particles, (NX, NY, NZ), m),
            "traffic_source": profiled_traffic_source(head["form"], args.particles, (NX, NY, NZ), m),
            "algorithmic_bytes": head["algo_bytes"],
            "steps_per_launch": m,
            "logical_gather_bytes": (head["gather_bytes_per_particle_step"] * head["active_particles"] * m
                                     if "gather_bytes_per_particle_step" in head else None),
            "note": (None if m == 1 else
                     f"per-step equivalent: algorithmic bytes are the per-particle-step figure times the {m} steps "
                     f"of each particle in one launch; the launch keeps state and unchanged stencils in registers "
                     f"and moves only 'traffic' through HBM, so this kernel is bound by VALU issue "
                     f"(roofline_valu), not by HBM"),
        },
        "roofline_valu": (valu_roofline(head["form"], args.particles, (NX, NY, NZ), m, head["active_particles"],
                                        head.get("kernel_ms")) if m > 1 else None),
        "cpu_baseline": cpu,
        "accuracy": acc,
        "batching": batching,
        "serving": serving,
        "isopycnal": isopycnal,
        "output_events": output_events,
        "forms": {k: {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                  for k, v in forms.items()},
    }
    for k, f in out["forms"].items():
        sl = f.get("steps_per_launch", 1)
        if sl > 1:
            f["roofline_valu"] = valu_roofline(k, args.particles, (NX, NY, NZ), sl, forms[k]["active_particles"],
                                               forms[k].get("kernel_ms"))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--particles", type=int, default=16_000_000)
    ap.add_argument("--grid", default="1024x768x50", help="NXxNYxNZ")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-particles", type=int, default=16_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="plain_fp32,plain_fp32_ordered,fused_fp32,fused_bf16,fused_rk2_fp32")
    ap.add_argument("--order-every", type=int, default=20,
                    help="steps between spatial reorders of the plain_fp32_ordered form")
    ap.add_argument("--rebin-every", type=int, default=25)
    ap.add_argument("--xlow", type=int, default=2)
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (use --xlow/--rebin-every)")
    ap.add_argument("--layout", choices=tuple(LAYOUTS), default=None,
                    help="stencil layout when not tuning (default: pair)")
    ap.add_argument("--block", type=int, default=128, help="workgroup size when not tuning")
    ap.add_argument("--accuracy", type=int, default=100_000, metavar="N",
                    help="also report deviation from the FP64 oracle on the first N particles")
    ap.add_argument("--accuracy-steps", type=int, default=1000)
    ap.add_argument("--batched", type=int, default=256, metavar="NREL",
                    help="also run the batching form with NREL releases (256 for the full case)")
    ap.add_argument("--batched-max-size", type=int, default=1_000_000)
    ap.add_argument("--mix-storage", default="fp32,bf16",
                    help="field storage of the batching and serving forms (comma-separated: fp32, bf16)")
    ap.add_argument("--serve-tiles", type=int, default=16, metavar="CELLS",
                    help="serving form: stream only the tiles (CELLS rho columns square) of each snapshot that "
                         "the particles can reach (0: whole snapshots)")
    ap.add_argument("--serve", type=float, default=5.0, metavar="SECONDS",
                    help="serving form: Poisson arrivals at 20/s for SECONDS on the 2048x1536x60 BF16 grid")
    ap.add_argument("--serve-grid", default="2048x1536x60", help="NXxNYxNZ of the serving form")
    ap.add_argument("--graph-steps", type=int, default=8, help="steps per captured HIP graph (0: no graphs)")
    ap.add_argument("--launch-steps", type=int, default=1, help="fused steps per launch (1..128)")
    ap.add_argument("--extras", default="iso,output",
                    help="also measure the isopycnal-following AB3 (iso), the output events (output) and, "
                         "on request, the Zarr snapshot ingest and trajectory writer on local disk (io) at the "
                         "same workload; empty for none")
    ap.add_argument("--output-rounds", type=int, default=10, help="timed output rounds of the output form")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (scale reference for rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus N>1 is not supported"}))
        sys.exit(2)

    import torch

    from offline_particles_amd import _lib
    from offline_particles_amd.workloads import turbulent_workload

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    _lib.load()
    device = torch.device("cuda", 0)
    NX, NY, NZ = (int(s) for s in args.grid.split("x"))
    t = time.perf_counter()
    # time axis: warm-up, the autotuner's trial steps (~600) and the timed window
    # (the requested steps rounded up to whole re-sort cycles of at most 448 steps)
    wl = turbulent_workload(NX=NX, NY=NY, NZ=NZ, nparticles=args.particles,
                            steps=args.steps + args.warmup + 1600, dtype=np.float32)
    log(f"workload generated in {time.perf_counter() - t:.1f}s: grid {NX}x{NY}x{NZ}, {args.particles} particles")

    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device=device)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 3 copies of 1 GiB (torch copy kernel)")
        del src, dst

    forms = {}
    wanted = args.forms.split(",")
    if "plain_fp32" in wanted:
        forms["plain_fp32"] = run_plain(wl, device, args.steps, args.warmup)
        log("plain_fp32", forms["plain_fp32"])
    if "plain_fp32_ordered" in wanted:
        forms["plain_fp32_ordered"] = run_plain(wl, device, args.steps, args.warmup, order_every=args.order_every)
        log("plain_fp32_ordered", forms["plain_fp32_ordered"])
    for name, dt, scheme in (("fused_fp32", torch.float32, "ab3"), ("fused_bf16", torch.bfloat16, "ab3"),
                             ("fused_rk2_fp32", torch.float32, "rk2"), ("fused_rk2_bf16", torch.bfloat16, "rk2")):
        if name in wanted:
            forms[name] = run_fused(wl, device, args.steps, args.warmup, dt, args.rebin_every, args.xlow,
                                    tune=not args.no_tune, layout=args.layout, block=args.block, scheme=scheme,
                                    launch_steps=args.launch_steps)
            log(name, forms[name])

    extras = set(filter(None, args.extras.split(",")))
    isopycnal = output_events = wl_iso = None
    if "iso" in extras:
        t = time.perf_counter()
        wl_iso = with_density(wl)
        log(f"density field generated in {time.perf_counter() - t:.1f}s")
        plain = run_plain_iso(wl_iso, device, args.steps, args.warmup)
        plain["traffic"] = profiled_traffic("plain_iso_fp32", args.particles, (NX, NY, NZ))
        log("isopycnal plain", plain)
        fused = run_fused(wl_iso, device, args.steps, args.warmup, torch.float32, args.rebin_every, args.xlow,
                          tune=not args.no_tune, layout=args.layout, block=args.block, scheme="ab3_iso",
                          launch_steps=args.launch_steps)
        fused["roofline_frac"] = fused["achieved_GBs"] / HBM_PEAK_GBS
        fused["traffic"] = profiled_traffic("fused_iso_fp32", args.particles, (NX, NY, NZ),
                                            fused["steps_per_launch"])
        if fused["steps_per_launch"] > 1:
            fused["roofline_valu"] = valu_roofline("fused_iso_fp32", args.particles, (NX, NY, NZ),
                                                   fused["steps_per_launch"], fused["active_particles"],
                                                   fused.get("kernel_ms"))
        log("isopycnal fused", fused)
        isopycnal = {"plain_iso_fp32": plain, "fused_iso_fp32": fused}
    if "output" in extras:
        plain = run_output(wl, device, args.output_rounds, 2)
        plain["traffic"] = profiled_traffic("output_fp32", args.particles, (NX, NY, NZ))
        log("output plain", plain)
        fused = run_output_fused(wl, device, args.output_rounds, 2)
        fused["traffic"] = profiled_traffic("output_fused_fp32", args.particles, (NX, NY, NZ))
        log("output fused", fused)
        output_events = {"output_fp32": plain, "output_fused_fp32": fused}
    io = None
    if "io" in extras:
        io = run_io(wl, device)
        log("io", io)

    storages = {"fp32": torch.float32, "bf16": torch.bfloat16}
    # the serving form runs before the batching form: after the batching form's
    # allocations (36M particles, 8 packed windows) the same serving run measured
    # 15-20% slower in one process (its step kernel too), while the batching
    # form measures the same either way
    serving = None
    if args.serve > 0:
        sg = tuple(int(v) for v in args.serve_grid.lower().split("x"))
        serving = {}
        wl_s = reqs = None
        for st in args.mix_storage.split(","):
            res, wl_s, reqs = run_serving(device, args.serve, grid=sg, graph_steps=args.graph_steps,
                                          tune=not args.no_tune, storage=storages[st], wl=wl_s, reqs=reqs,
                                          tiles=args.serve_tiles)
            serving[res["form"]] = res
            log("serving", res)
        del wl_s, reqs
    batching = None
    samples, wl_b = {}, None
    if args.batched:
        batching = {}
        for st in args.mix_storage.split(","):
            res, samples["batched_" + st], wl_b = run_batched(device, args.batched, args.batched_max_size,
                                                             storages[st], args.graph_steps)
            batching[res["form"]] = res
            log("batched", res)
    # CPU leg: the FP64 oracle is timed on a bounded sample of the workload
    # (the baseline), and run as the checker of every GPU form's positions on
    # a subsample over 1000 steps (the accuracy report)
    cpu = acc = None
    if not args.no_cpu:
        t = time.perf_counter()
        cpu = cpu_baseline(wl, args.cpu_seconds, args.cpu_particles)
        log(f"cpu baseline in {time.perf_counter() - t:.1f}s:", cpu)
        if isopycnal is not None:
            isopycnal["cpu_baseline"] = cpu_baseline(wl_iso, args.cpu_seconds, args.cpu_particles // 4,
                                                     scheme="ab3_iso", max_steps=5)
            log("isopycnal cpu baseline:", isopycnal["cpu_baseline"])
        if output_events is not None:
            output_events["cpu_baseline"] = cpu_output_baseline(wl, args.cpu_seconds, args.cpu_particles)
            log("output cpu baseline:", output_events["cpu_baseline"])
        if args.accuracy:
            wl_acc = wl
            if wl.time_array[-1] - wl.time_array[0] < (args.accuracy_steps + 5) * wl.dt:
                wl_acc = turbulent_workload(NX=NX, NY=NY, NZ=NZ, nparticles=args.accuracy,
                                            steps=args.accuracy_steps + 5, dtype=np.float32)
            acc = accuracy(wl_acc, device, args.accuracy, args.accuracy_steps)
            del wl_acc
        if samples:
            for name, a in batched_accuracy(wl_b, samples).items():
                batching[name]["accuracy"] = a
                batching[name]["meets_accuracy_bar"] = a["max"] <= ACCURACY_BAR
            log("batched accuracy:", {k: v.get("accuracy") for k, v in batching.items()})
    del wl_b, samples

    out = build_output(args, forms, (NX, NY, NZ), cpu, acc, batching, serving, isopycnal, output_events)
    if io is not None:
        out["io"] = io
    print(json.dumps(out))


if __name__ == "__main__":
    main()
