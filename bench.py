#!/usr/bin/env python3
"""Throughput benchmark of the advection hot path on one MI355X.

Workload (default): 16M particles, AB3 (Euler -> AB2 -> AB3 start-up), dt =
120 s, on a synthetic 1024 x 768 x 50 ROMS grid with 8 hourly snapshots cycled
(turbulent k^-3 flow, 2% land).  A "step" is one AB3 step of every particle:
validation, the AB3 stage and the post-step zidx update.

Forms measured in the same run:
- plain_fp32: the drop-in behind the reference API (Simulation.step ->
  validation, ab3_w_advection, ab3_post_w_advection HIP kernels), FP32 fields,
  FP64 positions and FP64 interpolation.

Prints ONE JSON line.  ``value`` is particle-steps/s of the headline form with
all inputs resident in HBM; particle-steps count the particles active during
the timed steps (average of the active counts before and after).
``roofline`` is computed for the dominant kernel from algorithmic bytes
(compulsory particle state + distinct field nodes touched) over its average
duration measured with HIP events on its stream.  ``cpu_baseline`` times the
reference-structured FP64 CPU oracle on the host cores over a bounded sample.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


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


def field_bytes_touched(wl, sim, storage_bytes_3d):
    """Bytes of distinct field values one AB3 launch must read (2 snapshots for u, v, w, zeta)."""
    from offline_particles_amd.workloads import OFFSETS

    P = sim.particles
    act = P["status"] < 128
    pos = [P[k][act] for k in ("zidx", "yidx", "xidx")]
    b2 = wl.fields["h"].dtype.itemsize
    total = 0
    for name in ("u", "v", "w"):
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


def run_plain(wl, device, steps, warmup):
    import torch

    from offline_particles_amd.kernels.roms import ab3_w_advection_kernel
    from offline_particles_amd.workloads import build_simulation

    tk = _TimedKernel(ab3_w_advection_kernel)
    sim = build_simulation(wl, "ab3", device=device, ab_kernel=tk.kernel)
    for _ in range(warmup):
        sim.step()
    torch.cuda.synchronize()
    a0 = active_count(sim)
    fbytes = field_bytes_touched(wl, sim, wl.fields["u"].dtype.itemsize)
    tk.active = True
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    a1 = active_count(sim)
    tk.active = False
    elapsed = t1 - t0
    act = 0.5 * (a0 + a1)
    kms = tk.mean_ms()
    algo = act * PLAIN_AB3_STATE_BYTES + fbytes
    res = {
        "form": "plain_fp32",
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


# ---------------------------------------------------------------------------
# CPU baseline: reference-structured FP64 oracle on the host cores
# ---------------------------------------------------------------------------

def cpu_baseline(wl, seconds, max_particles):
    from oracle import OracleField, OracleSimulation
    from offline_particles_amd.workloads import OFFSETS, TIME_DEPENDENT

    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        ncpu = os.cpu_count() or 1
    threads = max(1, min(16, ncpu))
    n = min(max_particles, wl.nparticles)
    fields = {k: OracleField(v, OFFSETS[k], k in TIME_DEPENDENT) for k, v in wl.fields.items()}
    sim = OracleSimulation(fields, wl.time_array, wl.dt, scheme="ab3", nparticles=n, hc=wl.hc, NZ=wl.NZ,
                           bounds=wl.bounds, threads=threads)
    for k, a in zip(("zidx", "yidx", "xidx"), wl.positions):
        sim.p[k][:] = a[:n]
    sim.p["status"][:] = 10
    sim.step()  # warm-up (start-up step)
    steps, t0 = 0, time.perf_counter()
    act = 0
    while True:
        act += int((sim.p["status"] < 128).sum())
        sim.step()
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 20:
            break
    return {
        "value": act / el,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} particles x {steps} AB3 steps of the same workload on the FP64 C/OpenMP oracle "
                  f"(full-grid time blend per launch, validation + AB3 + post per step), {el:.1f} s",
    }


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
    wl = turbulent_workload(NX=NX, NY=NY, NZ=NZ, nparticles=args.particles, steps=args.steps + args.warmup + 5,
                            dtype=np.float32)
    log(f"workload generated in {time.perf_counter() - t:.1f}s: grid {NX}x{NY}x{NZ}, {args.particles} particles")

    forms = {}
    forms["plain_fp32"] = run_plain(wl, device, args.steps, args.warmup)
    log("plain_fp32", forms["plain_fp32"])

    head = forms["plain_fp32"]
    cpu = None
    if not args.no_cpu:
        t = time.perf_counter()
        cpu = cpu_baseline(wl, args.cpu_seconds, args.cpu_particles)
        log(f"cpu baseline in {time.perf_counter() - t:.1f}s:", cpu)

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
        "dtype": "f64",
        "data": "synthetic (seeded turbulent ROMS fields, uniform random particles)",
        "config": {
            "workload": f"AB3 dt=120s, {args.particles} particles, {NX}x{NY}x{NZ} grid, 8 hourly FP32 snapshots",
            "form": head["form"],
            "particles": args.particles,
            "grid": [NX, NY, NZ],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": head["kernel"],
            "achieved": head["achieved_GBs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": head["achieved_GBs"] / HBM_PEAK_GBS,
            "traffic": None,
        },
        "cpu_baseline": cpu,
        "forms": {k: {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                  for k, v in forms.items()},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
