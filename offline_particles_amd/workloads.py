"""Seeded synthetic ROMS workloads for tests, smoke runs and benchmarks.

Two flow families on a ROMS C-grid (Vtransform = 2, Vstretching = 4,
theta_s = 5, theta_b = 0.4, hc = 10 m, dx = dy = 1 km):

- ``eddy``: 20 m + 180 m Gaussian seamount (sigma = 15 cells) under a Rankine
  eddy (solid body inside 20 km, 0.5 m/s peak, 1/r outside) with a linear
  vertical shear; the second snapshot is the first scaled by 0.8.
- ``turbulent``: random smooth bathymetry (50-4000 m, k^-2 amplitude spectrum with a
  Gaussian roll-off at 192 cells), horizontal velocity from a random-phase
  transport streamfunction (k^-3 amplitude spectrum with a 64-cell roll-off,
  times the depth; RMS 0.3 m/s) whose depth-integrated transport is exactly
  divergence free on the C-grid, modes advancing by 2 pi m/K per snapshot so the K
  snapshots repeat smoothly, and the shallowest 2% of columns masked as land
  (the streamfunction is tapered to zero around land, so u and v vanish on
  land faces and stay divergence free; w is zero in land columns).

In both, w is the parcel vertical velocity from continuity: omega integrated
upward from the seabed plus u.grad(z) of the sloping s-surfaces (so particles
follow the bathymetry), and the free surface is small and time-dependent.  Array shapes follow the ROMS staggering
(spatial_arrays.py:42-52 in the reference): u (NZ, NY, NX-1), v (NZ, NY-1, NX),
w (NZ+1, NY, NX), h/dx/dy/zeta (NY, NX), C = Cs_w (NZ+1).

The snapshot axis is cyclic: snapshot n is ``data[n % K]``; ``time_array``
is long enough for ``steps`` steps of ``dt``.
"""

from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

THETA_S = 5.0
THETA_B = 0.4
HC = 10.0

# Stagger offsets of the active dimensions (spatial_arrays.py:29-40).
OFFSETS = {
    "u": (0.0, 0.0, -0.5),
    "v": (0.0, -0.5, 0.0),
    "w": (0.5, 0.0, 0.0),
    "dx": (0.0, 0.0),
    "dy": (0.0, 0.0),
    "h": (0.0, 0.0),
    "zeta": (0.0, 0.0),
    "C": (0.5,),
    "rho": (0.0, 0.0, 0.0),
}
STAGGERS = {
    "u": ("center", "center", "inner"),
    "v": ("center", "inner", "center"),
    "w": ("outer", "center", "center"),
    "dx": ("invariant", "center", "center"),
    "dy": ("invariant", "center", "center"),
    "h": ("invariant", "center", "center"),
    "zeta": ("invariant", "center", "center"),
    "C": ("outer", "invariant", "invariant"),
    "rho": ("center", "center", "center"),
}
TIME_DEPENDENT = ("u", "v", "w", "zeta", "rho")


def stretching_cs_w(NZ: int, theta_s: float = THETA_S, theta_b: float = THETA_B) -> np.ndarray:
    """ROMS Vstretching = 4 (Shchepetkin 2010) at the NZ+1 w-levels."""
    sigma = (np.arange(NZ + 1, dtype=np.float64) - NZ) / NZ
    if theta_s > 0:
        C = (1.0 - np.cosh(theta_s * sigma)) / (np.cosh(theta_s) - 1.0)
    else:
        C = -(sigma**2)
    if theta_b > 0:
        C = (np.exp(theta_b * C) - 1.0) / (1.0 - np.exp(-theta_b))
    C[0], C[-1] = -1.0, 0.0
    return C


def layer_thickness(h: np.ndarray, C: np.ndarray, hc: float = HC) -> np.ndarray:
    """Hz (NZ, NY, NX) of the Vtransform = 2 grid with zeta = 0."""
    NZ = C.shape[0] - 1
    sigma_w = (np.arange(NZ + 1, dtype=np.float64) - NZ) / NZ
    S = (hc * sigma_w[:, None, None] + h[None] * C[:, None, None]) / (hc + h[None])
    z_w = h[None] * S
    return np.diff(z_w, axis=0)


class _ContinuityGeometry:
    """The snapshot-independent parts of ``_w_from_continuity`` for one grid:
    layer thicknesses at u and v faces and the slopes of the fixed s-surfaces
    at the NZ+1 w-levels (computed once, shared by every snapshot)."""

    def __init__(self, h: np.ndarray, C: np.ndarray, dx: float, dy: float, hc: float = HC):
        NZ = C.shape[0] - 1
        Hz = layer_thickness(h, C, hc).astype(np.float32)
        self.Hu = 0.5 * (Hz[:, :, 1:] + Hz[:, :, :-1])
        self.Hv = 0.5 * (Hz[:, 1:, :] + Hz[:, :-1, :])
        del Hz
        sigma_w = (np.arange(NZ + 1, dtype=np.float64) - NZ) / NZ
        Sw = ((hc * sigma_w[:, None, None] + h[None] * C[:, None, None]) / (hc + h[None])).astype(np.float32)
        # d z_w / dy, d z_w / dx per w-level
        self.slopes = [np.gradient((h * Sw[k]).astype(np.float32), dy, dx) for k in range(NZ + 1)]


def _w_from_continuity(u: np.ndarray, v: np.ndarray, h: np.ndarray, C: np.ndarray, dx: float,
                       dy: float, hc: float = HC, geom: _ContinuityGeometry | None = None) -> np.ndarray:
    """Physical w at the NZ+1 w-levels: omega from continuity plus u.grad(z_w).

    omega (the flux through sigma surfaces) is integrated upward from zero at
    the seabed, omega_{k+1} = omega_k - div(u Hz, v Hz); when the depth
    integrated transport is divergence free it also vanishes at the surface.
    The fixed s-surfaces z_w(x, y) slope, so the vertical velocity of a fluid
    parcel is w = omega + u.grad(z_w), which equals -u.grad(h) at the seabed.
    ``geom`` holds the grid's snapshot-independent terms (built here if absent).
    """
    NZ = C.shape[0] - 1
    if geom is None:
        geom = _ContinuityGeometry(h, C, dx, dy, hc)
    NY, NX = h.shape
    U = (u * geom.Hu).astype(np.float32)
    V = (v * geom.Hv).astype(np.float32)
    div = np.zeros((NZ, NY, NX), np.float32)
    div[:, :, :-1] += U / dx
    div[:, :, 1:] -= U / dx
    div[:, :-1, :] += V / dy
    div[:, 1:, :] -= V / dy
    del U, V
    w = np.zeros((NZ + 1, NY, NX), np.float32)
    np.cumsum(-div, axis=0, out=w[1:])
    del div
    # u, v at rho points and w-levels
    ur = np.zeros((NZ, NY, NX), np.float32)
    ur[:, :, 1:-1] = 0.5 * (u[:, :, :-1] + u[:, :, 1:])
    ur[:, :, 0], ur[:, :, -1] = u[:, :, 0], u[:, :, -1]
    vr = np.zeros((NZ, NY, NX), np.float32)
    vr[:, 1:-1, :] = 0.5 * (v[:, :-1, :] + v[:, 1:, :])
    vr[:, 0, :], vr[:, -1, :] = v[:, 0, :], v[:, -1, :]
    for k in range(NZ + 1):
        gy, gx = geom.slopes[k]
        if k == 0:
            uk, vk = ur[0], vr[0]
        elif k == NZ:
            uk, vk = ur[NZ - 1], vr[NZ - 1]
        else:
            uk, vk = 0.5 * (ur[k - 1] + ur[k]), 0.5 * (vr[k - 1] + vr[k])
        w[k] += uk * gx + vk * gy
    return w


@dataclass
class Workload:
    """A synthetic ROMS workload (host arrays; fields in ``dtype``)."""

    name: str
    NX: int
    NY: int
    NZ: int
    hc: float
    dt: float
    steps: int
    snapshot_dt: float
    fields: dict  # name -> array; time-dependent ones are (K, ...)
    time_array: np.ndarray
    positions: tuple  # (zidx, yidx, xidx) float64 arrays
    seed: int
    meta: dict = field(default_factory=dict)

    @property
    def nparticles(self) -> int:
        return int(self.positions[0].shape[0])

    @property
    def bounds(self):
        return (0.0, float(self.NZ - 1), 0.0, float(self.NY - 1), 0.0, float(self.NX - 1))

    def field_bytes(self) -> int:
        return int(sum(a.nbytes for a in self.fields.values()))


def _time_array(nsnap_needed: int, snapshot_dt: float) -> np.ndarray:
    return np.arange(nsnap_needed, dtype=np.float64) * snapshot_dt


def _positions(rng, n, NZ, NY, NX):
    z = rng.uniform(0.0, NZ - 1, n)
    y = rng.uniform(0.0, NY - 1, n)
    x = rng.uniform(0.0, NX - 1, n)
    return (z, y, x)


def eddy_workload(*, NX=128, NY=96, NZ=20, nparticles=10_000, steps=48, dt=300.0, seed=0,
                  dtype=np.float64, snapshot_dt=3600.0) -> Workload:
    """Seamount + Rankine eddy on a small grid (the CPU-runnable RK2 case)."""
    rng = np.random.default_rng(seed)
    dxm = dym = 1000.0
    yc, xc = (NY - 1) / 2.0, (NX - 1) / 2.0
    jj, ii = np.meshgrid(np.arange(NY, dtype=np.float64), np.arange(NX, dtype=np.float64), indexing="ij")
    r2 = (jj - yc) ** 2 + (ii - xc) ** 2
    h = 20.0 + 180.0 * np.exp(-r2 / (2 * 15.0**2))
    C = stretching_cs_w(NZ)
    sigma_r = (np.arange(NZ) + 0.5) / NZ - 1.0
    shear = (1.0 + sigma_r) / 2.0 + 0.5
    R = 20.0  # cells (20 km)

    def speed(r):
        return np.where(r < R, 0.5 * r / R, 0.5 * R / np.maximum(r, 1e-12))

    # u at (j, i + 1/2); v at (j + 1/2, i)
    ju, iu = np.meshgrid(np.arange(NY, dtype=np.float64), np.arange(NX - 1, dtype=np.float64) + 0.5, indexing="ij")
    ru = np.sqrt((ju - yc) ** 2 + (iu - xc) ** 2)
    u2 = -speed(ru) * (ju - yc) / np.maximum(ru, 1e-12)
    jv, iv = np.meshgrid(np.arange(NY - 1, dtype=np.float64) + 0.5, np.arange(NX, dtype=np.float64), indexing="ij")
    rv = np.sqrt((jv - yc) ** 2 + (iv - xc) ** 2)
    v2 = speed(rv) * (iv - xc) / np.maximum(rv, 1e-12)
    u = shear[:, None, None] * u2[None]
    v = shear[:, None, None] * v2[None]
    w = _w_from_continuity(u, v, h, C, dxm, dym).astype(np.float64)
    zeta = 0.05 * np.exp(-r2 / (2 * 20.0**2))
    amp = np.array([1.0, 0.8])
    nsnap = int(math.ceil(steps * dt / snapshot_dt)) + 3
    fields = {
        "u": (amp[:, None, None, None] * u[None]).astype(dtype),
        "v": (amp[:, None, None, None] * v[None]).astype(dtype),
        "w": (amp[:, None, None, None] * w[None]).astype(dtype),
        "zeta": (amp[:, None, None] * zeta[None]).astype(dtype),
        "h": h.astype(dtype),
        "dx": np.full((NY, NX), dxm, dtype),
        "dy": np.full((NY, NX), dym, dtype),
        "C": C.astype(np.float64),
    }
    return Workload("eddy", NX, NY, NZ, HC, float(dt), int(steps), float(snapshot_dt), fields,
                    _time_array(nsnap, snapshot_dt), _positions(rng, nparticles, NZ, NY, NX), seed)


def turbulent_workload(*, NX=1024, NY=768, NZ=50, nparticles=4_000_000, steps=1000, dt=120.0,
                       nsnap_unique=8, seed=1, dtype=np.float32, snapshot_dt=3600.0,
                       land_fraction=0.02, urms=0.3) -> Workload:
    """Random smooth bathymetry + divergence-free k^-3 flow (the at-scale AB3 case)."""
    rng = np.random.default_rng(seed)
    dxm = dym = 1000.0
    ky = np.fft.fftfreq(NY)[:, None]
    kx = np.fft.fftfreq(NX)[None, :]
    k = np.sqrt(kx**2 + ky**2)
    k0 = 1.0 / 64.0

    # bathymetry: k^-2 amplitude spectrum, rescaled to [50, 4000] m
    cb = (rng.standard_normal((NY, NX)) + 1j * rng.standard_normal((NY, NX))) * (k + 1.0 / 256) ** -2.0
    cb *= np.exp(-((k * 192.0) ** 2))
    cb[0, 0] = 0.0
    hb = np.real(np.fft.ifft2(cb))
    hb = (hb - hb.min()) / (hb.max() - hb.min())
    h = 50.0 + (4000.0 - 50.0) * hb

    land = h <= np.quantile(h, land_fraction) if land_fraction > 0 else np.zeros_like(h, bool)
    # streamfunction taper at psi (corner) points: 0 at any corner touching a
    # land cell, ramping to 1 over 32 cells, so u and v vanish on land faces
    # while staying exactly divergence free
    touch = land | np.roll(land, -1, 0) | np.roll(land, -1, 1) | np.roll(np.roll(land, -1, 0), -1, 1)
    # corners on and next to the domain edge (row/column -1 wraps to the last
    # one): no flow through the boundary
    touch[-1, :] = touch[0, :] = True
    touch[:, -1] = touch[:, 0] = True
    from scipy.ndimage import distance_transform_edt

    taper = np.clip(distance_transform_edt(~touch) / 32.0, 0.0, 1.0)
    # and over the first 300 m of depth above the coastline depth, which keeps
    # the shelf currents (transport / depth) bounded
    h_land = float(h[land].max()) if land.any() else float(h.min())
    h_corner = 0.25 * (h + np.roll(h, -1, 0) + np.roll(h, -1, 1) + np.roll(np.roll(h, -1, 0), -1, 1))
    taper = taper * np.clip((h_corner - h_land) / 300.0, 0.0, 1.0)
    C = stretching_cs_w(NZ)

    # streamfunction modes: k^-3 amplitude with roll-off, random phase and an
    # integer frequency m in [-2, 2] so snapshot n carries phase 2 pi m n / K
    amp = (k**2 + k0**2) ** -1.5
    amp[0, 0] = 0.0
    phase = rng.uniform(0, 2 * np.pi, (NY, NX))
    freq = rng.integers(-2, 3, (NY, NX))
    coef = amp * np.exp(1j * phase)
    # evaluate at psi (corner) points: shift by half a cell in x and y
    shift = np.exp(2j * np.pi * (kx * 0.5 + ky * 0.5))
    sigma_r = (np.arange(NZ) + 0.5) / NZ - 1.0
    prof = ((1.0 + sigma_r) / 2.0 + 0.5).astype(np.float32)
    # depth at corners and the profile-weighted depth at u and v faces
    h_c = 0.25 * (h + np.roll(h, -1, 0) + np.roll(h, -1, 1) + np.roll(np.roll(h, -1, 0), -1, 1))
    Hzf = layer_thickness(h, C)
    Pz = np.tensordot(prof.astype(np.float64), Hzf, axes=(0, 0))  # sum_k prof_k Hz_k at rho points
    Du = 0.5 * (Pz[:, 1:] + Pz[:, :-1])
    Dv = 0.5 * (Pz[1:, :] + Pz[:-1, :])
    del Hzf

    us, vs, ws, zs = [], [], [], []
    scale = None
    geom = _ContinuityGeometry(h, C, dxm, dym)
    for n in range(nsnap_unique):
        c = coef * np.exp(2j * np.pi * freq * n / nsnap_unique)
        # transport streamfunction at psi (corner) points (j + 1/2, i + 1/2)
        psi_c = np.real(np.fft.ifft2(c * shift)) * taper * h_c
        psi_r = np.real(np.fft.ifft2(c))
        # depth-integrated transport, exactly divergence free on the C-grid:
        # U(j, i+1/2) = -(Psi(j+1/2, i+1/2) - Psi(j-1/2, i+1/2)) / dy
        Ub = -(psi_c - np.roll(psi_c, 1, axis=0))[:, : NX - 1] / dym
        # V(j+1/2, i) = (Psi(j+1/2, i+1/2) - Psi(j+1/2, i-1/2)) / dx
        Vb = (psi_c - np.roll(psi_c, 1, axis=1))[: NY - 1, :] / dxm
        # vertical structure prof(sigma), normalised so sum_k u_k Hz_k = U
        u2 = Ub / Du
        v2 = Vb / Dv
        if scale is None:
            rms = np.sqrt(0.5 * (np.mean(u2**2) + np.mean(v2**2)))
            scale = urms / rms
        u = prof[:, None, None] * (u2 * scale).astype(np.float32)[None]
        v = prof[:, None, None] * (v2 * scale).astype(np.float32)[None]
        zeta = (psi_r * scale * 1e-4 / 9.81).astype(np.float64)
        zeta[land] = 0.0
        us.append(u)
        vs.append(v)
        zs.append(zeta.astype(dtype))
        del u, v

    # w of each snapshot from continuity, on host threads (numpy's loops run
    # without the GIL; each snapshot's arithmetic is unchanged)
    def w_of(n):
        w = _w_from_continuity(us[n], vs[n], h, C, dxm, dym, geom=geom)
        w[:, land] = 0.0
        return w.astype(dtype)

    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max(1, min(4, nsnap_unique, os.cpu_count() or 1))) as pool:
        ws = list(pool.map(w_of, range(nsnap_unique)))
    del geom
    us = [u.astype(dtype) for u in us]
    vs = [v.astype(dtype) for v in vs]
    nsnap = int(math.ceil(steps * dt / snapshot_dt)) + 3
    fields = {
        "u": np.stack(us),
        "v": np.stack(vs),
        "w": np.stack(ws),
        "zeta": np.stack(zs),
        "h": h.astype(dtype),
        "dx": np.full((NY, NX), dxm, dtype),
        "dy": np.full((NY, NX), dym, dtype),
        "C": C.astype(np.float64),
    }
    del us, vs, ws
    wl = Workload("turbulent", NX, NY, NZ, HC, float(dt), int(steps), float(snapshot_dt), fields,
                  _time_array(nsnap, snapshot_dt), _positions(rng, nparticles, NZ, NY, NX), seed)
    wl.meta["land_fraction"] = float(land.mean())
    return wl


def add_density(wl: Workload, *, rho0: float = 1026.0, drho: float = 2.0, amp: float = 0.05, seed: int = 11):
    """Add a time-dependent density field "rho" at rho points (NZ, NY, NX):
    stably stratified (denser with depth, ``drho`` over the column) plus a
    smooth horizontal perturbation that drifts between snapshots."""
    rng = np.random.default_rng(seed)
    K = wl.fields["u"].shape[0]
    NZ, NY, NX = wl.NZ, wl.NY, wl.NX
    zr = (np.arange(NZ) + 0.5) / NZ  # 0 at the bottom, 1 at the top
    y = np.arange(NY)[:, None]
    x = np.arange(NX)[None, :]
    ph = rng.uniform(0, 2 * np.pi, 2)
    snaps = []
    for n in range(K):
        pert = amp * np.sin(2 * np.pi * x / 37.0 + ph[0] + 2 * np.pi * n / K) * np.cos(2 * np.pi * y / 29.0 + ph[1])
        snaps.append(rho0 + drho * (1.0 - zr)[:, None, None] + pert[None] * np.sin(np.pi * zr)[:, None, None])
    wl.fields["rho"] = np.stack(snaps).astype(wl.fields["u"].dtype)
    return wl


def build_fieldset(wl: Workload, *, device=None, storage_dtype=None, cyclic: bool = True, constants=None):
    """A Fieldset on the device holding the workload's fields.

    ``storage_dtype`` (e.g. torch.bfloat16) overrides the storage type of
    u, v and w; the 2-D fields and C keep the workload's dtype.
    """
    from .fields import StaticField, TimeDependentField
    from .fieldset import Fieldset

    fields = {}
    for name, arr in wl.fields.items():
        zs, ys, xs = STAGGERS[name]
        if name in TIME_DEPENDENT:
            sd = storage_dtype if name in ("u", "v", "w") else None
            fields[name] = TimeDependentField(arr, zs, ys, xs, cyclic=cyclic, device=device, storage_dtype=sd)
        else:
            fields[name] = StaticField.from_numpy(arr, zs, ys, xs, device=device)
    consts = {"hc": wl.hc, "NZ": wl.NZ}
    consts.update(constants or {})
    return Fieldset(len(wl.time_array), wl.NZ, wl.NY, wl.NX, constants=consts, **fields)


def build_simulation(wl: Workload, scheme: str, *, device=None, storage_dtype=None, positions=None,
                     status=None, ab_kernel=None, spatial_order_every: int = 0, fieldset=None):
    """A drop-in Simulation for the workload with particles placed.

    ``scheme`` is "rk2", "ab3" or "ab3_rk2" (AB3 whose first two steps are RK2
    steps); AB3 particles start from MULTISTEP_1 (Euler -> AB2 -> AB3) unless
    ``status`` is given, RK2 and "ab3_rk2" ones from NORMAL.
    ``spatial_order_every`` > 0 keeps the particle arrays in the Morton order
    of the particles' cells, refreshed every that many steps (results and what
    the user sees are unchanged).  ``ab_kernel`` replaces the AB3
    kernel (same declarations), e.g. to time it.  ``fieldset`` replaces the
    one built from the workload (e.g. one of chunked fields).
    """
    from .kernels.roms import ab3_post_w_advection_kernel, ab3_w_advection_kernel
    from .kernels.validation import validation_kernel
    from .models import roms
    from .simulation import SimulationBuilder
    from .timesteppers import ABTimestepper

    pos = wl.positions if positions is None else positions
    fs = fieldset if fieldset is not None else build_fieldset(wl, device=device, storage_dtype=storage_dtype)
    dt = np.float64(wl.dt)
    if scheme == "rk2":
        ts = roms.rk2_w_advection_timestepper(wl.time_array, dt)
    elif scheme == "ab3_rk2":
        ts = roms.ab3_w_advection_timestepper(wl.time_array, dt, startup="rk2")
    elif scheme == "ab3":
        ts = ABTimestepper(wl.time_array, dt, ab_kernel=ab_kernel or ab3_w_advection_kernel, index_padding=5,
                           pre_step_kernel=validation_kernel, post_step_kernel=ab3_post_w_advection_kernel)
    else:
        raise ValueError(scheme)
    sim = SimulationBuilder(ts, fs).build_simulation(len(pos[0]), device=device,
                                                     spatial_order_every=spatial_order_every)
    sim.set_indices(*pos)
    if status is not None:
        sim.set_particle_field("status", status)
    elif scheme == "ab3":
        sim.set_particle_field("status", 10)
    return sim


@dataclass
class Release:
    """One particle release: seed positions, start time (s) and number of steps."""

    positions: tuple
    start_time: float
    nsteps: int

    @property
    def size(self) -> int:
        return int(self.positions[0].shape[0])


def batched_releases(wl: Workload, *, nrel=256, size_range=(1_000, 1_000_000), start_span=24 * 3600.0,
                     duration_range=(6 * 3600.0, 72 * 3600.0), box_cells=10, seed=3) -> list:
    """Releases of log-uniform size, uniform start time and duration, seeded in random boxes.

    Each release fills a random ``box_cells`` x ``box_cells`` horizontal box
    (10 km at 1 km spacing) over the full depth, uniformly in index space.
    """
    rng = np.random.default_rng(seed)
    out = []
    lo, hi = np.log(size_range[0]), np.log(size_range[1])
    for _ in range(nrel):
        m = int(np.exp(rng.uniform(lo, hi)))
        y0 = rng.uniform(0, wl.NY - 1 - box_cells)
        x0 = rng.uniform(0, wl.NX - 1 - box_cells)
        pos = (rng.uniform(0, wl.NZ - 1, m), y0 + rng.uniform(0, box_cells, m), x0 + rng.uniform(0, box_cells, m))
        t0 = float(rng.uniform(0, start_span))
        nsteps = int(round(rng.uniform(*duration_range) / wl.dt))
        out.append(Release(pos, t0, nsteps))
    return out
