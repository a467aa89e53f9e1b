"""Fused advection engine: the fast form of the hot path.

``AB3Engine`` advances particles with fused HIP kernels that take one or
several (``launch_steps``) whole steps per launch: AB3 (``opx_ab3_fused_steps``:
validation -> ab3_w_advection -> ab3_post_w_advection, the reference's
timesteppers.py:296-305 sequence), with ``scheme="rk2"`` RK2
(``opx_rk2_fused_steps``: validation -> _rk2_step_1 -> _rk2_step_2 ->
_rk2_update, timesteppers.py:241-257), or with ``scheme="ab3_iso"`` the
isopycnal-following AB3 (``opx_ab3_iso_fused_steps``: validation ->
ab3_isopycnal_following -> its post step, models/roms/__init__.py:109-140),
which reads the fieldset's density field "rho" in place of w, the constants
``w_max`` and ``rho_scale`` from the fieldset and each particle's target
density ``rho_t`` from ``add_release``.  Between the steps of a launch each
particle's state stays in registers and the stencil nodes it read are blended
again while its cells are unchanged; the results are the same bits as one step
per launch.  The engine keeps

- split positions (packed int32 cell + 32-bit fixed-point offsets, exact in
  FP64) and FP32 AB3 history;
- u, v, w snapshot rings in FP32 or BF16 (zeta FP32; h, dx, dy packed as one
  FP32 float4 per rho point; C in FP64) taken from a ``Fieldset`` built for the
  drop-in API;
- per-release FP64 clocks on the device (``opx_clock_advance`` restates
  Timestepper.get_time_index / advance_time, timesteppers.py:64-85, 157-161);
- Morton re-binning every ``rebin_every`` steps (``opx_rebin``), which also
  moves exited particles behind the active ones so they cost nothing.

Releases: particles are added in groups (``add_release``), each with its
own start time, start step and number of steps; one launch steps every active
release.  Results are read back in the order particles were added, or per
release (``extract``).  A finished release can be retired: the next re-bin
drops its particles and recycles its id, so a long-running service reuses
the same buffers (``offline_particles_amd.serving``).

Snapshot windows: by default every snapshot of u, v, w, zeta is resident.
With ``stream_slots`` the snapshots stay in pinned host memory and a ring of
that many device slots is filled on demand (a side stream copies the next
snapshot while the current ones are in use).  With ``stream_tiles`` as well,
a slot receives only the horizontal tiles (``stream_tiles`` rho columns
square, every level) that the particles can reach while it is read: the
reference's lazy chunked field loads only the chunks inside the particles'
padded bounding box (``ChunkedDaskArray.get_data_subset``,
spatial_arrays.py:260-295, with ``Launcher.construct_bbox``,
launcher.py:122-148); here the region is the set of tiles holding particles,
widened by a bound on how far they move, so separate clouds of particles do
not pull in the ocean between them.  The bound: m steps read the fields
within dt x R x (m + 4) cells of where a particle was before them for AB3
and within dt x R x (m s + |alpha| + s), s = |b1| + |b2|, for RK2, R the
largest |u| / dx or |v| / dy of any snapshot loaded so far (the
interpolation weights of a particle inside the grid are convex, so every
interpolated rate is at most R).  For AB3 the m steps' increments
dt (23 f[n] - 16 f[n-1] + 5 f[n-2]) / 12 sum to dt times a combination of the
rates in which each rate inside the run has weight 1 (23 - 16 + 5 = 12) and
the run's ends weights 23/12, 7/12, 11/12 and 5/12, so the total weight is
m + 22/12 for m >= 2 and 44/12 for one step (the Euler and AB2 start-up steps
weigh less); RK2 moves at most s dt R per step and its stage reads alpha dt R
further.  A tile map of the particles is computed
on the device after every run of steps and read back asynchronously; before
each run of steps every slot it reads must hold the tiles of the latest map
read back (and of releases added since), widened by the bound times the
steps since that map plus two cells of stencil reach.  Missing tiles are
copied (rectangular DMA copies straight into the slot) and the packed windows'
rows they touch are packed again.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib
from .fields import StaticField, TimeDependentField
from .fields import unpacked as fields_unpacked
from .kernels.status import ParticleStatus
from .snapshots import SnapshotSource
from .spatial_arrays import Stagger

_ROMS_STAGGERS = {
    "u": (Stagger.CENTER, Stagger.CENTER, Stagger.INNER),
    "v": (Stagger.CENTER, Stagger.INNER, Stagger.CENTER),
    "w": (Stagger.OUTER, Stagger.CENTER, Stagger.CENTER),
    "zeta": (Stagger.INVARIANT, Stagger.CENTER, Stagger.CENTER),
    "h": (Stagger.INVARIANT, Stagger.CENTER, Stagger.CENTER),
    "dx": (Stagger.INVARIANT, Stagger.CENTER, Stagger.CENTER),
    "dy": (Stagger.INVARIANT, Stagger.CENTER, Stagger.CENTER),
    "C": (Stagger.OUTER, Stagger.INVARIANT, Stagger.INVARIANT),
}

_STATE_F32 = ("fz", "fy", "fx", "hz1", "hy1", "hx1", "hz2", "hy2", "hx2")

# steps one fused launch may take (the release table holds one row per step)
MAX_LAUNCH_STEPS = 128

LAYOUT_NAMES = {_lib.OPX_LAYOUT_RING: "ring", _lib.OPX_LAYOUT_PAIR: "pair", _lib.OPX_LAYOUT_QUAD: "quad",
                _lib.OPX_LAYOUT_PAIR_LDS: "pair_lds", _lib.OPX_LAYOUT_OCT: "oct", _lib.OPX_LAYOUT_YPAIR: "ypair"}


def dilate_tiles(mask: np.ndarray, d: int) -> np.ndarray:
    """The tiles within d tiles (in y and in x) of a set tile of the 2-D bool `mask`."""
    mask = np.asarray(mask, dtype=bool)
    if d <= 0:
        return mask.copy()
    rows = mask.copy()
    for s in range(1, min(d, mask.shape[0] - 1) + 1):
        rows[s:] |= mask[:-s]
        rows[:-s] |= mask[s:]
    out = rows.copy()
    for s in range(1, min(d, mask.shape[1] - 1) + 1):
        out[:, s:] |= rows[:, :-s]
        out[:, :-s] |= rows[:, s:]
    return out


def run_reach(scheme: str, alpha: float = 2 / 3):
    """(s, c): m steps of `scheme` read the fields within dt x R x (m s + c) cells
    (horizontally) of a particle's position before them, R the largest |u| / dx,
    |v| / dy (the module docstring derives the constants)."""
    if scheme == "rk2":
        b1 = 1.0 / (2.0 * alpha)
        s = abs(b1) + abs(1.0 - b1)
        return s, abs(alpha) + s
    return 1.0, 4.0


def tile_boxes(mask: np.ndarray, tile: int, ny: int, nx: int, gap: int = 0):
    """Boxes covering the set tiles of `mask` (tile rows x tile columns of `tile`
    cells over an ny x nx grid) for rectangular copies.

    In each tile row, runs of set tiles separated by at most `gap` unset tiles
    form one run (longer rows copy faster); consecutive tile rows with the same
    run are stacked into one box.  Returns (boxes, covered): int32 (k, 4) boxes
    (y0, y1, x0, x1) in cells, clipped to the grid, in ascending order, and the
    bool mask of the tiles they cover (the set tiles and the bridged gaps).

    Array operations throughout (no loop over rows or runs): the streaming
    engine calls this for every snapshot load, top-up and window pack."""
    mask = np.asarray(mask, dtype=bool)
    nty, ntx = mask.shape
    edge = np.zeros((nty, 1), dtype=np.int8)
    d = np.diff(np.concatenate([edge, mask.view(np.int8), edge], axis=1), axis=1).ravel()
    # runs of set tiles, row-major (each row's in order), and their ends (exclusive)
    ty, a = np.divmod(np.flatnonzero(d == 1), ntx + 1)
    b = np.flatnonzero(d == -1) % (ntx + 1)
    if not ty.size:
        return np.zeros((0, 4), dtype=np.int32), mask.copy()
    covered = mask
    if gap > 0:
        # a run joins the one before it in its row when at most `gap` unset tiles
        # lie between them
        join = np.zeros(ty.size, dtype=bool)
        join[1:] = (ty[1:] == ty[:-1]) & (a[1:] - b[:-1] <= gap)
        if join.any():
            keep = ~join
            end = np.ones(ty.size, dtype=bool)
            end[:-1] = keep[1:]
            ty, a, b = ty[keep], a[keep], b[end]
            marks = np.zeros((nty, ntx + 1), dtype=np.int8)
            marks[ty, a] = 1
            marks[ty, b] = -1
            covered = np.cumsum(marks, axis=1, dtype=np.int8)[:, :ntx] > 0
    covered = covered.copy() if covered is mask else covered
    # a box is a run repeated in consecutive rows: group equal runs by row
    key = a.astype(np.int64) * (ntx + 1) + b
    o = np.lexsort((ty, key))
    ks, ts = key[o], ty[o]
    first = np.ones(ks.size, dtype=bool)
    first[1:] = (ks[1:] != ks[:-1]) | (ts[1:] != ts[:-1] + 1)
    last = np.ones(ks.size, dtype=bool)
    last[:-1] = first[1:]
    y0 = ts[first].astype(np.int64) * tile
    y1 = np.minimum((ts[last].astype(np.int64) + 1) * tile, ny)
    x0 = a[o][first].astype(np.int64) * tile
    x1 = np.minimum(b[o][first].astype(np.int64) * tile, nx)
    out = np.stack([y0, y1, x0, x1], axis=1)
    return out[np.lexsort((x1, x0, y1, y0))].astype(np.int32), covered


def _bits(n: int) -> int:
    return max(1, int(math.ceil(math.log2(max(n, 2)))))


def _ptr(t):
    return t.data_ptr() if t is not None else None


def _h2d(a, dtype, device):
    """Stream-ordered host->device copy through pinned memory (no host synchronisation)."""
    if isinstance(a, torch.Tensor):
        if a.device.type != "cpu":
            return a.to(device=device, dtype=dtype)
        a = a.numpy()
    host = torch.from_numpy(np.array(a, dtype=_NP[dtype])).pin_memory()
    return host.to(device, non_blocking=True)


_NP = {torch.float64: np.float64, torch.uint8: np.uint8}


class _PinnedStaging:
    """Pinned host buffers for host->device copies of release arrays, reused
    once the copy that last read them has finished (an event per buffer): the
    arrays are written once, straight into pinned memory, and the steady state
    allocates no pinned memory (an allocation of it can wait for the device)."""

    def __init__(self) -> None:
        self._bufs: list = []  # [pinned uint8 tensor, event of its last copy or None]

    def to_device(self, arrays, dtype, device) -> list:
        """The host arrays as device tensors of `dtype` (views of one copy, stream-ordered)."""
        npdt = np.dtype(_NP[dtype])
        flat = [np.asarray(a).reshape(-1) for a in arrays]
        sizes = [a.size for a in flat]
        nbytes = sum(sizes) * npdt.itemsize
        slot = next((b for b in self._bufs if b[0].numel() >= nbytes and (b[1] is None or b[1].query())), None)
        if slot is None:
            slot = [torch.empty(max(nbytes, 1 << 20), dtype=torch.uint8, pin_memory=True), None]
            self._bufs.append(slot)
        host = slot[0][:nbytes].view(dtype)
        hn = host.numpy()
        off = 0
        for a, n in zip(flat, sizes):
            np.copyto(hn[off:off + n], a, casting="unsafe")
            off += n
        dev = host.to(device, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        slot[1] = ev
        out, off = [], 0
        for n in sizes:
            out.append(dev[off:off + n])
            off += n
        return out


def _to_host(a) -> np.ndarray:
    """A numpy view or copy of an array or tensor (device tensors are copied back)."""
    if torch.is_tensor(a):
        return a.detach().to("cpu", torch.float64).numpy()
    return np.asarray(a, dtype=np.float64)


class SplitState:
    """Device SoA of the split-position state (opx_split_state)."""

    # status 1 + cell 4 + offsets 3 x 4 + FP32 history 6 x 4 + pid 4 + rel 2 + z 8
    BYTES_PER_PARTICLE = 55

    def __init__(self, capacity: int, device, *, with_rho_t: bool = False):
        n = int(capacity)
        self.capacity = n
        self.status = torch.full((n,), ParticleStatus.INACTIVE, dtype=torch.uint8, device=device)
        self.cell = torch.zeros(n, dtype=torch.int32, device=device)
        for name in _STATE_F32:
            # fz, fy, fx hold 32-bit fixed-point offsets (uint32 bits in int32 tensors)
            dt = torch.int32 if name in ("fz", "fy", "fx") else torch.float32
            setattr(self, name, torch.zeros(n, dtype=dt, device=device))
        self.pid = torch.zeros(n, dtype=torch.int32, device=device)
        self.rel = torch.zeros(n, dtype=torch.int16, device=device)  # read as uint16
        # z (m) after each particle's last step, the next step's start (NaN: from zidx)
        self.z = torch.full((n,), float("nan"), dtype=torch.float64, device=device)
        # target density of the isopycnal-following step (FP64, 8 B per particle)
        self.rho_t = torch.zeros(n, dtype=torch.float64, device=device) if with_rho_t else None

    def struct(self, n: int, start: int = 0) -> _lib.opx_split_state:
        s = _lib.opx_split_state()
        s.n = int(n)

        def p(t, size):
            return t.data_ptr() + start * size if t is not None else None

        s.status = p(self.status, 1)
        s.cell = p(self.cell, 4)
        for name in _STATE_F32:
            setattr(s, name, p(getattr(self, name), 4))
        s.pid = p(self.pid, 4)
        s.rel = p(self.rel, 2)
        s.z = p(self.z, 8)
        s.rho_t = p(self.rho_t, 8)
        return s


class AB3Engine:
    """Fused AB3 stepping of one or many particle releases on one device."""

    def __init__(self, fieldset, time_array, dt, *, storage_dtype=torch.float32, device=None,
                 capacity: int = 1 << 20, rebin_every: int = 25, xlow: int = 2,
                 max_releases: int = 256, stream_slots: int | None = None, scheme: str = "ab3",
                 alpha: float = 2 / 3, stream_tiles: int | None = None):
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        if storage_dtype not in (torch.float32, torch.bfloat16):
            raise TypeError("storage_dtype must be torch.float32 or torch.bfloat16")
        self.storage_dtype = storage_dtype
        if scheme not in ("ab3", "rk2", "ab3_iso"):
            raise ValueError(f"scheme must be 'ab3', 'rk2' or 'ab3_iso', not {scheme!r}")
        self.scheme = scheme
        if scheme == "ab3_iso":
            # the isopycnal-following step (ab3_isopycnal_following.pyx) reads
            # its scalars from the fieldset constants, as the reference's
            # launcher does; the density needs FP32 storage (BF16 resolves
            # 1026 kg/m^3 only to 8)
            if storage_dtype != torch.float32:
                raise TypeError("the isopycnal-following step needs storage_dtype=torch.float32")
            for k in ("w_max", "rho_scale"):
                if k not in fieldset.constants:
                    raise KeyError(k)
            self.w_max = float(fieldset.constants["w_max"])
            self.rho_scale = float(fieldset.constants["rho_scale"])
        self.alpha = float(alpha)
        if scheme == "rk2" and not self.alpha > 0:
            raise ValueError("RK2 needs alpha > 0")
        self.dt = float(np.float64(dt))
        self.time_array = np.asarray(time_array, dtype=np.float64)
        if np.any(self.time_array[1:] <= self.time_array[:-1]):
            raise ValueError("time_array must be strictly increasing.")
        self._ta = torch.from_numpy(self.time_array.copy()).to(self.device)
        self.rebin_every = int(rebin_every)
        self.xlow = int(xlow)
        self._stream_slots = None if stream_slots is None else int(stream_slots)
        if self._stream_slots is not None and self._stream_slots < 3:
            raise ValueError("a streamed window needs at least 3 slots")
        self._tile = None if stream_tiles is None else int(stream_tiles)
        if self._tile is not None:
            if self._stream_slots is None:
                raise ValueError("stream_tiles needs a streamed window (stream_slots)")
            if self._tile < 4:
                raise ValueError(f"stream_tiles must be at least 4 cells, not {self._tile}")
        self._build_window(fieldset)
        self._with_rho_t = scheme == "ab3_iso"
        self._state = self._new_state(capacity)
        self._spare = None
        self.n = 0
        # device-resident values the step reads, so captured graphs stay valid:
        # particles to launch, release count, global step, window descriptor
        self._n_active_dev = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._nrel_dev = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._step_dev = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._win_dev = torch.zeros(ctypes.sizeof(_lib.opx_window), dtype=torch.uint8, device=self.device)
        self._win_dirty = True
        self._err = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._clocks_host: list = []
        self._rel_cap = 0
        self._clocks = None
        self._table = None
        self._grow_releases(max(1, int(max_releases)))
        self._release_ranges: list = []
        # release ids: pid range (base, size) per id, ids free for reuse, and
        # retired releases whose particles the next re-bin drops
        self._release_pids: dict = {}
        self._free_ids: list = []
        self._retired: list = []
        self._retired_particles = 0
        self._next_pid = 0
        self._staging = _PinnedStaging()
        self._compacted = False
        self.global_step = 0
        self._since_rebin = 0
        self._work = None
        self._structs = None
        self._graphs = {}
        self._capture_stream = None
        self.graph_steps = 0  # > 0: replay captured graphs of this many steps
        # steps per fused launch (1..MAX_LAUNCH_STEPS): the state stays in
        # registers between them and unchanged stencils are not read again
        self.launch_steps = 1
        # optional list collecting (start, end, steps) around each fused launch
        # (steps 0: an output launch of sample())
        self.timer = None
        self.block = 128  # workgroup size of the fused kernel
        self.tuning = None
        # packed stencil windows (PAIR for FP32, QUAD for BF16 by default)
        self.layout = _lib.OPX_LAYOUT_PAIR if storage_dtype == torch.float32 else _lib.OPX_LAYOUT_QUAD
        # the pair ring: one packed window per ring slot s (slots s, s + 1),
        # with the snapshot pair each window was last packed from
        self._pair_bufs = None
        self._pair_strides = None
        self._pair_key = None
        self._packed_bytes = {}
        # host mirror of the device clocks (same FP64 arithmetic), used to
        # know which snapshot pairs the next step reads
        self._t = np.zeros(0)
        self._start = np.zeros(0, np.int64)
        self._left = np.zeros(0, np.int64)

    # ---------------------------------------------------------------- setup

    def _build_window(self, fs):
        # the third 3-D field: w at w-points, or for the isopycnal-following
        # step the density at rho points (kept in the same ring)
        third = "rho" if self.scheme == "ab3_iso" else "w"
        staggers = dict(_ROMS_STAGGERS)
        if third == "rho":
            del staggers["w"]
            staggers["rho"] = (Stagger.CENTER, Stagger.CENTER, Stagger.CENTER)
        for name, st in staggers.items():
            f = fs[name]
            if tuple(f.stagger) != st:
                raise ValueError(f"field '{name}' has staggering {tuple(map(str, f.stagger))}; the fused step "
                                 f"needs the ROMS C-grid {tuple(map(str, st))}")
        for name in ("u", "v", third, "zeta"):
            if not isinstance(fs[name], TimeDependentField):
                raise TypeError(f"field '{name}' must be time dependent")
        for name in ("h", "dx", "dy", "C"):
            if not isinstance(fs[name], StaticField):
                raise TypeError(f"field '{name}' must be static")
        NZ, NY, NX = fs.z_size, fs.y_size, fs.x_size
        self.NZ, self.NY, self.NX = NZ, NY, NX
        # the split state stores cells -2 .. N+1 per axis and clamps positions
        # beyond them; the domain check keeps flagging those only while every
        # bound lies in (-2, N+1]
        for axis, (lo, hi), n in (("zidx", fs.zidx_bounds, NZ), ("yidx", fs.yidx_bounds, NY),
                                  ("xidx", fs.xidx_bounds, NX)):
            if not all(-2.0 < float(b) <= n + 1 for b in (lo, hi)):
                raise ValueError(f"{axis} bounds ({float(lo)}, {float(hi)}) lie outside (-2, {n + 1}], the index "
                                 f"range the fused engine's split positions hold")
        flds ={n: fs[third if n == "w" else n] for n in ("u", "v", "w", "zeta")}
        lazy = [n for n, f in flds.items() if f.source is not None]
        if lazy and self._stream_slots is None:
            raise ValueError(f"fields {lazy} are read lazily from their source: give stream_slots to stream them")
        K = {int(f.num_snapshots) for f in flds.values()}
        if len(K) != 1:
            raise ValueError(f"u, v, w, zeta hold different snapshot counts {K}")
        self.nsnap = K.pop()
        cyc = {fs[n].cyclic for n in ("u", "v", third, "zeta")}
        if cyc == {False} and self.nsnap != len(self.time_array):
            raise ValueError("non-cyclic fields need one snapshot per time_array entry")
        self.nslots = self.nsnap
        dts = {"u": self.storage_dtype, "v": self.storage_dtype, "w": self.storage_dtype, "zeta": torch.float32}
        self._dts = dts
        if self._stream_slots is None:
            self._u, self._v, self._w, self._zeta = (flds[n].data.to(self.device, dts[n]).contiguous()
                                                     for n in ("u", "v", "w", "zeta"))
            self._src = None
        else:
            # host source (pinned, storage dtype, or a SnapshotSource decoded on
            # demand) and a ring of device slots; snapshot It lives in slot
            # It % nslots while it is needed
            self.nslots = self._stream_slots
            self._src = {n: (flds[n].source if flds[n].source is not None
                             else flds[n].data.to("cpu", dts[n]).contiguous().pin_memory())
                         for n in ("u", "v", "w", "zeta")}
            self._u, self._v, self._w, self._zeta = (
                torch.empty((self.nslots,) + tuple(flds[n].spatial_shape), dtype=dts[n], device=self.device)
                for n in ("u", "v", "w", "zeta"))
            self._ring_It = [None] * self.nslots
            self._ring_ready = [None] * self.nslots
            self._ring_shape = {n: tuple(flds[n].spatial_shape) for n in ("u", "v", "w", "zeta")}
            self._staged = {}
            self._copy_streams = [torch.cuda.Stream(self.device) for _ in range(3)]
            self.snapshot_loads = 0
            self.snapshot_bytes = 0
            # bytes whole-snapshot loads would have copied (for the sparse mode's ratio)
            self.snapshot_full_bytes = 0
            if self._tile is not None:
                self._init_tiles(fs, NY, NX)
        h = fs["h"].data.data.to(self.device, torch.float32)
        dx = fs["dx"].data.data.to(self.device, torch.float32)
        dy = fs["dy"].data.data.to(self.device, torch.float32)
        self._hdd = torch.stack([h, dx, dy, torch.zeros_like(h)], dim=-1).contiguous()
        self._dy = dy.contiguous()
        self._C = fs["C"].data.data.to(self.device, torch.float64).contiguous()
        hc = fs.constants["hc"]
        g = _lib.opx_grid()
        g.NZ, g.NY, g.NX = NZ, NY, NX
        g.xbits, g.ybits, g.zbits = _bits(NX + 5), _bits(NY + 5), _bits(NZ + 5)
        g.hc = float(hc)
        g.zidx_min, g.zidx_max = (float(b) for b in fs.zidx_bounds)
        g.yidx_min, g.yidx_max = (float(b) for b in fs.yidx_bounds)
        g.xidx_min, g.xidx_max = (float(b) for b in fs.xidx_bounds)
        self.grid = g
        w = _lib.opx_window()
        w.u, w.v, w.w, w.zeta = (_ptr(t) for t in (self._u, self._v, self._w, self._zeta))
        w.u_stride = self._u[0].numel()
        w.v_stride = self._v[0].numel()
        w.w_stride = self._w[0].numel()
        w.zeta_stride = self._zeta[0].numel()
        w.dtype3 = _lib.OPX_F32 if self.storage_dtype == torch.float32 else _lib.OPX_BF16
        w.nslots = self.nslots
        w.hdd = _ptr(self._hdd)
        w.C = _ptr(self._C)
        w.dy = _ptr(self._dy)
        w.wlevels = NZ if third == "rho" else NZ + 1
        self.window = w

    def _new_state(self, capacity):
        return SplitState(capacity, self.device, with_rho_t=self._with_rho_t)

    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def reserve(self, capacity: int) -> None:
        """Allocate the particle buffers, the re-bin target and the re-bin workspace
        for `capacity` particles up front (no allocation while stepping after this)."""
        self._ensure_capacity(int(capacity), exact=True)
        self._spare = self._new_state(self._state.capacity)
        self._alloc_work(self._state.capacity)
        self._graphs = {}

    def _alloc_work(self, n):
        need = ctypes.c_uint64()
        _lib.call("opx_rebin_workspace_bytes", int(n), ctypes.byref(need))
        self._work = torch.empty(int(need.value) + (1 << 20), dtype=torch.uint8, device=self.device)
        self._work_n = int(n)

    def packed_window_bytes(self, layout=None):
        """Bytes of one packed stencil window (u, v, w, 2-D) in `layout` (default: the current one).

        Raises OpxInvalidArgument where one window would exceed 4 GiB (the step
        kernels address a window with 32-bit byte offsets)."""
        key = self.layout if layout is None else int(layout)
        if key not in self._packed_bytes:
            bu, bv, bw, b2 = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
            _lib.call("opx_packed_bytes", ctypes.byref(self.grid), self.window.dtype3, key,
                      ctypes.byref(bu), ctypes.byref(bv), ctypes.byref(bw), ctypes.byref(b2))
            self._packed_bytes[key] = (bu.value, bv.value, bw.value, b2.value)
        return self._packed_bytes[key]

    def layout_fits(self, layout: int) -> bool:
        """Whether `layout` can be used on this grid: RING always, a packed layout
        when one window of it is at most 4 GiB."""
        if int(layout) == _lib.OPX_LAYOUT_RING:
            return True
        try:
            self.packed_window_bytes(layout)
        except _lib.OpxInvalidArgument:
            return False
        return True

    def _ensure_capacity(self, n, exact=False):
        if n <= self._state.capacity:
            return
        cap = n if exact else max(n, 2 * self._state.capacity)
        new = self._new_state(cap)
        old = self._state
        for name in (("status", "cell", "pid", "rel", "z") + _STATE_F32 + (("rho_t",) if self._with_rho_t else ())):
            getattr(new, name)[: self.n].copy_(getattr(old, name)[: self.n])
        self._state = new
        self._spare = None
        self._work = None
        self._graphs = {}

    def add_release(self, zidx, yidx, xidx, *, start_time=None, nsteps: int, start_step: int = 0,
                    status: int | None = None, rho_t=None) -> int:
        """Add a group of particles (a release); returns its id.

        The release's clock starts at ``start_time`` (default time_array[0])
        and it advances at global steps ``start_step`` .. while its
        ``nsteps`` steps last.  Particles start with ``status`` (default
        MULTISTEP_1 for AB3: Euler -> AB2 -> AB3 start-up; NORMAL for RK2).
        ``rho_t`` (a value or one per particle) is the target density of the
        isopycnal-following scheme, required there and refused otherwise.
        """
        if status is None:
            status = ParticleStatus.NORMAL if self.scheme == "rk2" else ParticleStatus.MULTISTEP_1
        if (rho_t is None) == self._with_rho_t:
            raise ValueError("rho_t is required by scheme 'ab3_iso'" if self._with_rho_t
                             else f"rho_t is only used by scheme 'ab3_iso', not {self.scheme!r}")
        r = self._free_ids.pop() if self._free_ids else len(self._clocks_host)
        if r >= 1 << 15:
            raise ValueError("too many releases")
        if any(torch.is_tensor(a) and a.device.type != "cpu" for a in (zidx, yidx, xidx)):
            z, y, x = (_h2d(a, torch.float64, self.device) for a in (zidx, yidx, xidx))
        else:
            z, y, x = self._staging.to_device((zidx, yidx, xidx), torch.float64, self.device)
        m = int(z.numel())
        if not (y.numel() == m and x.numel() == m):
            raise ValueError("zidx, yidx, xidx must have the same length")
        start = self.n
        self._ensure_capacity(start + m)
        S = self._state
        if m:
            _lib.call("opx_split_from_f64", ctypes.byref(S.struct(m, start)), ctypes.byref(self.grid),
                      z.data_ptr(), y.data_ptr(), x.data_ptr(), self._stream())
            S.status[start:start + m] = int(status)
            # pids count admissions (uint32, wrapping), stored as int32 bits
            base = self._next_pid
            pids = torch.arange(base, base + m, dtype=torch.int64, device=self.device)
            S.pid[start:start + m] = torch.where(pids >= 1 << 31, pids - (1 << 32), pids).to(torch.int32)
            S.rel[start:start + m] = r
            for name in _STATE_F32[3:]:
                getattr(S, name)[start:start + m] = 0.0
            if self._with_rho_t:
                rt = _h2d(rho_t if torch.is_tensor(rho_t) else np.asarray(rho_t, np.float64), torch.float64,
                          self.device).reshape(-1)
                if rt.numel() not in (1, m):
                    raise ValueError(f"rho_t has {rt.numel()} values for {m} particles")
                S.rho_t[start:start + m].copy_(rt.expand(m))
        self.n = start + m
        self._n_active_dev.fill_(self.n)
        self._release_pids[r] = (self._next_pid, m)
        self._next_pid = (self._next_pid + m) & 0xFFFFFFFF
        t0 = float(self.time_array[0]) if start_time is None else float(start_time)
        if r >= self._rel_cap:
            self._grow_releases(2 * self._rel_cap)
        if r == len(self._clocks_host):
            self._clocks_host.append((t0, int(start_step), int(nsteps)))
            self._t = np.append(self._t, t0)
            self._start = np.append(self._start, int(start_step))
            self._left = np.append(self._left, int(nsteps))
            self._release_ranges.append((start, m))
        else:
            self._clocks_host[r] = (t0, int(start_step), int(nsteps))
            self._t[r], self._start[r], self._left[r] = t0, int(start_step), int(nsteps)
            self._release_ranges[r] = (start, m)
        clk = np.frombuffer(bytes(_lib.opx_release_clock(t0, int(start_step), int(nsteps))), dtype=np.uint8)
        cs = ctypes.sizeof(_lib.opx_release_clock)
        self._clocks[r * cs:(r + 1) * cs].copy_(_h2d(clk, torch.uint8, self.device))
        self._nrel_dev.fill_(len(self._clocks_host))
        self._structs = None
        if self._tile is not None and m:
            # the release's tiles count until a map that includes it is read back
            self._occ_seq += 1
            self._occ_adds.append((self._occ_seq, self._tile_mask_of(_to_host(yidx), _to_host(xidx))))
        return r

    def steps_left(self, r: int) -> int:
        """Steps release r still has to run (host mirror of its device clock; -1 once retired)."""
        return int(self._left[r])

    def retire(self, r: int) -> None:
        """Give release r up: it stops stepping and the next re-bin drops its
        particles and frees its id.  Read its results (``extract``) first."""
        if r in self._retired or self._left[r] < 0:
            raise ValueError(f"release {r} is already retired")
        self._left[r] = -1
        cs = ctypes.sizeof(_lib.opx_release_clock)
        off = r * cs + _lib.opx_release_clock.steps_left.offset
        val = np.array([-1], dtype=np.int32).view(np.uint8)
        self._clocks[off:off + 4].copy_(_h2d(val, torch.uint8, self.device), non_blocking=True)
        self._retired.append(r)
        self._retired_particles += self._release_pids[r][1]

    @property
    def releases_in_use(self) -> int:
        """Release ids held: running, finished or retired but not yet dropped by a re-bin."""
        return len(self._clocks_host) - len(self._free_ids)

    @property
    def retired_releases(self) -> int:
        """Retired releases whose particles and ids the next re-bin frees."""
        return len(self._retired)

    @property
    def retired_particles(self) -> int:
        """Particles of retired releases still held (dropped by the next re-bin)."""
        return self._retired_particles

    def extract(self, r: int):
        """(zidx, yidx, xidx, status) of release r as device tensors, in admission order."""
        base, m = self._release_pids[r]
        out = [torch.empty(m, dtype=torch.float64, device=self.device) for _ in range(3)]
        st = torch.empty(m, dtype=torch.uint8, device=self.device)
        if m:
            _lib.call("opx_extract_release", ctypes.byref(self._state.struct(self.n)), ctypes.byref(self.grid), int(r),
                      int(base), int(m), out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(), st.data_ptr(),
                      self._stream())
        return out[0], out[1], out[2], st

    def _grow_releases(self, cap):
        """(Re)allocate the device clock and step tables for `cap` releases."""
        cs, ts = ctypes.sizeof(_lib.opx_release_clock), ctypes.sizeof(_lib.opx_release_step)
        clocks = torch.zeros(cap * cs, dtype=torch.uint8, device=self.device)
        if self._clocks is not None:
            clocks[: self._clocks.numel()].copy_(self._clocks)
        self._clocks = clocks
        # one row of `cap` entries per step of a launch
        self._table = torch.zeros(MAX_LAUNCH_STEPS * cap * ts, dtype=torch.uint8, device=self.device)
        self._rel_cap = cap
        self._graphs = {}

    # ---------------------------------------------------------------- stepping

    def _launch_span(self):
        """Particles the fused launch covers: a power of two >= n, capped at the capacity.

        The kernel steps min(span, *n_active) particles, so a graph captured for
        one span serves every particle count up to it; at most half the blocks
        of a launch find no particle."""
        return min(self._state.capacity, 1 << max(0, int(self.n) - 1).bit_length())

    def _launch_structs(self):
        span = self._launch_span()
        if self._structs is None or self._structs.n != span:
            self._structs = self._state.struct(span)
        return self._structs

    def rebin(self):
        """Sort particles into Morton order of their cells; exited particles go last."""
        if self.n == 0:
            return
        if self._spare is None or self._spare.capacity != self._state.capacity:
            self._spare = self._new_state(self._state.capacity)
            self._graphs = {}
        if self._work is None or self._work_n < self.n:
            self._alloc_work(self.n)
        src = self._state.struct(self.n)
        dst = self._spare.struct(self.n)
        _lib.call("opx_rebin", ctypes.byref(src), ctypes.byref(dst), ctypes.byref(self.grid), self.xlow,
                  self._clocks.data_ptr(), len(self._clocks_host), self._work.data_ptr(), self._work.numel(),
                  self._n_active_dev.data_ptr(), None, self._stream())
        self._state, self._spare = self._spare, self._state
        self._structs = None
        self._since_rebin = 0
        if self._retired:
            # retired releases' particles now sit behind every live one
            self.n -= self._retired_particles
            self._compacted = True
            for r in self._retired:
                del self._release_pids[r]
            self._free_ids.extend(sorted(self._retired, reverse=True))
            self._retired = []
            self._retired_particles = 0

    # ---------------------------------------------------------------- windows

    def _time_index(self, t):
        """Snapshot index It = floor(tidx) of times t and whether t is inside the time array
        (timesteppers.py:64-85)."""
        ta = self.time_array
        ok = (t >= ta[0]) & (t <= ta[-1])
        idx = np.searchsorted(ta, t, side="right") - 1
        ok &= idx + 1 < len(ta)
        idx = np.clip(idx, 0, len(ta) - 2)
        tidx = idx + (t - ta[idx]) / (ta[idx + 1] - ta[idx])
        return np.floor(tidx).astype(np.int64), ok

    def _advance_host_clocks(self):
        """Mirror opx_clock_advance; return the snapshot pairs (It, It + 1) the step reads
        (at t for the step, at t + dt for the post step)."""
        act = (self._left > 0) & (self.global_step >= self._start)
        if not act.any():
            return set()
        t0 = self._t[act]
        t1 = t0 + self.dt
        i0, ok0 = self._time_index(t0)
        i1, ok1 = self._time_index(t1)
        ok = ok0 & ok1
        parts = [i0, i1]
        if self.scheme == "rk2":
            # intermediate time of stage 2: time + alpha * dt (timesteppers.py:250)
            im, okm = self._time_index(t0 + self.alpha * self.dt)
            ok &= okm
            parts.append(im)
        idx = np.nonzero(act)[0]
        self._t[idx[ok]] = t1[ok]
        self._left[idx[ok]] -= 1
        self._left[idx[~ok]] = 0
        its = np.unique(np.concatenate([p[ok] for p in parts]))
        return {(int(i), int(i) + 1) for i in its}

    def _advance_host_clocks_run(self, k: int):
        """k calls of _advance_host_clocks (global steps global_step ..
        global_step + k - 1) at once: the same clocks afterwards and the union
        of the pairs.  A release's times over the run are a cumulative sum
        along its row, the sequential additions (the rounding) of k repeated
        t += dt; it runs from its start step while it has steps left, up to
        the first step whose times leave the array (which stops it)."""
        live = np.nonzero(self._left > 0)[0]
        if live.size == 0 or k <= 0:
            return set()
        j0 = np.maximum(self._start[live] - self.global_step, 0)
        n = np.minimum(np.maximum(k - j0, 0), self._left[live])
        keep = n > 0
        live, n = live[keep], n[keep]
        if live.size == 0:
            return set()
        R, m = live.size, int(n.max())
        T = np.empty((R, m + 1))
        T[:, 0] = self._t[live]
        T[:, 1:] = self.dt
        T = np.cumsum(T, axis=1)  # T[:, i]: the time after i steps
        it, ok = (a.reshape(R, m + 1) for a in self._time_index(T.ravel()))
        step_ok = ok[:, :-1] & ok[:, 1:]
        parts = [it[:, :-1], it[:, 1:]]
        if self.scheme == "rk2":
            # intermediate time of stage 2: time + alpha * dt (timesteppers.py:250)
            im, okm = (a.reshape(R, m) for a in self._time_index((T[:, :-1] + self.alpha * self.dt).ravel()))
            step_ok &= okm
            parts.append(im)
        steps = np.arange(m)
        fail = (steps[None, :] < n[:, None]) & ~step_ok
        done = np.where(fail.any(axis=1), fail.argmax(axis=1), n)
        ran = steps[None, :] < done[:, None]
        self._t[live] = T[np.arange(R), done]
        self._left[live] -= done
        self._left[live[done < n]] = 0
        its = np.unique(np.concatenate([p[ran] for p in parts]))
        return {(int(i), int(i) + 1) for i in its}

    # ---------------------------------------------------------------- streamed snapshots

    def _load_slot(self, It: int) -> None:
        """Copy snapshot It into ring slot It % nslots on the side streams, after the
        work already queued on the compute stream (which may still read the slot).
        With stream_tiles only the tiles the particles can reach are copied (zeta,
        2-D and small, is copied whole)."""
        slot = It % self.nslots
        src = It % self.nsnap
        compute = torch.cuda.current_stream(self.device)
        hosts = {name: self._host_snapshot(name, src) for name in ("u", "v", "w", "zeta")}
        boxes = None
        if self._tile is not None:
            self._rate = max(self._rate, self._snapshot_rate(src, hosts))
            region = self._tile_region(self.global_step + self.tile_margin_steps)
            boxes, covered = tile_boxes(region, self._tile, self.NY, self.NX, self.tile_gap)
            self._slot_tiles[slot] = covered
        # one side stream per 3-D field: the copies run on separate DMA queues
        events = []
        nbytes = 0
        for k, (dst, names) in enumerate(((self._u, ("u",)), (self._v, ("v",)), (self._w, ("w", "zeta")))):
            cs = self._copy_streams[k]
            cs.wait_stream(compute)
            with torch.cuda.stream(cs):
                for name in names:
                    ring = self._zeta if name == "zeta" else dst
                    host = hosts[name]
                    self.snapshot_full_bytes += host.nbytes
                    if boxes is None or name == "zeta":
                        ring[slot].copy_(host, non_blocking=True)
                        nbytes += host.nbytes
                    else:
                        if self._poison_unloaded:
                            # a checking aid: any read of a tile not loaded gives NaN
                            ring[slot].fill_(float("nan"))
                        nbytes += self._copy_boxes(ring[slot], host, name, boxes, cs.cuda_stream)
                ev = torch.cuda.Event()
                ev.record(cs)
            events.append(ev)
        if self._tile is not None:
            # the host copies of the slot's snapshot serve its top-ups; the ones it
            # held before stay referenced until their copies are done
            if self._slot_host[slot] is not None:
                self._host_keep.append(self._slot_host[slot])
            self._slot_host[slot] = (hosts, events)
            self._host_keep = [(h, e) for h, e in self._host_keep if not all(x.query() for x in e)]
        self._ring_It[slot] = It
        self._ring_ready[slot] = events
        self.snapshot_loads += 1
        self.snapshot_bytes += nbytes
        # pair-ring windows packed from the slot's old contents are repacked when
        # next read (their packed snapshot pair no longer matches the slots)

    # ---------------------------------------------------------------- sparse (tiled) streaming

    def _init_tiles(self, fs, NY, NX):
        T = self._tile
        self._nty, self._ntx = -(-NY // T), -(-NX // T)
        self._slot_tiles = [np.zeros((self._nty, self._ntx), dtype=bool) for _ in range(self.nslots)]
        self._slot_host = [None] * self.nslots
        self._host_keep: list = []
        dxm = float(torch.as_tensor(fs["dx"].data.data).min())
        dym = float(torch.as_tensor(fs["dy"].data.data).min())
        if not (dxm > 0 and dym > 0):
            raise ValueError("stream_tiles needs positive grid spacings dx, dy")
        self._inv_dx, self._inv_dy = 1.0 / dxm, 1.0 / dym
        self._reach, self._reach0 = run_reach(self.scheme, self.alpha)
        self._snap_rate: dict = {}
        self._rate = 0.0
        self._occ_dev = torch.zeros(self._nty * self._ntx, dtype=torch.uint8, device=self.device)
        self._occ_pending: list = []  # (seq, step, pinned map, event), oldest first
        self._occ_done = None  # (seq, step, bool map): the latest map read back
        self._occ_adds: list = []  # (seq, bool map) of releases added after it
        self._occ_seq = 0
        # steps of movement a load or top-up covers beyond the run of steps at hand
        self.tile_margin_steps = 8
        # unset tiles bridged between two runs of a tile row (longer copies)
        self.tile_gap = 2
        # wait for a newer map rather than widen the region past this many steps
        self.max_occupancy_lag = 64
        self.tile_topups = 0
        self._poison_unloaded = False
        self._gather_ok = True
        # how tiles reach a slot: "dma" (rectangular DMA copies), "gather" (a
        # kernel reading the pinned snapshot through its device mapping) or
        # "auto" (gather for 1- and 2-byte fields and for tiles narrower than
        # 32 cells, whose rows DMA copies move slowly; DMA otherwise)
        self.tile_copy = "auto"

    def _snapshot_rate(self, src: int, hosts: dict) -> float:
        """Largest |u| / dx, |v| / dy of snapshot src (index units per time unit)."""
        r = self._snap_rate.get(src)
        if r is None:
            r = 0.0
            for name, inv in (("u", self._inv_dx), ("v", self._inv_dy)):
                lo, hi = torch.aminmax(hosts[name].reshape(-1))
                r = max(r, max(abs(float(lo)), abs(float(hi))) * inv)
            if not math.isfinite(r):
                r = math.inf
            self._snap_rate[src] = r
        return r

    def _request_occupancy(self):
        """Tile map of the particles as of the work queued so far, read back asynchronously."""
        _lib.call("opx_tile_occupancy", ctypes.byref(self._state.struct(self.n)), ctypes.byref(self.grid),
                  self._tile, self._occ_dev.data_ptr(), self._nty, self._ntx, self._stream())
        buf = torch.empty(self._nty * self._ntx, dtype=torch.uint8, pin_memory=True)
        buf.copy_(self._occ_dev, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        self._occ_seq += 1
        self._occ_pending.append((self._occ_seq, self.global_step, buf, ev))

    def _fold_occupancy(self):
        seq, step, buf, ev = self._occ_pending.pop(0)
        self._occ_done = (seq, step, buf.numpy().reshape(self._nty, self._ntx) != 0)

    def _occupancy(self, horizon: int):
        """(step S, tile mask): the latest map read back (positions after S steps) with
        the releases added since; waits for a newer map when S lags `horizon` by more
        than max_occupancy_lag steps."""
        while self._occ_pending and self._occ_pending[0][3].query():
            self._fold_occupancy()
        while self._occ_pending and (self._occ_done is None
                                     or horizon - self._occ_done[1] > self.max_occupancy_lag):
            self._occ_pending[0][3].synchronize()
            self._fold_occupancy()
        if self._occ_done is None:
            self._request_occupancy()
            self._occ_pending[0][3].synchronize()
            self._fold_occupancy()
        seq, S, occ = self._occ_done
        mask = occ.copy()
        self._occ_adds = [(q, a) for q, a in self._occ_adds if q > seq]
        for _, a in self._occ_adds:
            mask |= a
        return S, mask

    def _tile_region(self, horizon: int) -> np.ndarray:
        """Tiles a particle can read up to step `horizon`: the tiles of the latest map
        widened by the bound on the reach of the steps since it, plus two cells of
        stencil reach."""
        S, occ = self._occupancy(horizon)
        m = max(0, horizon - S)
        r = (m * self._reach + self._reach0) * abs(self.dt) * self._rate + 2.0
        if not math.isfinite(r) or r >= self._tile * max(self._nty, self._ntx):
            return np.ones_like(occ)
        return dilate_tiles(occ, int(math.ceil(r / self._tile)))

    def _tile_mask_of(self, y, x) -> np.ndarray:
        """Tiles holding the finite positions (y, x) (cells outside the grid in the
        nearest edge tile), as the occupancy kernel counts them."""
        y = np.asarray(y, dtype=np.float64).reshape(-1)
        x = np.asarray(x, dtype=np.float64).reshape(-1)
        ok = np.isfinite(y) & np.isfinite(x)
        ty = np.clip(np.floor(y[ok]), 0, self.NY - 1).astype(np.int64) // self._tile
        tx = np.clip(np.floor(x[ok]), 0, self.NX - 1).astype(np.int64) // self._tile
        mask = np.zeros((self._nty, self._ntx), dtype=bool)
        mask[ty, tx] = True
        return mask

    def _copy_boxes(self, dst, host, name, boxes, stream) -> int:
        """Copy the rho-cell boxes (clipped to the field's own shape) of `host` into
        `dst`; returns the bytes copied."""
        nz, ny, nx = self._ring_shape[name]
        b = boxes.copy()
        b[:, 1] = np.minimum(b[:, 1], ny)
        b[:, 3] = np.minimum(b[:, 3], nx)
        b = np.ascontiguousarray(b[(b[:, 0] < b[:, 1]) & (b[:, 2] < b[:, 3])], dtype=np.int32)
        if not len(b):
            return 0
        if not (host.is_contiguous() and dst.is_contiguous()):
            raise ValueError("tile copies need contiguous snapshots")
        args = (dst.data_ptr(), host.data_ptr(), nz, ny, nx, host.element_size(), b.ctypes.data, len(b), stream)
        gather = self.tile_copy == "gather" or (self.tile_copy == "auto"
                                                 and (host.element_size() < 4 or self._tile < 32))
        if gather and self._gather_ok:
            # rectangular DMA copies of 2-byte rows run far below the link rate:
            # a kernel reads the boxes through the pinned memory's device mapping
            try:
                _lib.call("opx_gather_boxes_h2d", *args)
                return self._box_bytes(b, nz, host.element_size())
            except _lib.OpxInvalidArgument:
                self._gather_ok = False  # not mapped: DMA copies from now on
        _lib.call("opx_copy_boxes_h2d", *args)
        return self._box_bytes(b, nz, host.element_size())

    @staticmethod
    def _box_bytes(b, nz, eb) -> int:
        return int(((b[:, 1] - b[:, 0]).astype(np.int64) * (b[:, 3] - b[:, 2])).sum()) * nz * eb

    def _cover(self, need) -> None:
        """Make every slot of snapshots `need` hold the tiles the particles can reach
        by the end of the queued steps: copy the missing ones (on the compute
        stream, so the next steps see them) and pack their rows again."""
        req = self._tile_region(self.global_step)
        grow = None
        stream = self._stream()
        for It in need:
            slot = It % self.nslots
            miss = req & ~self._slot_tiles[slot]
            if not miss.any():
                continue
            if grow is None:
                grow = self._tile_region(self.global_step + self.tile_margin_steps) | req
            boxes, covered = tile_boxes(grow & ~self._slot_tiles[slot], self._tile, self.NY, self.NX, self.tile_gap)
            hosts = self._slot_host[slot][0]
            nbytes = 0
            for dst, name in ((self._u, "u"), (self._v, "v"), (self._w, "w")):
                nbytes += self._copy_boxes(dst[slot], hosts[name], name, boxes, stream)
            self._slot_tiles[slot] |= covered
            self.tile_topups += 1
            self.snapshot_bytes += nbytes
            self._repack_rows(slot, boxes)

    def _repack_rows(self, slot: int, boxes) -> None:
        """Pack again the nodes of `boxes` (and of the row above and the column
        left of each, whose y + 1 / x + 1 neighbours they are) in the pair-ring
        windows packed from this slot's current contents."""
        if self.layout == _lib.OPX_LAYOUT_RING or self._pair_bufs is None or not len(boxes):
            return
        for s0 in {(slot - 1) % self.nslots, slot}:
            key = self._pair_contents(s0)
            if self._pair_key[s0] != key or None in key:
                continue  # packed when next read
            self._pack_boxes(s0, boxes)

    def _pack_boxes(self, s0: int, boxes) -> None:
        """opx_pack_window_boxes of pair-ring window s0 over `boxes` (int32 (k, 4))."""
        b = np.ascontiguousarray(boxes, dtype=np.int32)
        bufs = [p.data_ptr() + s0 * st for p, st in zip(self._pair_bufs, self._pair_strides)]
        _lib.call("opx_pack_window_boxes", ctypes.byref(self.window), ctypes.byref(self.grid), self.layout, s0,
                  (s0 + 1) % self.nslots, b.ctypes.data, len(b), *bufs, self._stream())

    def _stage(self, name: str, src: int):
        """Start reading snapshot `src` of a lazily read field into pinned staging
        memory on the source's host threads (once); returns (future, pinned tensor)."""
        key = (name, src)
        job = self._staged.get(key)
        if job is None:
            s = self._src[name]
            pinned = torch.empty(self._ring_shape[name], dtype=self._dts[name], pin_memory=True)
            if pinned.dtype == torch.bfloat16 or np.dtype(s.dtype) != pinned.numpy().dtype:
                fut = s.submit(lambda: pinned.copy_(torch.from_numpy(s.read(src))))
            else:
                fut = s.submit(s.read_into, src, pinned.numpy())
            job = self._staged[key] = (fut, pinned)
        return job

    def _host_snapshot(self, name: str, src: int) -> torch.Tensor:
        """Snapshot `src` of a field in pinned host memory: a slice of the pinned copy,
        or read from the field's SnapshotSource into pinned staging memory (the
        next snapshot is then staged ahead on a host thread)."""
        s = self._src[name]
        if not isinstance(s, SnapshotSource):
            return s[src]
        fut, pinned = self._stage(name, src)
        fut.result()
        del self._staged[(name, src)]
        self._stage(name, (src + 1) % self.nsnap)
        return pinned

    def _snapshot_staged(self, It: int) -> bool:
        """Whether every lazily read field has snapshot It in pinned memory (starting
        the reads that have not started)."""
        ok = True
        for name, s in self._src.items():
            if isinstance(s, SnapshotSource):
                ok &= self._stage(name, It % self.nsnap)[0].done()
        return ok

    def _ensure_snapshots(self, pairs) -> None:
        """Make the snapshots of `pairs` resident; prefetch the one after them."""
        if self._src is None or not pairs:
            return
        need = sorted({i for p in pairs for i in p})
        slots = [i % self.nslots for i in need]
        if len(set(slots)) != len(slots):
            raise ValueError(f"active releases read snapshots {need[0]}..{need[-1]}, more than the "
                             f"{self.nslots}-slot streamed window holds")
        compute = torch.cuda.current_stream(self.device)
        for It, slot in zip(need, slots):
            if self._ring_It[slot] != It:
                self._load_slot(It)
            evs = self._ring_ready[slot]
            if evs is not None:
                for ev in evs:
                    compute.wait_event(ev)
                self._ring_ready[slot] = None
        if self._tile is not None:
            self._cover(need)
        nxt = need[-1] + 1
        slot = nxt % self.nslots
        cur = self._ring_It[slot]
        # the look-ahead load waits for nothing: a snapshot still being read from
        # its source is loaded by a later call
        if cur != nxt and (cur is None or cur < need[0]) and self._snapshot_staged(nxt):
            self._load_slot(nxt)

    def _pair_contents(self, s0):
        """The snapshot pair that ring slots s0 and s0 + 1 hold now."""
        s1 = (s0 + 1) % self.nslots
        if self._src is None:  # every snapshot resident: slot s holds snapshot s
            return (s0, s1)
        return (self._ring_It[s0], self._ring_It[s1])

    def _ensure_windows(self, pairs):
        """Pack the pair-ring windows of the snapshot pairs the next steps read, where stale.

        The pair ring holds one packed window per ring slot s, of the slot pair
        (s, s + 1); a step reading the pair with first slot s0 reads window s0,
        so every release reads packed stencils whatever its time."""
        w = self.window
        if self.layout == _lib.OPX_LAYOUT_RING:
            if w.npacked != 0:
                w.npacked = 0
                self._win_dirty = True
            return
        if self._pair_bufs is None:
            strides = [-(-b // 256) * 256 for b in self.packed_window_bytes()]
            self._pair_bufs = [torch.empty(self.nslots * st, dtype=torch.uint8, device=self.device) for st in strides]
            self._pair_strides = strides
            self._pair_key = [None] * self.nslots
            w.layout = self.layout
            w.npacked = self.nslots
            w.pu, w.pv, w.pw, w.p2 = (b.data_ptr() for b in self._pair_bufs)
            w.pu_stride, w.pv_stride, w.pw_stride, w.p2_stride = strides
            self._win_dirty = True
        for a, _ in sorted(pairs):
            s0 = a % self.nslots
            key = self._pair_contents(s0)
            if self._pair_key[s0] != key:
                if self._tile is not None:
                    # the slots hold only their loaded tiles: pack those (top-ups
                    # pack what they add)
                    s1 = (s0 + 1) % self.nslots
                    boxes, _ = tile_boxes(self._slot_tiles[s0] | self._slot_tiles[s1], self._tile, self.NY, self.NX)
                    self._pack_boxes(s0, boxes)
                else:
                    bufs = [b.data_ptr() + s0 * st for b, st in zip(self._pair_bufs, self._pair_strides)]
                    _lib.call("opx_pack_window", ctypes.byref(w), ctypes.byref(self.grid), self.layout, s0,
                              (s0 + 1) % self.nslots, *bufs, self._stream())
                self._pair_key[s0] = key

    def set_layout(self, layout: int) -> None:
        """Choose the stencil layout (OPX_LAYOUT_RING, _PAIR, _QUAD, _PAIR_LDS, _OCT or _YPAIR)."""
        if int(layout) != _lib.OPX_LAYOUT_RING:
            self.packed_window_bytes(layout)  # refuses a window above 4 GiB
        if layout != self.layout:
            self.layout = int(layout)
            self._pair_bufs = None
            self._pair_key = None
            self._win_dirty = True

    def _upload_window(self):
        if self._win_dirty:
            raw = np.frombuffer(bytes(self.window), dtype=np.uint8)
            self._win_dev.copy_(_h2d(raw, torch.uint8, self.device), non_blocking=True)
            self._win_dirty = False

    def _launch_steps(self, stream, m: int = 1):
        """m steps on `stream`: device clocks, then one fused launch (both graph-capturable)."""
        rk2 = self.scheme == "rk2"
        _lib.call("opx_clock_advance_steps", self._clocks.data_ptr(), self._table.data_ptr(), self._rel_cap, int(m),
                  self._rel_cap, self._nrel_dev.data_ptr(), self._ta.data_ptr(), int(self._ta.numel()), self.nslots,
                  0, self._step_dev.data_ptr(), self.dt, self.alpha if rk2 else 0.0, self._err.data_ptr(), stream)
        s = self._launch_structs()
        layout = self.layout
        if rk2:
            _lib.call("opx_rk2_fused_steps_dev", ctypes.byref(s), ctypes.byref(self.grid), self._win_dev.data_ptr(),
                      self.window.dtype3, layout, self._table.data_ptr(), self._rel_cap, int(m), self.dt, self.alpha,
                      self._n_active_dev.data_ptr(), self.block, self._err.data_ptr(), stream)
        elif self.scheme == "ab3_iso":
            _lib.call("opx_ab3_iso_fused_steps_dev", ctypes.byref(s), ctypes.byref(self.grid),
                      self._win_dev.data_ptr(), layout, self._table.data_ptr(), self._rel_cap, int(m), self.dt,
                      self.w_max, self.rho_scale, self._n_active_dev.data_ptr(), self.block, self._err.data_ptr(),
                      stream)
        else:
            _lib.call("opx_ab3_fused_steps_dev", ctypes.byref(s), ctypes.byref(self.grid), self._win_dev.data_ptr(),
                      self.window.dtype3, layout, self._table.data_ptr(), self._rel_cap, int(m), self.dt,
                      self._n_active_dev.data_ptr(), self.block, self._err.data_ptr(), stream)

    def _launch_sizes(self, k):
        """Steps per launch for a run of k steps."""
        m = self.launch_steps
        return [m] * (k // m) + ([k % m] if k % m else [])

    def _graph(self, k):
        """A captured graph of k steps for the current buffers and launch shape."""
        key = (self._state.cell.data_ptr(), self._launch_span(), k, self.layout, self.block, self.launch_steps)
        g = self._graphs.get(key)
        if g is None:
            # captured on a side stream ordered after the compute stream's work,
            # without torch.cuda.graph's device synchronisation and cache release
            # on entry (each capture would drain the device and re-allocate)
            if self._capture_stream is None:
                self._capture_stream = torch.cuda.Stream(device=self.device)
            cap = self._capture_stream
            cur = torch.cuda.current_stream(self.device)
            cap.wait_stream(cur)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.stream(cap):
                g.capture_begin()
                try:
                    for m in self._launch_sizes(k):
                        self._launch_steps(cap.cuda_stream, m)
                finally:
                    g.capture_end()
            cur.wait_stream(cap)
            self._graphs[key] = g
        return g

    def step(self, nsteps: int = 1) -> None:
        """Advance every active release by ``nsteps`` AB3 steps.

        With ``graph_steps`` > 0, runs of up to that many steps between
        re-bins are replayed from captured HIP graphs; the window descriptor,
        release count, active count and global step live in device memory and
        the fused kernel's grid covers a power-of-two span of particles, so a
        graph stays valid while windows are repacked, snapshots are streamed
        and releases are added or retired.
        """
        stream = self._stream()
        left = int(nsteps)
        while left > 0:
            if self.rebin_every > 0 and self._since_rebin >= self.rebin_every:
                self.rebin()
            k = min(self.launch_steps, left)
            if self.graph_steps > 0 and self.timer is None:
                k = min(max(self.graph_steps, k), left)
            if self.rebin_every > 0:
                k = max(1, min(k, self.rebin_every - self._since_rebin))
            pairs = self._advance_host_clocks_run(k)
            self.global_step += k
            self._ensure_snapshots(pairs)
            self._ensure_windows(pairs)
            self._upload_window()
            if self.graph_steps > 0 and self.timer is None and k > 1:
                self._graph(k).replay()
            else:
                for m in self._launch_sizes(k):
                    if self.timer is not None:
                        e0 = torch.cuda.Event(enable_timing=True)
                        e0.record()
                    self._launch_steps(stream, m)
                    if self.timer is not None:
                        e1 = torch.cuda.Event(enable_timing=True)
                        e1.record()
                        self.timer.append((e0, e1, m))
            if self._tile is not None:
                self._request_occupancy()
            self._since_rebin += k
            left -= k

    # ---------------------------------------------------------------- tuning

    def _time_steps(self, k: int) -> float:
        """Fused-kernel time (ms) per step over k real steps."""
        saved, self.timer = self.timer, []
        self.step(k)
        torch.cuda.synchronize(self.device)
        ms = (sum(a.elapsed_time(b) for a, b, _ in self.timer) / sum(m for _, _, m in self.timer)
              if self.timer else float("nan"))
        self.timer = saved
        return ms

    def autotune(self, *, xlows=(3, 4, 5), blocks=(128, 256), layouts=None, launch_steps=None,
                 trial_steps: int = 3, cadence_probe: int = 40, max_cadence: int = 400) -> dict:
        """Pick the bin shape, stencil layout, workgroup size, steps per launch and
        re-sort cadence by timing real steps.

        - bin shape: ``xlow`` low bits of the x cell kept contiguous in the sort
          key (Morton order above them), i.e. bins 2^xlow cells long in x;
        - stencil layout (ring / pair / quad / pair_lds / oct / ypair) and
          workgroup size of the fused kernel (timed at one step per launch);
        - AB3 steps per launch (``launch_steps`` candidates, default 1, 8, 16,
          32): more steps keep the state in registers longer and reload fewer
          stencils, but a launch needs the snapshots of all its steps;
        - re-sort cadence: after a rebin the step time grows as particles drift
          out of order; with rebin cost R and growth a per step, the average
          cost per step is minimal at N = sqrt(2 R / a), rounded up to whole
          launches.
        Tuning advances the simulation (results do not depend on these
        parameters); it needs about ``len(xlows) * (len(layouts) * len(blocks))
        * (trial_steps + 1) + 2 * sum(launch_steps) + cadence_probe`` steps of
        the active releases.
        """
        import time as _time

        if layouts is None:
            layouts = (_lib.OPX_LAYOUT_RING, _lib.OPX_LAYOUT_PAIR, _lib.OPX_LAYOUT_QUAD, _lib.OPX_LAYOUT_PAIR_LDS,
                       _lib.OPX_LAYOUT_OCT, _lib.OPX_LAYOUT_YPAIR)
            layouts = tuple(lay for lay in layouts if self.layout_fits(lay))
        if launch_steps is None:
            launch_steps = (1, 8, 16, 32)
        saved_every, self.rebin_every = self.rebin_every, 0
        self.launch_steps = 1
        trials = {}
        for xlow in xlows:
            self.xlow = int(xlow)
            self.rebin()
            for layout in layouts:
                self.set_layout(layout)
                # the LDS-tiled step runs one wave per workgroup
                for block in (blocks if layout != _lib.OPX_LAYOUT_PAIR_LDS else blocks[:1]):
                    self.block = int(block)
                    self._time_steps(1)  # first launch of a variant (and window packing)
                    trials[(self.xlow, self.layout, self.block)] = self._time_steps(trial_steps)
        (self.xlow, layout, self.block), best = min(trials.items(), key=lambda kv: kv[1])
        self.set_layout(layout)
        self.rebin()
        per_launch = {}
        for m in launch_steps:
            self.launch_steps = int(m)
            self._time_steps(int(m))
            per_launch[int(m)] = self._time_steps(max(trial_steps, int(m)))
        self.launch_steps, best = min(per_launch.items(), key=lambda kv: kv[1])
        m = self.launch_steps
        torch.cuda.synchronize(self.device)
        t0 = _time.perf_counter()
        self.rebin()
        torch.cuda.synchronize(self.device)
        rebin_ms = 1e3 * (_time.perf_counter() - t0)
        nprobe = max(3, -(-cadence_probe // m))
        ts = np.array([self._time_steps(m) for _ in range(nprobe)])
        k = m * np.arange(nprobe, dtype=np.float64)
        slope = float(np.polyfit(k, ts, 1)[0])
        cadence = int(np.clip(np.sqrt(2.0 * rebin_ms / slope), 5, max_cadence)) if slope > 0 else max_cadence
        cadence = m * max(1, -(-cadence // m))
        self.rebin_every = cadence if saved_every != 0 else 0
        self._since_rebin = m * nprobe
        names = LAYOUT_NAMES
        self.tuning = {
            "xlow": self.xlow, "layout": names[self.layout], "block": self.block, "launch_steps": m,
            "rebin_every": self.rebin_every, "kernel_ms_per_step": best, "rebin_ms": rebin_ms,
            "drift_ms_per_step": slope,
            "trials": {f"xlow={x},{names[lay]},block={b}": round(v, 4) for (x, lay, b), v in trials.items()},
            "launch_steps_trials": {str(k): round(v, 4) for k, v in per_launch.items()},
        }
        return self.tuning

    def check_clock(self) -> None:
        """Raise ValueError if a release's time left the time array (as the reference does)."""
        err = int(self._err.item())
        if err == _lib.OPX_FLAG_UNPACKED_PAIR:
            raise RuntimeError("a several-step launch met a snapshot pair that was not packed")
        if err:
            raise ValueError("Time is out of bounds of the time array.")

    # ---------------------------------------------------------------- results

    def positions(self):
        """(zidx, yidx, xidx) FP64 device tensors in the order particles were added."""
        if self._compacted:
            raise RuntimeError("releases were retired: read results per release with extract()")
        out = [torch.empty(self.n, dtype=torch.float64, device=self.device) for _ in range(3)]
        if self.n:
            _lib.call("opx_split_to_f64", ctypes.byref(self._state.struct(self.n)), ctypes.byref(self.grid),
                      out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(), self._stream())
        return tuple(out)

    def status(self):
        """Status codes (uint8 device tensor) in the order particles were added."""
        if self._compacted:
            raise RuntimeError("releases were retired: read results per release with extract()")
        out = torch.empty(self.n, dtype=torch.uint8, device=self.device)
        pid = self._state.pid[: self.n].long()
        out[pid] = self._state.status[: self.n]
        return out

    def sample(self, fieldset, fields=(), *, depth: str | None = None) -> dict:
        """Output event of the fused form: FP64 values at every particle, in the
        order of positions().

        ``fields`` names fields of ``fieldset`` (the drop-in Fieldset, e.g. the
        one the engine was built from): a 3-D field is interpolated trilinearly,
        a 2-D one bilinearly over (y, x) and a 1-D one linearly over z, like
        trilinear/bilinear/linear_interpolation_kernel
        (kernels/output/linear_interpolation.pyx:34-136); ``depth`` names an
        extra output holding z (compute_z_kernel, _roms_core.pyx:17-73).  Each
        release is sampled at its own current time (the time index of
        Timestepper.get_time_index, blend weights of the field's dtype).
        Inactive particles get NaN.  One kernel evaluates every output of a
        particle, in the re-binned order, so the gathers stay local; the values
        equal the drop-in kernels' at positions() bit for bit.  At most 8
        outputs per call; they are column views of one (n, 1/2/4/8) record
        tensor (each particle's outputs written together).
        """
        from .kernels._marshal import field_struct

        if self._compacted:
            raise RuntimeError("releases were retired: sample() follows positions(), which needs every release")
        names = list(dict.fromkeys(fields))
        if depth is not None:
            for k in ("h", "zeta", "C"):
                if k not in names:
                    names.append(k)
            for k in ("hc", "NZ"):
                if k not in fieldset.constants:
                    raise KeyError(k)
        if not names:
            raise ValueError("nothing to sample")
        nrel = len(self._clocks_host)
        rows = (_lib.opx_field * (max(1, nrel) * len(names)))()
        for r in range(nrel):
            t = self._t[r]
            ta = self.time_array
            if not (ta[0] <= t <= ta[-1]):
                raise ValueError("Time is out of bounds of the time array.")
            idx = np.searchsorted(ta, t, side="right") - 1
            tidx = idx + (t - ta[idx]) / (ta[idx + 1] - ta[idx])
            # the particles are sorted by cell: the plain layouts serve them better
            # than the drop-in's packed copies
            with fields_unpacked():
                for k, name in enumerate(names):
                    rows[r * len(names) + k] = field_struct(fieldset[name].get_field_data(tidx, None), name)
        index = {name: k for k, name in enumerate(names)}
        kinds = {1: _lib.OPX_SAMPLE_LINEAR, 2: _lib.OPX_SAMPLE_BILINEAR, 3: _lib.OPX_SAMPLE_TRILINEAR}
        outs, samples = [], []
        for name in dict.fromkeys(fields):
            ndim = rows[index[name]].ndim
            if ndim not in kinds:
                raise ValueError(f"field '{name}' has {ndim} dimensions")
            outs.append(name)
            samples.append(_lib.opx_sample(kinds[ndim], index[name], 0, 0))
        if depth is not None:
            outs.append(depth)
            samples.append(_lib.opx_sample(_lib.OPX_SAMPLE_COMPUTE_Z, index["h"], index["zeta"], index["C"]))
        if len(samples) > 8:
            raise ValueError(f"{len(samples)} outputs; one sample() call writes at most 8")
        # one record of 1, 2, 4 or 8 FP64 slots per particle, written whole
        width = next(w for w in (1, 2, 4, 8) if w >= len(samples))
        rec = torch.empty((self.n, width), dtype=torch.float64, device=self.device)
        if self.n and nrel:
            rows_dev = _h2d(np.frombuffer(bytes(rows), dtype=np.uint8), torch.uint8, self.device)
            smp = (_lib.opx_sample * len(samples))(*samples)
            smp_dev = _h2d(np.frombuffer(bytes(smp), dtype=np.uint8), torch.uint8, self.device)
            if self.timer is not None:
                e0 = torch.cuda.Event(enable_timing=True)
                e0.record()
            _lib.call("opx_sample_split", ctypes.byref(self._state.struct(self.n)), ctypes.byref(self.grid),
                      rows_dev.data_ptr(), len(names), nrel, smp_dev.data_ptr(), len(samples),
                      float(fieldset.constants.get("hc", self.grid.hc)), rec.data_ptr(), width, self._stream())
            if self.timer is not None:
                e1 = torch.cuda.Event(enable_timing=True)
                e1.record()
                self.timer.append((e0, e1, 0))
            # the tables live until the launch has read them
            self._sample_tables = (rows_dev, smp_dev)
        out = {name: rec[:, k] for k, name in enumerate(outs)}
        return out

    def release_slice(self, r: int) -> slice:
        """Positions of release r in positions()/status() (valid while no release was retired)."""
        if self._compacted:
            raise RuntimeError("releases were retired: read results per release with extract()")
        start, m = self._release_ranges[r]
        return slice(start, start + m)

    def active_count(self) -> int:
        return int((self._state.status[: self.n] < 128).sum().item())

    @property
    def state(self) -> SplitState:
        return self._state
