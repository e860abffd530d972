"""Device engine over libefm.so: wavefield storage, shot batches, forward runs
and exact-adjoint gradients over a forward tape that stores every state, keeps
the last states and √-checkpoints the first steps (hybrid), or keeps √nt
checkpoints, whichever fits the device memory.

torch is used only to own device memory and to name the current HIP stream;
every arithmetic operation on the path runs in the HIP kernels of libefm.so.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _capi
from ._capi import EfmDims, EfmMedium, EfmShots, EfmState, check

H_STENCIL = 2
FIELDS = ("vx", "vy", "vz", "txx", "tyy", "tzz", "txy", "txz", "tyz")
FIELD_PAD = 1088   # floats between fields of one allocation (keeps equal cells off power-of-two strides)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else ctypes.c_void_p(0)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@dataclass(frozen=True)
class Grid:
    """Grid and CPML geometry.  ``widths`` = (x_lo, x_hi, y_lo, y_hi, z_lo, z_hi)
    CPML cells per face; the reference uses W on all faces but z_lo (free top,
    M:1305-1306)."""
    nx: int
    ny: int
    nz: int
    dx: float
    dt: float
    widths: tuple

    @classmethod
    def make(cls, nx, ny, nz, dx, dt, W, free_top=True):
        return cls(int(nx), int(ny), int(nz), float(dx), float(dt),
                   (W, W, W, W, 0 if free_top else W, W))

    @property
    def shape(self):
        return (self.ny, self.nz, self.nx)

    @property
    def N(self):
        return self.nx * self.ny * self.nz

    @property
    def slab_lengths(self):
        w = self.widths
        return (w[0] + w[1], w[2] + w[3], w[4] + w[5])

    def interior(self, x, y, z):
        return (H_STENCIL <= x < self.nx - H_STENCIL and H_STENCIL <= y < self.ny - H_STENCIL
                and H_STENCIL <= z < self.nz - H_STENCIL)


class ShotBatch:
    """Per-shot sources, receivers, wavelets and CPML profiles, on the device.

    src (S, 3) integer (x, y, z); m6 (S, 6) (xx, yy, zz, xy, xz, yz); wavelet
    (S, ntmax); nt (S,); rec (S, nrec, 3) integer (x, y, z), -1 rows unused;
    cpml: dict of (S, n) arrays ax, bx, ay, by, az, bz."""

    def __init__(self, grid: Grid, src, m6, wavelet, nt, rec, cpml, device="cuda", fuse_receivers=True):
        self.grid = grid
        src = np.asarray(src, np.int64).reshape(-1, 3)
        S = src.shape[0]
        wavelet = np.asarray(wavelet, np.float32).reshape(S, -1)
        nt = np.asarray(nt, np.int64).reshape(S)
        rec = np.asarray(rec, np.int64).reshape(S, -1, 3)
        for s in range(S):
            if not grid.interior(*src[s]):
                raise ValueError(f"source {tuple(src[s])} of shot {s} is not an interior cell "
                                 f"(2 <= x,y,z < n-2)")
        if np.any(nt < 1) or np.any(nt > wavelet.shape[1]):
            raise ValueError("record lengths must lie in [1, wavelet length]")
        used = rec[..., 0] >= 0
        rx, ry, rz = rec[..., 0], rec[..., 1], rec[..., 2]
        if np.any(used & ((rx >= grid.nx) | (ry < 0) | (ry >= grid.ny) | (rz < 0) | (rz >= grid.nz))):
            raise ValueError("receiver outside the grid")
        lin = np.where(used, rx + grid.nx * rz + grid.nx * grid.nz * ry, -1)
        self.S, self.ntmax, self.nrec = S, wavelet.shape[1], rec.shape[1]
        dev = torch.device(device)
        # C-ordered, writable copies: the kernels index (S, n) arrays row-major
        # (a broadcast view copied in its own memory order would be column-major)
        f32 = lambda a: torch.as_tensor(np.array(a, np.float32, order="C"), device=dev)
        i32 = lambda a: torch.as_tensor(np.array(a, np.int32, order="C"), device=dev)
        self.src_host, self.nt_host, self.rec_host = src, nt, rec
        self.src, self.m6, self.wavelet = i32(src), f32(np.asarray(m6).reshape(S, 6)), f32(wavelet)
        self.nt, self.rec = i32(nt), i32(lin)
        # receiver bitmask (S, ceil(N/32)): the sweeps sample receivers and inject
        # residuals at their own cells instead of separate kernels each step
        self.recmask = None
        if fuse_receivers:
            words = (grid.N + 31) // 32
            mask = np.zeros((S, words), np.uint32)
            for s in range(S):
                cells = lin[s][lin[s] >= 0]
                np.bitwise_or.at(mask[s], cells // 32, (np.uint32(1) << (cells % 32).astype(np.uint32)))
            self.recmask = i32(mask.view(np.int32))
        prof = {}
        for k, n in (("ax", grid.nx), ("bx", grid.nx), ("ay", grid.ny), ("by", grid.ny),
                     ("az", grid.nz), ("bz", grid.nz)):
            a = np.asarray(cpml[k], np.float32)
            prof[k] = f32(np.broadcast_to(a, (S, n)))
        self.cpml = prof
        for t in (self.src, self.m6, self.wavelet, self.nt, self.rec, *prof.values()):
            assert t.is_contiguous()

    def set_m6(self, m6):
        self.m6.copy_(torch.as_tensor(np.asarray(m6, np.float32).reshape(self.S, 6)))

    def struct(self):
        c = self.cpml
        return EfmShots(_ptr(c["ax"]), _ptr(c["bx"]), _ptr(c["ay"]), _ptr(c["by"]), _ptr(c["az"]),
                        _ptr(c["bz"]), _ptr(self.src), _ptr(self.m6), _ptr(self.wavelet),
                        _ptr(self.nt), _ptr(self.rec), _ptr(self.recmask))


def component_pad():
    return int(_capi.lib().efm_component_pad())


def padded_components(ncomp, N, device, pad=None):
    """(ncomp, N) float32 zeros whose component stride is N + pad (the layout of
    the library's multi-component scratch arrays)."""
    pad = component_pad() if pad is None else pad
    buf = torch.zeros(ncomp * (N + pad), dtype=torch.float32, device=device)
    return torch.as_strided(buf, (ncomp, N), (N + pad, 1))


def available_bytes(device):
    """Device memory a new allocation can use: what the runtime reports free plus
    what torch's caching allocator holds reserved but unused (torch hands those
    blocks out again before asking the runtime)."""
    free, _ = torch.cuda.mem_get_info(device)
    cached = torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
    return free + max(0, cached)


class Wavefield:
    """Batched state: 9 volumes (S, ny, nz, nx) + three ψ slabs.  The 9 volumes
    share one allocation with a padded field stride (S·N + pad floats), so that
    equal cells of different fields are not a power of two apart."""

    def __init__(self, grid: Grid, S: int, device="cuda", psi_only_of=None):
        self.grid, self.S = grid, S
        Lx, Ly, Lz = grid.slab_lengths
        ny, nz, nx = grid.shape
        dev = torch.device(device)
        if psi_only_of is None:
            self.fields = _padded_fields(9, grid, S, dev)
        else:
            self.fields = psi_only_of.fields
        self.psi_x = torch.zeros((S, 6, ny, nz, Lx), dtype=torch.float32, device=dev)
        self.psi_y = torch.zeros((S, 6, Ly, nz, nx), dtype=torch.float32, device=dev)
        self.psi_z = torch.zeros((S, 6, ny, Lz, nx), dtype=torch.float32, device=dev)
        self._s = None

    @property
    def device(self):
        return self.fields.device

    def __getitem__(self, name):
        return self.fields[FIELDS.index(name)]

    def struct(self):
        if self._s is None:
            self._s = EfmState(*[_ptr(self.fields[k]) for k in range(9)],
                               _ptr(self.psi_x), _ptr(self.psi_y), _ptr(self.psi_z))
        return self._s

    def zero_(self):
        self.fields.zero_(); self.psi_x.zero_(); self.psi_y.zero_(); self.psi_z.zero_()
        return self

    def copy_(self, other):
        self.fields.copy_(other.fields); self.psi_x.copy_(other.psi_x)
        self.psi_y.copy_(other.psi_y); self.psi_z.copy_(other.psi_z)
        return self

    def clone(self):
        w = Wavefield(self.grid, self.S, self.fields.device)
        return w.copy_(self)

    def nbytes(self):
        return sum(t.numel() * 4 for t in (self.fields, self.psi_x, self.psi_y, self.psi_z))


def _padded_fields(n, grid, S, device):
    """(n, S, ny, nz, nx) float32 zeros in one allocation with the padded field
    stride of Wavefield."""
    ny, nz, nx = grid.shape
    per = S * grid.N
    buf = torch.zeros(n * (per + FIELD_PAD), dtype=torch.float32, device=device)
    return torch.as_strided(buf, (n, S, ny, nz, nx), (per + FIELD_PAD, grid.N, nz * nx, nx, 1))


class TapeState:
    """A forward state kept for the exact adjoint: its own velocities and CPML ψ
    slabs, and the stresses of one buffer that every state of the tape shares.
    The adjoint reads the stored states' v (the ρ products come from the
    velocity change, the λ/μ products from the strains) and ψ, never their
    stresses, so one live copy of τ carries the forward steps (the stress sweep
    updates each cell in place).  A tape state takes 3 volumes + ψ instead of 9
    + ψ.  `tau` is the shared (6, S, ny, nz, nx) buffer."""

    def __init__(self, grid: Grid, S: int, device, tau):
        self.grid, self.S = grid, S
        Lx, Ly, Lz = grid.slab_lengths
        ny, nz, nx = grid.shape
        dev = torch.device(device)
        self.v = _padded_fields(3, grid, S, dev)
        self.tau = tau
        self.psi_x = torch.zeros((S, 6, ny, nz, Lx), dtype=torch.float32, device=dev)
        self.psi_y = torch.zeros((S, 6, Ly, nz, nx), dtype=torch.float32, device=dev)
        self.psi_z = torch.zeros((S, 6, ny, Lz, nx), dtype=torch.float32, device=dev)
        self._s = None

    @property
    def device(self):
        return self.v.device

    def struct(self):
        if self._s is None:
            self._s = EfmState(*[_ptr(self.v[k]) for k in range(3)], *[_ptr(self.tau[k]) for k in range(6)],
                               _ptr(self.psi_x), _ptr(self.psi_y), _ptr(self.psi_z))
        return self._s

    def zero_(self):
        """Zero this state and the shared stresses (the initial state of a run)."""
        self.v.zero_(); self.tau.zero_(); self.psi_x.zero_(); self.psi_y.zero_(); self.psi_z.zero_()
        return self

    def copy_(self, other: "Wavefield"):
        """Load a full state: its v and ψ here, its stresses into the shared buffer."""
        if not isinstance(other, Wavefield):
            raise TypeError("a tape state loads a full Wavefield (a snapshot or a live state)")
        self.v.copy_(other.fields[:3]); self.tau.copy_(other.fields[3:])
        self.psi_x.copy_(other.psi_x); self.psi_y.copy_(other.psi_y); self.psi_z.copy_(other.psi_z)
        return self


def lame(vp, vs, rho):
    """λ, μ from Vp, Vs, ρ on the device (efm_model_lame)."""
    lam = torch.empty_like(vp); mu = torch.empty_like(vp)
    check(_capi.lib().efm_model_lame(_ptr(vp), _ptr(vs), _ptr(rho), _ptr(lam), _ptr(mu),
                                     vp.numel(), _stream()), "efm_model_lame")
    return lam, mu


def lame_adjoint(vp, vs, rho, glam, gmu, grho):
    gvp = torch.empty_like(vp); gvs = torch.empty_like(vp); gr = torch.empty_like(vp)
    check(_capi.lib().efm_model_lame_adjoint(_ptr(vp), _ptr(vs), _ptr(rho), _ptr(glam), _ptr(gmu),
                                             _ptr(grho), _ptr(gvp), _ptr(gvs), _ptr(gr), vp.numel(),
                                             _stream()), "efm_model_lame_adjoint")
    return gvp, gvs, gr


class Plan:
    """Owner of an EfmPlan (efm_plan_create/destroy): the HIP graphs of one shot
    batch's time loops, captured on first use of a set of buffers and replayed
    afterwards."""

    def __init__(self, prop, chunk):
        import weakref
        self.lib = prop.lib
        self.handle = ctypes.c_void_p()
        d, m, sh = prop.dims(), prop.medium(), prop.shots.struct()
        check(self.lib.efm_plan_create(ctypes.byref(d), ctypes.byref(m), ctypes.byref(sh), int(chunk),
                                       ctypes.byref(self.handle)), "efm_plan_create")
        self.chunk = int(chunk)
        self._fin = weakref.finalize(self, self.lib.efm_plan_destroy, self.handle)

    def graphs(self):
        return int(self.lib.efm_plan_graphs(self.handle))


class Propagator:
    """Forward modelling and exact-adjoint gradients for a shot batch in one medium."""

    def __init__(self, grid: Grid, lam, mu, rho, shots: ShotBatch, pool=None, graphs=False):
        """pool: an optional list of wavefields of this grid and shot count,
        shared with other propagators (the forward states of a gradient are
        taken from it and it grows as needed), so that repeated gradients do not
        allocate and clear them again.  graphs: run the time loops of records
        and gradients from a persistent plan of HIP graphs (captured on the
        first gradient, replayed by later ones; the workspaces the graphs use
        are then kept by the propagator between calls)."""
        self.graphs = bool(graphs)
        self._plan = None
        self._work = {}
        self.grid, self.shots = grid, shots
        for t in (lam, mu, rho):
            if t.dtype != torch.float32 or not t.is_contiguous() or tuple(t.shape) != grid.shape:
                raise ValueError(f"medium arrays must be contiguous float32 {grid.shape}")
        self.lam, self.mu, self.rho = lam, mu, rho
        self.lib = _capi.lib()
        if pool is not None:
            if any(w.grid != grid or w.S != shots.S or w.device != lam.device for w in pool[:1]):
                raise ValueError("the state pool belongs to another grid, shot count or device")
            self._pool = pool

    # -- descriptors --------------------------------------------------------
    def dims(self, sign=1.0):
        g, s = self.grid, self.shots
        w = g.widths
        return EfmDims(g.nx, g.ny, g.nz, s.S, w[0], w[1], w[2], w[3], w[4], w[5], s.ntmax, s.nrec,
                       float(sign * g.dt / g.dx))

    def medium(self):
        return EfmMedium(_ptr(self.lam), _ptr(self.mu), _ptr(self.rho))

    def new_wavefield(self):
        return Wavefield(self.grid, self.shots.S, self.lam.device)

    def new_gathers(self):
        s = self.shots
        return torch.zeros((s.S, 3, s.ntmax, s.nrec), dtype=torch.float32, device=self.lam.device)

    # -- forward ------------------------------------------------------------
    def forward(self, state=None, gathers=None, it0=0, nsteps=None, borders=None, W=0,
                graph_chunk=None):
        """Advance ``state`` (zeroed if None) through steps it0 .. it0+nsteps-1.
        ``graph_chunk`` = K replays one captured hipGraph of K steps (from a
        persistent plan of this propagator, so later calls on the same buffers
        capture nothing) instead of launching every kernel (no border save in
        that mode); so does every forward of a propagator made with graphs=True."""
        st = self.new_wavefield() if state is None else state
        g = self.new_gathers() if gathers is None else gathers
        n = self.shots.ntmax - it0 if nsteps is None else nsteps
        d, m, sh = self.dims(), self.medium(), self.shots.struct()
        if borders is None and (graph_chunk or self.graphs):
            plan = self.graph_plan(int(graph_chunk) if graph_chunk else None)
            check(self.lib.efm_plan_forward(plan.handle, ctypes.byref(st.struct()), _ptr(g), int(it0),
                                            int(n), _stream()), "efm_plan_forward")
            return g, st
        check(self.lib.efm_forward(ctypes.byref(d), ctypes.byref(m), ctypes.byref(sh),
                                   ctypes.byref(st.struct()), _ptr(g), _ptr(borders), int(W),
                                   int(it0), int(n), _stream()), "efm_forward")
        return g, st

    def graph_plan(self, chunk=None):
        """The propagator's persistent plan whose forward graphs cover `chunk`
        steps (default: the checkpoint interval, the records' segment length)."""
        chunk = self.checkpoint_interval() if chunk is None else int(chunk)
        if self._plan is None or self._plan.chunk != chunk:
            plans = self.__dict__.setdefault("_plans", {})
            if chunk not in plans:
                plans[chunk] = Plan(self, chunk)
            self._plan = plans[chunk]
        return self._plan

    def _kept(self, name, make):
        """A workspace kept between calls when the time loops run from graphs
        (a captured graph holds its buffers' addresses); else a new one."""
        if not self.graphs:
            return make()
        t = self._work.get(name)
        if t is None:
            t = self._work[name] = make()
        return t

    def _forward_states(self, arr, g, it0, nsteps):
        d, md, sh = self.dims(), self.medium(), self.shots.struct()
        if self.graphs:
            check(self.lib.efm_plan_forward_states(self.graph_plan().handle, arr, _ptr(g), int(it0), int(nsteps),
                                                   _stream()), "efm_plan_forward_states")
        else:
            check(self.lib.efm_forward_states(ctypes.byref(d), ctypes.byref(md), ctypes.byref(sh), arr, _ptr(g),
                                              int(it0), int(nsteps), _stream()), "efm_forward_states")

    def border_buffer(self, W):
        d = self.dims()
        per = self.lib.efm_border_floats(ctypes.byref(d), int(W))
        return torch.zeros((self.shots.S, self.shots.ntmax, per), dtype=torch.float32,
                           device=self.lam.device)

    # -- exact adjoint: tape (checkpoints or every state) + reverse sweep ----------
    def checkpoint_interval(self):
        return max(1, int(math.ceil(math.sqrt(self.shots.ntmax))))

    def _psi_bytes(self):
        g, S = self.grid, self.shots.S
        Lx, Ly, Lz = g.slab_lengths
        ny, nz, nx = g.shape
        return 4 * 6 * S * (ny * nz * Lx + Ly * nz * nx + ny * Lz * nx)

    def full_state_bytes(self):
        """Bytes of a full state (Wavefield: 9 volumes + ψ), as a checkpoint
        snapshot takes."""
        return 4 * 9 * (self.shots.S * self.grid.N + FIELD_PAD) + self._psi_bytes()

    def state_bytes(self):
        """Bytes of one kept forward state of a tape (TapeState: 3 volumes + ψ;
        the tape's stresses are one shared buffer of 6 volumes)."""
        return 4 * 3 * (self.shots.S * self.grid.N + FIELD_PAD) + self._psi_bytes()

    def _tau_bytes(self):
        return 4 * 6 * (self.shots.S * self.grid.N + FIELD_PAD)

    def choose_strategy(self, budget=None):
        """'store' keeps every forward state (no re-forward; (nt+1) states of
        HBM); 'hybrid' keeps the states of the last m steps and √-checkpoints
        the first nt-m (only those are run twice); 'checkpoint' keeps √nt
        snapshots and rebuilds each segment."""
        return self.plan(budget)[0]

    def plan(self, budget=None):
        """(strategy, m): m = number of final steps whose states are kept."""
        nt, sb, sf, tb = self.shots.ntmax, self.state_bytes(), self.full_state_bytes(), self._tau_bytes()
        have = getattr(self, "_pool", [])
        if len(have) >= nt + 1:
            return "store", nt
        if budget is None:
            # states already in the pool (and their shared stresses) are reused, so they count
            # as available too
            budget = 0.85 * (available_bytes(self.lam.device) + len(have) * sb + (tb if have else 0))
        if (nt + 1) * sb + tb <= budget:
            return "store", nt
        # hybrid: (m + 1) kept states + the full snapshots of the first nt - m steps
        # (the segment buffers reuse the kept states once their sweep is done),
        # with room left for the adjoint state, ψ ping-pong and accumulators (4 full states)
        room = budget - tb - 4 * sf
        m = int(room // sb)
        while m > 0:
            ncp = nt - m
            K = max(1, math.isqrt(max(ncp - 1, 0)) + 1)
            if (m + 1) * sb + -(-ncp // K) * sf <= room and m >= K:
                return "hybrid", m
            m -= max(1, m // 64)
        return "checkpoint", 0

    def _states(self, n):
        """At least n tape states (zeroed when allocated, sharing one stress
        buffer), allocated once per pool and reused (kernels write only interior
        cells, so the zero borders survive)."""
        pool = getattr(self, "_pool", None)
        if pool is None:
            pool = self._pool = []
        tau = pool[0].tau if pool else None
        while len(pool) < n:
            if tau is None:
                tau = _padded_fields(6, self.grid, self.shots.S, self.lam.device)
            pool.append(TapeState(self.grid, self.shots.S, self.lam.device, tau))
        return pool[:n]

    def release_states(self):
        """Drop this propagator's hold on its states, workspaces and graphs (a
        shared pool stays with its owner)."""
        self._pool = []
        self._arrays = {}
        self._work = {}
        self._plan = None
        self._plans = {}

    def _array(self, states):
        """ctypes array of EfmState pointers for `states` (cached per list of
        wavefields: building it for thousands of states costs milliseconds)."""
        key = (len(states),) + tuple(id(w) for w in states[:1] + states[-1:])
        cache = getattr(self, "_arrays", None)
        if cache is None:
            cache = self._arrays = {}
        hit = cache.get(key)
        if hit is not None and hit[0] == [id(w) for w in states]:
            return hit[1]
        arr = (ctypes.POINTER(EfmState) * len(states))(*[ctypes.pointer(w.struct()) for w in states])
        cache[key] = ([id(w) for w in states], arr, states)
        return arr

    def record(self, K=None, strategy="auto", m=None):
        """Forward run that keeps what the adjoint needs.  Returns a tape dict.
        K: checkpoint interval; m: steps kept by a hybrid tape (default: as many
        as fit).  With strategy "auto", a tape the plan sized but the device
        cannot hold after all (memory held elsewhere in the process, allocator
        fragmentation) falls back to √nt checkpoints instead of failing."""
        if strategy != "auto":
            return self._record(K, strategy, m)
        planned, m_fit = self.plan()
        if planned == "hybrid" and m is None:
            m = m_fit
        held = len(getattr(self, "_pool", None) or [])
        try:
            return self._record(K, planned, m)
        except torch.OutOfMemoryError:
            if planned == "checkpoint":
                raise
            pool = getattr(self, "_pool", None)
            if pool is not None:
                del pool[held:]                  # give back the states this call added
            if torch.cuda.is_initialized():
                torch.cuda.empty_cache()
            return self._record(K, "checkpoint", None)

    def _record(self, K, strategy, m):
        nt = self.shots.ntmax
        g = self._kept("gathers", self.new_gathers)
        if self.graphs:
            g.zero_()
        if strategy == "store":
            states = self._states(nt + 1)
            states[0].zero_()
            self._forward_states(self._array(states), g, 0, nt)
            return dict(strategy="store", gathers=g, states=states)
        if strategy == "hybrid":
            if m is None:
                planned, m = self.plan()
                if planned != "hybrid":            # asked for explicitly: keep half
                    m = nt // 2
            m = int(m)
            if not 0 < m < nt:
                raise ValueError("a hybrid tape keeps between 1 and nt - 1 steps")
            ncp = nt - m
            K = max(1, math.isqrt(max(ncp - 1, 0)) + 1) if K is None else int(K)
            if m < min(K, ncp):
                raise ValueError("a hybrid tape keeps at least one checkpoint segment of states")
            st = self._kept("record_state", self.new_wavefield)
            st.zero_()
            snaps = []
            for n0 in range(0, ncp, K):
                snaps.append(st.clone())
                self.forward(st, g, it0=n0, nsteps=min(K, ncp - n0))
            states = self._states(m + 1)
            states[0].copy_(st)
            del st
            self._forward_states(self._array(states[:m + 1]), g, ncp, m)
            return dict(strategy="hybrid", gathers=g, states=states[:m + 1], snaps=snaps, K=K, ncp=ncp)
        K = self.checkpoint_interval() if K is None else int(K)
        st = self._kept("record_state", self.new_wavefield)
        st.zero_()
        snaps = []
        for n0 in range(0, nt, K):
            snaps.append(st.clone())
            self.forward(st, g, it0=n0, nsteps=min(K, nt - n0))
        return dict(strategy="checkpoint", gathers=g, snaps=snaps, K=K)

    def misfit(self, gathers, obs, kind="l2"):
        """The misfit J and its residual ∂J/∂gathers on the device: ``l2``
        r = d - obs, J = ½‖r‖² (efm_residual, M:2128-2142); ``xcorr`` the
        per-trace normalised cross-correlation J = Σ (1 - ⟨d, obs⟩/(|d||obs|))
        and the reference's residual for it (efm_xcorr_residual, M:2145-2179)."""
        obs = obs.to(device=self.lam.device, dtype=torch.float32).contiguous()
        if obs.shape != gathers.shape:
            raise ValueError(f"observed data must be {tuple(gathers.shape)}")
        r = self._kept("resid", lambda: torch.empty_like(gathers))
        J = torch.zeros(1, dtype=torch.float64, device=self.lam.device)
        if kind == "l2":
            check(self.lib.efm_residual(_ptr(gathers), _ptr(obs), _ptr(r), _ptr(J), gathers.numel(),
                                        _stream()), "efm_residual")
        elif kind == "xcorr":
            S, _, nt, nrec = gathers.shape
            check(self.lib.efm_xcorr_residual(_ptr(gathers), _ptr(obs), _ptr(r), _ptr(J), S, nt, nrec,
                                              _stream()), "efm_xcorr_residual")
        else:
            raise ValueError(f"unknown misfit {kind!r} (l2 or xcorr)")
        return r, J

    def backprop(self, tape, resid):
        """∂J/∂λ, ∂J/∂μ, ∂J/∂ρ (volumes) and ∂J/∂M (S, 6) float64 for
        ∂J/∂gathers = resid (S, 3, ntmax, nrec)."""
        s, g = self.shots, self.grid
        dev = self.lam.device
        resid = resid.to(device=dev, dtype=torch.float32).contiguous()
        if tuple(resid.shape) != (s.S, 3, s.ntmax, s.nrec):
            raise ValueError(f"residual must be {(s.S, 3, s.ntmax, s.nrec)}")
        d, m, sh = self.dims(), self.medium(), s.struct()
        if self.graphs:
            kept = self._kept("resid", lambda: torch.empty_like(resid))
            if kept.data_ptr() != resid.data_ptr():
                kept.copy_(resid)
            resid = kept
        mbar = self._kept("mbar", lambda: padded_components(3, g.N, dev))
        bbar = self._kept("bbar", lambda: padded_components(3, g.N, dev))
        check(self.lib.efm_derived_medium(ctypes.byref(d), ctypes.byref(m), _ptr(mbar), _ptr(bbar),
                                          _stream()), "efm_derived_medium")
        nt = s.ntmax
        adj = self._kept("adj", self.new_wavefield)
        adj2 = self._kept("adj2", lambda: Wavefield(g, s.S, dev, psi_only_of=adj))
        acc = self._kept("acc", lambda: padded_components(s.S * 8, g.N, dev))
        gm6 = self._kept("gm6", lambda: torch.zeros((s.S, 6), dtype=torch.float64, device=dev))
        if self.graphs:                          # kept workspaces start each gradient at zero
            adj.zero_(); adj2.psi_x.zero_(); adj2.psi_y.zero_(); adj2.psi_z.zero_()
            # the whole buffer, padding included: the interleaved accumulators run across the
            # component strides of the padded view
            torch.as_strided(acc, (acc.shape[0] * acc.stride(0),), (1,)).zero_()
            gm6.zero_()

        def sweep(arr, n0, n1):
            if self.graphs:
                check(self.lib.efm_plan_adjoint(self.graph_plan().handle, _ptr(mbar), _ptr(bbar), arr,
                                                ctypes.byref(adj.struct()), ctypes.byref(adj2.struct()),
                                                _ptr(resid), _ptr(acc), _ptr(gm6), n0, n1, int(n0 == 0),
                                                _stream()), "efm_plan_adjoint")
                return
            check(self.lib.efm_adjoint(ctypes.byref(d), ctypes.byref(m), _ptr(mbar), _ptr(bbar),
                                       ctypes.byref(sh), arr, ctypes.byref(adj.struct()),
                                       ctypes.byref(adj2.struct()), _ptr(resid), _ptr(acc), _ptr(gm6),
                                       n0, n1, int(n0 == 0), _stream()), "efm_adjoint")

        if tape["strategy"] == "store":
            sweep(self._array(tape["states"][:nt + 1]), 0, nt)
        else:
            ncp = tape.get("ncp", nt)            # steps rebuilt from snapshots
            if tape["strategy"] == "hybrid":
                sweep(self._array(tape["states"]), ncp, nt)
            K, snaps = tape["K"], tape["snaps"]
            bufs = self._states(min(K, ncp) + 1)
            for seg in range(len(snaps) - 1, -1, -1):
                n0 = seg * K
                n1 = min(n0 + K, ncp)
                k = n1 - n0
                bufs[0].copy_(snaps[seg])
                arr = self._array(bufs[:k + 1])
                self._forward_states(arr, None, n0, k)
                sweep(arr, n0, n1)
        glam = torch.empty(g.shape, dtype=torch.float32, device=dev)
        gmu = torch.empty_like(glam)
        grho = torch.empty_like(glam)
        check(self.lib.efm_gradient_combine(ctypes.byref(d), ctypes.byref(m), _ptr(mbar), _ptr(bbar),
                                            _ptr(acc), _ptr(glam), _ptr(gmu), _ptr(grho), _stream()),
              "efm_gradient_combine")
        return glam, gmu, grho, (gm6.clone() if self.graphs else gm6)

    def forward_checkpointed(self, K=None):
        """Forward run keeping the pre-step state every K steps."""
        t = self.record(K, strategy="checkpoint")
        return t["gathers"], t["snaps"], t["K"]

    def gradient_from_snapshots(self, snaps, K, resid):
        return self.backprop(dict(strategy="checkpoint", snaps=snaps, K=K), resid)

    def gradient(self, obs, K=None, strategy="auto", m=None, misfit="l2"):
        """Misfit J (½‖d - obs‖², or the cross-correlation misfit with
        misfit='xcorr') and its exact gradient (the FWI gradient)."""
        tape = self.record(K, strategy, m)
        resid, J = self.misfit(tape["gathers"], obs, misfit)
        self.last_strategy = tape["strategy"]
        grads = self.backprop(tape, resid)       # queued before J is read back (no host sync)
        g = tape["gathers"].clone() if self.graphs else tape["gathers"]
        return (float(J), g) + grads
