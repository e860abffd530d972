"""Drop-in driver layer with the classes and methods of mainPropElastic3D_MT.py.

Each class keeps the reference's name, constructor arguments, attribute names
and array layouts, so code written against the reference runs unchanged; the
device work goes through libefm.so.  The reference's module globals (nx, ny, nz,
fq, nSismos, SrcIdx, threadBlock*, rank/comm, H_STENCIL) are explicit here:

* grid sizes come from the Model passed to Elastic3D;
* the Ricker exponent uses the wavelet's own frequency (the reference reads
  the global ``fq``, equal to it in every reference run, M:1270);
* the shot loop of the FWI driver runs every earthquake of the geometry on the
  one GPU (the reference deals them round-robin to MPI ranks, M:2383);
* block shapes are the autotuner's business (the reference hard-codes
  block=(32,2,4), M:2764-2766).

Plotting methods (graficar, save*Png, saveGif) are not provided: they are
matplotlib display code, not numerics.
"""
from __future__ import annotations

import copy
import ctypes
import os

import numpy as np
import torch

from . import _capi, conditioning as _C, model as _M
from .lbfgs import L_BFGS, compute_sk_yk
from .propagator import Grid, Propagator, ShotBatch, Wavefield, _ptr, _stream

H_STENCIL = 2


def steepest_alphas(lam, mu, rho, tl, tm, tr):
    """Steps of the fixed-α iterations of fwi_Vp_Vs_Rho (M:2669-2671): each block
    moves by a set fraction of its model's 2-norm at the image's maximum,
    α_λ = 1e-5·‖λ‖₂/max(tλ), α_μ = 1e-4·‖μ‖₂/max(tμ), α_ρ = 1e-6·‖ρ‖₂/max(tρ)
    (α_ρ is reported but ρ follows Vp, M:2679).  Returns (α_λ, α_μ, α_ρ)."""
    def a(frac, model, image):
        return float(frac * np.sqrt(np.sum(np.ravel(model).astype(np.float64) ** 2)) / np.amax(image))
    return a(1e-5, lam, tl), a(1e-4, mu, tm), a(1e-6, rho, tr)


def _device(device=None):
    return torch.device(device if device is not None else "cuda")


# ----------------------------------------------------------------------------
# geometry (M:536-731)
# ----------------------------------------------------------------------------
class Station:
    def __init__(self, x, y, z):
        self.x, self.y, self.z = x, y, z

    def __str__(self):
        return f"x:{self.x} y:{self.y}, z:{self.z}"

    def getX(self): return self.x
    def getY(self): return self.y
    def getZ(self): return self.z
    def setX(self, x): self.x = x
    def setY(self, y): self.y = y
    def setZ(self, z): self.z = z


class Earthquake:
    def __init__(self, x, y, z):
        self.x, self.y, self.z = x, y, z
        self.stations = []
        self.mag = None

    def __str__(self):
        return f"x:{self.x} y:{self.y}, z:{self.z}"

    def addStation(self, newStation): self.stations.append(newStation)
    def getLenStations(self): return len(self.stations)
    def getX(self): return self.x
    def getY(self): return self.y
    def getZ(self): return self.z
    def getMag(self): return self.mag
    def setX(self, v): self.x = v
    def setY(self, v): self.y = v
    def setZ(self, v): self.z = v
    def setMag(self, v): self.mag = v

    def printStationsSismo(self):
        print(f"{30 * '-'}\n{'# estacion':<10}| {'Coordenadas:':<20}\n{30 * '-'}")
        for i, st in enumerate(self.stations):
            print(f"{i + 1:< 10}| {st.__str__():<20}\n{30 * '-'}")


class Geometry:
    def __init__(self):
        self.sismos = []

    def addSismo(self, newSismo): self.sismos.append(newSismo)
    def getLenSismos(self): return len(self.sismos)

    def getStationsEarthquake(self, idxSismo):
        st = self.sismos[idxSismo].stations
        return (np.array([s.getX() for s in st], np.float64), np.array([s.getY() for s in st], np.float64),
                np.array([s.getZ() for s in st], np.float64))

    def printSismos(self):
        print(f"{30 * '-'}\n{'# sismo':<10}| {'Coordenadas:':<20}\n{30 * '-'}")
        for i, sm in enumerate(self.sismos):
            print(f"{i + 1:< 10}| {sm.__str__():<20}\n{30 * '-'}")

    def printStations(self, idxSismo=0):
        self.sismos[idxSismo].printStationsSismo()

    def save(self, directory="Results/Actual/Geometry"):
        """Geometry.save (M:698-731): sources and the concatenated per-quake station lists."""
        os.makedirs(directory, exist_ok=True)
        sx = np.array([s.getX() for s in self.sismos], np.float64)
        sy = np.array([s.getY() for s in self.sismos], np.float64)
        sz = np.array([s.getZ() for s in self.sismos], np.float64)
        st = [st for s in self.sismos for st in s.stations]
        np.save(os.path.join(directory, "sismos_pos_x.npy"), sx)
        np.save(os.path.join(directory, "sismos_pos_y.npy"), sy)
        np.save(os.path.join(directory, "sismos_pos_z.npy"), sz)
        np.save(os.path.join(directory, "sta_pos_x.npy"), np.array([s.getX() for s in st], np.float64))
        np.save(os.path.join(directory, "sta_pos_y.npy"), np.array([s.getY() for s in st], np.float64))
        np.save(os.path.join(directory, "sta_pos_z.npy"), np.array([s.getZ() for s in st], np.float64))


def station_grid(nx, ny, lenPML, nRxLine, nLinesRx, z=2):
    """Receiver layout of the reference run (M:2851-2861): x = linspace(W, nx-W-1,
    nRxLine) inner, y = linspace(W, ny-W-1, nLinesRx) outer, all at depth z."""
    xx = np.linspace(lenPML, nx - lenPML - 1, nRxLine).astype(int)
    yy = np.linspace(lenPML, ny - lenPML - 1, nLinesRx).astype(int)
    return [(int(x), int(y), z) for y in yy for x in xx]


# ----------------------------------------------------------------------------
# model, wavelet, CPML, containers (M:1132-1429)
# ----------------------------------------------------------------------------
class Model:
    def __init__(self, nz, ny, nx, dz, dy, dx):
        shape = (ny, nz, nx)
        self.Vp = np.zeros(shape, np.float32)
        self.Vs = np.zeros(shape, np.float32)
        self.Rho = np.zeros(shape, np.float32)
        self.Mu = np.zeros(shape, np.float32)
        self.Lambda = np.zeros(shape, np.float32)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dz, self.dy, self.dx = dz, dy, dx
        self.d_Vp = self.d_Vs = self.d_Rho = self.d_Mu = self.d_Lambda = None

    def compute_Lambda_Mu(self):
        self.Lambda, self.Mu = _M.lame_from_velocity(self.Vp, self.Vs, self.Rho)

    def compute_Vp_Vs(self):
        self.Vp, self.Vs = _M.velocity_from_lame(self.Lambda, self.Mu, self.Rho)

    def compute_Rho(self):
        self.Rho = _M.rho_nafe_drake(self.Vp)

    def MaxMinVp(self):
        self.VmaxVp = np.amax(self.Vp)
        self.VminVp = np.amin(self.Vp)

    def hostToDevice(self, device=None):
        dev = _device(device)
        f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)
        self.d_Vp, self.d_Vs, self.d_Rho = f(self.Vp), f(self.Vs), f(self.Rho)
        self.d_Mu, self.d_Lambda = f(self.Mu), f(self.Lambda)


class Wavelet:
    """Wavelet (M:1249-1262); dt defaults to the reference's hard-coded value."""

    def __init__(self, fq, tprop, cn=None, dt=_M.REF_DT):
        self.tprop, self.fq, self.dt = tprop, fq, dt
        self.nSamples = int(self.tprop / self.dt)
        self.waveform = np.zeros((self.nSamples, 1), np.float32)
        self.t = (np.arange(self.nSamples) * self.dt).astype(np.float32).reshape(-1, 1)
        self.d_waveform = None

    def hostToDevice(self, device=None):
        self.d_waveform = torch.as_tensor(self.waveform.astype(np.float32), device=_device(device))

    def ricker(self):
        self.waveform = _M.ricker(self.fq, self.dt, self.nSamples).reshape(-1, 1)

    def gaussian(self):
        self.waveform = _M.gaussian(self.fq, self.dt, self.nSamples).reshape(-1, 1)

    def __str__(self):
        return (f"---- Parametros Ondicula ---- \nSamples: {self.nSamples}\n"
                f"Sample time: {self.dt:e} [s] \nSimulation time: {self.tprop} [s]")


class CPML:
    """CPML (M:1298-1307): profiles from the wavelet frequency, dt and the
    model's maximum Vp.  ``free_top`` keeps the reference's free z-min face."""

    def __init__(self, lenPML, ondicula, cn, free_top=True):
        self.lenPML = lenPML
        self.free_top = free_top
        vmax = getattr(cn, "VmaxVp", None)
        vmax = float(np.amax(cn.Vp)) if vmax is None else float(vmax)
        c = _M.cpml_coefficients(lenPML, ondicula.fq, ondicula.dt, cn.dx, vmax, cn.nx, cn.ny, cn.nz,
                                 free_top=free_top, dy=cn.dy, dz=cn.dz)
        self.ax, self.bx, self.ay, self.by, self.az, self.bz = (c[k] for k in ("ax", "bx", "ay", "by", "az", "bz"))

    def profiles(self):
        return dict(ax=self.ax, bx=self.bx, ay=self.ay, by=self.by, az=self.az, bz=self.bz)


class StressTensor:
    """Mxx, Myy, Mzz, Mxy, Mxz, Myz (M:1374-1389)."""

    def __init__(self, Mxx, Myy, Mzz, Mxy, Mxz, Myz):
        self.xx, self.yy, self.zz, self.xy, self.xz, self.yz = Mxx, Myy, Mzz, Mxy, Mxz, Myz

    def as_array(self):
        return np.array([self.xx, self.yy, self.zz, self.xy, self.xz, self.yz], np.float64)

    def __str__(self):
        return (f"Mxx: {self.xx:.3f} Myy: {self.yy:.3f} Mzz: {self.zz:.3f} "
                f"Mxy: {self.xy:.3f} Mxz: {self.xz:.3f} Myz: {self.yz:.3f}")


class Gather:
    """(nSamples, nStations) float32 traces of vx, vy, vz (M:1132-1155)."""

    def __init__(self, nSamples, nStations):
        self.nSamples, self.nStations = nSamples, nStations
        self.setZero()

    def setZero(self):
        z = lambda: np.zeros((self.nSamples, self.nStations), np.float32)
        self.Vx, self.Vy, self.Vz = z(), z(), z()

    def stacked(self):
        return np.stack([self.Vx, self.Vy, self.Vz]).astype(np.float32)

    def hostToDevice(self, device=None):
        self.d = torch.as_tensor(self.stacked()[None], device=_device(device))
        return self.d


class Gradients:
    def __init__(self, mode, ny, nz, nx):
        self.setZero(mode, ny, nz, nx)

    def setZero(self, mode, ny, nz, nx):
        self.Mxx = self.Myy = self.Mzz = self.Mxy = self.Mxz = self.Myz = 0.0
        if mode == 1:
            self.Lambda = np.zeros((ny, nz, nx), np.float32)
            self.Mu = np.zeros((ny, nz, nx), np.float32)
            self.Rho = np.zeros((ny, nz, nx), np.float32)

    def __str__(self):
        return (f"gk_Mxx: {self.Mxx:.2f} gk_Myy: {self.Myy:.2f} gk_Mzz: {self.Mzz:.2f}"
                f"gk_Mxy: {self.Mxy:.2f} gk_Mxz: {self.Mxz:.2f} gk_Myz: {self.Myz:.2f}")


class Alpha:
    def __init__(self):
        self.Mxx = self.Myy = self.Mzz = self.Mxy = self.Mxz = self.Myz = 0.0

    def __str__(self):
        return (f"alpha.Mxx: {self.Mxx:.4f} alpha.Myy: {self.Myy:.4f} alpha.Mzz: {self.Mzz:.4f} \n"
                f"alpha.Mxy: {self.Mxy:.4f} alpha.Mxz: {self.Mxz:.4f} alpha.Myz: {self.Myz:.4f}")


# ----------------------------------------------------------------------------
# Elastic3D (M:1432-2719)
# ----------------------------------------------------------------------------
class Elastic3D:
    """mode 0: MT gradient runs; mode 1: λ/μ/ρ gradients with boundary saving."""

    def __init__(self, d_mk: Model, geometria: Geometry, ondicula: Wavelet, pml: CPML, mode=0,
                 device=None, results_dir=None):
        self.device = _device(device)
        self.mk = Model(d_mk.nz, d_mk.ny, d_mk.nx, d_mk.dz, d_mk.dy, d_mk.dx)
        for k in ("Mu", "Lambda", "Rho", "Vp", "Vs"):
            setattr(self.mk, k, getattr(d_mk, k).copy())
        self.nx, self.ny, self.nz = d_mk.nx, d_mk.ny, d_mk.nz
        self.dx, self.dy, self.dz = d_mk.dx, d_mk.dy, d_mk.dz
        self.geometria, self.ondicula, self.cpml = geometria, ondicula, pml
        self.nt = ondicula.nSamples
        self.lenPML = pml.lenPML
        self.mode = mode
        self.gk = Gradients(mode, self.ny, self.nz, self.nx)
        self.alpha = Alpha()
        self.costFunction = []
        self.results_dir = results_dir
        W = pml.lenPML
        self.grid = Grid(self.nx, self.ny, self.nz, self.dx, ondicula.dt,
                         (W, W, W, W, 0 if pml.free_top else W, W))
        self.M = StressTensor(0, 0, 0, 0, 0, 0)
        self.iteration = 0
        self._gk_device = None
        self.forwFields = None
        self._borders = None
        self._shots = None
        self._state_pool = []     # forward states of exact gradients, reused across calls
        # True: drive the reference-style time loops from Python one step at a time, with the
        # reference's per-step host transfers (its structure, M:1823, 1939-2054), so that the
        # plain path can be timed as the reference runs it; the results equal the C loops'
        self.host_loop = False
        self.mk.hostToDevice(self.device)

    # -- configuration ---------------------------------------------------------
    def setMode(self, mode):
        self.mode = mode
        self.gk = Gradients(mode, self.ny, self.nz, self.nx)

    def setIdxEarthquake(self, SrxIdx):
        """M:1467-1489 (the reference reads the global SrcIdx for the station count;
        here everything comes from SrxIdx)."""
        self.SrxIdx = SrxIdx
        eq = self.geometria.sismos[SrxIdx]
        self.LenStations = eq.getLenStations()
        for name in ("gather", "obsGather", "modGather", "resGather"):
            setattr(self, name, Gather(self.nt, self.LenStations))
        self.rec_pos_x, self.rec_pos_y, self.rec_pos_z = self.geometria.getStationsEarthquake(SrxIdx)
        self.Sx, self.Sy, self.Sz = int(eq.x), int(eq.y), int(eq.z)
        self._build_shots()

    def setStressTensor(self, M):
        self.M = copy.copy(M)
        if self._shots is not None:
            self._shots.set_m6(self.M.as_array()[None])

    def setModel(self, model):
        for k in ("Mu", "Lambda", "Rho", "Vp", "Vs"):
            setattr(self.mk, k, getattr(model, k).copy())
        self.mk.hostToDevice(self.device)

    def getModel(self):
        return self.mk

    def _build_shots(self):
        rec = np.stack([self.rec_pos_x, self.rec_pos_y, self.rec_pos_z], 1).astype(np.int64)
        w = np.asarray(self.ondicula.waveform, np.float32).reshape(1, -1)
        self._shots = ShotBatch(self.grid, [(self.Sx, self.Sy, self.Sz)], self.M.as_array()[None], w,
                                [self.nt], rec[None], self.cpml.profiles(), device=self.device)

    def release_states(self):
        """Free the forward states kept for exact gradients (no reference counterpart)."""
        self._state_pool.clear()

    def _propagator(self):
        return Propagator(self.grid, self.mk.d_Lambda, self.mk.d_Mu, self.mk.d_Rho, self._shots,
                          pool=self._state_pool)

    # -- forward (M:1812-1848) ---------------------------------------------------
    def forwardFields(self, strname="mod", saveFieldFlag=False):
        prop = self._propagator()
        self.gather.setZero()
        borders = None
        if strname == "mod" and self.mode == 1:
            borders = prop.border_buffer(self.lenPML)
        if self.host_loop:
            g, st = self._forward_per_step(prop, borders)
        else:
            g, st = prop.forward(borders=borders, W=self.lenPML)
        self.forwFields, self._borders = st, borders
        host = g[0].cpu().numpy()
        self.gather.Vx, self.gather.Vy, self.gather.Vz = host[0].copy(), host[1].copy(), host[2].copy()
        if strname == "mod":
            self.modGather.Vx, self.modGather.Vy, self.modGather.Vz = self.getGathers()
        if strname == "obs":
            self.obsGather.Vx, self.obsGather.Vy, self.obsGather.Vz = self.getGathers()

    def _forward_per_step(self, prop, borders):
        """The forward loop as the reference drives it (M:1823-1848): one host
        call per time step (stress, source, velocity, stations, border save)."""
        L = _capi.lib()
        st, g = prop.new_wavefield(), prop.new_gathers()
        d, m, sh, ss = prop.dims(), prop.medium(), self._shots.struct(), st.struct()
        by, gp, bp, W, hs = ctypes.byref, _ptr(g), _ptr(borders), int(self.lenPML), _stream()
        for it in range(self.nt):
            _capi.check(L.efm_forward(by(d), by(m), by(sh), by(ss), gp, bp, W, it, 1, hs), "efm_forward")
        return g, st

    def getGathers(self):
        return self.gather.Vx.copy(), self.gather.Vy.copy(), self.gather.Vz.copy()

    # -- misfit (M:2128-2142) ----------------------------------------------------
    def computeResidual(self):
        self.resGather.Vx = self.modGather.Vx - self.obsGather.Vx
        self.resGather.Vy = self.modGather.Vy - self.obsGather.Vy
        self.resGather.Vz = self.modGather.Vz - self.obsGather.Vz
        self.resGather.hostToDevice(self.device)

    def computeCostFunction(self, iter=None):
        r = self.resGather
        self.totalCostFunction = 0.5 * (np.sum(r.Vx ** 2) + np.sum(r.Vy ** 2) + np.sum(r.Vz ** 2))
        self.costFunction.append(float(self.totalCostFunction))

    # -- cross-correlation misfit (M:2145-2206) ---------------------------------------
    def computeResidualCrossCorrelation(self):
        """Per-trace normalised cross-correlation residual (M:2145-2179) on the
        device (efm_xcorr_residual); resGather gets it, as computeResidual's."""
        mod = torch.as_tensor(self.modGather.stacked()[None], device=self.device)
        obs = torch.as_tensor(self.obsGather.stacked()[None], device=self.device)
        r, J = self._propagator().misfit(mod, obs, "xcorr")
        host = r[0].cpu().numpy()
        self.resGather.Vx, self.resGather.Vy, self.resGather.Vz = host[0].copy(), host[1].copy(), host[2].copy()
        self.resGather.hostToDevice(self.device)
        self._xcorr_trace_cost = float(J)

    def computeCostFunctionCrossCorrelation(self, iter=None):
        """M:2181-2206: Σ over components of 1 - ⟨mod, obs⟩/(|mod||obs|) over the
        whole gathers (zero norms replaced by 1), appended to costFunction."""
        total = 0.0
        for c in ("Vx", "Vy", "Vz"):
            m = getattr(self.modGather, c).astype(np.float64)
            o = getattr(self.obsGather, c).astype(np.float64)
            nm = float(np.sqrt(np.sum(m * m))) or 1.0
            no = float(np.sqrt(np.sum(o * o))) or 1.0
            total += 1.0 - float(np.sum(m * o)) / (nm * no)
        self.totalCostFunction = total
        self.costFunction.append(float(total))

    # -- reference-style adjoints -------------------------------------------------
    def _mt_image(self, nsteps):
        prop = self._propagator()
        L = _capi.lib()
        adj = prop.new_wavefield()
        gm6 = torch.zeros((1, 6), dtype=torch.float64, device=self.device)
        res = self.resGather.hostToDevice(self.device)
        d, m, sh = prop.dims(), prop.medium(), self._shots.struct()
        _capi.check(L.efm_backward_mt_ref(ctypes.byref(d), ctypes.byref(m), ctypes.byref(sh),
                                          ctypes.byref(adj.struct()), _ptr(res), _ptr(gm6), int(nsteps),
                                          _stream()), "efm_backward_mt_ref")
        g = gm6[0].cpu().numpy()
        self.gk.Mxx, self.gk.Myy, self.gk.Mzz, self.gk.Mxy, self.gk.Mxz, self.gk.Myz = (float(v) for v in g)

    def backwardFields(self, saveFieldFlag=False):
        """M:2056-2082.  Mode 0 (its only use in the reference) forms the MT image."""
        self.gk.setZero(self.mode, self.ny, self.nz, self.nx)
        if self.mode != 0:
            raise NotImplementedError("backwardFields images λ/μ/ρ only through "
                                      "backwardFields_FowardFieldsRemake (mode 1)")
        self._mt_image(self.nt)

    def backwardFields_FowardFieldsRemake(self, saveFieldFlag=False):
        """M:1929-2054: adjoint with the forward operator, reverse-time
        reconstruction from the saved boundaries, imaging (nt-1 iterations)."""
        self.gk.setZero(self.mode, self.ny, self.nz, self.nx)
        if self.mode == 0:
            self._mt_image(self.nt - 1)
            return
        if self.forwFields is None or self._borders is None:
            raise RuntimeError("forwardFields('mod') in mode 1 must run first (it saves the boundaries)")
        prop = self._propagator()
        L = _capi.lib()
        adj = prop.new_wavefield()
        S = 1
        vpres = torch.zeros((3 * S * self.grid.N,), dtype=torch.float32, device=self.device)
        glam = torch.zeros(self.grid.shape, dtype=torch.float32, device=self.device)
        gmu, grho = torch.zeros_like(glam), torch.zeros_like(glam)
        res = self.resGather.hostToDevice(self.device)
        d, m, sh = prop.dims(), prop.medium(), self._shots.struct()
        by = ctypes.byref
        args = (by(d), by(m), by(sh), by(self.forwFields.struct()), _ptr(vpres), _ptr(self._borders),
                int(self.lenPML), by(adj.struct()), _ptr(res), float(self.ondicula.dt), float(self.dx),
                float(self.dy), float(self.dz), _ptr(glam), _ptr(gmu), _ptr(grho))
        if self.host_loop:
            # the reference's loop: one host call per iteration, each iteration's three images
            # copied to the host and summed there in float32 (M:1939-2054)
            hl, hm, hr = (np.zeros(self.grid.shape, np.float32) for _ in range(3))
            for it in range(self.nt - 1):
                glam.zero_(); gmu.zero_(); grho.zero_()
                _capi.check(L.efm_backward_remake_ref_steps(*args, it, 1, _stream()),
                            "efm_backward_remake_ref_steps")
                hl += glam.cpu().numpy()
                hm += gmu.cpu().numpy()
                hr += grho.cpu().numpy()
            self.forwFields = None
            self.gk.Lambda, self.gk.Mu, self.gk.Rho = hl, hm, hr
            f = lambda a: torch.as_tensor(a, device=self.device)
            self._gk_device = (f(hl), f(hm), f(hr))
            return
        _capi.check(L.efm_backward_remake_ref(*args, _stream()), "efm_backward_remake_ref")
        self.forwFields = None          # consumed by the reverse-time reconstruction
        self._gk_device = (glam, gmu, grho)
        self.gk.Lambda = glam.cpu().numpy()
        self.gk.Mu = gmu.cpu().numpy()
        self.gk.Rho = grho.cpu().numpy()

    # -- exact gradient (no reference counterpart) ----------------------------------
    def gradient_exact(self, obs=None, K=None, misfit="l2"):
        """Exact ∂J/∂λ, ∂J/∂μ, ∂J/∂ρ and ∂J/∂M of J = ½‖d - obs‖² (or of the
        cross-correlation misfit, misfit='xcorr') for the current earthquake
        (obs defaults to obsGather).  Returns (J, dict)."""
        prop = self._propagator()
        o = (obs if obs is not None else self.obsGather).hostToDevice(self.device)
        J, g, glam, gmu, grho, gm6 = prop.gradient(o, K, misfit=misfit)
        return J, dict(lam=glam, mu=gmu, rho=grho, m6=gm6[0].cpu().numpy(), gathers=g)

    # -- MT inversion (M:2084-2101, 2217-2346) ------------------------------------
    def computeAlpha(self):
        for k in ("Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz"):
            gv = getattr(self.gk, k)
            setattr(self.alpha, k, 0.0 if gv == 0 else 5e-2 / np.sqrt(np.sum(gv ** 2)))

    def updateStressTensor(self):
        a, g = self.alpha, self.gk
        self.M.xx += float(a.Mxx * g.Mxx)
        self.M.yy += float(a.Myy * g.Myy)
        self.M.zz += float(a.Mzz * g.Mzz)
        self.M.xy += float(0.5 * a.Mxy * g.Mxy)
        self.M.xz += float(0.5 * a.Mxz * g.Mxz)
        self.M.yz += float(0.5 * a.Myz * g.Myz)
        self.setStressTensor(self.M)

    def _trial(self):
        a, g, M = self.alpha, self.gk, self.M
        return StressTensor(M.xx + a.Mxx * g.Mxx, M.yy + a.Myy * g.Myy, M.zz + a.Mzz * g.Mzz,
                            M.xy + 0.5 * a.Mxy * g.Mxy, M.xz + 0.5 * a.Mxz * g.Mxz,
                            M.yz + 0.5 * a.Myz * g.Myz)

    def fwi_Mij(self, nIte, flagSave=False, tolFWI=1e-5):
        """MT inversion of the current earthquake: normalised first step (±0.05
        on the diagonal, ±0.025 off it), then trial steps accepted only when the
        misfit decreases, otherwise all α are halved (M:2217-2346)."""
        hist = {k: [float(getattr(self.M, k))] for k in ("xx", "yy", "zz", "xy", "xz", "yz")}
        modData = Gather(self.nt, self.LenStations)
        for it in range(nIte):
            self.setStressTensor(self.M)
            self.forwardFields("mod")
            modData.Vx, modData.Vy, modData.Vz = self.getGathers()
            if it == 0:
                self._save(f"Gathers/modDataIni{self.SrxIdx}", modData)
            self.computeResidual()
            self.computeCostFunction(it)
            self.backwardFields()
            cM = copy.copy(self.M)
            if it == 0:
                self.computeAlpha()
                self.updateStressTensor()
            else:
                trial = self._trial()
                self.setStressTensor(trial)
                self.forwardFields("mod")
                self.computeResidual()
                r = self.resGather
                J = 0.5 * (np.sum(r.Vx ** 2) + np.sum(r.Vy ** 2) + np.sum(r.Vz ** 2))
                if not np.isnan(self.costFunction[it]) and (J - self.costFunction[it]) < 0:
                    self.setStressTensor(trial)
                else:
                    for k in ("Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz"):
                        setattr(self.alpha, k, getattr(self.alpha, k) / 2)
                    self.setStressTensor(cM)
            for k in hist:
                hist[k].append(float(getattr(self.M, k)))
            self._save_mt(hist, modData)
        return hist

    # -- Vp/Vs/ρ FWI (M:2349-2719) -------------------------------------------------
    def _accumulate(self, acc, g, sign=1.0):
        """acc += sign·g on the device (the reference's float32 tempLambda += gk.Lambda,
        M:2466-2468)."""
        _C.axpy(sign, g, acc)

    def _all_shots_gradient(self, data_obs, listaTM, gradient):
        """Sum over earthquakes of the misfit and the λ/μ/ρ descent images
        (device volumes, float32 sums like the reference's temp* arrays)."""
        z = lambda: torch.zeros(self.grid.shape, dtype=torch.float32, device=self.device)
        tl, tm, tr = z(), z(), z()
        total = 0.0
        modData = [None] * self.geometria.getLenSismos()
        for i in range(self.geometria.getLenSismos()):
            self.setIdxEarthquake(i)
            self.setStressTensor(listaTM[i])
            self.obsGather.Vx, self.obsGather.Vy, self.obsGather.Vz = (
                data_obs[i].Vx.copy(), data_obs[i].Vy.copy(), data_obs[i].Vz.copy())
            if gradient == "exact":
                J, gr = self.gradient_exact(self.obsGather, misfit=getattr(self, "_misfit_kind", "l2"))
                total += J
                # the reference's update adds its image (a descent direction)
                self._accumulate(tl, gr["lam"], -1.0)
                self._accumulate(tm, gr["mu"], -1.0)
                self._accumulate(tr, gr["rho"], -1.0)
                host = gr["gathers"][0].cpu().numpy()
                d = Gather(self.nt, self.LenStations)
                d.Vx, d.Vy, d.Vz = host[0], host[1], host[2]
                modData[i] = d
                if self.iteration == 0:
                    self._save("Gathers/modData0", d)      # every shot, one name (M:2414-2417)
                continue
            self.forwardFields("mod")
            d = Gather(self.nt, self.LenStations)
            d.Vx, d.Vy, d.Vz = self.getGathers()
            modData[i] = d
            if self.iteration == 0:
                self._save("Gathers/modData0", d)          # every shot, one name (M:2414-2417)
            total += self._residual_and_cost()
            self.backwardFields_FowardFieldsRemake()
            gl, gm, gr_ = self._gk_device
            self._accumulate(tl, gl); self._accumulate(tm, gm); self._accumulate(tr, gr_)
        return total, tl, tm, tr, modData

    def _condition(self, g, sigma=1.0, top=10):
        """Gradient smoothing and masks (M:2479-2504) on the device: σ=1 Gaussian
        (scipy's defaults: mode='reflect', truncate=4), z < 10 and the CPML+2 zone
        of x, y and the bottom zeroed."""
        return _C.condition_gradient(g, sigma, 4.0, _C.reference_margins(self.nx, self.ny, self.nz,
                                                                           self.lenPML, top))

    def _apply_model(self, lam0, dlam, alpha_lam, mu0, dmu, alpha_mu, cRho):
        """λ = λ0 + α_λ·dλ, μ = μ0 + α_μ·dμ, then compute_Vp_Vs with the current ρ,
        compute_Rho and ρ restored in the top 10 cells (M:2577-2583, 2688-2694),
        on the device; the host copies follow."""
        lam, mu, vp, vs, rho = _C.model_update(lam0, mu0, self.mk.d_Rho, dlam, alpha_lam, dmu, alpha_mu,
                                               rho_keep=cRho, top=10)
        self._set_model_device(lam, mu, vp, vs, rho)

    def _set_model_device(self, lam, mu, vp, vs, rho):
        mk = self.mk
        mk.d_Lambda, mk.d_Mu, mk.d_Vp, mk.d_Vs, mk.d_Rho = lam, mu, vp, vs, rho
        mk.Lambda, mk.Mu, mk.Vp, mk.Vs, mk.Rho = (t.cpu().numpy() for t in (lam, mu, vp, vs, rho))

    def _misfit_all(self, data_obs, listaTM):
        total = 0.0
        for i in range(self.geometria.getLenSismos()):
            self.setIdxEarthquake(i)
            self.setStressTensor(listaTM[i])
            self.obsGather.Vx, self.obsGather.Vy, self.obsGather.Vz = (
                data_obs[i].Vx.copy(), data_obs[i].Vy.copy(), data_obs[i].Vz.copy())
            self.forwardFields("mod")
            total += self._residual_and_cost()
        return total

    def _residual_and_cost(self):
        """resGather and the shot's misfit: ½‖d - obs‖² (M:2128-2142), or the
        per-trace cross-correlation cost with its residual (M:2145-2179)."""
        if getattr(self, "_misfit_kind", "l2") == "xcorr":
            self.computeResidualCrossCorrelation()
            return self._xcorr_trace_cost
        self.computeResidual()
        r = self.resGather
        return 0.5 * (np.sum(r.Vx ** 2) + np.sum(r.Vy ** 2) + np.sum(r.Vz ** 2))

    def fwi_Vp_Vs_Rho(self, data_obs, listaTM, nIter, flagSave=False, tolFWI=1e-5, lbfgsLayers=10,
                      ite_cte=2, gradient="reference", nTry=10, misfit="l2"):
        """λ/μ FWI: per-iteration image over all earthquakes, σ=1 smoothing and
        masks, steepest steps for iter < ite_cte, then L-BFGS with α halving
        from 1 (at most nTry trials); ρ follows Vp through Nafe-Drake.
        ``gradient='exact'`` replaces the reference image by the exact
        (negated) gradient of the same misfit; ``misfit='xcorr'`` uses the
        cross-correlation residual (M:2145-2179) in place of d - obs (the
        per-trace cross-correlation cost is then the one minimised).  Images, conditioning and model
        maps stay on the device; the L-BFGS history is lbfgs.py's host arrays,
        as in the reference."""
        self.costFunction = []
        shape = (lbfgsLayers, self.ny, self.nz, self.nx)
        skL, ykL = np.zeros(shape, np.float32), np.zeros(shape, np.float32)
        skM, ykM = np.zeros(shape, np.float32), np.zeros(shape, np.float32)
        skR, ykR = np.zeros(shape, np.float32), np.zeros(shape, np.float32)
        self._misfit_kind = misfit
        for it in range(nIter):
            self.iteration = it
            total, tl_d, tm_d, tr_d, modData = self._all_shots_gradient(data_obs, listaTM, gradient)
            tr_d, tm_d, tl_d = self._condition(tr_d), self._condition(tm_d), self._condition(tl_d)
            tl, tm, tr = (np.asarray(t.cpu().numpy(), np.float32) for t in (tl_d, tm_d, tr_d))
            self._save_gathers_fq(modData)
            self.costFunction.append(float(total))
            compute_sk_yk(skL, ykL, self.mk.Lambda.copy(), tl, lbfgsLayers, it)
            compute_sk_yk(skM, ykM, self.mk.Mu.copy(), tm, lbfgsLayers, it)
            compute_sk_yk(skR, ykR, self.mk.Rho.copy(), tr, lbfgsLayers, it)
            mk = self.mk
            cL, cM, cR = mk.d_Lambda.clone(), mk.d_Mu.clone(), mk.d_Rho.clone()
            saved = (cL, cM, mk.d_Vp.clone(), mk.d_Vs.clone(), cR)
            if it >= ite_cte and it > 0:
                rL = torch.as_tensor(np.ascontiguousarray(L_BFGS(skL, ykL, tl, lbfgsLayers, it), np.float32),
                                     device=self.device)
                rM = torch.as_tensor(np.ascontiguousarray(L_BFGS(skM, ykM, tm, lbfgsLayers, it), np.float32),
                                     device=self.device)
                # halving line search, control flow of M:2547-2664: a NaN trial or the
                # nTry-th trial (whatever its misfit) restores the model and stops
                alpha, tries, best = 1.0, 0, 1e20
                while best > self.costFunction[it] or np.isnan(best):
                    tries += 1
                    alpha /= 2
                    self._apply_model(cL, rL, -alpha, cM, rM, -alpha, cR)
                    best = self._misfit_all(data_obs, listaTM)
                    if np.isnan(best) or tries >= nTry:
                        self._set_model_device(*saved)
                        if np.isnan(best) and tries < nTry:
                            continue
                        break
                else:
                    self._apply_model(cL, rL, -alpha, cM, rM, -alpha, cR)
                    self.costFunction[it] = best
            else:
                self.alpha.Lambda, self.alpha.Mu, self.alpha.Rho = steepest_alphas(
                    self.mk.Lambda, self.mk.Mu, self.mk.Rho, tl, tm, tr)
                self._apply_model(mk.d_Lambda, tl_d, self.alpha.Lambda, mk.d_Mu, tm_d, self.alpha.Mu, cR)
            self._save_fwi(tl, tm, tr)
        return self.costFunction

    # -- result files (the reference's names under results_dir) ---------------------
    def _save(self, stem, g):
        if self.results_dir:
            base = os.path.join(self.results_dir, stem)
            os.makedirs(os.path.dirname(base), exist_ok=True)
            for c in ("Vx", "Vy", "Vz"):
                np.save(f"{base}.{c}.npy", getattr(g, c))

    def _save_mt(self, hist, modData):
        if not self.results_dir:
            return
        i = self.SrxIdx
        for d in ("CostFunction", "TM", "Gathers"):
            os.makedirs(os.path.join(self.results_dir, d), exist_ok=True)
        np.save(os.path.join(self.results_dir, f"CostFunction/costFunction{i}.npy"), self.costFunction)
        for k, v in hist.items():
            np.save(os.path.join(self.results_dir, f"TM/Mmod{k}{i}.npy"), v)
        self._save(f"Gathers/modData{i}", modData)
        self._save(f"Gathers/resData{i}", self.resGather)

    def _save_gathers_fq(self, modData):
        """Gathers/modData_{fq}Hz.npy (M:2523): the list of per-earthquake Gather
        objects of this iteration, saved as a numpy object array (pickled, as
        np.save does for the reference's list; read back with allow_pickle=True)."""
        if not self.results_dir:
            return
        os.makedirs(os.path.join(self.results_dir, "Gathers"), exist_ok=True)
        arr = np.empty(len(modData), dtype=object)
        arr[:] = modData
        np.save(os.path.join(self.results_dir, f"Gathers/modData_{self.ondicula.fq}Hz.npy"), arr,
                allow_pickle=True)

    def _save_fwi(self, tl, tm, tr):
        if not self.results_dir:
            return
        fq = self.ondicula.fq
        for d in ("Models", "CostFunction", "Gradients"):
            os.makedirs(os.path.join(self.results_dir, d), exist_ok=True)
        for name, arr in (("Vp", self.mk.Vp), ("Vs", self.mk.Vs), ("Rho", self.mk.Rho),
                          ("Lambda", self.mk.Lambda), ("Mu", self.mk.Mu)):
            np.save(os.path.join(self.results_dir, f"Models/{name}_{fq}Hz.npy"), arr)
        np.save(os.path.join(self.results_dir, f"CostFunction/costFunction_{fq}Hz.npy"), self.costFunction)
        for name, arr in (("gRho", tr), ("gMu", tm), ("gLambda", tl)):
            np.save(os.path.join(self.results_dir, f"Gradients/{name}_{fq}Hz.npy"), arr)


# ----------------------------------------------------------------------------
# the reference's __main__ workflow for the Vp/Vs/ρ inversion (mode 1)
# ----------------------------------------------------------------------------
def initial_model(mTarget: Model, sigma=10.0, top=10, device=None):
    """Starting model of the Vp/Vs/ρ inversion (M:3073-3082): the target's Vp,
    Vs and ρ smoothed with a σ = 10 Gaussian (scipy defaults, here the device
    kernel), the top `top` cells of Vp and Vs from the target, ρ from the
    smoothed Vp by Nafe-Drake with the target's top cells, then λ, μ."""
    dev = _device(device)
    cn = Model(mTarget.nz, mTarget.ny, mTarget.nx, mTarget.dz, mTarget.dy, mTarget.dx)
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)
    smooth = lambda a: _C.condition_gradient(f(a), sigma, 4.0).cpu().numpy()
    cn.Vp, cn.Vs = smooth(mTarget.Vp), smooth(mTarget.Vs)
    cn.Vp[:, :top, :] = mTarget.Vp[:, :top, :]
    cn.Vs[:, :top, :] = mTarget.Vs[:, :top, :]
    cn.Rho = _C.rho_nafe_drake(f(cn.Vp)).cpu().numpy()
    cn.Rho[:, :top, :] = mTarget.Rho[:, :top, :]
    cn.compute_Lambda_Mu()
    return cn


def fwi_multifrequency(mTarget: Model, cn: Model, geometry: Geometry, listTM, Mobs, multFq, tprop,
                       lenPML, nIteFWI, dt=_M.REF_DT, device=None, results_dir=None, **fwi_kw):
    """Frequency continuation of the reference's mode-1 run (M:3134-3203): for
    each frequency of multFq, a Ricker wavelet and a CPML at that frequency (the
    CPML from the target model's Vmax), observed data modelled in mTarget with
    Mobs for every earthquake (saved as Gathers/obsData_{fq}Hz.npy), then
    fwi_Vp_Vs_Rho from the current model cn, whose λ, μ and ρ carry over to the
    next frequency (cn's Vp and Vs are left as they were, as in the reference,
    M:3188-3194).  Returns {fq: cost history} and leaves cn updated."""
    history = {}
    for fq in multFq:
        cn.MaxMinVp()
        wavelet = Wavelet(fq, tprop, mTarget, dt=dt)
        wavelet.ricker()
        cpml = CPML(lenPML, wavelet, mTarget)
        obs_obj = Elastic3D(mTarget, geometry, wavelet, cpml, mode=1, device=device, results_dir=results_dir)
        obs_obj.setStressTensor(Mobs)
        obs_obj.setModel(mTarget)
        obsData = []
        for i in range(geometry.getLenSismos()):
            obs_obj.setIdxEarthquake(i)
            obs_obj.forwardFields("obs")
            d = Gather(wavelet.nSamples, geometry.sismos[i].getLenStations())
            d.Vx, d.Vy, d.Vz = obs_obj.getGathers()
            obsData.append(d)
        if results_dir:
            os.makedirs(os.path.join(results_dir, "Gathers"), exist_ok=True)
            arr = np.empty(len(obsData), dtype=object)
            arr[:] = obsData
            np.save(os.path.join(results_dir, f"Gathers/obsData_{fq}Hz.npy"), arr, allow_pickle=True)
        del obs_obj
        fwi_obj = Elastic3D(cn, geometry, wavelet, cpml, mode=1, device=device, results_dir=results_dir)
        fwi_obj.setModel(cn)
        history[fq] = fwi_obj.fwi_Vp_Vs_Rho(obsData, listTM, nIteFWI, **fwi_kw)
        cn.Rho = fwi_obj.mk.Rho.copy()
        cn.Lambda = fwi_obj.mk.Lambda.copy()
        cn.Mu = fwi_obj.mk.Mu.copy()
    return history
