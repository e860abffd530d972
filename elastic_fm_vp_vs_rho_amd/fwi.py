"""Batched exact-gradient FWI + moment-tensor inversion (workload C4).

The reference inverts one earthquake per MPI rank and sums the λ/μ/ρ images
(fwi_Vp_Vs_Rho, mainPropElastic3D_MT.py M:2349-2719); its MT inversion runs per
earthquake (fwi_Mij M:2217-2346).  Here all shots of a survey go through the
batched propagator in sub-batches sized to the GPU's free HBM (√nt checkpoints
per shot when the full forward history does not fit), and Vp, Vs, ρ and every
shot's 6 MT components are inverted jointly with L-BFGS (lbfgs.py semantics)
on the exact discrete gradient.

    python -m elastic_fm_vp_vs_rho_amd.fwi --iters 20
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from . import autotune, conditioning as _C, lbfgs, workloads
from .propagator import Grid, Propagator, ShotBatch, available_bytes, lame, lame_adjoint


class Survey:
    """All shots of a workload on one grid, served in sub-batches."""

    def __init__(self, wl, device="cuda"):
        self.wl, self.dev = wl, torch.device(device)
        self.grid = Grid(wl.nx, wl.ny, wl.nz, wl.dx, wl.dt, (wl.W,) * 6)
        self.cp = wl.cpml()
        self.wav = wl.wavelets()
        self.S = wl.S
        self.pools = {}      # shot count -> forward states kept between gradients
        self.last_strategy = None

    def batch(self, idx, m6):
        idx = np.asarray(idx)
        wl = self.wl
        sh = ShotBatch(self.grid, wl.src[idx], m6[idx], self.wav[idx], wl.nt[idx], wl.rec[idx],
                       {k: v[idx] for k, v in self.cp.items()}, device=self.dev)
        return sh

    def shot_bytes(self):
        """HBM one shot needs for a checkpointed gradient (full-state snapshots,
        one segment of tape states and their shared stresses, adjoint state,
        accumulators, gathers)."""
        probe = Propagator(self.grid, *self._dummy_medium(), self.batch([0], self.wl.m6))
        K = probe.checkpoint_interval()
        sf, sb = probe.full_state_bytes(), probe.state_bytes()
        return ((K + 4) * sf + (K + 1) * sb + probe._tau_bytes() + 8 * 4 * self.grid.N
                + 4 * 4 * 3 * self.wl.ntmax * self.wl.rec.shape[1])

    def _dummy_medium(self):
        z = torch.ones(self.grid.shape, device=self.dev)
        return z, z, z

    def max_batch(self, fraction=0.8):
        free = available_bytes(self.dev)
        return max(1, min(self.S, autotune.shots_per_dispatch(self.grid),
                           int(fraction * free // self.shot_bytes())))

    def default_batch(self):
        """One shot per sub-batch when a single shot already fills the GPU (2^21
        cells or more: batching gains nothing there, while a lone shot can keep
        the most forward states and re-run the fewest steps); else as many as
        fit with √nt checkpoints."""
        return 1 if self.grid.N >= 1 << 21 else self.max_batch()

    def release_states(self):
        self.pools.clear()


def _tensor(a, dev):
    return torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)


def sub_batches(S, batch):
    """Shot index ranges of at most `batch` shots, as equal as possible (16 shots
    in batches of at most 12 go as 8 + 8), so that all sub-batches share one
    pool of kept forward states."""
    nb = -(-S // max(1, batch))
    cuts = [round(i * S / nb) for i in range(nb + 1)]
    return [np.arange(cuts[i], cuts[i + 1]) for i in range(nb) if cuts[i + 1] > cuts[i]]


def gradient(survey: Survey, vp, vs, rho, m6, obs, batch):
    """Misfit J = ½Σ‖d - obs‖² over all shots and its exact gradient with respect
    to Vp, Vs, ρ (device volumes) and every shot's MT ((S, 6) float64), computed
    `batch` shots at a time."""
    lam, mu = lame(vp, vs, rho)
    gl = torch.zeros_like(vp)
    gm = torch.zeros_like(vp)
    gr = torch.zeros_like(vp)
    gm6 = np.zeros((survey.S, 6))
    J = 0.0
    for idx in sub_batches(survey.S, batch):
        for k in [k for k in survey.pools if k != len(idx)]:
            del survey.pools[k]                  # one pool at a time: never two sets of states
        prop = Propagator(survey.grid, lam, mu, rho, survey.batch(idx, m6),
                          pool=survey.pools.setdefault(len(idx), []))
        Jb, _, a, b, c, d6 = prop.gradient(obs[idx[0]:idx[-1] + 1], strategy="auto")
        survey.last_strategy = prop.last_strategy
        prop.release_states()
        J += Jb
        gl += a
        gm += b
        gr += c
        gm6[idx] = d6.cpu().numpy()
        del prop
    gvp, gvs, grho = lame_adjoint(vp, vs, rho, gl, gm, gr)
    return J, gvp, gvs, grho, gm6


def misfit(survey: Survey, vp, vs, rho, m6, obs, batch):
    lam, mu = lame(vp, vs, rho)
    J = 0.0
    for idx in sub_batches(survey.S, batch):
        prop = Propagator(survey.grid, lam, mu, rho, survey.batch(idx, m6))
        g, _ = prop.forward()
        J += float(prop.misfit(g, obs[idx[0]:idx[-1] + 1])[1])
        del prop
    return J


def observed(survey: Survey, vp, vs, rho, m6, batch):
    lam, mu = lame(vp, vs, rho)
    wl = survey.wl
    out = torch.empty((survey.S, 3, wl.ntmax, wl.rec.shape[1]), device=survey.dev)
    for idx in sub_batches(survey.S, batch):
        prop = Propagator(survey.grid, lam, mu, rho, survey.batch(idx, m6))
        out[idx[0]:idx[-1] + 1] = prop.forward()[0]
        del prop
    return out


def update_mask(grid: Grid, margin=2):
    """1 inside the CPML-free interior shrunk by `margin`, else 0 (model updates
    stay out of the absorbing layers, like the reference's masks, M:2479-2504)."""
    m = torch.zeros(grid.shape)
    w = grid.widths
    m[w[2] + margin:grid.ny - w[3] - margin, w[4] + margin:grid.nz - w[5] - margin,
      w[0] + margin:grid.nx - w[1] - margin] = 1.0
    return m


def invert(survey: Survey, vp0, vs0, rho0, m60, obs, iters, batch, layers=7,
           step=(20.0, 12.0, 10.0, 0.05), tries=6, log=None, smooth=0.0, callback=None):
    """Joint L-BFGS on (Vp, Vs, ρ, MT).  The search variable z maps to the model
    as x = x0 + P z with a fixed diagonal P per block, set at the first gradient
    so that the first steepest step (α = 1) changes Vp, Vs, ρ at most by
    `step[0:3]` (m/s, m/s, kg/m³) and the MTs by `step[3]` of their mean
    Frobenius norm.  The line search
    halves from α = 1 until the misfit decreases (at most `tries` trials).
    `layers` is lbfgs.py's history size (layers - 2 pairs are used).
    The Vp, Vs, ρ gradients are conditioned on the device (efm_condition_gradient):
    zeroed in the CPML + 2 cells of every face and, when `smooth` > 0, smoothed
    with a Gaussian of σ = `smooth` cells (the reference smooths with σ = 1,
    M:2479-2482)."""
    dev = survey.dev
    mask = update_mask(survey.grid).to(dev)
    w = survey.grid.widths
    margins = tuple(v + 2 for v in w)

    def condition(gv):
        return _C.condition_gradient(gv, smooth, 4.0, margins)
    nv = vp0.numel()
    x0 = [vp0, vs0, rho0]
    m60 = np.asarray(m60, np.float64)
    n = 3 * nv + m60.size
    z = np.zeros(n)
    sk = np.zeros((layers, n))
    yk = np.zeros((layers, n))
    P = None
    hist = []

    def model(zv):
        parts = []
        for k in range(3):
            d = torch.as_tensor(zv[k * nv:(k + 1) * nv].reshape(vp0.shape), dtype=torch.float32, device=dev)
            parts.append(x0[k] + float(P[k]) * mask * d)
        m6 = m60 + P[3] * zv[3 * nv:].reshape(m60.shape)
        return parts[0], parts[1], parts[2], m6

    for it in range(iters + 1):
        if P is None:
            J, gvp, gvs, grho, gm6 = gradient(survey, vp0, vs0, rho0, m60, obs, batch)
            # a steepest step of unit length moves x by P²·g: size P per block so
            # that it moves each block by its `step`
            gvp, gvs, grho = condition(gvp), condition(gvs), condition(grho)
            fro = np.sqrt((m60[:, :3] ** 2).sum(1) + 2 * (m60[:, 3:] ** 2).sum(1)).mean()
            P = [math.sqrt(step[0] / max(float(gvp.abs().max()), 1e-30)),
                 math.sqrt(step[1] / max(float(gvs.abs().max()), 1e-30)),
                 math.sqrt(step[2] / max(float(grho.abs().max()), 1e-30)),
                 math.sqrt(step[3] * fro / max(float(np.abs(gm6).max()), 1e-30))]
        else:
            vp, vs, rho, m6 = model(z)
            J, gvp, gvs, grho, gm6 = gradient(survey, vp, vs, rho, m6, obs, batch)
            gvp, gvs, grho = condition(gvp), condition(gvs), condition(grho)
        hist.append(J)
        if callback:
            callback(it, J)
        if log:
            log(f"[fwi] iteration {it}: J = {J:.6e}")
        if it == iters:
            break
        g = np.concatenate([float(P[0]) * gvp.double().cpu().numpy().ravel(),
                            float(P[1]) * gvs.double().cpu().numpy().ravel(),
                            float(P[2]) * grho.double().cpu().numpy().ravel(),
                            P[3] * gm6.ravel()])
        lbfgs.compute_sk_yk(sk, yk, z, g, layers, it)
        p = -(g if it == 0 else lbfgs.L_BFGS(sk, yk, g, layers, it))
        if it > 0 and np.dot(p, g) >= 0:       # not a descent direction: restart
            p = -g
        alpha, accepted = 1.0, False
        for _ in range(tries):
            vp, vs, rho, m6 = model(z + alpha * p)
            Jt = misfit(survey, vp, vs, rho, m6, obs, batch)
            if np.isfinite(Jt) and Jt < J:
                z = z + alpha * p
                accepted = True
                break
            alpha *= 0.5
        if not accepted:
            if log:
                log(f"[fwi] iteration {it}: line search failed, stopping")
            break
    vp, vs, rho, m6 = model(z)
    return hist, (vp, vs, rho, m6)


def plain_driver(wl, obs, device):
    """The reference-structured path on a workload: driver.Elastic3D in mode 1
    (one earthquake at a time, boundary-saving forward, reverse-time
    reconstruction with the reference's imaging conditions, host misfit and
    L-BFGS, M:2349-2719), with the workload's sources, receivers, true moment
    tensors, wavelet and CPML on every face.  Returns (Elastic3D, observed
    Gathers, listTM)."""
    from . import driver as D
    mdl = D.Model(wl.nz, wl.ny, wl.nx, wl.dx, wl.dx, wl.dx)
    mdl.Vp, mdl.Vs, mdl.Rho = wl.vp.copy(), wl.vs.copy(), wl.rho.copy()
    mdl.compute_Lambda_Mu()
    mdl.VmaxVp = wl.vmax
    wav = D.Wavelet(float(wl.f0[0]), (wl.ntmax + 0.5) * wl.dt, mdl, dt=wl.dt)
    wav.ricker()
    cpml = D.CPML(wl.W, wav, mdl, free_top=wl.free_top)
    geo = D.Geometry()
    for s in range(wl.S):
        e = D.Earthquake(*(int(v) for v in wl.src[s]))
        for r in wl.rec[s]:
            e.addStation(D.Station(*(int(v) for v in r)))
        geo.addSismo(e)
    e3 = D.Elastic3D(mdl, geo, wav, cpml, mode=1, device=device)
    host = obs.cpu().numpy()
    gathers = []
    for s in range(wl.S):
        g = D.Gather(wl.ntmax, wl.rec.shape[1])
        g.Vx, g.Vy, g.Vz = host[s, 0].copy(), host[s, 1].copy(), host[s, 2].copy()
        gathers.append(g)
    listTM = [D.StressTensor(*(float(v) for v in wl.m6[s])) for s in range(wl.S)]
    return e3, gathers, listTM


def compare_plain(nx=256, ny=256, nz=128, S=8, nt=600, plain_iters=5, iters=12, log=None,
                  host_loop=True, host_iters=0, step=None):
    """FWI on a 256×256×128 survey: the plain single-shot path (plain_driver,
    one earthquake at a time, host-driven, the reference's iteration rules),
    with the reference's imaging conditions and with the exact gradient, against
    this package's batched joint L-BFGS on the exact gradient (invert); all from
    the smooth start with the true moment tensors.  Reports s per FWI gradient
    of each path, the misfit histories and wall times, and the time the batched
    inversion needs to reach each plain run's final misfit (when that run
    lowered the misfit at all).  host_loop: also time one gradient of the plain
    path driven the way the reference drives it (one host call per time step,
    each reverse iteration's three images copied to the host and summed there,
    M:1823, 1939-2054); host_iters > 0 also runs that many iterations of the
    host-driven plain inversion (its wall time to its final misfit).  step:
    the batched inversion's first-step sizes (invert's `step`; default its own)."""
    dev = torch.device("cuda:0")
    wt = workloads.c4(nx=nx, ny=ny, nz=nz, S=S, nt=nt, anomalies=True)
    w0 = workloads.c4(nx=nx, ny=ny, nz=nz, S=S, nt=nt, anomalies=False)
    survey = Survey(w0, dev)
    batch = survey.default_batch()
    f = lambda a: _tensor(a, dev)
    lam0, mu0 = lame(f(w0.vp), f(w0.vs), f(w0.rho))
    autotune.tune(Propagator(survey.grid, lam0, mu0, f(w0.rho), survey.batch(np.arange(batch), w0.m6)))
    del lam0, mu0
    obs = observed(survey, f(wt.vp), f(wt.vs), f(wt.rho), wt.m6, batch)
    J0 = misfit(survey, f(w0.vp), f(w0.vs), f(w0.rho), w0.m6, obs, batch)
    out = {"workload": f"{nx}x{ny}x{nz}, {S} shots x {nt} steps (record length bounded from 2500), "
                       f"C4-style layered model with two anomalies", "J0": J0, "plain": {}}

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    for grad in ("reference", "exact"):
        e3, gathers, listTM = plain_driver(w0, obs, dev)
        e3._misfit_kind = "l2"
        e3.iteration = 1
        e3._all_shots_gradient(gathers, listTM, grad)            # warm-up: allocations, captures
        _, t_grad = timed(lambda: e3._all_shots_gradient(gathers, listTM, grad))
        cost, t_inv = timed(lambda: e3.fwi_Vp_Vs_Rho(gathers, listTM, plain_iters, lbfgsLayers=7, ite_cte=2,
                                                     gradient=grad))
        J_end = float(e3._misfit_all(gathers, listTM))             # misfit of its final model
        out["plain"][grad] = {"path": f"driver.Elastic3D.fwi_Vp_Vs_Rho(gradient={grad!r}): one earthquake "
                                      f"at a time, host misfit, lbfgs.py on host arrays",
                              "s_per_gradient": t_grad, "iterations": plain_iters, "wall_s": t_inv,
                              "J_history": [float(c) for c in cost], "J_final": J_end}
        if log:
            log(f"[compare] plain ({grad}): {t_grad:.2f} s per gradient; {plain_iters} iterations in "
                f"{t_inv:.1f} s, J {J0:.4e} -> {J_end:.4e}")
        e3.release_states()
        del e3
        torch.cuda.empty_cache()
    if host_loop:
        e3, gathers, listTM = plain_driver(w0, obs, dev)
        e3._misfit_kind = "l2"
        e3.iteration = 1
        e3.host_loop = True
        _, t_host = timed(lambda: e3._all_shots_gradient(gathers, listTM, "reference"))
        out["plain_host_driven"] = {
            "path": "driver.Elastic3D with host_loop=True, gradient='reference': one host call per time "
                    "step, each reverse iteration's three images copied to the host and summed there "
                    "(the reference's loop structure); one gradient timed, its iterates are those of the "
                    "'reference' plain run",
            "s_per_gradient": t_host}
        if log:
            log(f"[compare] plain (reference, host-driven): {t_host:.2f} s per gradient")
        if host_iters > 0:
            cost, t_inv = timed(lambda: e3.fwi_Vp_Vs_Rho(gathers, listTM, host_iters, lbfgsLayers=7,
                                                         ite_cte=2, gradient="reference"))
            J_end = float(e3._misfit_all(gathers, listTM))
            out["plain_host_driven"].update(iterations=host_iters, wall_s=t_inv,
                                            J_history=[float(c) for c in cost], J_final=J_end)
            if log:
                log(f"[compare] plain (reference, host-driven): {host_iters} iterations in {t_inv:.1f} s, "
                    f"J {J0:.4e} -> {J_end:.4e}")
        e3.release_states()
        del e3
        torch.cuda.empty_cache()
    args = (survey, f(w0.vp), f(w0.vs), f(w0.rho), w0.m6.copy(), obs, batch)
    gradient(*args)                                              # warm-up
    _, t_bgrad = timed(lambda: gradient(*args))
    survey.release_states()
    marks = []
    t1 = [0.0]

    def mark(it, J):
        marks.append((it, J, time.perf_counter() - t1[0]))

    torch.cuda.synchronize()
    t1[0] = time.perf_counter()
    kw = {} if step is None else {"step": tuple(step)}
    hist, _ = invert(survey, f(w0.vp), f(w0.vs), f(w0.rho), w0.m6.copy(), obs, iters, batch, log=log,
                     callback=mark, **kw)
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - t1[0]
    out["batched"] = {"path": "fwi.invert: joint L-BFGS on Vp, Vs, rho and MT, exact gradient, "
                              f"{batch} shot(s) per sub-batch",
                      "step": list(step) if step is not None else "default",
                      "s_per_gradient": t_bgrad, "iterations": len(hist) - 1, "wall_s": t_batched,
                      "J_history": [float(J) for J in hist]}
    for grad, pl in out["plain"].items():
        target = pl["J_final"]
        reach = next((m for m in marks if m[1] <= target), None) if target < J0 else None
        pl["batched_time_to_this_misfit_s"] = reach[2] if reach else None
        pl["speedup_to_equal_misfit"] = pl["wall_s"] / reach[2] if reach else None
        pl["plain_over_batched_s_per_gradient"] = pl["s_per_gradient"] / t_bgrad
    if "plain_host_driven" in out:
        ph = out["plain_host_driven"]
        ph["plain_over_batched_s_per_gradient"] = ph["s_per_gradient"] / t_bgrad
        if "J_final" in ph:
            target = ph["J_final"]
            reach = next((m for m in marks if m[1] <= target), None) if target < J0 else None
            ph["batched_time_to_this_misfit_s"] = reach[2] if reach else None
            ph["speedup_to_equal_misfit"] = ph["wall_s"] / reach[2] if reach else None
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--shots", type=int, default=16)
    ap.add_argument("--nt", type=int, default=2500)
    ap.add_argument("--nx", type=int, default=200)
    ap.add_argument("--nz", type=int, default=100)
    ap.add_argument("--batch", type=int, default=0, help="shots per sub-batch (0: as many as fit)")
    ap.add_argument("--no-plain", action="store_true", help="skip timing the one-shot-at-a-time gradient")
    ap.add_argument("--no-tune", action="store_true", help="keep the installed tiles")
    ap.add_argument("--gradient-only", action="store_true",
                    help="time two batched gradients and stop (no inversion)")
    ap.add_argument("--compare-plain", action="store_true",
                    help="time the reference-structured single-shot FWI against the batched inversion "
                         "to equal misfit (defaults: 256x256x128, 8 shots, 600 steps)")
    ap.add_argument("--plain-iters", type=int, default=5)
    ap.add_argument("--no-host-loop", action="store_true",
                    help="skip timing the plain gradient driven one step at a time from the host")
    ap.add_argument("--host-iters", type=int, default=0,
                    help="also run this many iterations of the host-driven plain inversion")
    ap.add_argument("--step", default=None,
                    help="first-step sizes of the batched inversion: Vp,Vs,rho,MT (default: invert's own)")
    args = ap.parse_args(argv)
    log = lambda s: print(s, file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    if args.compare_plain:
        print(json.dumps(compare_plain(nx=args.nx, ny=args.nx, nz=args.nz, S=args.shots, nt=args.nt,
                                       plain_iters=args.plain_iters, iters=args.iters, log=log,
                                       host_loop=not args.no_host_loop, host_iters=args.host_iters,
                                       step=[float(v) for v in args.step.split(",")] if args.step else None)))
        return
    wt = workloads.c4(nx=args.nx, ny=args.nx, nz=args.nz, S=args.shots, nt=args.nt, anomalies=True)
    w0 = workloads.c4(nx=args.nx, ny=args.nx, nz=args.nz, S=args.shots, nt=args.nt, anomalies=False)
    survey = Survey(w0, dev)
    batch = args.batch or survey.default_batch()
    f = lambda a: _tensor(a, dev)
    if not args.no_tune:
        # tiles for this grid and sub-batch size (cached per shape)
        lam0, mu0 = lame(f(w0.vp), f(w0.vs), f(w0.rho))
        tune_prop = Propagator(survey.grid, lam0, mu0, f(w0.rho), survey.batch(np.arange(batch), w0.m6))
        log(f"[fwi] tiles: { {k: v['tile'] for k, v in autotune.tune(tune_prop).items()} }")
        del tune_prop, lam0, mu0
        torch.cuda.empty_cache()
    obs = observed(survey, f(wt.vp), f(wt.vs), f(wt.rho), wt.m6, batch)
    m6_start = np.tile(np.array([0.5, 0.0, 0.5, 0.0, 0.5, 0.0]), (w0.S, 1))
    vp0, vs0, rho0 = f(w0.vp), f(w0.vs), f(w0.rho)
    if args.gradient_only:
        out = {"workload": w0.name, "shots_per_batch": batch, "s_per_gradient": []}
        for _ in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            gradient(survey, vp0, vs0, rho0, m6_start, obs, batch)
            torch.cuda.synchronize()
            out["s_per_gradient"].append(time.perf_counter() - t0)
        print(json.dumps(out))
        return
    # s per FWI gradient, batched (sub-batches of `batch`) vs plain (one shot at a time);
    # the second of two calls is timed (the first allocates the kept forward states)
    times = {}
    configs = [("batched", batch)]
    if not args.no_plain and batch != 1:
        configs.append(("plain", 1))
    for name, b in configs:
        survey.release_states()
        torch.cuda.empty_cache()
        torch.cuda.reset_peak_memory_stats(dev)
        secs = []
        for _ in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            Jg = gradient(survey, vp0, vs0, rho0, m6_start, obs, b)[0]
            torch.cuda.synchronize()
            secs.append(time.perf_counter() - t0)
        times[name] = {"s_per_gradient": secs[1], "first_call_s": secs[0], "shots_per_batch": b,
                       "forward_states": survey.last_strategy,
                       "peak_hbm_gb": torch.cuda.max_memory_allocated(dev) / 1e9, "J": Jg}
        log(f"[fwi] {name}: {times[name]}")
    if len(configs) > 1:
        survey.release_states()
        torch.cuda.empty_cache()
    t0 = time.perf_counter()
    hist, _ = invert(survey, vp0, vs0, rho0, m6_start, obs, args.iters, batch, log=log)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cell_steps = float(survey.grid.N) * float(w0.nt.sum())
    print(json.dumps({"workload": w0.name, "shots": w0.S, "nt": int(w0.ntmax),
                      "gradient": times, "batched_over_plain":
                      times["plain"]["s_per_gradient"] / times["batched"]["s_per_gradient"]
                      if "plain" in times else None,
                      "gradient_cell_steps_per_s": cell_steps / times["batched"]["s_per_gradient"],
                      "iterations": len(hist) - 1, "J": hist, "J_final_over_J0": hist[-1] / hist[0],
                      "inversion_wall_s": wall, "lbfgs_layers": 7, "lbfgs_pairs": 5}))


if __name__ == "__main__":
    main()
