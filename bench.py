#!/usr/bin/env python3
"""Benchmark: exact FWI gradient (forward + adjoint) of the 128³ three-layer
model with a Gaussian anomaly, 2000 steps, 64 receivers, one MT source, on one
MI355X.  A bench step is one full gradient evaluation: forward run keeping
√nt checkpoints of the forward wavefield (not the full history), misfit,
reverse sweep (each checkpoint segment re-run forward, then adjoint + imaging),
gradient assembly.  The same gradient with every forward state kept in HBM
("store") is timed beside it.

Prints one JSON line.  ``value`` is gradient cell-steps per second (grid cells
× time steps of one gradient, per second of wall time); one gradient cell-step
is one cell carried through one forward and one adjoint time step.  The line
also carries the seismogram error of the timed configuration against a
committed FP64 fixture, the gradient error against the FP64 oracle's gradient
(a committed fixture) on a small member of the same workload family, and a
bounded sample of the C4 gradient.  The oracle itself runs only in the CPU
baseline leg.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-pt updates/s + s per FWI gradient, batched vs plain, at ≤1e-5 err vs FP64"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per interior cell for one launch of each kernel (each field
# read once and written once; neighbours reused on chip; ψ, source, receiver
# traffic excluded):
#   stress     v×3, λ, μ read; τ×6 read+write                    = 17 words
#   velocity   τ×6, ρ read; v×3 read+write                        = 13 words
#   adj_vel    v̄×3, forward v^{n+1}, v^n ×3 each, b̄×3 read; τ̄×6 and
#              the ρ sum r+w                                       = 26 words
#   adj_str    τ̄×6, forward v×3, λ, μ, μ̄×3 read; v̄×3, λμ-sums×5 r+w = 30 words
ALG_BYTES = {"stress": 68, "velocity": 52, "adj_velocity": 104, "adj_stress": 120}
# The byte model of the metric counts an adjoint time step as 120 B (the adjoint
# state's sweeps) + 24 B (three imaging accumulators read and written) = 144 B
# per interior cell; the two-sweep design moves 224 B (its τ̄ crosses HBM
# between the sweeps, it reads the derived medium and keeps 6 sums).
ADJ_STEP_MODEL_BYTES = 144
FIXTURE_C2 = os.path.join(ROOT, "tests", "golden", "c2_gathers_400.npz")
FIXTURE_C2_GRADIENT = os.path.join(ROOT, "tests", "golden", "c2_gradient_40.npz")

# HBM bytes per dispatch on the 128^3 C2 grid, by tile "bx,bz,ny,map": rocprofv3
# FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (the factors that make the counters read
# 256 MiB each way for a 256 MiB copy), medians over the dispatches of
# `bench.py --profile-kernels`.  Only valid for the kernels' current sources and
# the 128^3 grid; any other size reports traffic null.
MEASURED_TRAFFIC = {
    "stress": {
        "128,1,1,0": 191197637, "128,2,1,0": 171724689, "64,1,1,0": 208424926, "64,2,1,0": 187271551,
        "64,4,1,0": 177759425, "64,4,2,0": 183068655, "128,1,1,8": 155800165, "128,2,1,8": 155800773,
        "256,1,1,8": 155798693, "64,1,1,8": 155800069, "64,2,1,8": 155797926, "64,4,1,8": 155799461,
        "64,4,2,8": 156048560, "64,8,1,8": 155785703, "128,1,1,16": 155861316, "128,1,2,16": 157314835,
        "256,1,1,16": 155860004, "64,1,1,16": 155954500,
    },
    "velocity": {
        "128,1,1,0": 161603185, "128,2,1,0": 140184167, "64,1,1,0": 174893508, "64,2,1,0": 154752523,
        "64,4,1,0": 145708997, "64,4,2,0": 148252102, "128,1,1,8": 124300601, "128,2,1,8": 124302857,
        "256,1,1,8": 124293337, "64,1,1,8": 124300505, "64,2,1,8": 124301241, "64,4,1,8": 124293081,
        "64,4,2,8": 124414799, "64,8,1,8": 124302617, "128,1,1,16": 124601857, "128,1,2,16": 125905375,
        "256,1,1,16": 124322808, "64,1,1,16": 124376568,
    },
    "adj_velocity": {
        "128,1,1,0": 297996656, "128,2,1,0": 260773961, "64,1,1,0": 321601941, "64,2,1,0": 283630696,
        "64,4,1,0": 265993074, "64,4,2,0": 312446312, "128,1,1,8": 231463820, "128,2,1,8": 231460153,
        "256,1,1,8": 230669693, "64,1,1,8": 231400148, "64,2,1,8": 231437298, "64,4,1,8": 231500798,
        "64,4,2,8": 253276609, "64,8,1,8": 232402444, "128,1,1,16": 239307653, "128,1,2,16": 266988764,
        "256,1,1,16": 237318627, "64,1,1,16": 239244164,
    },
    "adj_stress": {
        "128,1,1,0": 388869317, "128,2,1,0": 333875468, "64,1,1,0": 422968550, "64,2,1,0": 370303011,
        "64,4,1,0": 343344816, "64,4,2,0": 474192917, "128,1,1,8": 304226871, "128,2,1,8": 302532747,
        "256,1,1,8": 302041446, "64,1,1,8": 305047935, "64,2,1,8": 304884462, "64,4,1,8": 304141087,
        "64,4,2,8": 412439331, "64,8,1,8": 302609413, "128,1,1,16": 327087078, "128,1,2,16": 468075640,
        "256,1,1,16": 323230582, "64,1,1,16": 327790592,
    },
}


def measured_traffic(kernel, wl):
    """HBM bytes per launch of ``kernel`` with the installed tiles (C2 grid only)."""
    from elastic_fm_vp_vs_rho_amd import autotune
    if (wl.nx, wl.ny, wl.nz) != (128, 128, 128):
        return None
    names = ("adj_velocity", "adj_stress") if kernel == "adjoint_step" else (kernel,)
    total = 0
    for k in names:
        key = ",".join(str(v) for v in autotune.get_tile(autotune.KERNELS.index(k)))
        if key not in MEASURED_TRAFFIC[k]:
            return None
        total += MEASURED_TRAFFIC[k][key]
    return total


def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def time_kernels(prop, reps=30):
    """Average duration (ms) of each sweep kernel with the installed tiles,
    measured with HIP events on the stream the library launches on (torch's
    current stream).  The adjoint pair is timed together (one adjoint step =
    points kernel + adjoint velocity + adjoint stress sweeps)."""
    from elastic_fm_vp_vs_rho_amd.autotune import _Bench
    B = _Bench(prop, warm_steps=100)
    return {"stress": B.time(0, reps), "velocity": B.time(1, reps), "adjoint_step": B.time(2, reps)}


def sample_propagator(wl, S, steps):
    """Propagator for the first S shots of a workload, every record cut to
    `steps` steps, in the workload's smooth model (a kernel-profiling sample)."""
    import numpy as np
    import torch
    from elastic_fm_vp_vs_rho_amd.propagator import Grid, Propagator, ShotBatch, lame
    dev = torch.device("cuda:0")
    idx = np.arange(S)
    grid = Grid(wl.nx, wl.ny, wl.nz, wl.dx, wl.dt, (wl.W,) * 6)
    cp = {k: v[idx] for k, v in wl.cpml().items()}
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)
    rho = f(wl.rho)
    lam, mu = lame(f(wl.vp), f(wl.vs), rho)
    sh = ShotBatch(grid, wl.src[idx], wl.m6[idx], wl.wavelets()[idx][:, :steps], np.full(S, steps),
                   wl.rec[idx], cp, device=dev)
    return Propagator(grid, lam, mu, rho, sh)


def profile_kernels(prop, name):
    """A few dispatches of each sweep kernel (after tuning, which dispatches
    every variant) and three 256 MiB copies, for hardware-counter runs; prints
    the kernel times and the interior cells per launch."""
    import torch
    from elastic_fm_vp_vs_rho_amd import autotune
    autotune.tune(prop, use_cache=False, reps=3)
    kt = time_kernels(prop, reps=4)
    src = torch.ones(64 << 20, dtype=torch.float32, device=prop.lam.device)   # 256 MiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    g = prop.grid
    cells = (g.nx - 4) * (g.ny - 4) * (g.nz - 4) * prop.shots.S
    print(json.dumps({"workload": name, "profile_kernels_ms": kt, "interior_cells_per_launch": cells,
                      "tiles": {k: autotune.get_tile(i) for i, k in enumerate(autotune.KERNELS)},
                      "copy_bytes_each_way": src.numel() * 4}))


def batched_vs_plain(wl, S, steps, chunk=50, tune=False):
    """Forward grid-point updates/s on a bounded sample of a multi-shot
    workload: its first S shots (record lengths scaled to at most ``steps``,
    staying uneven), advanced together in one batch (hipGraph time loop) and
    one shot at a time (eager launches).  tune: each path runs with the forward
    tiles the autotuner picks for its own batch size."""
    import numpy as np
    import torch
    from elastic_fm_vp_vs_rho_amd.propagator import Grid, Propagator, ShotBatch, lame
    dev = torch.device("cuda:0")
    idx = np.arange(S)
    nt = np.maximum(1, wl.nt[idx] * steps // wl.nt[idx].max())
    grid = Grid(wl.nx, wl.ny, wl.nz, wl.dx, wl.dt, (wl.W,) * 6)
    cp = {k: v[idx] for k, v in wl.cpml().items()}
    wav = wl.wavelets()[idx][:, :int(nt.max())]
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    rho = f(wl.rho)
    lam, mu = lame(f(wl.vp), f(wl.vs), rho)

    def prop(sel):
        sh = ShotBatch(grid, wl.src[idx][sel], wl.m6[idx][sel], wav[sel], nt[sel], wl.rec[idx][sel],
                       {k: v[sel] for k, v in cp.items()}, device=dev)
        return Propagator(grid, lam, mu, rho, sh)

    from elastic_fm_vp_vs_rho_amd import autotune
    batch = prop(np.arange(S))
    if tune:
        autotune.tune(batch, kernels=(0, 1))
    batch.forward(nsteps=chunk, graph_chunk=chunk)           # warm-up (graph capture path)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.forward(graph_chunk=chunk)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    del batch
    singles = [prop(np.array([s])) for s in range(S)]
    if tune:
        autotune.tune(singles[0], kernels=(0, 1))
    singles[0].forward(nsteps=chunk)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p in singles:
        p.forward()
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    work = float(grid.N) * float(nt.sum())
    return {"sample": f"{wl.nx}x{wl.ny}x{wl.nz}, {S} shots, uneven record lengths <= {steps} steps "
                      f"({int(nt.sum())} shot-steps)",
            "batched_updates_per_s": work / tb, "plain_updates_per_s": work / tp,
            "batched_over_plain": tp / tb}


def cpu_baseline(wl, steps):
    """FP64 NumPy oracle: the same gradient (forward with stored history + exact
    adjoint) on the first ``steps`` time steps of the workload, one thread."""
    import numpy as np
    from oracle import elastic_fp64 as O
    lam, mu = wl.lame()
    med = O.Medium(lam, mu, wl.rho)
    cp = O.Cpml.from_dict({k: v[0] for k, v in wl.cpml().items()})
    w = wl.wavelets()[0]
    rec = wl.rec[0]
    c = wl.dt / wl.dx
    t0 = time.perf_counter()
    hist = []
    g, _ = O.forward(med, cp, c, tuple(wl.src[0]), wl.m6[0], w, rec, steps, threads=1, history=hist)
    O.exact_adjoint(med, cp, c, tuple(wl.src[0]), w, rec, steps, g, hist)
    dt = time.perf_counter() - t0
    out = dict(value=wl.nx * wl.ny * wl.nz * steps / dt, unit="gradient cell-steps/s", cores=1,
               kind="port", sample=f"first {steps} of {wl.ntmax} time steps of the same gradient "
               f"(forward with stored history + exact adjoint), {wl.nx}x{wl.ny}x{wl.nz}, FP64 NumPy, "
               f"{dt:.1f} s")
    # the forward alone split over the host cores (y slabs on a thread pool; at most 16,
    # the CPU share of one GPU)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    t0 = time.perf_counter()
    O.forward(med, cp, c, tuple(wl.src[0]), wl.m6[0], w, rec, steps, threads=threads)
    dtf = time.perf_counter() - t0
    out["forward_all_cores"] = dict(value=wl.nx * wl.ny * wl.nz * steps / dtf,
                                    unit="grid-pt updates/s", cores=threads,
                                    sample=f"first {steps} forward steps, {dtf:.1f} s")
    return out


def gradient_vs_oracle():
    """Relative L2 error of the device gradient (λ, μ, ρ per field and the 6 MT
    components) against the FP64 oracle's exact adjoint on a small member of the
    C2 workload family (same layers, source and receiver layout, 40³ cells, 200
    steps), each side with its own misfit residual against the anomaly model's
    data; the oracle's gradient is the committed fixture
    tests/golden/c2_gradient_40.npz."""
    import numpy as np
    import torch
    from elastic_fm_vp_vs_rho_amd import workloads
    from elastic_fm_vp_vs_rho_amd.propagator import Grid, Propagator, ShotBatch, lame
    ref = np.load(FIXTURE_C2_GRADIENT)
    n, nt = int(ref["n"]), int(ref["nt"])
    wl, wt = workloads.c2(n=n, nt=nt), workloads.c2(n=n, nt=nt, anomaly=True)
    dev = torch.device("cuda:0")
    grid = Grid(wl.nx, wl.ny, wl.nz, wl.dx, wl.dt, (wl.W,) * 6)
    shots = ShotBatch(grid, wl.src, wl.m6, wl.wavelets(), wl.nt, wl.rec, wl.cpml(), device=dev)
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    lt, mt = lame(f(wt.vp), f(wt.vs), f(wt.rho))
    obs, _ = Propagator(grid, lt, mt, f(wt.rho), shots).forward()
    lam, mu = lame(f(wl.vp), f(wl.vs), f(wl.rho))
    _, _, glam, gmu, grho, gm6 = Propagator(grid, lam, mu, f(wl.rho), shots).gradient(obs)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return {"gradient_rel_l2": {"lam": rel(glam.double().cpu().numpy(), ref["lam"].astype(np.float64)),
                                "mu": rel(gmu.double().cpu().numpy(), ref["mu"].astype(np.float64)),
                                "rho": rel(grho.double().cpu().numpy(), ref["rho"].astype(np.float64)),
                                "m6": rel(gm6[0].cpu().numpy(), ref["m6"])},
            "gradient_sample": f"C2 family at {n}^3 x {nt} steps (same layers, source and receiver "
                               f"layout), each side with its own residual, vs the FP64 oracle's "
                               f"exact adjoint (committed fixture)"}


def plan_sample():
    """Seconds per exact gradient, eager launches against replays from a
    persistent plan of HIP graphs (after one capturing call), on a batch whose
    steps are long (96^3 x 64 shots, 256 steps) and on one whose steps are a few
    microseconds (40^3 x 1 shot, 512 steps); tiles tuned per batch."""
    import numpy as np
    import torch
    from elastic_fm_vp_vs_rho_amd import autotune, workloads
    out = {}
    for n, S, nt in ((96, 64, 256), (40, 1, 512)):
        wl = workloads.c5(n=n, S=S, nt=nt)
        res = {}
        for graphs in (False, True):
            prop = sample_propagator(wl, S, nt)
            prop.graphs = graphs
            if not graphs:
                autotune.tune(prop)
            obs = torch.zeros_like(prop.forward()[0])
            prop.gradient(obs)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(3):
                prop.gradient(obs)
            torch.cuda.synchronize()
            res["graphs" if graphs else "eager"] = (time.perf_counter() - t0) / 3
            res["strategy"] = prop.last_strategy
            prop.release_states()
            del prop
            torch.cuda.empty_cache()
        res["eager_over_graphs"] = res["eager"] / res["graphs"]
        out[f"{n}^3 x {S} shot(s) x {nt} steps"] = res
    return out


def c4_gradient_sample(steps):
    """Seconds per exact gradient of the C4 survey at full size (200x200x100, 16
    shots, one per sub-batch as the FWI driver runs it) with every record cut to
    `steps` steps, after one untimed call; tiles tuned for one shot."""
    import numpy as np
    import torch
    from elastic_fm_vp_vs_rho_amd import autotune, fwi, workloads
    from elastic_fm_vp_vs_rho_amd.propagator import Propagator, lame
    dev = torch.device("cuda:0")
    w0 = workloads.c4(nt=steps, anomalies=False)
    wt = workloads.c4(nt=steps, anomalies=True)
    survey = fwi.Survey(w0, dev)
    f = lambda a: fwi._tensor(a, dev)
    batch = survey.default_batch()
    saved = [autotune.get_tile(k) for k in range(4)]
    lam0, mu0 = lame(f(w0.vp), f(w0.vs), f(w0.rho))
    autotune.tune(Propagator(survey.grid, lam0, mu0, f(w0.rho), survey.batch(np.arange(batch), w0.m6)))
    obs = fwi.observed(survey, f(wt.vp), f(wt.vs), f(wt.rho), wt.m6, batch)
    m6 = np.tile([0.5, 0.0, 0.5, 0.0, 0.5, 0.0], (w0.S, 1))
    args = (survey, f(w0.vp), f(w0.vs), f(w0.rho), m6, obs, batch)
    fwi.gradient(*args)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fwi.gradient(*args)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    strategy = survey.last_strategy
    survey.release_states()
    for k, t in enumerate(saved):
        autotune.set_tile(k, t)
    return {"sample": f"{w0.name}: 16 shots x {steps} steps (bounded from 2500), full grid",
            "s_per_gradient": sec, "shots_per_batch": batch, "forward_wavefields": strategy,
            "gradient_cell_steps_per_s": survey.grid.N * float(w0.nt.sum()) / sec}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128, help="grid size (default 128: the C2 workload)")
    ap.add_argument("--nt", type=int, default=2000)
    ap.add_argument("--cpu-steps", type=int, default=16,
                    help="time steps of the FP64 CPU baseline sample (about 10 s of CPU work at 128^3)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-kernels", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--c3-full", action="store_true",
                    help="only measure the whole C3 workload (32 shots of 256x256x128, record lengths "
                         "1500-3000) batched vs one shot at a time, with tuned tiles, and print that")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-vs-plain forward comparison on a C3 sample")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only run a few dispatches of each sweep kernel and a 256 MiB copy "
                         "(for hardware-counter runs)")
    ap.add_argument("--profile-workload", default="c2", choices=("c2", "c3x4", "c4x1"),
                    help="with --profile-kernels: the C2 grid, 4 shots of the 256x256x128 C3 "
                         "grid, or one shot of the 200x200x100 C4 grid")
    ap.add_argument("--strategy", default="checkpoint", choices=("auto", "store", "hybrid", "checkpoint"),
                    help="forward wavefields for the adjoint: sqrt(nt) checkpoints with segment "
                         "re-forward (default), every state kept in HBM, the last states kept and "
                         "the first steps checkpointed, or whichever fits (auto)")
    ap.add_argument("--no-store", action="store_true", help="skip timing the store strategy beside it")
    ap.add_argument("--no-parity", action="store_true", help="skip the errors against FP64")
    ap.add_argument("--no-c4", action="store_true", help="skip the bounded C4 gradient sample")
    ap.add_argument("--no-graphs", action="store_true",
                    help="skip timing the same gradient from a persistent plan of HIP graphs")
    ap.add_argument("--plan-sample", action="store_true",
                    help="only time gradients eager vs from a persistent graph plan on 96^3 x 64 "
                         "shots and on 40^3 x 1 shot, and print that")
    ap.add_argument("--c4-steps", type=int, default=300,
                    help="record length of the bounded C4 gradient sample (16 shots, full grid)")
    ap.add_argument("--keep", type=int, default=None,
                    help="steps whose states a hybrid tape keeps (default: as many as fit)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU (the propagator shards by shot, with no cross-GPU "
              "work inside a gradient); run it with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from elastic_fm_vp_vs_rho_amd import workloads
    from elastic_fm_vp_vs_rho_amd.propagator import Grid, Propagator, ShotBatch, lame

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.c3_full:
        c3 = workloads.c3()
        print(json.dumps({"workload": c3.name, "batched_vs_plain":
                          batched_vs_plain(c3, c3.S, int(c3.nt.max()), tune=True)}))
        return
    if args.plan_sample:
        print(json.dumps({"plan_vs_eager": plan_sample()}))
        return
    if args.profile_kernels and args.profile_workload != "c2":
        big = workloads.c3(S=4) if args.profile_workload == "c3x4" else workloads.c4(S=1)
        profile_kernels(sample_propagator(big, big.S, 120), big.name + f", {big.S} shot(s), 120 steps")
        return
    wl = workloads.c2(n=args.n, nt=args.nt, anomaly=False)
    wt = workloads.c2(n=args.n, nt=args.nt, anomaly=True)
    grid = Grid(wl.nx, wl.ny, wl.nz, wl.dx, wl.dt, (wl.W,) * 6)
    shots = ShotBatch(grid, wl.src, wl.m6, wl.wavelets(), wl.nt, wl.rec, wl.cpml(), device=dev)
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    lam_t, mu_t = lame(f(wt.vp), f(wt.vs), f(wt.rho))
    true = Propagator(grid, lam_t, mu_t, f(wt.rho), shots)
    obs, _ = true.forward()
    lam, mu = lame(f(wl.vp), f(wl.vs), f(wl.rho))
    prop = Propagator(grid, lam, mu, f(wl.rho), shots)
    K = prop.checkpoint_interval()
    if args.profile_kernels:
        profile_kernels(prop, wl.name)
        return
    tiles = None
    if not args.no_tune:
        from elastic_fm_vp_vs_rho_amd import autotune
        rep = autotune.tune(prop, use_cache=False, verbose=True)
        tiles = {k: v["tile"] for k, v in rep.items()}
        print(json.dumps({"autotune": rep}), file=sys.stderr)

    strategy = args.strategy
    for _ in range(args.warmup):
        prop.gradient(obs, K, strategy, args.keep)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        J, g, glam, gmu, grho, gm6 = prop.gradient(obs, K, strategy, args.keep)
    torch.cuda.synchronize()
    T = time.perf_counter() - t0
    s_per_grad = T / max(args.steps, 1)
    N = grid.N
    nt = wl.ntmax
    value = N * nt * wl.S / s_per_grad

    line = {
        "metric": METRIC, "value": value, "unit": "gradient cell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * s_per_grad,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": wl.name, "grid": [wl.nx, wl.ny, wl.nz], "nt": nt, "shots": wl.S,
                   "receivers": int(wl.rec.shape[1]), "checkpoint_interval": K,
                   "gradient": "exact discrete adjoint", "tiles": tiles,
                   "forward_wavefields": prop.last_strategy,
                   "peak_hbm_gb": torch.cuda.max_memory_allocated() / 1e9},
        "s_per_gradient": s_per_grad,
        "forward_grid_pt_updates_per_s": None,
        "misfit": J,
    }
    # forward-only throughput of the same model (plain modelling path)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prop.forward()
    torch.cuda.synchronize()
    line["forward_grid_pt_updates_per_s"] = N * nt / (time.perf_counter() - t1)

    if not args.no_kernels:
        kt = time_kernels(prop)
        interior = (wl.nx - 4) * (wl.ny - 4) * (wl.nz - 4)
        per = {k: ALG_BYTES[k] * interior / (kt[k] * 1e-3) / 1e9 for k in ("stress", "velocity")}
        adj_gbs = ADJ_STEP_MODEL_BYTES * interior / (kt["adjoint_step"] * 1e-3) / 1e9
        design = ALG_BYTES["adj_velocity"] + ALG_BYTES["adj_stress"]
        dom = max(kt, key=lambda k: kt[k])
        if dom == "adjoint_step":
            achieved, name, alg = adj_gbs, "adjoint step (adj_velocity + adj_stress + points)", ADJ_STEP_MODEL_BYTES
        else:
            achieved, name, alg = per[dom], dom, ALG_BYTES[dom]
        traffic = measured_traffic(dom, wl)
        line["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                            "kernel": name,
                            "alg_bytes_per_cell": alg, "cells_per_launch": interior,
                            "traffic_over_alg": traffic / (alg * interior) if traffic else None,
                            "avg_ms": kt[dom], "all_kernels_ms": kt,
                            "all_kernels_gbs": dict(per, adjoint_step=adj_gbs)}
        if dom == "adjoint_step":
            line["roofline"]["design_bytes_per_cell"] = design
            line["roofline"]["design_gbs"] = design * interior / (kt[dom] * 1e-3) / 1e9
    if not args.no_store and strategy != "store":
        prop.gradient(obs, K, "store")
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        t1 = time.perf_counter()
        ns = max(1, min(args.steps, 2))
        for _ in range(ns):
            prop.gradient(obs, K, "store")
        torch.cuda.synchronize()
        ts = (time.perf_counter() - t1) / ns
        line["store"] = {"s_per_gradient": ts, "value": N * nt * wl.S / ts,
                         "peak_hbm_gb": torch.cuda.max_memory_allocated() / 1e9,
                         "note": "every forward state kept in HBM (no re-forward)"}
    if not args.no_graphs:
        gprop = Propagator(grid, lam, mu, f(wl.rho), shots, graphs=True)
        gprop._pool = prop._pool                     # the same forward-state buffers
        gprop.gradient(obs, K, strategy, args.keep)  # captures the plan's graphs
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ng = max(1, min(args.steps, 3))
        for _ in range(ng):
            gprop.gradient(obs, K, strategy, args.keep)
        torch.cuda.synchronize()
        tg = (time.perf_counter() - t1) / ng
        line["graphs"] = {"s_per_gradient": tg, "value": N * nt * wl.S / tg,
                          "graphs_held": gprop.graph_plan().graphs(),
                          "note": "same gradient, time loops replayed from a persistent plan of HIP "
                                  "graphs (bit-identical results)"}
        gprop.release_states()
        del gprop
    if not args.no_parity and os.path.exists(FIXTURE_C2):
        ref = np.load(FIXTURE_C2)["gathers"]
        n400 = ref.shape[1]
        got = g[0, :, :n400].double().cpu().numpy()
        line["err_vs_fp64"] = {
            "seismogram_rel_l2": float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
            "seismogram_sample": f"timed configuration, first {n400} of {nt} steps, all 64 receivers "
                                 f"x 3 components, vs the committed FP64 oracle fixture"}
    prop.release_states()
    if not args.no_c4:
        line["c4_gradient_sample"] = c4_gradient_sample(args.c4_steps)
    if not args.no_batched:
        c5 = workloads.c5(S=64, nt=256)
        c5.nt = np.array([128 + (37 * s) % 129 for s in range(c5.S)])      # uneven, 128..256
        line["batched_vs_plain"] = {"C3": batched_vs_plain(workloads.c3(), 8, 300),
                                    "C5": batched_vs_plain(c5, 64, int(c5.nt.max()))}
    if not args.no_parity:
        line.setdefault("err_vs_fp64", {}).update(gradient_vs_oracle())
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(wl, args.cpu_steps)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
