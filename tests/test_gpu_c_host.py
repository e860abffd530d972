"""The C host example (efm_host_example: a program without Python or torch
that links ROCm's HIP runtime and libefm.so) against the Python path on the
same float32 inputs: one shot's gathers, its exact gradient through the
allocation-free entry point and ∂J/∂M bit for bit, the misfit to the order of
its atomic sums.  The host runs the receivers through the separate per-step
kernels (no receiver mask), the Python path inside the sweeps."""
import os
import subprocess

import numpy as np
import pytest
import torch

from tests.gpu_helpers import propagator_of
from tests.problems import tiny

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "elastic_fm_vp_vs_rho_amd", "efm_host_example")


def _write_input(path, prop, obs):
    d, sh = prop.dims(), prop.shots
    cpu = lambda t: t.detach().cpu().numpy()
    with open(path, "wb") as f:
        np.array([d.nx, d.ny, d.nz, d.wx_lo, d.wx_hi, d.wy_lo, d.wy_hi, d.wz_lo, d.wz_hi,
                  d.ntmax, d.nrec], np.int32).tofile(f)
        np.array([d.dtdx], np.float32).tofile(f)
        for t in (prop.lam, prop.mu, prop.rho):
            cpu(t).astype(np.float32).tofile(f)
        for k in ("ax", "bx", "ay", "by", "az", "bz"):
            cpu(sh.cpml[k][0]).astype(np.float32).tofile(f)
        cpu(sh.src[0]).astype(np.int32).tofile(f)
        cpu(sh.m6[0]).astype(np.float32).tofile(f)
        cpu(sh.wavelet[0]).astype(np.float32).tofile(f)
        cpu(sh.rec[0]).astype(np.int32).tofile(f)
        cpu(obs[0]).astype(np.float32).tofile(f)


def test_c_host_matches_python_path(gpu_lib, tmp_path):
    assert os.path.exists(EXE), "efm_host_example is not built (run __graft_entry__.build())"
    p = tiny(n=(30, 26, 22), nt=70, W=4, seed=43, nrec=7, free_top=True)
    prop = propagator_of(p)
    tape = prop.record(strategy="store")
    gathers = tape["gathers"]
    obs = (0.6 * gathers).contiguous()
    resid, J = prop.misfit(gathers, obs)
    glam, gmu, grho, gm6 = prop.backprop(tape, resid)
    torch.cuda.synchronize()

    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(fin, prop, obs)
    print(f"efm_host_example: grid {p.nx}x{p.ny}x{p.nz}, nt {p.nt}", flush=True)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout

    ng, N = gathers[0].numel(), prop.grid.N
    raw = fout.read_bytes()
    off = 0

    def take(n, dt):
        nonlocal off
        a = np.frombuffer(raw, dt, n, off)
        off += n * np.dtype(dt).itemsize
        return a

    g_c = take(ng, np.float32)
    J_c = float(take(1, np.float64)[0])
    vols = [take(N, np.float32) for _ in range(3)]
    gm6_c = take(6, np.float64)
    assert off == len(raw)
    assert np.array_equal(g_c, gathers[0].cpu().numpy().ravel())
    assert abs(J_c - float(J)) <= 1e-12 * float(J)
    for got, want, name in zip(vols, (glam, gmu, grho), ("lam", "mu", "rho")):
        assert np.array_equal(got, want.cpu().numpy().ravel()), name
    assert np.array_equal(gm6_c, gm6[0].cpu().numpy())
    assert np.abs(vols[0]).max() > 0 and np.abs(gm6_c).max() > 0
