"""Writes tests/golden/c1_gathers.npz: FP64 oracle seismograms of the C1
workload (64³, dx 10 m, homogeneous Vp 3000 / Vs 1730 / ρ 2400, Courant 0.5,
10-cell CPML on all faces, MT source at the centre, 10 Hz Ricker, 1000 steps,
4×4 receivers one level below the top CPML).

    python tests/golden/make_c1_fixture.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from elastic_fm_vp_vs_rho_amd import workloads  # noqa: E402
from oracle import elastic_fp64 as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_gathers.npz")


def main():
    wl = workloads.c1()
    lam, mu = wl.lame()
    med = O.Medium(lam, mu, wl.rho)
    cp = O.Cpml.from_dict({k: v[0] for k, v in wl.cpml().items()})
    w = wl.wavelets()[0]
    t0 = time.time()
    g, _ = O.forward(med, cp, wl.dt / wl.dx, tuple(wl.src[0]), wl.m6[0], w, wl.rec[0], wl.ntmax,
                     threads=O.default_threads())
    np.savez_compressed(OUT, gathers=g, m6=wl.m6[0], src=wl.src[0], rec=wl.rec[0], dt=wl.dt)
    print(f"wrote {OUT} in {time.time() - t0:.1f} s; max |v| {np.abs(g).max():.3e}")


if __name__ == "__main__":
    main()
