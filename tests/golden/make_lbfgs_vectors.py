"""Writes tests/golden/lbfgs_vectors.npz from the reference's own lbfgs.py.

Run in the build container only (the reference tree is not present on GPU
machines):  python tests/golden/make_lbfgs_vectors.py

The history shift of compute_sk_yk (slot L-1 never written) and the
slot-1 scaling of L_BFGS are conventions no closed form pins, so the fixture
records the reference's outputs for a seeded sequence of iterations.
"""
import os
import sys

import numpy as np

REF = "/root/reference/Elastic3DGPU_MT_VP_VS_RHO_MPI"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lbfgs_vectors.npz")


def main():
    sys.path.insert(0, REF)
    import lbfgs as ref  # noqa: E402  (the reference module, imported only here)
    rng = np.random.default_rng(2024)
    data = {}
    for L in (2, 4, 5):
        shape = (3, 4, 5)
        sk = np.zeros((L,) + shape, np.float32)
        yk = np.zeros((L,) + shape, np.float32)
        ms, gs, rs, sks, yks = [], [], [], [], []
        m = rng.standard_normal(shape).astype(np.float32)
        for it in range(7):
            g = (0.5 * m + 0.1 * rng.standard_normal(shape)).astype(np.float32)
            ref.compute_sk_yk(sk, yk, m.copy(), g.copy(), L, it)
            r = ref.L_BFGS(sk, yk, g.copy(), L, it) if it > 0 else np.zeros(shape, np.float32)
            ms.append(m.copy()); gs.append(g.copy()); rs.append(np.asarray(r, np.float64))
            sks.append(sk.copy()); yks.append(yk.copy())
            m = (m - 0.3 * np.asarray(r, np.float32) - 0.05 * g).astype(np.float32)
        data[f"L{L}_m"] = np.stack(ms); data[f"L{L}_g"] = np.stack(gs); data[f"L{L}_r"] = np.stack(rs)
        data[f"L{L}_sk"] = np.stack(sks); data[f"L{L}_yk"] = np.stack(yks)
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, sorted(data))


if __name__ == "__main__":
    main()
