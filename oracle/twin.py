"""ctypes wrapper of twin.c (scalar-loop FP64 forward step).  Test infrastructure."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from .elastic_fp64 import FIELDS, PSI_T, PSI_V

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "libtwin.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
        P = ctypes.POINTER
        D = P(ctypes.c_double)
        _lib.twin_step.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double] + [P(D)] * 3 + [D] * 9
                                   + [ctypes.c_int] * 3 + [D, ctypes.c_double])
        _lib.twin_step.restype = None
    return _lib


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def step(F, lam, mu, rho, cp, c, src_pos, m6, amp):
    """One forward step on oracle ``Fields`` F (in place)."""
    lib = _load()
    ny, nz, nx = F.shape
    D = ctypes.POINTER(ctypes.c_double)
    f = (D * 9)(*[_dp(F.f[k]) for k in FIELDS])
    pv = (D * 9)(*[_dp(F.psi[k]) for k in PSI_V])
    pt = (D * 9)(*[_dp(F.psi[k]) for k in PSI_T])
    arrs = [np.ascontiguousarray(a, np.float64) for a in (lam, mu, rho, cp.ax, cp.bx, cp.ay, cp.by, cp.az, cp.bz)]
    m = np.ascontiguousarray(m6, np.float64)
    lib.twin_step(nx, ny, nz, float(c), f, pv, pt, *[_dp(a) for a in arrs],
                  int(src_pos[0]), int(src_pos[1]), int(src_pos[2]), _dp(m), float(amp))
