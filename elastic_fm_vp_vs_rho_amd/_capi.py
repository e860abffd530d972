"""ctypes binding of libefm.so (the C ABI declared in include/efm.h).

The library is linked without a HIP runtime dependency; it must share the HIP
runtime that torch already loaded, so that torch's streams and device pointers
are valid inside it.  ``lib()`` promotes torch's runtime to global symbol scope,
loads libefm.so, and verifies that exactly one HIP runtime is mapped in the
process.  There is no fallback: if the library is missing, every entry point of
this package raises ``EfmError``.
"""
from __future__ import annotations

import ctypes
import os
import threading

c_int, c_float, c_ll, c_void_p = ctypes.c_int, ctypes.c_float, ctypes.c_longlong, ctypes.c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libefm.so")


class EfmError(RuntimeError):
    pass


class EfmDims(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("nz", c_int), ("nshot", c_int),
                ("wx_lo", c_int), ("wx_hi", c_int), ("wy_lo", c_int), ("wy_hi", c_int),
                ("wz_lo", c_int), ("wz_hi", c_int), ("ntmax", c_int), ("nrec", c_int),
                ("dtdx", c_float)]


class EfmMedium(ctypes.Structure):
    _fields_ = [("lam", c_void_p), ("mu", c_void_p), ("rho", c_void_p)]


class EfmState(ctypes.Structure):
    _fields_ = [(k, c_void_p) for k in ("vx", "vy", "vz", "txx", "tyy", "tzz", "txy", "txz", "tyz",
                                        "psi_x", "psi_y", "psi_z")]


class EfmSmooth(ctypes.Structure):
    _fields_ = [("sigma", c_float), ("truncate", c_float), ("margin", c_int * 6)]


class EfmShots(ctypes.Structure):
    _fields_ = [(k, c_void_p) for k in ("ax", "bx", "ay", "by", "az", "bz", "src", "m6",
                                        "wavelet", "nt", "rec", "recmask")]


P = ctypes.POINTER
_SIGS = {
    "efm_version": ([], c_int),
    "efm_last_error": ([ctypes.c_char_p, c_int], c_int),
    "efm_set_tile": ([c_int, c_int, c_int, c_int, c_int], c_int),
    "efm_get_tile": ([c_int, P(c_int), P(c_int), P(c_int), P(c_int)], c_int),
    "efm_component_pad": ([], c_int),
    "efm_tile_variants": ([P(c_int), c_int], c_int),
    "efm_residual": ([c_void_p] * 4 + [c_ll, c_void_p], c_int),
    "efm_model_lame": ([c_void_p] * 5 + [c_ll, c_void_p], c_int),
    "efm_model_lame_adjoint": ([c_void_p] * 9 + [c_ll, c_void_p], c_int),
    "efm_stress_step": ([P(EfmDims), P(EfmMedium), P(EfmShots), P(EfmState), P(EfmState), c_int,
                         c_void_p], c_int),
    "efm_velocity_step": ([P(EfmDims), P(EfmMedium), P(EfmShots), P(EfmState), P(EfmState), c_int,
                           c_void_p], c_int),
    "efm_record": ([P(EfmDims), P(EfmShots), P(EfmState), c_void_p, c_int, c_void_p], c_int),
    "efm_forward": ([P(EfmDims), P(EfmMedium), P(EfmShots), P(EfmState), c_void_p, c_void_p, c_int,
                     c_int, c_int, c_void_p], c_int),
    "efm_forward_graph": ([P(EfmDims), P(EfmMedium), P(EfmShots), P(EfmState), c_void_p, c_int, c_int,
                           c_int, c_void_p, c_void_p], c_int),
    "efm_forward_states": ([P(EfmDims), P(EfmMedium), P(EfmShots), P(P(EfmState)), c_void_p, c_int,
                            c_int, c_void_p], c_int),
    "efm_derived_medium": ([P(EfmDims), P(EfmMedium), c_void_p, c_void_p, c_void_p], c_int),
    "efm_adjoint": ([P(EfmDims), P(EfmMedium), c_void_p, c_void_p, P(EfmShots), P(P(EfmState)),
                     P(EfmState), P(EfmState), c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                     c_void_p], c_int),
    "efm_gradient_static": ([P(EfmDims), P(EfmMedium), P(EfmShots), P(P(EfmState)), P(EfmState)]
                            + [c_void_p] * 6, c_int),
    "efm_gradient_combine": ([P(EfmDims), P(EfmMedium), c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_void_p], c_int),
    "efm_inject_residual": ([P(EfmDims), P(EfmShots), P(EfmState), c_void_p, c_int, c_void_p], c_int),
    "efm_image_mt_ref": ([P(EfmDims), P(EfmMedium), P(EfmShots), P(EfmState), c_void_p, c_int,
                          c_void_p], c_int),
    "efm_backward_mt_ref": ([P(EfmDims), P(EfmMedium), P(EfmShots), P(EfmState), c_void_p, c_void_p,
                             c_int, c_void_p], c_int),
    "efm_backward_remake_ref": ([P(EfmDims), P(EfmMedium), P(EfmShots), P(EfmState), c_void_p,
                                 c_void_p, c_int, P(EfmState), c_void_p, c_float, c_float, c_float,
                                 c_float, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "efm_backward_remake_ref_steps": ([P(EfmDims), P(EfmMedium), P(EfmShots), P(EfmState), c_void_p,
                                       c_void_p, c_int, P(EfmState), c_void_p, c_float, c_float, c_float,
                                       c_float, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p],
                                      c_int),
    "efm_border_floats": ([P(EfmDims), c_int], c_ll),
    "efm_condition_gradient": ([P(EfmSmooth), c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                c_void_p], c_int),
    "efm_model_update": ([c_void_p, c_void_p, ctypes.c_double, c_void_p, c_void_p, ctypes.c_double,
                          c_void_p, c_void_p] + [c_int] * 4 + [c_void_p] * 6, c_int),
    "efm_model_rho_nafe_drake": ([c_void_p, c_void_p, c_ll, c_void_p], c_int),
    "efm_volume_axpy": ([c_float, c_void_p, c_void_p, c_ll, c_void_p], c_int),
    "efm_xcorr_residual": ([c_void_p] * 4 + [c_int] * 3 + [c_void_p], c_int),
    "efm_plan_create": ([P(EfmDims), P(EfmMedium), P(EfmShots), c_int, P(c_void_p)], c_int),
    "efm_plan_destroy": ([c_void_p], c_int),
    "efm_plan_graphs": ([c_void_p], c_int),
    "efm_plan_forward": ([c_void_p, P(EfmState), c_void_p, c_int, c_int, c_void_p], c_int),
    "efm_plan_forward_states": ([c_void_p, P(P(EfmState)), c_void_p, c_int, c_int, c_void_p], c_int),
    "efm_plan_adjoint": ([c_void_p, c_void_p, c_void_p, P(P(EfmState)), P(EfmState), P(EfmState), c_void_p,
                          c_void_p, c_void_p, c_int, c_int, c_int, c_void_p], c_int),
    # initializer network (include/efm_initnet.h)
    "efm_net_cast_bf16": ([c_void_p, c_void_p, c_ll, c_void_p], c_int),
    "efm_net_gemm": ([c_int, c_int, c_int, c_void_p, c_ll, c_ll, c_void_p, c_ll, c_ll, c_void_p, c_ll,
                      c_void_p, c_void_p, c_int, c_void_p], c_int),
    "efm_net_im2col": ([c_void_p, c_void_p] + [c_int] * 6 + [c_void_p], c_int),
    "efm_net_col2im": ([c_void_p, c_void_p] + [c_int] * 6 + [c_void_p], c_int),
    "efm_net_relu_backward": ([c_void_p] * 4 + [c_ll, c_void_p], c_int),
    "efm_net_bias_grad": ([c_void_p, c_void_p, c_int, c_ll, c_void_p], c_int),
    "efm_net_mse": ([c_void_p] * 5 + [c_ll, c_float, c_void_p], c_int),
    "efm_net_adam": ([c_void_p] * 5 + [c_ll] + [c_float] * 6 + [c_void_p], c_int),
    "efm_net_flatten": ([c_void_p, c_void_p, c_int, c_int, c_int, c_void_p], c_int),
    "efm_net_unflatten": ([c_void_p, c_void_p, c_int, c_int, c_int, c_void_p], c_int),
    "efm_net_pack_gathers": ([c_void_p] + [c_int] * 5 + [c_void_p, c_ll, c_void_p], c_int),
    "efm_net_gather_batch": ([c_void_p, c_void_p, c_ll, c_void_p] + [c_int] * 4 + [c_void_p] * 3, c_int),
}

_lock = threading.Lock()
_lib = None


def _hip_runtimes_mapped():
    try:
        with open("/proc/self/maps") as f:
            return sorted({ln.split()[-1] for ln in f if "libamdhip64" in ln})
    except OSError:
        return []


def lib():
    """Load (once) and return the ctypes handle of libefm.so."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        import torch  # noqa: F401  (loads torch's HIP runtime)
        tdir = os.path.join(os.path.dirname(torch.__file__), "lib")
        hip = os.path.join(tdir, "libamdhip64.so")
        if not os.path.exists(hip):
            raise EfmError(f"torch's HIP runtime not found at {hip}: a ROCm build of torch is required")
        ctypes.CDLL(hip, mode=ctypes.RTLD_GLOBAL)
        if not os.path.exists(LIB_PATH):
            raise EfmError(f"{LIB_PATH} is not built; run `python -m elastic_fm_vp_vs_rho_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        rts = _hip_runtimes_mapped()
        if len(rts) > 1:
            raise EfmError(f"more than one HIP runtime is mapped into the process: {rts}")
        _lib = L
        return L


def last_error():
    buf = ctypes.create_string_buffer(1024)
    lib().efm_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc, what=""):
    if rc != 0:
        raise EfmError(f"{what}: {last_error()} (code {rc})")
    return rc


def exported_symbols():
    return list(_SIGS)
