"""CPU oracle for the elastic FDTD propagator and its adjoint.

THIS PACKAGE IS TEST INFRASTRUCTURE.  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg may import it, and only as the checker
(or, for the CPU baseline, as the timed CPU restatement).  The product path in
``elastic_fm_vp_vs_rho_amd`` never imports anything from here.

Modules
-------
``elastic_fp64``  FP64 NumPy restatement of mainPropElastic3D_MT.py's forward
                  propagator, its reference-style adjoint/imaging, the exact
                  discrete adjoint and the host-side model/CPML/wavelet maps.
``twin``          ctypes wrapper of ``twin.c``: a scalar-loop FP64 restatement of
                  one forward step, used to pin the vectorised NumPy version.
"""
