"""gfx950 (MI355X) elastic velocity–stress FDTD propagator and exact adjoint.

Modules: ``driver`` (the reference's Elastic3D, Model, Wavelet, CPML, Geometry,
… classes over the HIP library), ``lbfgs`` (the reference's L-BFGS interface),
``propagator`` (grids, shot batches, wavefields, forward runs and exact-adjoint
gradients), ``ops`` (torch.autograd function), ``fwi`` (batched joint FWI + MT
inversion), ``autotune`` (tile autotuner), ``initnet`` / ``initnet_data``
(learned MT/location initializer), ``model`` (CPML profiles, wavelets, Lamé
maps), ``workloads`` (the C1–C5 configurations), ``_capi`` (ctypes binding of
libefm.so) and ``build`` (compiles the HIP sources in place).
Every call that would launch a kernel raises EfmError when libefm.so is not
built; there is no CPU fallback.
"""
