"""Seismograms as a differentiable torch function of (Vp, Vs, ρ, M).

``elastic_modeling(vp, vs, rho, m6, grid, shots)`` returns the gathers
(S, 3, ntmax, nrec) of a shot batch.  Its backward is the exact discrete
adjoint of the HIP propagator (adjoint.hip): the forward keeps every state when
they fit in HBM, else the states of the last steps plus √-checkpoints of the
first (or √nt checkpoints alone), and the backward rebuilds the checkpointed
segments; the backward runs the transposed sweeps, which
cross-correlate forward and adjoint wavefields into ∂/∂λ, ∂/∂μ, ∂/∂ρ and
∂/∂M; efm_model_lame_adjoint maps them to Vp, Vs, ρ.  No torch arithmetic is
involved on either pass.
"""
from __future__ import annotations

import torch

from .propagator import Grid, Propagator, ShotBatch, lame, lame_adjoint

# Idle forward-state pools, one per (grid, shot count, device).  A forward that
# records a tape takes the pool (so no other call can overwrite its states) and
# its backward gives it back; repeated gradients then neither allocate nor clear
# their states again.
_IDLE_POOLS = {}


def _take_pool(key):
    return _IDLE_POOLS.pop(key, [])


def _give_pool(key, pool):
    if pool:
        _IDLE_POOLS[key] = pool


def release_state_pools():
    """Free the forward states kept between differentiable calls."""
    _IDLE_POOLS.clear()


class ElasticModeling(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, vs, rho, m6, grid: Grid, shots: ShotBatch, K):
        for t in (vp, vs, rho):
            if t.dtype != torch.float32 or not t.is_cuda:
                raise TypeError("vp, vs, rho must be float32 device tensors")
        vp, vs, rho = vp.contiguous(), vs.contiguous(), rho.contiguous()
        shots.m6.copy_(m6.detach().reshape(shots.S, 6).to(torch.float32))
        lam, mu = lame(vp, vs, rho)
        if not any(ctx.needs_input_grad[:4]):
            return Propagator(grid, lam, mu, rho, shots).forward()[0]   # no backward follows
        key = (grid, shots.S, vp.device)
        pool = _take_pool(key)
        prop = Propagator(grid, lam, mu, rho, shots, pool=pool)
        tape = prop.record(K, strategy="checkpoint" if K is not None else "auto")
        ctx.prop, ctx.tape, ctx.pool = prop, tape, (key, pool)
        ctx.save_for_backward(vp, vs, rho)
        ctx.m6_dtype, ctx.m6_shape = m6.dtype, m6.shape
        # the backward may re-run forward segments from checkpoints: it needs the
        # moment tensors of this call even if the batch is reused meanwhile
        ctx.m6_used = shots.m6.clone()
        return tape["gathers"]

    @staticmethod
    def backward(ctx, grad_g):
        vp, vs, rho = ctx.saved_tensors
        shots = ctx.prop.shots
        m6_now = shots.m6.clone()
        shots.m6.copy_(ctx.m6_used)
        try:
            glam, gmu, grho, gm6 = ctx.prop.backprop(ctx.tape, grad_g)
        finally:
            shots.m6.copy_(m6_now)
        ctx.tape = None
        ctx.prop.release_states()
        _give_pool(*ctx.pool)
        ctx.pool = None
        gvp, gvs, grho_total = lame_adjoint(vp, vs, rho, glam, gmu, grho)
        return (gvp, gvs, grho_total, gm6.to(ctx.m6_dtype).reshape(ctx.m6_shape), None, None, None)


def elastic_modeling(vp, vs, rho, m6, grid: Grid, shots: ShotBatch, K=None):
    """Differentiable gathers of ``shots`` in the medium (vp, vs, rho)."""
    return ElasticModeling.apply(vp, vs, rho, m6, grid, shots, K)
