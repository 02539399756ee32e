"""Multi-step basal-friction inversion: the SSA step as a torch autograd
function with an explicit warm start, time-step checkpointing, and the
velocity-misfit objective of the adjoint-inversion workload.

The reference differentiates one step (glide/torch.py:4-59; the Greenland
example chains it with torch, examples/greenland/greenland_inverse.py:119-168).
Over many steps the states a backward pass needs are the forward solutions of
every step.  :func:`rollout` keeps only the state at every ``checkpoint``-th
step boundary and recomputes a segment's steps when the backward pass reaches
it (torch.utils.checkpoint).  A recomputed step must reproduce the stored one
bit for bit, so the step's initial iterate (velocities and active set) is an
input of the step here (:class:`GlideStepWarm`), not whatever the model last
held; the native solves are deterministic run to run.  The gradient then
equals the one of the uncheckpointed chain exactly.

Memory: a step keeps u, v, H, the active set and the grounded fraction (five
grid fields) for its backward; a checkpointed rollout of n steps keeps the
boundary states of n / checkpoint segments plus one segment's steps.
"""
from __future__ import annotations

from typing import Callable, Optional

import torch
from torch.utils.checkpoint import checkpoint as _checkpoint


class GlideStepWarm(torch.autograd.Function):
    """GlideStep (glide/torch.py:4-59) with the initial iterate as an input.

    ``GlideStepWarm.apply(t, dt, model, level, H_prev, bed, beta, smb, u0, v0,
    mask0, B=None) -> (u, v, H, mask)``: the forward solve starts from u0, v0
    (the velocities), H = H_prev and the active set mask0; the backward solves
    the adjoint at the returned state and returns the gradients with respect
    to H_prev, bed, beta and smb, and B when it is given (the rate factor, no
    counterpart in the reference's GlideStep; otherwise the model's own B is
    used and gets no gradient).  The converged solution does not depend on the
    initial iterate, so u0, v0 and mask0 get none; mask is not
    differentiable.  The adjoint solve starts from zero multipliers, and the
    gradients are those of the discrete problem (GLIDE_GRAD_EXACT: the
    reference's gradient kernel drops the boundary-facet terms of the last
    column and row).  The adjoint holds the grounded fraction fixed, as the
    reference's does, so the gradient is exact where the grounded fraction
    does not vary with H (fully grounded or fully floating ice)."""

    @staticmethod
    def forward(ctx, t, dt, model, level, H_prev, bed, beta, smb, u0, v0, mask0, B=None):
        ctx.t, ctx.dt, ctx.model, ctx.level = t, dt, model, level
        model.set_top_level(level)
        mg = model.mg
        fields = [(mg.state.H_prev, H_prev), (mg.state.H, H_prev), (mg.geometry.bed, bed), (mg.sliding.beta, beta),
                  (mg.forcing.smb, smb), (mg.state.u, u0), (mg.state.v, v0), (mg.state.mask, mask0)]
        if B is not None:
            fields.append((mg.rheology.B, B))
        for field, value in fields:
            field.set(value.detach(), start_level=level)
        model.forward(t, dt, update_geometry=False)
        st = mg[level].state
        u, v, H = st.u.data.clone(), st.v.data.clone(), st.H.data.clone()
        mask, phi = st.mask.data.clone(), st.phi.data.clone()
        ctx.has_B = B is not None
        ctx.save_for_backward(u, v, H, mask, phi, H_prev, bed, beta, smb,
                              B if B is not None else mg[level].rheology.B.data.clone())
        ctx.mark_non_differentiable(mask)
        return u, v, H, mask

    @staticmethod
    def backward(ctx, gu, gv, gH, gmask):
        model, level, dt = ctx.model, ctx.level, ctx.dt
        u, v, H, mask, phi, H_prev, bed, beta, smb, B = ctx.saved_tensors
        mg = model.mg
        for field, value in ((mg.state.H_prev, H_prev), (mg.geometry.bed, bed), (mg.sliding.beta, beta),
                             (mg.forcing.smb, smb), (mg.rheology.B, B), (mg.state.u, u), (mg.state.v, v),
                             (mg.state.H, H), (mg.state.phi, phi), (mg.state.mask, mask)):
            field.set(value.detach(), start_level=level)
        # the adjoint solve starts from zero multipliers, so that a backward
        # pass does not depend on what the model solved before it
        adj = mg[level].adjoint
        for lam in (adj.lambda_u, adj.lambda_v, adj.lambda_H):
            lam.data.zero_()
        z = torch.zeros_like
        model.backward(ctx.t, dt, dJdu=gu if gu is not None else z(u), dJdv=gv if gv is not None else z(v),
                       dJdH=gH if gH is not None else z(H), exact_gradients=True, compute_B_grad=ctx.has_B)
        g = mg[level]
        return (None, None, None, None, g.state.H_prev.grad.clone(), g.geometry.bed.grad.clone(),
                g.sliding.beta.grad.clone(), g.forcing.smb.grad.clone(), None, None, None,
                g.rheology.B.grad.clone() if ctx.has_B else None)


def misfit(u, v, u_obs, v_obs):
    """0.5 sum((u - u_obs)^2 + (v - v_obs)^2) over the facets."""
    return 0.5 * (((u - u_obs) ** 2).sum() + ((v - v_obs) ** 2).sum())


def gradient_penalty(log_beta, weight):
    """weight |grad log beta|^2 with one-sided differences between cell
    centres (per unit grid spacing)."""
    gy = torch.diff(log_beta, dim=0)
    gx = torch.diff(log_beta, dim=1)
    return weight * ((gy ** 2).sum() + (gx ** 2).sum())


def rollout(model, H0, bed, beta, smb, n_steps: int, dt: float, step_loss: Callable, t0: float = 0.0,
            level: int = 0, checkpoint: Optional[int] = 8, u0=None, v0=None, mask0=None, B=None):
    """sum over k of step_loss(k, u_k, v_k, H_k) over n_steps implicit steps
    from H0, each step's H the next one's H_prev, differentiable with respect
    to H0, bed, beta and smb, and the rate factor B when it is given.  With
    checkpoint = c, the rollout keeps the state at every c-th step boundary
    and recomputes each segment in the backward pass; checkpoint None or 0
    keeps every step's state.  Returns (loss, (u, v, H, mask) after the last
    step)."""
    g = model.mg[level]
    z = torch.zeros_like
    u = z(g.state.u.data) if u0 is None else u0
    v = z(g.state.v.data) if v0 is None else v0
    mask = z(g.state.mask.data) if mask0 is None else mask0

    def segment(k0, k1, H, u, v, mask, bed, beta, smb, B):
        loss = H.new_zeros(())
        for k in range(k0, k1):
            u, v, H, mask = GlideStepWarm.apply(t0 + k * dt, dt, model, level, H, bed, beta, smb, u, v, mask, B)
            loss = loss + step_loss(k, u, v, H)
        return loss, H, u, v, mask

    total = H0.new_zeros(())
    H = H0
    c = int(checkpoint) if checkpoint else n_steps
    for k0 in range(0, n_steps, c):
        k1 = min(k0 + c, n_steps)
        if checkpoint:
            loss, H, u, v, mask = _checkpoint(segment, k0, k1, H, u, v, mask, bed, beta, smb, B, use_reentrant=False)
        else:
            loss, H, u, v, mask = segment(k0, k1, H, u, v, mask, bed, beta, smb, B)
        total = total + loss
    return total, (u, v, H, mask)


def coupled_rollout(model, thermal, H0, E0, bed, beta, smb, Q_geo, E_surface, n_steps: int, dt: float,
                    step_loss: Callable, t0: float = 0.0, level: int = 0, checkpoint: Optional[int] = 8, u0=None,
                    v0=None, mask0=None, B0=None):
    """The coupled step of glide/model.py:156-213 over n_steps, differentiable
    end to end: each step solves the momentum balance with the current rate
    factor B (GlideStepWarm, H_prev the previous step's H), then the thermal
    step from that step's velocities and thickness (ThermalStep, E_prev the
    previous step's enthalpy), whose new rate factor (when
    ``thermal.update_rheology``) is the next momentum step's B.  ``thermal``
    is a ThermalModel on the grid of ``model.mg[level]``; dt is in years (the
    thermal step gets it in seconds).  Returns
    (sum over k of step_loss(k, u_k, v_k, H_k, E_k, B_k), (u, v, H, mask, E,
    B) after the last step); the loss is differentiable with respect to H0,
    E0, bed, beta, smb, Q_geo, E_surface and B0 (the first step's rate
    factor; when None, the model's current one without a gradient, which
    after a rollout is its last thermal step's, so repeated evaluations pass
    B0).  The
    gradient carries every coupling: the rate factor through the momentum
    solve, the velocities, thickness and sliding coefficient through the
    enthalpy (advection, omega, frictional and strain heating) and back.
    Checkpointing as in :func:`rollout` (the thermal step's backward rebuilds
    its state from its saved inputs, so recomputed segments and interleaved
    steps are exact)."""
    from .coupled import SEC_PER_YR
    from .torch import ThermalStep
    g = model.mg[level]
    if thermal.ops.grid is not g:
        raise ValueError("thermal must be a ThermalModel on the grid of model.mg[level]")
    z = torch.zeros_like
    u = z(g.state.u.data) if u0 is None else u0
    v = z(g.state.v.data) if v0 is None else v0
    mask = z(g.state.mask.data) if mask0 is None else mask0
    B = g.rheology.B.data.clone() if B0 is None else B0

    def segment(k0, k1, H, E, u, v, mask, B, bed, beta, smb, Q_geo, E_surface):
        loss = H.new_zeros(())
        for k in range(k0, k1):
            H_prev = H
            u, v, H, mask = GlideStepWarm.apply(t0 + k * dt, dt, model, level, H_prev, bed, beta, smb, u, v, mask, B)
            E, B_new = ThermalStep.apply(thermal, dt * SEC_PER_YR, E, Q_geo, E_surface, u, v, H, H_prev, beta, B)
            if thermal.update_rheology:
                B = B_new
            loss = loss + step_loss(k, u, v, H, E, B)
        return loss, H, E, u, v, mask, B

    total = H0.new_zeros(())
    H, E = H0, E0
    c = int(checkpoint) if checkpoint else n_steps
    for k0 in range(0, n_steps, c):
        k1 = min(k0 + c, n_steps)
        args = (k0, k1, H, E, u, v, mask, B, bed, beta, smb, Q_geo, E_surface)
        if checkpoint:
            loss, H, E, u, v, mask, B = _checkpoint(segment, *args, use_reentrant=False)
        else:
            loss, H, E, u, v, mask, B = segment(*args)
        total = total + loss
    return total, (u, v, H, mask, E, B)


__all__ = ["GlideStepWarm", "rollout", "coupled_rollout", "misfit", "gradient_penalty"]
