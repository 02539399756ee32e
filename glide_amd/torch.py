"""torch.autograd surfaces: one SSA step (glide/torch.py:4-59) and one
thermal step.

``GlideStep.apply(t, dt, model, level, H_prev, bed, beta, smb) -> (u, v, H)``
runs the native forward solve; its backward runs the native adjoint solve and
returns the gradients with respect to (H_prev, bed, beta, smb).  As in the
reference, the forward solve is warm-started from the velocities the model
currently holds, and H starts from H_prev.
"""
from __future__ import annotations

import torch


class GlideStep(torch.autograd.Function):

    @staticmethod
    def forward(ctx, t, dt, model, level, H_prev, bed, beta, smb):
        ctx.t, ctx.dt, ctx.model, ctx.level = t, dt, model, level
        model.set_top_level(level)
        mg = model.mg
        mg.state.H_prev.set(H_prev.detach(), start_level=level)
        mg.state.H.set(H_prev.detach(), start_level=level)
        mg.geometry.bed.set(bed.detach(), start_level=level)
        mg.sliding.beta.set(beta.detach(), start_level=level)
        mg.forcing.smb.set(smb.detach(), start_level=level)
        model.forward(t, dt, update_geometry=False)
        st = mg[level].state
        u, v, H = st.u.data.clone(), st.v.data.clone(), st.H.data.clone()
        mask, phi = st.mask.data.clone(), st.phi.data.clone()
        ctx.save_for_backward(u, v, H, mask, phi, H_prev, bed, beta, smb)
        return u, v, H

    @staticmethod
    def backward(ctx, gu, gv, gH):
        model, level, dt = ctx.model, ctx.level, ctx.dt
        u, v, H, mask, phi, H_prev, bed, beta, smb = ctx.saved_tensors
        mg = model.mg
        mg.state.H_prev.set(H_prev.detach(), start_level=level)
        mg.geometry.bed.set(bed.detach(), start_level=level)
        mg.sliding.beta.set(beta.detach(), start_level=level)
        mg.forcing.smb.set(smb.detach(), start_level=level)
        mg.state.u.set(u, start_level=level)
        mg.state.v.set(v, start_level=level)
        mg.state.H.set(H, start_level=level)
        mg.state.phi.set(phi, start_level=level)
        mg.state.mask.set(mask, start_level=level)
        zeros = lambda t: torch.zeros_like(t)
        model.backward(ctx.t, dt, dJdu=gu if gu is not None else zeros(u), dJdv=gv if gv is not None else zeros(v),
                       dJdH=gH if gH is not None else zeros(H))
        g = mg[level]
        return (None, None, None, None, g.state.H_prev.grad.clone(), g.geometry.bed.grad.clone(),
                g.sliding.beta.grad.clone(), g.forcing.smb.grad.clone())


class ThermalStep(torch.autograd.Function):
    """torch.autograd surface of one thermal step (no counterpart in the
    reference, whose enthalpy solve has no adjoint).

    ``ThermalStep.apply(thermal, dt, E_prev, Q_geo, E_surface[, u, v, H,
    H_prev, beta, B]) -> (E, B_new)`` sets the previous enthalpy (also the
    initial iterate), the geothermal flux and the surface enthalpy of
    ``thermal`` (a ``ThermalModel``), and, where given, the SSA velocities u
    (ny, nx+1) and v (ny+1, nx) in m/a, the new thickness H, the pre-momentum
    thickness H_prev, the sliding coefficient beta and the rate factor B (head
    units) the step reads; the others keep the model's current values (after
    ``pre_momentum`` and the momentum step).  It runs ``thermal.step(dt)`` and
    returns the new enthalpy and rate factor (zeros when ``update_rheology``
    is off).  The backward restores the step's inputs and converged enthalpy
    into the model (``rebuild_step``, so other steps may run in between) and
    solves the adjoint of the converged solve, so it is exact when the
    forward sweeps converge (tighten ``thermal.ops.smoother_config``
    tolerances and ``thermal.n_smooth`` for gradient work); it returns the
    gradients of every tensor input, the momentum step's outputs u, v, H
    included, which closes the adjoint of the coupled step with ``GlideStep``.
    It leaves the model's grid holding the step's inputs."""

    @staticmethod
    def forward(ctx, thermal, dt, E_prev, Q_geo, E_surface, u=None, v=None, H=None, H_prev=None, beta=None, B=None):
        ops = thermal.ops
        grid = ops.grid
        ops.enthalpy_state.E_prev.copy_(E_prev.detach())
        ops.enthalpy_state.E.copy_(E_prev.detach())
        ops.enthalpy_forcing.Q_geo.copy_(Q_geo.detach())
        ops.enthalpy_forcing.E_surface.copy_(E_surface.detach())
        for given, dst in ((u, grid.state.u.data), (v, grid.state.v.data), (H, grid.state.H.data),
                           (H_prev, ops.H_prev), (beta, grid.sliding.beta.data), (B, grid.rheology.B.data)):
            if given is not None:
                dst.copy_(given.detach())
        inputs = tuple(t.clone() for t in (grid.state.u.data, grid.state.v.data, grid.state.H.data, ops.H_prev,
                                            grid.sliding.beta.data, grid.rheology.B.data))
        thermal.step(dt)
        E = ops.enthalpy_state.E.clone()
        B_new = grid.rheology.B.data.clone() if thermal.update_rheology else torch.zeros_like(Q_geo)
        ctx.thermal, ctx.dt = thermal, dt
        ctx.save_for_backward(E_prev, Q_geo, E_surface, E, *inputs)
        return E, B_new

    @staticmethod
    def backward(ctx, gE, gB):
        thermal = ctx.thermal
        ops = thermal.ops
        grid = ops.grid
        E_prev, Q_geo, E_surface, E, u, v, H, H_prev, beta, B_in = ctx.saved_tensors
        ops.enthalpy_state.E_prev.copy_(E_prev.detach())
        ops.enthalpy_forcing.Q_geo.copy_(Q_geo.detach())
        ops.enthalpy_forcing.E_surface.copy_(E_surface.detach())
        for src, dst in ((u, grid.state.u.data), (v, grid.state.v.data), (H, grid.state.H.data), (H_prev, ops.H_prev),
                         (beta, grid.sliding.beta.data), (B_in, grid.rheology.B.data)):
            dst.copy_(src)
        thermal.rebuild_step(ctx.dt, E)
        if gE is None:
            gE = torch.zeros_like(E)
        use_B = gB is not None and thermal.update_rheology
        g = thermal.adjoint(ctx.dt, gE.contiguous(), gB.contiguous() if use_B else None, B_in=B_in)
        need = ctx.needs_input_grad  # one entry per argument given

        def pick(idx, key):
            return g[key] if idx < len(need) and need[idx] and g.get(key) is not None else None
        grads = (None, None, g["E_prev"], g["Q_geo"], g["E_surface"], pick(5, "u"), pick(6, "v"), pick(7, "H"),
                 pick(8, "H_prev"), pick(9, "beta"), pick(10, "B"))
        return grads[:len(need)]
