"""One coupled momentum + mass + enthalpy step and its adjoint on one GPU.

The reference's drivers call the pieces in turn (ThermalModel.pre_momentum,
IceDynamics.forward, ThermalModel.step, then IceDynamics.backward:
glide/model.py:65-94, 156-213; examples/greenland/greenland_inverse.py:119-168).
:class:`CoupledStep` runs the same sequence and, by default, the adjoint of
the converged coupled step of a loss of the step's velocities, thickness,
enthalpy and new rate factor: the enthalpy adjoint first, then the thermal
step's gradients with respect to its inputs (ThermalModel.adjoint), whose
velocity and thickness parts join the momentum adjoint's seeds and whose
sliding-coefficient and previous-thickness parts join the momentum
gradients.  With ``coupled_adjoint=False`` the enthalpy adjoint is solved on
its own seeds and, with ``concurrent``, on a second stream beside the
momentum adjoint (both then depend only on the forward solution).  The new
rate factor goes to a buffer of its own and is written to the grid at the end,
so the momentum adjoint sees the rate factor of its forward solve.  The step
runs on a stream of its own rather than the null stream, whose implicit
synchronisation with blocking streams would serialise the two.  Every stream
keeps the default priority.
"""
from __future__ import annotations

from typing import Callable, Optional

import torch

SEC_PER_YR = 365.25 * 86400.0


def velocity_loss_seeds(grid):
    """dJ/du, dJ/dv, dJ/dH of J = 0.5 |u|^2 + 0.5 |v|^2."""
    return grid.state.u.data, grid.state.v.data, None


class CoupledStep:
    """step(t, dt) advances the coupled state by dt years and leaves the
    momentum multipliers and parameter gradients of the step in the grid.

    model: IceDynamics; thermal: ThermalModel on the model's finest grid, or
    None (momentum only).  seeds(grid) gives the momentum adjoint's dJ/du,
    dJ/dv, dJ/dH (seeds=None: no momentum adjoint); thermal_seeds(thermal,
    grid) the enthalpy adjoint's dL/dE and dL/dB (default E and ones; a
    callable returning None skips the enthalpy adjoint).  forward, when
    given, replaces model.forward(t, dt, update_geometry=True).
    coupled_adjoint (with both adjoints): the momentum adjoint's seeds and
    gradients take the enthalpy's dependence on u, v, H, H_prev and beta, so
    the multipliers and gradients are those of the coupled step's loss.
    concurrent (decoupled adjoints only) puts the enthalpy work on the second
    stream."""

    def __init__(self, model, thermal=None, seeds: Callable = velocity_loss_seeds,
                 thermal_seeds: Optional[Callable] = None, enthalpy_adjoint_rtol: float = 1e-10,
                 enthalpy_adjoint_iters: int = 64, forward: Optional[Callable] = None, concurrent: bool = True,
                 coupled_adjoint: bool = True):
        self.model, self.thermal = model, thermal
        self.seeds = seeds
        self.coupled_adjoint = bool(coupled_adjoint)
        self.thermal_grads = None
        self.thermal_seeds = thermal_seeds if thermal_seeds is not None else self._default_thermal_seeds
        self.enthalpy_adjoint_rtol = float(enthalpy_adjoint_rtol)
        self.enthalpy_adjoint_iters = int(enthalpy_adjoint_iters)
        self.forward = forward
        self.concurrent = bool(concurrent)
        g = model.mg[0]
        dev = g.state.u.data.device
        self.main = torch.cuda.Stream(device=dev)
        self.side = torch.cuda.Stream(device=dev) if thermal is not None else None
        self.B_next = torch.empty_like(g.rheology.B.data) if thermal is not None else None
        self._ones = torch.ones_like(g.rheology.B.data) if thermal is not None else None
        self.ev_fwd = torch.cuda.Event()
        self.ev_thermal = torch.cuda.Event()
        self.enthalpy_report = None

    def _default_thermal_seeds(self, thermal, grid):
        return thermal.ops.enthalpy_state.E, self._ones

    def __call__(self, t: float, dt: float, report: bool = False):
        caller = torch.cuda.current_stream()
        self.main.wait_stream(caller)
        with torch.cuda.stream(self.main):
            self._step(t, dt, report)
        caller.wait_stream(self.main)

    def _thermal(self, dt, report, coupled):
        th = self.thermal
        g = self.model.mg[0]
        self.enthalpy_report = th.step(dt * SEC_PER_YR, report=report, B_out=self.B_next)
        seeds = self.thermal_seeds(th, g)
        self.thermal_grads = None
        if seeds is not None:
            dL_dE, dL_dB = seeds
            if coupled:  # the grid still holds the rate factor the step read
                self.thermal_grads = th.adjoint(dt * SEC_PER_YR, dL_dE, dL_dB, B_in=g.rheology.B.data,
                                                n_iter=self.enthalpy_adjoint_iters,
                                                relative_tolerance=self.enthalpy_adjoint_rtol)
            else:
                th.ops.adjoint_solve(dt * SEC_PER_YR, dL_dE=dL_dE, dL_dB=dL_dB, B_scale=th.B_scale,
                                     relative_tolerance=self.enthalpy_adjoint_rtol,
                                     n_iter=self.enthalpy_adjoint_iters, u2d=g.state.u.data, v2d=g.state.v.data,
                                     vel_div=SEC_PER_YR)

    def _step(self, t, dt, report):
        model, th = self.model, self.thermal
        if th is not None:
            th.pre_momentum()
        if self.forward is not None:
            self.forward(t, dt)
        else:
            model.forward(t, dt, update_geometry=True)
        g = model.mg[0]
        main = torch.cuda.current_stream()
        coupled = self.coupled_adjoint and th is not None and self.seeds is not None
        side = th is not None and self.concurrent and not coupled
        if th is not None:
            if side:
                self.ev_fwd.record(main)
                self.side.wait_event(self.ev_fwd)
                with torch.cuda.stream(self.side):
                    self._thermal(dt, report, False)
                    self.ev_thermal.record(self.side)
            else:
                self._thermal(dt, report, coupled)
        if self.seeds is not None:
            dJdu, dJdv, dJdH = self.seeds(g)
            tg = self.thermal_grads if coupled else None
            if tg is not None:
                def add(seed, extra):
                    return extra if seed is None else seed + extra
                dJdu, dJdv, dJdH = add(dJdu, tg["u"]), add(dJdv, tg["v"]), add(dJdH, tg["H"])
            model.backward(t, dt, dJdu=dJdu, dJdv=dJdv, dJdH=dJdH)
            if tg is not None:
                if tg["beta"] is not None:
                    g.sliding.beta.grad.add_(tg["beta"])
                g.state.H_prev.grad.add_(tg["H_prev"])
        if th is not None:
            if side:
                main.wait_event(self.ev_thermal)
            g.rheology.B.data.copy_(self.B_next)


__all__ = ["CoupledStep", "velocity_loss_seeds", "SEC_PER_YR"]
