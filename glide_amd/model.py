"""Public forward/adjoint API of the SSA step (glide/model.py:24-94)."""
from __future__ import annotations

import torch

from .grid import Grid
from .multigrid import FASAdjointSolver, FASCDSolver, Multigrid


class IceDynamics:
    """Momentum + mass conservation step solved by FAS-CD multigrid.

    Construct from (a) a Multigrid, (b) a Grid and a level count, or (c)
    ny, nx, dx and a level count (glide/model.py:25-39)."""

    def __init__(self, mg=None, n_levels=None, grid=None, ny=None, nx=None, dx=None, x0=0.0, y0=0.0, crs=None,
                 device=None):
        if mg is not None:
            self.mg = mg
        elif grid is not None and n_levels is not None:
            self.mg = Multigrid(n_levels, finest_grid=grid)
        elif ny and nx and dx and n_levels:
            self.mg = Multigrid(n_levels, ny=ny, nx=nx, dx=dx, x0=x0, y0=y0, crs=crs, device=device)
        else:
            raise ValueError("Must supply either (a) a multigrid object (b) a grid and number of levels "
                             "(c) ny/nx/dx and number of levels")
        self._forward_solver = None
        self._adjoint_solver = None
        self.top_level = 0
        self._post_forward_hooks = []

    @property
    def forward_solver(self) -> FASCDSolver:
        if self._forward_solver is None:
            self._forward_solver = FASCDSolver(self.mg)
        return self._forward_solver

    @property
    def adjoint_solver(self) -> FASAdjointSolver:
        if self._adjoint_solver is None:
            self._adjoint_solver = FASAdjointSolver(self.mg)
        return self._adjoint_solver

    def set_top_level(self, level):
        self.top_level = level

    def register_post_forward_hook(self, hook):
        self._post_forward_hooks.append(hook)

    def forward(self, t, dt, update_geometry=True):
        """Solve one implicit step (glide/model.py:65-72)."""
        self.forward_solver.solve(dt, start_level=self.top_level)
        top = self.mg.levels[self.top_level]
        if update_geometry:
            top.state.H_prev.data.copy_(top.state.H.data)
        for f in self._post_forward_hooks:
            f(t + dt)

    def backward(self, t, dt, dJdu=None, dJdv=None, dJdH=None, compute_beta_grad=True, compute_bed_grad=True,
                 compute_H_prev_grad=True, compute_smb_grad=True):
        """Adjoint solve Jᵀλ = -dJ/dx and parameter gradients (glide/model.py:74-94)."""
        top = self.mg.levels[self.top_level]
        ao = top.adjoint_operators
        for buf, g in ((ao.f_u, dJdu), (ao.f_v, dJdv), (ao.f_H, dJdH)):
            if g is None:
                buf.zero_()
            else:
                buf.copy_(-torch.as_tensor(g, device=buf.device, dtype=buf.dtype))
        self.adjoint_solver.solve(dt, start_level=self.top_level)
        which = ((1 if compute_beta_grad else 0) | (2 if compute_bed_grad else 0) |
                 (4 if compute_H_prev_grad else 0) | (8 if compute_smb_grad else 0))
        if which:
            ao._grad(which, dt)
