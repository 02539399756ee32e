"""Mixed-precision forward and adjoint solves (no counterpart in the reference).

The reference solves in FP32 only (glide/operators.py:24-29); this package's
plain form is FP64.  The mixed form keeps FP64 where accuracy is decided and
runs the bulk of the work in FP32:

* an FP64 defect-correction loop on the finest level: the FP64 residual (with
  the active set) decides convergence and the FP64 state owns the solution
  and the active set;
* per outer iteration, FP32 FAS-CD V-cycles (glide/multigrid.py:536-645) on an
  FP32 copy of the hierarchy, smoothed by red-black (multicolour) Vanka passes
  whose local Newton iteration takes a few relaxed full steps, with the
  coarsest level's passes run in FP64;
* the adjoint likewise: FP64 transpose-operator residual, FP32 FAS adjoint
  V-cycles (glide/multigrid.py:905-973) for the correction with red-black
  adjoint passes, FP64 accumulation.

The FP32 copy of the state holds the solution only to FP32 rounding of the
state (~6e-8 relative), so once an FP32 iteration gains less than a factor
0.6 the forward solve finishes with FP64 V-cycles of the same red-black form.
An outer iteration that does not reduce the FP64 residual is undone and
replaced by one FP64 V-cycle of the reference form.  The FP32 buffers are
allocated on first use, per level, beside the FP64 ones (they are the FP32
mirror of every buffer a V-cycle touches).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import ptr, ptr32

F32 = torch.float32

_FORWARD_CELL = ("H", "H_prev", "phi", "mask", "bed", "B", "beta", "f_H", "r_H", "F_H", "gamma", "w_H", "y_H",
                 "z_H", "chi", "cphi")
_FORWARD_U = ("u", "f_u", "r_u", "F_u", "w_u", "y_u", "z_u")
_FORWARD_V = ("v", "f_v", "r_v", "F_v", "w_v", "y_v", "z_v")
_ADJ_CELL = ("lambda_H", "af_H", "ar_H", "vjp_H", "wl_H", "zl_H")
_ADJ_U = ("lambda_u", "af_u", "ar_u", "vjp_u", "wl_u", "zl_u")
_ADJ_V = ("lambda_v", "af_v", "ar_v", "vjp_v", "wl_v", "zl_v")


class Level32:
    """FP32 buffers of one level, shaped like the FP64 level's."""

    def __init__(self, grid, adjoint: bool):
        ny, nx, dev = grid.ny, grid.nx, grid.device

        def z(shape):
            return torch.zeros(shape, dtype=F32, device=dev)

        self.grid = grid
        self.buf = {}
        for names, shape in ((_FORWARD_CELL, (ny, nx)), (_FORWARD_U, (ny, nx + 1)), (_FORWARD_V, (ny + 1, nx))):
            for n in names:
                self.buf[n] = z(shape)
        self.buf["work"] = z((5, ny, nx))
        self.adjoint = False
        if adjoint:
            self.add_adjoint()

    def add_adjoint(self):
        if self.adjoint:
            return
        g = self.grid
        ny, nx, dev = g.ny, g.nx, g.device
        for names, shape in ((_ADJ_CELL, (ny, nx)), (_ADJ_U, (ny, nx + 1)), (_ADJ_V, (ny + 1, nx))):
            for n in names:
                self.buf[n] = torch.zeros(shape, dtype=F32, device=dev)
        self.buf["adj_cache"] = torch.zeros((26, ny, nx), dtype=F32, device=dev)
        self.adjoint = True

    def struct(self) -> _lib.Level:
        g = self.grid
        L = _lib.Level()
        L.ny, L.nx, L.dx = g.ny, g.nx, g.dx
        L.phys = g.physics()
        for n, t in self.buf.items():
            setattr(L, n, ptr32(t))
        return L


@dataclass
class MixedConfig:
    """Settings of the mixed form.  The outer loop stops once the FP64 residual
    norm is below relative_tolerance times the initial one (or
    absolute_tolerance), after at most max_outer iterations; each iteration
    runs inner_vcycles FP32 V-cycles with the given counts; the smoother's
    local Newton iteration takes newton_steps steps of `relaxation` (forward)
    and the adjoint passes move the multipliers by adjoint_omega times their
    patch solution; fp64_coarsest runs the coarsest level's passes in FP64.
    asynchronous with both tolerances <= 0 runs exactly max_outer FP32
    iterations without a host wait and without a report (no norms, no
    fallback); the solve then returns None.  krylov_dim > 0 runs the adjoint
    outer iteration as flexible GMRES restarted every krylov_dim iterations
    (the FP32 correction V-cycles precondition it) and mixes the forward
    solve's FP64 finishing cycles by Anderson acceleration of depth
    min(krylov_dim, 6); 0 keeps the plain loops.  The forward and adjoint
    solves of a hierarchy share the workspace, so they must not overlap.
    fp64_newton_steps / fp64_relaxation set the local Newton iteration of
    the forward solve's FP64 finishing cycles (default one full step: near
    the solution the local problems are close to linear; 0 keeps
    newton_steps / relaxation).
    compat_vprolong is off: the mixed form uses the v-facet prolongation."""
    max_outer: int = 30
    inner_vcycles: int = 1
    relative_tolerance: float = 1e-10
    absolute_tolerance: float = 0.0
    fp64_coarsest: bool = True
    coarsest_steps: int = 5
    pre_steps: int = 4
    post_steps: int = 4
    finest_steps: int = 8
    newton_steps: int = 3
    relaxation: float = 0.5
    adjoint_omega: float = 0.85
    compat_vprolong: bool = False
    asynchronous: bool = False
    krylov_dim: int = 12
    # the forward solve's FP64 finishing cycles: one full local Newton step
    # per cell (0 keeps newton_steps / relaxation)
    fp64_newton_steps: int = 1
    fp64_relaxation: float = 1.0

    def to_c(self) -> _lib.MixedConfigC:
        if not 0 <= int(self.max_outer) <= 64:
            raise ValueError(f"max_outer must be in [0, 64] (got {self.max_outer})")
        if not 0 <= int(self.krylov_dim) <= 16:
            raise ValueError(f"krylov_dim must be in [0, 16] (got {self.krylov_dim})")
        return _lib.MixedConfigC(max_outer=int(self.max_outer), inner_vcycles=int(self.inner_vcycles),
                                 relative_tolerance=float(self.relative_tolerance),
                                 absolute_tolerance=float(self.absolute_tolerance),
                                 fp64_coarsest=int(bool(self.fp64_coarsest)), krylov_dim=0, krylov_work=None,
                                 fp64_newton_steps=int(self.fp64_newton_steps),
                                 fp64_relaxation=float(self.fp64_relaxation))

    def fas_c(self, freeze_coarse_calving=True) -> _lib.FasConfigC:
        return _lib.FasConfigC(freeze_coarse_calving=int(bool(freeze_coarse_calving)), freeze_coarse_phi=1,
                               coarsest_steps=int(self.coarsest_steps), pre_steps=int(self.pre_steps),
                               post_steps=int(self.post_steps), finest_steps=int(self.finest_steps),
                               maximum_vcycles=0, relative_tolerance=0.0, absolute_tolerance=0.0,
                               compat_vprolong=int(bool(self.compat_vprolong)))

    def vanka_c(self, adjoint: bool) -> _lib.VankaConfigC:
        # omega: the update weight of the adjoint passes (the forward
        # red-black passes write at full weight)
        return _lib.VankaConfigC(omega=float(self.adjoint_omega), relax_phi=0.0,
                                 newton_steps=int(self.newton_steps), relaxation=float(self.relaxation),
                                 ssa_damping=0.0, mc_damping=0.0, local_fp32=0, multicolour=1, newton_tol=0.0)


class MixedHierarchy:
    """The FP32 mirror of a multigrid hierarchy for the mixed solves."""

    def __init__(self, levels):
        self.levels = levels  # the FP64 Grid objects, finest first
        self.l32 = [None] * len(levels)
        self._krylov = {}

    def krylov_work(self, start_level: int, m: int) -> torch.Tensor:
        """The flexible GMRES workspace of an adjoint solve from start_level."""
        key = (start_level, m)
        if key not in self._krylov:
            g = self.levels[start_level]
            n = int(_lib.load().glide_krylov_work_doubles(g.ny, g.nx, m))
            if n <= 0:
                raise ValueError(f"invalid Krylov workspace request ({g.ny} x {g.nx}, m = {m})")
            self._krylov[key] = torch.zeros(n, dtype=torch.float64, device=g.device)
        return self._krylov[key]

    def structs(self, start_level: int, adjoint: bool):
        arr = (_lib.Level * len(self.levels))()
        for k, g in enumerate(self.levels):
            if k < start_level:
                continue
            if self.l32[k] is None:
                self.l32[k] = Level32(g, adjoint)
            elif adjoint:
                self.l32[k].add_adjoint()
            arr[k] = self.l32[k].struct()
        return arr


def _run(name, levels64, mh, start_level, dt, cfg, adjoint, stream):
    n = len(levels64)
    l32 = mh.structs(start_level, adjoint)
    vcs = (_lib.VankaConfigC * n)()
    for k in range(n):
        vcs[k] = cfg.vanka_c(adjoint)
    fixed = not (cfg.relative_tolerance > 0.0) and not (cfg.absolute_tolerance > 0.0)
    rep = None if (cfg.asynchronous and fixed) else _lib.SolveReport()
    stats = (ctypes.c_int * 3)()
    fas = cfg.fas_c()
    mc = cfg.to_c()
    if int(cfg.krylov_dim) > 0 and (rep is not None or not fixed):
        mc.krylov_dim = int(cfg.krylov_dim)
        mc.krylov_work = ptr(mh.krylov_work(start_level, int(cfg.krylov_dim)))
    _lib.call(name, levels64, l32, n, int(start_level), float(dt), ctypes.byref(fas), vcs, ctypes.byref(mc),
              ctypes.byref(rep) if rep is not None else None, stats, stream)
    return rep, {"fallbacks": stats[0], "fp32": stats[1], "fp64": stats[2]}


def forward_solve(levels64, mh: MixedHierarchy, start_level: int, dt: float, cfg: MixedConfig, stream):
    """glide_fascd_solve_mixed on the FP64 level descriptors `levels64`."""
    return _run("glide_fascd_solve_mixed", levels64, mh, start_level, dt, cfg, False, stream)


def adjoint_solve(levels64, mh: MixedHierarchy, start_level: int, dt: float, cfg: MixedConfig, stream):
    """glide_fas_adjoint_solve_mixed on the FP64 level descriptors `levels64`."""
    return _run("glide_fas_adjoint_solve_mixed", levels64, mh, start_level, dt, cfg, True, stream)


__all__ = ["MixedConfig", "MixedHierarchy", "Level32", "forward_solve", "adjoint_solve", "ptr"]
