"""Per-level operator objects of the drop-in API (glide/operators.py).

:class:`ForwardOperators` owns the residual / forcing / coarse-operator
buffers of one level and calls the gfx950 residual and Vanka kernels;
:class:`AdjointOperators` does the same for the transpose operator, the
adjoint Vanka smoother and the parameter gradients.  Every call enqueues on
torch's current stream; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Callable

import torch

from . import _lib
from ._lib import FREEZE_CALVING, FREEZE_PHI, USE_FORCING, USE_MASK, call, ptr, stream_handle


@dataclass
class NewtonConfig:
    """Local Newton settings of the forward smoother (glide/operators.py:314-319).

    The forward kernel uses a fixed local relaxation of 0.5, as the reference
    kernel does; `relaxation` is kept for API parity."""
    steps: int = 30
    relaxation: float = 0.5
    ssa_damping: float = 0.01
    mc_damping: float = 1.0


@dataclass
class VankaConfig:
    """glide/operators.py:321-327.  ``local_fp32`` (not in the reference)
    selects the mixed form of the forward smoother: each cell's patch Jacobian
    is evaluated and factorised in FP32 while the residual, the local stopping
    test and the Newton update stay FP64."""
    omega: float = 0.5
    newton_config: NewtonConfig = field(default_factory=NewtonConfig)
    relax_phi: float = 0.0
    hook_interval: int = 1
    hook_func: Callable[[int], None] | None = None
    local_fp32: bool = False
    # (an addition) red-black passes instead of the additive one; their local
    # Newton iteration uses newton_config.relaxation and stops once
    # sum r^2 <= newton_tol
    multicolour: bool = False
    newton_tol: float = 1e-6

    def to_c(self) -> _lib.VankaConfigC:
        n = self.newton_config
        return _lib.VankaConfigC(omega=float(self.omega), relax_phi=float(self.relax_phi),
                                 newton_steps=int(n.steps), relaxation=float(n.relaxation),
                                 ssa_damping=float(n.ssa_damping), mc_damping=float(n.mc_damping),
                                 local_fp32=1 if self.local_fp32 else 0, multicolour=1 if self.multicolour else 0,
                                 newton_tol=float(self.newton_tol))


def _norms(a, b, c, stream):
    out = (ctypes.c_double * 3)()
    call("glide_norms3", ptr(a), a.numel(), ptr(b), b.numel(), ptr(c), c.numel(), out, stream)
    return out[0], out[1], out[2]


class ForwardOperators:
    """Residual, grounded fraction and Vanka smoothing of one level."""

    def __init__(self, grid):
        self.grid = grid
        z = grid.zeros
        ny, nx = grid.ny, grid.nx
        self.r_u, self.r_v, self.r_H = z((ny, nx + 1)), z((ny + 1, nx)), z((ny, nx))
        self.f_u, self.f_v, self.f_H = z((ny, nx + 1)), z((ny + 1, nx)), z((ny, nx))
        self.F_u, self.F_v, self.F_H = z((ny, nx + 1)), z((ny + 1, nx)), z((ny, nx))
        self.gamma = z((ny, nx)).fill_(grid.geometry.thklim.value)
        self.vanka_config = VankaConfig()
        self._var = None  # direction and product of compute_jvp, allocated on first use
        self._jvp = None

    def _level(self):
        return self.grid.level_struct()

    def _pair(self, which):
        g = self.grid
        if getattr(self, which) is None:
            setattr(self, which, (g.zeros((g.ny, g.nx + 1)), g.zeros((g.ny + 1, g.nx)), g.zeros((g.ny, g.nx))))
        return getattr(self, which)

    @property
    def var_u(self):
        """Direction of compute_jvp, u-facets (glide/operators.py:278-282)."""
        return self._pair("_var")[0]

    @property
    def var_v(self):
        return self._pair("_var")[1]

    @property
    def var_H(self):
        return self._pair("_var")[2]

    @property
    def jvp_u(self):
        """J var of the last compute_jvp, u-facets (glide/operators.py:296-312)."""
        return self._pair("_jvp")[0]

    @property
    def jvp_v(self):
        return self._pair("_jvp")[1]

    @property
    def jvp_H(self):
        return self._pair("_jvp")[2]

    def compute_jvp(self, dt, use_mask=True, freeze_calving=False, freeze_phi=False, return_norms=False):
        """jvp = J(x) var, the residual linearised at the current state along
        (var_u, var_v, var_H) with phi, mask, bed, B and beta frozen
        (glide/operators.py:129-171, glide/cuda/residuals.cu:340-660).  Rows of
        Dirichlet facets are zero and the thickness row is (1 - mask) times
        the mass-balance derivative, as in the reference."""
        st = stream_handle(self.grid.device)
        if not freeze_phi:
            self.compute_phi()
        L = self._level()
        flags = (USE_MASK if use_mask else 0) | (FREEZE_CALVING if freeze_calving else 0)
        d = self._pair("_var")
        out = self._pair("_jvp")
        call("glide_ssa_jvp", ctypes.byref(L), float(dt), flags, ptr(d[0]), ptr(d[1]), ptr(d[2]), ptr(out[0]),
             ptr(out[1]), ptr(out[2]), st)
        if return_norms:
            return _norms(out[0], out[1], out[2], st)
        return None

    def compute_residual(self, dt, use_mask=True, operator_only=False, freeze_calving=False, freeze_phi=False,
                         return_norms=False):
        """glide/operators.py:71-127."""
        st = stream_handle(self.grid.device)
        if not freeze_phi:
            self.compute_phi()
        L = self._level()
        flags = (USE_MASK if use_mask else 0) | (FREEZE_CALVING if freeze_calving else 0)
        if operator_only:
            out = (self.F_u, self.F_v, self.F_H)
        else:
            out = (self.r_u, self.r_v, self.r_H)
            flags |= USE_FORCING
        call("glide_ssa_residual", ctypes.byref(L), float(dt), flags, ptr(out[0]), ptr(out[1]), ptr(out[2]), st)
        if return_norms:
            return _norms(out[0], out[1], out[2], st)
        return None

    def compute_phi(self, relaxation=0.0):
        """Grounded fraction from H and bed (glide/operators.py:173-192)."""
        L = self._level()
        call("glide_compute_grounded", ctypes.byref(L), float(relaxation), stream_handle(self.grid.device))

    def vanka_smooth(self, dt, freeze_calving=False, freeze_phi=False):
        """One additive Vanka pass including the damped update."""
        self.vanka_sweep(dt, 1, freeze_calving=freeze_calving, freeze_phi=freeze_phi)

    def vanka_sweep(self, dt, n_iter, freeze_calving=False, freeze_phi=False):
        """glide/operators.py:243-251: n_iter passes; the hook (if any) runs after each."""
        L = self._level()
        vc = self.vanka_config.to_c()
        flags = (FREEZE_CALVING if freeze_calving else 0) | (FREEZE_PHI if freeze_phi else 0)
        st = stream_handle(self.grid.device)
        hook = self.vanka_config.hook_func
        if hook is None:
            call("glide_vanka_sweeps", ctypes.byref(L), float(dt), int(n_iter), ctypes.byref(vc), flags, st)
            return
        for i in range(int(n_iter)):
            call("glide_vanka_sweeps", ctypes.byref(L), float(dt), 1, ctypes.byref(vc), flags, st)
            hook(i)

    def vanka_dump(self, dt):
        """Local 5x5 systems of every cell: J (ny*nx, 25), r (ny*nx, 5)."""
        g = self.grid
        J = g.zeros((g.ny * g.nx, 25))
        r = g.zeros((g.ny * g.nx, 5))
        L = self._level()
        call("glide_vanka_dump", ctypes.byref(L), float(dt), ptr(J), ptr(r), stream_handle(g.device))
        return J, r

    def set_rhs(self, dt):
        """f_H = H_prev/dt + smb (glide/operators.py:275-276)."""
        g = self.grid
        torch.add(g.state.H_prev.data / dt, g.forcing.smb.data, out=self.f_H)


class AdjointOperators:
    """Transpose operator, adjoint Vanka smoothing and gradients of one level."""

    def __init__(self, grid):
        self.grid = grid
        z = grid.zeros
        ny, nx = grid.ny, grid.nx
        self.r_u, self.r_v, self.r_H = z((ny, nx + 1)), z((ny + 1, nx)), z((ny, nx))
        self.f_u, self.f_v, self.f_H = z((ny, nx + 1)), z((ny + 1, nx)), z((ny, nx))
        self.vjp_u, self.vjp_v, self.vjp_H = z((ny, nx + 1)), z((ny + 1, nx)), z((ny, nx))
        self.gamma = z((ny, nx)).fill_(grid.geometry.thklim.value)
        self.vanka_config = VankaConfig()
        # Adjoint smoother cache (transposed patch LU factors and cell
        # viscosity, refreshed by every adjoint solve or sweep call for the
        # state at that moment); use_cache = False recomputes them per pass.
        self.use_cache = True
        self.cache = z((26, ny, nx))
        grid.adjoint  # the multipliers exist whenever the operators do

    def _level(self):
        return self.grid.level_struct()

    def _lam(self):
        a = self.grid.adjoint
        return ptr(a.lambda_u.data), ptr(a.lambda_v.data), ptr(a.lambda_H.data)

    def compute_residual(self, dt, use_mask=True, freeze_calving=False, return_norms=False):
        """r = Jᵀλ - f (glide/operators.py:379-425); masked cells use the state mask."""
        L = self._level()
        st = stream_handle(self.grid.device)
        flags = USE_FORCING | (FREEZE_CALVING if freeze_calving else 0)
        call("glide_ssa_vjp", ctypes.byref(L), float(dt), flags, *self._lam(), ptr(self.r_u), ptr(self.r_v),
             ptr(self.r_H), st)
        if return_norms:
            return _norms(self.r_u, self.r_v, self.r_H, st)
        return None

    def compute_vjp(self, dt, use_mask=True, use_forcing=False, freeze_calving=False):
        """vjp = Jᵀλ (glide/operators.py:429-472)."""
        L = self._level()
        flags = FREEZE_CALVING if freeze_calving else 0
        call("glide_ssa_vjp", ctypes.byref(L), float(dt), flags, *self._lam(), ptr(self.vjp_u), ptr(self.vjp_v),
             ptr(self.vjp_H), stream_handle(self.grid.device))

    def vanka_smooth(self, dt, freeze_calving=False):
        self.vanka_sweep(dt, 1, freeze_calving=freeze_calving)

    def vanka_sweep(self, dt, n_iter, freeze_calving=False):
        """glide/operators.py:517-524: residual, transposed local solves, λ -= ω δλ."""
        L = self._level()
        vc = self.vanka_config.to_c()
        call("glide_vanka_adjoint_sweeps", ctypes.byref(L), float(dt), int(n_iter), ctypes.byref(vc),
             FREEZE_CALVING if freeze_calving else 0, stream_handle(self.grid.device))

    def _grad(self, which, dt=1.0):
        g = self.grid
        if which & 1:
            g.sliding.beta.grad
        if which & 2:
            g.geometry.bed.grad
        if which & 4:
            g.state.H_prev.grad
        if which & 8:
            g.forcing.smb.grad
        L = self._level()
        call("glide_gradients", ctypes.byref(L), float(dt), int(which), stream_handle(g.device))

    def compute_gradient_B(self, dt=1.0):
        """rheology.B.grad = lambda^T dR/dB (no counterpart in the reference,
        whose gradients cover beta, bed, H_prev and smb)."""
        g = self.grid
        L = self._level()
        call("glide_gradient_rate_factor", ctypes.byref(L), float(dt), ptr(g.rheology.B.grad), stream_handle(g.device))

    def compute_gradient_beta(self):
        self._grad(1)

    def compute_gradient_bed(self):
        self._grad(2)

    def compute_gradient_H_prev(self, dt):
        self._grad(4, dt)

    def compute_gradient_smb(self):
        self._grad(8)
