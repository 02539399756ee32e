"""Enthalpy column solver of the coupled step (glide/enthalpy.py).

Conservative sigma-coordinate enthalpy equation (time term on H E, vertical
diffusion with a cold/temperate diffusivity, upwind sigma advection,
Lax-Friedrichs horizontal advection with the momentum solver's mass flux,
temperate-ice drainage), solved by column Newton sweeps with a tridiagonal
inner solve alternated with pointwise Jacobi layer sweeps.  The fields are
float64 torch tensors on the ROCm device; every operator is a gfx950 kernel
behind ``include/glide_enthalpy.h``.  Enthalpy is stored scaled by
``E_SCALE``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Callable

import torch

from . import _lib
from ._lib import call, ptr, stream_handle
from .field import Constant

# glide/enthalpy.py:46-65
RHO_I = 910.0
RHO_W = 1000.0
C_I = 2009.0
K_I = 2.1
L_HEAT = 3.34e5
T_REF = 223.15
T_MELT = 273.15
BETA_CC = 7.9e-8
GRAVITY = 9.81
K_COLD = K_I / C_I
K_TEMP_FACTOR = 1e-1
K_TEMP = K_COLD * K_TEMP_FACTOR
E_SCALE = C_I * (T_MELT - T_REF)
SEC_PER_YR = 365.25 * 86400.0
MASS_FLUX_REG = 10.0 / (SEC_PER_YR ** 2)
MASS_FLUX_REG_YR = 10.0


def water_content_from_enthalpy(E, depth=0.0):
    """Water content of (unscaled) enthalpy E at a depth (glide/enthalpy.py:68-72)."""
    T_pmp = T_MELT - BETA_CC * RHO_I * GRAVITY * depth
    E_pmp = C_I * (T_pmp - T_REF)
    return torch.clamp(E - E_pmp, min=0.0) / L_HEAT


def make_sigma(nz, device=None):
    """Uniform sigma nodes from the bed (0) to the surface (1) (glide/enthalpy.py:75-90)."""
    return torch.linspace(0.0, 1.0, nz, dtype=torch.float64, device=device)


@dataclass
class EnthalpyState:
    """glide/enthalpy.py:96-105."""
    E: torch.Tensor | None = None
    E_prev: torch.Tensor | None = None


@dataclass
class EnthalpyForcing:
    """Boundary data and forcing (glide/enthalpy.py:108-144)."""
    E_surface: torch.Tensor | None = None
    Q_geo: torch.Tensor | None = None
    Q_fh: torch.Tensor | None = None
    phi_strain: torch.Tensor | None = None
    drain_rate: Constant = field(default_factory=lambda: Constant(
        value=0.01 / (365.25 * 86400.0), name="drain_rate", units="s^{-1}",
        attrs={"long_name": "Water drainage rate in temperate ice (0.01 a^{-1} default)"}))
    h_thin: Constant = field(default_factory=lambda: Constant(
        value=25.0, name="h_thin", units="m",
        attrs={"long_name": "Thickness below which the column is clamped to the surface enthalpy (Dirichlet)"}))


@dataclass
class EnthalpyVelocity:
    """glide/enthalpy.py:147-157."""
    u3d: torch.Tensor | None = None
    v3d: torch.Tensor | None = None
    omega: torch.Tensor | None = None


@dataclass
class EnthalpyTermFlags:
    """glide/enthalpy.py:160-179."""
    horizontal_advection: bool = True
    omega: bool = True
    drainage: bool = True

    @property
    def bitmask(self) -> int:
        return ((1 if self.horizontal_advection else 0) | ((1 if self.omega else 0) << 1)
                | ((1 if self.drainage else 0) << 2))


@dataclass
class ColumnSmootherConfig:
    """glide/enthalpy.py:182-198.  ``hook_func=None`` (the default) lets the
    whole column sweep run as one native call with the stopping test on the
    device; a hook makes the sweep a host loop that calls it each iteration."""
    omega: float = 1.0
    n_newton: int = 3
    relaxation: float = 1.0
    relative_tolerance: float = 1e-5
    absolute_tolerance: float = 1e-2
    lf_c: float = 1e-12
    report_norms: bool = False
    hook_func: Callable[[int], None] | None = None
    # (additions) red-black column ordering of the native sweep (block
    # Gauss-Seidel over columns: fewer sweeps; near the pressure-melting point
    # the discrete system can have more than one solution, and this order can
    # settle on a different one than the reference's), and whether the
    # alternating sweep keeps its pointwise layer pass
    multicolour: bool = False
    layer_sweeps: bool = True

    def to_c(self, alternating=True) -> _lib.EnthSweepConfigC:
        return _lib.EnthSweepConfigC(omega=float(self.omega), relaxation=float(self.relaxation),
                                     relative_tolerance=float(self.relative_tolerance),
                                     absolute_tolerance=float(self.absolute_tolerance),
                                     n_newton=int(self.n_newton),
                                     alternating=1 if (alternating and self.layer_sweeps) else 0,
                                     multicolour=1 if self.multicolour else 0)


def _as_field(value, shape, device) -> torch.Tensor:
    """Scalar / broadcastable array -> contiguous float64 tensor of `shape`."""
    t = torch.as_tensor(value, dtype=torch.float64, device=device)
    return t.expand(shape).contiguous() if tuple(t.shape) != tuple(shape) else t.contiguous()


class EnthalpyOperators:
    """Operators of the enthalpy equation on one grid (glide/enthalpy.py:205-643)."""

    def __init__(self, grid, nz=11, use_fast_math=True):
        if not 2 <= nz <= 64:
            raise ValueError("nz must be in [2, 64]")
        self.grid = grid
        self.nz = nz
        self.sigma = make_sigma(nz, grid.device)
        ny, nx = grid.ny, grid.nx
        z3 = lambda: grid.zeros((ny, nx, nz))  # noqa: E731
        z2 = lambda: grid.zeros((ny, nx))  # noqa: E731
        self.enthalpy_state = EnthalpyState(E=z3(), E_prev=z3())
        self.enthalpy_velocity = EnthalpyVelocity(u3d=grid.zeros((nz, ny, nx + 1)),
                                                  v3d=grid.zeros((nz, ny + 1, nx)), omega=z3())
        self.enthalpy_forcing = EnthalpyForcing(E_surface=z2(), Q_geo=z2(), Q_fh=z2(), phi_strain=z3())
        self.H_prev = z2()
        self.r_E = z3()
        self.f_E = z3()
        self.delta_E = z3()
        self._E_work = z3()
        # level-major copies of E, E_prev, f_E and omega for the fused sweep of the thermal step
        self._sweep_work = grid.zeros((4, nz, ny, nx))
        self._dh_dt = z2()
        self.smoother_config = ColumnSmootherConfig()
        self.term_flags = EnthalpyTermFlags()
        self.last_report = None

    # ------------------------------------------------------------ plumbing
    def _column(self) -> _lib.EnthColumn:
        g = self.grid
        st, vel, frc = self.enthalpy_state, self.enthalpy_velocity, self.enthalpy_forcing
        C = _lib.EnthColumn(ny=g.ny, nx=g.nx, nz=self.nz, term_flags=int(self.term_flags.bitmask), dx=float(g.dx),
                            drain_rate=float(frc.drain_rate.value), h_thin=float(frc.h_thin.value),
                            lf_c=float(self.smoother_config.lf_c))
        bufs = dict(E=st.E, E_prev=st.E_prev, f_E=self.f_E, omega=vel.omega, phi_strain=frc.phi_strain,
                    r_E=self.r_E, delta_E=self.delta_E, E_work=self._E_work, u3d=vel.u3d, v3d=vel.v3d,
                    H=g.state.H.data, H_prev=self.H_prev, E_surface=frc.E_surface, Q_geo=frc.Q_geo, Q_fh=frc.Q_fh,
                    dh_dt=self._dh_dt, sweep_work=self._sweep_work)
        for name, t in bufs.items():
            setattr(C, name, ptr(t))
        return C

    def _stream(self):
        return stream_handle(self.grid.device)

    # ------------------------------------------------------------ operators
    def broadcast_velocity(self):
        """u3d[k], v3d[k] = depth-averaged SSA velocity (glide/enthalpy.py:289-301)."""
        C = self._column()
        call("glide_enth_broadcast_velocity", ctypes.byref(C), ptr(self.grid.state.u.data),
             ptr(self.grid.state.v.data), 1.0, self._stream())

    def compute_omega(self, dh_dt, bmb=None):
        """omega = H sigma_dot by upward integration (glide/enthalpy.py:303-339)."""
        shape = (self.grid.ny, self.grid.nx)
        dh = _as_field(dh_dt, shape, self.grid.device)
        b = None if bmb is None else _as_field(bmb, shape, self.grid.device)
        C = self._column()
        call("glide_enth_omega", ctypes.byref(C), ptr(dh), ptr(b), 1.0, self._stream())

    def compute_residual(self, dt):
        """r_E and max |r_E| (glide/enthalpy.py:341-364)."""
        C = self._column()
        out = ctypes.c_double(0.0)
        call("glide_enth_residual", ctypes.byref(C), float(dt), 1, ctypes.byref(out), self._stream())
        return out.value

    def column_smooth(self, dt):
        """delta_E from one column Newton solve (glide/enthalpy.py:366-393)."""
        C = self._column()
        cfg = self.smoother_config
        call("glide_enth_column_smooth", ctypes.byref(C), float(dt), int(cfg.n_newton), float(cfg.relaxation),
             self._stream())

    def layer_smooth(self, dt):
        """delta_E = -r / J_kk (glide/enthalpy.py:395-424)."""
        C = self._column()
        call("glide_enth_layer_smooth", ctypes.byref(C), float(dt), self._stream())

    def _apply(self, omega):
        C = self._column()
        call("glide_enth_apply", ctypes.byref(C), float(omega), self._stream())

    def column_sweep(self, dt, n_iter, alternating=True):
        """Alternating column / layer sweeps until the residual test passes
        (glide/enthalpy.py:426-471)."""
        cfg = self.smoother_config
        if cfg.hook_func is None:
            C = self._column()
            rep = _lib.EnthReport() if cfg.report_norms else None
            if rep is not None:
                self.last_report = rep  # kept when the call raises (a non-finite norm)
            call("glide_enth_column_sweep", ctypes.byref(C), float(dt), int(n_iter), ctypes.byref(cfg.to_c(alternating)),
                 ctypes.byref(rep) if rep is not None else None, self._stream())
            if rep is not None:
                print(f"  Enthalpy initial: |r0| = {rep.r0:.2e}")
                for it in range(rep.sweeps):
                    r = rep.norms[it]
                    rel = r / rep.r0 if rep.r0 > 0 else 0.0
                    print(f"  Enthalpy sweep {it}: |r|/|r0| = {rel:.2e}, |r| = {r:.2e}")
            return
        r0 = self.compute_residual(dt)
        if cfg.report_norms:
            print(f"  Enthalpy initial: |r0| = {r0:.2e}")
        for iteration in range(n_iter):
            self.column_smooth(dt)
            self._apply(cfg.omega)
            if alternating:
                self.layer_smooth(dt)
                self._apply(cfg.omega)
            cfg.hook_func(iteration)
            r = self.compute_residual(dt)
            rel = r / r0 if r0 > 0 else 0.0
            if cfg.report_norms:
                print(f"  Enthalpy sweep {iteration}: |r|/|r0| = {rel:.2e}, |r| = {r:.2e}")
            if rel < cfg.relative_tolerance or r < cfg.absolute_tolerance:
                break

    def set_rhs(self, snapshot=True):
        """f_E from strain heating and the basal heat fluxes (glide/enthalpy.py:473-519)."""
        if snapshot:
            self.enthalpy_state.E_prev.copy_(self.enthalpy_state.E)
            self.H_prev.copy_(self.grid.state.H.data)
        C = self._column()
        call("glide_enth_set_rhs", ctypes.byref(C), self._stream())

    def compute_frictional_heating(self, m, u_reg):
        """Q_fh = beta |u_b|^(m+1) from the bed-level velocities (glide/model.py:215-228)."""
        C = self._column()
        call("glide_enth_frictional_heating", ctypes.byref(C), ptr(self.grid.sliding.beta.data), float(m),
             float(u_reg), self._stream())

    def compute_strain_heating(self, rho_i=917.0):
        """phi_strain = 4 eta eps_e^2 rho_i g / a on every level from the grid's
        depth-averaged velocity and rate factor B (head units), eta and eps_e^2
        being the momentum solver's cell viscosity and strain-rate invariant.
        The reference leaves phi_strain to the caller; this supplies the SSA
        dissipation for the coupled step."""
        g = self.grid
        C = self._column()
        call("glide_enth_strain_heating", ctypes.byref(C), ptr(g.state.u.data), ptr(g.state.v.data),
             ptr(g.rheology.B.data), float(g.rheology.n.value), float(g.rheology.eps_reg.value), float(rho_i),
             self._stream())

    def set_surface_enthalpy_from_temperature(self, T_surface):
        """glide/enthalpy.py:521-532."""
        Ts = _as_field(T_surface, (self.grid.ny, self.grid.nx), self.grid.device)
        C = self._column()
        call("glide_enth_surface_from_temperature", ctypes.byref(C), ptr(Ts), ptr(self.enthalpy_forcing.E_surface),
             self._stream())

    def initialize_from_temperature(self, T_field):
        """E from a scalar, (ny, nx) or (ny, nx, nz) temperature capped at T_pmp
        (glide/enthalpy.py:534-563)."""
        ny, nx, nz = self.grid.ny, self.grid.nx, self.nz
        T = torch.as_tensor(T_field, dtype=torch.float64, device=self.grid.device)
        if T.ndim == 2:
            T = T[:, :, None]
        T = T.expand(ny, nx, nz).contiguous()
        C = self._column()
        call("glide_enth_from_temperature", ctypes.byref(C), ptr(T), ptr(self.enthalpy_state.E), self._stream())
        self.enthalpy_state.E_prev.copy_(self.enthalpy_state.E)
        self.H_prev.copy_(self.grid.state.H.data)

    def get_temperature(self):
        """glide/enthalpy.py:565-581."""
        T = torch.empty_like(self.enthalpy_state.E)
        C = self._column()
        call("glide_enth_diagnostics", ctypes.byref(C), ptr(T), None, self._stream())
        return T

    def get_water_content(self):
        """glide/enthalpy.py:583-598."""
        w = torch.empty_like(self.enthalpy_state.E)
        C = self._column()
        call("glide_enth_diagnostics", ctypes.byref(C), None, ptr(w), self._stream())
        return w

    def adjoint_solve(self, dt, dL_dE, dL_dB=None, B_scale=1.0, n_iter=200, relative_tolerance=1e-12,
                      absolute_tolerance=0.0, omega=1.0, report=False, u2d=None, v2d=None, vel_div=1.0):
        """Adjoint of the converged solve at the current E (no counterpart in
        the reference): solves J^T lambda = dL/dE + (dB/dE)^T dL/dB with J the
        exact Jacobian of the system the sweeps solve, B the rate factor
        divided by ``B_scale``, and returns ``{"lambda", "E_prev",
        "phi_strain", "Q_geo", "Q_fh", "E_surface"}`` gradients (``Q_fh`` and
        ``Q_geo`` share one).  Exact for a converged forward solve.  u2d, v2d
        (optional): the depth-constant velocity whose quotient by vel_div the
        velocity fields hold on every level (the coupled step's broadcast SSA
        velocity); the sweeps then compute the horizontal couplings from it."""
        g = self.grid
        ny, nx, nz = g.ny, g.nx, self.nz
        if not hasattr(self, "_adj"):
            plane = 2 * nz * ny * ((nx + 1) // 2)  # GLIDE_ENTH_ADJ_PLANE
            self._adj = {"rows": g.zeros((7, plane)), "g": g.zeros((plane,)), "lambda_work": g.zeros((plane,))}
        W = self._adj
        lam = g.zeros((ny, nx, nz))
        out = {"E_prev": g.zeros((ny, nx, nz)), "phi_strain": g.zeros((ny, nx, nz)), "Q_geo": g.zeros((ny, nx)),
               "E_surface": g.zeros((ny, nx))}
        X = _lib.EnthAdjointC(rows=ptr(W["rows"]), g=ptr(W["g"]), lambda_=ptr(lam), lambda_work=ptr(W["lambda_work"]),
                              r=None, g_E_prev=ptr(out["E_prev"]), g_phi_strain=ptr(out["phi_strain"]),
                              g_Q_geo=ptr(out["Q_geo"]), g_E_surface=ptr(out["E_surface"]),
                              u2d=ptr(u2d) if u2d is not None else None,
                              v2d=ptr(v2d) if v2d is not None else None, vel_div=float(vel_div))
        gE = _as_field(dL_dE, (ny, nx, nz), g.device)
        gB = None if dL_dB is None else _as_field(dL_dB, (ny, nx), g.device)
        cfg = ColumnSmootherConfig(omega=omega, relative_tolerance=relative_tolerance,
                                   absolute_tolerance=absolute_tolerance).to_c(True)
        rep = _lib.EnthReport() if report else None
        C = self._column()
        call("glide_enth_adjoint_solve", ctypes.byref(C), ctypes.byref(X), float(dt), ptr(gE), ptr(gB),
             float(g.rheology.n.value), float(B_scale), int(n_iter), ctypes.byref(cfg),
             ctypes.byref(rep) if rep is not None else None, self._stream())
        out["lambda"] = lam
        out["Q_fh"] = out["Q_geo"]
        if rep is not None:
            out["report"] = rep
        return out

    def get_arrhenius_factor(self):
        """Depth-averaged B = A^(-1/n) (SI) from the Paterson-Budd law with the
        water-content enhancement (glide/enthalpy.py:600-643)."""
        B = self.grid.zeros((self.grid.ny, self.grid.nx))
        C = self._column()
        call("glide_enth_arrhenius", ctypes.byref(C), float(self.grid.rheology.n.value), 1.0, ptr(B), self._stream())
        return B
