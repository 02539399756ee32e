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
                 device=None, precision="fp64", dtype=None):
        if dtype is not None:
            precision = self.precision_from_dtype(dtype)
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
        # (an addition) "fp64": the reference's solves in FP64; "mixed": the
        # FP32 solves with FP64 defect correction (glide_amd.mixed)
        self.precision = precision

    @staticmethod
    def precision_from_dtype(dtype) -> str:
        """The solve form of a `dtype` argument (an addition: the reference
        solves in FP32 only): torch.float64 / "float64" -> the FP64 form,
        "mixed" -> FP32 multigrid with FP64 defect correction.  A plain FP32
        form is not built (its accuracy is that of the mixed form's inner
        iteration, ~1e-7), so float32 is refused."""
        name = str(dtype).replace("torch.", "")
        if name in ("float64", "fp64", "double"):
            return "fp64"
        if name == "mixed":
            return "mixed"
        if name in ("float32", "fp32", "float"):
            raise ValueError("dtype float32: a plain FP32 form is not built; use dtype='mixed' "
                             "(FP32 multigrid with FP64 defect correction) or torch.float64")
        raise ValueError(f"dtype must be torch.float64, 'float64' or 'mixed' (got {dtype!r})")

    @property
    def forward_solver(self) -> FASCDSolver:
        if self._forward_solver is None:
            self._forward_solver = FASCDSolver(self.mg)
            self._forward_solver.precision = self.precision
        return self._forward_solver

    @property
    def adjoint_solver(self) -> FASAdjointSolver:
        if self._adjoint_solver is None:
            self._adjoint_solver = FASAdjointSolver(self.mg)
            self._adjoint_solver.precision = self.precision
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
                 compute_H_prev_grad=True, compute_smb_grad=True, exact_gradients=False, compute_B_grad=False):
        """Adjoint solve Jᵀλ = -dJ/dx and parameter gradients (glide/model.py:74-94).

        exact_gradients (an addition): the gradient of the discrete problem,
        with the boundary-facet terms the reference's gradient kernel drops and
        zero H_prev / smb gradients in active-set cells (GLIDE_GRAD_EXACT).
        compute_B_grad (an addition): the rate factor's gradient into
        rheology.B.grad."""
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
        if which and exact_gradients:
            which |= 16
        if which:
            ao._grad(which, dt)
        if compute_B_grad:
            ao.compute_gradient_B(dt)


class ThermalModel:
    """Enthalpy solver coupled to the momentum step (glide/model.py:97-280).

    ``pre_momentum()`` before ``IceDynamics.forward``, ``step(dt)`` after it:
    the step broadcasts the SSA velocity to the sigma levels, derives omega
    from the realised thickness change, adds basal frictional heating, runs
    the column sweep and feeds the depth-averaged rate factor back into
    ``grid.rheology.B`` (head units).  Without a smoother hook the whole step
    is one native call."""

    SEC_PER_YR = 365.25 * 86400.0

    def __init__(self, grid, nz=21, n_smooth=10, update_rheology=True, frictional_heating=True, rho_i=917.0,
                 strain_heating=False):
        from .enthalpy import EnthalpyOperators
        self.ops = EnthalpyOperators(grid, nz=nz)
        self.n_smooth = n_smooth
        self.update_rheology = update_rheology
        self.frictional_heating = frictional_heating
        self.rho_i = rho_i
        # SSA strain heating from the momentum solution before each sweep
        # (an addition: the reference's phi_strain is caller-supplied)
        self.strain_heating = strain_heating
        self.g = 9.81
        self.last_report = None

    def initialize(self, T_surface, T_field=None, Q_geo=None):
        """Initial temperature (K), surface temperature (Dirichlet) and
        geothermal flux (W/m^2) (glide/model.py:134-154)."""
        if T_field is None:
            T_field = T_surface
        self.ops.initialize_from_temperature(T_field)
        self.ops.set_surface_enthalpy_from_temperature(T_surface)
        if Q_geo is not None:
            self.ops.enthalpy_forcing.Q_geo.copy_(torch.as_tensor(Q_geo, dtype=torch.float64).expand(
                self.ops.grid.ny, self.ops.grid.nx))

    def pre_momentum(self):
        """Snapshot E and H before the momentum step (glide/model.py:156-165)."""
        ops = self.ops
        ops.enthalpy_state.E_prev.copy_(ops.enthalpy_state.E)
        ops.H_prev.copy_(ops.grid.state.H.data)

    def step(self, dt, report=False, B_out=None):
        """Advance the enthalpy one step (glide/model.py:167-213).

        B_out (optional, (ny, nx) float64 tensor): receive the new rate factor
        there instead of overwriting ``grid.rheology.B``, which then stays the
        one the momentum solve used (an addition for callers that run the
        momentum adjoint concurrently and apply the feedback afterwards)."""
        import ctypes

        from . import _lib
        from ._lib import call, ptr
        ops = self.ops
        if ops.smoother_config.hook_func is not None:
            if B_out is not None:
                raise ValueError("B_out needs the fused native step (no smoother hook)")
            return self._step_host(dt)
        grid = ops.grid
        B = grid.rheology.B.data
        if B_out is not None:
            if tuple(B_out.shape) != tuple(B.shape) or B_out.dtype != B.dtype or B_out.device != B.device:
                raise ValueError("B_out must match grid.rheology.B in shape, dtype and device")
            B_out.copy_(B)  # strain heating reads the current rate factor
            B = B_out
        tc = self._config(dt)
        want = report or ops.smoother_config.report_norms
        rep = _lib.EnthReport() if want else None
        C = ops._column()
        call("glide_thermal_step", ctypes.byref(C), ptr(grid.state.u.data), ptr(grid.state.v.data),
             ptr(grid.sliding.beta.data),
             ptr(B) if (self.update_rheology or self.strain_heating) else None,
             ctypes.byref(tc), ctypes.byref(ops.smoother_config.to_c(True)),
             ctypes.byref(rep) if rep is not None else None, ops._stream())
        self.last_report = rep
        return rep

    def _config(self, dt, n_smooth=None, update_rheology=None):
        from . import _lib
        grid = self.ops.grid
        upd = self.update_rheology if update_rheology is None else update_rheology
        return _lib.ThermalConfigC(dt=float(dt), m=float(grid.sliding.m.value), u_reg=float(grid.sliding.u_reg.value),
                                   rho_i=float(self.rho_i), n_glen=float(grid.rheology.n.value),
                                   eps_reg=float(grid.rheology.eps_reg.value),
                                   n_smooth=int(self.n_smooth if n_smooth is None else n_smooth),
                                   frictional_heating=1 if self.frictional_heating else 0,
                                   update_rheology=1 if upd else 0, strain_heating=1 if self.strain_heating else 0)

    def rebuild_step(self, dt, E):
        """Recompute the fields ``step(dt)`` derives from its inputs (velocity
        planes, thickness tendency, omega, Q_fh, phi_strain, f_E) from the
        grid's current u, v, H, beta and B and the snapshot E_prev / H_prev,
        without sweeping and without touching B, and set the enthalpy to E (the
        converged state of that step): after other steps have overwritten the
        operators' buffers, ``adjoint`` then sees the step again exactly (no
        counterpart in the reference)."""
        import ctypes

        from . import _lib
        from ._lib import call, ptr
        ops = self.ops
        grid = ops.grid
        B = grid.rheology.B.data if self.strain_heating else None
        C = ops._column()
        call("glide_thermal_step", ctypes.byref(C), ptr(grid.state.u.data), ptr(grid.state.v.data),
             ptr(grid.sliding.beta.data), ptr(B), ctypes.byref(self._config(dt, n_smooth=0, update_rheology=False)),
             ctypes.byref(ops.smoother_config.to_c(True)), None, ops._stream())
        ops.enthalpy_state.E.copy_(E)

    def adjoint(self, dt, dL_dE, dL_dB=None, B_in=None, n_iter=200, relative_tolerance=1e-12,
                absolute_tolerance=0.0, report=False, flat=True):
        """Adjoint of the last ``step(dt)`` (no counterpart in the reference,
        whose enthalpy solve has no adjoint): solves the transposed system of
        the converged solve for dL/dE and, when ``update_rheology``, dL/dB of
        the new rate factor (head units), and returns the gradients
        ``{"lambda", "E_prev", "phi_strain", "Q_geo", "Q_fh", "E_surface", "u",
        "v", "H", "H_prev", "beta", "B"}``: u (ny, nx+1) and v (ny+1, nx) are
        the step's SSA velocities (m/a), H / H_prev the new / pre-momentum
        thickness, beta the sliding coefficient (through the frictional
        heating; None without it) and B the rate factor the strain heating
        read (None without strain heating).  Every input's path counts: the
        time, diffusion, drainage and forcing terms, the facet mass fluxes of
        the horizontal advection, omega, Q_fh and phi_strain, and the new rate
        factor's dependence on H.  The grid and the operators must hold the
        step's state (``step`` just ran, or ``rebuild_step``); ``B_in`` is the
        rate factor that step read, needed with strain heating when the step
        overwrote ``grid.rheology.B``.  flat: the adjoint sweeps read the
        depth-constant SSA velocity (the step's, equal results) instead of the
        per-level velocity planes."""
        import ctypes

        from . import _lib
        from ._lib import call, ptr
        ops = self.ops
        grid = ops.grid
        ny, nx = grid.ny, grid.nx
        if self.strain_heating and B_in is None:
            if self.update_rheology:
                raise ValueError("B_in (the rate factor the step read) is needed with strain heating when the step "
                                 "updated grid.rheology.B")
            B_in = grid.rheology.B.data
        gB = dL_dB if (dL_dB is not None and self.update_rheology) else None
        vel = dict(u2d=grid.state.u.data, v2d=grid.state.v.data, vel_div=self.SEC_PER_YR) if flat else {}
        out = ops.adjoint_solve(dt, dL_dE, gB, B_scale=self.B_scale, n_iter=n_iter,
                                relative_tolerance=relative_tolerance, absolute_tolerance=absolute_tolerance,
                                report=report, **vel)
        z = grid.zeros
        n = int(_lib.load().glide_thermal_vjp_work_doubles(ny, nx))
        if getattr(self, "_vjp_work", None) is None or self._vjp_work.numel() != n:
            self._vjp_work = z((n,))
        g = {"u": z((ny, nx + 1)), "v": z((ny + 1, nx)), "H": z((ny, nx)), "H_prev": z((ny, nx)),
             "beta": z((ny, nx)) if self.frictional_heating else None,
             "B": z((ny, nx)) if self.strain_heating else None}
        gBt = None if gB is None else torch.as_tensor(gB, dtype=torch.float64, device=grid.device).expand(
            ny, nx).contiguous()
        V = _lib.ThermalVjpC(lambda_=ptr(out["lambda"]), g_phi_strain=ptr(out["phi_strain"]), g_Q=ptr(out["Q_geo"]),
                             dL_dB=ptr(gBt), work=ptr(self._vjp_work), g_u=ptr(g["u"]), g_v=ptr(g["v"]),
                             g_H=ptr(g["H"]), g_H_prev=ptr(g["H_prev"]), g_beta=ptr(g["beta"]), g_B=ptr(g["B"]))
        C = ops._column()
        call("glide_thermal_step_vjp", ctypes.byref(C), ptr(grid.state.u.data), ptr(grid.state.v.data),
             ptr(grid.sliding.beta.data), ptr(B_in), ctypes.byref(self._config(dt)), ctypes.byref(V),
             ops._stream())
        out.update(g)
        return out

    def _step_host(self, dt):
        ops = self.ops
        ops.broadcast_velocity()
        self._compute_omega(dt)
        ops.enthalpy_velocity.u3d /= self.SEC_PER_YR
        ops.enthalpy_velocity.v3d /= self.SEC_PER_YR
        if self.frictional_heating:
            self._compute_frictional_heating()
        if self.strain_heating:
            ops.compute_strain_heating(self.rho_i)
        ops.set_rhs(snapshot=False)
        ops.column_sweep(dt, self.n_smooth)
        if self.update_rheology:
            ops.grid.rheology.B.data.copy_(ops.get_arrhenius_factor() / self.B_scale)
        return None

    def _compute_frictional_heating(self):
        """glide/model.py:215-228."""
        s = self.ops.grid.sliding
        self.ops.compute_frictional_heating(s.m.value, s.u_reg.value)

    def _compute_omega(self, dt):
        """omega from (H - H_prev) / dt in m/a, converted to m/s (glide/model.py:230-254)."""
        ops = self.ops
        dt_yr = dt / self.SEC_PER_YR
        ops.compute_omega((ops.grid.state.H.data - ops.H_prev) / dt_yr)
        ops.enthalpy_velocity.omega /= self.SEC_PER_YR

    @property
    def B_scale(self):
        """B_glide = B_SI / B_scale (glide/model.py:256-260)."""
        n = float(self.ops.grid.rheology.n.value)
        return self.rho_i * self.g * self.SEC_PER_YR ** (1.0 / n)

    @property
    def temperature(self):
        return self.ops.get_temperature()

    @property
    def water_content(self):
        return self.ops.get_water_content()

    @property
    def enthalpy(self):
        return self.ops.enthalpy_state.E

    @property
    def sigma(self):
        return self.ops.sigma
