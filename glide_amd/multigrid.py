"""Multigrid hierarchy, field managers and the FAS solvers (glide/multigrid.py).

The hierarchy and the user-facing option objects keep the reference's names
and defaults.  `FASCDSolver.solve` / `FASAdjointSolver.solve` hand the whole
solve to the native drivers (one C call per solve; the V-cycles, smoothing
passes and transfers are enqueued from C++ on torch's current stream).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import call, ptr, stream_handle
from .field import BroadcastOption, LocalOption
from .grid import Grid
from .operators import NewtonConfig, VankaConfig


class Multigrid:
    """Grid hierarchy built by halving (glide/multigrid.py:8-87)."""

    def __init__(self, n_levels: int, finest_grid: Grid | None = None, ny=None, nx=None, dx=None, x0=0.0, y0=0.0,
                 crs=None, device=None, verbose: bool = False):
        if finest_grid is not None:
            self.finest_grid = finest_grid
        else:
            self.finest_grid = Grid(ny, nx, dx, x0=x0, y0=y0, crs=crs, device=device)
        if verbose:
            print("Instantiating multigrid", flush=True)
        self.n_levels = n_levels
        if n_levels is not None:
            self.create_grid_hierarchy(n_levels, restrict_fields=True)
        self.state = MGStateManager(self)
        self.geometry = MGGeometryManager(self)
        self.rheology = MGRheologyManager(self)
        self.sliding = MGSlidingManager(self)
        self.calving = MGCalvingManager(self)
        self.forcing = MGForcingManager(self)
        self._adjoint = None

    @property
    def adjoint(self):
        if self._adjoint is None:
            self._adjoint = MGAdjointManager(self)
        return self._adjoint

    @property
    def device(self):
        return self.finest_grid.device

    def create_grid_hierarchy(self, n_levels, restrict_fields=True):
        g = self.finest_grid
        f = 2 ** (n_levels - 1)
        if g.ny % f or g.nx % f:
            raise ValueError(f"grid ({g.ny}, {g.nx}) is not divisible by 2^(n_levels-1) = {f}")
        self.levels = [self.finest_grid]
        for _ in range(1, n_levels):
            self.levels.append(self.create_coarse_grid(self.levels[-1], restrict_fields=restrict_fields))
        return self.levels

    def create_coarse_grid(self, parent_grid, restrict_fields=True):
        child = Grid(parent_grid.ny // 2, parent_grid.nx // 2, parent_grid.dx * 2,
                     x0=parent_grid.x0 + parent_grid.dx / 2, y0=parent_grid.y0 - parent_grid.dx / 2,
                     crs=parent_grid.crs, parent=parent_grid, device=parent_grid.device)
        parent_grid.child = child
        if restrict_fields and child.device.type == "cuda":
            self.restrict_state(parent_grid, child)
            self.restrict_geometry(parent_grid, child)
            self.restrict_rheology(parent_grid, child)
            self.restrict_sliding(parent_grid, child)
            self.restrict_calving(parent_grid, child)
            self.restrict_forcing(parent_grid, child)
        elif restrict_fields:
            self._copy_constants(parent_grid, child)
        return child

    @staticmethod
    def _copy_constants(f, c):
        c.geometry.thklim.set(f.geometry.thklim.value)
        c.geometry.flotation_reg_driving.set(f.geometry.flotation_reg_driving.value)
        c.rheology.n.set(f.rheology.n.value)
        c.rheology.eps_reg.set(f.rheology.eps_reg.value)
        for k in ("m", "u_reg", "water_drag", "flotation_reg_sliding"):
            getattr(c.sliding, k).set(getattr(f.sliding, k).value)
        c.calving.calving_rate.set(f.calving.calving_rate.value)
        c.calving.flotation_reg_calving.set(f.calving.flotation_reg_calving.value)

    def restrict_state(self, fine_grid, coarse_grid):
        self.restrict_vfacet(fine_grid.state.u.data, coarse_grid.state.u.data)
        self.restrict_hfacet(fine_grid.state.v.data, coarse_grid.state.v.data)
        self.restrict_cell(fine_grid.state.H.data, coarse_grid.state.H.data)
        self.restrict_cell(fine_grid.state.H_prev.data, coarse_grid.state.H_prev.data)
        self.restrict_cell(fine_grid.state.phi.data, coarse_grid.state.phi.data)
        self.restrict_cell(fine_grid.state.mask.data, coarse_grid.state.mask.data, method="max")

    def restrict_geometry(self, f, c):
        self.restrict_cell(f.geometry.bed.data, c.geometry.bed.data)
        c.geometry.thklim.set(f.geometry.thklim.value)
        c.geometry.flotation_reg_driving.set(f.geometry.flotation_reg_driving.value)

    def restrict_rheology(self, f, c):
        self.restrict_cell(f.rheology.B.data, c.rheology.B.data)
        c.rheology.n.set(f.rheology.n.value)
        c.rheology.eps_reg.set(f.rheology.eps_reg.value)

    def restrict_sliding(self, f, c):
        self.restrict_cell(f.sliding.beta.data, c.sliding.beta.data)
        for k in ("m", "u_reg", "water_drag", "flotation_reg_sliding"):
            getattr(c.sliding, k).set(getattr(f.sliding, k).value)

    def restrict_calving(self, f, c):
        c.calving.calving_rate.set(f.calving.calving_rate.value)
        c.calving.flotation_reg_calving.set(f.calving.flotation_reg_calving.value)

    def restrict_forcing(self, f, c):
        self.restrict_cell(f.forcing.smb.data, c.forcing.smb.data)

    def restrict_residual(self, f, c):
        fo, co = f.forward_operators, c.forward_operators
        self.restrict_vfacet(fo.r_u, co.r_u)
        self.restrict_hfacet(fo.r_v, co.r_v)
        self.restrict_cell(fo.r_H, co.r_H)

    # ---------------------------------------------------------------- transfers
    def _st(self):
        return stream_handle(self.device)

    def restrict_vfacet(self, fine, coarse=None):
        ny, nx1 = fine.shape
        nyc, nxc = ny // 2, (nx1 - 1) // 2
        if coarse is None:
            coarse = torch.empty((nyc, nxc + 1), dtype=fine.dtype, device=fine.device)
        call("glide_restrict_vfacet", ptr(fine), ptr(coarse), nyc, nxc, self._st())
        return coarse

    def restrict_hfacet(self, fine, coarse=None):
        ny1, nx = fine.shape
        nyc, nxc = (ny1 - 1) // 2, nx // 2
        if coarse is None:
            coarse = torch.empty((nyc + 1, nxc), dtype=fine.dtype, device=fine.device)
        call("glide_restrict_hfacet", ptr(fine), ptr(coarse), nyc, nxc, self._st())
        return coarse

    def restrict_cell(self, fine, coarse=None, method="avg"):
        if method not in ("avg", "max"):
            raise TypeError("Valid restriction methods: [avg, max]")
        ny, nx = fine.shape
        if coarse is None:
            coarse = torch.empty((ny // 2, nx // 2), dtype=fine.dtype, device=fine.device)
        call("glide_restrict_cell", ptr(fine), ptr(coarse), ny // 2, nx // 2, 1 if method == "max" else 0, self._st())
        return coarse

    def prolongate_vfacet(self, coarse, fine=None, method="bilinear"):
        if method != "bilinear":
            raise TypeError("Valid facet prolongation method: bilinear")
        ny, nx1 = coarse.shape
        if fine is None:
            fine = torch.empty((2 * ny, 2 * (nx1 - 1) + 1), dtype=coarse.dtype, device=coarse.device)
        call("glide_prolong_vfacet_bilinear", ptr(coarse), ptr(fine), 2 * ny, 2 * (nx1 - 1), self._st())
        return fine

    def prolongate_hfacet(self, coarse, fine=None, method="bilinear", compat=True):
        """v-facet prolongation.  compat=True reproduces the reference, which
        routes v through the u-facet routine on reinterpreted memory
        (glide/multigrid.py:212-235; square grids only)."""
        if method != "bilinear":
            raise TypeError("Valid facet prolongation method: bilinear")
        ny1, nx = coarse.shape
        nyf, nxf = 2 * (ny1 - 1), 2 * nx
        if fine is None:
            fine = torch.empty((nyf + 1, nxf), dtype=coarse.dtype, device=coarse.device)
        if compat:
            if nyf != nxf:
                raise ValueError("compat v prolongation is defined for square grids only")
            call("glide_prolong_vfacet_bilinear", ptr(coarse), ptr(fine), nyf, nxf, self._st())
        else:
            call("glide_prolong_hfacet_bilinear", ptr(coarse), ptr(fine), nyf, nxf, self._st())
        return fine

    def prolongate_cell(self, coarse, fine=None, method="injection"):
        if method not in ("injection", "bilinear"):
            raise TypeError("Valid prolongation methods: [injection, bilinear]")
        ny, nx = coarse.shape
        if fine is None:
            fine = torch.empty((2 * ny, 2 * nx), dtype=coarse.dtype, device=coarse.device)
        call("glide_prolong_cell", ptr(coarse), ptr(fine), 2 * ny, 2 * nx, 1 if method == "bilinear" else 0,
             self._st())
        return fine

    def __getitem__(self, key):
        return self.levels[key]


class HierarchyFieldManager:
    """Set a field on one level and restrict it down (glide/multigrid.py:264-283)."""

    def __init__(self, levels, getter, restrict, name=None):
        self._levels = levels
        self._getter = getter
        self._restrict = restrict
        self._name = name

    def set(self, value, start_level=0):
        self._getter(self._levels[start_level]).set(value)
        self.restrict_down(start_level)

    def restrict_down(self, start_level):
        for l in range(start_level, len(self._levels) - 1):
            self._restrict(self._getter(self._levels[l]), self._getter(self._levels[l + 1]))

    def set_level(self, level, value):
        self._getter(self._levels[level]).set(value)


def _field_manager(mg, getter, kind, name):
    if kind == "u":
        r = lambda f, c: mg.restrict_vfacet(f.data, c.data)
    elif kind == "v":
        r = lambda f, c: mg.restrict_hfacet(f.data, c.data)
    elif kind == "max":
        r = lambda f, c: mg.restrict_cell(f.data, c.data, method="max")
    elif kind == "const":
        r = lambda f, c: c.set(f.value)
    else:
        r = lambda f, c: mg.restrict_cell(f.data, c.data, method="avg")
    return HierarchyFieldManager(mg.levels, getter=getter, restrict=r, name=name)


class MGStateManager:
    def __init__(self, mg):
        self.mg = mg
        self.u = _field_manager(mg, lambda g: g.state.u, "u", "u")
        self.v = _field_manager(mg, lambda g: g.state.v, "v", "v")
        self.H = _field_manager(mg, lambda g: g.state.H, "avg", "H")
        self.H_prev = _field_manager(mg, lambda g: g.state.H_prev, "avg", "H_prev")
        self.phi = _field_manager(mg, lambda g: g.state.phi, "avg", "phi")
        self.mask = _field_manager(mg, lambda g: g.state.mask, "max", "mask")


class MGAdjointManager:
    def __init__(self, mg):
        self.mg = mg
        self.lambda_u = _field_manager(mg, lambda g: g.adjoint.lambda_u, "u", "lambda_u")
        self.lambda_v = _field_manager(mg, lambda g: g.adjoint.lambda_v, "v", "lambda_v")
        self.lambda_H = _field_manager(mg, lambda g: g.adjoint.lambda_H, "avg", "lambda_H")


class MGGeometryManager:
    def __init__(self, mg):
        self.mg = mg
        self.bed = _field_manager(mg, lambda g: g.geometry.bed, "avg", "bed")
        self.flotation_reg_driving = _field_manager(mg, lambda g: g.geometry.flotation_reg_driving, "const",
                                                    "flotation_reg_driving")
        self.thklim = _field_manager(mg, lambda g: g.geometry.thklim, "const", "thklim")


class MGRheologyManager:
    def __init__(self, mg):
        self.mg = mg
        self.B = _field_manager(mg, lambda g: g.rheology.B, "avg", "B")
        self.n = _field_manager(mg, lambda g: g.rheology.n, "const", "n")
        self.eps_reg = _field_manager(mg, lambda g: g.rheology.eps_reg, "const", "eps_reg")


class MGSlidingManager:
    def __init__(self, mg):
        self.mg = mg
        self.beta = _field_manager(mg, lambda g: g.sliding.beta, "avg", "beta")
        self.m = _field_manager(mg, lambda g: g.sliding.m, "const", "m")
        self.u_reg = _field_manager(mg, lambda g: g.sliding.u_reg, "const", "u_reg")
        self.water_drag = _field_manager(mg, lambda g: g.sliding.water_drag, "const", "water_drag")
        self.flotation_reg_sliding = _field_manager(mg, lambda g: g.sliding.flotation_reg_sliding, "const",
                                                    "flotation_reg_sliding")


class MGCalvingManager:
    def __init__(self, mg):
        self.mg = mg
        self.calving_rate = _field_manager(mg, lambda g: g.calving.calving_rate, "const", "calving_rate")
        self.flotation_reg_calving = _field_manager(mg, lambda g: g.calving.flotation_reg_calving, "const",
                                                    "flotation_reg_calving")


class MGForcingManager:
    def __init__(self, mg):
        self.mg = mg
        self.smb = _field_manager(mg, lambda g: g.forcing.smb, "avg", "smb")


# ====================================================================== configs
def _check_vcycles(n) -> None:
    if not 0 <= int(n) <= 64:
        raise ValueError(f"maximum_vcycles must be in [0, 64] (got {n})")


@dataclass
class FASCDConfig:
    """glide/multigrid.py:671-682; compat_vprolong reproduces the reference's
    v-correction prolongation (square grids)."""
    freeze_coarse_calving: bool = True
    freeze_coarse_phi: bool = True
    coarsest_steps: int = 200
    pre_steps: int = 10
    post_steps: int = 20
    finest_steps: int = 50
    maximum_vcycles: int = 10
    relative_tolerance: float = 1e-3
    absolute_tolerance: float = 5.0
    report_norms: bool = True
    compat_vprolong: bool = True
    # (an addition) fixed-cycle solves (both tolerances <= 0) return at once,
    # without waiting for the GPU, computing residual norms or filling a
    # report (the residual buffers then hold the last cycle's inner residual);
    # solve() returns None
    asynchronous: bool = False

    def to_c(self) -> _lib.FasConfigC:
        _check_vcycles(self.maximum_vcycles)
        return _lib.FasConfigC(
            freeze_coarse_calving=int(bool(self.freeze_coarse_calving)),
            freeze_coarse_phi=int(bool(self.freeze_coarse_phi)), coarsest_steps=int(self.coarsest_steps),
            pre_steps=int(self.pre_steps), post_steps=int(self.post_steps), finest_steps=int(self.finest_steps),
            maximum_vcycles=int(self.maximum_vcycles), relative_tolerance=float(self.relative_tolerance),
            absolute_tolerance=float(self.absolute_tolerance), compat_vprolong=int(bool(self.compat_vprolong)))


@dataclass
class FASAdjointConfig:
    """glide/multigrid.py:992-1001."""
    coarsest_steps: int = 200
    pre_steps: int = 10
    post_steps: int = 20
    finest_steps: int = 50
    maximum_vcycles: int = 10
    relative_tolerance: float = 1e-3
    absolute_tolerance: float = 5.0
    report_norms: bool = True
    compat_vprolong: bool = True
    # (an addition) fixed-cycle solves (both tolerances <= 0) return at once,
    # without waiting for the GPU, computing residual norms or filling a
    # report (the residual buffers then hold the last cycle's inner residual);
    # solve() returns None
    asynchronous: bool = False

    def to_c(self) -> _lib.FasConfigC:
        _check_vcycles(self.maximum_vcycles)
        return _lib.FasConfigC(
            freeze_coarse_calving=0, freeze_coarse_phi=1, coarsest_steps=int(self.coarsest_steps),
            pre_steps=int(self.pre_steps), post_steps=int(self.post_steps), finest_steps=int(self.finest_steps),
            maximum_vcycles=int(self.maximum_vcycles), relative_tolerance=float(self.relative_tolerance),
            absolute_tolerance=float(self.absolute_tolerance), compat_vprolong=int(bool(self.compat_vprolong)))


def _no_report(cfg, report_norms) -> bool:
    """An asynchronous fixed-cycle solve that prints nothing needs no report."""
    fixed = not (cfg.relative_tolerance > 0.0) and not (cfg.absolute_tolerance > 0.0)
    return bool(cfg.asynchronous) and fixed and not (cfg.report_norms and report_norms)


def dataclass_field_names(cls):
    """Option names of a configuration dataclass, in declaration order."""
    return list(cls.__dataclass_fields__)


def validate_kwargs(config_cls, kwargs):
    """Refuse option names the configuration does not define (AttributeError,
    as the reference's option setters raise)."""
    known = config_cls.__dataclass_fields__
    bad = [k for k in kwargs if k not in known]
    if bad:
        raise AttributeError(f"{config_cls.__name__} has no option {', '.join(map(repr, bad))} "
                             f"(options: {', '.join(known)})")


class _ConfigOptions:
    """User-facing wrapper around one solver config (glide/multigrid.py:696-776)."""
    _cls = None
    options: list = []

    def __init__(self, config):
        self._config = config
        for name in dataclass_field_names(self._cls):
            setattr(self, name, LocalOption(getter=(lambda n=name: getattr(self._config, n)),
                                            setter=(lambda v, n=name: setattr(self._config, n, v)), name=name))

    def set(self, **kwargs):
        validate_kwargs(self._cls, kwargs)
        for k, v in kwargs.items():
            setattr(self._config, k, v)

    def __repr__(self):
        return "fas_options={" + ", ".join(f"{getattr(self, o)}" for o in self.options) + "}"


class FASCDOptions(_ConfigOptions):
    _cls = FASCDConfig
    options = ["freeze_coarse_calving", "freeze_coarse_phi", "coarsest_steps", "pre_steps", "post_steps",
               "finest_steps"]


class FASAdjointOptions(_ConfigOptions):
    _cls = FASAdjointConfig
    options = ["coarsest_steps", "pre_steps", "post_steps", "finest_steps"]


class NewtonOptions:
    """Broadcast NewtonConfig across levels (glide/multigrid.py:820-852)."""

    def __init__(self, levels, getter):
        self._levels = levels
        self._getter = getter
        self.options = ["steps", "relaxation", "ssa_damping", "mc_damping"]
        for o in self.options:
            setattr(self, o, BroadcastOption(levels, getter, o))

    def set(self, **kwargs):
        validate_kwargs(NewtonConfig, kwargs)
        for lev in self._levels:
            for k, v in kwargs.items():
                setattr(self._getter(lev), k, v)

    def set_level(self, level_index, **kwargs):
        validate_kwargs(NewtonConfig, kwargs)
        for k, v in kwargs.items():
            setattr(self._getter(self._levels[level_index]), k, v)

    def __repr__(self):
        return "newton_options={" + ", ".join(f"{getattr(self, o)}" for o in self.options) + "}"


class VankaOptions:
    """Broadcast VankaConfig across levels (glide/multigrid.py:780-816)."""

    def __init__(self, levels, getter):
        self._levels = levels
        self._getter = getter
        self.options = ["omega", "relax_phi", "local_fp32", "multicolour", "newton_tol"]
        for name in self.options:
            setattr(self, name, BroadcastOption(levels, getter, name))
        self.newton_options = NewtonOptions(levels, getter=lambda lev: getter(lev).newton_config)

    def set(self, **kwargs):
        validate_kwargs(VankaConfig, kwargs)
        for lev in self._levels:
            for k, v in kwargs.items():
                setattr(self._getter(lev), k, v)

    def set_level(self, level_index, **kwargs):
        validate_kwargs(VankaConfig, kwargs)
        for k, v in kwargs.items():
            setattr(self._getter(self._levels[level_index]), k, v)

    def __repr__(self):
        return "vanka_options={" + ", ".join(f"{getattr(self, o)}" for o in self.options) + "}"


# ====================================================================== solvers
class FASCDScratch:
    """glide/multigrid.py:647-664."""

    def __init__(self, grid):
        z = grid.zeros
        ny, nx = grid.ny, grid.nx
        self.w_u, self.w_v, self.w_H = z((ny, nx + 1)), z((ny + 1, nx)), z((ny, nx))
        self.y_u, self.y_v, self.y_H = z((ny, nx + 1)), z((ny + 1, nx)), z((ny, nx))
        self.z_u, self.z_v, self.z_H = z((ny, nx + 1)), z((ny + 1, nx)), z((ny, nx))
        self.chi, self.phi = z((ny, nx)), z((ny, nx))


@dataclass
class FASCDLevel:
    grid: Grid
    scratch: FASCDScratch


class FASAdjointScratch:
    """glide/multigrid.py:975-985."""

    def __init__(self, grid):
        z = grid.zeros
        ny, nx = grid.ny, grid.nx
        self.w_lambda_u, self.w_lambda_v, self.w_lambda_H = z((ny, nx + 1)), z((ny + 1, nx)), z((ny, nx))
        self.z_lambda_u, self.z_lambda_v, self.z_lambda_H = z((ny, nx + 1)), z((ny + 1, nx)), z((ny, nx))


@dataclass
class FASAdjointLevel:
    grid: Grid
    scratch: FASAdjointScratch


def _print_report(rep, n3_final):
    print(f"  Initial:   |r0|     = {rep.initial_norm:.2e}", flush=True)
    for k in range(rep.vcycles):
        rel = rep.norms[k] / rep.initial_norm if rep.initial_norm > 0 else 0.0
        print(f"  V-cycle {k}: |r|/|r0| = {rel:.2e}, |r| = {rep.norms[k]:.2e}", flush=True)
    print(f"  final |r_u| = {n3_final[0]:.2e}, |r_v| = {n3_final[1]:.2e}, |r_H| = {n3_final[2]:.2e}", flush=True)


def _mixed_hierarchy(multigrid):
    """The FP32 mirror of the hierarchy, shared by the forward and adjoint solvers."""
    from .mixed import MixedHierarchy
    mh = getattr(multigrid, "_mixed_hierarchy", None)
    if mh is None:
        mh = MixedHierarchy(multigrid.levels)
        multigrid._mixed_hierarchy = mh
    return mh


class FASCDSolver:
    """FAS constrained-descent solver (glide/multigrid.py:486-645).

    ``precision`` (an addition): "fp64" (default) runs the reference's solve in
    FP64; "mixed" runs the defect-corrected FP32 solve of glide_amd.mixed with
    ``mixed_options`` (a MixedConfig) and ignores fas_options/vanka_options."""

    def __init__(self, multigrid):
        from .mixed import MixedConfig
        self.multigrid = multigrid
        self.levels = [FASCDLevel(g, FASCDScratch(g)) for g in multigrid.levels]
        self._fas_config = FASCDConfig()
        self.fas_options = FASCDOptions(self._fas_config)
        self.vanka_options = VankaOptions(self.levels, getter=lambda lev: lev.grid.forward_operators.vanka_config)
        self.n_levels = len(self.levels)
        self.dt = None
        self.last_report = None
        self.precision = "fp64"
        self.mixed_options = MixedConfig()
        self.last_fallbacks = 0
        self.last_stats = None

    def _structs(self):
        arr = (_lib.Level * self.n_levels)()
        for k, lev in enumerate(self.levels):
            lev.grid.forward_operators  # allocate
            arr[k] = lev.grid.level_struct(fas_scratch=lev.scratch)
        vcs = (_lib.VankaConfigC * self.n_levels)()
        for k, lev in enumerate(self.levels):
            vcs[k] = lev.grid.forward_operators.vanka_config.to_c()
        return arr, vcs

    def _check_square(self, start_level):
        if self._fas_config.compat_vprolong:
            for lev in self.levels[start_level:]:
                if lev.grid.ny != lev.grid.nx:
                    raise ValueError("compat_vprolong needs square levels; set fas_options compat_vprolong=False")

    def _solve_mixed(self, start_level, report_norms):
        from . import mixed
        arr, _ = self._structs()
        rep, self.last_stats = mixed.forward_solve(arr, _mixed_hierarchy(self.multigrid), start_level, self.dt,
                                                   self.mixed_options, stream_handle(self.multigrid.device))
        self.last_fallbacks = self.last_stats["fallbacks"]
        self.last_report = rep
        if rep is not None and self._fas_config.report_norms and report_norms:
            _print_report(rep, list(rep.norms3))
        return rep

    def solve(self, dt, start_level=0, report_norms=True):
        """glide/multigrid.py:502-533 (one native call)."""
        self.dt = float(dt)
        if self.precision == "mixed":
            return self._solve_mixed(start_level, report_norms)
        if self.precision != "fp64":
            raise ValueError(f"precision must be 'fp64' or 'mixed' (got {self.precision!r})")
        self._check_square(start_level)
        arr, vcs = self._structs()
        cfg = self._fas_config.to_c()
        rep = None if _no_report(self._fas_config, report_norms) else _lib.SolveReport()
        st = stream_handle(self.multigrid.device)
        self.last_report = rep
        call("glide_fascd_solve", arr, self.n_levels, int(start_level), self.dt, ctypes.byref(cfg), vcs,
             ctypes.byref(rep) if rep is not None else None, st)
        if rep is not None and self._fas_config.report_norms and report_norms:
            _print_report(rep, list(rep.norms3))
        return rep

    def vcycle(self, l, finest=False):
        """One V-cycle from level l (glide/multigrid.py:536-645)."""
        if self.dt is None:
            raise RuntimeError("vcycle needs dt: call solve() first or set solver.dt")
        arr, vcs = self._structs()
        cfg = self._fas_config.to_c()
        call("glide_fascd_vcycle", arr, self.n_levels, int(l), int(bool(finest)), self.dt, ctypes.byref(cfg), vcs,
             stream_handle(self.multigrid.device))


class FASAdjointSolver:
    """FAS solver of the adjoint system (glide/multigrid.py:857-973); ``precision``
    and ``mixed_options`` as for FASCDSolver (the mixed form needs linear drag)."""

    def __init__(self, multigrid):
        from .mixed import MixedConfig
        self.multigrid = multigrid
        self.levels = [FASAdjointLevel(g, FASAdjointScratch(g)) for g in multigrid.levels]
        self._fas_config = FASAdjointConfig()
        self.fas_options = FASAdjointOptions(self._fas_config)
        self.vanka_options = VankaOptions(self.levels, getter=lambda lev: lev.grid.adjoint_operators.vanka_config)
        self.n_levels = len(self.levels)
        self.dt = None
        self.last_report = None
        self.precision = "fp64"
        # 4 finest-level passes per mixed adjoint V-cycle (8 in the forward):
        # at 512² the flexible GMRES adjoint takes one iteration more and the
        # coupled step is ~2 ms shorter
        self.mixed_options = MixedConfig(finest_steps=4)
        self.last_fallbacks = 0
        self.last_stats = None

    def _structs(self):
        arr = (_lib.Level * self.n_levels)()
        for k, lev in enumerate(self.levels):
            lev.grid.forward_operators
            lev.grid.adjoint_operators
            arr[k] = lev.grid.level_struct(adj_scratch=lev.scratch)
        vcs = (_lib.VankaConfigC * self.n_levels)()
        for k, lev in enumerate(self.levels):
            vcs[k] = lev.grid.adjoint_operators.vanka_config.to_c()
        return arr, vcs

    def solve(self, dt, start_level=0, report_norms=True):
        self.dt = float(dt)
        if self.precision == "mixed":
            from . import mixed
            arr, _ = self._structs()
            rep, self.last_stats = mixed.adjoint_solve(arr, _mixed_hierarchy(self.multigrid), start_level,
                                                       self.dt, self.mixed_options,
                                                       stream_handle(self.multigrid.device))
            self.last_fallbacks = self.last_stats["fallbacks"]
            self.last_report = rep
            if rep is not None and self._fas_config.report_norms and report_norms:
                _print_report(rep, list(rep.norms3))
            return rep
        if self.precision != "fp64":
            raise ValueError(f"precision must be 'fp64' or 'mixed' (got {self.precision!r})")
        if self._fas_config.compat_vprolong:
            for lev in self.levels[start_level:]:
                if lev.grid.ny != lev.grid.nx:
                    raise ValueError("compat_vprolong needs square levels; set fas_options compat_vprolong=False")
        arr, vcs = self._structs()
        cfg = self._fas_config.to_c()
        rep = None if _no_report(self._fas_config, report_norms) else _lib.SolveReport()
        self.last_report = rep
        call("glide_fas_adjoint_solve", arr, self.n_levels, int(start_level), self.dt, ctypes.byref(cfg), vcs,
             ctypes.byref(rep) if rep is not None else None, stream_handle(self.multigrid.device))
        if rep is not None and self._fas_config.report_norms and report_norms:
            _print_report(rep, list(rep.norms3))
        return rep

    def vcycle(self, l, finest=False):
        if self.dt is None:
            raise RuntimeError("vcycle needs dt: call solve() first or set solver.dt")
        arr, vcs = self._structs()
        cfg = self._fas_config.to_c()
        call("glide_fas_adjoint_vcycle", arr, self.n_levels, int(l), int(bool(finest)), self.dt, ctypes.byref(cfg),
             vcs, stream_handle(self.multigrid.device))
