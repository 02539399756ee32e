"""Per-level MAC grid with every field of the SSA path (glide/grid.py:147-372).

Shapes are the reference's: u (ny, nx+1), v (ny+1, nx), cells (ny, nx); the
dtype is float64 and the arrays are torch tensors on the ROCm device.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import torch

from .field import Constant, Field, GridEntity, SubgridField

DTYPE = torch.float64


def default_device() -> torch.device:
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


@dataclass
class State:
    u: Field | None = None
    v: Field | None = None
    H: Field | None = None
    H_prev: Field | None = None
    phi: Field | None = None
    mask: Field | None = None


@dataclass
class AdjointState:
    lambda_u: Field | None = None
    lambda_v: Field | None = None
    lambda_H: Field | None = None


@dataclass
class Geometry:
    bed: SubgridField | None = None
    thklim: Constant = field(default_factory=lambda: Constant(0.1, "thklim", "m"))
    flotation_reg_driving: Constant = field(default_factory=lambda: Constant(0.1, "flotation_reg_driving", "m^{-1}"))


@dataclass
class Rheology:
    B: Field | None = None
    n: Constant = field(default_factory=lambda: Constant(3.0, "n", ""))
    eps_reg: Constant = field(default_factory=lambda: Constant(1e-6, "eps_reg", "s^{-2}"))


@dataclass
class Sliding:
    beta: Field | None = None
    m: Constant = field(default_factory=lambda: Constant(1.0, "m", ""))
    u_reg: Constant = field(default_factory=lambda: Constant(1.0, "u_reg", "m a^{-1}"))
    water_drag: Constant = field(default_factory=lambda: Constant(1e-5, "water_drag", ""))
    flotation_reg_sliding: Constant = field(default_factory=lambda: Constant(0.1, "flotation_reg_sliding", "m"))


@dataclass
class Calving:
    calving_rate: Constant = field(default_factory=lambda: Constant(0.0, "calving_rate", "m a^{-1}"))
    flotation_reg_calving: Constant = field(default_factory=lambda: Constant(0.1, "flotation_reg_calving", "m"))


@dataclass
class Forcing:
    smb: Field | None = None


class Grid:
    """Single level of the multigrid hierarchy (glide/grid.py:147-252)."""

    def __init__(self, ny: int, nx: int, dx: float, x0: float = 0.0, y0: float = 0.0, crs=None, parent=None,
                 state: State | None = None, adjoint: AdjointState | None = None, geometry: Geometry | None = None,
                 rheology: Rheology | None = None, sliding: Sliding | None = None, calving: Calving | None = None,
                 forcing: Forcing | None = None, device: torch.device | str | None = None):
        self.parent = parent
        self.child = None
        self.ny, self.nx = int(ny), int(nx)
        self.dx = float(dx)
        self.x0, self.y0 = float(x0), float(y0)
        self.crs = crs
        self.device = torch.device(device) if device is not None else (
            parent.device if parent is not None else default_device())
        if self.device.type == "cuda" and parent is None:
            from . import _lib
            _lib.init_device(self.device)
        self.x_cell = x0 + dx * torch.arange(nx, dtype=DTYPE)
        self.y_cell = y0 - dx * torch.arange(ny, dtype=DTYPE)
        self.x_vfacet = x0 - dx / 2 + dx * torch.arange(nx + 1, dtype=DTYPE)
        self.y_hfacet = y0 + dx / 2 - dx * torch.arange(ny + 1, dtype=DTYPE)
        self.nu = ny * (nx + 1)
        self.nv = (ny + 1) * nx
        self.nh = ny * nx
        self.n_total = self.nu + self.nv + self.nh
        self.state = state if state is not None else self._allocate_state()
        self.geometry = geometry if geometry is not None else Geometry(bed=self._cell("bed", "m", Sub=True))
        self.rheology = rheology if rheology is not None else Rheology(B=self._cell("B", ""))
        self.sliding = sliding if sliding is not None else Sliding(beta=self._cell("beta", ""))
        self.calving = calving if calving is not None else Calving()
        self.forcing = forcing if forcing is not None else Forcing(smb=self._cell("smb", "m a^{-1}"))
        self._adjoint = adjoint
        self._forward_operators = None
        self._adjoint_operators = None
        self._work = None

    # ---------------------------------------------------------------- allocation
    def zeros(self, shape) -> torch.Tensor:
        return torch.zeros(shape, dtype=DTYPE, device=self.device)

    def _u(self, name, units=None):
        return Field(self.zeros((self.ny, self.nx + 1)), GridEntity.VERTICAL_FACET, self.dx, self, name, units)

    def _v(self, name, units=None):
        return Field(self.zeros((self.ny + 1, self.nx)), GridEntity.HORIZONTAL_FACET, self.dx, self, name, units)

    def _cell(self, name, units=None, Sub=False):
        cls = SubgridField if Sub else Field
        return cls(self.zeros((self.ny, self.nx)), GridEntity.CELL, self.dx, self, name, units)

    def _allocate_state(self) -> State:
        return State(u=self._u("u", "m a^{-1}"), v=self._v("v", "m a^{-1}"), H=self._cell("H", "m"),
                     H_prev=self._cell("H_prev", "m"), phi=self._cell("phi", ""), mask=self._cell("mask", ""))

    @property
    def adjoint(self) -> AdjointState:
        if self._adjoint is None:
            self._adjoint = AdjointState(lambda_u=self._u("lambda_u"), lambda_v=self._v("lambda_v"),
                                         lambda_H=self._cell("lambda_H"))
        return self._adjoint

    def has_adjoint(self) -> bool:
        return self._adjoint is not None

    @property
    def forward_operators(self):
        if self._forward_operators is None:
            from .operators import ForwardOperators
            self._forward_operators = ForwardOperators(self)
        return self._forward_operators

    @property
    def adjoint_operators(self):
        if self._adjoint_operators is None:
            from .operators import AdjointOperators
            self._adjoint_operators = AdjointOperators(self)
        return self._adjoint_operators

    @property
    def work(self) -> torch.Tensor:
        """Smoother workspace: per-cell patch increments, 5 planes of (ny, nx)."""
        if self._work is None:
            self._work = torch.zeros((5, self.ny, self.nx), dtype=DTYPE, device=self.device)
        return self._work

    def spawn_child(self) -> "Grid":
        child = Grid(self.ny // 2, self.nx // 2, self.dx * 2, parent=self, device=self.device)
        self.child = child
        return child

    def physics(self):
        """Scalar parameters of this level as the native struct."""
        from ._lib import Physics
        return Physics(n=self.rheology.n.value, eps_reg=self.rheology.eps_reg.value,
                       flotation_reg_driving=self.geometry.flotation_reg_driving.value,
                       m=self.sliding.m.value, u_reg=self.sliding.u_reg.value,
                       water_drag=self.sliding.water_drag.value,
                       flotation_reg_sliding=self.sliding.flotation_reg_sliding.value,
                       calving_rate=self.calving.calving_rate.value,
                       flotation_reg_calving=self.calving.flotation_reg_calving.value,
                       thklim=self.geometry.thklim.value)

    def level_struct(self, fas_scratch=None, adj_scratch=None):
        """Native level descriptor pointing at this grid's tensors."""
        from ._lib import Level, ptr
        L = Level()
        L.ny, L.nx, L.dx = self.ny, self.nx, self.dx
        L.phys = self.physics()
        s = self.state
        L.u, L.v, L.H, L.H_prev = ptr(s.u.data), ptr(s.v.data), ptr(s.H.data), ptr(s.H_prev.data)
        L.phi, L.mask = ptr(s.phi.data), ptr(s.mask.data)
        L.bed, L.B = ptr(self.geometry.bed.data), ptr(self.rheology.B.data)
        L.beta, L.smb = ptr(self.sliding.beta.data), ptr(self.forcing.smb.data)
        L.work = ptr(self.work)
        fo = self._forward_operators
        if fo is not None:
            for n in ("f_u", "f_v", "f_H", "r_u", "r_v", "r_H", "F_u", "F_v", "F_H", "gamma"):
                setattr(L, n, ptr(getattr(fo, n)))
        if fas_scratch is not None:
            for n in ("w_u", "w_v", "w_H", "y_u", "y_v", "y_H", "z_u", "z_v", "z_H", "chi"):
                setattr(L, n, ptr(getattr(fas_scratch, n)))
            L.cphi = ptr(fas_scratch.phi)
        if self._adjoint is not None:
            a = self._adjoint
            L.lambda_u, L.lambda_v, L.lambda_H = ptr(a.lambda_u.data), ptr(a.lambda_v.data), ptr(a.lambda_H.data)
        ao = self._adjoint_operators
        if ao is not None:
            L.af_u, L.af_v, L.af_H = ptr(ao.f_u), ptr(ao.f_v), ptr(ao.f_H)
            L.ar_u, L.ar_v, L.ar_H = ptr(ao.r_u), ptr(ao.r_v), ptr(ao.r_H)
            L.vjp_u, L.vjp_v, L.vjp_H = ptr(ao.vjp_u), ptr(ao.vjp_v), ptr(ao.vjp_H)
            if ao.use_cache:
                L.adj_cache = ptr(ao.cache)
        if adj_scratch is not None:
            L.wl_u, L.wl_v, L.wl_H = ptr(adj_scratch.w_lambda_u), ptr(adj_scratch.w_lambda_v), ptr(
                adj_scratch.w_lambda_H)
            L.zl_u, L.zl_v, L.zl_H = ptr(adj_scratch.z_lambda_u), ptr(adj_scratch.z_lambda_v), ptr(
                adj_scratch.z_lambda_H)
        if self.sliding.beta.has_grad():
            L.g_beta = ptr(self.sliding.beta.grad)
        if self.geometry.bed.has_grad():
            L.g_bed = ptr(self.geometry.bed.grad)
        if self.state.H_prev.has_grad():
            L.g_H_prev = ptr(self.state.H_prev.grad)
        if self.forcing.smb.has_grad():
            L.g_smb = ptr(self.forcing.smb.grad)
        return L
