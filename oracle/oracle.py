"""ctypes wrapper of the FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).

Only the tests, the smoke check and the CPU-baseline leg of the benchmark use
this module, always as the checker.  The product (glide_amd) never imports it.
Arrays are float64 numpy arrays in the reference's layouts.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p


class Phys(ctypes.Structure):
    _fields_ = [(n, _D) for n in ("n", "eps_reg", "flotation_reg_driving", "m", "u_reg", "water_drag",
                                  "flotation_reg_sliding", "calving_rate", "flotation_reg_calving", "thklim")]


class FasConfig(ctypes.Structure):
    _fields_ = [("freeze_coarse_calving", _I), ("freeze_coarse_phi", _I), ("coarsest_steps", _I),
                ("pre_steps", _I), ("post_steps", _I), ("finest_steps", _I), ("maximum_vcycles", _I),
                ("relative_tolerance", _D), ("absolute_tolerance", _D), ("compat_hfacet", _I),
                ("project_pinned", _I)]


class VankaConfig(ctypes.Structure):
    _fields_ = [("omega", _D), ("relax_phi", _D), ("newton_steps", _I), ("ssa_damping", _D), ("mc_damping", _D),
                ("multicolour", _I), ("relaxation", _D), ("newton_tol", _D)]


class Report(ctypes.Structure):
    _fields_ = [("vcycles", _I), ("initial_norm", _D), ("final_norm", _D), ("norms", _D * 64)]


# field identifiers of the oracle hierarchy (ssa_oracle.cpp FieldId)
FIELD_IDS = {n: k for k, n in enumerate([
    "u", "v", "H", "H_prev", "phi", "mask", "bed", "B", "beta", "smb",
    "lambda_u", "lambda_v", "lambda_H",
    "f_u", "f_v", "f_H", "af_u", "af_v", "af_H",
    "g_beta", "g_bed", "g_H_prev", "g_smb",
    "r_u", "r_v", "r_H", "ar_u", "ar_v", "ar_H", "mask_last"])}

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(str(LIB))
        _lib.orc_mg_create.restype = _P
        _lib.orc_mg_create.argtypes = [_I, _I, _D, _I, ctypes.POINTER(Phys)]
        for name in ("orc_mg_destroy", "orc_mg_set", "orc_mg_put", "orc_mg_get", "orc_fascd_solve",
                     "orc_fas_adjoint_solve", "orc_gradients", "orc_mg_levels"):
            getattr(_lib, name).restype = None if name != "orc_mg_levels" else _I
        _lib.orc_mg_destroy.argtypes = [_P]
        _lib.orc_mg_set.argtypes = [_P, _I, _I, _P]
        _lib.orc_mg_put.argtypes = [_P, _I, _I, _P]
        _lib.orc_mg_get.argtypes = [_P, _I, _I, _P]
        _lib.orc_fascd_solve.argtypes = [_P, _D, _I, ctypes.POINTER(FasConfig), ctypes.POINTER(VankaConfig),
                                         ctypes.POINTER(Report)]
        _lib.orc_fas_adjoint_solve.argtypes = _lib.orc_fascd_solve.argtypes
        _lib.orc_gradients.argtypes = [_P, _I, _D]
        _lib.orc_num_threads.restype = _I
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous float64"
    return ctypes.c_void_p(a.ctypes.data)


def phys(n=3.0, eps_reg=1e-6, flotation_reg_driving=0.1, m=1.0, u_reg=1.0, water_drag=1e-5,
         flotation_reg_sliding=0.1, calving_rate=0.0, flotation_reg_calving=0.1, thklim=0.1) -> Phys:
    return Phys(n, eps_reg, flotation_reg_driving, m, u_reg, water_drag, flotation_reg_sliding, calving_rate,
                flotation_reg_calving, thklim)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def num_threads() -> int:
    return lib().orc_num_threads()


# ---------------------------------------------------------------- kernels
def grounded(H, bed, c=0.1, relax=0.0, phi=None):
    ny, nx = H.shape
    out = np.zeros_like(H) if phi is None else phi.copy()
    lib().orc_grounded(_p(out), _p(H), _p(bed), _D(c), _D(relax), ny, nx)
    return out


def residual(F, P: Phys, dx, dt, use_forcing=True, use_mask=True):
    """F: dict with u, v, H, phi, mask, f_u, f_v, f_H, bed, B, beta, gamma."""
    ny, nx = F["H"].shape
    ru, rv, rH = np.zeros((ny, nx + 1)), np.zeros((ny + 1, nx)), np.zeros((ny, nx))
    lib().orc_residual(_p(ru), _p(rv), _p(rH), *[_p(F[k]) for k in ("u", "v", "H", "phi", "mask", "f_u", "f_v",
                                                                      "f_H", "bed", "B", "beta", "gamma")],
                       _I(int(use_forcing)), _I(int(use_mask)), ctypes.byref(P), _D(dx), _D(dt), ny, nx)
    return ru, rv, rH


def jvp(F, du, dv, dH, P: Phys, dx, dt, use_mask=True):
    """J d of the coupled residual with phi, mask, bed, B, beta frozen; zero
    rows on Dirichlet facets and (1 - mask) times the mass-balance derivative
    (glide/cuda/residuals.cu:340-660).  Pinned by central differences of
    residual() and by the dot product with vjp() (tests/test_oracle_ssa.py)."""
    ny, nx = F["H"].shape
    du = np.ascontiguousarray(du, dtype=np.float64)
    dv = np.ascontiguousarray(dv, dtype=np.float64)
    dH = np.ascontiguousarray(dH, dtype=np.float64)
    ou, ov, oH = np.zeros((ny, nx + 1)), np.zeros((ny + 1, nx)), np.zeros((ny, nx))
    lib().orc_jvp(_p(ou), _p(ov), _p(oH), _p(F["u"]), _p(F["v"]), _p(F["H"]), _p(du), _p(dv), _p(dH), _p(F["phi"]),
                  _p(F["mask"]), _p(F["bed"]), _p(F["B"]), _p(F["beta"]), _I(int(use_mask)), ctypes.byref(P),
                  _D(dx), _D(dt), ny, nx)
    return ou, ov, oH


def viscosity(u, v, B, n, eps_reg, dx, du=None, dv=None):
    ny, nx = B.shape
    eta, etad = np.zeros((ny, nx)), np.zeros((ny, nx))
    lib().orc_viscosity(_p(eta), _p(etad) if du is not None else None, _p(u), _p(v), _p(du), _p(dv), _p(B), _D(n),
                        _D(eps_reg), _D(dx), ny, nx)
    return (eta, etad) if du is not None else eta


def strain_heating(u, v, B, n, eps_reg, dx, rho_i=917.0, g=9.81):
    """SSA strain heating 4 eta eps_e^2 rho_i g / SEC_PER_YR (W/m^3), (ny, nx).
    Parity pinned by the closed forms of uniform stretching and simple shear
    (tests/test_oracle_enthalpy.py)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    ny, nx = B.shape
    phi = np.zeros((ny, nx))
    lib().orc_strain_heating(_p(phi), _p(u), _p(v), _p(B), _D(n), _D(eps_reg), _D(dx), _D(rho_i * g), ny, nx)
    return phi


def vanka_dump(F, P: Phys, dx, dt):
    ny, nx = F["H"].shape
    J, r = np.zeros((ny * nx, 25)), np.zeros((ny * nx, 5))
    lib().orc_vanka_dump(_p(J), _p(r), *[_p(F[k]) for k in ("u", "v", "H", "phi", "f_u", "f_v", "f_H", "bed", "B",
                                                               "beta", "gamma")],
                         ctypes.byref(P), _D(dx), _D(dt), ny, nx)
    return J, r


def local_system(x, F, P: Phys, dx, dt, i, j):
    ny, nx = F["H"].shape
    J, r = np.zeros(25), np.zeros(5)
    x = np.ascontiguousarray(x, dtype=np.float64)
    lib().orc_local_system(_p(J), _p(r), _p(x), *[_p(F[k]) for k in ("u", "v", "H", "phi", "bed", "B", "beta")],
                           ctypes.byref(P), _D(dx), _D(dt), ny, nx, i, j)
    return J.reshape(5, 5), r


def lu5(A, b):
    x = np.zeros(5)
    lib().orc_lu5(_p(np.ascontiguousarray(A, dtype=np.float64).reshape(25)), _p(np.ascontiguousarray(b, np.float64)),
                  _p(x))
    return x


def vanka_sweep(F, P: Phys, dx, dt, newton_steps=30, ssa_damping=0.01, mc_damping=1.0, omega=0.5):
    """One additive pass; returns the new (u, v, H, mask) and Newton iteration counts."""
    ny, nx = F["H"].shape
    u, v, H, mask = F["u"].copy(), F["v"].copy(), F["H"].copy(), F["mask"].copy()
    it = np.zeros((ny, nx))
    lib().orc_vanka_sweep(_p(u), _p(v), _p(H), _p(mask), *[_p(F[k]) for k in ("phi", "f_u", "f_v", "f_H", "bed", "B",
                                                                             "beta", "gamma")],
                          ctypes.byref(P), _D(dx), _D(dt), ny, nx, newton_steps, _D(ssa_damping), _D(mc_damping),
                          _D(omega), _p(it))
    return u, v, H, mask, it


def mc_sweep(F, P: Phys, dx, dt, newton_steps=4, relaxation=1.0, newton_tol=0.0, ssa_damping=0.0, mc_damping=0.0):
    """One multicolour (red-black) pass, no counterpart in the reference:
    returns the new (u, v, H, mask) and the Newton iteration counts."""
    ny, nx = F["H"].shape
    u, v, H, mask = F["u"].copy(), F["v"].copy(), F["H"].copy(), F["mask"].copy()
    it = np.zeros((ny, nx))
    lib().orc_mc_sweep(_p(u), _p(v), _p(H), _p(mask), *[_p(F[k]) for k in ("phi", "f_u", "f_v", "f_H", "bed", "B",
                                                                          "beta", "gamma")],
                       ctypes.byref(P), _D(dx), _D(dt), ny, nx, int(newton_steps), _D(relaxation), _D(newton_tol),
                       _D(ssa_damping), _D(mc_damping), _p(it))
    return u, v, H, mask, it


def mc_adjoint_sweep(F, lam, af, P: Phys, dx, dt, ssa_damping=0.0, mc_damping=0.0, omega=1.0):
    """One multicolour adjoint pass (no counterpart in the reference) on the
    multipliers lam = (l_u, l_v, l_H) with adjoint forcing af; returns the new
    multipliers."""
    ny, nx = F["H"].shape
    lu, lv, lH = (np.ascontiguousarray(x, dtype=np.float64).copy() for x in lam)
    fu, fv, fH = (np.ascontiguousarray(x, dtype=np.float64) for x in af)
    lib().orc_mc_adjoint_sweep(_p(lu), _p(lv), _p(lH), _p(F["u"]), _p(F["v"]), _p(F["H"]), _p(F["phi"]),
                               _p(F["mask"]), _p(fu), _p(fv), _p(fH), _p(F["bed"]), _p(F["B"]), _p(F["beta"]),
                               ctypes.byref(P), _D(dx), _D(dt), ny, nx, _D(ssa_damping), _D(mc_damping), _D(omega))
    return lu, lv, lH


def vjp(F, lam_u, lam_v, lam_H, P: Phys, dx, dt, use_forcing=False, af=None):
    ny, nx = F["H"].shape
    ou, ov, oH = np.zeros((ny, nx + 1)), np.zeros((ny + 1, nx)), np.zeros((ny, nx))
    z = (np.zeros((ny, nx + 1)), np.zeros((ny + 1, nx)), np.zeros((ny, nx)))
    fu, fv, fH = af if af is not None else z
    lib().orc_vjp(_p(ou), _p(ov), _p(oH), _p(F["u"]), _p(F["v"]), _p(F["H"]), _p(lam_u), _p(lam_v), _p(lam_H),
                  _p(F["phi"]), _p(F["mask"]), _p(fu), _p(fv), _p(fH), _p(F["bed"]), _p(F["B"]), _p(F["beta"]),
                  _I(int(use_forcing)), ctypes.byref(P), _D(dx), _D(dt), ny, nx)
    return ou, ov, oH


def vanka_adjoint_sweep(F, lam, r, P: Phys, dx, dt, ssa_damping=0.01, mc_damping=1.0, omega=0.5):
    ny, nx = F["H"].shape
    lu, lv, lH = (a.copy() for a in lam)
    lib().orc_vanka_adjoint_sweep(_p(lu), _p(lv), _p(lH), _p(r[0]), _p(r[1]), _p(r[2]), _p(F["u"]), _p(F["v"]),
                                  _p(F["H"]), _p(F["phi"]), _p(F["mask"]), _p(F["bed"]), _p(F["B"]), _p(F["beta"]),
                                  ctypes.byref(P), _D(dx), _D(dt), ny, nx, _D(ssa_damping), _D(mc_damping),
                                  _D(omega))
    return lu, lv, lH


def grad_beta(F, lam_u, lam_v, P: Phys):
    ny, nx = F["H"].shape
    g = np.zeros((ny, nx))
    lib().orc_grad_beta(_p(g), _p(F["u"]), _p(F["v"]), _p(F["H"]), _p(F["phi"]), _p(F["beta"]), _p(lam_u),
                        _p(lam_v), ctypes.byref(P), ny, nx)
    return g


def grad_bed(F, lam_u, lam_v, dx):
    ny, nx = F["H"].shape
    g = np.zeros((ny, nx))
    lib().orc_grad_bed(_p(g), _p(F["H"]), _p(F["phi"]), _p(F["bed"]), _p(lam_u), _p(lam_v), _D(dx), ny, nx)
    return g


def restrict_vfacet(f):
    nyc, nxc = f.shape[0] // 2, (f.shape[1] - 1) // 2
    c = np.zeros((nyc, nxc + 1))
    lib().orc_restrict_vfacet(_p(f), _p(c), nyc, nxc)
    return c


def restrict_hfacet(f):
    nyc, nxc = (f.shape[0] - 1) // 2, f.shape[1] // 2
    c = np.zeros((nyc + 1, nxc))
    lib().orc_restrict_hfacet(_p(f), _p(c), nyc, nxc)
    return c


def restrict_cell(f, use_max=False):
    nyc, nxc = f.shape[0] // 2, f.shape[1] // 2
    c = np.zeros((nyc, nxc))
    lib().orc_restrict_cell(_p(f), _p(c), nyc, nxc, int(use_max))
    return c


def prolong_vfacet_bilinear(c, nyf, nxf):
    f = np.zeros((nyf, nxf + 1))
    lib().orc_prolong_vfacet_bilinear(_p(c), _p(f), nyf, nxf)
    return f


def prolong_hfacet_bilinear(c, nyf, nxf):
    f = np.zeros((nyf + 1, nxf))
    lib().orc_prolong_hfacet_bilinear(_p(c), _p(f), nyf, nxf)
    return f


def prolong_cell(c, nyf, nxf, bilinear=False):
    f = np.zeros((nyf, nxf))
    fn = lib().orc_prolong_cell_bilinear if bilinear else lib().orc_prolong_cell_injection
    fn(_p(c), _p(f), nyf, nxf)
    return f


# ---------------------------------------------------------------- hierarchy
class OracleMG:
    """Oracle multigrid hierarchy (FAS-CD and FAS adjoint drivers)."""

    def __init__(self, ny, nx, dx, n_levels, P: Phys | None = None):
        self.P = P if P is not None else phys()
        self.n_levels = n_levels
        self.shapes = [(ny >> l, nx >> l) for l in range(n_levels)]
        self.h = lib().orc_mg_create(ny, nx, _D(dx), n_levels, ctypes.byref(self.P))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mg_destroy(self.h)
            self.h = None

    def _shape(self, name, level):
        ny, nx = self.shapes[level]
        if name in ("u", "lambda_u", "f_u", "af_u", "r_u", "ar_u"):
            return ny, nx + 1
        if name in ("v", "lambda_v", "f_v", "af_v", "r_v", "ar_v"):
            return ny + 1, nx
        return ny, nx

    def set(self, name, value, level=0):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=np.float64), self._shape(name, level)))
        lib().orc_mg_set(self.h, FIELD_IDS[name], level, _p(a))

    def put(self, name, value, level=0):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=np.float64), self._shape(name, level)))
        lib().orc_mg_put(self.h, FIELD_IDS[name], level, _p(a))

    def get(self, name, level=0):
        out = np.zeros(self._shape(name, level))
        lib().orc_mg_get(self.h, FIELD_IDS[name], level, _p(out))
        return out

    @staticmethod
    def fas_config(coarsest=200, pre=10, post=20, finest=50, vcycles=10, rtol=1e-3, atol=5.0,
                   freeze_coarse_calving=True, freeze_coarse_phi=True, compat=True) -> FasConfig:
        return FasConfig(int(freeze_coarse_calving), int(freeze_coarse_phi), coarsest, pre, post, finest, vcycles,
                         rtol, atol, int(compat))

    def vanka_configs(self, omega=0.5, relax_phi=0.0, newton_steps=30, ssa_damping=0.01, mc_damping=1.0,
                      multicolour=False, relaxation=1.0, newton_tol=1e-6):
        """Per-level smoother settings; multicolour selects the red-black
        passes (mc_sweep / mc_adjoint_sweep) with the given local Newton
        relaxation and stopping threshold instead of the additive pass."""
        arr = (VankaConfig * self.n_levels)()
        for k in range(self.n_levels):
            arr[k] = VankaConfig(omega, relax_phi, newton_steps, ssa_damping, mc_damping, int(bool(multicolour)),
                                 relaxation, newton_tol)
        return arr

    def fascd_solve(self, dt, cfg: FasConfig, vcs=None, start_level=0) -> Report:
        rep = Report()
        vcs = vcs if vcs is not None else self.vanka_configs()
        lib().orc_fascd_solve(self.h, _D(dt), start_level, ctypes.byref(cfg), vcs, ctypes.byref(rep))
        return rep

    def fas_adjoint_solve(self, dt, cfg: FasConfig, vcs=None, start_level=0) -> Report:
        rep = Report()
        vcs = vcs if vcs is not None else self.vanka_configs()
        lib().orc_fas_adjoint_solve(self.h, _D(dt), start_level, ctypes.byref(cfg), vcs, ctypes.byref(rep))
        return rep

    def gradients(self, dt, level=0):
        lib().orc_gradients(self.h, level, _D(dt))


# ---------------------------------------------------------------- enthalpy
class EnthParams(ctypes.Structure):
    _fields_ = [("dx", _D), ("dt", _D), ("drain_rate", _D), ("h_thin", _D), ("lf_c", _D), ("term_flags", _I)]


RHO_I_E, RHO_W_E, C_I, K_I, L_HEAT = 910.0, 1000.0, 2009.0, 2.1, 3.34e5
T_REF, T_MELT, BETA_CC, GRAVITY = 223.15, 273.15, 7.9e-8, 9.81
K_COLD = K_I / C_I
K_TEMP = K_COLD * 0.1
E_SCALE = C_I * (T_MELT - T_REF)
SEC_PER_YR = 365.25 * 86400.0


def _enth_lib():
    L = lib()
    if not getattr(L, "_enth_typed", False):
        L.orc_enth_residual.restype = _D
        L.orc_enth_column_sweep.restype = _I
        for f in ("orc_E_pmp", "orc_K", "orc_water"):
            getattr(L, f).restype = _D
            getattr(L, f).argtypes = [_D, _D]
        L._enth_typed = True
    return L


class OracleThermal:
    """FP64 mirror of the reference's EnthalpyOperators for one grid
    (glide/enthalpy.py:205-643); arrays are float64 numpy."""

    def __init__(self, ny, nx, nz, dx, H):
        self.ny, self.nx, self.nz, self.dx = ny, nx, nz, float(dx)
        self.sigma = np.linspace(0.0, 1.0, nz)
        self.H = np.ascontiguousarray(H, dtype=np.float64).copy()
        self.H_prev = self.H.copy()
        z3 = lambda: np.zeros((ny, nx, nz))
        self.E, self.E_prev, self.f_E, self.omega = z3(), z3(), z3(), z3()
        self.u3d = np.zeros((nz, ny, nx + 1))
        self.v3d = np.zeros((nz, ny + 1, nx))
        self.E_surface = np.zeros((ny, nx))
        self.Q_geo = np.zeros((ny, nx))
        self.Q_fh = np.zeros((ny, nx))
        self.phi_strain = z3()
        self.drain_rate = 0.01 / SEC_PER_YR
        self.h_thin = 25.0
        self.lf_c = 1e-12
        self.flags = 7
        self.omega_relax, self.n_newton, self.relaxation = 1.0, 3, 1.0
        self.rtol, self.atol = 1e-5, 1e-2

    def params(self, dt):
        return EnthParams(self.dx, float(dt), self.drain_rate, self.h_thin, self.lf_c, int(self.flags))

    def T_pmp(self):
        depth = (1.0 - self.sigma)[None, None, :] * self.H[:, :, None]
        return T_MELT - BETA_CC * RHO_I_E * GRAVITY * depth

    def initialize_from_temperature(self, T):
        T = np.broadcast_to(np.asarray(T, dtype=np.float64) if np.ndim(T) != 2 else np.asarray(T)[:, :, None],
                            (self.ny, self.nx, self.nz))
        T = np.minimum(T, self.T_pmp())
        self.E = np.ascontiguousarray(C_I * (T - T_REF) / E_SCALE)
        self.E_prev = self.E.copy()
        self.H_prev = self.H.copy()

    def set_surface_enthalpy_from_temperature(self, Ts):
        Ts = np.minimum(np.broadcast_to(np.asarray(Ts, dtype=np.float64), (self.ny, self.nx)), T_MELT)
        self.E_surface = np.ascontiguousarray(C_I * (Ts - T_REF) / E_SCALE)

    def set_rhs(self, snapshot=True):
        if snapshot:
            self.E_prev = self.E.copy()
            self.H_prev = self.H.copy()
        dsig_half = 0.5 / (self.nz - 1)
        f = self.H[:, :, None] * self.phi_strain / E_SCALE
        f[:, :, 0] += (self.Q_geo + self.Q_fh) / (dsig_half * E_SCALE)
        f[:, :, self.nz - 1] = 0.0
        self.f_E = np.ascontiguousarray(f)

    def _args(self):
        return (_p(self.E_prev), _p(self.f_E), _p(self.u3d), _p(self.v3d), _p(self.omega), _p(self.H),
                _p(self.H_prev), _p(self.E_surface))

    def compute_residual(self, dt, use_forcing=True):
        r = np.zeros_like(self.E)
        P = self.params(dt)
        mx = _enth_lib().orc_enth_residual(_p(r), _p(self.E), *self._args(), ctypes.byref(P), int(use_forcing),
                                           self.ny, self.nx, self.nz)
        self.r_E = r
        return mx

    def column_smooth(self, dt):
        d = np.zeros_like(self.E)
        P = self.params(dt)
        _enth_lib().orc_enth_column_smooth(_p(d), _p(self.E), *self._args(), ctypes.byref(P), self.n_newton,
                                           _D(self.relaxation), self.ny, self.nx, self.nz)
        self.delta_E = d
        return d

    def layer_smooth(self, dt):
        d = np.zeros_like(self.E)
        P = self.params(dt)
        _enth_lib().orc_enth_layer_smooth(_p(d), _p(self.E), *self._args(), ctypes.byref(P), self.ny, self.nx,
                                          self.nz)
        self.delta_E = d
        return d

    def column_sweep(self, dt, n_iter, alternating=True):
        P = self.params(dt)
        r0, r = _D(0.0), _D(0.0)
        it = _enth_lib().orc_enth_column_sweep(_p(self.E), *self._args(), ctypes.byref(P), int(n_iter),
                                               int(alternating), _D(self.omega_relax), self.n_newton,
                                               _D(self.relaxation), _D(self.rtol), _D(self.atol), ctypes.byref(r0),
                                               ctypes.byref(r), self.ny, self.nx, self.nz)
        return it, r0.value, r.value

    def compute_omega(self, dh_dt, bmb=None):
        bmb = np.zeros((self.ny, self.nx)) if bmb is None else np.ascontiguousarray(bmb, dtype=np.float64)
        dh = np.ascontiguousarray(np.broadcast_to(np.asarray(dh_dt, dtype=np.float64), (self.ny, self.nx)))
        om = np.zeros_like(self.E)
        _enth_lib().orc_enth_omega(_p(om), _p(self.u3d), _p(self.v3d), _p(self.H), _p(dh), _p(bmb), _D(self.dx),
                                   self.ny, self.nx, self.nz)
        self.omega = om
        return om

    def get_temperature(self):
        return np.minimum(self.E * E_SCALE / C_I + T_REF, self.T_pmp())

    def get_water_content(self):
        E_pmp = C_I * (self.T_pmp() - T_REF)
        return np.maximum(self.E * E_SCALE - E_pmp, 0.0) / L_HEAT

    # ---------------------------------------------------------- adjoint pieces
    def system_residual(self, dt):
        """G(E): the residual on interior nodes, E - E_surface on the surface
        node, E - min(E_surface, E_pmp) on thin columns."""
        out = np.zeros_like(self.E)
        P = self.params(dt)
        _enth_lib().orc_enth_system_residual(_p(out), _p(self.E), *self._args(), ctypes.byref(P), self.ny, self.nx,
                                             self.nz)
        return out

    def jacobian_rows(self, dt):
        """Exact dG/dE rows (a, b, c, xm, xp, ym, yp) per node."""
        rows = np.zeros((self.ny, self.nx, self.nz, 7))
        P = self.params(dt)
        _enth_lib().orc_enth_jacobian_rows(_p(rows), _p(self.E), _p(self.u3d), _p(self.v3d), _p(self.omega),
                                           _p(self.H), _p(self.H_prev), ctypes.byref(P), self.ny, self.nx, self.nz)
        return rows

    def jacobian(self, dt):
        """dG/dE as a scipy CSR matrix over the (ny, nx, nz) node order."""
        import scipy.sparse as sp
        ny, nx, nz = self.ny, self.nx, self.nz
        rows = self.jacobian_rows(dt)
        idx = np.arange(ny * nx * nz).reshape(ny, nx, nz)
        ii, jj, kk = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
        targets = [(ii, jj, kk - 1), (ii, jj, kk), (ii, jj, kk + 1), (ii, jj - 1, kk), (ii, jj + 1, kk),
                   (ii - 1, jj, kk), (ii + 1, jj, kk)]
        R, C, V = [], [], []
        for q, (ti, tj, tk) in enumerate(targets):
            val = rows[..., q]
            ok = (val != 0.0) & (ti >= 0) & (ti < ny) & (tj >= 0) & (tj < nx) & (tk >= 0) & (tk < nz)
            R.append(idx[ok]); C.append(idx[ti[ok], tj[ok], tk[ok]]); V.append(val[ok])
        n = ny * nx * nz
        return sp.csr_matrix((np.concatenate(V), (np.concatenate(R), np.concatenate(C))), shape=(n, n))

    def adjoint_solve(self, dt, g):
        """lambda with J^T lambda = g (direct sparse solve)."""
        import scipy.sparse.linalg as spla
        J = self.jacobian(dt)
        lam = spla.spsolve(J.T.tocsc(), np.ascontiguousarray(g).ravel())
        return lam.reshape(self.E.shape)

    def adjoint_gradients(self, dt, lam):
        """dL/dp = -lambda^T dG/dp for E_prev, Q_geo, Q_fh, phi_strain, E_surface."""
        nz = self.nz
        dsig = 1.0 / (nz - 1)
        thin = self.H < self.h_thin
        live = ~thin[:, :, None] & (np.arange(nz) < nz - 1)[None, None, :]
        g = {}
        g["E_prev"] = np.where(live, lam * RHO_I_E * self.H_prev[:, :, None] / dt, 0.0)
        g["Q_geo"] = np.where(~thin, lam[:, :, 0] / (0.5 * dsig * E_SCALE), 0.0)
        g["Q_fh"] = g["Q_geo"].copy()
        g["phi_strain"] = np.where(live, lam * self.H[:, :, None] / E_SCALE, 0.0)
        sig = np.linspace(0.0, 1.0, nz)
        Epmp = C_I * (T_MELT - BETA_CC * RHO_I_E * GRAVITY * (1.0 - sig)[None, None, :] * self.H[:, :, None]
                      - T_REF) / E_SCALE
        below = self.E_surface[:, :, None] < Epmp
        g["E_surface"] = np.where(thin, np.sum(np.where(below, lam, 0.0), axis=2), lam[:, :, nz - 1])
        return g

    def arrhenius_vjp(self, gB, n_glen=3.0, scale=1.0):
        gE = np.zeros_like(self.E)
        gB = np.ascontiguousarray(gB, dtype=np.float64)
        _enth_lib().orc_enth_arrhenius_vjp(_p(gE), _p(gB), _p(self.E), _p(self.H), _D(n_glen), _D(scale), self.ny,
                                           self.nx, self.nz)
        return gE

    def get_arrhenius_factor(self, n_glen=3.0):
        B = np.zeros((self.ny, self.nx))
        _enth_lib().orc_enth_arrhenius(_p(B), _p(self.E), _p(self.H), _D(n_glen), self.ny, self.nx, self.nz)
        return B


def thermal_step(th: OracleThermal, u, v, H_new, beta, dt, n_smooth=10, m=1.0, u_reg=1.0, update_rheology=True,
                 frictional_heating=True, rho_i=917.0, n_glen=3.0, strain_heating=None):
    """ThermalModel.step (glide/model.py:167-254) after pre_momentum (E_prev,
    H_prev snapshot) and the momentum solve that produced u, v (m/a) and H_new.
    Returns B_glide when update_rheology else None."""
    th.H = np.ascontiguousarray(H_new, dtype=np.float64).copy()
    for k in range(th.nz):
        th.u3d[k] = u
        th.v3d[k] = v
    dt_yr = dt / SEC_PER_YR
    dh_dt_yr = (th.H - th.H_prev) / dt_yr
    th.compute_omega(dh_dt_yr)
    th.omega /= SEC_PER_YR
    th.u3d /= SEC_PER_YR
    th.v3d /= SEC_PER_YR
    if frictional_heating:
        ub, vb = th.u3d[0], th.v3d[0]
        uc = 0.5 * (ub[:, 1:] + ub[:, :-1])
        vc = 0.5 * (vb[1:, :] + vb[:-1, :])
        speed = np.sqrt(uc ** 2 + vc ** 2 + u_reg ** 2)
        th.Q_fh = np.ascontiguousarray(beta * speed ** (m + 1.0))
    if strain_heating is not None:
        th.phi_strain = np.ascontiguousarray(strain_heating)
    th.set_rhs(snapshot=False)
    th.column_sweep(dt, n_smooth)
    if update_rheology:
        scale = rho_i * GRAVITY * SEC_PER_YR ** (1.0 / n_glen)
        return th.get_arrhenius_factor(n_glen) / scale
    return None
