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
                ("relative_tolerance", _D), ("absolute_tolerance", _D), ("compat_hfacet", _I)]


class VankaConfig(ctypes.Structure):
    _fields_ = [("omega", _D), ("relax_phi", _D), ("newton_steps", _I), ("ssa_damping", _D), ("mc_damping", _D)]


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


def viscosity(u, v, B, n, eps_reg, dx, du=None, dv=None):
    ny, nx = B.shape
    eta, etad = np.zeros((ny, nx)), np.zeros((ny, nx))
    lib().orc_viscosity(_p(eta), _p(etad) if du is not None else None, _p(u), _p(v), _p(du), _p(dv), _p(B), _D(n),
                        _D(eps_reg), _D(dx), ny, nx)
    return (eta, etad) if du is not None else eta


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

    def vanka_configs(self, omega=0.5, relax_phi=0.0, newton_steps=30, ssa_damping=0.01, mc_damping=1.0):
        arr = (VankaConfig * self.n_levels)()
        for k in range(self.n_levels):
            arr[k] = VankaConfig(omega, relax_phi, newton_steps, ssa_damping, mc_damping)
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
