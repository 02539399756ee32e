"""ctypes binding of ``libglide_hip.so`` (declared in ``include/glide_hip.h`` and
``include/glide_enthalpy.h``).

The shared library is loaded on first use.  There is no fallback: when the
library is missing, or a call is made with tensors that are not on a ROCm
device, a :class:`GlideError` is raised.
"""
from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

import torch

# GLIDE_LIB names another build of the library (A/B measurements)
LIB_PATH = Path(os.environ.get("GLIDE_LIB") or Path(__file__).resolve().parent / "libglide_hip.so")

USE_FORCING = 1
USE_MASK = 2
FREEZE_CALVING = 4
FREEZE_PHI = 8


class GlideError(RuntimeError):
    """Raised when a native entry point reports a failure."""


class Physics(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "n", "eps_reg", "flotation_reg_driving", "m", "u_reg", "water_drag", "flotation_reg_sliding",
        "calving_rate", "flotation_reg_calving", "thklim")]


LEVEL_POINTERS = (
    "u", "v", "H", "H_prev", "phi", "mask",
    "bed", "B", "beta", "smb",
    "f_u", "f_v", "f_H", "r_u", "r_v", "r_H", "F_u", "F_v", "F_H", "gamma",
    "w_u", "w_v", "w_H", "y_u", "y_v", "y_H", "z_u", "z_v", "z_H", "chi", "cphi",
    "lambda_u", "lambda_v", "lambda_H",
    "af_u", "af_v", "af_H", "ar_u", "ar_v", "ar_H", "vjp_u", "vjp_v", "vjp_H",
    "wl_u", "wl_v", "wl_H", "zl_u", "zl_v", "zl_H",
    "g_beta", "g_bed", "g_H_prev", "g_smb",
    "work", "adj_cache",
)


class Level(ctypes.Structure):
    _fields_ = ([("ny", ctypes.c_int32), ("nx", ctypes.c_int32), ("dx", ctypes.c_double), ("phys", Physics)]
                + [(n, ctypes.c_void_p) for n in LEVEL_POINTERS])


class VankaConfigC(ctypes.Structure):
    _fields_ = [("omega", ctypes.c_double), ("relax_phi", ctypes.c_double), ("newton_steps", ctypes.c_int32),
                ("relaxation", ctypes.c_double), ("ssa_damping", ctypes.c_double), ("mc_damping", ctypes.c_double),
                ("local_fp32", ctypes.c_int32), ("multicolour", ctypes.c_int32), ("newton_tol", ctypes.c_double)]


class FasConfigC(ctypes.Structure):
    _fields_ = [("freeze_coarse_calving", ctypes.c_int32), ("freeze_coarse_phi", ctypes.c_int32),
                ("coarsest_steps", ctypes.c_int32), ("pre_steps", ctypes.c_int32), ("post_steps", ctypes.c_int32),
                ("finest_steps", ctypes.c_int32), ("maximum_vcycles", ctypes.c_int32),
                ("relative_tolerance", ctypes.c_double), ("absolute_tolerance", ctypes.c_double),
                ("compat_vprolong", ctypes.c_int32)]


class MixedConfigC(ctypes.Structure):
    _fields_ = [("max_outer", ctypes.c_int32), ("inner_vcycles", ctypes.c_int32),
                ("relative_tolerance", ctypes.c_double), ("absolute_tolerance", ctypes.c_double),
                ("fp64_coarsest", ctypes.c_int32), ("krylov_dim", ctypes.c_int32), ("krylov_work", ctypes.c_void_p),
                ("fp64_newton_steps", ctypes.c_int32), ("fp64_relaxation", ctypes.c_double)]


class SolveReport(ctypes.Structure):
    _fields_ = [("vcycles", ctypes.c_int32), ("initial_norm", ctypes.c_double), ("final_norm", ctypes.c_double),
                ("norms", ctypes.c_double * 64), ("norms3", ctypes.c_double * 3)]


ENTH_POINTERS = ("E", "E_prev", "f_E", "omega", "phi_strain", "r_E", "delta_E", "E_work", "u3d", "v3d",
                 "H", "H_prev", "E_surface", "Q_geo", "Q_fh", "dh_dt", "sweep_work")
ENTH_MAX_SWEEPS = 256


class EnthColumn(ctypes.Structure):
    _fields_ = [("ny", ctypes.c_int32), ("nx", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("term_flags", ctypes.c_int32), ("dx", ctypes.c_double), ("drain_rate", ctypes.c_double),
                ("h_thin", ctypes.c_double), ("lf_c", ctypes.c_double)] + [(n, ctypes.c_void_p)
                                                                           for n in ENTH_POINTERS]


class EnthSweepConfigC(ctypes.Structure):
    _fields_ = [("omega", ctypes.c_double), ("relaxation", ctypes.c_double),
                ("relative_tolerance", ctypes.c_double), ("absolute_tolerance", ctypes.c_double),
                ("n_newton", ctypes.c_int32), ("alternating", ctypes.c_int32), ("multicolour", ctypes.c_int32)]


class EnthReport(ctypes.Structure):
    _fields_ = [("sweeps", ctypes.c_int32), ("pad", ctypes.c_int32), ("r0", ctypes.c_double),
                ("r", ctypes.c_double), ("norms", ctypes.c_double * ENTH_MAX_SWEEPS)]


ENTH_ADJOINT_POINTERS = ("rows", "g", "lambda_", "lambda_work", "r", "g_E_prev", "g_phi_strain", "g_Q_geo",
                         "g_E_surface", "u2d", "v2d")


class EnthAdjointC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ENTH_ADJOINT_POINTERS] + [("vel_div", ctypes.c_double)]


class ThermalConfigC(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("m", ctypes.c_double), ("u_reg", ctypes.c_double),
                ("rho_i", ctypes.c_double), ("n_glen", ctypes.c_double), ("eps_reg", ctypes.c_double),
                ("n_smooth", ctypes.c_int32), ("frictional_heating", ctypes.c_int32),
                ("update_rheology", ctypes.c_int32), ("strain_heating", ctypes.c_int32)]


THERMAL_VJP_POINTERS = ("lambda_", "g_phi_strain", "g_Q", "dL_dB", "work", "g_u", "g_v", "g_H", "g_H_prev", "g_beta",
                        "g_B")


class ThermalVjpC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in THERMAL_VJP_POINTERS]


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_S = ctypes.c_size_t
_LP = ctypes.POINTER(Level)
_CP = ctypes.POINTER(EnthColumn)

# name -> (restype, argtypes); every symbol the header declares.
SIGNATURES = {
    "glide_version": (ctypes.c_char_p, []),
    "glide_init": (_I, []),
    "glide_check_status": (_I, [_P]),
    "glide_persistent_spin_limit": (_I, [_I]),
    "glide_last_error": (_I, [ctypes.c_char_p, _S]),
    "glide_compute_grounded": (_I, [_LP, _D, _P]),
    "glide_ssa_residual": (_I, [_LP, _D, _I, _P, _P, _P, _P]),
    "glide_ssa_jvp": (_I, [_LP, _D, _I, _P, _P, _P, _P, _P, _P, _P]),
    "glide_norms3": (_I, [_P, _S, _P, _S, _P, _S, ctypes.POINTER(_D), _P]),
    "glide_vanka_sweeps": (_I, [_LP, _D, _I, ctypes.POINTER(VankaConfigC), _I, _P]),
    "glide_vanka_dump": (_I, [_LP, _D, _P, _P, _P]),
    "glide_restrict_vfacet": (_I, [_P, _P, _I, _I, _P]),
    "glide_restrict_hfacet": (_I, [_P, _P, _I, _I, _P]),
    "glide_restrict_cell": (_I, [_P, _P, _I, _I, _I, _P]),
    "glide_prolong_vfacet_bilinear": (_I, [_P, _P, _I, _I, _P]),
    "glide_prolong_hfacet_bilinear": (_I, [_P, _P, _I, _I, _P]),
    "glide_prolong_cell": (_I, [_P, _P, _I, _I, _I, _P]),
    "glide_fascd_solve": (_I, [_LP, _I, _I, _D, ctypes.POINTER(FasConfigC), ctypes.POINTER(VankaConfigC),
                               ctypes.POINTER(SolveReport), _P]),
    "glide_fascd_vcycle": (_I, [_LP, _I, _I, _I, _D, ctypes.POINTER(FasConfigC), ctypes.POINTER(VankaConfigC), _P]),
    "glide_fascd_solve_mixed": (_I, [_LP, _LP, _I, _I, _D, ctypes.POINTER(FasConfigC), ctypes.POINTER(VankaConfigC),
                                     ctypes.POINTER(MixedConfigC), ctypes.POINTER(SolveReport),
                                     ctypes.POINTER(_I), _P]),
    "glide_fas_adjoint_solve_mixed": (_I, [_LP, _LP, _I, _I, _D, ctypes.POINTER(FasConfigC),
                                           ctypes.POINTER(VankaConfigC), ctypes.POINTER(MixedConfigC),
                                           ctypes.POINTER(SolveReport), ctypes.POINTER(_I), _P]),
    "glide_ensemble_create": (_I, [ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(_LP),
                                   ctypes.POINTER(_LP), ctypes.POINTER(FasConfigC), ctypes.POINTER(FasConfigC),
                                   ctypes.POINTER(VankaConfigC),
                                   ctypes.POINTER(VankaConfigC), ctypes.POINTER(MixedConfigC), ctypes.c_int32,
                                   ctypes.POINTER(_P)]),
    "glide_ensemble_step": (_I, [_P, _D, _P]),
    "glide_ensemble_launches": (ctypes.c_int32, [_P]),
    "glide_ensemble_levels": (ctypes.c_int32, [_P]),
    "glide_ensemble_set_tiles": (_I, [_P, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]),
    "glide_ensemble_destroy": (_I, [_P]),
    "glide_ssa_vjp": (_I, [_LP, _D, _I, _P, _P, _P, _P, _P, _P, _P]),
    "glide_vanka_adjoint_sweeps": (_I, [_LP, _D, _I, ctypes.POINTER(VankaConfigC), _I, _P]),
    "glide_fas_adjoint_solve": (_I, [_LP, _I, _I, _D, ctypes.POINTER(FasConfigC), ctypes.POINTER(VankaConfigC),
                                     ctypes.POINTER(SolveReport), _P]),
    "glide_fas_adjoint_vcycle": (_I, [_LP, _I, _I, _I, _D, ctypes.POINTER(FasConfigC),
                                      ctypes.POINTER(VankaConfigC), _P]),
    "glide_gradients": (_I, [_LP, _D, _I, _P]),
    "glide_gradient_rate_factor": (_I, [_LP, _D, _P, _P]),
    "glide_krylov_work_doubles": (ctypes.c_int64, [_I, _I, _I]),
    "glide_graphs_enable": (_I, [_I]),
    "glide_graphs_capacity": (_I, [_I]),
    "glide_adjoint_fused_enable": (_I, [_I]),
    "glide_forward_fused_enable": (_I, [_I]),
    "glide_forward_persistent_enable": (_I, [_I]),
    "glide_adjoint_coarse_enable": (_I, [_I]),
    "glide_adjoint_pass_fused_enable": (_I, [_I]),
    "glide_adjoint_persistent_enable": (_I, [_I]),
    "glide_persistent_fenced_enable": (_I, [_I]),
    "glide_graphs_clear": (_I, []),
    "glide_timing_enable": (_I, [_I]),
    "glide_timing_read": (_I, [_I, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "glide_timing_read_units": (_I, [_I, ctypes.c_long, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(_D)]),
    # include/glide_enthalpy.h
    "glide_enth_init": (_I, []),
    "glide_enth_residual": (_I, [_CP, _D, _I, ctypes.POINTER(_D), _P]),
    "glide_enth_column_smooth": (_I, [_CP, _D, _I, _D, _P]),
    "glide_enth_layer_smooth": (_I, [_CP, _D, _P]),
    "glide_enth_apply": (_I, [_CP, _D, _P]),
    "glide_enth_column_sweep": (_I, [_CP, _D, _I, ctypes.POINTER(EnthSweepConfigC), ctypes.POINTER(EnthReport),
                                     _P]),
    "glide_enth_column_sweep_noalloc": (_I, [_CP, _D, _I, ctypes.POINTER(EnthSweepConfigC),
                                             ctypes.POINTER(EnthReport), _P]),
    "glide_enth_elementary": (_I, [_P, ctypes.c_int64, _P, _P, _P]),
    "glide_enth_omega": (_I, [_CP, _P, _P, _D, _P]),
    "glide_enth_broadcast_velocity": (_I, [_CP, _P, _P, _D, _P]),
    "glide_enth_frictional_heating": (_I, [_CP, _P, _D, _D, _P]),
    "glide_enth_strain_heating": (_I, [_CP, _P, _P, _P, _D, _D, _D, _P]),
    "glide_enth_set_rhs": (_I, [_CP, _P]),
    "glide_enth_arrhenius": (_I, [_CP, _D, _D, _P, _P]),
    "glide_enth_diagnostics": (_I, [_CP, _P, _P, _P]),
    "glide_enth_from_temperature": (_I, [_CP, _P, _P, _P]),
    "glide_enth_surface_from_temperature": (_I, [_CP, _P, _P, _P]),
    "glide_enth_adjoint_solve": (_I, [_CP, ctypes.POINTER(EnthAdjointC), _D, _P, _P, _D, _D, _I,
                                      ctypes.POINTER(EnthSweepConfigC), ctypes.POINTER(EnthReport), _P]),
    "glide_thermal_step": (_I, [_CP, _P, _P, _P, _P, ctypes.POINTER(ThermalConfigC),
                                ctypes.POINTER(EnthSweepConfigC), ctypes.POINTER(EnthReport), _P]),
    "glide_thermal_vjp_work_doubles": (ctypes.c_int64, [_I, _I]),
    "glide_thermal_step_vjp": (_I, [_CP, _P, _P, _P, _P, ctypes.POINTER(ThermalConfigC), ctypes.POINTER(ThermalVjpC),
                                    _P]),
}

TIMED_KERNELS = {"vanka_cell": 0, "vanka_apply": 1, "adjoint_cell": 2, "vjp": 3, "residual": 4, "adjoint_apply": 5,
                 "enth_column": 6, "enth_layer": 7, "enth_residual": 8, "adjoint_factor": 9, "vanka_mc32": 10,
                 "vanka_mc64": 11, "adjoint_mc32": 12, "adjoint_mc64": 13}


def graphs_enable(on: bool) -> None:
    """Replay fixed-cycle FAS solves from cached HIP graphs (default on)."""
    call("glide_graphs_enable", int(bool(on)))


def graphs_capacity(n: int) -> None:
    """Number of cached graphs kept per device (default 8)."""
    call("glide_graphs_capacity", int(n))


def adjoint_fused_enable(on: bool) -> None:
    """Cached adjoint passes as one fused launch (default) or solve + apply."""
    call("glide_adjoint_fused_enable", int(bool(on)))


def forward_fused_enable(on: bool) -> None:
    """Forward passes of small levels as one fused launch (default) or solve + apply."""
    call("glide_forward_fused_enable", int(bool(on)))


def forward_persistent_enable(on: bool) -> None:
    """Forward passes of levels at most 64 wide, 96 high as one persistent launch per
    pass sequence (default) or one launch per pass."""
    call("glide_forward_persistent_enable", int(bool(on)))


def adjoint_coarse_enable(on: bool) -> None:
    """Cached adjoint passes of levels of at most 512 cells (linear drag) in
    one workgroup (default) or one launch pair per pass."""
    call("glide_adjoint_coarse_enable", int(bool(on)))


def adjoint_pass_fused_enable(on: bool) -> None:
    """Cached adjoint passes of levels up to 512 x 512 cells (linear drag) as one
    launch each, transpose operator included (default), or two launches."""
    call("glide_adjoint_pass_fused_enable", int(bool(on)))


def adjoint_persistent_enable(on: bool) -> None:
    """Cached adjoint passes of levels at most 64 wide, 32 high (linear drag) as one
    persistent launch per call (default) or one fused launch per pass."""
    call("glide_adjoint_persistent_enable", int(bool(on)))


def persistent_fenced_enable(on: bool) -> None:
    """Persistent launches with the cross-XCD cache maintenance at every
    barrier even when their workgroups share one XCD (testing switch)."""
    call("glide_persistent_fenced_enable", int(bool(on)))


_initialised_devices: set = set()


def init_device(device: torch.device | None = None) -> None:
    """Allocate the library's per-device scratch once per device (glide_init),
    so that later calls allocate nothing, e.g. inside a caller's stream capture."""
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    if idx in _initialised_devices:
        return
    with torch.cuda.device(idx):
        call("glide_init")
    _initialised_devices.add(idx)


def check_status(device: torch.device | None = None) -> None:
    """Synchronise torch's current stream and raise if a persistent launch's
    barrier timed out since the last check (glide_check_status)."""
    call("glide_check_status", stream_handle(device))


def persistent_spin_limit(polls: int) -> None:
    """Polls of a persistent barrier wait before it gives up (testing switch;
    negative gives up at once)."""
    call("glide_persistent_spin_limit", int(polls))


def graphs_clear() -> None:
    call("glide_graphs_clear")


def timing_enable(on: bool) -> None:
    call("glide_timing_enable", int(bool(on)))


def timing_read_units(kernel: str, units: int):
    """(launch count, summed ms) of a timed kernel's launches over exactly `units` cells."""
    c, ms = ctypes.c_long(0), ctypes.c_double(0.0)
    call("glide_timing_read_units", TIMED_KERNELS[kernel], int(units), ctypes.byref(c), ctypes.byref(ms))
    return c.value, ms.value


def timing_read(kernel: str):
    """(launch count, summed ms, summed cells) of a timed kernel since enable."""
    c, ms, cells = ctypes.c_long(0), ctypes.c_double(0.0), ctypes.c_double(0.0)
    call("glide_timing_read", TIMED_KERNELS[kernel], ctypes.byref(c), ctypes.byref(ms), ctypes.byref(cells))
    return c.value, ms.value, cells.value

_lock = threading.Lock()
_lib = None


def load(path: Path | str | None = None):
    """Load (once) and return the native library with typed signatures."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path is not None else LIB_PATH
        if not p.exists():
            raise GlideError(f"native library not built: {p} (run glide_amd.build.build())")
        lib = ctypes.CDLL(str(p))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def last_error() -> str:
    lib = load()
    buf = ctypes.create_string_buffer(1024)
    lib.glide_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int, what: str) -> None:
    if status != 0:
        raise GlideError(f"{what} failed ({status}): {last_error()}")


def call(name: str, *args) -> None:
    """Call a native entry point and raise on a non-zero status."""
    fn = getattr(load(), name)
    check(fn(*args), name)


def ptr(t: torch.Tensor | None, dtype: torch.dtype = torch.float64) -> int | None:
    """Device address of a contiguous ROCm tensor of `dtype` (None stays NULL);
    float64 everywhere except the FP32 buffers of the mixed-precision solves."""
    if t is None:
        return None
    if not t.is_cuda:
        raise GlideError("glide_amd kernels need tensors on a ROCm device (got a CPU tensor)")
    if not t.is_contiguous():
        raise GlideError("glide_amd kernels need contiguous tensors")
    if t.dtype != dtype:
        raise GlideError(f"this glide_amd buffer must be {dtype} (got {t.dtype})")
    return t.data_ptr()


def ptr32(t: torch.Tensor | None) -> int | None:
    return ptr(t, torch.float32)


def stream_handle(device: torch.device | None = None) -> int:
    """hipStream_t of torch's current stream on `device`."""
    return torch.cuda.current_stream(device).cuda_stream
