"""C ABI and host-side logic (CPU only; no kernel is launched)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest
import torch

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    syms = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        syms.update(re.findall(r"\b(glide_[a-z0-9_]+)\s*\(", text))
    return sorted(syms)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(ROOT / "glide_amd" / "libglide_hip.so"))
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    from glide_amd import _lib
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_struct_layouts_match_header():
    from glide_amd import _lib
    # glide_level: 2 int32 + dx + 10-double physics + pointers, 8-byte aligned
    n_ptr = len(_lib.LEVEL_POINTERS)
    assert ctypes.sizeof(_lib.Level) == 8 + 8 + 80 + 8 * n_ptr
    header = (ROOT / "include" / "glide_hip.h").read_text()
    body = header[header.index("typedef struct glide_level"):header.index("} glide_level;")]
    ptrs = re.findall(r"\*(\w+)", body)
    assert tuple(ptrs) == _lib.LEVEL_POINTERS
    assert ctypes.sizeof(_lib.SolveReport) == 8 + 8 + 8 + 64 * 8 + 3 * 8
    eh = (ROOT / "include" / "glide_enthalpy.h").read_text()
    body = eh[eh.index("typedef struct glide_enth_column"):eh.index("} glide_enth_column;")]
    assert tuple(re.findall(r"\*(\w+)", body)) == _lib.ENTH_POINTERS
    assert ctypes.sizeof(_lib.EnthColumn) == 16 + 32 + 8 * len(_lib.ENTH_POINTERS)
    assert ctypes.sizeof(_lib.EnthSweepConfigC) == 4 * 8 + 4 * 4  # three int32 fields, padded to 8
    assert ctypes.sizeof(_lib.EnthReport) == 8 + 16 + 8 * _lib.ENTH_MAX_SWEEPS
    assert ctypes.sizeof(_lib.ThermalConfigC) == 6 * 8 + 4 * 4
    assert ctypes.sizeof(_lib.EnthAdjointC) == 11 * 8 + 8
    body = eh[eh.index("typedef struct glide_enth_adjoint"):eh.index("} glide_enth_adjoint;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    # lambda is a Python keyword: the ctypes field is lambda_
    assert tuple(re.findall(r"\*\s*(\w+)\s*[;,]", body)) == tuple(
        f.rstrip("_") for f in _lib.ENTH_ADJOINT_POINTERS)
    assert "double vel_div" in body
    assert f"GLIDE_ENTH_MAX_SWEEPS {_lib.ENTH_MAX_SWEEPS}" in eh
    body = eh[eh.index("typedef struct glide_thermal_vjp"):eh.index("} glide_thermal_vjp;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    assert tuple(re.findall(r"\*\s*(\w+)\s*[;,]", body)) == tuple(
        f.rstrip("_") for f in _lib.THERMAL_VJP_POINTERS)
    assert ctypes.sizeof(_lib.ThermalVjpC) == 8 * len(_lib.THERMAL_VJP_POINTERS)


def test_version_and_error_text():
    from glide_amd import _lib
    lib = _lib.load()
    assert b"gfx950" in lib.glide_version()
    # invalid arguments are rejected on the host before any device work
    st = lib.glide_ssa_residual(None, 1.0, 0, None, None, None, None)
    assert st == -1
    assert "NULL" in _lib.last_error()
    assert lib.glide_ssa_jvp(None, 1.0, 0, None, None, None, None, None, None, None) == -1
    L = _lib.Level()
    L.ny, L.nx, L.dx = 1, 8, 100.0
    assert lib.glide_compute_grounded(ctypes.byref(L), 0.0, None) == -2
    assert "ny >= 2" in _lib.last_error()


def test_hierarchy_checks_on_host():
    from glide_amd import _lib
    lib = _lib.load()
    levels = (_lib.Level * 2)()
    cfg = _lib.FasConfigC(1, 1, 10, 1, 1, 1, 3, 0.0, 0.0, 1)
    vc = (_lib.VankaConfigC * 2)()
    # buffers missing -> invalid argument, nothing launched
    levels[0].ny, levels[0].nx, levels[0].dx = 8, 8, 1.0
    levels[1].ny, levels[1].nx, levels[1].dx = 4, 4, 2.0
    assert lib.glide_fascd_solve(levels, 2, 0, 1.0, ctypes.byref(cfg), vc, None, None) == -1
    cfg.maximum_vcycles = 65
    assert lib.glide_fascd_solve(levels, 2, 0, 1.0, ctypes.byref(cfg), vc, None, None) == -1
    assert "maximum_vcycles" in _lib.last_error()
    # a negative count is rejected too (it would size the norm read-back)
    cfg.maximum_vcycles = -2
    for fn in (lib.glide_fascd_solve, lib.glide_fas_adjoint_solve):
        assert fn(levels, 2, 0, 1.0, ctypes.byref(cfg), vc, None, None) == -1
        assert "maximum_vcycles" in _lib.last_error()


def test_python_configs_reject_bad_vcycle_counts():
    from glide_amd.multigrid import FASAdjointConfig, FASCDConfig
    for cls in (FASCDConfig, FASAdjointConfig):
        for bad in (-1, 65):
            with pytest.raises(ValueError, match="maximum_vcycles"):
                cls(maximum_vcycles=bad).to_c()
        assert cls(maximum_vcycles=0).to_c().maximum_vcycles == 0


def test_status_codes_in_header():
    header = (ROOT / "include" / "glide_hip.h").read_text()
    assert "GLIDE_ERR_NONFINITE = -4" in header
    assert "GLIDE_ERR_TIMEOUT = -5" in header


def test_enthalpy_argument_checks_on_host():
    from glide_amd import _lib
    lib = _lib.load()
    C = _lib.EnthColumn(ny=4, nx=4, nz=65, term_flags=7, dx=100.0)
    assert lib.glide_enth_residual(ctypes.byref(C), 1.0, 1, None, None) == -2
    assert "nz" in _lib.last_error()
    C.nz = 11
    assert lib.glide_enth_residual(ctypes.byref(C), 1.0, 1, None, None) == -1
    assert "NULL" in _lib.last_error()
    C.term_flags = 8
    assert lib.glide_enth_set_rhs(ctypes.byref(C), None) == -1
    assert "flag" in _lib.last_error()
    C.term_flags = 7
    cfg = _lib.EnthSweepConfigC(1.0, 1.0, 1e-5, 1e-2, 3, 1)
    assert lib.glide_enth_column_sweep(ctypes.byref(C), 1.0, 257, ctypes.byref(cfg), None, None) == -1
    assert "n_iter" in _lib.last_error()
    assert lib.glide_thermal_step(ctypes.byref(C), None, None, None, None, None, None, None, None) == -1
    assert lib.glide_thermal_vjp_work_doubles(0, 4) == -1
    assert lib.glide_thermal_vjp_work_doubles(3, 5) == 14 * 15
    tc = _lib.ThermalConfigC(dt=1.0, m=1.0, u_reg=1.0, rho_i=917.0, n_glen=3.0, eps_reg=1e-6, n_smooth=10,
                             frictional_heating=1, update_rheology=1, strain_heating=1)
    V = _lib.ThermalVjpC()
    assert lib.glide_thermal_step_vjp(ctypes.byref(C), None, None, None, None, ctypes.byref(tc), None, None) == -1
    assert "NULL" in _lib.last_error()
    assert lib.glide_thermal_step_vjp(ctypes.byref(C), None, None, None, None, ctypes.byref(tc), ctypes.byref(V),
                                      None) == -1


def test_enthalpy_host_objects_on_cpu():
    from glide_amd import Grid, ThermalModel
    from glide_amd._lib import GlideError
    from glide_amd.enthalpy import ColumnSmootherConfig, EnthalpyTermFlags, make_sigma
    g = Grid(8, 8, 1000.0, device="cpu")
    tm = ThermalModel(g, nz=5)
    ops = tm.ops
    assert ops.enthalpy_state.E.shape == (8, 8, 5)
    assert ops.enthalpy_velocity.u3d.shape == (5, 8, 9)
    assert ops.enthalpy_velocity.v3d.shape == (5, 9, 8)
    assert torch.equal(make_sigma(5), torch.tensor([0.0, 0.25, 0.5, 0.75, 1.0], dtype=torch.float64))
    assert EnthalpyTermFlags().bitmask == 7
    assert EnthalpyTermFlags(horizontal_advection=False, drainage=False).bitmask == 2
    c = ColumnSmootherConfig().to_c(alternating=False)
    assert (c.n_newton, c.alternating, c.omega) == (3, 0, 1.0)
    assert tm.B_scale == pytest.approx(917.0 * 9.81 * (365.25 * 86400.0) ** (1.0 / 3.0))
    with pytest.raises(GlideError):
        ops.compute_residual(1.0)  # CPU tensors are refused, nothing runs on the host


def test_ptr_rejects_cpu_and_wrong_dtype():
    from glide_amd._lib import GlideError, ptr
    with pytest.raises(GlideError):
        ptr(torch.zeros(4, dtype=torch.float64))
    assert ptr(None) is None


def test_grid_layouts_and_defaults():
    from glide_amd.grid import Grid
    g = Grid(6, 10, 250.0, device="cpu")
    assert tuple(g.state.u.data.shape) == (6, 11)
    assert tuple(g.state.v.data.shape) == (7, 10)
    assert tuple(g.state.H.data.shape) == (6, 10)
    assert g.state.u.data.dtype == torch.float64
    assert g.geometry.thklim.value == 0.1 and g.rheology.n.value == 3.0 and g.sliding.m.value == 1.0
    assert g.sliding.water_drag.value == 1e-5 and g.calving.calving_rate.value == 0.0
    g.state.H.set(np.ones((6, 10)))
    assert float(g.state.H.data.sum()) == 60.0
    g.state.H.set(2.0)
    assert float(g.state.H.data.mean()) == 2.0
    with pytest.raises(ValueError):
        g.state.H.set(np.ones((5, 10)))
    assert not g.sliding.beta.has_grad()
    assert tuple(g.sliding.beta.grad.shape) == (6, 10)
    assert tuple(g.work.shape) == (5, 6, 10)


def test_multigrid_hierarchy_on_cpu():
    from glide_amd.multigrid import Multigrid
    mg = Multigrid(4, ny=32, nx=48, dx=100.0, device="cpu")
    assert [(l.ny, l.nx, l.dx) for l in mg.levels] == [(32, 48, 100.0), (16, 24, 200.0), (8, 12, 400.0),
                                                        (4, 6, 800.0)]
    mg.sliding.m.set(1.0 / 3.0)
    assert all(abs(l.sliding.m.value - 1.0 / 3.0) < 1e-15 for l in mg.levels)
    with pytest.raises(ValueError):
        Multigrid(4, ny=20, nx=20, dx=1.0, device="cpu")


def test_options_validation_and_defaults():
    from glide_amd.multigrid import FASCDConfig, FASCDOptions, validate_kwargs
    cfg = FASCDConfig()
    assert (cfg.coarsest_steps, cfg.pre_steps, cfg.post_steps, cfg.finest_steps) == (200, 10, 20, 50)
    assert (cfg.maximum_vcycles, cfg.relative_tolerance, cfg.absolute_tolerance) == (10, 1e-3, 5.0)
    opts = FASCDOptions(cfg)
    opts.set(pre_steps=3)
    assert cfg.pre_steps == 3 and opts.pre_steps.get() == 3
    opts.report_norms.set(False)
    assert opts.report_norms.get() is False  # the getter reads report_norms
    with pytest.raises(AttributeError):
        validate_kwargs(FASCDConfig, {"bogus": 1})
    c = cfg.to_c()
    assert c.pre_steps == 3 and c.compat_vprolong == 1


def test_vanka_options_broadcast():
    from glide_amd.multigrid import Multigrid, VankaOptions
    from glide_amd.operators import VankaConfig
    mg = Multigrid(3, ny=16, nx=16, dx=1.0, device="cpu")
    cfgs = {id(l): VankaConfig() for l in mg.levels}
    vo = VankaOptions(mg.levels, getter=lambda l: cfgs[id(l)])
    vo.omega.set(0.25)
    vo.newton_options.steps.set(7)
    vo.newton_options.set_level(1, ssa_damping=0.1)
    assert all(c.omega == 0.25 and c.newton_config.steps == 7 for c in cfgs.values())
    assert cfgs[id(mg.levels[1])].newton_config.ssa_damping == 0.1
    assert cfgs[id(mg.levels[0])].newton_config.ssa_damping == 0.01
    c = cfgs[id(mg.levels[1])].to_c()
    assert c.newton_steps == 7 and abs(c.omega - 0.25) < 1e-15


def test_icedynamics_constructor_forms():
    from glide_amd import IceDynamics
    from glide_amd.grid import Grid
    from glide_amd.multigrid import Multigrid
    m1 = IceDynamics(ny=16, nx=16, dx=10.0, n_levels=2, device="cpu")
    assert m1.mg.n_levels == 2
    m2 = IceDynamics(grid=Grid(8, 8, 1.0, device="cpu"), n_levels=2)
    assert m2.mg[1].ny == 4
    m3 = IceDynamics(mg=Multigrid(1, ny=4, nx=4, dx=1.0, device="cpu"))
    assert m3.mg.n_levels == 1
    with pytest.raises(ValueError):
        IceDynamics()


def test_synthetic_generators_are_seeded():
    from glide_amd.synthetic import dome_problem, gaussian_random_field
    a = dome_problem(64, 1500.0, seed=3)
    b = dome_problem(64, 1500.0, seed=3)
    assert np.array_equal(a.bed, b.bed) and np.array_equal(a.beta, b.beta)
    f = gaussian_random_field(np.random.default_rng(0), 128, 128, 1.0, 6.0)
    assert abs(f.mean()) < 1e-12 and abs(f.std() - 1.0) < 1e-12
    assert a.H.max() <= 2000.0 and a.H.min() == 0.0
    # dome radius 0.45 W: the corner cells carry no ice
    assert a.H[0, 0] == 0.0 and a.H[32, 32] > 1900.0


def test_ensemble_rejects_invalid_arguments():
    """glide_ensemble_create checks its arguments on the host (no device work):
    no domains, a missing level descriptor, a NULL level buffer."""
    from glide_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    fas, vc, mc = _lib.FasConfigC(), _lib.VankaConfigC(), _lib.MixedConfigC(max_outer=3, inner_vcycles=1)
    assert lib.glide_ensemble_create(0, None, None, None, ctypes.byref(fas), ctypes.byref(fas), ctypes.byref(vc), ctypes.byref(vc),
                                     ctypes.byref(mc), 15, ctypes.byref(h)) == -1
    assert not h.value
    nl = (ctypes.c_int32 * 1)(2)
    p = (ctypes.POINTER(_lib.Level) * 1)()
    assert lib.glide_ensemble_create(1, nl, p, p, ctypes.byref(fas), ctypes.byref(fas), ctypes.byref(vc), ctypes.byref(vc),
                                     ctypes.byref(mc), 15, ctypes.byref(h)) == -1
    assert "levels" in _lib.last_error()
    levels = (_lib.Level * 2)()
    for k, n in enumerate((16, 8)):
        levels[k].ny = levels[k].nx = n
        levels[k].dx = 100.0
    p[0] = ctypes.cast(levels, ctypes.POINTER(_lib.Level))
    assert lib.glide_ensemble_create(1, nl, p, p, ctypes.byref(fas), ctypes.byref(fas), ctypes.byref(vc), ctypes.byref(vc),
                                     ctypes.byref(mc), 15, ctypes.byref(h)) == -1
    assert "NULL" in _lib.last_error()
    mc.max_outer = -1
    assert lib.glide_ensemble_create(1, nl, p, p, ctypes.byref(fas), ctypes.byref(fas), ctypes.byref(vc), ctypes.byref(vc),
                                     ctypes.byref(mc), 15, ctypes.byref(h)) == -1
    assert lib.glide_ensemble_step(None, 1.0, None) == -1
    assert lib.glide_ensemble_launches(None) == -1
    assert lib.glide_ensemble_destroy(None) == 0


def test_dtype_argument_selects_the_solve_form():
    """IceDynamics(dtype=...): float64 -> the FP64 form, 'mixed' -> the mixed
    form; float32 is refused (not built)."""
    import torch
    from glide_amd import IceDynamics
    assert IceDynamics.precision_from_dtype(torch.float64) == "fp64"
    assert IceDynamics.precision_from_dtype("float64") == "fp64"
    assert IceDynamics.precision_from_dtype("mixed") == "mixed"
    m = IceDynamics(ny=16, nx=16, dx=10.0, n_levels=2, device="cpu", dtype="mixed")
    assert m.precision == "mixed"
    for bad in (torch.float32, "float32", "int8"):
        try:
            IceDynamics.precision_from_dtype(bad)
        except ValueError:
            pass
        else:
            raise AssertionError(bad)
