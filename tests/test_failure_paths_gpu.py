"""Failure reporting of the native path: non-finite residuals and persistent
barriers whose bounded wait gives up.

The reference lets NaNs propagate silently (CuPy norms, glide/operators.py:126,
glide/enthalpy.py:364).  Here a NaN reaches every reported norm, a sweep with a
NaN residual never passes its stopping test, and a synchronising call returns
GLIDE_ERR_NONFINITE (-4).  A persistent launch whose barrier gives up latches a
status word that the next synchronising call reports as GLIDE_ERR_TIMEOUT (-5).
The timeout is forced from the host (a negative spin limit gives up at the
first barrier without waiting), so every kernel still exits normally.
"""
import numpy as np
import pytest
import torch

from tests.util import enth_case

pytestmark = pytest.mark.gpu

DT = 10.0 * 365.25 * 86400.0


def _thermal_ops(nz=11):
    from tests.test_parity_enthalpy_gpu import make_pair
    th, ops = make_pair(enth_case(8, 9, nz, 5), nz)
    return ops


def test_nan_enthalpy_reaches_the_sweep_report():
    from glide_amd._lib import GlideError
    print("nan enthalpy sweep", flush=True)
    ops = _thermal_ops()
    ops.enthalpy_state.E[3, 4, 5] = float("nan")
    ops.smoother_config.report_norms = True
    n_iter = 6
    with pytest.raises(GlideError, match=r"\(-4\)"):
        ops.column_sweep(DT, n_iter)
    rep = ops.last_report
    assert np.isnan(rep.r0) and np.isnan(rep.r)
    assert rep.sweeps == n_iter  # a NaN norm never passes the stopping test


def test_nan_adjoint_seed_reaches_the_adjoint_report():
    from glide_amd._lib import GlideError
    print("nan enthalpy adjoint", flush=True)
    ops = _thermal_ops()
    ops.compute_residual(DT)
    seed = torch.zeros_like(ops.enthalpy_state.E)
    seed[2, 2, 3] = float("nan")
    with pytest.raises(GlideError, match=r"\(-4\)"):
        ops.adjoint_solve(DT, seed, n_iter=5, relative_tolerance=1e-12, report=True)


def _model(n=32, L=3, seed=0):
    from glide_amd import IceDynamics
    from glide_amd.synthetic import dome_problem, install
    p = dome_problem(n, 1500.0, seed=seed)
    model = IceDynamics(ny=n, nx=n, dx=1500.0, n_levels=L, device="cuda")
    install(model.mg, p)
    return model


def test_nan_state_fails_the_reported_solve():
    from glide_amd._lib import GlideError
    print("nan fascd solve", flush=True)
    model = _model()
    model.mg[0].state.H.data[5, 7] = float("nan")
    fs = model.forward_solver
    fs.fas_options.set(maximum_vcycles=2, relative_tolerance=1e-6, absolute_tolerance=0.0, report_norms=False)
    with pytest.raises(GlideError, match=r"\(-4\)"):
        fs.solve(10.0)


def test_forced_barrier_timeout_is_reported_and_cleared():
    from glide_amd import _lib
    from glide_amd._lib import GlideError
    print("forced persistent barrier timeout", flush=True)
    model = _model()
    fs = model.forward_solver
    fs.fas_options.set(maximum_vcycles=1, relative_tolerance=0.0, absolute_tolerance=0.0, report_norms=False)
    _lib.graphs_enable(False)
    try:
        _lib.check_status()  # nothing latched before
        _lib.persistent_spin_limit(-1)
        # a reported solve synchronises and reports the latch
        with pytest.raises(GlideError, match=r"\(-5\)"):
            fs.solve(10.0)
        _lib.check_status()  # the report cleared it
        # an asynchronous solve leaves it for the next check
        fs.fas_options.set(asynchronous=True)
        fs.solve(10.0)
        with pytest.raises(GlideError, match=r"\(-5\)"):
            _lib.check_status()
        _lib.check_status()
        assert not torch.isfinite(model.mg[0].state.H.data).all()
    finally:
        _lib.persistent_spin_limit(1 << 20)
        _lib.graphs_enable(True)
    # with the normal limit the same solve succeeds from a fresh state
    model = _model()
    fs = model.forward_solver
    fs.fas_options.set(maximum_vcycles=1, relative_tolerance=0.0, absolute_tolerance=0.0, report_norms=False)
    fs.solve(10.0)
    _lib.check_status()
    assert torch.isfinite(model.mg[0].state.H.data).all()
