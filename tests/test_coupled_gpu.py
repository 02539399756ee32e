"""Coupled momentum + mass + enthalpy steps on gfx950 (glide_amd.coupled,
the sequence of glide/model.py:156-213 around IceDynamics.forward).

* Three coupled steps of the 128 x 128, 5-level dome with 41 sigma levels
  (FAS-CD solve of 3 V-cycles advancing the geometry, thermal step with strain
  and frictional heating, rate-factor feedback) against the golden fixture the
  FP64 oracle wrote (tests/golden/make_fixtures.py): u, v, H, B and the
  enthalpy to 1e-10 relative, the active set exactly.
* Run-to-run determinism of the production step (fixed-cycle solves replayed
  from HIP graphs, the enthalpy work on a second stream beside the momentum
  adjoint): two fresh runs of 20 steps agree bitwise, in the plain FP64 form
  and in the mixed form.
"""
from pathlib import Path

import numpy as np
import pytest
import torch

from tests.util import rel_err

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
DX = 1500.0
DT = 10.0


def N(t):
    return t.detach().cpu().numpy()


def _coupled(n, levels, nz, seed, precision="fp64", sweeps=None):
    from glide_amd import IceDynamics, ThermalModel
    from glide_amd.synthetic import dome_problem, install, thermal_forcing
    p = dome_problem(n, DX, seed=seed)
    Ts, qg = thermal_forcing(p)
    model = IceDynamics(ny=n, nx=n, dx=DX, n_levels=levels, device="cuda", precision=precision)
    install(model.mg, p)
    th = ThermalModel(model.mg[0], nz=nz, n_smooth=sweeps or 10, strain_heating=True)
    th.initialize(torch.as_tensor(Ts, device="cuda"), Q_geo=torch.as_tensor(qg, device="cuda"))
    if sweeps is not None:  # exactly `sweeps` column sweeps
        th.ops.smoother_config.relative_tolerance = 0.0
        th.ops.smoother_config.absolute_tolerance = 0.0
    return p, Ts, qg, model, th


def test_coupled_steps_match_oracle_fixture():
    from tests.golden.make_fixtures import COUPLED, coupled_checksums
    from glide_amd.coupled import CoupledStep
    c = COUPLED
    f = dict(np.load(GOLD / "coupled_dome128_l5_nz41_3steps.npz"))
    print(f"coupled {c['n']}^2 x {c['nz']}, {c['levels']} levels, {c['steps']} steps", flush=True)
    p, Ts, qg, model, th = _coupled(c["n"], c["levels"], c["nz"], c["seed"], sweeps=c["sweeps"])
    assert np.allclose(coupled_checksums(p, Ts, qg), f["checksums"], rtol=1e-14, atol=0.0)
    model.forward_solver.fas_options.set(maximum_vcycles=3, relative_tolerance=0.0, absolute_tolerance=0.0,
                                         report_norms=False)
    step = CoupledStep(model, th, seeds=None, thermal_seeds=lambda th_, g: None)
    t = 0.0
    for _ in range(c["steps"]):
        step(t, DT)
        t += DT
    torch.cuda.synchronize()
    g = model.mg[0]
    E = N(th.ops.enthalpy_state.E)
    s = c["sub"]
    got = {"u": N(g.state.u.data), "v": N(g.state.v.data), "H": N(g.state.H.data), "B": N(g.rheology.B.data),
           "E_sub": E[::s, ::s, :], "E_colsum": E.sum(axis=2)}
    errs = {k: rel_err(v, f[k]) for k, v in got.items()}
    print("  errors", {k: f"{e:.1e}" for k, e in errs.items()}, flush=True)
    for k, e in errs.items():
        assert e < 1e-10, (k, e)
    assert np.array_equal(N(g.state.mask.data), f["mask"])


def _run(precision, steps):
    from glide_amd.coupled import CoupledStep
    p, Ts, qg, model, th = _coupled(128, 5, 41, seed=11, precision=precision)
    for s in (model.forward_solver, model.adjoint_solver):
        if precision == "fp64":
            s.fas_options.set(maximum_vcycles=3, relative_tolerance=0.0, absolute_tolerance=0.0,
                              report_norms=False, asynchronous=True)
        else:
            s.mixed_options.relative_tolerance = 1e-9
    step = CoupledStep(model, th)
    t = 0.0
    for _ in range(steps):
        step(t, DT)
        t += DT
    torch.cuda.synchronize()
    g = model.mg[0]
    return [N(x) for x in (g.state.u.data, g.state.v.data, g.state.H.data, g.state.mask.data, g.rheology.B.data,
                           th.ops.enthalpy_state.E, g.adjoint.lambda_u.data, g.adjoint.lambda_v.data,
                           g.adjoint.lambda_H.data, g.sliding.beta.grad, g.geometry.bed.grad)]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_coupled_stepper_is_deterministic(precision):
    print(f"coupled stepper determinism ({precision}), 2 x 20 steps at 128^2", flush=True)
    a = _run(precision, 20)
    b = _run(precision, 20)
    for k, (x, y) in enumerate(zip(a, b)):
        assert np.all(np.isfinite(x)), k
        assert np.array_equal(x, y), k
