"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_fixtures.py
from the FP64 oracle): the oracle must still reproduce them (CPU), and the
gfx950 path must match them (GPU; the coupled fixture's GPU test is in
tests/test_coupled_gpu.py)."""
from pathlib import Path

import numpy as np
import pytest

from tests.util import rel_err

GOLD = Path(__file__).resolve().parent / "golden"
DX, DT, SEC = 1500.0, 10.0, 365.25 * 86400.0


def _load(name):
    return dict(np.load(GOLD / name))


def test_oracle_reproduces_ssa_fixture(orc):
    f = _load("ssa_dome32_l3_v3.npz")
    n = f["H0"].shape[0]
    om = orc.OracleMG(n, n, DX, 3)
    for k in ("bed", "B", "beta", "smb"):
        om.set(k, f[k])
    om.set("H", f["H0"])
    om.set("H_prev", f["H0"])
    cfg = orc.OracleMG.fas_config(vcycles=3, rtol=0.0, atol=0.0)
    om.fascd_solve(DT, cfg)
    for k in ("u", "v", "H"):
        assert rel_err(om.get(k), f[k]) < 1e-12, k
    om.put("af_u", -om.get("u"))
    om.put("af_v", -om.get("v"))
    om.put("af_H", np.zeros((n, n)))
    om.fas_adjoint_solve(DT, cfg)
    om.gradients(DT)
    for k in ("lambda_u", "lambda_v", "lambda_H", "g_beta", "g_bed", "g_H_prev", "g_smb"):
        assert rel_err(om.get(k), f[k]) < 1e-12, k


def test_oracle_reproduces_thermal_fixture(orc):
    f = _load("thermal_dome16_nz11.npz")
    n, nz = f["H_prev"].shape[0], f["E"].shape[2]
    th = orc.OracleThermal(n, n, nz, DX, f["H_prev"])
    th.initialize_from_temperature(f["Ts"])
    th.set_surface_enthalpy_from_temperature(f["Ts"])
    th.Q_geo = f["Q_geo"].copy()
    assert rel_err(th.E, f["E_prev"]) < 1e-14
    th.rtol, th.atol = 1e-15, 0.0
    B = orc.thermal_step(th, f["u"], f["v"], f["H"], f["beta"], DT * SEC, n_smooth=200)
    assert rel_err(th.E, f["E"]) < 1e-12
    assert rel_err(B, f["B"]) < 1e-12
    lam = th.adjoint_solve(DT * SEC, f["w"])
    assert rel_err(lam, f["lam"]) < 1e-10


def test_oracle_reproduces_coupled_fixture(orc):
    """Three coupled 128^2 x 41, 5-level steps (about 20 s on 8 cores)."""
    from tests.golden.make_fixtures import coupled_oracle
    f = _load("coupled_dome128_l5_nz41_3steps.npz")
    out = coupled_oracle()
    assert np.array_equal(out["checksums"], f["checksums"])
    for k in ("u", "v", "H", "B", "E_sub", "E_colsum"):
        assert rel_err(out[k], f[k]) < 1e-12, k
    assert np.array_equal(out["mask"], f["mask"])


@pytest.mark.gpu
def test_gpu_matches_ssa_fixture():
    import torch
    from glide_amd import IceDynamics
    f = _load("ssa_dome32_l3_v3.npz")
    n = f["H0"].shape[0]
    print("golden ssa 32x32", flush=True)
    model = IceDynamics(ny=n, nx=n, dx=DX, n_levels=3, device="cuda")
    mg = model.mg
    mg.geometry.bed.set(f["bed"])
    mg.rheology.B.set(f["B"])
    mg.sliding.beta.set(f["beta"])
    mg.forcing.smb.set(f["smb"])
    mg.state.H.set(f["H0"])
    mg.state.H_prev.set(f["H0"])
    for s in (model.forward_solver, model.adjoint_solver):
        s.fas_options.set(maximum_vcycles=3, relative_tolerance=0.0, absolute_tolerance=0.0, report_norms=False)
    model.forward(0.0, DT, update_geometry=False)
    g = mg[0]
    model.backward(0.0, DT, dJdu=g.state.u.data, dJdv=g.state.v.data, dJdH=None)
    torch.cuda.synchronize()
    got = {"u": g.state.u.data, "v": g.state.v.data, "H": g.state.H.data, "lambda_u": g.adjoint.lambda_u.data,
           "lambda_v": g.adjoint.lambda_v.data, "lambda_H": g.adjoint.lambda_H.data, "g_beta": g.sliding.beta.grad,
           "g_bed": g.geometry.bed.grad, "g_H_prev": g.state.H_prev.grad, "g_smb": g.forcing.smb.grad}
    for k, t in got.items():
        assert rel_err(t.detach().cpu().numpy(), f[k]) < 1e-10, k
    assert np.array_equal(g.state.mask.data.cpu().numpy(), f["mask"])


@pytest.mark.gpu
def test_gpu_matches_thermal_fixture():
    import torch
    from glide_amd import Grid, ThermalModel
    f = _load("thermal_dome16_nz11.npz")
    n, nz = f["H_prev"].shape[0], f["E"].shape[2]
    print("golden thermal 16x16x11", flush=True)
    g = Grid(n, n, DX, device="cuda")
    g.state.H.set(f["H_prev"])
    g.state.u.set(f["u"])
    g.state.v.set(f["v"])
    g.sliding.beta.set(f["beta"])
    tm = ThermalModel(g, nz=nz, n_smooth=200)
    tm.ops.smoother_config.relative_tolerance = 1e-15
    tm.ops.smoother_config.absolute_tolerance = 0.0
    tm.initialize(torch.as_tensor(f["Ts"], device="cuda"), Q_geo=torch.as_tensor(f["Q_geo"], device="cuda"))
    tm.pre_momentum()
    g.state.H.set(f["H"])
    tm.step(DT * SEC)
    torch.cuda.synchronize()
    assert rel_err(tm.ops.enthalpy_state.E.cpu().numpy(), f["E"]) < 1e-10
    assert rel_err(g.rheology.B.data.cpu().numpy(), f["B"]) < 1e-10
    out = tm.ops.adjoint_solve(DT * SEC, torch.as_tensor(f["w"], device="cuda"), relative_tolerance=1e-14,
                               n_iter=250)
    assert rel_err(out["lambda"].cpu().numpy(), f["lam"]) < 1e-9
    for k in ("E_prev", "Q_geo", "E_surface", "phi_strain"):
        assert rel_err(out[k].cpu().numpy(), f["g_" + k]) < 1e-9, k
