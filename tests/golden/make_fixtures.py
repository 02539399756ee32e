"""Write the golden fixtures of the path from the FP64 oracle (seeded inputs).

Each .npz holds inputs and expected outputs at sizes the oracle finishes in
seconds: a 32x32, 3-level FAS-CD forward solve (3 V-cycles) with the adjoint
solve and the four gradients for the loss 0.5|u|^2 + 0.5|v|^2, a 16x16x11
thermal step on that dome with its adjoint gradients, and three coupled steps
(forward solve of 3 V-cycles advancing the geometry, thermal step with strain
and frictional heating, rate-factor feedback) of the 128x128, 5-level dome
with 41 sigma levels.  The coupled fixture keeps the outputs only (its inputs
come from glide_amd.synthetic with the stored seed; input checksums guard
that) and the enthalpy on every 8th column.  The oracle
itself is pinned by tests/test_oracle_*.py; these fixtures freeze its outputs
so that a change of either implementation shows up against fixed numbers.

    python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

DX = 1500.0
DT = 10.0
SEC = 365.25 * 86400.0


def ssa_fixture():
    from glide_amd.synthetic import dome_problem
    from oracle import oracle as O
    n, L = 32, 3
    p = dome_problem(n, DX, seed=21)
    om = O.OracleMG(n, n, DX, L)
    for k in ("bed", "B", "beta", "smb"):
        om.set(k, getattr(p, k))
    om.set("H", p.H)
    om.set("H_prev", p.H)
    cfg = O.OracleMG.fas_config(vcycles=3, rtol=0.0, atol=0.0)
    om.fascd_solve(DT, cfg)
    u, v, H = om.get("u"), om.get("v"), om.get("H")
    om.put("af_u", -u)
    om.put("af_v", -v)
    om.put("af_H", np.zeros((n, n)))
    om.fas_adjoint_solve(DT, cfg)
    om.gradients(DT)
    out = {"bed": p.bed, "B": p.B, "beta": p.beta, "smb": p.smb, "H0": p.H, "u": u, "v": v, "H": H,
           "mask": om.get("mask")}
    for k in ("lambda_u", "lambda_v", "lambda_H", "g_beta", "g_bed", "g_H_prev", "g_smb"):
        out[k] = om.get(k)
    np.savez_compressed(HERE / "ssa_dome32_l3_v3.npz", **out)


def thermal_fixture():
    from glide_amd.synthetic import dome_problem, thermal_forcing
    from oracle import oracle as O
    n, nz = 16, 11
    p = dome_problem(n, DX, seed=22)
    rng = np.random.default_rng(23)
    u = 40.0 * rng.standard_normal((n, n + 1))
    v = 30.0 * rng.standard_normal((n + 1, n))
    H_new = p.H + 2.0 * rng.standard_normal((n, n)) * (p.H > 30.0)
    Ts, qg = thermal_forcing(p)
    th = O.OracleThermal(n, n, nz, DX, p.H)
    th.initialize_from_temperature(Ts)
    th.set_surface_enthalpy_from_temperature(Ts)
    th.Q_geo = qg.copy()
    E_prev = th.E.copy()
    th.rtol, th.atol = 1e-15, 0.0
    B = O.thermal_step(th, u, v, H_new, p.beta, DT * SEC, n_smooth=200)
    w = rng.standard_normal(th.E.shape)
    lam = th.adjoint_solve(DT * SEC, w)
    g = th.adjoint_gradients(DT * SEC, lam)
    np.savez_compressed(HERE / "thermal_dome16_nz11.npz", H_prev=p.H, H=H_new, u=u, v=v, beta=p.beta, Ts=Ts,
                        Q_geo=qg, E_prev=E_prev, E=th.E, B=B, w=w, lam=lam, g_E_prev=g["E_prev"],
                        g_Q_geo=g["Q_geo"], g_E_surface=g["E_surface"], g_phi_strain=g["phi_strain"])


COUPLED = dict(n=128, levels=5, nz=41, steps=3, seed=7, sweeps=12, sub=8)


def coupled_inputs(n, seed):
    """The dome workload of the coupled fixture and its thermal forcing."""
    from glide_amd.synthetic import dome_problem, thermal_forcing
    p = dome_problem(n, DX, seed=seed)
    Ts, qg = thermal_forcing(p)
    return p, Ts, qg


def coupled_checksums(p, Ts, qg):
    return np.array([p.bed.sum(), p.B.sum(), p.beta.sum(), p.smb.sum(), p.H.sum(), Ts.sum(), qg.sum()])


def coupled_oracle(c=COUPLED):
    """c['steps'] coupled steps on the oracle: pre_momentum, FAS-CD solve of 3
    V-cycles, H_prev = H, thermal step (exactly c['sweeps'] column sweeps),
    B = the new rate factor (the coarse levels keep the rate factor restricted
    at set-up, as in the reference)."""
    from oracle import oracle as O
    n, L, nz = c["n"], c["levels"], c["nz"]
    p, Ts, qg = coupled_inputs(n, c["seed"])
    om = O.OracleMG(n, n, DX, L)
    for k in ("bed", "B", "beta", "smb"):
        om.set(k, getattr(p, k))
    om.set("H", p.H)
    om.set("H_prev", p.H)
    th = O.OracleThermal(n, n, nz, DX, p.H)
    th.initialize_from_temperature(Ts)
    th.set_surface_enthalpy_from_temperature(Ts)
    th.Q_geo = qg.copy()
    th.rtol, th.atol = 0.0, 0.0
    cfg = O.OracleMG.fas_config(vcycles=3, rtol=0.0, atol=0.0)
    for _ in range(c["steps"]):
        th.E_prev = th.E.copy()
        th.H_prev = om.get("H").copy()
        om.fascd_solve(DT, cfg)
        H = om.get("H")
        # the reference writes H_prev and the new rate factor on the finest
        # level only (glide/model.py:65-72, 204-212): put, not set
        om.put("H_prev", H)
        u, v, B = om.get("u"), om.get("v"), om.get("B")
        phi = np.repeat(O.strain_heating(u, v, B, 3.0, 1e-6, DX)[:, :, None], nz, axis=2)
        om.put("B", O.thermal_step(th, u, v, H, p.beta, DT * SEC, n_smooth=c["sweeps"], strain_heating=phi))
    s = c["sub"]
    return {"u": om.get("u"), "v": om.get("v"), "H": om.get("H"), "B": om.get("B"), "mask": om.get("mask"),
            "E_sub": th.E[::s, ::s, :].copy(), "E_colsum": th.E.sum(axis=2), "checksums": coupled_checksums(p, Ts, qg)}


def coupled_fixture():
    out = coupled_oracle()
    np.savez_compressed(HERE / "coupled_dome128_l5_nz41_3steps.npz", **out)


if __name__ == "__main__":
    ssa_fixture()
    thermal_fixture()
    coupled_fixture()
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size, "bytes")
