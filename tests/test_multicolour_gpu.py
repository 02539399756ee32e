"""The multicolour (red-black) Vanka smoother on gfx950 against the oracle's
restatement (mc_sweep, FASCD with multicolour configs), and against the
additive smoother's fixed point.

The smoother has no counterpart in the reference: its cells solve the
reference's 5x5 patch system (glide/cuda/vanka.cu:90-475), but red cells,
then black cells, write their unknowns at full weight.  Bars: one pass 1e-12
relative, a few passes 1e-11, whole solves 1e-10; the fixed point of a
converged multicolour solve equals the additive solve's to 1e-9.
"""
import numpy as np
import pytest
import torch

from tests.test_parity_ssa_gpu import N, grid_from
from tests.util import dome_fields, rel_err

pytestmark = pytest.mark.gpu

FAST = dict(relaxation=1.0, steps=3, ssa_damping=0.0, mc_damping=0.0, tol=0.0)
ROBUST = dict(relaxation=0.5, steps=30, ssa_damping=0.01, mc_damping=1.0, tol=1e-6)
# the mixed solver's setting: three half steps, no dampings (stable from cold
# starts, where undamped full steps diverge)
MIXED = dict(relaxation=0.5, steps=3, ssa_damping=0.0, mc_damping=0.0, tol=0.0)


def _configure(vc, s):
    vc.multicolour = True
    vc.newton_tol = s["tol"]
    n = vc.newton_config
    n.steps, n.relaxation, n.ssa_damping, n.mc_damping = s["steps"], s["relaxation"], s["ssa_damping"], s["mc_damping"]


def _oracle_kw(s):
    return dict(newton_steps=s["steps"], relaxation=s["relaxation"], newton_tol=s["tol"],
                ssa_damping=s["ssa_damping"], mc_damping=s["mc_damping"])


@pytest.mark.parametrize("shape", [(32, 32), (24, 40), (66, 130), (17, 129), (3, 5)])
@pytest.mark.parametrize("mode", ["fast", "robust"])
@pytest.mark.parametrize("m", [1.0, 1.0 / 3.0])
def test_mc_passes_match_oracle(orc, shape, mode, m):
    ny, nx = shape
    s = FAST if mode == "fast" else ROBUST
    # undamped full Newton steps diverge from these unbalanced random states
    # (it is the smoother for states near a solution), so it gets one pass
    n_pass = 1 if mode == "fast" else 3
    print(f"mc passes shape={shape} mode={mode} m={m:.3f}", flush=True)
    F = dome_fields(ny=ny, nx=nx, seed=31, mask_frac=0.1)
    P = orc.phys(m=m, calving_rate=0.2)
    g = grid_from(F)
    g.sliding.m.set(m)
    g.calving.calving_rate.set(0.2)
    fo = g.forward_operators
    _configure(fo.vanka_config, s)
    Fo = dict(F)
    for k in range(n_pass):
        fo.vanka_sweep(10.0, 1, freeze_phi=True)
        torch.cuda.synchronize()
        u, v, H, mask, _ = orc.mc_sweep(Fo, P, 1500.0, 10.0, **_oracle_kw(s))
        Fo.update(u=u, v=v, H=H, mask=mask)
        tol = 1e-12 if k == 0 else 1e-11
        assert rel_err(N(g.state.u.data), u) < tol, k
        assert rel_err(N(g.state.v.data), v) < tol, k
        assert rel_err(N(g.state.H.data), H) < tol, k
        assert np.array_equal(N(g.state.mask.data), mask), k


def _warm_models(orc, n, L, warm=2, ny=None, compat=True):
    """GPU model and oracle hierarchy after `warm` additive V-cycles (matching
    to 1e-10), then H_prev = H (the start of a next time step)."""
    from tests.test_parity_ssa_gpu import _dome_models
    p, model, om = _dome_models(orc, n, L)
    fs = model.forward_solver
    fs.fas_options.set(maximum_vcycles=warm, relative_tolerance=0.0, absolute_tolerance=0.0, report_norms=False,
                       compat_vprolong=compat)
    fs.solve(10.0)
    cfg = orc.OracleMG.fas_config(vcycles=warm, rtol=0.0, atol=0.0)
    cfg.compat_hfacet = int(compat)
    om.fascd_solve(10.0, cfg)
    model.mg.state.H_prev.set(model.mg[0].state.H.data.clone())
    om.set("H_prev", om.get("H"))
    return model, om


@pytest.mark.parametrize("n,L", [(64, 4), (32, 3)])
def test_mc_fascd_solve_matches_oracle(orc, n, L):
    print(f"mc fascd solve n={n} L={L}", flush=True)
    model, om = _warm_models(orc, n, L)
    fs = model.forward_solver
    for lev in fs.levels:
        _configure(lev.grid.forward_operators.vanka_config, MIXED)
    fs.fas_options.set(maximum_vcycles=3, coarsest_steps=40, pre_steps=4, post_steps=4, finest_steps=8,
                       compat_vprolong=False)
    rep = fs.solve(10.0)
    torch.cuda.synchronize()
    cfg = orc.OracleMG.fas_config(coarsest=40, pre=4, post=4, finest=8, vcycles=3, rtol=0.0, atol=0.0)
    cfg.compat_hfacet = 0
    orep = om.fascd_solve(10.0, cfg, om.vanka_configs(multicolour=True, **_oracle_kw(MIXED)))
    g = model.mg[0]
    for k in ("u", "v", "H"):
        assert rel_err(N(getattr(g.state, k).data), om.get(k)) < 1e-10, k
    assert np.array_equal(N(g.state.mask.data), om.get("mask"))
    assert rep.norms[2] < 1e-3 * rep.initial_norm
    for k in range(3):
        assert abs(rep.norms[k] - orep.norms[k]) <= 1e-8 * orep.norms[0]


def test_mc_non_square_solve_matches_oracle(orc):
    """64 x 32 with three levels and the v-facet prolongation (the reference's
    quirk needs square grids), forward solve against the oracle."""
    from glide_amd import IceDynamics
    from glide_amd.synthetic import dome_problem
    print("mc fascd solve 32x64 L3", flush=True)
    ny, nx, L = 32, 64, 3
    p = dome_problem(nx, 1500.0, seed=3)
    sl = (slice(0, ny), slice(0, nx))
    model = IceDynamics(ny=ny, nx=nx, dx=1500.0, n_levels=L, device="cuda")
    mg = model.mg
    mg.geometry.bed.set(p.bed[sl].copy())
    mg.rheology.B.set(p.B[sl].copy())
    mg.sliding.beta.set(p.beta[sl].copy())
    mg.forcing.smb.set(p.smb[sl].copy())
    mg.state.H.set(p.H[sl].copy())
    mg.state.H_prev.set(p.H[sl].copy())
    om = orc.OracleMG(ny, nx, 1500.0, L)
    for k in ("bed", "B", "beta", "smb", "H"):
        om.set(k, getattr(p, k)[sl].copy())
    om.set("H_prev", p.H[sl].copy())
    fs = model.forward_solver
    for lev in fs.levels:
        _configure(lev.grid.forward_operators.vanka_config, ROBUST)
    fs.fas_options.set(maximum_vcycles=2, coarsest_steps=40, pre_steps=4, post_steps=4, finest_steps=8,
                       compat_vprolong=False, relative_tolerance=0.0, absolute_tolerance=0.0, report_norms=False)
    fs.solve(10.0)
    torch.cuda.synchronize()
    cfg = orc.OracleMG.fas_config(coarsest=40, pre=4, post=4, finest=8, vcycles=2, rtol=0.0, atol=0.0)
    cfg.compat_hfacet = 0
    om.fascd_solve(10.0, cfg, om.vanka_configs(multicolour=True, **_oracle_kw(ROBUST)))
    g = model.mg[0]
    for k in ("u", "v", "H"):
        assert rel_err(N(getattr(g.state, k).data), om.get(k)) < 1e-10, k


def test_mc_solve_is_deterministic_and_reaches_the_additive_fixed_point(orc):
    """Two multicolour solves are bitwise identical; converged to a tight
    tolerance, they reach the additive smoother's solution."""
    out = []
    for mc in (True, True, False):
        model, _ = _warm_models(orc, 64, 4)
        fs = model.forward_solver
        if mc:
            for lev in fs.levels:
                _configure(lev.grid.forward_operators.vanka_config, MIXED)
            fs.fas_options.set(coarsest_steps=40, pre_steps=4, post_steps=4, finest_steps=8)
        fs.fas_options.set(maximum_vcycles=40 if not mc else 20, relative_tolerance=1e-13, absolute_tolerance=0.0,
                           compat_vprolong=True)
        rep = fs.solve(10.0)
        print(f"  {'mc' if mc else 'additive'}: {rep.vcycles} V-cycles to "
              f"{rep.final_norm / rep.initial_norm:.2e}", flush=True)
        g = model.mg[0]
        out.append([N(t.data) for t in (g.state.u, g.state.v, g.state.H, g.state.mask)])
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)
    for x, y in zip(out[0][:3], out[2][:3]):
        assert rel_err(x, y) < 1e-9
    assert np.array_equal(out[0][3], out[2][3])
