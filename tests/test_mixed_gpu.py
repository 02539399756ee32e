"""The mixed-precision forward and adjoint solves (FP32 multigrid with FP64
defect correction; no counterpart in the reference) against converged FP64
oracle solutions of the same discrete systems.

The FP64 outer residual drives both solves to the FP64 solution, so the bar is
the FP64 solution itself: relative L2 distance <= 1e-9 for u, v, H, the
multipliers and the gradients once the outer loop reaches a relative residual
of 1e-12 (the metric's target is 1e-8), with the active set equal.
"""
import numpy as np
import pytest
import torch

from tests.test_multicolour_gpu import _warm_models
from tests.test_parity_ssa_gpu import N, _dome_models
from tests.util import l2_rel

pytestmark = pytest.mark.gpu

BAR = 1e-9


def _oracle_converged(orc, om, vcycles=60, rtol=1e-13):
    cfg = orc.OracleMG.fas_config(vcycles=vcycles, rtol=rtol, atol=0.0)
    rep = om.fascd_solve(10.0, cfg)
    assert rep.final_norm <= 1e-12 * rep.initial_norm or rep.final_norm < 1e-9, (rep.vcycles, rep.final_norm)
    return rep


def _mixed(model, rtol=1e-12):
    for s in (model.forward_solver, model.adjoint_solver):
        s.precision = "mixed"
        s.mixed_options.relative_tolerance = rtol
        s.mixed_options.max_outer = 40
        s.fas_options.set(report_norms=False)


@pytest.mark.parametrize("n,L", [(64, 4), (32, 3)])
def test_mixed_forward_warm_step_reaches_fp64_solution(orc, n, L):
    print(f"mixed forward warm n={n} L={L}", flush=True)
    model, om = _warm_models(orc, n, L)
    _mixed(model)
    rep = model.forward_solver.solve(10.0)
    torch.cuda.synchronize()
    _oracle_converged(orc, om)
    g = model.mg[0]
    errs = {k: l2_rel(N(getattr(g.state, k).data), om.get(k)) for k in ("u", "v", "H")}
    print(f"  {rep.vcycles} outer iterations {model.forward_solver.last_stats}, "
          f"rel residual {rep.final_norm / rep.initial_norm:.2e}, errors {errs}", flush=True)
    for k, e in errs.items():
        assert e < BAR, (k, e)
    assert np.array_equal(N(g.state.mask.data), om.get("mask"))
    assert rep.final_norm <= 1e-12 * rep.initial_norm
    # a warm step needs no reference-form fallback cycle
    assert model.forward_solver.last_stats["fallbacks"] == 0


def test_mixed_forward_cold_start_reaches_fp64_solution(orc):
    """From the initial state (u = v = 0), where undamped smoothers diverge."""
    print("mixed forward cold 64 L4", flush=True)
    p, model, om = _dome_models(orc, 64, 4)
    _mixed(model)
    rep = model.forward_solver.solve(10.0)
    torch.cuda.synchronize()
    _oracle_converged(orc, om)
    g = model.mg[0]
    errs = {k: l2_rel(N(getattr(g.state, k).data), om.get(k)) for k in ("u", "v", "H")}
    print(f"  {rep.vcycles} outer iterations {model.forward_solver.last_stats}, errors {errs}", flush=True)
    for k, e in errs.items():
        assert e < BAR, (k, e)
    assert np.array_equal(N(g.state.mask.data), om.get("mask"))


def test_mixed_adjoint_and_gradients_reach_fp64_solution(orc):
    """Against the oracle's adjoint solve with the multipliers of pinned rows
    kept at zero: the transpose operator does not determine them, and the
    reference's coarse-grid correction leaves path-dependent values there."""
    print("mixed adjoint 64 L4", flush=True)
    model, om = _warm_models(orc, 64, 4)
    _oracle_converged(orc, om)
    g = model.mg[0]
    # the converged FP64 forward state on both sides
    for k in ("u", "v", "H", "mask", "phi"):
        getattr(g.state, k).set(om.get(k))
    rng = np.random.default_rng(5)
    dJdu = rng.standard_normal((64, 65))
    dJdv = rng.standard_normal((65, 64))
    _mixed(model)
    model.backward(0.0, 10.0, dJdu=dJdu, dJdv=dJdv, dJdH=None)
    torch.cuda.synchronize()
    rep = model.adjoint_solver.last_report
    om.put("af_u", -dJdu)
    om.put("af_v", -dJdv)
    om.put("af_H", np.zeros((64, 64)))
    acfg = orc.OracleMG.fas_config(vcycles=60, rtol=1e-13, atol=0.0)
    acfg.project_pinned = 1
    orep = om.fas_adjoint_solve(10.0, acfg)
    om.gradients(10.0)
    errs = {k: l2_rel(N(getattr(g.adjoint, k).data), om.get(k)) for k in ("lambda_u", "lambda_v", "lambda_H")}
    errs["g_beta"] = l2_rel(N(g.sliding.beta.grad), om.get("g_beta"))
    errs["g_bed"] = l2_rel(N(g.geometry.bed.grad), om.get("g_bed"))
    print(f"  {rep.vcycles} outer iterations (oracle {orep.vcycles} V-cycles) {model.adjoint_solver.last_stats}, "
          f"errors {errs}", flush=True)
    for k, e in errs.items():
        assert e < BAR, (k, e)
    assert rep.final_norm <= 1e-12 * rep.initial_norm


def test_mixed_solves_are_deterministic():
    """Two mixed forward + adjoint solves from the same state are bitwise equal."""
    from glide_amd import IceDynamics
    from glide_amd.synthetic import dome_problem, install
    out = []
    for _ in range(2):
        p = dome_problem(64, 1500.0, seed=4)
        model = IceDynamics(ny=64, nx=64, dx=1500.0, n_levels=4, device="cuda", precision="mixed")
        install(model.mg, p)
        model.forward_solver.fas_options.set(report_norms=False)
        model.adjoint_solver.fas_options.set(report_norms=False)
        model.forward(0.0, 10.0, update_geometry=False)
        g = model.mg[0]
        model.backward(0.0, 10.0, dJdu=g.state.u.data.clone(), dJdv=g.state.v.data.clone())
        torch.cuda.synchronize()
        out.append([N(t) for t in (g.state.u.data, g.state.H.data, g.adjoint.lambda_u.data, g.sliding.beta.grad)])
    for x, y in zip(*out):
        assert np.array_equal(x, y)


def test_mixed_forward_with_floating_ice_reaches_fp64_solution(orc):
    """The 2048² workload's grounding line at 64² (dome_problem(grounding_line=
    True), seed 1: beta = 0 on the floating ice over bed below -300 m, 44 ice
    cells and 684 ice-free ones): the mixed forward solve reaches the FP64
    solution as on grounded ice."""
    from glide_amd import IceDynamics
    from glide_amd.synthetic import dome_problem, install
    n, L, dx = 64, 4, 1500.0
    print(f"mixed forward, floating ice, {n}x{n} L{L}", flush=True)
    p = dome_problem(n, dx, seed=1, grounding_line=True)
    q = dome_problem(n, dx, seed=1)
    assert int(((p.beta == 0) & (q.beta > 0) & (p.H > 0)).sum()) == 44
    model = IceDynamics(ny=n, nx=n, dx=dx, n_levels=L, device="cuda")
    install(model.mg, p)
    om = orc.OracleMG(n, n, dx, L)
    for k in ("bed", "B", "beta", "smb"):
        om.set(k, getattr(p, k))
    om.set("H", p.H)
    om.set("H_prev", p.H)
    fs = model.forward_solver
    fs.fas_options.set(maximum_vcycles=2, relative_tolerance=0.0, absolute_tolerance=0.0, report_norms=False)
    fs.solve(10.0)
    om.fascd_solve(10.0, orc.OracleMG.fas_config(vcycles=2, rtol=0.0, atol=0.0))
    model.mg.state.H_prev.set(model.mg[0].state.H.data.clone())
    om.set("H_prev", om.get("H"))
    _mixed(model)
    rep = fs.solve(10.0)
    torch.cuda.synchronize()
    _oracle_converged(orc, om)
    g = model.mg[0]
    errs = {k: l2_rel(N(getattr(g.state, k).data), om.get(k)) for k in ("u", "v", "H")}
    print(f"  {rep.vcycles} outer iterations {fs.last_stats}, errors {errs}", flush=True)
    for k, e in errs.items():
        assert e < BAR, (k, e)
    assert np.array_equal(N(g.state.mask.data), om.get("mask"))
