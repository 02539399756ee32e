"""Adjoint of the coupled momentum + enthalpy step over several steps
(glide_amd.inversion.coupled_rollout): each step solves the momentum balance
with the current rate factor, runs the thermal step from its velocities and
thickness, and feeds the new rate factor into the next momentum solve
(glide/model.py:156-213); the chain is differentiated end to end through
GlideStepWarm and ThermalStep with time-step checkpointing.

* The checkpointed gradient equals the stored-state one: recomputed steps
  reproduce the stored ones bit for bit.
* Taylor remainders |J(m + e d) - J(m) - e <dJ/dm, d>| in seeded random
  directions of log beta and of the initial enthalpy fall as e^2 until the
  solves' rounding floor: a missing coupling (the rate factor into the
  momentum solve, the velocities, thickness and sliding coefficient into the
  enthalpy) would leave an O(e) remainder.  The dome's bed is raised so that
  all ice is grounded with the grounded fraction at its clamp (the momentum
  adjoint holds that fraction fixed, as the reference's does).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DX, DT = 1500.0, 5.0


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=torch.float64)


def _setup(n=32, L=3, nz=11, seed=13):
    from glide_amd import IceDynamics, ThermalModel
    from glide_amd.synthetic import dome_problem, install, thermal_forcing
    p = dome_problem(n, DX, seed=seed)
    p.bed = np.maximum(p.bed, 0.0) + 250.0
    Ts, qg = thermal_forcing(p)
    model = IceDynamics(ny=n, nx=n, dx=DX, n_levels=L, device="cuda", precision="fp64")
    install(model.mg, p)
    for s in (model.forward_solver, model.adjoint_solver):
        s.fas_options.set(maximum_vcycles=64, relative_tolerance=1e-13, absolute_tolerance=0.0, report_norms=False)
    th = ThermalModel(model.mg[0], nz=nz, n_smooth=250, strain_heating=True)
    th.initialize(T(Ts), Q_geo=T(qg))
    cfg = th.ops.smoother_config
    cfg.relative_tolerance, cfg.absolute_tolerance = 1e-15, 0.0
    frc = th.ops.enthalpy_forcing
    # every rollout starts from the installed rate factor (the model's B is
    # the last thermal step's after a rollout)
    B0 = model.mg[0].rheology.B.data.clone()
    return (model, th, T(p.H), th.ops.enthalpy_state.E.clone(), T(p.bed), T(np.log(p.beta)), T(p.smb),
            frc.Q_geo.clone(), frc.E_surface.clone(), B0)


def _objective(model, th, H0, bed, smb, Q, Es, B0, obs, wE, steps, checkpoint):
    from glide_amd.inversion import coupled_rollout, misfit

    def step_loss(k, u, v, H, E, B):
        return misfit(u, v, *obs[k]) + 100.0 * (wE * E).sum()

    def J(log_beta, E0):
        loss, _ = coupled_rollout(model, th, H0, E0, bed, torch.exp(log_beta), smb, Q, Es, steps, DT, step_loss,
                                  checkpoint=checkpoint, B0=B0)
        return loss
    return J


def _problem(steps=4):
    from glide_amd.inversion import coupled_rollout
    model, th, H0, E0, bed, lb, smb, Q, Es, B0 = _setup()
    gen = torch.Generator(device="cuda").manual_seed(7)
    lb_true = lb + 0.3 * torch.randn(lb.shape, generator=gen, device="cuda", dtype=lb.dtype)
    obs = []

    def record(k, u, v, H, E, B):
        obs.append((u.detach().clone(), v.detach().clone()))
        return u.new_zeros(())

    with torch.no_grad():
        coupled_rollout(model, th, H0, E0, bed, torch.exp(lb_true), smb, Q, Es, steps, DT, record, checkpoint=None,
                        B0=B0)
    wE = torch.randn(E0.shape, generator=gen, device="cuda", dtype=E0.dtype)
    return model, th, H0, E0, bed, lb, smb, Q, Es, B0, obs, wE, steps, gen


def test_coupled_checkpointed_gradient_equals_stored_gradient():
    print("coupled adjoint 32^2 L3 x 11 levels, 4 steps: checkpoint every 2 vs stored", flush=True)
    model, th, H0, E0, bed, lb, smb, Q, Es, B0, obs, wE, steps, _ = _problem()
    grads = []
    for ck in (None, 2):
        J = _objective(model, th, H0, bed, smb, Q, Es, B0, obs, wE, steps, ck)
        m, e0 = lb.clone().requires_grad_(), E0.clone().requires_grad_()
        L = J(m, e0)
        L.backward()
        grads.append((float(L.detach()), m.grad.clone(), e0.grad.clone()))
        print(f"  checkpoint {ck}: J = {grads[-1][0]:.12e}, |g_beta| = {float(m.grad.norm()):.6e}, "
              f"|g_E0| = {float(e0.grad.norm()):.6e}", flush=True)
    (J0, gb0, ge0), (J1, gb1, ge1) = grads
    assert abs(J1 - J0) <= 1e-14 * abs(J0)  # the segments' partial sums round differently
    assert float((gb1 - gb0).abs().max()) <= 1e-12 * float(gb0.abs().max())
    assert float((ge1 - ge0).abs().max()) <= 1e-12 * float(ge0.abs().max())
    assert float(ge0.abs().max()) > 0.0


def test_coupled_taylor_remainders_are_second_order():
    print("coupled adjoint 32^2 L3 x 11 levels, 4 steps, checkpoint every 2: Taylor test", flush=True)
    model, th, H0, E0, bed, lb, smb, Q, Es, B0, obs, wE, steps, gen = _problem()
    J = _objective(model, th, H0, bed, smb, Q, Es, B0, obs, wE, steps, 2)
    m, e0 = lb.clone().requires_grad_(), E0.clone().requires_grad_()
    L = J(m, e0)
    L.backward()
    gb, ge = m.grad.clone(), e0.grad.clone()
    J0 = float(L.detach())
    eps = [1e-3, 1e-4, 1e-5, 1e-6]
    floor = 1e-11 * abs(J0)
    orders, rel = [], []
    for q, kind in enumerate(("log beta", "E0", "log beta", "E0")):
        if kind == "log beta":
            db = torch.randn(lb.shape, generator=gen, device="cuda", dtype=lb.dtype)
            de = torch.zeros_like(E0)
        else:
            db = torch.zeros_like(lb)
            de = 0.05 * torch.randn(E0.shape, generator=gen, device="cuda", dtype=E0.dtype)
        dJ = float((gb * db).sum() + (ge * de).sum())
        with torch.no_grad():
            rem = [abs(float(J(lb + e * db, E0 + e * de)) - J0 - e * dJ) for e in eps]
        print(f"  direction {q} ({kind}): <g,d> = {dJ:.4e}, remainders " + " ".join(f"{r:.2e}" for r in rem),
              flush=True)
        for k in range(len(eps) - 1):
            if rem[k + 1] > 10 * floor:
                orders.append(np.log10(rem[k] / rem[k + 1]))
        rel.append(rem[-1] / (eps[-1] * abs(dJ)))
    print(f"  observed orders above the floor: {np.round(np.array(orders), 2).tolist()}", flush=True)
    print(f"  relative mismatch at e = 1e-6: {np.round(np.array(rel), 6).tolist()}", flush=True)
    assert max(rel) < 1e-3
    assert len(orders) >= 4
    assert min(orders) > 1.7


def test_coupled_step_adjoint_matches_step_differences():
    """CoupledStep's adjoint of one coupled step (the enthalpy adjoint, the
    thermal step's input gradients into the momentum adjoint's seeds and the
    basal-drag gradient) against central differences of whole coupled steps of
    L = 0.5 |u|^2 + 0.5 |v|^2 + c (0.5 |E|^2 + sum B_new) in a random
    direction of beta (interior cells: the momentum gradient in the
    reference's form drops the boundary-facet terms of the last row and
    column).  The decoupled adjoint (enthalpy and momentum adjoints on their
    own seeds) misses the derivative by far more than the tolerance, so the
    coupling terms are what the agreement checks."""
    from glide_amd.coupled import CoupledStep
    print("coupled step adjoint 32^2 L3 x 11 levels against step differences in beta", flush=True)
    model, th, H0, E0, bed, lb, smb, Q, Es, B0 = _setup()
    g = model.mg[0]
    ops = th.ops
    c = 1e3

    def tensors():
        return [g.state.u.data, g.state.v.data, g.state.H.data, g.state.H_prev.data, g.state.mask.data,
                g.state.phi.data, g.rheology.B.data, g.sliding.beta.data, ops.enthalpy_state.E,
                ops.enthalpy_state.E_prev, ops.H_prev]

    snap = [t.clone() for t in tensors()]

    def run(beta, **kw):
        for t, s in zip(tensors(), snap):
            t.copy_(s)
        model.mg.sliding.beta.set(beta)
        step = CoupledStep(model, th, thermal_seeds=lambda th_, gr: (c * th_.ops.enthalpy_state.E,
                                                                     c * torch.ones_like(gr.rheology.B.data)),
                           enthalpy_adjoint_rtol=1e-14, enthalpy_adjoint_iters=250, **kw)
        step(0.0, DT)
        torch.cuda.synchronize()
        E = ops.enthalpy_state.E
        L = (0.5 * (g.state.u.data ** 2).sum() + 0.5 * (g.state.v.data ** 2).sum()
             + c * (0.5 * (E ** 2).sum() + g.rheology.B.data.sum()))
        return float(L), (g.sliding.beta.grad.clone() if kw.get("seeds", 1) is not None else None)

    beta0 = torch.exp(lb)
    L0, grad = run(beta0)
    _, grad_dec = run(beta0, coupled_adjoint=False)
    gen = torch.Generator(device="cuda").manual_seed(17)
    d = torch.randn(beta0.shape, generator=gen, device="cuda", dtype=beta0.dtype)
    d[:2, :] = 0.0
    d[-2:, :] = 0.0
    d[:, :2] = 0.0
    d[:, -2:] = 0.0
    d = d * beta0
    h = 1e-5
    Lp, _ = run(beta0 + h * d, seeds=None)
    Lm, _ = run(beta0 - h * d, seeds=None)
    fd = (Lp - Lm) / (2 * h)
    ad = float((grad * d).sum())
    dec = float((grad_dec * d).sum())
    print(f"  L = {L0:.10e}: coupled adjoint {ad:.10e}, decoupled {dec:.10e}, step differences {fd:.10e}",
          flush=True)
    tol = 1e-6 * abs(fd)
    assert abs(ad - fd) <= tol
    assert abs(dec - fd) > 10 * tol
