"""Multi-step basal-friction inversion on gfx950 (glide_amd.inversion): the
gradient of a velocity misfit over several implicit steps with respect to
log(beta), through the native forward and adjoint solves chained by torch
(glide/torch.py:4-59, examples/greenland/greenland_inverse.py:119-168).

* Time-step checkpointing recomputes the forward solves bit for bit, so the
  checkpointed gradient equals the stored-state gradient exactly.
* Taylor test in 8 seeded random directions (4 for the rate factor B): the remainder
  |J(m + e d) - J(m) - e <dJ/dm, d>| falls by ~100x per decade of e (second
  order) until it reaches the solves' rounding floor.  The adjoint holds the
  grounded fraction fixed (as the reference's does) while the forward solve
  lets it follow H, so the gradient is exact only where the grounded
  fraction cannot vary: the test runs on the dome over a bed raised so that
  all ice is grounded with the fraction at its clamp (second order
  asserted), and reports the remainders on the marine dome, whose
  grounding-zone cells make the gradient first-order accurate only.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DX, DT = 1500.0, 5.0


def _setup(n=64, L=4, steps=6, seed=13, precision="fp64", grounded=True):
    from glide_amd import IceDynamics
    from glide_amd.synthetic import dome_problem, gaussian_random_field, install
    p = dome_problem(n, DX, seed=seed)
    if grounded:  # bed + 0.917 H >= 250 m everywhere: grounded fraction at its clamp
        p.bed = np.maximum(p.bed, 0.0) + 250.0
    model = IceDynamics(ny=n, nx=n, dx=DX, n_levels=L, device="cuda", precision=precision)
    install(model.mg, p)
    for s in (model.forward_solver, model.adjoint_solver):
        s.fas_options.set(maximum_vcycles=64, relative_tolerance=1e-13, absolute_tolerance=0.0, report_norms=False)
        s.mixed_options.relative_tolerance = 1e-12
        s.mixed_options.max_outer = 64
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=torch.float64)
    rng = np.random.default_rng(seed + 1)
    log_beta = np.log(p.beta)
    true_log_beta = log_beta + 0.5 * gaussian_random_field(rng, n, n, DX, 20e3)
    return model, T(p.H), T(p.bed), T(log_beta), T(true_log_beta), T(p.smb), steps


def _observations(model, H0, bed, true_log_beta, smb, steps, noise=5.0, seed=3):
    from glide_amd.inversion import rollout
    obs = []

    def record(k, u, v, H):
        obs.append((u.detach().clone(), v.detach().clone()))
        return u.new_zeros(())

    with torch.no_grad():
        rollout(model, H0, bed, torch.exp(true_log_beta), smb, steps, DT, record, checkpoint=None)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    return [(u + noise * torch.randn(u.shape, generator=gen, device="cuda", dtype=u.dtype),
             v + noise * torch.randn(v.shape, generator=gen, device="cuda", dtype=v.dtype)) for u, v in obs]


def _objective(model, H0, bed, smb, obs, steps, checkpoint, reg=1e-3):
    from glide_amd.inversion import gradient_penalty, misfit, rollout

    def J(log_beta):
        loss, _ = rollout(model, H0, bed, torch.exp(log_beta), smb, steps, DT,
                          lambda k, u, v, H: misfit(u, v, *obs[k]), checkpoint=checkpoint)
        return loss + gradient_penalty(log_beta, reg)
    return J


def test_checkpointed_gradient_equals_stored_gradient():
    print("inversion 64^2 L4: checkpointed vs stored gradient, 6 steps", flush=True)
    model, H0, bed, lb, lb_true, smb, steps = _setup()
    obs = _observations(model, H0, bed, lb_true, smb, steps)
    grads = []
    for ck in (None, 2, 4):
        m = lb.clone().requires_grad_()
        J = _objective(model, H0, bed, smb, obs, steps, ck)(m)
        J.backward()
        grads.append((float(J.detach()), m.grad.clone()))
    (J0, g0) = grads[0]
    for J1, g1 in grads[1:]:
        # the forward and adjoint solves are recomputed bit for bit; the loss
        # and the gradient are summed over the segments in another order
        assert abs(J1 - J0) <= 1e-14 * abs(J0)
        assert float((g1 - g0).norm()) <= 1e-13 * float(g0.norm())
    assert torch.isfinite(g0).all() and float(g0.abs().max()) > 0


def _taylor(grounded, directions):
    model, H0, bed, lb, lb_true, smb, steps = _setup(grounded=grounded)
    obs = _observations(model, H0, bed, lb_true, smb, steps)
    Jf = _objective(model, H0, bed, smb, obs, steps, 2)
    m = lb.clone().requires_grad_()
    J0 = Jf(m)
    J0.backward()
    g = m.grad.clone()
    J0 = float(J0.detach())
    # cells whose grounded fraction is off its +-20 clamp (varies with H)
    g0 = model.mg[0]
    z = float(g0.geometry.flotation_reg_driving.value) * (g0.geometry.bed.data + 0.917 * g0.state.H.data)
    partial = int(((z > -20.0) & (z < 20.0)).sum())
    print(f"  J = {J0:.6e}, |g| = {float(g.norm()):.3e}, partly grounded cells after the last step: {partial}",
          flush=True)
    gen = torch.Generator(device="cuda").manual_seed(29)
    eps = [1e-2, 1e-3, 1e-4, 1e-5, 1e-6]
    floor = 1e-11 * abs(J0)
    orders, rel = [], []
    for q in range(directions):
        d = torch.randn(m.shape, generator=gen, device="cuda", dtype=m.dtype)
        dJ = float((g * d).sum())
        rem = []
        with torch.no_grad():
            for e in eps:
                rem.append(abs(float(Jf(lb + e * d)) - J0 - e * dJ))
        print(f"  direction {q}: <g,d> = {dJ:.4e}, remainders " + " ".join(f"{r:.2e}" for r in rem), flush=True)
        # second order on the decades whose remainder is above the floor
        for k in range(len(eps) - 1):
            if rem[k + 1] > 10 * floor:
                orders.append(np.log10(rem[k] / rem[k + 1]))
        rel.append(rem[-1] / (eps[-1] * abs(dJ)))
    print(f"  observed orders above the floor: {np.round(np.array(orders), 2).tolist()}", flush=True)
    return orders, rel


def test_taylor_remainders_are_second_order_on_grounded_ice():
    print("inversion 64^2 L4, grounded dome: Taylor test, 8 directions, 6 steps, checkpoint every 2", flush=True)
    orders, rel = _taylor(True, 8)
    # the first-order term is matched: at e = 1e-6 the remainder is far
    # below e |<g,d>|
    assert max(rel) < 1e-3
    assert len(orders) >= 8
    assert min(orders) > 1.7


def test_taylor_remainders_on_the_marine_dome_are_reported():
    """The grounding zone: the adjoint's fixed grounded fraction (as in the
    reference) misses the grounded fraction's change with H, so the
    remainder keeps a first-order part e |<g,d> - dJ/de|: it decays one
    decade per decade of e once the second-order part is below it.  The
    relative mismatch of the directional derivative is reported."""
    print("inversion 64^2 L4, marine dome: Taylor test, 4 directions (reported)", flush=True)
    orders, rel = _taylor(False, 4)
    print(f"  relative mismatch of the directional derivative (e = 1e-6): {np.round(rel, 3).tolist()}", flush=True)
    assert all(np.isfinite(rel))
    assert min(orders) > 0.2 and np.median(orders) > 0.9


def test_taylor_remainders_in_the_rate_factor_are_second_order():
    """The misfit's gradient with respect to log B (the rate factor, no
    counterpart in the reference's GlideStep) through the checkpointed
    rollout, on the grounded dome: second-order Taylor remainders."""
    from glide_amd.inversion import misfit, rollout
    print("inversion 64^2 L4, grounded dome: Taylor test in log B, 4 directions, 6 steps", flush=True)
    model, H0, bed, lb, lb_true, smb, steps = _setup(grounded=True)
    beta = torch.exp(lb)
    B0 = model.mg[0].rheology.B.data.clone()
    gen = torch.Generator(device="cuda").manual_seed(41)
    logB_true = torch.log(B0) + 0.3 * torch.randn(B0.shape, generator=gen, device="cuda", dtype=B0.dtype)
    obs = []

    def record(k, u, v, H):
        obs.append((u.detach().clone(), v.detach().clone()))
        return u.new_zeros(())

    with torch.no_grad():
        rollout(model, H0, bed, beta, smb, steps, DT, record, checkpoint=None, B=torch.exp(logB_true))

    def Jf(logB):
        loss, _ = rollout(model, H0, bed, beta, smb, steps, DT, lambda k, u, v, H: misfit(u, v, *obs[k]),
                          checkpoint=2, B=torch.exp(logB))
        return loss

    m = torch.log(B0).clone().requires_grad_()
    J0 = Jf(m)
    J0.backward()
    g = m.grad.clone()
    J0 = float(J0.detach())
    eps = [1e-2, 1e-3, 1e-4, 1e-5]
    floor = 1e-11 * abs(J0)
    orders, rel = [], []
    for q in range(4):
        d = torch.randn(m.shape, generator=gen, device="cuda", dtype=m.dtype)
        dJ = float((g * d).sum())
        with torch.no_grad():
            rem = [abs(float(Jf(torch.log(B0) + e * d)) - J0 - e * dJ) for e in eps]
        print(f"  direction {q}: <g,d> = {dJ:.4e}, remainders " + " ".join(f"{r:.2e}" for r in rem), flush=True)
        for k in range(len(eps) - 1):
            if rem[k + 1] > 10 * floor:
                orders.append(np.log10(rem[k] / rem[k + 1]))
        rel.append(rem[-1] / (eps[-1] * abs(dJ)))
    print(f"  observed orders above the floor: {np.round(np.array(orders), 2).tolist()}", flush=True)
    # second order down to e = 1e-5: a directional-derivative error d would
    # add d e to the remainder, so |d| < rem(1e-5) / 1e-5, a bound that the
    # curvature in log B keeps near 1e-3 |<g,d>|
    assert max(rel) < 1e-2
    assert len(orders) >= 4
    assert min(orders) > 1.7
