"""Gradients of the converged thermal step with respect to the inputs the
momentum step hands it: the SSA velocities u, v (m/a), the new and
pre-momentum thicknesses H, H_prev, the sliding coefficient beta (frictional
heating) and the rate factor B the strain heating reads
(glide_thermal_step_vjp, ThermalModel.adjoint, ThermalStep).

The reference has no enthalpy adjoint; the bar is the derivative of the
reference's own thermal step (glide/model.py:167-254, restated by the
oracle): directional derivatives <g, d> against central differences of the
oracle's converged step in seeded random directions, every input on its own
so that a wrong term shows in its group, and second-order Taylor remainders of
the GPU step through the autograd surface.
"""
import numpy as np
import pytest
import torch

from tests.test_parity_enthalpy_gpu import DT, DX, N, T, _thermal_pair

pytestmark = pytest.mark.gpu

KEYS = ("u", "v", "H", "H_prev", "beta", "B")


def _converged(tm):
    tm.n_smooth = 250
    cfg = tm.ops.smoother_config
    cfg.relative_tolerance, cfg.absolute_tolerance = 1e-15, 0.0


def _oracle_loss(base, E0, Ts, Qg, nz, w, wB, strain, frictional, pert=None, h=0.0):
    """<w, E> + <wB, B> of the oracle's converged thermal step at the inputs
    `base` moved by h along pert[key]."""
    from oracle.oracle import OracleThermal, strain_heating, thermal_step
    x = {k: base[k] + (h * pert[k] if pert is not None and k in pert else 0.0) for k in KEYS}
    ny, nx = x["H"].shape
    th = OracleThermal(ny, nx, nz, DX, x["H_prev"])
    th.E = E0.copy()
    th.E_prev = E0.copy()
    th.H_prev = x["H_prev"].copy()
    th.set_surface_enthalpy_from_temperature(Ts)
    th.Q_geo = Qg.copy()
    th.rtol, th.atol = 1e-15, 0.0
    phi = None
    if strain:
        phi = np.repeat(strain_heating(x["u"], x["v"], x["B"], 3.0, 1e-6, DX, 917.0)[:, :, None], nz, axis=2)
    B = thermal_step(th, x["u"], x["v"], x["H"], x["beta"], DT, n_smooth=400, frictional_heating=frictional,
                     strain_heating=phi)
    return float((w * th.E).sum() + (wB * B).sum())


@pytest.mark.parametrize("n,nz,seed,strain,frictional", [(12, 11, 21, True, True), (10, 21, 22, False, True),
                                                         (11, 9, 23, True, False)])
def test_input_gradients_match_oracle_differences(n, nz, seed, strain, frictional):
    print(f"thermal input gradients n={n} nz={nz} seed={seed} strain={strain} frictional={frictional}", flush=True)
    F, g, tm, _ = _thermal_pair(n, nz, seed, frictional, strain)
    _converged(tm)
    ops = tm.ops
    E0 = N(ops.enthalpy_state.E_prev).copy()
    Qg = N(ops.enthalpy_forcing.Q_geo).copy()
    Ts = 243.15 + 6.5e-3 * np.maximum(2000.0 - (F["bed"] + F["H"]), 0.0) * 0.2
    base = {"u": F["u"], "v": F["v"], "H": F["H"], "H_prev": F["H_prev"], "beta": F["beta"], "B": F["B"]}
    B_in = g.rheology.B.data.clone()
    rep = tm.step(DT, report=True)
    print(f"  forward sweeps {rep.sweeps}, |r| {rep.r:.2e} of {rep.r0:.2e}", flush=True)
    rng = np.random.default_rng(seed + 50)
    w = rng.standard_normal((n, n, nz))
    wB = 1e-3 * rng.standard_normal((n, n))
    out = tm.adjoint(DT, T(w), T(wB), B_in=B_in, relative_tolerance=1e-14, n_iter=250, report=True)
    print(f"  adjoint sweeps {out['report'].sweeps}", flush=True)
    H = F["H"]
    # directions: velocities on their own scale, thicknesses relative (kept
    # clear of the thin-column threshold), beta and B relative
    dirs = {"u": rng.standard_normal(F["u"].shape) * np.abs(F["u"]).max(),
            "v": rng.standard_normal(F["v"].shape) * np.abs(F["v"]).max(),
            "H": rng.standard_normal(H.shape) * H * (np.abs(H - 25.0) > 1.0),
            "H_prev": rng.standard_normal(H.shape) * F["H_prev"],
            "beta": rng.standard_normal(H.shape) * F["beta"],
            "B": rng.standard_normal(H.shape) * F["B"]}
    h = 1e-6
    for key in KEYS:
        gk = out[key]
        if gk is None:
            assert (key == "beta" and not frictional) or (key == "B" and not strain)
            continue
        ad = float((N(gk) * dirs[key]).sum())
        Lp = _oracle_loss(base, E0, Ts, Qg, nz, w, wB, strain, frictional, {key: dirs[key]}, h)
        Lm = _oracle_loss(base, E0, Ts, Qg, nz, w, wB, strain, frictional, {key: dirs[key]}, -h)
        fd = (Lp - Lm) / (2 * h)
        print(f"  {key}: adjoint {ad:.10e}  oracle fd {fd:.10e}", flush=True)
        assert abs(ad - fd) <= 2e-6 * max(abs(fd), abs(ad)) + 1e-12, key


def test_thermal_step_autograd_taylor_remainders_in_the_momentum_outputs():
    """ThermalStep with u, v, H, H_prev, beta and B as leaves: the Taylor
    remainder |L(p + e d) - L(p) - e <grad L, d>| falls as e^2 along a
    combined random direction (a wrong or missing term leaves an O(e)
    remainder), and another step run between the forward and the backward
    does not change the gradient (the backward rebuilds the step)."""
    print("thermal step autograd, momentum inputs", flush=True)
    from glide_amd.torch import ThermalStep
    F, g, tm, _ = _thermal_pair(14, 11, 31, frictional=True, strain=True)
    _converged(tm)
    ops = tm.ops
    E_prev0 = ops.enthalpy_state.E.clone()
    Q0 = ops.enthalpy_forcing.Q_geo.clone()
    Es0 = ops.enthalpy_forcing.E_surface.clone()
    x0 = {"u": T(F["u"]), "v": T(F["v"]), "H": T(F["H"]), "H_prev": ops.H_prev.clone(), "beta": T(F["beta"]),
          "B": T(F["B"])}
    rng = np.random.default_rng(5)
    w = T(rng.standard_normal(tuple(E_prev0.shape)))
    wB = T(rng.standard_normal(tuple(Q0.shape)))

    def loss(x):
        E, B = ThermalStep.apply(tm, DT, E_prev0, Q0, Es0, x["u"], x["v"], x["H"], x["H_prev"], x["beta"], x["B"])
        return (w * E).sum() + 1e-3 * (wB * B).sum()

    leaves = {k: t.clone().requires_grad_(True) for k, t in x0.items()}
    L0 = loss(leaves)
    # another step in between overwrites the operators' buffers
    with torch.no_grad():
        loss({k: t * (1.0 + 1e-2) for k, t in x0.items()})
    L0.backward()
    grads = {k: leaves[k].grad.clone() for k in KEYS}
    # the same gradient without the step in between
    leaves2 = {k: t.clone().requires_grad_(True) for k, t in x0.items()}
    loss(leaves2).backward()
    for k in KEYS:
        ref = N(leaves2[k].grad)
        assert np.abs(N(grads[k]) - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-300), k
    Hn = F["H"]
    d = {"u": T(rng.standard_normal(F["u"].shape) * np.abs(F["u"]).max()),
         "v": T(rng.standard_normal(F["v"].shape) * np.abs(F["v"]).max()),
         "H": T(rng.standard_normal(Hn.shape) * Hn * (np.abs(Hn - 25.0) > 1.0)),
         "H_prev": T(rng.standard_normal(Hn.shape)) * x0["H_prev"],
         "beta": T(rng.standard_normal(Hn.shape) * F["beta"]), "B": T(rng.standard_normal(Hn.shape) * F["B"])}
    slope = sum(float((grads[k] * d[k]).sum()) for k in KEYS)
    L0v = float(L0.detach())
    rems = []
    for e in (1e-4, 1e-5, 1e-6):
        with torch.no_grad():
            Le = float(loss({k: x0[k] + e * d[k] for k in KEYS}))
        rems.append(abs(Le - L0v - e * slope))
        print(f"  e={e:.0e}: remainder {rems[-1]:.3e} (first-order term {abs(e * slope):.3e})", flush=True)
    orders = [np.log10(rems[q] / rems[q + 1]) for q in range(2)]
    print(f"  orders {orders}", flush=True)
    assert rems[0] < 1e-2 * abs(1e-4 * slope)
    assert min(orders) > 1.7
