"""Parity of the gfx950 enthalpy kernels, the fused column sweep and the fused
thermal step with the FP64 CPU oracle.

Every test runs the HIP path through the C ABI (glide_amd.enthalpy /
glide_amd.model.ThermalModel) and the oracle on the same seeded columns:
cold and temperate ice, thin and ice-free columns, depth-varying velocities.
Bar: single kernels agree to 1e-12 relative to the field's scale, whole
sweeps and thermal steps to 1e-10 relative (enthalpy, omega, forcing, B).
"""
import numpy as np
import pytest
import torch

from tests.util import enth_case, rel_err

pytestmark = pytest.mark.gpu

KTOL = 1e-12   # one kernel
STOL = 1e-10   # sweeps / thermal steps
DT = 10.0 * 365.25 * 86400.0
DX = 1500.0


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=torch.float64)


def N(t):
    return t.detach().cpu().numpy()


def make_pair(C, nz, flags=7):
    """(oracle, EnthalpyOperators) holding the same state, f_E from the oracle's
    set_rhs copied to the device."""
    from glide_amd.enthalpy import EnthalpyOperators
    from glide_amd.grid import Grid
    from oracle.oracle import OracleThermal
    ny, nx = C["H"].shape
    th = OracleThermal(ny, nx, nz, DX, C["H"])
    th.flags = flags
    for k in ("E", "E_prev", "u3d", "v3d", "omega", "E_surface", "Q_geo", "Q_fh", "phi_strain", "H_prev"):
        setattr(th, k, np.ascontiguousarray(C[k], dtype=np.float64).copy())
    th.set_rhs(snapshot=False)
    g = Grid(ny, nx, DX, device="cuda")
    g.state.H.set(C["H"])
    ops = EnthalpyOperators(g, nz=nz)
    ops.term_flags.horizontal_advection = bool(flags & 1)
    ops.term_flags.omega = bool(flags & 2)
    ops.term_flags.drainage = bool(flags & 4)
    st, vel, frc = ops.enthalpy_state, ops.enthalpy_velocity, ops.enthalpy_forcing
    st.E.copy_(T(C["E"]))
    st.E_prev.copy_(T(C["E_prev"]))
    vel.u3d.copy_(T(C["u3d"]))
    vel.v3d.copy_(T(C["v3d"]))
    vel.omega.copy_(T(C["omega"]))
    for k in ("E_surface", "Q_geo", "Q_fh", "phi_strain"):
        getattr(frc, k).copy_(T(C[k]))
    ops.H_prev.copy_(T(C["H_prev"]))
    ops.set_rhs(snapshot=False)
    return th, ops


CASES = [(12, 10, 11, 0), (9, 14, 41, 1), (6, 5, 64, 2), (5, 7, 2, 3), (1, 3, 21, 4)]


@pytest.mark.parametrize("ny,nx,nz,seed", CASES)
def test_set_rhs(ny, nx, nz, seed):
    print(f"set_rhs {ny}x{nx}x{nz} seed={seed}", flush=True)
    th, ops = make_pair(enth_case(ny, nx, nz, seed), nz)
    assert rel_err(N(ops.f_E), th.f_E) < KTOL


@pytest.mark.parametrize("flags", [7, 6, 5, 3, 1, 0])
@pytest.mark.parametrize("ny,nx,nz,seed", CASES)
def test_residual(ny, nx, nz, seed, flags):
    print(f"residual {ny}x{nx}x{nz} seed={seed} flags={flags}", flush=True)
    th, ops = make_pair(enth_case(ny, nx, nz, seed), nz, flags)
    m_gpu = ops.compute_residual(DT)
    m_orc = th.compute_residual(DT)
    assert rel_err(N(ops.r_E), th.r_E) < KTOL
    assert abs(m_gpu - m_orc) <= KTOL * m_orc


@pytest.mark.parametrize("flags", [7, 2, 0])
@pytest.mark.parametrize("ny,nx,nz,seed", CASES)
def test_column_smooth(ny, nx, nz, seed, flags):
    print(f"column_smooth {ny}x{nx}x{nz} seed={seed} flags={flags}", flush=True)
    th, ops = make_pair(enth_case(ny, nx, nz, seed), nz, flags)
    ops.column_smooth(DT)
    d = th.column_smooth(DT)
    # scale: the enthalpy itself (the correction solves a diagonally dominant system)
    assert np.max(np.abs(N(ops.delta_E) - d)) < 1e-11 * max(np.max(np.abs(th.E)), 1.0)


@pytest.mark.parametrize("ny,nx,nz,seed", CASES)
def test_layer_smooth(ny, nx, nz, seed):
    print(f"layer_smooth {ny}x{nx}x{nz} seed={seed}", flush=True)
    th, ops = make_pair(enth_case(ny, nx, nz, seed), nz)
    ops.layer_smooth(DT)
    d = th.layer_smooth(DT)
    assert rel_err(N(ops.delta_E), d) < KTOL


@pytest.mark.parametrize("ny,nx,nz,seed", CASES)
def test_omega(ny, nx, nz, seed):
    print(f"omega {ny}x{nx}x{nz} seed={seed}", flush=True)
    from oracle.oracle import SEC_PER_YR
    C = enth_case(ny, nx, nz, seed)
    th, ops = make_pair(C, nz)
    rng = np.random.default_rng(seed + 7)
    dh = 0.5 * rng.standard_normal((ny, nx))
    bmb = 0.01 * rng.random((ny, nx))
    # velocities in m/a, as in the coupled step
    ops.enthalpy_velocity.u3d.mul_(SEC_PER_YR)
    ops.enthalpy_velocity.v3d.mul_(SEC_PER_YR)
    th.u3d = np.ascontiguousarray(N(ops.enthalpy_velocity.u3d))
    th.v3d = np.ascontiguousarray(N(ops.enthalpy_velocity.v3d))
    ops.compute_omega(T(dh), T(bmb))
    om = th.compute_omega(dh, bmb)
    assert rel_err(N(ops.enthalpy_velocity.omega), om) < KTOL
    ops.compute_omega(T(dh))
    om0 = th.compute_omega(dh)
    assert rel_err(N(ops.enthalpy_velocity.omega), om0) < KTOL


@pytest.mark.parametrize("alternating", [True, False])
@pytest.mark.parametrize("ny,nx,nz,seed", CASES)
def test_column_sweep(ny, nx, nz, seed, alternating):
    print(f"column_sweep {ny}x{nx}x{nz} seed={seed} alternating={alternating}", flush=True)
    th, ops = make_pair(enth_case(ny, nx, nz, seed), nz)
    ops.smoother_config.report_norms = True
    ops.smoother_config.absolute_tolerance = 0.0   # the default 1e-2 stops every case after one sweep
    th.atol = 0.0
    ops.column_sweep(DT, 6, alternating=alternating)
    rep = ops.last_report
    it, r0, r = th.column_sweep(DT, 6, alternating=alternating)
    assert rep.sweeps == it
    assert abs(rep.r0 - r0) <= KTOL * r0
    assert abs(rep.r - r) <= 1e-8 * r0
    assert rel_err(N(ops.enthalpy_state.E), th.E) < STOL


def test_column_sweep_stops_early_on_device():
    """A loose tolerance stops after the first sweep; later launches are skipped
    on the device and leave E untouched."""
    print("column_sweep early stop", flush=True)
    th, ops = make_pair(enth_case(10, 12, 21, 5), 21)
    ops.smoother_config.relative_tolerance = 0.9
    th.rtol = 0.9
    ops.smoother_config.report_norms = True
    ops.column_sweep(DT, 8)
    it, _, _ = th.column_sweep(DT, 8)
    assert ops.last_report.sweeps == it < 8
    assert rel_err(N(ops.enthalpy_state.E), th.E) < STOL


def test_column_sweep_host_loop_with_hook_matches_fused():
    print("column_sweep hook path", flush=True)
    C = enth_case(8, 9, 15, 6)
    _, ops_a = make_pair(C, 15)
    th, ops_b = make_pair(C, 15)
    calls = []
    ops_b.smoother_config.hook_func = calls.append
    ops_a.column_sweep(DT, 4)
    ops_b.column_sweep(DT, 4)
    th.column_sweep(DT, 4)
    assert calls == list(range(len(calls))) and len(calls) >= 1
    assert rel_err(N(ops_a.enthalpy_state.E), th.E) < STOL
    assert rel_err(N(ops_b.enthalpy_state.E), th.E) < STOL


@pytest.mark.parametrize("ny,nx,nz,seed", CASES)
def test_diagnostics_and_rate_factor(ny, nx, nz, seed):
    print(f"diagnostics {ny}x{nx}x{nz} seed={seed}", flush=True)
    th, ops = make_pair(enth_case(ny, nx, nz, seed), nz)
    assert rel_err(N(ops.get_temperature()), th.get_temperature()) < KTOL
    w_orc = th.get_water_content()
    assert np.max(np.abs(N(ops.get_water_content()) - w_orc)) < KTOL * max(np.max(w_orc), 1e-3)
    assert rel_err(N(ops.get_arrhenius_factor()), th.get_arrhenius_factor(3.0)) < KTOL


@pytest.mark.parametrize("ny,nx,nz,seed", CASES[:3])
def test_initialisation_from_temperature(ny, nx, nz, seed):
    print(f"initialise {ny}x{nx}x{nz} seed={seed}", flush=True)
    rng = np.random.default_rng(seed)
    C = enth_case(ny, nx, nz, seed)
    th, ops = make_pair(C, nz)
    T3 = 230.0 + 45.0 * rng.random((ny, nx, nz))   # partly above T_pmp: capped
    Ts = 240.0 + 40.0 * rng.random((ny, nx))       # partly above T_MELT: capped
    ops.initialize_from_temperature(T(T3))
    th.initialize_from_temperature(T3)
    assert rel_err(N(ops.enthalpy_state.E), th.E) < KTOL
    assert rel_err(N(ops.enthalpy_state.E_prev), th.E) < KTOL
    ops.initialize_from_temperature(T(Ts))         # 2-D field broadcast over the levels
    th.initialize_from_temperature(Ts)
    assert rel_err(N(ops.enthalpy_state.E), th.E) < KTOL
    ops.set_surface_enthalpy_from_temperature(T(Ts))
    th.set_surface_enthalpy_from_temperature(Ts)
    assert rel_err(N(ops.enthalpy_forcing.E_surface), th.E_surface) < KTOL


@pytest.mark.parametrize("n,nz,seed", [(16, 11, 0), (24, 41, 1), (13, 2, 2)])
def test_strain_heating(n, nz, seed):
    print(f"strain_heating n={n} nz={nz} seed={seed}", flush=True)
    from glide_amd.enthalpy import EnthalpyOperators
    from glide_amd.grid import Grid
    from oracle.oracle import strain_heating
    from tests.util import dome_fields
    F = dome_fields(n=n, seed=seed)
    g = Grid(n, n, DX, device="cuda")
    g.state.H.set(F["H"])
    g.state.u.set(F["u"])
    g.state.v.set(F["v"])
    g.rheology.B.set(F["B"])
    ops = EnthalpyOperators(g, nz=nz)
    ops.compute_strain_heating(917.0)
    phi = strain_heating(F["u"], F["v"], F["B"], 3.0, 1e-6, DX, 917.0)
    got = N(ops.enthalpy_forcing.phi_strain)
    for k in range(nz):
        assert rel_err(got[:, :, k], phi) < KTOL


def _thermal_pair(n, nz, seed, frictional=True, strain=False):
    """ThermalModel on a dome state after a momentum step, and the oracle."""
    from glide_amd.grid import Grid
    from glide_amd.model import ThermalModel
    from oracle.oracle import OracleThermal
    from tests.util import dome_fields
    F = dome_fields(n=n, seed=seed)
    ny, nx = F["H"].shape
    g = Grid(ny, nx, DX, device="cuda")
    g.state.H.set(F["H_prev"])
    g.state.u.set(F["u"])
    g.state.v.set(F["v"])
    g.sliding.beta.set(F["beta"])
    g.rheology.B.set(F["B"])
    tm = ThermalModel(g, nz=nz, n_smooth=10, frictional_heating=frictional, strain_heating=strain)
    rng = np.random.default_rng(seed + 3)
    Ts = 243.15 + 6.5e-3 * np.maximum(2000.0 - (F["bed"] + F["H"]), 0.0) * 0.2
    Qg = 0.05 + 0.01 * rng.random((ny, nx))
    tm.initialize(T(Ts), Q_geo=T(Qg))
    th = OracleThermal(ny, nx, nz, DX, F["H_prev"])
    th.initialize_from_temperature(Ts)
    th.set_surface_enthalpy_from_temperature(Ts)
    th.Q_geo = Qg.copy()
    # momentum step realised: H_prev snapshot, new H
    tm.pre_momentum()
    g.state.H.set(F["H"])
    return F, g, tm, th


@pytest.mark.parametrize("n,nz,seed,frictional,strain", [(16, 11, 0, True, False), (24, 41, 1, True, True),
                                                          (20, 21, 2, False, False), (18, 41, 3, False, True)])
def test_thermal_step(n, nz, seed, frictional, strain):
    print(f"thermal_step n={n} nz={nz} seed={seed} frictional={frictional} strain={strain}", flush=True)
    from oracle.oracle import strain_heating, thermal_step
    F, g, tm, th = _thermal_pair(n, nz, seed, frictional, strain)
    rep = tm.step(DT, report=True)
    phi = None
    if strain:
        phi = np.repeat(strain_heating(F["u"], F["v"], F["B"], 3.0, 1e-6, DX, 917.0)[:, :, None], nz, axis=2)
    B_orc = thermal_step(th, F["u"], F["v"], F["H"], F["beta"], DT, n_smooth=10, frictional_heating=frictional,
                         strain_heating=phi)
    ops = tm.ops
    assert rel_err(N(ops.enthalpy_velocity.omega), th.omega) < KTOL
    assert rel_err(N(ops.enthalpy_velocity.u3d), th.u3d) < KTOL
    if frictional:
        assert rel_err(N(ops.enthalpy_forcing.Q_fh), th.Q_fh) < KTOL
    assert rel_err(N(ops.f_E), th.f_E) < KTOL
    assert rel_err(N(ops.enthalpy_state.E), th.E) < STOL
    assert rel_err(N(g.rheology.B.data), B_orc) < STOL
    assert rep.sweeps >= 1


def test_red_black_column_sweep_reaches_the_same_fixed_point():
    """The red-black column ordering (no counterpart in the reference), with
    and without the layer passes, converges to the alternating sweep's fixed
    point (with them in no more sweeps)."""
    print("thermal_step red-black columns", flush=True)
    out = []
    for mc, layer in ((False, True), (True, True), (True, False)):
        F, g, tm, th = _thermal_pair(24, 41, 6, strain=True)
        tm.n_smooth = 200
        sc = tm.ops.smoother_config
        sc.relative_tolerance, sc.absolute_tolerance, sc.multicolour, sc.layer_sweeps = 1e-13, 0.0, mc, layer
        rep = tm.step(DT, report=True)
        out.append((N(tm.ops.enthalpy_state.E), N(g.rheology.B.data), rep.sweeps))
    E0, B0, n0 = out[0]
    print("  sweeps: alternating, red-black + layer, red-black:", [o[2] for o in out], flush=True)
    for E1, B1, n1 in out[1:]:
        assert rel_err(E1, E0) < 1e-10
        assert rel_err(B1, B0) < 1e-10
    assert out[1][2] <= n0


def test_thermal_step_host_path_matches_fused():
    print("thermal_step host path", flush=True)
    F, g, tm, th = _thermal_pair(16, 11, 4, strain=True)
    F2, g2, tm2, _ = _thermal_pair(16, 11, 4, strain=True)
    tm2.ops.smoother_config.hook_func = lambda i: None
    tm.step(DT)
    tm2.step(DT)
    assert rel_err(N(tm2.ops.enthalpy_state.E), N(tm.ops.enthalpy_state.E)) < STOL
    assert rel_err(N(g2.rheology.B.data), N(g.rheology.B.data)) < STOL


def test_thermal_step_deferred_rate_factor():
    """step(B_out=...) leaves grid.rheology.B as the momentum solve used it and
    returns the same new rate factor (and enthalpy) as the in-place step, also
    when run on a second stream."""
    print("thermal_step deferred B", flush=True)
    F, g, tm, _ = _thermal_pair(16, 21, 5, strain=True)
    F2, g2, tm2, _ = _thermal_pair(16, 21, 5, strain=True)
    B0 = g2.rheology.B.data.clone()
    B_out = torch.empty_like(B0)
    tm.step(DT)
    s = torch.cuda.Stream()
    ev = torch.cuda.Event()
    ev.record(torch.cuda.current_stream())
    s.wait_event(ev)
    with torch.cuda.stream(s):
        tm2.step(DT, B_out=B_out)
    torch.cuda.synchronize()
    assert torch.equal(g2.rheology.B.data, B0)
    assert torch.equal(B_out, g.rheology.B.data)
    assert torch.equal(tm2.ops.enthalpy_state.E, tm.ops.enthalpy_state.E)
    with pytest.raises(ValueError):
        tm2.step(DT, B_out=torch.empty(3, device="cuda", dtype=torch.float64))


# ---------------------------------------------------------------- adjoint
B_SCALE = 917.0 * 9.81 * (365.25 * 86400.0) ** (1.0 / 3.0)


@pytest.mark.parametrize("n,nz,seed", [(16, 21, 6), (23, 41, 7)])
def test_adjoint_depth_constant_velocity_path(n, nz, seed):
    """After a thermal step (whose velocity fields are the broadcast SSA
    velocity over seconds per year), the adjoint with the 2-D velocity
    (horizontal couplings computed in the sweeps) equals the one reading the
    stored coupling planes."""
    print(f"adjoint 2-D velocity n={n} nz={nz} seed={seed}", flush=True)
    F, g, tm, _ = _thermal_pair(n, nz, seed, strain=True)
    tm.step(DT)
    ops = tm.ops
    gE = ops.enthalpy_state.E.clone()
    gB = torch.ones_like(g.rheology.B.data)
    kw = dict(B_scale=B_SCALE, n_iter=200, relative_tolerance=1e-12, report=True)
    a = ops.adjoint_solve(DT, gE, gB, **kw)
    b = ops.adjoint_solve(DT, gE, gB, u2d=g.state.u.data, v2d=g.state.v.data, vel_div=365.25 * 86400.0, **kw)
    torch.cuda.synchronize()
    print(f"  sweeps {a['report'].sweeps} / {b['report'].sweeps}", flush=True)
    for k in ("lambda", "E_prev", "phi_strain", "Q_geo", "E_surface"):
        assert rel_err(N(b[k]), N(a[k])) < 1e-10, k


@pytest.mark.parametrize("ny,nx,nz,seed", CASES)
def test_adjoint_solve_matches_oracle(ny, nx, nz, seed):
    """GPU adjoint (iterative transposed column/layer passes) against the
    oracle's direct sparse solve of J^T lambda = g at the same converged E."""
    print(f"adjoint {ny}x{nx}x{nz} seed={seed}", flush=True)
    th, ops = make_pair(enth_case(ny, nx, nz, seed), nz)
    ops.smoother_config.relative_tolerance = 1e-15
    ops.smoother_config.absolute_tolerance = 0.0
    ops.column_sweep(DT, 200)
    th.E = np.ascontiguousarray(N(ops.enthalpy_state.E))
    rng = np.random.default_rng(seed + 11)
    w = rng.standard_normal((ny, nx, nz))
    gB = rng.standard_normal((ny, nx))
    out = ops.adjoint_solve(DT, T(w), T(gB), B_scale=B_SCALE, relative_tolerance=1e-14, n_iter=250, report=True)
    rep = out["report"]
    g = w + th.arrhenius_vjp(gB, 3.0, B_SCALE)
    lam = th.adjoint_solve(DT, g)
    grads = th.adjoint_gradients(DT, lam)
    print(f"  adjoint sweeps {rep.sweeps}, |r0| {rep.r0:.3e} -> {rep.r:.3e}", flush=True)
    assert rel_err(N(out["lambda"]), lam) < 1e-9
    for k in ("E_prev", "phi_strain", "Q_geo", "E_surface"):
        assert rel_err(N(out[k]), grads[k]) < 1e-9, k


def test_thermal_step_autograd_matches_finite_differences():
    """ThermalStep backward (adjoint of the converged solve) against central
    differences of complete forward steps in Q_geo, E_prev and E_surface,
    with the rate-factor output in the loss."""
    print("thermal step autograd", flush=True)
    from glide_amd.torch import ThermalStep
    F, g, tm, th = _thermal_pair(16, 11, 6)
    tm.n_smooth = 200
    cfg = tm.ops.smoother_config
    cfg.relative_tolerance, cfg.absolute_tolerance = 1e-15, 0.0
    ops = tm.ops
    E_prev0 = ops.enthalpy_state.E.clone()
    Q0 = ops.enthalpy_forcing.Q_geo.clone()
    Es0 = ops.enthalpy_forcing.E_surface.clone()
    rng = np.random.default_rng(4)
    w, wB = T(rng.standard_normal(tuple(E_prev0.shape))), T(rng.standard_normal(tuple(Q0.shape)))

    def loss(Ep, Q, Es):
        E, B = ThermalStep.apply(tm, DT, Ep, Q, Es)
        return (w * E).sum() + 1e-3 * (wB * B).sum()

    Ep, Q, Es = (t.clone().requires_grad_(True) for t in (E_prev0, Q0, Es0))
    L = loss(Ep, Q, Es)
    L.backward()
    H = N(g.state.H.data)
    ti, tj = [int(v) for v in np.argwhere(H > 500.0)[0]]
    checks = [("Q_geo", Q, Q0, (ti, tj), 1e-4), ("E_prev", Ep, E_prev0, (ti, tj, 3), 1e-5),
              ("E_surface", Es, Es0, (ti, tj), 1e-5)]
    for name, leaf, base, idx, h in checks:
        def L_at(delta):
            args = [E_prev0.clone(), Q0.clone(), Es0.clone()]
            pos = {"E_prev": 0, "Q_geo": 1, "E_surface": 2}[name]
            args[pos][idx] += delta
            with torch.no_grad():
                return float(loss(*args))
        fd = (L_at(h) - L_at(-h)) / (2 * h)
        ad = float(leaf.grad[idx])
        print(f"  {name}{idx}: adjoint {ad:.10e}  fd {fd:.10e}", flush=True)
        assert abs(fd - ad) <= 1e-5 * max(abs(fd), abs(ad)), name


# ------------------------------------------------------ elementary functions
def test_row_elementary_functions_within_a_few_ulp():
    """The exponential and logarithm the enthalpy rows use (written for the
    rows' argument ranges) against numpy's: exp over the rows' clamped
    arguments [-20, 20] and over [-700, 700], log over 1 + e^a of the
    drainage softplus (a in [-20, 10]), over [1e-300, 1e300] and at the
    reduction's breakpoints, within 8 ulp (the largest is printed)."""
    from glide_amd import _lib
    rng = np.random.default_rng(5)
    s = np.sqrt(0.5)
    xe = np.concatenate([rng.uniform(-20.0, 20.0, 200000), rng.uniform(-700.0, 700.0, 50000),
                         [0.0, -20.0, 20.0, 1e-300, -1e-18, np.log(2.0), -np.log(2.0) / 2]])
    xl = np.concatenate([1.0 + np.exp(rng.uniform(-20.0, 10.0, 200000)), 10.0 ** rng.uniform(-300.0, 300.0, 50000),
                         [1.0, 2.0, 0.5, s, np.nextafter(s, 0.0), np.nextafter(s, 1.0), 1.0 + 2.0 ** -52]])
    ulp = 2.0 ** -52
    for x, which in ((xe, "exp"), (xl, "log")):
        xt = torch.as_tensor(x, device="cuda")
        out = torch.empty_like(xt)
        args = (out, None) if which == "exp" else (None, out)
        _lib.call("glide_enth_elementary", _lib.ptr(xt), int(x.size), _lib.ptr(args[0]), _lib.ptr(args[1]),
                  _lib.stream_handle())
        torch.cuda.synchronize()
        want = np.exp(x) if which == "exp" else np.log(x)
        got = N(out)
        nz = want != 0.0
        err = np.abs(got - want)
        rel = np.max(err[nz] / (np.abs(want[nz]) * ulp))
        print(f"  {which}: largest error {rel:.2f} ulp over {x.size} values", flush=True)
        assert np.all(np.isfinite(got))
        assert rel <= 8.0
        assert np.all(err[~nz] == 0.0)
