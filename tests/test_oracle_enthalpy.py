"""Pins of the FP64 enthalpy oracle (CPU only).

Restates the reference's own analytic tests (tests/test_enthalpy_cold_column.py,
tests/test_enthalpy_advection.py, tests/test_enthalpy_stefan.py,
tests/test_enthalpy_horiz_advection.py) with their tolerances, plus closed
forms for the omega integration and the column Newton/Thomas solve.
"""
import numpy as np
import pytest
from scipy.optimize import brentq

from oracle import oracle as O

SEC_PER_KYR = 1000.0 * O.SEC_PER_YR
H_ICE = 1000.0


def slab(nz, T_s, Q, ny=4, nx=4, horiz=False):
    th = O.OracleThermal(ny, nx, nz, 1000.0, np.full((ny, nx), H_ICE))
    th.flags = (1 if horiz else 0) | 2 | 4
    th.initialize_from_temperature(T_s)
    th.set_surface_enthalpy_from_temperature(T_s)
    th.Q_geo[:] = Q
    return th


def test_cold_column_steady_state():
    th = slab(21, 243.15, 0.04)
    dt = SEC_PER_KYR
    for _ in range(200):
        th.set_rhs()
        th.column_sweep(dt, 20)
    E = th.E[2, 2] * O.E_SCALE
    Ea = O.C_I * (243.15 - O.T_REF) + O.C_I * 0.04 * H_ICE / O.K_I * (1 - th.sigma)
    err = np.max(np.abs(E[:-1] - Ea[:-1]) / np.abs(Ea[:-1]))
    assert err < 5e-3          # the reference's bar
    assert err < 1e-6          # FP64 steady state


def test_cold_column_residual_convergence():
    th = slab(21, 243.15, 0.04)
    dt = SEC_PER_KYR
    th.set_rhs()
    res = []
    for _ in range(50):
        res.append(th.compute_residual(dt))
        d = th.column_smooth(dt)
        th.E += th.omega_relax * d
    assert res[0] / res[-1] > 100


def _run_to_steady(th, conv_tol):
    dt, dt_max = 100.0 * SEC_PER_KYR, 1e6 * SEC_PER_KYR
    for step in range(1, 5001):
        E_old = th.E[2, 2].copy()
        th.set_rhs()
        th.column_sweep(dt, 20)
        rel = np.max(np.abs(th.E[2, 2] - E_old)) / (np.max(np.abs(th.E[2, 2])) + 1.0)
        if rel < 1e-2:
            dt = min(dt * 1.5, dt_max)
        if rel < conv_tol:
            return True
    return False


def test_constant_sigma_advection_exponential_profile():
    Pe = -5.0
    th = slab(41, 243.15, 0.04)
    th.drain_rate = 0.0
    th.omega[:] = Pe * O.K_COLD / (O.RHO_I_E * H_ICE)
    Es = O.C_I * (243.15 - O.T_REF)
    B = -H_ICE * 0.04 / (O.K_COLD * Pe)
    A = Es - B * np.exp(Pe)
    Ea = A + B * np.exp(Pe * th.sigma)
    assert _run_to_steady(th, 1e-6)
    E = th.E[2, 2] * O.E_SCALE
    assert np.max(np.abs(E[:-1] - Ea[:-1])) / np.max(np.abs(Ea[:-1])) < 0.02


def _E_pmp_bed():
    return O.C_I * (O.T_MELT - O.BETA_CC * O.RHO_I_E * O.GRAVITY * H_ICE - O.T_REF)


def test_temperate_bed_residual_is_the_geothermal_flux():
    th = slab(21, 223.15, 0.5)
    th.drain_rate = 0.0
    th.E[:] = (_E_pmp_bed() + 1000.0) / O.E_SCALE
    th.E_prev[:] = th.E
    th.set_rhs()
    th.compute_residual(SEC_PER_KYR)
    dsig_half = 0.5 / 20
    expected = -0.5 / (dsig_half * O.E_SCALE)
    assert np.isclose(th.r_E[2, 2, 0], expected, rtol=1e-3, atol=1e-2)


def test_temperate_bed_is_not_clamped():
    th = slab(21, 223.15, 0.5)
    th.drain_rate = 0.0
    th.E[:] = (_E_pmp_bed() + 1000.0) / O.E_SCALE
    th.E_prev[:] = th.E
    th.set_rhs()
    th.column_sweep(SEC_PER_KYR, 5)
    assert th.E[0, 0, 0] * O.E_SCALE > _E_pmp_bed()


def test_stefan_two_zone_steady_profile():
    th = slab(41, 223.15, 0.5)
    th.drain_rate = 0.0
    s = th.sigma
    T_pmp = lambda g: O.T_MELT - O.BETA_CC * O.RHO_I_E * O.GRAVITY * (1 - g) * H_ICE
    E_pmp = lambda g: O.C_I * (T_pmp(g) - O.T_REF)
    sig_star = brentq(lambda g: 0.5 * H_ICE * (1 - g) / O.K_I - (T_pmp(g) - 223.15), 0.0, 1.0 - 1e-10)
    Ea = np.where(s <= sig_star, E_pmp(sig_star) + 0.5 * H_ICE / O.K_TEMP * (sig_star - s),
                  O.C_I * (223.15 - O.T_REF) + 0.5 * H_ICE / O.K_COLD * (1 - s))
    Ta = np.minimum(Ea / O.C_I + O.T_REF, T_pmp(s))
    wa = np.maximum(Ea - E_pmp(s), 0.0) / O.L_HEAT
    assert _run_to_steady(th, 1e-5)
    E = th.E[2, 2] * O.E_SCALE
    T = th.get_temperature()[2, 2]
    w = th.get_water_content()[2, 2]
    assert np.max(np.abs(T[:-1] - Ta[:-1])) < 0.1
    assert np.max(np.abs(E[:-1] - Ea[:-1])) / np.max(np.abs(Ea[:-1])) < 0.02
    assert abs(w[0] - wa[0]) / wa[0] < 0.02


def _flat(ny, nx, nz):
    th = O.OracleThermal(ny, nx, nz, 1000.0, np.full((ny, nx), H_ICE))
    th.flags = 1
    th.drain_rate = 0.0
    return th


def test_horizontal_divergent_velocity_stays_bounded():
    th = _flat(8, 8, 9)
    th.initialize_from_temperature(248.15)
    th.set_surface_enthalpy_from_temperature(248.15)
    th.u3d[:] = (1e-10 * 1000.0 * np.arange(9))[None, None, :]
    th.set_rhs()
    th.column_sweep(100 * O.SEC_PER_YR, 5)
    # the reference only runs this case; here: finite and bounded
    assert np.all(np.isfinite(th.E))
    assert np.max(np.abs(th.E)) < 10.0


def test_horizontal_step_profile_keeps_y_symmetry():
    th = _flat(8, 16, 9)
    T = np.full((8, 16, 9), 240.0)
    T[:, 8:, :] = 260.0
    th.E = np.ascontiguousarray(O.C_I * (T - O.T_REF) / O.E_SCALE)
    th.E_prev = th.E.copy()
    th.set_surface_enthalpy_from_temperature(240.0)
    th.u3d[:] = 100.0 / O.SEC_PER_YR
    for _ in range(3):
        th.set_rhs()
        th.column_sweep(10 * O.SEC_PER_YR, 20)
    mid = th.E[:, :, 4]
    assert np.max(np.ptp(mid, axis=0)) < 1.0
    assert np.max(np.ptp(mid, axis=0)) < 1e-12     # exact symmetry in FP64


def test_horizontal_interior_energy_conservation():
    th = _flat(8, 16, 5)
    Tj = 250.0 + 10.0 * np.sin(2 * np.pi * np.arange(16) / 16)
    th.E = np.ascontiguousarray(np.broadcast_to((O.C_I * (Tj - O.T_REF) / O.E_SCALE)[None, :, None],
                                                (8, 16, 5)).copy())
    th.E_prev = th.E.copy()
    th.set_surface_enthalpy_from_temperature(250.0)
    th.u3d[:] = 50.0 / O.SEC_PER_YR
    before = np.sum(H_ICE * th.E[2:-2, 2:-2, :-1])
    th.set_rhs()
    th.column_sweep(O.SEC_PER_YR, 20)
    after = np.sum(H_ICE * th.E[2:-2, 2:-2, :-1])
    assert abs(after - before) / abs(before) < 0.01


def test_column_newton_step_solves_the_linear_column_system():
    """With horizontal coupling off, cold ice (K constant to 1e-9) and no
    drainage, one Newton/Thomas step with unit relaxation zeroes the column
    residual: pins the tridiagonal rows and the Thomas solve together."""
    th = slab(17, 240.0, 0.06)
    th.drain_rate = 0.0
    rng = np.random.default_rng(0)
    th.E[:, :, :-1] += 0.01 * rng.standard_normal((4, 4, 16))
    th.omega[:] = -1e-9 * rng.random((4, 4, 17))
    th.set_rhs()
    th.n_newton = 1
    r_before = th.compute_residual(10 * O.SEC_PER_YR)
    th.E += th.column_smooth(10 * O.SEC_PER_YR)
    r_after = th.compute_residual(10 * O.SEC_PER_YR)
    assert r_after < 1e-7 * r_before


def test_thomas_matches_banded_solver():
    """The column smoother's tridiagonal system (extracted by one Newton step
    from a zero-residual-shifted state) equals scipy.linalg.solve_banded."""
    from scipy.linalg import solve_banded
    th = slab(9, 245.0, 0.05, ny=2, nx=2)
    th.drain_rate = 0.0
    th.set_rhs()
    dt = 50 * O.SEC_PER_YR
    th.compute_residual(dt)
    r = th.r_E[1, 1].copy()
    d = th.column_smooth(dt)[1, 1]
    # rebuild the tridiagonal matrix by probing the linear residual
    nz = 9
    J = np.zeros((nz, nz))
    base = th.E.copy()
    for k in range(nz):
        h = 1e-6
        th.E = base.copy()
        th.E[:, :, k] += h
        th.compute_residual(dt)
        rp = th.r_E[1, 1].copy()
        J[:, k] = (rp - r) / h
    th.E = base
    J[-1, :] = 0.0
    J[-1, -1] = 1.0
    rhs = -r.copy()
    rhs[-1] = -(base[1, 1, -1] - th.E_surface[1, 1])
    ab = np.zeros((3, nz))
    ab[0, 1:] = np.diag(J, 1)
    ab[1] = np.diag(J)
    ab[2, :-1] = np.diag(J, -1)
    ref = solve_banded((1, 1), ab, rhs)
    assert np.allclose(d, ref, rtol=1e-5, atol=1e-9 * np.max(np.abs(ref)))


def test_omega_closed_form_uniform_divergence():
    ny, nx, nz = 6, 8, 11
    th = O.OracleThermal(ny, nx, nz, 500.0, np.full((ny, nx), 800.0))
    a = 2e-3
    th.u3d[:] = (a * 500.0 * np.arange(nx + 1))[None, None, :]
    dh = 0.3
    bmb = np.full((ny, nx), 0.05)
    om = th.compute_omega(dh, bmb)
    # interior columns: div(Hu) = H a exactly (equal thickness, no LF dissipation)
    ref = -0.05 - th.sigma * (dh + 800.0 * a)
    assert np.allclose(om[2:-2, 2:-2], ref[None, None, :], rtol=1e-12, atol=1e-12)
    thin = O.OracleThermal(2, 2, 5, 1.0, np.zeros((2, 2)))
    assert np.all(thin.compute_omega(1.0) == 0.0)


def test_thin_columns_are_driven_to_surface_enthalpy():
    th = O.OracleThermal(3, 3, 6, 1000.0, np.array([[10.0, 500.0, 500.0]] * 3))
    th.initialize_from_temperature(250.0)
    th.set_surface_enthalpy_from_temperature(230.0)
    th.set_rhs()
    d = th.column_smooth(O.SEC_PER_YR)
    E_pmp = O.C_I * (th.T_pmp()[0, 0] - O.T_REF) / O.E_SCALE
    assert np.allclose(th.E[0, 0] + d[0, 0], np.minimum(th.E_surface[0, 0], E_pmp))
    th.compute_residual(O.SEC_PER_YR)
    assert np.all(th.r_E[:, 0] == 0.0)


def test_arrhenius_factor_closed_form():
    th = O.OracleThermal(2, 2, 5, 1.0, np.full((2, 2), 100.0))
    th.initialize_from_temperature(250.0)
    B = th.get_arrhenius_factor(3.0)
    A = 3.985e-13 * np.exp(-60e3 / (8.314 * 250.0))
    assert np.allclose(B, A ** (-1.0 / 3.0), rtol=1e-12)


def test_strain_heating_closed_forms(orc):
    """SSA strain heating Phi = 4 eta eps_e^2 rho_i g / a with
    eta = B/2 (eps_e^2 + eps_reg)^((1-n)/(2n)): uniform stretching u = a x gives
    eps_e^2 = a^2 in every cell; simple shear u = c y gives eps_e^2 = c^2/4 in
    interior cells (four unmasked corners) and Phi = 0 for rigid translation."""
    ny, nx, dx, n, eps = 12, 10, 1000.0, 3.0, 1e-6
    sec = 365.25 * 86400.0
    rho_g = 917.0 * 9.81
    B = np.full((ny, nx), 0.5) + 0.1 * np.arange(nx)[None, :]
    a = 2e-3
    x_u = dx * np.arange(nx + 1)
    u = np.broadcast_to(a * x_u, (ny, nx + 1)).copy()
    v = np.zeros((ny + 1, nx))
    phi = orc.strain_heating(u, v, B, n, eps, dx)
    p = (1 - n) / (2 * n)
    expect = 4.0 * (0.5 * B * (a * a + eps) ** p) * a * a * rho_g / sec
    np.testing.assert_allclose(phi, expect, rtol=1e-13)
    c = 1.5e-3
    y_u = -dx * np.arange(ny)             # rows run towards -y
    u = np.broadcast_to((c * y_u)[:, None], (ny, nx + 1)).copy()
    phi = orc.strain_heating(u, v, B, n, eps, dx)
    e2 = c * c / 4.0
    expect = 4.0 * (0.5 * B * (e2 + eps) ** p) * e2 * rho_g / sec
    np.testing.assert_allclose(phi[1:-1, 1:-1], expect[1:-1, 1:-1], rtol=1e-12)
    assert np.all(phi[0, 1:-1] < expect[0, 1:-1])   # boundary rows lose masked corners
    u = np.full((ny, nx + 1), 37.0)
    v = np.full((ny + 1, nx), -12.0)
    assert np.max(np.abs(orc.strain_heating(u, v, B, n, eps, dx))) == 0.0


# ---------------------------------------------------------------- adjoint pins
def _adjoint_case(orc, ny=6, nx=5, nz=9, seed=3):
    from tests.util import enth_case
    C = enth_case(ny, nx, nz, seed)
    th = orc.OracleThermal(ny, nx, nz, 1500.0, C["H"])
    for k in ("E", "E_prev", "u3d", "v3d", "omega", "E_surface", "Q_geo", "Q_fh", "phi_strain", "H_prev"):
        setattr(th, k, np.ascontiguousarray(C[k], dtype=np.float64).copy())
    th.set_rhs(snapshot=False)
    return th


DT_ADJ = 10.0 * 365.25 * 86400.0


def test_exact_jacobian_matches_central_differences(orc):
    """dG/dE rows (diffusivity and drainage derivatives included) against
    central differences of the system residual in a random direction."""
    th = _adjoint_case(orc)
    rng = np.random.default_rng(0)
    J = th.jacobian(DT_ADJ)
    E0 = th.E.copy()
    for trial in range(3):
        v = rng.standard_normal(E0.shape)
        eps = 1e-6
        th.E = E0 + eps * v
        gp = th.system_residual(DT_ADJ)
        th.E = E0 - eps * v
        gm = th.system_residual(DT_ADJ)
        th.E = E0.copy()
        fd = (gp - gm) / (2 * eps)
        jv = (J @ v.ravel()).reshape(E0.shape)
        assert np.max(np.abs(fd - jv)) < 1e-6 * np.max(np.abs(jv)), trial


def _solve(th, n_iter=400):
    th.rtol, th.atol = 1e-15, 0.0
    th.column_sweep(DT_ADJ, n_iter)
    return th


def test_adjoint_gradients_match_finite_differences_of_the_solve(orc):
    """Gradients of L = sum(w * E) through the converged solve: the adjoint
    (J^T lambda = w) against central differences of complete solves in
    E_prev, Q_geo, phi_strain and E_surface (thick and thin columns)."""
    base = _adjoint_case(orc, seed=5)
    E_init = base.E.copy()
    rng = np.random.default_rng(1)
    w = rng.standard_normal(E_init.shape)
    th = _solve(_adjoint_case(orc, seed=5))
    G = th.system_residual(DT_ADJ)
    assert np.max(np.abs(G)) < 1e-10
    lam = th.adjoint_solve(DT_ADJ, w)
    grads = th.adjoint_gradients(DT_ADJ, lam)

    def L_after(mutate):
        t = _adjoint_case(orc, seed=5)
        mutate(t)
        t.set_rhs(snapshot=False)
        t.E = E_init.copy()
        _solve(t)
        return float(np.sum(w * t.E))

    thin = np.argwhere(th.H < th.h_thin)
    thick = np.argwhere(th.H >= th.h_thin)
    i, j = thick[0]
    checks = []
    for name, (ii, jj, kk), size in (("E_prev", (i, j, 2), 1e-3), ("phi_strain", (i, j, 3), 1e-9),
                                    ("Q_geo", (i, j, None), 1e-3), ("E_surface", (i, j, None), 1e-3)):
        def mut(t, s, name=name, ii=ii, jj=jj, kk=kk):
            arr = getattr(t, name)
            if kk is None:
                arr[ii, jj] += s
            else:
                arr[ii, jj, kk] += s
        fd = (L_after(lambda t: mut(t, size)) - L_after(lambda t: mut(t, -size))) / (2 * size)
        ad = grads[name][ii, jj] if kk is None else grads[name][ii, jj, kk]
        checks.append((name, fd, ad))
        assert abs(fd - ad) <= 1e-5 * max(abs(fd), abs(ad), 1e-30), (name, fd, ad)
    if len(thin):
        ti, tj = thin[0]
        fd = (L_after(lambda t: t.E_surface.__setitem__((ti, tj), t.E_surface[ti, tj] + 1e-3))
              - L_after(lambda t: t.E_surface.__setitem__((ti, tj), t.E_surface[ti, tj] - 1e-3))) / 2e-3
        assert abs(fd - grads["E_surface"][ti, tj]) <= 1e-5 * max(abs(fd), 1e-30)


def test_arrhenius_vjp_matches_finite_differences(orc):
    th = _adjoint_case(orc, seed=7)
    rng = np.random.default_rng(2)
    gB = rng.standard_normal((th.ny, th.nx))
    scale = 917.0 * 9.81 * (365.25 * 86400.0) ** (1.0 / 3.0)
    gE = th.arrhenius_vjp(gB, 3.0, scale)
    v = rng.standard_normal(th.E.shape)
    E0 = th.E.copy()
    eps = 1e-7
    th.E = E0 + eps * v
    Bp = th.get_arrhenius_factor(3.0) / scale
    th.E = E0 - eps * v
    Bm = th.get_arrhenius_factor(3.0) / scale
    th.E = E0
    fd = np.sum(gB * (Bp - Bm)) / (2 * eps)
    assert abs(fd - np.sum(gE * v)) <= 1e-6 * abs(fd)
