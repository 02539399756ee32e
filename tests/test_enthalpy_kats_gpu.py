"""The reference's analytic enthalpy tests on the GPU path, through the same
operator API the reference's tests drive (EnthalpyOperators): the cold column
and its residual decay (tests/test_enthalpy_cold_column.py), the constant
sigma-advection exponential profile (tests/test_enthalpy_advection.py), the
temperate bed and the two-zone Stefan profile (tests/test_enthalpy_stefan.py)
and horizontal advection (tests/test_enthalpy_horiz_advection.py), each with the
reference's bar and, where FP64 allows, a tighter one.  The same cases pin the
oracle on the CPU (tests/test_oracle_enthalpy.py)."""
import numpy as np
import pytest
import torch
from scipy.optimize import brentq

pytestmark = pytest.mark.gpu

H_ICE = 1000.0


def _ops(nz, ny=4, nx=4, horizontal=False):
    from glide_amd.enthalpy import EnthalpyOperators
    from glide_amd.grid import Grid
    g = Grid(ny, nx, 1000.0, device="cuda")
    g.state.H.set(np.full((ny, nx), H_ICE))
    g.state.H_prev.set(np.full((ny, nx), H_ICE))
    g.geometry.bed.set(np.zeros((ny, nx)))
    g.forcing.smb.set(np.zeros((ny, nx)))
    g.sliding.beta.set(np.ones((ny, nx)))
    g.sliding.m.set(1.0)
    g.sliding.u_reg.set(1.0)
    ops = EnthalpyOperators(g, nz=nz)
    ops.term_flags.horizontal_advection = horizontal
    return ops


def _slab(nz, T_s, Q):
    ops = _ops(nz)
    ops.initialize_from_temperature(T_s)
    ops.set_surface_enthalpy_from_temperature(np.full((4, 4), T_s))
    ops.enthalpy_forcing.Q_geo[:, :] = Q
    return ops


def _np(t):
    return t.detach().cpu().numpy()


def _sec_per_kyr():
    from glide_amd.enthalpy import SEC_PER_YR
    return 1000.0 * SEC_PER_YR


def test_cold_column_steady_state():
    from glide_amd import enthalpy as en
    print("enthalpy KAT: cold column", flush=True)
    ops = _slab(21, 243.15, 0.04)
    dt = _sec_per_kyr()
    for _ in range(200):
        ops.set_rhs()
        ops.column_sweep(dt, 20)
    E = _np(ops.enthalpy_state.E[2, 2]) * en.E_SCALE
    s = _np(ops.sigma)
    Ea = en.C_I * (243.15 - en.T_REF) + en.C_I * 0.04 * H_ICE / en.K_I * (1 - s)
    err = np.max(np.abs(E[:-1] - Ea[:-1]) / np.abs(Ea[:-1]))
    print(f"  max relative error {err:.3e}", flush=True)
    assert err < 5e-3  # the reference's bar
    assert err < 1e-6  # FP64 steady state


def test_cold_column_residual_decreases():
    print("enthalpy KAT: cold column residual", flush=True)
    ops = _slab(21, 243.15, 0.04)
    dt = _sec_per_kyr()
    ops.set_rhs()
    res = []
    for _ in range(50):
        res.append(float(ops.compute_residual(dt)))
        ops.column_smooth(dt)
        ops.enthalpy_state.E[:] += ops.smoother_config.omega * ops.delta_E
    print(f"  residual {res[0]:.3e} -> {res[-1]:.3e}", flush=True)
    assert res[0] / res[-1] > 100


def _run_to_steady(ops, conv_tol):
    dt, dt_max = 100.0 * _sec_per_kyr(), 1e6 * _sec_per_kyr()
    for _ in range(5000):
        E_old = _np(ops.enthalpy_state.E[2, 2]).copy()
        ops.set_rhs()
        ops.column_sweep(dt, 20)
        E_new = _np(ops.enthalpy_state.E[2, 2])
        rel = np.max(np.abs(E_new - E_old)) / (np.max(np.abs(E_new)) + 1.0)
        if rel < 1e-2:
            dt = min(dt * 1.5, dt_max)
        if rel < conv_tol:
            return True
    return False


def test_constant_sigma_advection_exponential_profile():
    from glide_amd import enthalpy as en
    print("enthalpy KAT: Pe = -5 advection-diffusion", flush=True)
    Pe = -5.0
    ops = _slab(41, 243.15, 0.04)
    ops.enthalpy_forcing.drain_rate.value = 0.0
    ops.enthalpy_velocity.omega.fill_(Pe * en.K_COLD / (en.RHO_I * H_ICE))
    s = _np(ops.sigma)
    Es = en.C_I * (243.15 - en.T_REF)
    B = -H_ICE * 0.04 / (en.K_COLD * Pe)
    Ea = (Es - B * np.exp(Pe)) + B * np.exp(Pe * s)
    assert _run_to_steady(ops, 1e-6)
    E = _np(ops.enthalpy_state.E[2, 2]) * en.E_SCALE
    err = np.max(np.abs(E[:-1] - Ea[:-1])) / np.max(np.abs(Ea[:-1]))
    print(f"  max relative error {err:.3e}", flush=True)
    assert err < 0.02  # the reference's bar (upwind discretisation error)


def _E_pmp_bed():
    from glide_amd import enthalpy as en
    return en.C_I * (en.T_MELT - en.BETA_CC * en.RHO_I * en.GRAVITY * H_ICE - en.T_REF)


def test_temperate_bed_residual_is_the_geothermal_flux():
    from glide_amd import enthalpy as en
    print("enthalpy KAT: temperate bed flux", flush=True)
    ops = _slab(21, 223.15, 0.5)
    ops.enthalpy_forcing.drain_rate.value = 0.0
    ops.enthalpy_state.E.fill_((_E_pmp_bed() + 1000.0) / en.E_SCALE)
    ops.enthalpy_state.E_prev.copy_(ops.enthalpy_state.E)
    ops.set_rhs()
    ops.compute_residual(_sec_per_kyr())
    expected = -0.5 / ((0.5 / 20) * en.E_SCALE)
    got = float(ops.r_E[2, 2, 0])
    print(f"  bed residual {got:.6e} expected {expected:.6e}", flush=True)
    assert np.isclose(got, expected, rtol=1e-3, atol=1e-2)


def test_temperate_bed_is_not_clamped():
    from glide_amd import enthalpy as en
    print("enthalpy KAT: temperate bed not clamped", flush=True)
    ops = _slab(21, 223.15, 0.5)
    ops.enthalpy_forcing.drain_rate.value = 0.0
    ops.enthalpy_state.E.fill_((_E_pmp_bed() + 1000.0) / en.E_SCALE)
    ops.enthalpy_state.E_prev.copy_(ops.enthalpy_state.E)
    ops.set_rhs()
    ops.column_sweep(_sec_per_kyr(), 5)
    assert float(ops.enthalpy_state.E[0, 0, 0]) * en.E_SCALE > _E_pmp_bed()


def test_stefan_two_zone_steady_profile():
    from glide_amd import enthalpy as en
    print("enthalpy KAT: Stefan two-zone profile", flush=True)
    ops = _slab(41, 223.15, 0.5)
    ops.enthalpy_forcing.drain_rate.value = 0.0
    s = _np(ops.sigma)

    def T_pmp(g):
        return en.T_MELT - en.BETA_CC * en.RHO_I * en.GRAVITY * (1 - g) * H_ICE

    def E_pmp(g):
        return en.C_I * (T_pmp(g) - en.T_REF)

    sig_star = brentq(lambda g: 0.5 * H_ICE * (1 - g) / en.K_I - (T_pmp(g) - 223.15), 0.0, 1.0 - 1e-10)
    Ea = np.where(s <= sig_star, E_pmp(sig_star) + 0.5 * H_ICE / en.K_TEMP * (sig_star - s),
                  en.C_I * (223.15 - en.T_REF) + 0.5 * H_ICE / en.K_COLD * (1 - s))
    Ta = np.minimum(Ea / en.C_I + en.T_REF, T_pmp(s))
    wa = np.maximum(Ea - E_pmp(s), 0.0) / en.L_HEAT
    assert _run_to_steady(ops, 1e-5)
    E = _np(ops.enthalpy_state.E[2, 2]) * en.E_SCALE
    T = _np(ops.get_temperature()[2, 2])
    w = _np(ops.get_water_content()[2, 2])
    print(f"  max |T - Ta| {np.max(np.abs(T[:-1] - Ta[:-1])):.3e} K, bed water {w[0]:.4e} vs {wa[0]:.4e}",
          flush=True)
    assert np.max(np.abs(T[:-1] - Ta[:-1])) < 0.1
    assert np.max(np.abs(E[:-1] - Ea[:-1])) / np.max(np.abs(Ea[:-1])) < 0.02
    assert abs(w[0] - wa[0]) / wa[0] < 0.02


def _flat(ny, nx, nz):
    ops = _ops(nz, ny, nx, horizontal=True)
    ops.term_flags.omega = False
    ops.term_flags.drainage = False
    ops.enthalpy_forcing.drain_rate.value = 0.0
    return ops


def test_horizontal_divergent_velocity_stays_bounded():
    from glide_amd.enthalpy import SEC_PER_YR
    print("enthalpy KAT: divergent horizontal velocity", flush=True)
    ops = _flat(8, 8, 9)
    ops.initialize_from_temperature(248.15)
    ops.set_surface_enthalpy_from_temperature(np.full((8, 8), 248.15))
    ops.enthalpy_velocity.u3d[:] = torch.as_tensor(1e-10 * 1000.0 * np.arange(9), device="cuda")[None, None, :]
    ops.set_rhs()
    ops.column_sweep(100 * SEC_PER_YR, 5)
    E = _np(ops.enthalpy_state.E)
    assert np.all(np.isfinite(E))
    assert np.max(np.abs(E)) < 10.0


def test_horizontal_step_profile_keeps_y_symmetry():
    from glide_amd import enthalpy as en
    print("enthalpy KAT: step profile y symmetry", flush=True)
    ops = _flat(8, 16, 9)
    T = np.full((8, 16, 9), 240.0)
    T[:, 8:, :] = 260.0
    ops.enthalpy_state.E[:] = torch.as_tensor(en.C_I * (T - en.T_REF) / en.E_SCALE, device="cuda")
    ops.enthalpy_state.E_prev.copy_(ops.enthalpy_state.E)
    ops.set_surface_enthalpy_from_temperature(np.full((8, 16), 240.0))
    ops.enthalpy_velocity.u3d.fill_(100.0 / en.SEC_PER_YR)
    for _ in range(3):
        ops.set_rhs()
        ops.column_sweep(10 * en.SEC_PER_YR, 20)
    mid = _np(ops.enthalpy_state.E[:, :, 4])
    spread = np.max(np.ptp(mid, axis=0))
    print(f"  largest spread across rows {spread:.3e}", flush=True)
    assert spread < 1.0      # the reference's bar
    assert spread < 1e-12    # exact symmetry in FP64


def test_horizontal_interior_energy_conservation():
    from glide_amd import enthalpy as en
    print("enthalpy KAT: interior energy conservation", flush=True)
    ops = _flat(8, 16, 5)
    Tj = 250.0 + 10.0 * np.sin(2 * np.pi * np.arange(16) / 16)
    E0 = np.broadcast_to((en.C_I * (Tj - en.T_REF) / en.E_SCALE)[None, :, None], (8, 16, 5))
    ops.enthalpy_state.E[:] = torch.as_tensor(np.ascontiguousarray(E0), device="cuda")
    ops.enthalpy_state.E_prev.copy_(ops.enthalpy_state.E)
    ops.set_surface_enthalpy_from_temperature(np.full((8, 16), 250.0))
    ops.enthalpy_velocity.u3d.fill_(50.0 / en.SEC_PER_YR)
    before = np.sum(H_ICE * _np(ops.enthalpy_state.E)[2:-2, 2:-2, :-1])
    ops.set_rhs()
    ops.column_sweep(en.SEC_PER_YR, 20)
    after = np.sum(H_ICE * _np(ops.enthalpy_state.E)[2:-2, 2:-2, :-1])
    print(f"  relative change of interior energy {abs(after - before) / abs(before):.3e}", flush=True)
    assert abs(after - before) / abs(before) < 0.01


def test_conservative_balance_uses_the_pre_momentum_thickness():
    """tests/test_conservative_balance.py: uniform E on a column thinned by
    1 m in a year leaves a smaller residual with the pre-momentum H_prev than
    with the thinned one.  The script derives omega with a compute_omega_ssa
    that the reference's EnthalpyOperators no longer defines; omega here comes
    from compute_omega with the realised thinning rate."""
    from glide_amd.enthalpy import EnthalpyOperators
    from glide_amd.grid import Grid
    print("enthalpy KAT: conservative balance", flush=True)
    ny, nx, nz = 4, 4, 11
    H_old, H_new = 1000.0, 999.0
    dt = 365.25 * 86400.0
    smb_si = -1.0 / dt
    g = Grid(ny, nx, 5000.0, device="cuda")
    g.state.H.set(np.full((ny, nx), H_new))
    g.state.H_prev.set(np.full((ny, nx), H_new))
    g.geometry.bed.set(np.zeros((ny, nx)))
    g.forcing.smb.set(np.zeros((ny, nx)))
    g.sliding.beta.set(np.ones((ny, nx)))
    g.sliding.m.set(1.0)
    g.sliding.u_reg.set(1.0)
    ops = EnthalpyOperators(g, nz=nz)
    ops.term_flags.horizontal_advection = False
    ops.term_flags.drainage = False
    ops.initialize_from_temperature(253.15)
    ops.set_surface_enthalpy_from_temperature(np.full((ny, nx), 253.15))
    ops.enthalpy_state.E_prev.copy_(ops.enthalpy_state.E)
    ops.compute_omega(np.full((ny, nx), smb_si))
    ops.H_prev.fill_(H_old)
    r_correct = float(ops.compute_residual(dt))
    ops.H_prev.fill_(H_new)
    r_wrong = float(ops.compute_residual(dt))
    print(f"  |r| with pre-momentum H_prev {r_correct:.3e}, with the thinned one {r_wrong:.3e}", flush=True)
    assert r_correct < r_wrong
