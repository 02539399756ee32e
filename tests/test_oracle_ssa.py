"""Pins of the FP64 CPU oracle of the SSA path (CPU only).

The reference ships no golden vectors for this path (its tests need CuPy on an
NVIDIA GPU), so the oracle is pinned by closed forms, invariants, independent
numpy/scipy re-computations, finite differences and a dot-product test —
each chosen so that a dropped term, a wrong sign or a transposed index fails.
"""
import numpy as np
import pytest

from tests.util import dome_fields, rel_err

DX, DT = 1500.0, 10.0


def test_lu5_matches_numpy(orc):
    rng = np.random.default_rng(0)
    for _ in range(20):
        A = rng.standard_normal((5, 5)) + 8.0 * np.eye(5)
        b = rng.standard_normal(5)
        assert np.allclose(orc.lu5(A, b), np.linalg.solve(A, b), rtol=1e-13, atol=1e-14)


def test_restrictions_match_numpy(orc):
    rng = np.random.default_rng(1)
    nyc, nxc = 6, 9
    fu = rng.random((2 * nyc, 2 * nxc + 1))
    fv = rng.random((2 * nyc + 1, 2 * nxc))
    fc = rng.random((2 * nyc, 2 * nxc))
    # u-facets: mean of the two fine faces stacked in y on every other column
    ref_u = 0.5 * (fu[0::2, 0::2] + fu[1::2, 0::2])
    ref_v = 0.5 * (fv[0::2, 0::2] + fv[0::2, 1::2])
    blocks = fc.reshape(nyc, 2, nxc, 2)
    assert np.allclose(orc.restrict_vfacet(fu), ref_u, rtol=0, atol=1e-15)
    assert np.allclose(orc.restrict_hfacet(fv), ref_v, rtol=0, atol=1e-15)
    assert np.allclose(orc.restrict_cell(fc), blocks.mean(axis=(1, 3)), rtol=0, atol=1e-15)
    assert np.array_equal(orc.restrict_cell(fc, use_max=True), blocks.max(axis=(1, 3)))


def _clamped_bilinear(c, ys, xs):
    """Independent bilinear interpolation of c at fractional coordinates
    (ys, xs) clamped to the array, via scipy."""
    from scipy.interpolate import RegularGridInterpolator
    ny, nx = c.shape
    f = RegularGridInterpolator((np.arange(ny), np.arange(nx)), c, method="linear")
    Y, X = np.meshgrid(np.clip(ys, 0, ny - 1), np.clip(xs, 0, nx - 1), indexing="ij")
    return f(np.stack([Y.ravel(), X.ravel()], axis=-1)).reshape(Y.shape)


def test_prolongations_match_scipy(orc):
    rng = np.random.default_rng(2)
    nyc, nxc = 5, 7
    nyf, nxf = 2 * nyc, 2 * nxc
    cu = rng.random((nyc, nxc + 1))
    ref = _clamped_bilinear(cu, (np.arange(nyf) - 0.5) / 2, np.arange(nxf + 1) / 2)
    assert np.allclose(orc.prolong_vfacet_bilinear(cu, nyf, nxf), ref, atol=1e-14)
    cv = rng.random((nyc + 1, nxc))
    ref = _clamped_bilinear(cv, np.arange(nyf + 1) / 2, (np.arange(nxf) - 0.5) / 2)
    assert np.allclose(orc.prolong_hfacet_bilinear(cv, nyf, nxf), ref, atol=1e-14)
    cc = rng.random((nyc, nxc))
    ref = _clamped_bilinear(cc, (np.arange(nyf) - 0.5) / 2, (np.arange(nxf) - 0.5) / 2)
    assert np.allclose(orc.prolong_cell(cc, nyf, nxf, bilinear=True), ref, atol=1e-14)
    assert np.array_equal(orc.prolong_cell(cc, nyf, nxf), np.kron(cc, np.ones((2, 2))))


def _uniform_strain(ny=12, nx=14, a=2e-3, b=-1.5e-3, H=800.0):
    """u = a x, v = b * (height above the bottom face): exx = a, eyy = b, exy = 0."""
    F = dome_fields(ny=ny, nx=nx, seed=0, with_velocity=False)
    F["u"] = np.tile(a * DX * np.arange(nx + 1), (ny, 1))
    F["v"] = np.tile((b * DX * (ny - np.arange(ny + 1)))[:, None], (1, nx))
    F["H"] = np.full((ny, nx), H)
    F["bed"] = np.zeros((ny, nx))
    F["beta"] = np.zeros((ny, nx))
    F["phi"] = np.ones((ny, nx))
    F["mask"] = np.zeros((ny, nx))
    F["f_H"] = F["H"] / DT
    return F


def test_viscosity_closed_form(orc):
    F = _uniform_strain()
    a, b = 2e-3, -1.5e-3
    eta = orc.viscosity(F["u"], F["v"], F["B"], 3.0, 1e-6, DX)
    ref = 0.5 * F["B"] * (a * a + b * b + a * b + 1e-6) ** (-1.0 / 3.0)
    assert rel_err(eta, ref) < 1e-13


def test_residual_closed_form_uniform_strain(orc):
    F = _uniform_strain()
    a, b, H = 2e-3, -1.5e-3, 800.0
    P = orc.phys()
    ru, rv, rH = orc.residual(F, P, DX, DT, use_forcing=False, use_mask=False)
    ny, nx = F["H"].shape
    # interior momentum rows: stresses are uniform, only the water drag remains
    assert rel_err(ru[1:-1, 1:-1], -P.water_drag * F["u"][1:-1, 1:-1]) < 1e-9
    assert rel_err(rv[1:-1, 1:-1], -P.water_drag * F["v"][1:-1, 1:-1]) < 1e-9
    # Dirichlet rows return the facet velocity
    assert np.array_equal(ru[:, 0], F["u"][:, 0]) and np.array_equal(ru[:, -1], F["u"][:, -1])
    assert np.array_equal(rv[0], F["v"][0]) and np.array_equal(rv[-1], F["v"][-1])
    # thickness rows away from the edge: H/dt + H (a + b)
    assert rel_err(rH[1:-1, 1:-1], H / DT + H * (a + b) * np.ones((ny - 2, nx - 2))) < 1e-12


def test_mass_residual_telescopes(orc):
    F = dome_fields(n=20, seed=3)
    P = orc.phys()
    _, _, rH = orc.residual(F, P, DX, DT, use_forcing=False, use_mask=False)
    div = rH - F["H"] / DT
    assert abs(div.sum()) < 1e-10 * np.abs(div).sum()


def test_local_system_matches_global_residual(orc):
    """Rows of every interior cell's local 5x5 system at the current state equal
    the rows of the global operator (two independent code paths)."""
    F = dome_fields(ny=10, nx=12, seed=4)
    P = orc.phys(m=1.0 / 3.0, calving_rate=0.3)
    Fu, Fv, FH = orc.residual(F, P, DX, DT, use_forcing=False, use_mask=False)
    ny, nx = F["H"].shape
    for i in range(1, ny - 1):
        for j in range(1, nx - 1):
            x = [F["u"][i, j], F["u"][i, j + 1], F["v"][i, j], F["v"][i + 1, j], F["H"][i, j]]
            _, r = orc.local_system(x, F, P, DX, DT, i, j)
            ref = [Fu[i, j], Fu[i, j + 1], Fv[i, j], Fv[i + 1, j], FH[i, j]]
            assert np.allclose(r, ref, rtol=1e-12, atol=1e-10)


# entries of the reference's local Jacobian that are complete derivatives of
# the local residual (the others drop cross-couplings by design)
COMPLETE = [(0, 0), (0, 1), (0, 4), (1, 0), (1, 1), (1, 4), (2, 2), (2, 4), (3, 3), (3, 4),
            (4, 0), (4, 1), (4, 2), (4, 3), (4, 4)]


@pytest.mark.parametrize("cell", [(1, 1), (4, 5), (0, 3), (7, 10)])
@pytest.mark.parametrize("m", [1.0, 1.0 / 3.0])
def test_local_jacobian_vs_central_fd(orc, cell, m):
    F = dome_fields(ny=8, nx=11, seed=5)
    P = orc.phys(m=m, calving_rate=0.2)
    i, j = cell
    x0 = np.array([F["u"][i, j], F["u"][i, j + 1], F["v"][i, j], F["v"][i + 1, j], F["H"][i, j]])
    J, _ = orc.local_system(x0, F, P, DX, DT, i, j)
    for row, col in COMPLETE:
        h = 1e-4 * max(1.0, abs(x0[col]))
        xp, xm = x0.copy(), x0.copy()
        xp[col] += h
        xm[col] -= h
        fd = (orc.local_system(xp, F, P, DX, DT, i, j)[1][row] - orc.local_system(xm, F, P, DX, DT, i, j)[1][row]) / (2 * h)
        assert abs(fd - J[row, col]) <= 1e-6 * max(1.0, abs(J[row, col])), (row, col, fd, J[row, col])


def _fd_jvp(orc, F, P, d, eps):
    Fp, Fm = dict(F), dict(F)
    for k, dk in zip(("u", "v", "H"), d):
        Fp[k] = F[k] + eps * dk
        Fm[k] = F[k] - eps * dk
    rp = orc.residual(Fp, P, DX, DT, use_forcing=False, use_mask=False)
    rm = orc.residual(Fm, P, DX, DT, use_forcing=False, use_mask=False)
    return [(a - b) / (2 * eps) for a, b in zip(rp, rm)]


@pytest.mark.parametrize("m", [1.0, 1.0 / 3.0])
def test_vjp_dot_product(orc, m):
    """<J^T lam, d> == <lam, J d> with J d by central differences of the
    residual; multipliers and directions vanish on Dirichlet facets and masked
    cells (adjoint_test.py restated with an assertion)."""
    F = dome_fields(ny=14, nx=18, seed=6, mask_frac=0.2)
    P = orc.phys(m=m)
    rng = np.random.default_rng(7)
    ny, nx = F["H"].shape

    def pinned(shape_u, shape_v, shape_h):
        a, b, c = rng.standard_normal(shape_u), rng.standard_normal(shape_v), rng.standard_normal(shape_h)
        a[:, 0] = a[:, -1] = 0.0
        b[0] = b[-1] = 0.0
        c[F["mask"] > 0.5] = 0.0
        return a, b, c

    lam = pinned((ny, nx + 1), (ny + 1, nx), (ny, nx))
    d = pinned((ny, nx + 1), (ny + 1, nx), (ny, nx))
    vu, vv, vH = orc.vjp(F, *lam, P, DX, DT)
    Jd = _fd_jvp(orc, F, P, d, 1e-4)
    lhs = (vu * d[0]).sum() + (vv * d[1]).sum() + (vH * d[2]).sum()
    rhs = sum((l * j).sum() for l, j in zip(lam, Jd))
    assert abs(lhs - rhs) <= 1e-7 * abs(rhs)


@pytest.mark.parametrize("m", [1.0, 1.0 / 3.0])
@pytest.mark.parametrize("calving", [0.0, 0.7])
def test_jvp_vs_central_fd(orc, m, calving):
    """J d against central differences of the residual with phi frozen
    (jvp_test.py restated with an assertion): interior facet rows and every
    thickness row; the error of the difference quotient is O(eps^2)."""
    F = dome_fields(ny=13, nx=17, seed=21)
    P = orc.phys(m=m, calving_rate=calving)
    rng = np.random.default_rng(22)
    ny, nx = F["H"].shape
    d = (rng.standard_normal((ny, nx + 1)), rng.standard_normal((ny + 1, nx)), rng.standard_normal((ny, nx)))
    ju, jv, jH = orc.jvp(F, *d, P, DX, DT, use_mask=False)
    fu, fv, fH = _fd_jvp(orc, F, P, d, 1e-3)
    assert rel_err(ju[:, 1:-1], fu[:, 1:-1]) < 1e-7
    assert rel_err(jv[1:-1], fv[1:-1]) < 1e-7
    assert rel_err(jH, fH) < 1e-7
    # each term family moves the result: dropping the thickness direction changes every row kind
    ju0, jv0, jH0 = orc.jvp(F, d[0], d[1], np.zeros_like(d[2]), P, DX, DT, use_mask=False)
    assert rel_err(ju0[:, 1:-1], ju[:, 1:-1]) > 1e-3 and rel_err(jH0, jH) > 1e-3


def test_jvp_reference_conventions(orc):
    """Dirichlet facet rows are zero (not d), and a masked cell's thickness row
    is zero even where d_H is not (glide/cuda/residuals.cu:437,567,659)."""
    F = dome_fields(ny=12, nx=11, seed=23, mask_frac=0.3)
    P = orc.phys()
    rng = np.random.default_rng(24)
    ny, nx = F["H"].shape
    d = (rng.standard_normal((ny, nx + 1)), rng.standard_normal((ny + 1, nx)), rng.standard_normal((ny, nx)))
    ju, jv, jH = orc.jvp(F, *d, P, DX, DT, use_mask=True)
    assert np.all(ju[:, 0] == 0) and np.all(ju[:, -1] == 0)
    assert np.all(jv[0] == 0) and np.all(jv[-1] == 0)
    m = F["mask"] > 0.5
    assert m.any() and np.all(jH[m] == 0)
    _, _, jH_nomask = orc.jvp(F, *d, P, DX, DT, use_mask=False)
    assert np.array_equal(jH[~m], jH_nomask[~m])


@pytest.mark.parametrize("m", [1.0, 1.0 / 3.0])
def test_jvp_vjp_dot_product(orc, m):
    """<lam, J d> == <J^T lam, d> between the oracle's own JVP and VJP
    (adjoint_test.py restated with an assertion)."""
    F = dome_fields(ny=14, nx=18, seed=25, mask_frac=0.2)
    P = orc.phys(m=m, calving_rate=0.4)
    rng = np.random.default_rng(26)
    ny, nx = F["H"].shape

    def pinned():
        a, b, c = rng.standard_normal((ny, nx + 1)), rng.standard_normal((ny + 1, nx)), rng.standard_normal((ny, nx))
        a[:, 0] = a[:, -1] = 0.0
        b[0] = b[-1] = 0.0
        c[F["mask"] > 0.5] = 0.0
        return a, b, c

    lam, d = pinned(), pinned()
    ju, jv, jH = orc.jvp(F, *d, P, DX, DT, use_mask=True)
    vu, vv, vH = orc.vjp(F, *lam, P, DX, DT)
    lhs = (lam[0] * ju).sum() + (lam[1] * jv).sum() + (lam[2] * jH).sum()
    rhs = (vu * d[0]).sum() + (vv * d[1]).sum() + (vH * d[2]).sum()
    assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + abs(rhs))


def test_vjp_forcing_is_subtracted(orc):
    F = dome_fields(n=10, seed=8)
    P = orc.phys()
    rng = np.random.default_rng(9)
    lam = [rng.standard_normal(F[k].shape) for k in ("u", "v", "H")]
    af = [rng.standard_normal(F[k].shape) for k in ("u", "v", "H")]
    a = orc.vjp(F, *lam, P, DX, DT, use_forcing=False, af=af)
    b = orc.vjp(F, *lam, P, DX, DT, use_forcing=True, af=af)
    assert np.allclose(b[0][:, 1:-1], a[0][:, 1:-1] - af[0][:, 1:-1], atol=1e-12)
    assert np.allclose(b[1][1:-1], a[1][1:-1] - af[1][1:-1], atol=1e-12)
    assert np.allclose(b[2], a[2] - af[2], atol=1e-12)


@pytest.mark.parametrize("m", [1.0, 1.0 / 3.0])
def test_gradients_vs_fd(orc, m):
    """<g_beta, db> == <lam, dR/dbeta db> (and bed), with lam zero on the facets
    whose right/bottom-cell contribution the reference's guards skip."""
    F = dome_fields(ny=12, nx=15, seed=10)
    P = orc.phys(m=m)
    rng = np.random.default_rng(11)
    ny, nx = F["H"].shape
    lu, lv = rng.standard_normal((ny, nx + 1)), rng.standard_normal((ny + 1, nx))
    lu[:, 0] = lu[:, nx - 1] = lu[:, nx] = 0.0
    lv[0] = lv[ny - 1] = lv[ny] = 0.0
    for name, g in (("beta", orc.grad_beta(F, lu, lv, P)), ("bed", orc.grad_bed(F, lu, lv, DX))):
        d = rng.standard_normal((ny, nx))
        eps = 1e-3 * np.abs(F[name]).max()
        Fp, Fm = dict(F), dict(F)
        Fp[name] = F[name] + eps * d
        Fm[name] = F[name] - eps * d
        rp = orc.residual(Fp, P, DX, DT, use_forcing=False, use_mask=False)
        rm = orc.residual(Fm, P, DX, DT, use_forcing=False, use_mask=False)
        rhs = ((rp[0] - rm[0]) * lu).sum() / (2 * eps) + ((rp[1] - rm[1]) * lv).sum() / (2 * eps)
        lhs = (g * d).sum()
        assert abs(lhs - rhs) <= 1e-7 * max(abs(rhs), 1e-30), name


def test_gradient_guard_quirk(orc):
    """The reference drops the right-cell term of facet nx-1 (grad.cu:62)."""
    F = dome_fields(ny=6, nx=8, seed=12)
    P = orc.phys()
    lu = np.zeros((6, 9))
    lv = np.zeros((7, 8))
    lu[2, 7] = 1.0  # facet nx-1: only its left cell (2, 6) receives a term
    g = orc.grad_beta(F, lu, lv, P)
    assert g[2, 7] == 0.0 and g[2, 6] != 0.0
    lu[2, 7], lu[2, 6] = 0.0, 1.0  # an ordinary facet feeds both neighbours
    g = orc.grad_beta(F, lu, lv, P)
    assert g[2, 5] != 0.0 and g[2, 6] != 0.0


def _dome_mg(orc, n, L, seed=0):
    from glide_amd.synthetic import dome_problem
    p = dome_problem(n, DX, seed=seed)
    om = orc.OracleMG(n, n, DX, L)
    for k in ("bed", "B", "beta", "smb"):
        om.set(k, getattr(p, k))
    om.set("H", p.H)
    om.set("H_prev", p.H)
    return p, om


def test_fascd_contracts(orc):
    _, om = _dome_mg(orc, 32, 3)
    rep = om.fascd_solve(DT, orc.OracleMG.fas_config(vcycles=4, rtol=0.0, atol=0.0))
    norms = [rep.initial_norm] + [rep.norms[k] for k in range(rep.vcycles)]
    assert rep.vcycles == 4
    for a, b in zip(norms[:-1], norms[1:]):
        assert b < 0.5 * a
    assert norms[3] < 1e-4 * norms[0]
    assert om.get("H").min() >= 0.1 - 1e-12  # thickness constraint


def test_vanka_fixed_point_at_solution(orc):
    """At a converged state an additive pass is (nearly) the identity."""
    _, om = _dome_mg(orc, 16, 2, seed=1)
    om.fascd_solve(DT, orc.OracleMG.fas_config(vcycles=12, rtol=0.0, atol=0.0))
    F = {k: om.get(k) for k in ("u", "v", "H", "phi", "mask", "bed", "B", "beta")}
    F["f_u"], F["f_v"] = np.zeros_like(F["u"]), np.zeros_like(F["v"])
    F["f_H"] = om.get("H_prev") / DT + om.get("smb")
    F["gamma"] = np.full_like(F["H"], 0.1)
    u, v, H, mask, _ = orc.vanka_sweep(F, orc.phys(), DX, DT, newton_steps=60)
    assert rel_err(u, F["u"]) < 1e-5 and rel_err(H, F["H"]) < 1e-6


def test_mc_pass_fixed_point_at_solution(orc):
    """At a converged state a red-black pass with full local Newton solves
    is the identity: its patch solves see zero residual rows (no counterpart
    in the reference; pins mc_sweep against the residual's zero)."""
    _, om = _dome_mg(orc, 16, 2, seed=1)
    rep = om.fascd_solve(DT, orc.OracleMG.fas_config(vcycles=40, rtol=1e-13, atol=0.0))
    assert rep.final_norm < 1e-11 * rep.initial_norm
    F = {k: om.get(k) for k in ("u", "v", "H", "phi", "mask", "bed", "B", "beta")}
    F["f_u"], F["f_v"] = np.zeros_like(F["u"]), np.zeros_like(F["v"])
    F["f_H"] = om.get("H_prev") / DT + om.get("smb")
    F["gamma"] = np.full_like(F["H"], 0.1)
    u, v, H, mask, _ = orc.mc_sweep(F, orc.phys(), DX, DT, newton_steps=30, relaxation=1.0, newton_tol=0.0)
    assert rel_err(u, F["u"]) < 1e-9 and rel_err(v, F["v"]) < 1e-9 and rel_err(H, F["H"]) < 1e-12
    assert np.array_equal(mask, F["mask"])
    # and away from it the pass moves the state
    F2 = dict(F)
    F2["u"] = F["u"] * 1.01
    u2, _, _, _, _ = orc.mc_sweep(F2, orc.phys(), DX, DT, newton_steps=30, relaxation=1.0, newton_tol=0.0)
    assert rel_err(u2, F2["u"]) > 1e-4


def test_mc_adjoint_pass_fixed_point_at_solution(orc):
    """At a converged adjoint solution (pinned multipliers zero) a red-black
    adjoint pass leaves the multipliers unchanged: its patch solves see the
    zero residual of the transpose operator."""
    _, om = _dome_mg(orc, 16, 2, seed=5)
    om.fascd_solve(DT, orc.OracleMG.fas_config(vcycles=40, rtol=1e-13, atol=0.0))
    rng = np.random.default_rng(9)
    af = (rng.standard_normal((16, 17)), rng.standard_normal((17, 16)), np.zeros((16, 16)))
    for k, a in zip(("af_u", "af_v", "af_H"), af):
        om.put(k, a)
    cfg = orc.OracleMG.fas_config(vcycles=60, rtol=1e-14, atol=0.0)
    cfg.project_pinned = 1
    om.fas_adjoint_solve(DT, cfg)
    F = {k: om.get(k) for k in ("u", "v", "H", "phi", "mask", "bed", "B", "beta")}
    lam = tuple(om.get(k) for k in ("lambda_u", "lambda_v", "lambda_H"))
    out = orc.mc_adjoint_sweep(F, lam, af, orc.phys(), DX, DT, omega=1.0)
    for a, b in zip(out, lam):
        assert rel_err(a, b) < 1e-9
    lam2 = (lam[0] * 1.01, lam[1], lam[2])
    out2 = orc.mc_adjoint_sweep(F, lam2, af, orc.phys(), DX, DT, omega=1.0)
    assert rel_err(out2[0], lam2[0]) > 1e-5


def test_adjoint_solve_contracts(orc):
    _, om = _dome_mg(orc, 32, 3, seed=2)
    cfg = orc.OracleMG.fas_config(vcycles=3, rtol=0.0, atol=0.0)
    om.fascd_solve(DT, cfg)
    rng = np.random.default_rng(3)
    om.put("af_u", rng.standard_normal((32, 33)))
    om.put("af_v", rng.standard_normal((33, 32)))
    om.put("af_H", np.zeros((32, 32)))
    rep = om.fas_adjoint_solve(DT, cfg)
    assert rep.norms[rep.vcycles - 1] < 1e-2 * rep.initial_norm


# ---------------------------------------------------------------- closed-form value pins
# Each pins a term's value (not only its consistency with the oracle's own
# derivatives): a wrong factor or sign in the basal drag, driving stress,
# stress coefficients, grounded sigmoid or the omitted Jacobian couplings
# fails one of them.

def _slab(ny, nx, seed=0):
    F = dome_fields(ny=ny, nx=nx, seed=seed, with_velocity=False)
    for k in ("f_u", "f_v", "f_H"):
        F[k] = np.zeros_like(F[k])
    F["mask"] = np.zeros((ny, nx))
    return F


@pytest.mark.parametrize("m", [1.0, 1.0 / 3.0])
@pytest.mark.parametrize("axis", ["x", "y"])
def test_plug_flow_slab_closed_form(orc, m, axis):
    """Uniform thickness on a bed of constant slope s (downhill towards +x, or
    towards the top row for y), uniform beta and grounded fraction phi, and a
    uniform plug velocity w along the slope: every strain rate vanishes, so an
    interior momentum row is the basal drag minus the driving stress,
        r(w) = H phi s - (beta phi (w^2 + u_reg)^((m-1)/2) + water_drag) w,
    (head units, glide/cuda/stress.cu:253-548), zero at the sliding speed w*
    (a scalar root); the cross-slope rows vanish."""
    from scipy.optimize import brentq
    ny, nx = 8, 12
    P = orc.phys(m=m)
    H, s, beta, phi = 600.0, 2e-3, 7.3, 0.7
    F = _slab(ny, nx)
    F["H"] = np.full((ny, nx), H)
    F["beta"] = np.full((ny, nx), beta)
    F["phi"] = np.full((ny, nx), phi)
    if axis == "x":
        F["bed"] = np.tile(1000.0 - s * (np.arange(nx) + 0.5) * DX, (ny, 1))
    else:  # row i is the top: the bed rises with i, downhill is towards the top (+v)
        F["bed"] = np.tile((1000.0 + s * (np.arange(ny) + 0.5) * DX)[:, None], (1, nx))

    def r_closed(w):
        return H * phi * s - (beta * phi * (w * w + P.u_reg) ** ((m - 1.0) / 2.0) + P.water_drag) * w

    w_star = brentq(r_closed, 0.0, 1e6, xtol=1e-14, rtol=1e-15)
    for w in (0.0, 0.37 * w_star, w_star, 1.9 * w_star):
        F["u"] = np.full((ny, nx + 1), w if axis == "x" else 0.0)
        F["v"] = np.full((ny + 1, nx), w if axis == "y" else 0.0)
        ru, rv, _ = orc.residual(F, P, DX, DT, use_forcing=False, use_mask=False)
        along, across = (ru[:, 1:-1], rv[1:-1]) if axis == "x" else (rv[1:-1], ru[:, 1:-1])
        scale = H * phi * s
        assert np.abs(along - r_closed(w)).max() <= 1e-12 * scale, w
        assert np.abs(across).max() <= 1e-14 * scale
    assert r_closed(w_star) == pytest.approx(0.0, abs=1e-12 * H * phi * s) and w_star > 0


def test_manufactured_quadratic_velocity_n1(orc):
    """n = 1 makes eta = B/2 independent of the strain rate.  For quadratic
    u = a x^2 + c y^2, v = d x^2 + e y^2 (y up: row i is the top) with uniform
    H, B, no basal drag (beta = 0) and a flat grounded bed, the centred
    differences are exact, so interior rows equal the continuum operator
        r_u = d/dx[2 eta H (2 u_x + v_y)] + d/dy[eta H (u_y + v_x)] - wd u = eta H (8a + 2c) - wd u
        r_v = d/dy[2 eta H (u_x + 2 v_y)] + d/dx[eta H (u_y + v_x)] - wd v = eta H (8e + 2d) - wd v
    which fixes the sigma_xx / sigma_yy / sigma_xy coefficients and signs."""
    ny, nx = 10, 12
    P = orc.phys(n=1.0)
    rng = np.random.default_rng(1)
    a, c, d, e = 1e-7 * rng.standard_normal(4)
    Hc, Bc = 700.0, 3.0
    F = _slab(ny, nx)
    F["H"], F["B"] = np.full((ny, nx), Hc), np.full((ny, nx), Bc)
    F["bed"], F["beta"], F["phi"] = np.zeros((ny, nx)), np.zeros((ny, nx)), np.ones((ny, nx))
    xu, yu = np.arange(nx + 1) * DX, -(np.arange(ny) + 0.5) * DX
    xv, yv = (np.arange(nx) + 0.5) * DX, -np.arange(ny + 1) * DX
    F["u"] = a * xu[None, :] ** 2 + c * yu[:, None] ** 2
    F["v"] = d * xv[None, :] ** 2 + e * yv[:, None] ** 2
    ru, rv, _ = orc.residual(F, P, DX, DT, use_forcing=False, use_mask=False)
    K = 0.5 * Bc * Hc
    ref_u = K * (8 * a + 2 * c) - P.water_drag * F["u"]
    ref_v = K * (8 * e + 2 * d) - P.water_drag * F["v"]
    # rows whose stencil stays off the boundary vertices (sigma_xy = 0 there)
    assert rel_err(ru[2:-2, 2:-2], ref_u[2:-2, 2:-2]) < 1e-12
    assert rel_err(rv[2:-2, 2:-2], ref_v[2:-2, 2:-2]) < 1e-12


def test_grounded_fraction_closed_forms(orc):
    """phi = 1 / (1 + exp(-clamp(c (bed + 0.917 H), -20, 20))): one half at
    flotation, the clamp at +-20, and a value inside the clamp range
    (glide/cuda/common.cu:66-82, glide/cuda/viscosity.cu:16-36)."""
    c = 0.1
    H = np.array([[100.0, 200.0, 300.0], [10.0, 2500.0, 55.0]])
    assert np.array_equal(orc.grounded(H, -0.917 * H, c=c), np.full_like(H, 0.5))
    bed = np.array([[1000.0, -1000.0, 3.0], [-5.0, -2500.0, -60.0]])
    z = np.clip(c * (bed + 0.917 * H), -20.0, 20.0)
    got = orc.grounded(H, bed, c=c)
    assert got[0, 0] == pytest.approx(1.0 / (1.0 + np.exp(-20.0)), rel=1e-15)
    assert got[0, 1] == pytest.approx(1.0 / (1.0 + np.exp(20.0)), rel=1e-14)
    assert rel_err(got, 1.0 / (1.0 + np.exp(-z))) < 1e-15


def test_driving_stress_flotation_closed_form(orc):
    """At rest (u = v = 0: no strain, no drag) a momentum row is minus the
    driving stress -H_avg (S_B - S_A)/dx with the flotation-blended surface
    S = phi bed - (1 - phi) 0.917 H + H (x: A left, B right; y: A bottom,
    B top), which is (1 - 0.917) H for floating ice (phi = 0) and bed + H for
    grounded ice (glide/cuda/stress.cu:418-548)."""
    ny, nx = 8, 12
    F = _slab(ny, nx, seed=3)
    rng = np.random.default_rng(4)
    F["phi"] = rng.random((ny, nx))
    F["phi"][:, :3] = 0.0     # a floating strip
    F["phi"][:, -3:] = 1.0    # a grounded strip
    F["beta"] = 10.0 * rng.random((ny, nx))
    F["u"], F["v"] = np.zeros((ny, nx + 1)), np.zeros((ny + 1, nx))
    ru, rv, _ = orc.residual(F, orc.phys(), DX, DT, use_forcing=False, use_mask=False)
    H, b, p = F["H"], F["bed"], F["phi"]
    S = np.where(p == 0.0, (1.0 - 0.917) * H, np.where(p == 1.0, b + H, p * b - (1 - p) * 0.917 * H + H))
    ref_u = -0.5 * (H[:, 1:] + H[:, :-1]) * (S[:, 1:] - S[:, :-1]) / DX
    ref_v = -0.5 * (H[1:] + H[:-1]) * (S[:-1] - S[1:]) / DX
    assert rel_err(ru[:, 1:-1], ref_u) < 1e-13
    assert rel_err(rv[1:-1], ref_v) < 1e-13


@pytest.mark.parametrize("cell", [(1, 1), (3, 4), (5, 8)])
def test_local_jacobian_omits_exactly_the_shear_cross_couplings(orc, cell):
    """The reference's 5x5 patch Jacobian leaves out the dependence of the
    corner shear stresses on the other velocity component (glide/cuda/vanka.cu
    assembles only the own-component derivative of sigma_xy).  With n = 1
    (eta = B/2), uniform B and linear drag, the complete derivative of the
    local residual (central differences, viscosity frozen) minus the
    reference's entry is, for the four corner vertices v with
    eta H_v = (B/8) (sum of the four adjacent H),
        d(u_l, v_t) = d(v_t, u_l) = D^2 eta H_tl,  d(u_l, v_b) = d(v_b, u_l) = -D^2 eta H_bl,
        d(u_r, v_t) = d(v_t, u_r) = -D^2 eta H_tr, d(u_r, v_b) = d(v_b, u_r) = D^2 eta H_br,
    and zero for every other entry (D = 1/dx)."""
    ny, nx = 8, 11
    F = dome_fields(ny=ny, nx=nx, seed=5)
    F["B"] = np.full((ny, nx), 2.5)
    P = orc.phys(n=1.0)
    i, j = cell
    x0 = np.array([F["u"][i, j], F["u"][i, j + 1], F["v"][i, j], F["v"][i + 1, j], F["H"][i, j]])
    J, _ = orc.local_system(x0, F, P, DX, DT, i, j)
    full = np.zeros((5, 5))
    for col in range(5):
        h = 1e-4 * max(1.0, abs(x0[col]))
        xp, xm = x0.copy(), x0.copy()
        xp[col] += h
        xm[col] -= h
        full[:, col] = (orc.local_system(xp, F, P, DX, DT, i, j)[1] - orc.local_system(xm, F, P, DX, DT, i, j)[1]) / (2 * h)
    Hf = F["H"]

    def eH(i0, j0):  # vertex at the top-left corner of cell (i0, j0)
        return 2.5 / 8.0 * (Hf[i0 - 1, j0 - 1] + Hf[i0 - 1, j0] + Hf[i0, j0 - 1] + Hf[i0, j0])

    D2 = 1.0 / DX ** 2
    tl, tr, bl, br = eH(i, j), eH(i, j + 1), eH(i + 1, j), eH(i + 1, j + 1)
    dropped = np.zeros((5, 5))
    dropped[0, 2] = dropped[2, 0] = D2 * tl
    dropped[0, 3] = dropped[3, 0] = -D2 * bl
    dropped[1, 2] = dropped[2, 1] = -D2 * tr
    dropped[1, 3] = dropped[3, 1] = D2 * br
    scale = np.abs(J).max()
    assert np.abs((full - J) - dropped).max() <= 1e-7 * scale
    # the omitted couplings are not negligible against the velocity block
    assert np.abs(dropped).max() > 1e-3 * np.abs(J[:4, :4]).max()
    # the reference's velocity block (and the complete one) is symmetric for linear drag
    assert np.abs(J[:4, :4] - J[:4, :4].T).max() <= 1e-15 * scale
