"""The reference's own GPU test scripts, run through this build's API with the
same setups: the linearised operator against central differences
(tests/jvp_test.py), the dot-product test of the transpose operator
(tests/adjoint_test.py) and the twin-experiment gradient check
(tests/grad_test.py).  The reference scripts draw unseeded CuPy random numbers
in FP32 and only print their comparisons (the gradient script asserts 5e-2);
here the draws are seeded NumPy draws in FP64 and each comparison has a bar
(reasons in the tests)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

RHO_I, G = 917.0, 9.81


def _slab(ny, nx, L):
    """tests/jvp_test.py:8-36: a 1000 m slab on a 0.1 degree slope with
    sinusoidal drag (the grids of adjoint_test.py and grad_test.py too)."""
    x = np.linspace(0, 7 * L, nx)  # 7 wavelengths in each direction
    y = np.linspace(0, 7 * L, ny)
    dx = x[1] - x[0]
    X, Y = np.meshgrid(x, y)
    srf = 1000.0 * np.ones((ny, nx)) - np.tan(np.deg2rad(0.1)) * X + 10000
    bed = srf - 1000
    thk = srf - bed
    beta = (1000 * np.sin(2 * np.pi * X / L) * np.sin(2 * np.pi * Y / L) + 1000) / (RHO_I * G)
    B = np.full((ny, nx), (1e-16 ** -(1.0 / 3)) / (RHO_I * G))
    return dict(dx=dx, bed=bed, thk=thk, beta=beta, B=B)


def _slab_grid(n, L):
    from glide_amd.grid import Grid
    s = _slab(n, n, L)
    grid = Grid(n, n, s["dx"], device="cuda")
    grid.geometry.bed.set(s["bed"])
    grid.rheology.B.set(s["B"])
    grid.sliding.beta.set(s["beta"])
    grid.sliding.m.set(1.0)
    grid.sliding.u_reg.set(1.0)
    grid.state.H.set(s["thk"])
    grid.state.H_prev.set(s["thk"])
    return grid


def _directions(grid, rng):
    """Random mask and directions with the Dirichlet rows and masked cells
    zeroed (jvp_test.py:46-60, adjoint_test.py:46-72)."""
    ny, nx = grid.ny, grid.nx
    mask = rng.integers(0, 2, size=(ny, nx)).astype(np.float64)
    grid.state.mask.data[:, :] = torch.as_tensor(mask, device="cuda")
    du = rng.standard_normal((ny, nx + 1))
    du[:, 0] = 0
    du[:, -1] = 0
    dv = rng.standard_normal((ny + 1, nx))
    dv[0] = 0
    dv[-1] = 0
    dH = rng.standard_normal((ny, nx))
    dH[mask > 0.5] = 0
    return du, dv, dH


def _t(a):
    return torch.as_tensor(a, dtype=torch.float64, device="cuda")


@pytest.fixture(scope="module")
def settled_grid():
    """224 x 224 slab after set_rhs and 4000 smoothing passes (jvp_test.py:45-46)."""
    print("reference scripts: 224^2 slab, 4000 passes", flush=True)
    grid = _slab_grid(224, 10000.0)
    dt = 1.0
    grid.forward_operators.set_rhs(dt)
    grid.forward_operators.vanka_sweep(dt, 4000)
    torch.cuda.synchronize()
    return grid


def test_jvp_script(settled_grid):
    """tests/jvp_test.py: J var against central differences of the residual
    with eps = 1e-2 (the slab is grounded everywhere, so the grounded fraction
    does not move)."""
    grid = settled_grid
    fo = grid.forward_operators
    dt = 1.0
    rng = np.random.default_rng(41)
    u0, v0, H0, m0 = (grid.state.u.data.clone(), grid.state.v.data.clone(), grid.state.H.data.clone(),
                      grid.state.mask.data.clone())
    du, dv, dH = _directions(grid, rng)
    fo.var_u[:, :] = _t(du)
    fo.var_v[:, :] = _t(dv)
    fo.var_H[:, :] = _t(dH)
    fo.compute_jvp(dt)
    jvp = [fo.jvp_u.clone(), fo.jvp_v.clone(), fo.jvp_H.clone()]
    errs = {}
    for eps in (1e-2, 1e-4):
        res = []
        for sgn in (1.0, -1.0):
            grid.state.u.data[:, :] = u0 + sgn * eps * fo.var_u
            grid.state.v.data[:, :] = v0 + sgn * eps * fo.var_v
            grid.state.H.data[:, :] = H0 + sgn * eps * fo.var_H
            fo.compute_residual(dt)
            res.append([fo.r_u.clone(), fo.r_v.clone(), fo.r_H.clone()])
        for name, j, r1, r0 in zip("uvH", jvp, res[0], res[1]):
            fd = (r1 - r0) / (2 * eps)
            errs[name, eps] = float(torch.linalg.norm(fd - j) / torch.linalg.norm(j))
            print(f"  jvp {name} eps {eps:g}: relative norm against finite differences {errs[name, eps]:.3e}",
                  flush=True)
    grid.state.u.data.copy_(u0)
    grid.state.v.data.copy_(v0)
    grid.state.H.data.copy_(H0)
    grid.state.mask.data.copy_(m0)
    # the script's eps = 1e-2 is outside the linear range of the viscosity on
    # this nearly uniform flow (white-noise directions against small strain
    # rates), so the check is the central difference's O(eps^2) decay
    for name in "uvH":
        assert errs[name, 1e-4] < 1e-5 or errs[name, 1e-4] < 1e-3 * errs[name, 1e-2], name


def test_adjoint_script(settled_grid):
    """tests/adjoint_test.py: <J^T lambda, var> = <J var, lambda>."""
    grid = settled_grid
    fo, ao = grid.forward_operators, grid.adjoint_operators
    dt = 1.0
    rng = np.random.default_rng(42)
    m0 = grid.state.mask.data.clone()
    du, dv, dH = _directions(grid, rng)
    fo.var_u[:, :] = _t(du)
    fo.var_v[:, :] = _t(dv)
    fo.var_H[:, :] = _t(dH)
    lu, lv, lH = _directions_like(grid, rng)
    grid.adjoint.lambda_u.data[:, :] = _t(lu)
    grid.adjoint.lambda_v.data[:, :] = _t(lv)
    grid.adjoint.lambda_H.data[:, :] = _t(lH)
    fo.compute_jvp(dt)
    ao.compute_vjp(dt)
    t1 = float((ao.vjp_u * fo.var_u).sum() + (ao.vjp_v * fo.var_v).sum() + (ao.vjp_H * fo.var_H).sum())
    t2 = float((fo.jvp_u * grid.adjoint.lambda_u.data).sum() + (fo.jvp_v * grid.adjoint.lambda_v.data).sum()
               + (fo.jvp_H * grid.adjoint.lambda_H.data).sum())
    grid.state.mask.data.copy_(m0)
    rel = abs(t1 - t2) / abs(0.5 * (t1 + t2))
    print(f"  dot products {t1:.12e} {t2:.12e} relative difference {rel:.3e}", flush=True)
    assert rel < 1e-11


def _directions_like(grid, rng):
    """Multipliers with the Dirichlet rows and the current mask's cells zeroed
    (adjoint_test.py:62-72)."""
    ny, nx = grid.ny, grid.nx
    mask = grid.state.mask.data.cpu().numpy()
    lu = rng.standard_normal((ny, nx + 1))
    lu[:, 0] = 0
    lu[:, -1] = 0
    lv = rng.standard_normal((ny + 1, nx))
    lv[0] = 0
    lv[-1] = 0
    lH = rng.standard_normal((ny, nx))
    lH[mask > 0.5] = 0
    return lu, lv, lH


def test_grad_script():
    """tests/grad_test.py: twin experiment on the 448 x 448 slab with six
    levels; the beta gradient's projection on a random direction against
    central differences of the L1 velocity misfit, with the script's solver
    options except the stopping test: with its relative tolerance of 1e-3
    (three V-cycles here) the perturbed solves' own error dominates the
    difference quotient (16% measured), so the solves here run to 1e-9
    (nine forward, seven adjoint V-cycles), where the two agree to ~2e-3 (the
    L1 misfit's kinks bound it)."""
    from glide_amd.multigrid import FASAdjointSolver, FASCDSolver, Multigrid
    print("reference scripts: grad_test 448^2, 6 levels", flush=True)
    s = _slab(448, 448, 20000.0)
    dt = 1.0
    mg = Multigrid(6, ny=448, nx=448, dx=s["dx"], device="cuda")
    grid = mg.levels[0]
    mg.geometry.bed.set(s["bed"])
    mg.rheology.B.set(s["B"])
    mg.sliding.beta.set(s["beta"])
    mg.sliding.m.set(1.0)
    mg.sliding.u_reg.set(1.0)
    mg.state.H.set(s["thk"])
    mg.state.H_prev.set(s["thk"])

    solver = FASCDSolver(mg)
    solver.vanka_options.omega.set(0.5)
    solver.vanka_options.newton_options.relaxation.set(0.5)
    solver.vanka_options.newton_options.steps.set(30)
    solver.fas_options.coarsest_steps.set(200)
    solver.fas_options.pre_steps.set(10)
    solver.fas_options.post_steps.set(50)
    solver.fas_options.finest_steps.set(150)
    solver.fas_options.maximum_vcycles.set(60)
    solver.fas_options.relative_tolerance.set(1e-9)
    solver.fas_options.absolute_tolerance.set(0.0)
    solver.fas_options.report_norms.set(False)
    solver.solve(dt)
    u_obs = grid.state.u.data.clone()
    v_obs = grid.state.v.data.clone()

    beta_mean = float(np.mean(s["beta"]))
    mg.sliding.beta.set(np.full((448, 448), beta_mean))
    solver.solve(dt)
    u = grid.state.u.data.clone()
    v = grid.state.v.data.clone()
    H = grid.state.H.data.clone()

    dJdu = torch.sign(u - u_obs)
    dJdv = torch.sign(v - v_obs)
    grid.adjoint_operators.f_u[:, :] = -dJdu
    grid.adjoint_operators.f_v[:, :] = -dJdv
    adjoint_solver = FASAdjointSolver(mg)
    adjoint_solver.fas_options.coarsest_steps.set(200)
    adjoint_solver.fas_options.pre_steps.set(10)
    adjoint_solver.fas_options.post_steps.set(50)
    adjoint_solver.fas_options.finest_steps.set(150)
    adjoint_solver.fas_options.maximum_vcycles.set(60)
    adjoint_solver.fas_options.absolute_tolerance.set(0.0)
    adjoint_solver.fas_options.relative_tolerance.set(1e-9)
    adjoint_solver.fas_options.report_norms.set(False)
    adjoint_solver.vanka_options.newton_options.ssa_damping.set(0.1)
    adjoint_solver.vanka_options.omega.set(0.5)
    adjoint_solver.solve(dt)
    grid.adjoint_operators.compute_gradient_beta()

    rng = np.random.default_rng(43)
    beta_pert = _t(rng.standard_normal((448, 448)))
    eps = 1e-3
    beta_0 = grid.sliding.beta.data.clone()
    J = []
    for sgn in (1.0, -1.0):
        grid.sliding.beta.data[:, :] = beta_0 + sgn * eps * beta_pert
        grid.state.u.data[:, :] = u
        grid.state.v.data[:, :] = v
        grid.state.H.data[:, :] = H
        solver.solve(dt)
        J.append(float((grid.state.u.data - u_obs).abs().sum() + (grid.state.v.data - v_obs).abs().sum()))
    gvp_fd = (J[0] - J[1]) / (2 * eps)
    gvp_ad = float((grid.sliding.beta.grad * beta_pert).sum())
    rel = abs(gvp_fd - gvp_ad) / abs(gvp_ad)
    print(f"  FD: {gvp_fd:.8e}, Adj: {gvp_ad:.8e}, Rel. Err.: {rel:.3e}", flush=True)
    assert rel < 1e-2
