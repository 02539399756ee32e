"""Parity of the gfx950 SSA kernels and drivers with the FP64 CPU oracle.

Every test runs the HIP path through the C ABI (via the glide_amd API) and the
oracle on the same seeded inputs.  Bar: FP64 results agree to 1e-12 relative
for single kernels and 1e-10 relative for whole solves (u, v, H, lambda,
gradients); active-set masks agree exactly.
"""
import numpy as np
import pytest
import torch

from tests.util import dome_fields, rel_err

pytestmark = pytest.mark.gpu

KTOL = 1e-12   # one kernel
STOL = 1e-10   # whole FAS solves


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=torch.float64)


def N(t):
    return t.detach().cpu().numpy()


def grid_from(F, dx=1500.0):
    from glide_amd.grid import Grid
    ny, nx = F["H"].shape
    g = Grid(ny, nx, dx, device="cuda")
    for k in ("u", "v", "H", "H_prev", "phi", "mask"):
        getattr(g.state, k).set(F[k])
    g.geometry.bed.set(F["bed"])
    g.rheology.B.set(F["B"])
    g.sliding.beta.set(F["beta"])
    g.forcing.smb.set(F["smb"])
    fo = g.forward_operators
    fo.f_u.copy_(T(F["f_u"]))
    fo.f_v.copy_(T(F["f_v"]))
    fo.f_H.copy_(T(F["f_H"]))
    fo.gamma.copy_(T(F["gamma"]))
    return g


SHAPES = [(32, 32), (24, 40), (66, 130)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("use_forcing,use_mask", [(True, True), (True, False), (False, False)])
def test_residual(orc, shape, use_forcing, use_mask):
    ny, nx = shape
    print(f"residual shape={shape} forcing={use_forcing} mask={use_mask}", flush=True)
    F = dome_fields(ny=ny, nx=nx, seed=1, mask_frac=0.2)
    P = orc.phys()
    g = grid_from(F)
    fo = g.forward_operators
    fo.compute_residual(10.0, use_mask=use_mask, operator_only=not use_forcing, freeze_phi=True)
    torch.cuda.synchronize()
    ru, rv, rH = orc.residual(F, P, 1500.0, 10.0, use_forcing=use_forcing, use_mask=use_mask)
    out = (fo.r_u, fo.r_v, fo.r_H) if use_forcing else (fo.F_u, fo.F_v, fo.F_H)
    for a, b in zip(out, (ru, rv, rH)):
        assert rel_err(N(a), b) < KTOL


def test_residual_nonlinear_drag(orc):
    F = dome_fields(n=40, seed=2, mask_frac=0.1)
    P = orc.phys(m=1.0 / 3.0, u_reg=0.5, calving_rate=0.2)
    g = grid_from(F)
    g.sliding.m.set(1.0 / 3.0)
    g.sliding.u_reg.set(0.5)
    g.calving.calving_rate.set(0.2)
    fo = g.forward_operators
    fo.compute_residual(5.0, freeze_phi=True)
    ru, rv, rH = orc.residual(F, P, 1500.0, 5.0)
    for a, b in zip((fo.r_u, fo.r_v, fo.r_H), (ru, rv, rH)):
        assert rel_err(N(a), b) < KTOL


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("m,calving,use_mask", [(1.0, 0.0, True), (1.0 / 3.0, 0.3, True), (1.0, 0.2, False)])
def test_jvp(orc, shape, m, calving, use_mask):
    """ForwardOperators.compute_jvp against the oracle's J d (pinned by
    central differences and by the dot product with the VJP)."""
    ny, nx = shape
    print(f"jvp shape={shape} m={m:.3f} calving={calving} mask={use_mask}", flush=True)
    F = dome_fields(ny=ny, nx=nx, seed=31, mask_frac=0.2)
    P = orc.phys(m=m, calving_rate=calving)
    g = grid_from(F)
    g.sliding.m.set(m)
    g.calving.calving_rate.set(calving)
    rng = np.random.default_rng(32)
    d = (rng.standard_normal((ny, nx + 1)), rng.standard_normal((ny + 1, nx)), rng.standard_normal((ny, nx)))
    fo = g.forward_operators
    fo.var_u.copy_(T(d[0]))
    fo.var_v.copy_(T(d[1]))
    fo.var_H.copy_(T(d[2]))
    fo.compute_jvp(10.0, use_mask=use_mask, freeze_phi=True)
    torch.cuda.synchronize()
    ref = orc.jvp(F, *d, P, 1500.0, 10.0, use_mask=use_mask)
    for a, b in zip((fo.jvp_u, fo.jvp_v, fo.jvp_H), ref):
        assert rel_err(N(a), b) < KTOL
    assert np.all(N(fo.jvp_u)[:, [0, -1]] == 0) and np.all(N(fo.jvp_v)[[0, -1]] == 0)


def test_jvp_freeze_calving_and_phi_refresh(orc):
    """freeze_calving drops the calving term; without freeze_phi the grounded
    fraction is recomputed from H and bed first."""
    F = dome_fields(ny=20, nx=28, seed=33)
    P0 = orc.phys(calving_rate=0.0)
    g = grid_from(F)
    g.calving.calving_rate.set(0.5)
    rng = np.random.default_rng(34)
    d = (rng.standard_normal((20, 29)), rng.standard_normal((21, 28)), rng.standard_normal((20, 28)))
    fo = g.forward_operators
    for t, a in zip((fo.var_u, fo.var_v, fo.var_H), d):
        t.copy_(T(a))
    fo.compute_jvp(10.0, freeze_calving=True)
    torch.cuda.synchronize()
    F2 = dict(F)
    F2["phi"] = orc.grounded(F["H"], F["bed"], 0.1)
    ref = orc.jvp(F2, *d, P0, 1500.0, 10.0)
    for a, b in zip((fo.jvp_u, fo.jvp_v, fo.jvp_H), ref):
        assert rel_err(N(a), b) < KTOL


def test_grounded(orc):
    F = dome_fields(n=48, seed=3)
    g = grid_from(F)
    g.forward_operators.compute_phi()
    ref = orc.grounded(F["H"], F["bed"], 0.1)
    assert rel_err(N(g.state.phi.data), ref) < KTOL


@pytest.mark.parametrize("shape", SHAPES)
def test_vanka_dump(orc, shape):
    ny, nx = shape
    F = dome_fields(ny=ny, nx=nx, seed=4)
    g = grid_from(F)
    J, r = g.forward_operators.vanka_dump(10.0)
    Jo, ro = orc.vanka_dump(F, orc.phys(), 1500.0, 10.0)
    assert rel_err(N(J), Jo) < KTOL
    assert rel_err(N(r), ro) < KTOL


@pytest.mark.parametrize("shape", SHAPES)
def test_vanka_single_pass(orc, shape):
    ny, nx = shape
    print(f"vanka pass shape={shape}", flush=True)
    F = dome_fields(ny=ny, nx=nx, seed=5)
    g = grid_from(F)
    g.forward_operators.vanka_sweep(10.0, 1, freeze_phi=True)
    u, v, H, mask, _ = orc.vanka_sweep(F, orc.phys(), 1500.0, 10.0)
    assert rel_err(N(g.state.u.data), u) < KTOL
    assert rel_err(N(g.state.v.data), v) < KTOL
    assert rel_err(N(g.state.H.data), H) < KTOL
    assert np.array_equal(N(g.state.mask.data), mask)


@pytest.mark.parametrize("shape", SHAPES[:2])
def test_vanka_single_pass_nonlinear_drag_calving(orc, shape):
    ny, nx = shape
    print(f"vanka pass nonlinear drag shape={shape}", flush=True)
    F = dome_fields(ny=ny, nx=nx, seed=9, mask_frac=0.1)
    P = orc.phys(m=1.0 / 3.0, u_reg=0.5, calving_rate=0.2)
    g = grid_from(F)
    g.sliding.m.set(1.0 / 3.0)
    g.sliding.u_reg.set(0.5)
    g.calving.calving_rate.set(0.2)
    g.forward_operators.vanka_sweep(5.0, 1, freeze_phi=True)
    u, v, H, mask, _ = orc.vanka_sweep(F, P, 1500.0, 5.0)
    assert rel_err(N(g.state.u.data), u) < KTOL
    assert rel_err(N(g.state.v.data), v) < KTOL
    assert rel_err(N(g.state.H.data), H) < KTOL
    assert np.array_equal(N(g.state.mask.data), mask)


@pytest.mark.parametrize("shape,n_pass,m", [((32, 32), 6, 1.0), ((24, 90), 5, 1.0), ((17, 129), 3, 1.0 / 3.0),
                                             ((48, 48), 2, 1.0), ((128, 128), 3, 1.0), ((96, 130), 2, 1.0)])
def test_vanka_fused_passes_match_two_launch_passes(orc, shape, n_pass, m):
    """Forward passes of a small level as one launch each (state alternating
    with the workspace, odd counts copied back) give the state and active set
    of the solve + apply passes (the same arithmetic; the compiler may contract
    a product into a different FMA in the other kernel), and match the oracle."""
    from glide_amd import _lib
    ny, nx = shape
    print(f"vanka fused shape={shape} passes={n_pass} m={m:.3f}", flush=True)
    F = dome_fields(ny=ny, nx=nx, seed=41, mask_frac=0.1)
    out = []
    for fused in (True, False):
        _lib.forward_fused_enable(fused)
        g = grid_from(F)
        g.sliding.m.set(m)
        g.calving.calving_rate.set(0.2)
        g.forward_operators.vanka_sweep(10.0, n_pass, freeze_phi=True)
        torch.cuda.synchronize()
        out.append([N(t.data) for t in (g.state.u, g.state.v, g.state.H, g.state.mask)])
    _lib.forward_fused_enable(True)
    for x, y in zip(out[0][:3], out[1][:3]):
        print(f"  fused vs two-launch max rel diff {rel_err(x, y):.3e}", flush=True)
        assert rel_err(x, y) < 1e-13
    assert np.array_equal(out[0][3], out[1][3])
    if n_pass <= 3:
        P = orc.phys(m=m, calving_rate=0.2)
        Fo = dict(F)
        for _ in range(n_pass):
            u, v, H, mask, _ = orc.vanka_sweep(Fo, P, 1500.0, 10.0)
            Fo.update(u=u, v=v, H=H, mask=mask)
        for x, y in zip(out[0][:3], (Fo["u"], Fo["v"], Fo["H"])):
            assert rel_err(x, y) < 1e-11
        assert np.array_equal(out[0][3], Fo["mask"])


@pytest.mark.parametrize("shape,n_pass,m", [((32, 32), 200, 1.0), ((32, 32), 7, 1.0 / 3.0), ((17, 29), 6, 1.0),
                                             ((5, 64), 4, 1.0), ((48, 64), 3, 1.0), ((2, 3), 5, 1.0),
                                             ((64, 64), 30, 1.0), ((96, 41), 5, 1.0 / 3.0)])
def test_vanka_persistent_passes_match_fused_launches(shape, n_pass, m):
    """All passes of a level at most 64 wide, 96 high in one persistent launch (rows
    split over workgroups, device-wide barriers between passes) give bitwise
    the state and active set of one fused launch per pass."""
    from glide_amd import _lib
    ny, nx = shape
    print(f"vanka persistent shape={shape} passes={n_pass} m={m:.3f}", flush=True)
    F = dome_fields(ny=ny, nx=nx, seed=43, mask_frac=0.1)
    out = []
    for persistent in (True, False):
        _lib.forward_persistent_enable(persistent)
        g = grid_from(F)
        g.sliding.m.set(m)
        g.calving.calving_rate.set(0.2)
        g.forward_operators.vanka_sweep(10.0, n_pass, freeze_phi=True)
        torch.cuda.synchronize()
        out.append([N(t.data) for t in (g.state.u, g.state.v, g.state.H, g.state.mask)])
    _lib.forward_persistent_enable(True)
    assert np.isfinite(out[0][2]).all()
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("shape,n_pass", [((32, 32), 30), ((64, 64), 5), ((128, 128), 3), ((17, 29), 4)])
def test_vanka_mixed_one_launch_forms_match_two_launch_passes(shape, n_pass):
    """The mixed form (FP32 patch Jacobian and LU) through the one-launch
    (fused, persistent) forward passes gives the state and active set of its
    solve + apply passes (the same arithmetic; the compiler may contract a
    product into a different FMA in the other kernel)."""
    from glide_amd import _lib
    ny, nx = shape
    print(f"vanka mixed one-launch shape={shape} passes={n_pass}", flush=True)
    F = dome_fields(ny=ny, nx=nx, seed=45, mask_frac=0.1)
    out = []
    for one_launch in (True, False):
        _lib.forward_fused_enable(one_launch)
        g = grid_from(F)
        g.sliding.m.set(1.0)
        g.calving.calving_rate.set(0.2)
        fo = g.forward_operators
        fo.vanka_config.local_fp32 = True
        fo.vanka_sweep(10.0, n_pass, freeze_phi=True)
        torch.cuda.synchronize()
        out.append([N(t.data) for t in (g.state.u, g.state.v, g.state.H, g.state.mask)])
    _lib.forward_fused_enable(True)
    for x, y in zip(out[0][:3], out[1][:3]):
        print(f"  one-launch vs two-launch max rel diff {rel_err(x, y):.3e}", flush=True)
        assert rel_err(x, y) < 1e-13
    assert np.array_equal(out[0][3], out[1][3])


def test_vanka_passes_with_phi_refresh(orc):
    F = dome_fields(n=32, seed=6)
    g = grid_from(F)
    g.forward_operators.vanka_sweep(10.0, 6)
    P = orc.phys()
    for _ in range(6):
        F["phi"] = orc.grounded(F["H"], F["bed"], 0.1)
        F["u"], F["v"], F["H"], F["mask"], _ = orc.vanka_sweep(F, P, 1500.0, 10.0)
    assert rel_err(N(g.state.u.data), F["u"]) < 1e-11
    assert rel_err(N(g.state.v.data), F["v"]) < 1e-11
    assert rel_err(N(g.state.H.data), F["H"]) < 1e-11
    assert rel_err(N(g.state.phi.data), F["phi"]) < 1e-11


def test_transfers(orc):
    from glide_amd.multigrid import Multigrid
    rng = np.random.default_rng(7)
    mg = Multigrid(2, ny=16, nx=16, dx=100.0, device="cuda")
    for nyc, nxc in ((8, 8), (6, 10)):
        nyf, nxf = 2 * nyc, 2 * nxc
        fu, fv, fc = rng.random((nyf, nxf + 1)), rng.random((nyf + 1, nxf)), rng.random((nyf, nxf))
        assert rel_err(N(mg.restrict_vfacet(T(fu))), orc.restrict_vfacet(fu)) < KTOL
        assert rel_err(N(mg.restrict_hfacet(T(fv))), orc.restrict_hfacet(fv)) < KTOL
        assert rel_err(N(mg.restrict_cell(T(fc))), orc.restrict_cell(fc)) < KTOL
        assert rel_err(N(mg.restrict_cell(T(fc), method="max")), orc.restrict_cell(fc, True)) < KTOL
        cu, cv, cc = rng.random((nyc, nxc + 1)), rng.random((nyc + 1, nxc)), rng.random((nyc, nxc))
        assert rel_err(N(mg.prolongate_vfacet(T(cu))), orc.prolong_vfacet_bilinear(cu, nyf, nxf)) < KTOL
        assert rel_err(N(mg.prolongate_hfacet(T(cv), compat=False)), orc.prolong_hfacet_bilinear(cv, nyf, nxf)) < KTOL
        assert rel_err(N(mg.prolongate_cell(T(cc))), orc.prolong_cell(cc, nyf, nxf)) < KTOL
        assert rel_err(N(mg.prolongate_cell(T(cc), method="bilinear")), orc.prolong_cell(cc, nyf, nxf, True)) < KTOL
    # the reference's v path through the u routine (square grids)
    cv = rng.random((9, 8))
    got = N(mg.prolongate_hfacet(T(cv), compat=True))
    ref = orc.prolong_vfacet_bilinear(np.ascontiguousarray(cv.reshape(8, 9)), 16, 16).reshape(17, 16)
    assert rel_err(got, ref) < KTOL


def _lam(F, seed):
    rng = np.random.default_rng(seed)
    ny, nx = F["H"].shape
    return (rng.standard_normal((ny, nx + 1)), rng.standard_normal((ny + 1, nx)), rng.standard_normal((ny, nx)))


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("m", [1.0, 1.0 / 3.0])
@pytest.mark.parametrize("use_forcing", [False, True])
def test_vjp(orc, shape, m, use_forcing):
    ny, nx = shape
    print(f"vjp shape={shape} m={m} forcing={use_forcing}", flush=True)
    F = dome_fields(ny=ny, nx=nx, seed=8, mask_frac=0.15)
    P = orc.phys(m=m)
    lu, lv, lH = _lam(F, 9)
    g = grid_from(F)
    g.sliding.m.set(m)
    a = g.adjoint
    a.lambda_u.set(lu)
    a.lambda_v.set(lv)
    a.lambda_H.set(lH)
    ao = g.adjoint_operators
    rng = np.random.default_rng(10)
    af = (rng.standard_normal(lu.shape), rng.standard_normal(lv.shape), rng.standard_normal(lH.shape))
    ao.f_u.copy_(T(af[0]))
    ao.f_v.copy_(T(af[1]))
    ao.f_H.copy_(T(af[2]))
    if use_forcing:
        ao.compute_residual(10.0)
        out = (ao.r_u, ao.r_v, ao.r_H)
    else:
        ao.compute_vjp(10.0)
        out = (ao.vjp_u, ao.vjp_v, ao.vjp_H)
    ref = orc.vjp(F, lu, lv, lH, P, 1500.0, 10.0, use_forcing=use_forcing, af=af)
    for x, y in zip(out, ref):
        assert rel_err(N(x), y) < KTOL


@pytest.mark.parametrize("use_cache", [True, False])
def test_vanka_adjoint_pass(orc, use_cache):
    print(f"adjoint pass cache={use_cache}", flush=True)
    F = dome_fields(n=40, seed=11, mask_frac=0.1)
    P = orc.phys()
    lam = _lam(F, 12)
    g = grid_from(F)
    a = g.adjoint
    a.lambda_u.set(lam[0])
    a.lambda_v.set(lam[1])
    a.lambda_H.set(lam[2])
    ao = g.adjoint_operators
    ao.use_cache = use_cache
    rng = np.random.default_rng(13)
    af = (rng.standard_normal(lam[0].shape), rng.standard_normal(lam[1].shape), rng.standard_normal(lam[2].shape))
    ao.f_u.copy_(T(af[0]))
    ao.f_v.copy_(T(af[1]))
    ao.f_H.copy_(T(af[2]))
    ao.vanka_sweep(10.0, 1)
    r = orc.vjp(F, *lam, P, 1500.0, 10.0, use_forcing=True, af=af)
    ref = orc.vanka_adjoint_sweep(F, lam, r, P, 1500.0, 10.0)
    for x, y in zip((a.lambda_u.data, a.lambda_v.data, a.lambda_H.data), ref):
        assert rel_err(N(x), y) < KTOL


@pytest.mark.parametrize("shape", [(36, 52), (70, 130)])
def test_adjoint_cache_matches_uncached_passes(orc, shape):
    """Several adjoint passes from the cached patch factors, fused into one
    launch per pass or as solve + apply, agree with passes that rebuild and
    factorise every patch (mask, calving and nonlinear drag on, non-square
    levels, partial edge tiles); the cache is filled by the sweep call."""
    from glide_amd import _lib
    ny, nx = shape
    print(f"adjoint cache shape={shape}", flush=True)
    F = dome_fields(ny=ny, nx=nx, seed=14, mask_frac=0.15)
    out = []
    for use_cache, fused in ((True, True), (True, False), (False, True)):
        _lib.adjoint_fused_enable(fused)
        g = grid_from(F)
        g.sliding.m.set(1.0 / 3.0)
        g.calving.calving_rate.set(0.3)
        lam = _lam(F, 15)
        for t, x in zip((g.adjoint.lambda_u, g.adjoint.lambda_v, g.adjoint.lambda_H), lam):
            t.set(x)
        ao = g.adjoint_operators
        ao.use_cache = use_cache
        rng = np.random.default_rng(16)
        for t in (ao.f_u, ao.f_v, ao.f_H):
            t.copy_(T(rng.standard_normal(tuple(t.shape))))
        ao.vanka_sweep(10.0, 6)
        torch.cuda.synchronize()
        out.append([N(t.data) for t in (g.adjoint.lambda_u, g.adjoint.lambda_v, g.adjoint.lambda_H)])
    _lib.adjoint_fused_enable(True)
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)
    for x, y in zip(out[0], out[2]):
        assert rel_err(x, y) < 1e-13


@pytest.mark.parametrize("shape", [(16, 32), (17, 29), (8, 6)])
def test_adjoint_coarse_workgroup_matches_launches(shape):
    """All cached adjoint passes of a level of at most 512 cells in one
    workgroup agree with the transpose operator and the fused pass launched
    once per pass (linear drag, mask and calving on, non-square levels)."""
    from glide_amd import _lib
    ny, nx = shape
    print(f"adjoint coarse workgroup shape={shape}", flush=True)
    F = dome_fields(ny=ny, nx=nx, seed=21, mask_frac=0.15)
    out = []
    for coarse in (True, False):
        _lib.adjoint_coarse_enable(coarse)
        g = grid_from(F)
        g.sliding.m.set(1.0)
        g.calving.calving_rate.set(0.3)
        for t, x in zip((g.adjoint.lambda_u, g.adjoint.lambda_v, g.adjoint.lambda_H), _lam(F, 22)):
            t.set(x)
        ao = g.adjoint_operators
        ao.use_cache = True
        rng = np.random.default_rng(23)
        for t in (ao.f_u, ao.f_v, ao.f_H):
            t.copy_(T(rng.standard_normal(tuple(t.shape))))
        ao.vanka_sweep(10.0, 7)
        torch.cuda.synchronize()
        out.append([N(t.data) for t in (g.adjoint.lambda_u, g.adjoint.lambda_v, g.adjoint.lambda_H)])
    _lib.adjoint_coarse_enable(True)
    for x, y in zip(out[0], out[1]):
        assert rel_err(x, y) < 1e-13


@pytest.mark.parametrize("shape", [(36, 52), (70, 130), (17, 29), (1030, 200)])
def test_adjoint_pass_fused_matches_launches(shape):
    """A cached adjoint pass as one launch (transpose operator, patch solves
    and update, each tile recomputing its left/top neighbours' residuals)
    agrees with the transpose operator and the fused pass launched separately
    (linear drag, mask and calving on, non-square levels, partial tiles)."""
    from glide_amd import _lib
    ny, nx = shape
    print(f"adjoint pass fused shape={shape}", flush=True)
    F = dome_fields(ny=ny, nx=nx, seed=24, mask_frac=0.15)
    out = []
    _lib.adjoint_coarse_enable(False)
    try:
        for fused in (True, False):
            _lib.adjoint_pass_fused_enable(fused)
            g = grid_from(F)
            g.sliding.m.set(1.0)
            g.calving.calving_rate.set(0.3)
            for t, x in zip((g.adjoint.lambda_u, g.adjoint.lambda_v, g.adjoint.lambda_H), _lam(F, 25)):
                t.set(x)
            ao = g.adjoint_operators
            ao.use_cache = True
            rng = np.random.default_rng(26)
            for t in (ao.f_u, ao.f_v, ao.f_H):
                t.copy_(T(rng.standard_normal(tuple(t.shape))))
            ao.vanka_sweep(10.0, 5)
            torch.cuda.synchronize()
            out.append([N(t.data) for t in (g.adjoint.lambda_u, g.adjoint.lambda_v, g.adjoint.lambda_H)])
    finally:
        _lib.adjoint_pass_fused_enable(True)
        _lib.adjoint_coarse_enable(True)
    for x, y in zip(out[0], out[1]):
        assert rel_err(x, y) < 1e-13


@pytest.mark.parametrize("shape,n_pass", [((32, 32), 200), ((17, 29), 6), ((32, 64), 7), ((5, 64), 3), ((2, 3), 4)])
def test_adjoint_persistent_matches_fused_launches(shape, n_pass):
    """All cached adjoint passes of a level at most 64 wide and 32 high in one persistent
    launch give the multipliers of one fused launch per pass (the same
    arithmetic; the compiler may contract a product into a different FMA in
    the other kernel, so the bar is relative round-off after up to 200
    passes)."""
    from glide_amd import _lib
    ny, nx = shape
    print(f"adjoint persistent shape={shape} passes={n_pass}", flush=True)
    F = dome_fields(ny=ny, nx=nx, seed=27, mask_frac=0.15)
    out = []
    _lib.adjoint_coarse_enable(False)
    try:
        for persistent in (True, False):
            _lib.adjoint_persistent_enable(persistent)
            g = grid_from(F)
            g.sliding.m.set(1.0)
            g.calving.calving_rate.set(0.3)
            for t, x in zip((g.adjoint.lambda_u, g.adjoint.lambda_v, g.adjoint.lambda_H), _lam(F, 28)):
                t.set(x)
            ao = g.adjoint_operators
            ao.use_cache = True
            rng = np.random.default_rng(29)
            for t in (ao.f_u, ao.f_v, ao.f_H):
                t.copy_(T(rng.standard_normal(tuple(t.shape))))
            ao.vanka_sweep(10.0, n_pass)
            torch.cuda.synchronize()
            out.append([N(t.data) for t in (g.adjoint.lambda_u, g.adjoint.lambda_v, g.adjoint.lambda_H)])
    finally:
        _lib.adjoint_persistent_enable(True)
        _lib.adjoint_coarse_enable(True)
    assert all(np.isfinite(x).all() for x in out[0])
    for x, y in zip(out[0], out[1]):
        print(f"  persistent vs fused launches rel diff {rel_err(x, y):.3e}", flush=True)
        assert rel_err(x, y) < 1e-12


@pytest.mark.parametrize("shape,n_pass,m", [((32, 32), 60, 1.0), ((96, 41), 5, 1.0 / 3.0)])
def test_persistent_fenced_barriers_match_launches(shape, n_pass, m):
    """The persistent launches' cross-XCD barrier path (cache write-back before
    each arrival, invalidation after each wait), forced on: the forward passes
    equal one fused launch per pass bitwise, the adjoint passes to round-off."""
    from glide_amd import _lib
    ny, nx = shape
    print(f"persistent fenced shape={shape} passes={n_pass}", flush=True)
    F = dome_fields(ny=ny, nx=nx, seed=47, mask_frac=0.1)
    fwd, adj = [], []
    _lib.adjoint_coarse_enable(False)
    try:
        for persistent in (True, False):
            _lib.persistent_fenced_enable(persistent)
            _lib.forward_persistent_enable(persistent)
            _lib.adjoint_persistent_enable(persistent)
            g = grid_from(F)
            g.sliding.m.set(m)
            g.calving.calving_rate.set(0.2)
            g.forward_operators.vanka_sweep(10.0, n_pass, freeze_phi=True)
            for t, x in zip((g.adjoint.lambda_u, g.adjoint.lambda_v, g.adjoint.lambda_H), _lam(F, 48)):
                t.set(x)
            ao = g.adjoint_operators
            ao.use_cache = True
            rng = np.random.default_rng(49)
            for t in (ao.f_u, ao.f_v, ao.f_H):
                t.copy_(T(rng.standard_normal(tuple(t.shape))))
            ao.vanka_sweep(10.0, n_pass)
            torch.cuda.synchronize()
            fwd.append([N(t.data) for t in (g.state.u, g.state.v, g.state.H, g.state.mask)])
            adj.append([N(t.data) for t in (g.adjoint.lambda_u, g.adjoint.lambda_v, g.adjoint.lambda_H)])
    finally:
        _lib.persistent_fenced_enable(False)
        _lib.forward_persistent_enable(True)
        _lib.adjoint_persistent_enable(True)
        _lib.adjoint_coarse_enable(True)
    for x, y in zip(fwd[0], fwd[1]):
        assert np.array_equal(x, y)
    for x, y in zip(adj[0], adj[1]):
        assert np.isfinite(x).all()
        assert rel_err(x, y) < 1e-12


@pytest.mark.parametrize("shape", SHAPES)
def test_gradients(orc, shape):
    ny, nx = shape
    F = dome_fields(ny=ny, nx=nx, seed=14)
    P = orc.phys(m=1.0 / 3.0)
    lu, lv, lH = _lam(F, 15)
    g = grid_from(F)
    g.sliding.m.set(1.0 / 3.0)
    a = g.adjoint
    a.lambda_u.set(lu)
    a.lambda_v.set(lv)
    a.lambda_H.set(lH)
    ao = g.adjoint_operators
    ao.compute_gradient_beta()
    ao.compute_gradient_bed()
    ao.compute_gradient_H_prev(10.0)
    ao.compute_gradient_smb()
    assert rel_err(N(g.sliding.beta.grad), orc.grad_beta(F, lu, lv, P)) < KTOL
    assert rel_err(N(g.geometry.bed.grad), orc.grad_bed(F, lu, lv, 1500.0)) < KTOL
    assert rel_err(N(g.state.H_prev.grad), -lH / 10.0) < KTOL
    assert rel_err(N(g.forcing.smb.grad), -lH) < KTOL


@pytest.mark.parametrize("shape,m", [((32, 32), 1.0), ((24, 40), 1.0 / 3.0)])
def test_rate_factor_gradient_matches_residual_differences(orc, shape, m):
    """lambda^T dR/dB (no counterpart in the reference) against central
    differences of lambda^T R(B) with the oracle's residual, in random
    directions (active-set cells included; Dirichlet rows do not depend on B)."""
    ny, nx = shape
    print(f"rate-factor gradient shape={shape} m={m}", flush=True)
    F = dome_fields(ny=ny, nx=nx, seed=21, mask_frac=0.1)
    P = orc.phys(m=m)
    lu, lv, lH = _lam(F, 22)
    g = grid_from(F)
    g.sliding.m.set(m)
    a = g.adjoint
    a.lambda_u.set(lu)
    a.lambda_v.set(lv)
    a.lambda_H.set(lH)
    g.adjoint_operators.compute_gradient_B(10.0)
    torch.cuda.synchronize()
    gB = N(g.rheology.B.grad)
    rng = np.random.default_rng(23)
    eps = 1e-6
    for _ in range(3):
        d = rng.standard_normal((ny, nx)) * F["B"]
        Fp, Fm = dict(F), dict(F)
        Fp["B"], Fm["B"] = F["B"] + eps * d, F["B"] - eps * d
        rp, rm = orc.residual(Fp, P, 1500.0, 10.0), orc.residual(Fm, P, 1500.0, 10.0)
        fd = sum(float(np.sum(lam * (x - y))) for lam, x, y in zip((lu, lv, lH), rp, rm)) / (2.0 * eps)
        an = float(np.sum(gB * d))
        print(f"  <g_B, d> = {an:.10e}, central difference {fd:.10e}", flush=True)
        assert abs(an - fd) <= 1e-7 * abs(fd)


def _dome_models(orc, n, L, seed=0, dx=1500.0):
    from glide_amd import IceDynamics
    from glide_amd.synthetic import dome_problem, install
    p = dome_problem(n, dx, seed=seed)
    model = IceDynamics(ny=n, nx=n, dx=dx, n_levels=L, device="cuda")
    install(model.mg, p)
    om = orc.OracleMG(n, n, dx, L)
    for k in ("bed", "B", "beta", "smb"):
        om.set(k, getattr(p, k))
    om.set("H", p.H)
    om.set("H_prev", p.H)
    return p, model, om


@pytest.mark.parametrize("n,L", [(64, 4), (32, 3)])
def test_fascd_solve(orc, n, L):
    print(f"fascd solve n={n} L={L}", flush=True)
    p, model, om = _dome_models(orc, n, L)
    fs = model.forward_solver
    fs.fas_options.set(maximum_vcycles=3, relative_tolerance=0.0, absolute_tolerance=0.0, report_norms=False)
    rep = fs.solve(10.0)
    torch.cuda.synchronize()
    orep = om.fascd_solve(10.0, orc.OracleMG.fas_config(vcycles=3, rtol=0.0, atol=0.0))
    assert rep.vcycles == orep.vcycles == 3
    g = model.mg[0]
    for k in ("u", "v", "H"):
        assert rel_err(N(getattr(g.state, k).data), om.get(k)) < STOL, k
    assert np.array_equal(N(g.state.mask.data), om.get("mask"))
    for k in range(3):
        assert abs(rep.norms[k] - orep.norms[k]) <= 1e-9 * orep.norms[0]
    assert rep.norms[2] < 1e-2 * rep.initial_norm


@pytest.mark.parametrize("n,L", [(64, 4)])
def test_mixed_form_solve_against_fp64_oracle(orc, n, L):
    """Mixed form (FP32 patch Jacobian and LU, FP64 residual/stopping/update):
    three V-cycles stay within the stated distance of the FP64 oracle's, and
    converge as the FP64 form does."""
    print(f"mixed fascd solve n={n} L={L}", flush=True)
    p, model, om = _dome_models(orc, n, L)
    fs = model.forward_solver
    fs.fas_options.set(maximum_vcycles=3, relative_tolerance=0.0, absolute_tolerance=0.0, report_norms=False)
    fs.vanka_options.set(local_fp32=True)
    rep = fs.solve(10.0)
    torch.cuda.synchronize()
    orep = om.fascd_solve(10.0, orc.OracleMG.fas_config(vcycles=3, rtol=0.0, atol=0.0))
    g = model.mg[0]
    errs = {k: rel_err(N(getattr(g.state, k).data), om.get(k)) for k in ("u", "v", "H")}
    print("mixed vs FP64 oracle max rel err:", errs, "norms", list(rep.norms[:3]), list(orep.norms[:3]), flush=True)
    for k, e in errs.items():
        assert e < 1e-4, (k, e)
    assert rep.norms[2] < 1e-2 * rep.initial_norm


def test_asynchronous_fixed_cycle_solves_match(orc):
    """asynchronous=True (fixed cycles, no report) returns None without
    waiting and leaves the same state, multipliers and gradients."""
    out = []
    for asynchronous in (False, True):
        p, model, om = _dome_models(orc, 64, 4, seed=7)
        for s in (model.forward_solver, model.adjoint_solver):
            s.fas_options.set(maximum_vcycles=2, relative_tolerance=0.0, absolute_tolerance=0.0, report_norms=False,
                              asynchronous=asynchronous)
        rep = model.forward_solver.solve(10.0)
        assert (rep is None) == asynchronous
        g = model.mg[0]
        model.backward(0.0, 10.0, dJdu=g.state.u.data.clone(), dJdv=g.state.v.data.clone(), dJdH=None)
        torch.cuda.synchronize()
        out.append([N(t) for t in (g.state.u.data, g.state.H.data, g.adjoint.lambda_u.data, g.sliding.beta.grad)])
    for x, y in zip(*out):
        assert np.array_equal(x, y)


def test_fascd_solve_tolerance_stop(orc):
    p, model, om = _dome_models(orc, 32, 3, seed=3)
    fs = model.forward_solver
    fs.fas_options.set(maximum_vcycles=10, relative_tolerance=1e-3, absolute_tolerance=5.0, report_norms=False)
    rep = fs.solve(10.0)
    orep = om.fascd_solve(10.0, orc.OracleMG.fas_config(vcycles=10, rtol=1e-3, atol=5.0))
    assert rep.vcycles == orep.vcycles
    assert rel_err(N(model.mg[0].state.u.data), om.get("u")) < STOL


@pytest.mark.parametrize("n,L,use_cache", [(64, 4, True), (64, 4, False)])
def test_forward_adjoint_gradients(orc, n, L, use_cache):
    print(f"forward+adjoint n={n} L={L} cache={use_cache}", flush=True)
    p, model, om = _dome_models(orc, n, L, seed=5)
    for lev in range(L):
        model.mg[lev].adjoint_operators.use_cache = use_cache
    model.forward_solver.fas_options.set(maximum_vcycles=3, relative_tolerance=0.0, absolute_tolerance=0.0,
                                         report_norms=False)
    model.adjoint_solver.fas_options.set(maximum_vcycles=3, relative_tolerance=0.0, absolute_tolerance=0.0,
                                         report_norms=False)
    model.forward(0.0, 10.0, update_geometry=False)
    cfg = orc.OracleMG.fas_config(vcycles=3, rtol=0.0, atol=0.0)
    om.fascd_solve(10.0, cfg)
    g = model.mg[0]
    u, v = N(g.state.u.data), N(g.state.v.data)
    rng = np.random.default_rng(17)
    u_obs = om.get("u") + 5.0 * rng.standard_normal(u.shape)
    v_obs = om.get("v") + 5.0 * rng.standard_normal(v.shape)
    model.backward(0.0, 10.0, dJdu=T(u - u_obs), dJdv=T(v - v_obs), dJdH=None)
    torch.cuda.synchronize()
    om.put("af_u", -(om.get("u") - u_obs))
    om.put("af_v", -(om.get("v") - v_obs))
    om.put("af_H", np.zeros((n, n)))
    om.fas_adjoint_solve(10.0, cfg)
    om.gradients(10.0)
    a = g.adjoint
    for name, t in (("lambda_u", a.lambda_u.data), ("lambda_v", a.lambda_v.data), ("lambda_H", a.lambda_H.data)):
        assert rel_err(N(t), om.get(name)) < STOL, name
    assert rel_err(N(g.sliding.beta.grad), om.get("g_beta")) < STOL
    assert rel_err(N(g.geometry.bed.grad), om.get("g_bed")) < STOL
    assert rel_err(N(g.state.H_prev.grad), om.get("g_H_prev")) < STOL
    assert rel_err(N(g.forcing.smb.grad), om.get("g_smb")) < STOL


def test_glidestep_autograd(orc):
    from glide_amd.torch import GlideStep
    p, model, om = _dome_models(orc, 32, 3, seed=9)
    for s in (model.forward_solver, model.adjoint_solver):
        s.fas_options.set(maximum_vcycles=2, relative_tolerance=0.0, absolute_tolerance=0.0, report_norms=False)
    H_prev = T(p.H).requires_grad_(True)
    bed = T(p.bed).requires_grad_(True)
    beta = T(p.beta).requires_grad_(True)
    smb = T(p.smb).requires_grad_(True)
    u, v, H = GlideStep.apply(0.0, 10.0, model, 0, H_prev, bed, beta, smb)
    loss = 0.5 * (u ** 2).sum() + 0.5 * (v ** 2).sum()
    loss.backward()
    cfg = orc.OracleMG.fas_config(vcycles=2, rtol=0.0, atol=0.0)
    om.fascd_solve(10.0, cfg)
    assert rel_err(N(u), om.get("u")) < STOL
    om.put("af_u", -om.get("u"))
    om.put("af_v", -om.get("v"))
    om.put("af_H", np.zeros((32, 32)))
    om.fas_adjoint_solve(10.0, cfg)
    om.gradients(10.0)
    assert rel_err(N(beta.grad), om.get("g_beta")) < STOL
    assert rel_err(N(bed.grad), om.get("g_bed")) < STOL
    assert rel_err(N(H_prev.grad), om.get("g_H_prev")) < STOL
    assert rel_err(N(smb.grad), om.get("g_smb")) < STOL


def test_full_size_kernels_512(orc):
    """At the benchmark size: one residual, one smoothing pass and one VJP
    against the oracle, plus run-to-run bitwise determinism of the pass."""
    print("512^2 kernels", flush=True)
    F = dome_fields(n=512, seed=21, mask_frac=0.05)
    P = orc.phys()
    g = grid_from(F)
    fo = g.forward_operators
    fo.compute_residual(10.0, freeze_phi=True)
    ru, rv, rH = orc.residual(F, P, 1500.0, 10.0)
    for a, b in zip((fo.r_u, fo.r_v, fo.r_H), (ru, rv, rH)):
        assert rel_err(N(a), b) < KTOL
    fo.vanka_sweep(10.0, 1, freeze_phi=True)
    u1 = N(g.state.u.data).copy()
    ref = orc.vanka_sweep(F, P, 1500.0, 10.0)
    assert rel_err(u1, ref[0]) < KTOL
    assert rel_err(N(g.state.H.data), ref[2]) < KTOL
    g2 = grid_from(F)
    g2.forward_operators.vanka_sweep(10.0, 1, freeze_phi=True)
    assert np.array_equal(N(g2.state.u.data), u1)


@pytest.mark.parametrize("n_pass", [5, 6])
def test_adjoint_passes_512_deterministic(n_pass):
    """At the benchmark size, cached adjoint passes in every launch form give
    bitwise the same multipliers run to run (tiles read their neighbours'
    multipliers, so an update in place would race), and the one-launch form
    agrees with the separate transpose operator and pass."""
    from glide_amd import _lib
    print(f"512^2 adjoint passes n={n_pass}", flush=True)
    F = dome_fields(n=512, seed=27, mask_frac=0.05)
    lam = _lam(F, 28)
    rng = np.random.default_rng(29)
    forcing = None

    def run(fused):
        nonlocal forcing
        _lib.adjoint_pass_fused_enable(fused)
        g = grid_from(F)
        g.sliding.m.set(1.0)
        for t, x in zip((g.adjoint.lambda_u, g.adjoint.lambda_v, g.adjoint.lambda_H), lam):
            t.set(x)
        ao = g.adjoint_operators
        ao.use_cache = True
        if forcing is None:
            forcing = [rng.standard_normal(tuple(t.shape)) for t in (ao.f_u, ao.f_v, ao.f_H)]
        for t, x in zip((ao.f_u, ao.f_v, ao.f_H), forcing):
            t.copy_(T(x))
        ao.vanka_sweep(10.0, n_pass)
        torch.cuda.synchronize()
        return [N(t.data).copy() for t in (g.adjoint.lambda_u, g.adjoint.lambda_v, g.adjoint.lambda_H)]

    try:
        a, b, c = run(True), run(True), run(False)
    finally:
        _lib.adjoint_pass_fused_enable(True)
    for x, y, z in zip(a, b, c):
        assert np.array_equal(x, y)
        assert rel_err(x, z) < 1e-13
