"""Shared helpers of the test suite: seeded states and error metrics."""
from __future__ import annotations

import numpy as np


def rel_err(a, b) -> float:
    """max|a - b| / max(max|b|, tiny)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(float(np.max(np.abs(b))) if b.size else 0.0, 1e-300)
    return float(np.max(np.abs(a - b))) / den if a.size else 0.0


def l2_rel(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(float(np.linalg.norm(b)), 1e-300)
    return float(np.linalg.norm(a - b)) / den


def dome_fields(n=32, seed=0, dx=1500.0, with_velocity=True, mask_frac=0.0, ny=None, nx=None):
    """A state of the dome workload as a dict of float64 numpy arrays in the
    reference's layouts, with smooth non-zero velocities when requested."""
    from glide_amd.synthetic import dome_problem, gaussian_random_field
    ny = ny or n
    nx = nx or n
    p = dome_problem(max(ny, nx), dx, seed=seed)
    rng = np.random.default_rng(seed + 100)
    sl = (slice(0, ny), slice(0, nx))
    F = {"bed": p.bed[sl].copy(), "H": p.H[sl].copy(), "B": p.B[sl].copy(), "beta": p.beta[sl].copy(),
         "smb": p.smb[sl].copy()}
    F["H"] = np.maximum(F["H"], 0.1) + 5.0 * rng.random((ny, nx))
    if with_velocity:
        gu = gaussian_random_field(rng, ny, nx + 1, dx, 8 * dx)
        gv = gaussian_random_field(rng, ny + 1, nx, dx, 8 * dx)
        F["u"], F["v"] = 100.0 * gu, 80.0 * gv
    else:
        F["u"], F["v"] = np.zeros((ny, nx + 1)), np.zeros((ny + 1, nx))
    F["phi"] = 1.0 / (1.0 + np.exp(-np.clip(0.1 * (F["bed"] + 0.917 * F["H"]), -20, 20)))
    F["mask"] = (rng.random((ny, nx)) < mask_frac).astype(np.float64)
    F["f_u"] = 0.3 * rng.standard_normal((ny, nx + 1))
    F["f_v"] = 0.3 * rng.standard_normal((ny + 1, nx))
    F["H_prev"] = F["H"] + rng.standard_normal((ny, nx))
    F["f_H"] = F["H_prev"] / 10.0 + F["smb"]
    F["gamma"] = np.full((ny, nx), 0.1) + 0.05 * rng.random((ny, nx))
    return F


def enth_case(ny, nx, nz, seed, temperate_frac=0.3):
    """Seeded enthalpy state in the reference's layouts (numpy, SI, scaled E)."""
    from oracle.oracle import BETA_CC, C_I, E_SCALE, GRAVITY, RHO_I_E, SEC_PER_YR, T_MELT, T_REF
    rng = np.random.default_rng(seed)
    H = 200.0 + 1800.0 * rng.random((ny, nx))
    r = rng.random((ny, nx))
    H[r < 0.08] = 20.0 * rng.random(int((r < 0.08).sum()))      # thin columns (< h_thin = 25 m)
    H[r < 0.03] = 0.0                                           # ice-free
    sigma = np.linspace(0.0, 1.0, nz)
    T_pmp = T_MELT - BETA_CC * RHO_I_E * GRAVITY * (1.0 - sigma)[None, None, :] * H[:, :, None]
    Temp = T_pmp - 35.0 * rng.random((ny, nx, nz)) * sigma[None, None, :] - 0.5 * rng.random((ny, nx, nz))
    E = C_I * (Temp - T_REF) / E_SCALE
    warm = rng.random((ny, nx)) < temperate_frac
    E[warm, :max(2, nz // 4)] += 0.02 * rng.random((int(warm.sum()), max(2, nz // 4)))
    C = {"H": H, "H_prev": H + 5.0 * rng.standard_normal((ny, nx)) * (H > 0), "E": E,
         "E_prev": E + 0.01 * rng.standard_normal(E.shape)}
    shape_u = (1.0 + sigma)[:, None, None] / 2.0
    C["u3d"] = shape_u * (60.0 * rng.standard_normal((1, ny, nx + 1))) / SEC_PER_YR
    C["v3d"] = shape_u * (45.0 * rng.standard_normal((1, ny + 1, nx))) / SEC_PER_YR
    C["omega"] = 0.3 * rng.standard_normal((ny, nx, nz)) / SEC_PER_YR
    Ts = 235.0 + 30.0 * rng.random((ny, nx))
    C["E_surface"] = C_I * (np.minimum(Ts, T_MELT) - T_REF) / E_SCALE
    C["Q_geo"] = 0.05 + 0.02 * rng.random((ny, nx))
    C["Q_fh"] = 0.02 * rng.random((ny, nx))
    C["phi_strain"] = 2e-6 * rng.random((ny, nx, nz))
    return C
