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
