"""Seeded synthetic workloads of the SSA step.

Dome on a Gaussian-random-field bed (glide's head units: stresses / (rho_i g),
velocities in m/a, time in years; B converted as
B_SI / (rho_i g SEC_PER_YR^(1/n)), glide/model.py:206-213):

* bed  b = -100 m + 300 m * GRF(squared-exponential covariance, 20 km)
* thickness H = 2000 m * max(0, 1 - (r/R)^(4/3))^(3/8), R = 0.45 * domain width
* Glen n = 3, B = 1.7e8 Pa s^(1/3)
* friction beta = 1e3 Pa a/m * exp(GRF, sigma = 1)
* SMB = 0.5 m/a - 1e-3/a * surface elevation (flotation surface)

The GRF is spectral synthesis of seeded white noise with the covariance's
spectral density, normalised to zero mean and unit variance.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

SEC_PER_YR = 365.25 * 86400.0
RHO_I = 917.0
GRAVITY = 9.81


def gaussian_random_field(rng: np.random.Generator, ny: int, nx: int, dx: float, corr_len: float) -> np.ndarray:
    """Periodic GRF with covariance exp(-r^2 / (2 l^2)), unit variance."""
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=dx)
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    amp = np.exp(-k2 * corr_len ** 2 / 4.0)  # sqrt of the spectral density
    noise = rng.standard_normal((ny, nx))
    f = np.fft.ifft2(np.fft.fft2(noise) * amp).real
    f -= f.mean()
    s = f.std()
    return f / s if s > 0 else f


def dome_thickness(ny: int, nx: int, dx: float, height: float = 2000.0, radius_frac: float = 0.45) -> np.ndarray:
    y = (np.arange(ny) + 0.5) * dx - 0.5 * ny * dx
    x = (np.arange(nx) + 0.5) * dx - 0.5 * nx * dx
    r = np.sqrt(y[:, None] ** 2 + x[None, :] ** 2)
    R = radius_frac * nx * dx
    return height * np.maximum(0.0, 1.0 - (r / R) ** (4.0 / 3.0)) ** (3.0 / 8.0)


def flotation_surface(bed: np.ndarray, H: np.ndarray) -> np.ndarray:
    return np.maximum(bed + H, (1.0 - 0.917) * H)


@dataclass
class DomeProblem:
    ny: int
    nx: int
    dx: float
    bed: np.ndarray
    H: np.ndarray
    B: np.ndarray
    beta: np.ndarray
    smb: np.ndarray
    seed: int


def dome_problem(n: int = 64, dx: float = 1500.0, seed: int = 0, corr_len: float = 20e3,
                 grounding_line: bool = False) -> DomeProblem:
    """The dome workload on an n x n grid (fields in head-year units).

    grounding_line: beta = 0 where the bed lies below -300 m and the ice (or
    an ice-free cell) is afloat there, 0.917 H + b < 0: floating ice over the
    deep parts of the seeded bed carries no basal drag.  (Zeroing beta under
    grounded ice as well makes frictionless grounded patches whose forward
    solve does not converge at 2048 x 2048.)"""
    rng = np.random.default_rng(seed)
    bed = -100.0 + 300.0 * gaussian_random_field(rng, n, n, dx, corr_len)
    H = dome_thickness(n, n, dx)
    B = np.full((n, n), 1.7e8 / (RHO_I * GRAVITY * SEC_PER_YR ** (1.0 / 3.0)))
    beta = 1e3 * np.exp(gaussian_random_field(rng, n, n, dx, corr_len)) / (RHO_I * GRAVITY)
    if grounding_line:
        beta = np.where((bed < -300.0) & (0.917 * H + bed < 0.0), 0.0, beta)
    smb = 0.5 - 1e-3 * flotation_surface(bed, H)
    return DomeProblem(n, n, dx, bed, H, B, beta, smb, seed)


def install(mg, prob: DomeProblem, level: int = 0) -> None:
    """Write the problem into a Multigrid (restricting down) with u = v = 0."""
    mg.geometry.bed.set(prob.bed, start_level=level)
    mg.rheology.B.set(prob.B, start_level=level)
    mg.sliding.beta.set(prob.beta, start_level=level)
    mg.forcing.smb.set(prob.smb, start_level=level)
    mg.state.H.set(prob.H, start_level=level)
    mg.state.H_prev.set(prob.H, start_level=level)
    mg.state.u.set(0.0, start_level=level)
    mg.state.v.set(0.0, start_level=level)
    mg.state.mask.set(0.0, start_level=level)


def thermal_forcing(prob: DomeProblem, lapse: float = 6.5e-3, T_ref_surface: float = 273.15 - 30.0,
                    z_ref: float = 2000.0, q_geo: float = 0.05):
    """Surface temperature T_s = -30 C + 6.5e-3 K/m (2000 m - s) on the
    flotation surface s (K) and a uniform geothermal flux (W/m^2) of the
    coupled enthalpy workload."""
    s = flotation_surface(prob.bed, prob.H)
    Ts = T_ref_surface + lapse * (z_ref - s)
    return Ts, np.full_like(Ts, q_geo)
