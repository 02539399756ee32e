"""An ensemble of independent glacier domains of uneven sizes stepped
together on one GPU (no counterpart in the reference, which runs one domain
per process: glide/model.py:25-39, glide/multigrid.py:212-235).

Each domain is an ``IceDynamics`` of its own size and level count.  A step
enqueues every domain's forward and adjoint solve with a fixed number of
iterations and no host wait (``asynchronous``), replayed from the library's
cached HIP graphs (one per domain and solve), round-robin over a few streams,
so that the small domains' latency-bound launches overlap.  The persistent
coarse-level kernels of the FP64 form assume their workgroups are resident
together, which concurrent streams do not guarantee, so the ensemble turns
them off for its steps (the red-black passes of the mixed form have none).

:func:`valley_glacier` is the seeded mountain-glacier workload: a valley bed
b = z0 - slope x + 150 m (2 y / W - 1)^2 + 30 m x GRF, initial thickness 0.3 of
the valley depth, SMB linear in elevation about a seeded equilibrium-line
altitude.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Sequence

import numpy as np
import torch

from . import _lib
from .model import IceDynamics
from .synthetic import GRAVITY, RHO_I, SEC_PER_YR, gaussian_random_field


@dataclass
class GlacierProblem:
    n: int
    dx: float
    bed: np.ndarray
    H: np.ndarray
    B: np.ndarray
    beta: np.ndarray
    smb: np.ndarray
    seed: int


def valley_glacier(n: int, dx: float = 100.0, seed: int = 0) -> GlacierProblem:
    """An n x n valley glacier (x along the valley, y across it; head-year
    units as glide_amd.synthetic): z0 in [2500, 3500] m, slope in
    [0.05, 0.15], a 150 m deep parabolic cross-section, 30 m of roughness
    (GRF, 1 km correlation), H = 0.3 x the depth below the valley rim,
    beta = 1e3 Pa a/m x exp(0.5 GRF), SMB = 5e-3 /a x (s - ELA) within
    [-4, 2] m/a with the ELA seeded around the mean surface."""
    rng = np.random.default_rng(seed)
    W = n * dx
    x = (np.arange(n) + 0.5) * dx
    y = (np.arange(n) + 0.5) * dx
    z0 = 2500.0 + 1000.0 * rng.random()
    slope = 0.05 + 0.1 * rng.random()
    across = (2.0 * y / W - 1.0) ** 2
    bed = z0 - slope * x[None, :] + 150.0 * across[:, None] + 30.0 * gaussian_random_field(rng, n, n, dx, 1000.0)
    depth = 150.0 * (1.0 - across)
    H = np.broadcast_to(0.3 * depth[:, None], (n, n)).copy()
    B = np.full((n, n), 1.7e8 / (RHO_I * GRAVITY * SEC_PER_YR ** (1.0 / 3.0)))
    beta = 1e3 * np.exp(0.5 * gaussian_random_field(rng, n, n, dx, 1000.0)) / (RHO_I * GRAVITY)
    s = bed + H
    ela = float(s.mean()) + 200.0 * (rng.random() - 0.5)
    smb = np.clip(5e-3 * (s - ela), -4.0, 2.0)
    return GlacierProblem(n, dx, bed, H, B, beta, smb, seed)


def ensemble_sizes(count: int, lo: int = 48, hi: int = 384, seed: int = 0, multiple: int = 8) -> List[int]:
    """Side lengths drawn log-uniformly in [lo, hi] (seeded), rounded to a
    multiple of `multiple` cells."""
    rng = np.random.default_rng(seed)
    raw = np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))
    return [int(max(lo, min(hi, multiple * round(v / multiple)))) for v in raw]


def levels_for(n: int, max_levels: int = 5, min_coarse: int = 6) -> int:
    """The most levels (at most max_levels) whose coarsest level keeps at
    least min_coarse cells per side on an exact halving."""
    L = 1
    while L < max_levels and n % (1 << L) == 0 and (n >> L) >= min_coarse:
        L += 1
    return L


def install_glacier(mg, p: GlacierProblem) -> None:
    mg.geometry.bed.set(p.bed)
    mg.rheology.B.set(p.B)
    mg.sliding.beta.set(p.beta)
    mg.forcing.smb.set(p.smb)
    mg.state.H.set(p.H)
    mg.state.H_prev.set(p.H)
    mg.state.u.set(0.0)
    mg.state.v.set(0.0)
    mg.state.mask.set(0.0)


class Ensemble:
    """Independent domains stepped together: ``step(t, dt)`` runs one
    forward solve (advancing the geometry) and one adjoint solve with the
    gradients of 0.5 |u|^2 + 0.5 |v|^2 on every domain, each with
    `iterations` fixed iterations (mixed: FP32 defect-correction iterations;
    fp64: reference-form V-cycles), spread over `streams` streams."""

    def __init__(self, problems: Sequence[GlacierProblem], precision: str = "mixed", iterations: int = 3,
                 streams: int = 8, device: str = "cuda"):
        self.problems = list(problems)
        self.precision = precision
        self.models = []
        for p in self.problems:
            m = IceDynamics(ny=p.n, nx=p.n, dx=p.dx, n_levels=levels_for(p.n), device=device, precision=precision)
            install_glacier(m.mg, p)
            for s in (m.forward_solver, m.adjoint_solver):
                s.fas_options.set(maximum_vcycles=iterations, relative_tolerance=0.0, absolute_tolerance=0.0,
                                  report_norms=False, asynchronous=True)
                mo = s.mixed_options
                mo.max_outer = iterations
                mo.relative_tolerance = 0.0
                mo.absolute_tolerance = 0.0
                mo.asynchronous = True
            self.models.append(m)
        dev = self.models[0].mg[0].state.u.data.device
        self.streams = [torch.cuda.Stream(device=dev) for _ in range(max(1, int(streams)))]
        # two solves per domain, plus the graphs of other callers
        _lib.graphs_capacity(max(8, 2 * len(self.models) + 8))

    def step(self, t: float, dt: float) -> None:
        caller = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(caller)
        _lib.forward_persistent_enable(False)
        _lib.adjoint_persistent_enable(False)
        try:
            for k, m in enumerate(self.models):
                with torch.cuda.stream(self.streams[k % len(self.streams)]):
                    m.forward(t, dt, update_geometry=True)
                    g = m.mg[0]
                    m.backward(t, dt, dJdu=g.state.u.data, dJdv=g.state.v.data, dJdH=None)
        finally:
            _lib.forward_persistent_enable(True)
            _lib.adjoint_persistent_enable(True)
        for s in self.streams:
            caller.wait_stream(s)

    @property
    def cells(self) -> int:
        return sum(p.n * p.n for p in self.problems)


__all__ = ["GlacierProblem", "valley_glacier", "ensemble_sizes", "levels_for", "install_glacier", "Ensemble"]
