"""An ensemble of independent glacier domains of uneven sizes stepped
together on one GPU (no counterpart in the reference, which runs one domain
per process: glide/model.py:25-39, glide/multigrid.py:212-235).

Each domain is an ``IceDynamics`` of its own size and level count.  A step
runs every domain's forward and adjoint solve with a fixed number of
iterations and no host wait (``asynchronous``).  Batched
(``Ensemble(batched=True)``, the mixed form): every stage of the step is one
launch over a device table of the domains taking part in it
(glide_ensemble_step: the V-cycles level by level, all coarsest-level passes
together, the whole step one HIP graph), bitwise the per-domain results;
:func:`autotune_tiles` picks the red-black phases' workgroup shapes from timed
steps.  Per domain (the default): each domain's solves are replayed from the
library's cached HIP graphs (one per domain and solve), round-robin over a
few streams, so that the small domains' latency-bound launches overlap; the
persistent coarse-level kernels of the FP64 form assume their workgroups are
resident together, which concurrent streams do not guarantee, so that path
turns them off for its steps (the red-black passes of the mixed form have
none).

:func:`valley_glacier` is the seeded mountain-glacier workload: a valley bed
b = z0 - slope x + 150 m (2 y / W - 1)^2 + 30 m x GRF, initial thickness 0.3 of
the valley depth, SMB linear in elevation about a seeded equilibrium-line
altitude.
"""
from __future__ import annotations

import ctypes
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
    fp64: reference-form V-cycles).

    batched (mixed only): every stage of the step is one launch over all the
    domains taking part in it (glide_ensemble_step: domain tables, the
    V-cycles level by level, the whole step replayed as one HIP graph);
    per domain bitwise the results of its own solves.  Otherwise each
    domain's solves are replayed from their own cached graphs, round-robin
    over `streams` streams."""

    def __init__(self, problems: Sequence[GlacierProblem], precision: str = "mixed", iterations: int = 3,
                 streams: int = 8, device: str = "cuda", batched: bool = False):
        self.problems = list(problems)
        self.precision = precision
        self.iterations = int(iterations)
        self.batched = bool(batched)
        if self.batched and precision != "mixed":
            raise ValueError("the batched ensemble runs the mixed form (precision='mixed')")
        self._handle = None
        self.tiles = None
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
        self.device = dev
        if self.batched:
            self.streams = []
            self._create_batched()
        else:
            self.streams = [torch.cuda.Stream(device=dev) for _ in range(max(1, int(streams)))]
            # two solves per domain, plus the graphs of other callers
            _lib.graphs_capacity(max(8, 2 * len(self.models) + 8))

    def _create_batched(self) -> None:
        """Level descriptors of every domain (FP64 with the forward and adjoint
        buffers and the gradients, and the FP32 mirror) for glide_ensemble_create."""
        from .multigrid import _mixed_hierarchy
        n = len(self.models)
        nl = (ctypes.c_int32 * n)()
        p64 = (ctypes.POINTER(_lib.Level) * n)()
        p32 = (ctypes.POINTER(_lib.Level) * n)()
        self._descriptors = []
        for d, m in enumerate(self.models):
            fs, ads = m.forward_solver, m.adjoint_solver
            top = m.mg[0]
            for t in (top.sliding.beta, top.geometry.bed, top.state.H_prev, top.forcing.smb):
                t.grad  # allocate the gradient buffers
            L = len(fs.levels)
            a64 = (_lib.Level * L)()
            for k in range(L):
                g = fs.levels[k].grid
                g.forward_operators
                g.adjoint_operators
                a64[k] = g.level_struct(fas_scratch=fs.levels[k].scratch, adj_scratch=ads.levels[k].scratch)
            a32 = _mixed_hierarchy(m.mg).structs(0, adjoint=True)
            self._descriptors.append((a64, a32))
            nl[d] = L
            p64[d] = ctypes.cast(a64, ctypes.POINTER(_lib.Level))
            p32[d] = ctypes.cast(a32, ctypes.POINTER(_lib.Level))
        fo, ao = self.models[0].forward_solver.mixed_options, self.models[0].adjoint_solver.mixed_options
        fas, afas, vcf, vca, mc = fo.fas_c(), ao.fas_c(), fo.vanka_c(False), ao.vanka_c(True), fo.to_c()
        ac = ao.to_c()
        if (ac.inner_vcycles, ac.fp64_coarsest) != (mc.inner_vcycles, mc.fp64_coarsest):
            raise ValueError("the batched ensemble needs the same inner_vcycles and fp64_coarsest in both solves")
        mc.max_outer = self.iterations
        mc.relative_tolerance = 0.0
        mc.absolute_tolerance = 0.0
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.call("glide_ensemble_create", n, nl, p64, p32, ctypes.byref(fas), ctypes.byref(afas), ctypes.byref(vcf),
                      ctypes.byref(vca), ctypes.byref(mc), 15, ctypes.byref(h))
        self._handle = h

    def set_tiles(self, rows: Sequence[int]) -> None:
        """Rows of cell pairs per block of the batched red-black phases: one
        entry per level (the most levels of any domain) and one for the
        coarsest-level passes (glide_ensemble_set_tiles).  The results do not
        depend on them."""
        if not self._handle:
            raise ValueError("tiles apply to the batched ensemble")
        r = (ctypes.c_int32 * len(rows))(*[int(x) for x in rows])
        _lib.call("glide_ensemble_set_tiles", self._handle, r, len(rows))
        self.tiles = [int(x) for x in rows]

    @property
    def levels(self) -> int:
        return int(_lib.load().glide_ensemble_levels(self._handle)) if self._handle else 0

    @property
    def launches_per_step(self) -> int:
        """Launches of one batched step (recorded at the first step)."""
        return int(_lib.load().glide_ensemble_launches(self._handle)) if self._handle else 0

    def close(self) -> None:
        if self._handle:
            _lib.load().glide_ensemble_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, t: float, dt: float) -> None:
        if self.batched:
            with torch.cuda.device(self.device):
                _lib.call("glide_ensemble_step", self._handle, float(dt), _lib.stream_handle(self.device))
            return
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


TILE_CANDIDATES = (1, 2, 4, 8)


def autotune_tiles(problems: Sequence[GlacierProblem], iterations: int = 3, dt: float = 1.0, reps: int = 3,
                   candidates: Sequence[int] = TILE_CANDIDATES, cache: str | None = None, device: str = "cuda"):
    """Tile shapes of the batched ensemble's red-black phases (rows of cell
    pairs per block, per level and for the coarsest-level passes), chosen by
    coordinate search over timed steps of a scratch ensemble of the same
    domains: each entry in turn takes the candidate with the shortest step
    (`reps` graph replays after one warm-up step), the others held at their
    current choice.  The results do not depend on the tiles, so the search
    only measures.  With `cache` (a JSON file) the choice is stored under the
    device name, the domain sizes and the iteration count, and reused.
    Returns (tiles, {candidate tiles as text: ms per step})."""
    import json
    import os
    sizes = [p.n for p in problems]
    key = f"{torch.cuda.get_device_name(torch.device(device))}|{sizes}|{int(iterations)}"
    table = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
        if key in table:
            return list(table[key]["tiles"]), dict(table[key]["timings"])
    ens = Ensemble(problems, precision="mixed", iterations=iterations, device=device, batched=True)
    tiles = [4] * (ens.levels + 1)
    timings = {}

    def timed(rows):
        k = ",".join(str(x) for x in rows)
        if k not in timings:
            ens.set_tiles(rows)
            ens.step(0.0, dt)  # records the tables and captures the graph
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                ens.step(0.0, dt)
            e1.record()
            torch.cuda.synchronize()
            timings[k] = e0.elapsed_time(e1) / reps
        return timings[k]

    for slot in range(len(tiles)):
        best = None
        for c in candidates:
            trial = list(tiles)
            trial[slot] = int(c)
            ms = timed(trial)
            if best is None or ms < best[0]:
                best = (ms, int(c))
        tiles[slot] = best[1]
    ens.close()
    if cache:
        table[key] = {"tiles": tiles, "timings": timings}
        with open(cache, "w") as f:
            json.dump(table, f, indent=1)
    return tiles, timings


__all__ = ["GlacierProblem", "valley_glacier", "ensemble_sizes", "levels_for", "install_glacier", "Ensemble",
           "autotune_tiles", "TILE_CANDIDATES"]
