#!/usr/bin/env python3
"""Benchmark of the coupled SSA step on one MI355X.

Workload: the dome on a Gaussian-random-field bed at 512 x 512 cells,
dx = 1500 m, 5 multigrid levels, dt = 10 a (glide_amd.synthetic), coupled to
the enthalpy solve on its 512 x 512 columns with 41 sigma levels (surface
temperature -30 C + 6.5e-3 K/m (2000 m - s), geothermal flux 0.05 W/m^2,
strain heating from the SSA velocities, frictional heating, rate factor fed
back to the momentum solver).  One step is a forward FAS-CD solve of exactly
3 V-cycles (advancing the geometry), the thermal step (column Newton sweeps
with the device-side stopping test), the adjoint FAS solve of exactly 3
V-cycles for the loss 0.5 * |u|^2 + 0.5 * |v|^2, and the four parameter
gradients.  Inputs live in HBM before timing starts.

Prints one JSON line.  Besides the contract fields it carries
  roofline     - the dominant kernel's algorithmic bytes / measured duration
                 (HIP events around each launch over a second run of the same
                 number of steps, launched eagerly) vs HBM peak
  cpu_baseline - the FP64 oracle on the host cores for a bounded sample of the
                 same workload
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "forward+adjoint steps/sec at 512² MAC grid, mixed vs FP64 at equal 1e-8 error"
HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFS = 78.6  # MI355X vector FP64
WORD = 8  # bytes per FP64 value

# Algorithmic bytes per cell of one launch (each field counted once at 8 B):
#   forward smoothing pass (patch solve + increment apply): 12 inputs + 4 outputs
#   transpose operator: 15 inputs + 3 outputs
#   adjoint smoothing pass: 12 inputs + 3 outputs
#   residual: 12 inputs + 3 outputs
#   enthalpy column / layer smoother, per column of nz nodes: E, E_prev, f_E,
#   omega in and E out (5 nz) + five 2-D fields (u, v, H, H_prev, E_surface);
#   enthalpy residual + max-norm: 4 nz + 5 per column (set in main from nz)
ALG_FIELDS = {"vanka_sweep": 16, "vjp": 18, "adjoint_sweep": 15, "residual": 15}
SEC_PER_YR = 365.25 * 86400.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--vcycles", type=int, default=3)
    ap.add_argument("--dt", type=float, default=10.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--nz", type=int, default=41, help="sigma levels of the enthalpy columns (0: no thermal step)")
    ap.add_argument("--n-smooth", type=int, default=10, help="maximum enthalpy sweeps per step")
    ap.add_argument("--enth-adjoint-rtol", type=float, default=1e-8,
                    help="stopping test of the enthalpy adjoint sweeps (relative to the first residual)")
    ap.add_argument("--serial-enthalpy", dest="concurrent", action="store_false",
                    help="run the enthalpy step and its adjoint on the momentum stream (no overlap)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", dest="mixed", action="store_false",
                    help="skip the mixed-precision run (FP32 patch Jacobian) and its error against FP64")
    ap.add_argument("--cpu-sample-n", type=int, default=128,
                    help="grid side of the oracle sample (same level count and cycle counts)")
    return ap.parse_args()


def build_model(args, torch):
    from glide_amd import IceDynamics, ThermalModel
    from glide_amd.synthetic import dome_problem, install, thermal_forcing
    p = dome_problem(args.n, 1500.0, seed=args.seed)
    model = IceDynamics(ny=args.n, nx=args.n, dx=1500.0, n_levels=args.levels, device="cuda")
    install(model.mg, p)
    for s in (model.forward_solver, model.adjoint_solver):
        s.fas_options.set(maximum_vcycles=args.vcycles, relative_tolerance=0.0, absolute_tolerance=0.0,
                          report_norms=False, asynchronous=True)
    thermal = None
    if args.nz > 0:
        thermal = ThermalModel(model.mg[0], nz=args.nz, n_smooth=args.n_smooth, strain_heating=True)
        Ts, qg = thermal_forcing(p)
        thermal.initialize(torch.as_tensor(Ts, device="cuda"), Q_geo=torch.as_tensor(qg, device="cuda"))
    return model, thermal


def run_timed(stepper, args, torch):
    """Warm-up steps, then the timed steps on the production path (fixed-cycle
    solves replayed from the library's cached HIP graphs, enthalpy work on its
    own stream); returns (t, seconds)."""
    t = 0.0
    for _ in range(args.warmup):
        stepper(t, args.dt, concurrent=args.concurrent)
        t += args.dt
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stepper(t, args.dt, concurrent=args.concurrent)
        t += args.dt
    torch.cuda.synchronize()
    return t, time.perf_counter() - t0


def snapshot(model, thermal):
    g = model.mg[0]
    out = {"u": g.state.u.data, "v": g.state.v.data, "H": g.state.H.data, "lambda_u": g.adjoint.lambda_u.data,
           "lambda_v": g.adjoint.lambda_v.data, "lambda_H": g.adjoint.lambda_H.data, "grad_beta": g.sliding.beta.grad}
    if thermal is not None:
        out["E"] = thermal.ops.enthalpy_state.E
    return {k: v.detach().clone() for k, v in out.items()}


def rel_l2(a, b):
    den = float(b.norm())
    return float((a - b).norm()) / den if den > 0 else float((a - b).norm())


def read_traffic(args, kernel):
    """HBM bytes per finest-level launch of the dominant pass from the committed
    FETCH_SIZE / WRITE_SIZE summary (copy-kernel calibrated); None when the
    summary does not cover this workload."""
    path = ROOT / "profiles" / "vanka_pass_512_hbm_traffic.md"
    if kernel != "vanka_sweep" or args.n != 512 or not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def read_fp64_flops(args):
    """FP64 operations per finest-level forward patch-solve launch from the
    committed VALU instruction-count summary (64 lanes x (add + mul + trans +
    2 fma) per wave instruction, masked lanes included); None when absent."""
    path = ROOT / "profiles" / "vanka_cell_512_fp64_valu.md"
    if args.n != 512 or not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("fp64_flops_per_launch:"):
            return float(line.split(":")[1])
    return None


class Stepper:
    """One coupled step: momentum + mass forward solve, enthalpy step and its
    adjoint, momentum adjoint + gradients, then the rate-factor feedback.

    The enthalpy step and its adjoint depend only on the forward solution, and
    the momentum adjoint only on the forward state and the rate factor the
    forward used, so with ``concurrent`` the enthalpy work runs on a second
    stream beside the momentum adjoint; the new rate factor is written to the
    grid after both.  The step runs on a stream of its own rather than the
    null stream, whose implicit synchronisation with blocking streams could
    serialise the two."""

    def __init__(self, model, thermal, torch, adj_rtol=1e-8):
        self.model, self.thermal, self.torch = model, thermal, torch
        self.adj_rtol = adj_rtol
        dev = model.mg[0].state.u.data.device
        self.main = torch.cuda.Stream(device=dev)
        if thermal is not None:
            g = model.mg[0]
            self.stream = torch.cuda.Stream(device=dev)
            self.B_next = torch.empty_like(g.rheology.B.data)
            self.dL_dB = torch.ones_like(g.rheology.B.data)
            self.ev_fwd, self.ev_thermal = torch.cuda.Event(), torch.cuda.Event()
        self.adjoint_sweeps = None

    def _thermal(self, dt, report):
        th = self.thermal
        th.step(dt * SEC_PER_YR, B_out=self.B_next)
        # objective 0.5 |E|^2 + sum(B): the seeds only set the right-hand side;
        # a report synchronises the host, so only the sequential steps ask for one
        g = self.model.mg[0]
        res = th.ops.adjoint_solve(dt * SEC_PER_YR, dL_dE=th.ops.enthalpy_state.E, dL_dB=self.dL_dB,
                                   B_scale=th.B_scale, report=report, relative_tolerance=self.adj_rtol,
                                   n_iter=64, u2d=g.state.u.data, v2d=g.state.v.data, vel_div=SEC_PER_YR)
        if report:
            self.adjoint_sweeps = res["report"].sweeps

    def __call__(self, t, dt, concurrent=True, report=False):
        caller = self.torch.cuda.current_stream()
        self.main.wait_stream(caller)
        with self.torch.cuda.stream(self.main):
            self._step(t, dt, concurrent, report)
        caller.wait_stream(self.main)

    def _step(self, t, dt, concurrent, report):
        torch, model, thermal = self.torch, self.model, self.thermal
        if thermal is not None:
            thermal.pre_momentum()
        model.forward(t, dt, update_geometry=True)
        g = model.mg[0]
        main = torch.cuda.current_stream()
        if thermal is not None:
            if concurrent:
                self.ev_fwd.record(main)
                self.stream.wait_event(self.ev_fwd)
                with torch.cuda.stream(self.stream):
                    self._thermal(dt, report)
                    self.ev_thermal.record(self.stream)
            else:
                self._thermal(dt, report)
        model.backward(t, dt, dJdu=g.state.u.data, dJdv=g.state.v.data, dJdH=None)
        if thermal is not None:
            if concurrent:
                main.wait_event(self.ev_thermal)
            g.rheology.B.data.copy_(self.B_next)


def cpu_baseline(args):
    """FP64 oracle on the host: one forward and one adjoint V-cycle of the same
    dome workload at cpu_sample_n^2 with the same level count, scaled to a full
    step (x vcycles) and to 512^2 by the cell ratio."""
    import numpy as np
    from glide_amd.synthetic import dome_problem
    from oracle import oracle as O
    n = args.cpu_sample_n
    p = dome_problem(n, 1500.0, seed=args.seed)
    om = O.OracleMG(n, n, 1500.0, args.levels)
    for k in ("bed", "B", "beta", "smb"):
        om.set(k, getattr(p, k))
    om.set("H", p.H)
    om.set("H_prev", p.H)
    cfg = O.OracleMG.fas_config(vcycles=1, rtol=0.0, atol=0.0)
    t0 = time.perf_counter()
    om.fascd_solve(args.dt, cfg)
    t_fwd = time.perf_counter() - t0
    om.put("af_u", -om.get("u"))
    om.put("af_v", -om.get("v"))
    om.put("af_H", np.zeros((n, n)))
    t0 = time.perf_counter()
    om.fas_adjoint_solve(args.dt, cfg)
    om.gradients(args.dt)
    t_adj = time.perf_counter() - t0
    t_th = t_tha = 0.0
    if args.nz > 0:
        from glide_amd.synthetic import thermal_forcing
        Ts, qg = thermal_forcing(p)
        th = O.OracleThermal(n, n, args.nz, 1500.0, p.H)
        th.initialize_from_temperature(Ts)
        th.set_surface_enthalpy_from_temperature(Ts)
        th.Q_geo = qg
        B = om.get("B")
        u, v = om.get("u"), om.get("v")
        phi = np.repeat(O.strain_heating(u, v, B, 3.0, 1e-6, 1500.0)[:, :, None], args.nz, axis=2)
        t0 = time.perf_counter()
        O.thermal_step(th, u, v, om.get("H"), p.beta, args.dt * SEC_PER_YR, n_smooth=args.n_smooth,
                       strain_heating=phi)
        t_th = time.perf_counter() - t0
        # the enthalpy adjoint with the GPU step's seeds (dL/dE = E, dL/dB = 1):
        # the oracle solves J^T lambda = g directly (sparse LU), whose fill-in
        # grows faster than the cell count, so it is timed on a na^2 sample and
        # scaled by the cell ratio (which favours the CPU)
        na = min(n, 24)
        pa = dome_problem(na, 1500.0, seed=args.seed)
        Tsa, qga = thermal_forcing(pa)
        tha = O.OracleThermal(na, na, args.nz, 1500.0, pa.H)
        tha.initialize_from_temperature(Tsa)
        tha.set_surface_enthalpy_from_temperature(Tsa)
        tha.Q_geo = qga
        O.thermal_step(tha, np.zeros((na, na + 1)), np.zeros((na + 1, na)), pa.H, pa.beta, args.dt * SEC_PER_YR,
                       n_smooth=args.n_smooth)
        t0 = time.perf_counter()
        bs = 917.0 * 9.81 * SEC_PER_YR ** (1.0 / 3.0)
        gE = tha.E + tha.arrhenius_vjp(np.ones((na, na)), 3.0, bs)
        lam = tha.adjoint_solve(args.dt * SEC_PER_YR, gE)
        tha.adjoint_gradients(args.dt * SEC_PER_YR, lam)
        t_tha = (time.perf_counter() - t0) * (n / na) ** 2
    scale = (args.n / n) ** 2
    step_s = (args.vcycles * (t_fwd + t_adj) + t_th + t_tha) * scale
    return {"value": 1.0 / step_s, "unit": "steps/s", "cores": O.num_threads(), "kind": "port",
            "sample": (f"FP64 oracle, 1 forward + 1 adjoint V-cycle of the dome workload at {n}^2 "
                       f"({args.levels} levels): {t_fwd:.2f} s + {t_adj:.2f} s, x{args.vcycles} V-cycles per step; "
                       f"one thermal step on {n}^2 x {args.nz} columns: {t_th:.2f} s, its adjoint "
                       f"(sparse direct, 24^2 sample scaled to {n}^2): {t_tha:.2f} s; "
                       f"x{scale:g} cell ratio to {args.n}^2"),
            "sample_seconds": round(t_fwd + t_adj + t_th + t_tha, 3)}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from glide_amd import _lib
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm device", file=sys.stderr)
        sys.exit(1)
    _lib.load()
    model, thermal = build_model(args, torch)
    stepper = Stepper(model, thermal, torch, args.enth_adjoint_rtol)
    t, elapsed = run_timed(stepper, args, torch)
    fp64_state = snapshot(model, thermal)
    mixed = None
    if args.mixed:
        # the mixed form (FP32 patch Jacobian/LU in the forward smoother) on a
        # fresh copy of the workload, same steps; error against the FP64 state
        m2, th2 = build_model(args, torch)
        m2.forward_solver.vanka_options.set(local_fp32=True)
        _, el2 = run_timed(Stepper(m2, th2, torch, args.enth_adjoint_rtol), args, torch)
        errs = {k: rel_l2(a, b) for (k, a), b in zip(snapshot(m2, th2).items(), fp64_state.values())}
        mixed = {"value": round(args.steps / el2, 4), "ms_per_step": round(1000.0 * el2 / args.steps, 3),
                 "speedup_vs_fp64": round(elapsed / el2, 4), "rel_l2_err_vs_fp64": errs,
                 "form": "forward patch Jacobian and LU in FP32; residuals, stopping tests, updates, transfers, "
                         "adjoint and enthalpy in FP64"}
        del m2, th2
    # per-kernel durations: the same number of steps again with HIP events
    # around every hot launch (events inside graphs cannot be timed, so these
    # steps launch eagerly, and on one stream so that kernels do not share the
    # GPU while timed; kernel bodies are identical)
    _lib.timing_enable(True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        stepper(t, args.dt, concurrent=False, report=True)
        t += args.dt
    torch.cuda.synchronize()
    elapsed_eager = time.perf_counter() - t1
    kern = {}
    for name in ("vanka_cell", "vanka_apply", "adjoint_cell", "adjoint_apply", "vjp", "residual", "enth_column",
                 "enth_layer", "enth_residual"):
        kern[name] = _lib.timing_read(name)
    # finest-level launches only (args.n^2 cells each), read before disabling clears the totals
    fine = args.n * args.n
    fine_kern = {name: _lib.timing_read_units(name, fine)
                 for name in ("vanka_cell", "vanka_apply", "adjoint_cell", "adjoint_apply", "vjp")}
    _lib.timing_enable(False)

    ms_step = 1000.0 * elapsed / args.steps
    value = args.steps / elapsed

    def bw(fields, ms, cells):
        return fields * WORD * cells / (ms * 1e-3) / 1e9 if ms > 0 else 0.0

    # a smoothing pass is the patch-solve launch plus the increment-apply launch
    passes = {
        "vanka_sweep": (kern["vanka_cell"][0], kern["vanka_cell"][1] + kern["vanka_apply"][1], kern["vanka_cell"][2]),
        "adjoint_sweep": (kern["adjoint_cell"][0], kern["adjoint_cell"][1] + kern["adjoint_apply"][1],
                          kern["adjoint_cell"][2]),
        "vjp": kern["vjp"],
    }
    dominant = max(passes, key=lambda k: passes[k][1])
    # roofline of the dominant pass over the finest-level launches (args.n^2 cells each)
    pair = {"vanka_sweep": ("vanka_cell", "vanka_apply"), "adjoint_sweep": ("adjoint_cell", "adjoint_apply"),
            "vjp": ("vjp",)}[dominant]
    f_launches = fine_kern[pair[0]][0]
    f_ms = sum(fine_kern[k][1] for k in pair)
    avg_ms = f_ms / max(f_launches, 1)
    alg_bytes = ALG_FIELDS[dominant] * WORD * fine
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": read_traffic(args, dominant),
                "kernel": dominant, "level": f"{args.n}x{args.n}", "launches": f_launches,
                "avg_launch_us": round(1000.0 * avg_ms, 3), "alg_bytes_per_launch": alg_bytes,
                "all_levels_alg_GBps": round(bw(ALG_FIELDS[dominant], passes[dominant][1], passes[dominant][2]), 2)}
    # the patch solve's own bound: FP64 VALU issue (per-launch FP64 operations
    # from the committed counter summary over its measured finest-level duration)
    flops = read_fp64_flops(args)
    c_launches, c_ms = fine_kern["vanka_cell"]
    valu = None
    if flops is not None and c_launches > 0 and c_ms > 0:
        tf = flops / (c_ms / c_launches * 1e-3) / 1e12
        valu = {"bound": "fp64 valu", "kernel": "vanka_cell", "level": f"{args.n}x{args.n}",
                "achieved": round(tf, 2), "peak": FP64_VALU_PEAK_TFS, "unit": "TFLOP/s",
                "frac": round(tf / FP64_VALU_PEAK_TFS, 4), "fp64_flops_per_launch": flops,
                "avg_launch_us": round(1000.0 * c_ms / c_launches, 3)}
    roofline["valu"] = valu
    kernels = {k: {"launches": v[0], "total_ms": round(v[1], 3), "cells": v[2]} for k, v in kern.items()}
    for k, (c, ms, cells) in passes.items():
        kernels[k] = {"launches": c, "total_ms": round(ms, 3), "alg_GBps": round(bw(ALG_FIELDS[k], ms, cells), 2)}
    if args.nz > 0:
        ALG_FIELDS.update({"enth_column": 5 + 5 / args.nz, "enth_layer": 5 + 5 / args.nz,
                           "enth_residual": 4 + 5 / args.nz})
    for k in ("enth_column", "enth_layer", "enth_residual") if args.nz > 0 else ():
        c, ms, units = kern[k]
        kernels[k]["alg_GBps_incl_skipped"] = round(bw(ALG_FIELDS[k], ms, units), 2)

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"error": repr(e)}

    line = {"metric": METRIC, "value": round(value, 4), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": (f"dome on GRF bed, {args.n}x{args.n} MAC grid, dx=1500 m, {args.levels} levels, "
                                    f"{args.vcycles}+{args.vcycles} V-cycles forward+adjoint per step, dt={args.dt} a"
                                    + (f", coupled enthalpy on {args.n}x{args.n}x{args.nz} columns "
                                       f"(<= {args.n_smooth} sweeps, strain + frictional heating)"
                                       if args.nz > 0 else "")),
                       "sigma_levels": args.nz,
                       "precision": "plain FP64 form", "grid": f"{args.n}x{args.n}", "levels": args.levels,
                       "vcycles": args.vcycles, "dt_years": args.dt, "parallelism": "single GPU"},
            "roofline": roofline, "cpu_baseline": cpu, "mixed": mixed, "kernels": kernels,
            "ms_per_step_eager_timed": round(1000.0 * elapsed_eager / args.steps, 3),
            "enthalpy_adjoint_sweeps": stepper.adjoint_sweeps,
            "omp_threads": os.environ.get("OMP_NUM_THREADS")}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
