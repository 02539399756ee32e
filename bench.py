#!/usr/bin/env python3
"""Benchmark of the coupled SSA step on one MI355X: the mixed-precision form
against the plain FP64 form at equal error.

Workload: the dome on a Gaussian-random-field bed at 512 x 512 cells,
dx = 1500 m, 5 multigrid levels, dt = 10 a (glide_amd.synthetic), coupled to
the enthalpy solve on its 512 x 512 columns with 41 sigma levels (surface
temperature -30 C + 6.5e-3 K/m (2000 m - s), geothermal flux 0.05 W/m^2,
strain and frictional heating, rate factor fed back to the momentum solver).

One step: the forward FAS-CD solve (advancing the geometry), the thermal step
and the adjoint of the converged coupled step for the loss
0.5 |u|^2 + 0.5 |v|^2 + 0.5 |E|^2 + sum(B_new): the enthalpy adjoint, the
thermal step's gradients with respect to u, v, H, H_prev and beta
(glide_thermal_step_vjp), the momentum adjoint seeded with the velocity terms
plus those, and the parameter gradients (the basal-drag gradient includes the
frictional-heating path).  Inputs live in HBM before timing starts.  The two
forms:

* fp64  - the reference's algorithm in FP64 (additive Vanka smoother with its
          default counts) for the forward and the adjoint solves;
* mixed - glide_amd.mixed: FP32 multigrid with red-black smoothing inside
          FP64 outer iterations (FP64 residual, coarsest level and
          accumulation): defect correction then Anderson-mixed FP64
          red-black V-cycles for the forward solve, flexible GMRES
          preconditioned by the FP32 V-cycles for the adjoint.

Both forms share the FP64 enthalpy solve with the reference's alternating
sweep order: near the pressure-melting point the discrete enthalpy system can
have more than one solution, and another sweep order (the red-black column
option of glide_amd.enthalpy) can settle on a different one.  Equal error: before timing, each
form's stopping tolerances (forward, enthalpy, adjoint) are set to the loosest
values of a ladder for which steps from the warmed-up trajectory land within
--target / 2 (relative L2) of the truth on two steps (the last warm-up step
and the next one): the same step with FP64 solves driven to
the residual floor (forward: FP64 red-black V-cycles to 3e-13; adjoint: FP64
reference-form V-cycles to 1e-14; enthalpy sweeps to 1e-13; enthalpy adjoint
sweeps to 1e-14, at most 250).  After
timing, every timed step is replayed from the saved state (the solves are
deterministic) and its u, v, H, E, multipliers and basal-drag gradient are
compared with the truth of that step; the line reports the largest errors.
If a replayed step misses the target, the stages feeding the quantities over
it are tightened by one notch of their ladders and the timing is run again
(at most four timed runs; `timed_runs` counts them), so the reported rate is
the one of the tolerances the replay passed.

Prints one JSON line.  Besides the contract fields it carries
  fp64         - the FP64 form's rate, tolerances and errors at the same target
  roofline     - the dominant kernel of the mixed step: algorithmic bytes per
                 launch / its mean duration (HIP events around each launch in
                 an eager replay of the timed steps) vs HBM peak
  cpu_baseline - the FP64 oracle on the host cores: one forward and one
                 adjoint V-cycle at the full 512 x 512 size, scaled by the
                 V-cycle counts the FP64 form needed, plus one thermal step
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "forward+adjoint steps/sec at 512² MAC grid, mixed vs FP64 at equal 1e-8 error"
HBM_PEAK_GBS = 8000.0
SEC_PER_YR = 365.25 * 86400.0
FORWARD_LADDER = (1e-8, 3e-9, 1e-9, 3e-10, 1e-10, 3e-11, 1e-11, 3e-12, 1e-12, 3e-13)
ENTHALPY_LADDER = (1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-10, 1e-11, 1e-12)
ADJOINT_LADDER = FORWARD_LADDER
ENTHALPY_SWEEPS = 100  # cap of the column sweeps per thermal step (the tolerance stops them)
TRUTH_FORWARD = 3e-13  # the truth's forward and enthalpy stopping tests (near the FP64 floors)
TRUTH_ENTHALPY = 1e-13
TRUTH_ADJOINT = 1e-14
CALIBRATION_MARGIN = 0.5  # the calibration steps must land within this fraction of the target
VERIFY_ROUNDS = 4  # timed runs at most, each tightening the stages whose quantities missed the target

# Algorithmic bytes of one launch per cell of its level, in values of the
# launch's precision (each plane counted once):
#   red-black colour phase: reads u, v, H, the frozen viscosity, phi, beta,
#     bed, f_u, f_v (every cell), f_H, gamma and the active-set flag (the
#     solving half); writes u, v, H (every cell) and the flag (the solving
#     half): 10.5 + 3.5
#   red-black adjoint colour phase: the transpose operator reads the three
#     multiplier planes, the three right-hand sides, the cached viscosity and
#     u, v, H, phi, beta, bed and the flag; the solving half reads its 25 LU
#     values; writes the three multiplier planes: 16 + 12.5 + 3
#   FP64 residual: 12 inputs + 3 outputs
ALG_VALUES = {"vanka_mc32": 14.0, "vanka_mc64": 14.0, "adjoint_mc32": 31.5, "adjoint_mc64": 31.5,
              "residual": 15.0, "vjp": 18.0}
# bytes per value; at the finest level the transpose operator runs in FP64
# (the mixed adjoint's Krylov products and residuals)
WORD = {"vanka_mc32": 4, "vanka_mc64": 8, "adjoint_mc32": 4, "adjoint_mc64": 8, "residual": 8, "vjp": 8}
TIMED = ("vanka_mc32", "vanka_mc64", "adjoint_mc32", "adjoint_mc64", "residual", "vanka_cell", "vanka_apply",
         "adjoint_cell", "adjoint_apply", "vjp", "adjoint_factor", "enth_column", "enth_layer", "enth_residual")


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--dt", type=float, default=10.0)
    ap.add_argument("--dx", type=float, default=1500.0, help="cell size (m)")
    ap.add_argument("--grounding-line", action="store_true",
                    help="beta = 0 for floating ice where the bed lies below -300 m (the 2048^2 workload)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--nz", type=int, default=41, help="sigma levels of the enthalpy columns (0: no thermal step)")
    ap.add_argument("--target", type=float, default=1e-8, help="relative L2 error bound against the truth")
    ap.add_argument("--no-fp64", dest="fp64", action="store_false", help="skip the FP64 form")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--verbose", action="store_true", help="calibration progress on stderr")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only run the mixed step's solves a few times for counter collection (no JSON line)")
    ap.add_argument("--inversion", action="store_true",
                    help="the multi-step adjoint-inversion workload instead (own JSON line, see inversion())")
    ap.add_argument("--inv-steps", type=int, default=40)
    ap.add_argument("--inv-checkpoint", type=int, default=8)
    ap.add_argument("--inv-dt", type=float, default=5.0)
    ap.add_argument("--inv-directions", type=int, default=8)
    ap.add_argument("--inv-grounded", action="store_true",
                    help="raise the bed so that all ice is grounded (grounded fraction at its clamp)")
    ap.add_argument("--ensemble", action="store_true",
                    help="the batched-ensemble workload instead (own JSON line, see ensemble())")
    ap.add_argument("--ens-domains", type=int, default=96)
    ap.add_argument("--ens-streams", type=int, default=8)
    ap.add_argument("--ens-tiles", default="",
                    help="red-black tiles of the batched ensemble (comma-separated rows per block); empty: autotune")
    ap.add_argument("--ens-forms", default="batched,mixed,fp64",
                    help="comma-separated ensemble forms to time (batched, mixed, fp64)")
    return ap.parse_args()


def log(args, msg):
    if args.verbose:
        print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ the forms
class Form:
    """One model of the workload with its solve settings."""

    def __init__(self, args, torch, kind):
        from glide_amd import IceDynamics, ThermalModel
        from glide_amd.synthetic import dome_problem, install, thermal_forcing
        self.args, self.torch, self.kind = args, torch, kind
        p = dome_problem(args.n, args.dx, seed=args.seed, grounding_line=args.grounding_line)
        prec = "mixed" if kind == "mixed" else "fp64"
        self.model = IceDynamics(ny=args.n, nx=args.n, dx=args.dx, n_levels=args.levels, device="cuda",
                                 precision=prec)
        install(self.model.mg, p)
        self.thermal = None
        if args.nz > 0:
            self.thermal = ThermalModel(self.model.mg[0], nz=args.nz, n_smooth=10, strain_heating=True)
            Ts, qg = thermal_forcing(p)
            self.thermal.initialize(torch.as_tensor(Ts, device="cuda"), Q_geo=torch.as_tensor(qg, device="cuda"))

        fs, ad = self.model.forward_solver, self.model.adjoint_solver
        for s in (fs, ad):
            s.fas_options.set(maximum_vcycles=64, absolute_tolerance=0.0, report_norms=False)
            s.mixed_options.max_outer = 64
            s.mixed_options.absolute_tolerance = 0.0
        if kind == "truth":
            self.set_tolerances(TRUTH_FORWARD, TRUTH_ENTHALPY, TRUTH_ADJOINT)
        else:
            self.set_tolerances(1e-8, 1e-6, 1e-8)
        from glide_amd.coupled import CoupledStep
        truth = kind == "truth"
        self.stepper = CoupledStep(self.model, self.thermal, enthalpy_adjoint_rtol=TRUTH_ADJOINT if truth else 1e-10,
                                   enthalpy_adjoint_iters=250 if truth else 64,
                                   forward=(lambda t, dt: self._truth_forward(dt)) if truth else None)

    def set_tolerances(self, fwd, enth, adj):
        self.tol = {"forward": fwd, "enthalpy": enth, "adjoint": adj}
        fs, ad = self.model.forward_solver, self.model.adjoint_solver
        fs.fas_options.set(relative_tolerance=fwd)
        fs.mixed_options.relative_tolerance = fwd
        ad.fas_options.set(relative_tolerance=adj)
        ad.mixed_options.relative_tolerance = adj
        if self.thermal is not None:
            sc = self.thermal.ops.smoother_config
            sc.relative_tolerance = enth
            sc.absolute_tolerance = 0.0
            self.thermal.n_smooth = ENTHALPY_SWEEPS

    def _truth_forward(self, dt):
        """FP64 to the residual floor: reference-form V-cycles (robust from any
        state) to 1e-6, then FP64 red-black V-cycles (the fastest FP64 route
        to the floor) to TRUTH_FORWARD of the first residual."""
        fs = self.model.forward_solver
        fs.vanka_options.set(multicolour=False, newton_tol=1e-6)
        fs.vanka_options.newton_options.set(steps=30, relaxation=0.5, ssa_damping=0.01, mc_damping=1.0)
        fs.fas_options.set(relative_tolerance=1e-6, absolute_tolerance=0.0, coarsest_steps=200, pre_steps=10,
                           post_steps=20, finest_steps=50, compat_vprolong=True)
        rep = fs.solve(dt)
        fs.vanka_options.set(multicolour=True, newton_tol=0.0)
        fs.vanka_options.newton_options.set(steps=3, relaxation=0.5, ssa_damping=0.0, mc_damping=0.0)
        fs.fas_options.set(relative_tolerance=0.0, absolute_tolerance=TRUTH_FORWARD * rep.initial_norm,
                           coarsest_steps=40,
                           pre_steps=4, post_steps=4, finest_steps=8, compat_vprolong=False)
        fs.solve(dt)
        top = self.model.mg[0]
        top.state.H_prev.data.copy_(top.state.H.data)

    def step(self, t, dt):
        """One coupled step and its adjoint (glide_amd.coupled.CoupledStep:
        the enthalpy adjoint seeded with dL/dE = E and dL/dB = 1, then the
        thermal step's input gradients into the momentum adjoint's seeds)."""
        self.stepper(t, dt)

    def tensors(self):
        g = self.model.mg[0]
        d = {"u": g.state.u.data, "v": g.state.v.data, "H": g.state.H.data, "H_prev": g.state.H_prev.data,
             "mask": g.state.mask.data, "phi": g.state.phi.data, "B": g.rheology.B.data,
             "lambda_u": g.adjoint.lambda_u.data, "lambda_v": g.adjoint.lambda_v.data,
             "lambda_H": g.adjoint.lambda_H.data}
        if self.thermal is not None:
            ops = self.thermal.ops
            d.update(E=ops.enthalpy_state.E, E_prev=ops.enthalpy_state.E_prev, H_prev_thermal=ops.H_prev)
        return d

    def snapshot(self):
        return {k: v.detach().clone() for k, v in self.tensors().items()}

    def restore(self, snap):
        for k, v in self.tensors().items():
            v.copy_(snap[k])

    def results(self):
        """What the error report compares (multipliers of the rows the
        transpose operator does not determine are zeroed by the caller)."""
        g = self.model.mg[0]
        out = {"u": g.state.u.data, "v": g.state.v.data, "H": g.state.H.data, "lambda_u": g.adjoint.lambda_u.data,
               "lambda_v": g.adjoint.lambda_v.data, "lambda_H": g.adjoint.lambda_H.data,
               "grad_beta": g.sliding.beta.grad, "mask": g.state.mask.data}
        if self.thermal is not None:
            out["E"] = self.thermal.ops.enthalpy_state.E
        return {k: v.detach().clone() for k, v in out.items()}


def rel_l2(a, b):
    den = float(b.norm())
    return float((a - b).norm()) / den if den > 0 else float((a - b).norm())


def errors(res, truth):
    """Relative L2 errors against the truth; the multipliers compared on the
    rows the transpose operator determines (not the domain-edge facets nor the
    active-set cells of the truth)."""
    e = {k: rel_l2(res[k], truth[k]) for k in ("u", "v", "H", "grad_beta") + (("E",) if "E" in truth else ())}
    keep_H = truth["mask"] <= 0.5

    def pinned(d):
        lu, lv, lH = d["lambda_u"].clone(), d["lambda_v"].clone(), d["lambda_H"].clone()
        lu[:, 0] = 0.0
        lu[:, -1] = 0.0
        lv[0, :] = 0.0
        lv[-1, :] = 0.0
        lH = lH * keep_H
        return lu, lv, lH

    for k, a, b in zip(("lambda_u", "lambda_v", "lambda_H"), pinned(res), pinned(truth)):
        e[k] = rel_l2(a, b)
    return e


FORWARD_KEYS = ("u", "v", "H")
ADJOINT_KEYS = ("lambda_u", "lambda_v", "lambda_H", "grad_beta")


def calibrate(form, truth_form, snaps, args):
    """The loosest tolerances of the ladders for which one step from each
    (state, time) of `snaps` (the last warm-up states) is within
    CALIBRATION_MARGIN x args.target of its truth in every compared quantity
    (the multipliers and gradients inherit the forward solution's error, so
    the forward tolerance is chosen against all of them): first the forward
    tolerance with tight enthalpy and adjoint solves, then the adjoint
    tolerance, then the enthalpy tolerance."""
    truths = []
    for snap, t in snaps:
        truth_form.restore(snap)
        truth_form.step(t, args.dt)
        truths.append(truth_form.results())
    tight = ENTHALPY_LADDER[-1], ADJOINT_LADDER[-1]
    bound = CALIBRATION_MARGIN * args.target

    def trial(fwd, enth, adj, limits):
        """Per compared quantity, the largest error over the calibration
        steps (stops early once one exceeds its limit)."""
        form.set_tolerances(fwd, enth, adj)
        worst = {}
        for (snap, t), T in zip(snaps, truths):
            form.restore(snap)
            form.step(t, args.dt)
            e = errors(form.results(), T)
            for k, v in e.items():
                worst[k] = max(worst.get(k, 0.0), v)
            log(args, f"  {form.kind} t={t:g} forward {fwd:.0e} enthalpy {enth:.0e} adjoint {adj:.0e}: "
                      + " ".join(f"{k} {v:.1e}" for k, v in e.items()))
            if any(worst[k] > limits[k] for k in worst):
                break
        return worst

    def loosest(ladder, make, limits):
        """The first ladder value within the limits, else the one with the
        smallest largest error (and its errors)."""
        best, best_e = ladder[-1], None
        for v in ladder:
            e = trial(*make(v), limits)
            if all(e[k] <= limits[k] for k in e):
                return v, e, True
            if best_e is None or max(e.values()) < max(best_e.values()):
                best, best_e = v, e
        return best, best_e, False

    keys = FORWARD_KEYS + ADJOINT_KEYS + (("E",) if form.thermal is not None else ())
    fwd, e_fwd, ok = loosest(FORWARD_LADDER, lambda v: (v, tight[0], tight[1]), {k: bound for k in keys})
    # when no forward tolerance meets the bound, the later stages may not add
    # more than 10% to the error each quantity reached at the best forward
    # tolerance (nor exceed the bound where that error is below it)
    limits = {k: max(bound, 1.1 * e_fwd.get(k, 0.0)) for k in keys}
    adj, _, _ = loosest(ADJOINT_LADDER, lambda v: (fwd, tight[0], v), limits)
    enth = tight[0]
    if form.thermal is not None:
        enth, _, _ = loosest(ENTHALPY_LADDER, lambda v: (fwd, v, adj), limits)
    form.set_tolerances(fwd, enth, adj)
    return {"forward": fwd, "enthalpy": enth, "adjoint": adj}, ok


def run_form(form, truth_form, args, torch):
    """Warm-up, calibration, timed steps, then the replayed error check."""
    t = 0.0
    snaps = []
    for _ in range(args.warmup):
        snaps.append((form.snapshot(), t))
        form.step(t, args.dt)
        t += args.dt
    torch.cuda.synchronize()
    snap = form.snapshot()
    # calibration on the last warm-up step and the first step after it
    cal = (snaps[-1:] if snaps else []) + [(snap, t)]
    tols, calibrated = calibrate(form, truth_form, cal, args)
    form.restore(snap)
    log(args, f"{form.kind}: tolerances {tols} (met on the calibration steps: {calibrated})")
    for rnd in range(VERIFY_ROUNDS):
        form.set_tolerances(tols["forward"], tols["enthalpy"], tols["adjoint"])
        # one untimed step with the chosen tolerances (graphs of its solves recorded)
        form.restore(snap)
        form.step(t, args.dt)
        form.restore(snap)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tt = t
        for _ in range(args.steps):
            form.step(tt, args.dt)
            tt += args.dt
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        stats = _solver_stats(form)
        # replay: the same steps from the same state, each against its truth
        # (a truth depends only on the state it starts from)
        form.restore(snap)
        worst = {}
        tt = t
        for _ in range(args.steps):
            before = form.snapshot()
            form.step(tt, args.dt)
            res = form.results()
            truth_form.restore(before)
            truth_form.step(tt, args.dt)
            e = errors(res, truth_form.results())
            for q, v in e.items():
                worst[q] = max(worst.get(q, 0.0), v)
            tt += args.dt
        torch.cuda.synchronize()
        over = [q for q, v in worst.items() if not v <= args.target]
        log(args, f"{form.kind}: round {rnd}, tolerances {tols}, {1e3 * elapsed / args.steps:.1f} ms/step, largest "
                  f"errors " + " ".join(f"{q} {v:.1e}" for q, v in worst.items()))
        if not over or rnd + 1 == VERIFY_ROUNDS:
            break
        # a timed step missed the target: tighten the stages that feed the
        # quantities over it by one notch of their ladders and time again
        if any(q in FORWARD_KEYS or q in ADJOINT_KEYS for q in over):
            tols["forward"] = _tighter(FORWARD_LADDER, tols["forward"])
        if any(q in ADJOINT_KEYS for q in over):
            tols["adjoint"] = _tighter(ADJOINT_LADDER, tols["adjoint"])
        if "E" in over:
            tols["enthalpy"] = _tighter(ENTHALPY_LADDER, tols["enthalpy"])
    return {"elapsed": elapsed, "tolerances": tols, "calibrated": calibrated, "errors": worst, "snap": snap,
            "t": t, "stats": stats, "verify_rounds": rnd + 1}


def _tighter(ladder, v):
    """The next tighter value of a ladder (the tightest stays)."""
    for x in ladder:
        if x < v * (1 - 1e-9):
            return x
    return ladder[-1]


def _solver_stats(form):
    fs, ad = form.model.forward_solver, form.model.adjoint_solver
    out = {}
    for name, s in (("forward", fs), ("adjoint", ad)):
        rep = getattr(s, "last_report", None)
        if rep is not None:
            out[name + "_iterations"] = int(rep.vcycles)
        if getattr(s, "last_stats", None):
            out[name + "_phases"] = dict(s.last_stats)
    return out


# -------------------------------------------------------------- measurement
def kernel_times(form, res, args, torch, _lib):
    """Replay the timed steps eagerly with HIP events around every hot launch
    (events inside graphs cannot be timed; kernel bodies are identical), the
    enthalpy work in turn rather than beside the momentum adjoint so that no
    timed kernel shares the GPU."""
    form.restore(res["snap"])
    torch.cuda.synchronize()
    form.stepper.concurrent = False
    _lib.timing_enable(True)
    tt = res["t"]
    for _ in range(args.steps):
        form.step(tt, args.dt)
        tt += args.dt
    torch.cuda.synchronize()
    form.stepper.concurrent = True
    fine = args.n * args.n
    tot = {k: _lib.timing_read(k) for k in TIMED}
    fine_t = {k: _lib.timing_read_units(k, fine) for k in TIMED}
    _lib.timing_enable(False)
    return tot, fine_t


def roofline_of(tot, fine_t, args):
    cands = [k for k in ALG_VALUES if tot[k][0] > 0]
    dominant = max(cands, key=lambda k: tot[k][1])
    launches, ms = fine_t[dominant]
    avg_ms = ms / max(launches, 1)
    alg = ALG_VALUES[dominant] * WORD[dominant] * args.n * args.n
    achieved = alg / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": read_traffic(dominant, args),
            "kernel": dominant, "level": f"{args.n}x{args.n}", "launches": launches,
            "avg_launch_us": round(1000.0 * avg_ms, 3), "alg_bytes_per_launch": int(alg)}


def read_traffic(kernel, args):
    """HBM bytes per finest-level launch of `kernel` from the committed
    FETCH_SIZE / WRITE_SIZE summary (copy-kernel calibrated), if it covers
    this grid."""
    path = ROOT / "profiles" / f"mixed_step_{args.n}_hbm_traffic.md"
    if not path.exists():
        return None
    tag = f"traffic_bytes_per_launch[{kernel}]:"
    for line in path.read_text().splitlines():
        if line.startswith(tag):
            return int(line[len(tag):])
    return None


def hbm_fractions(fine_t, args):
    """Algorithmic HBM rate of every kernel with a byte model, finest level."""
    out = {}
    for k in ALG_VALUES:
        launches, ms = fine_t[k]
        if launches > 0 and ms > 0:
            gbs = ALG_VALUES[k] * WORD[k] * args.n * args.n / (ms / launches * 1e-3) / 1e9
            out[k] = {"launches": launches, "avg_us": round(1000.0 * ms / launches, 3), "alg_GBps": round(gbs, 1),
                      "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
            traffic = read_traffic(k, args)
            if traffic is not None:  # measured bytes over the same mean duration
                out[k]["traffic_GBps"] = round(traffic / (ms / launches * 1e-3) / 1e9, 1)
    return out


def cpu_info():
    model = None
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"os_cpu_count": os.cpu_count(), "lscpu_model": model, "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(args, cycles):
    """The FP64 oracle on the host cores at the full size: one forward and one
    adjoint V-cycle (the reference's counts) and one thermal step, scaled by
    the V-cycle counts the FP64 form needed per step on the GPU."""
    import numpy as np
    from glide_amd.synthetic import dome_problem, thermal_forcing
    from oracle import oracle as O
    n = args.n
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
    t_th = 0.0
    if args.nz > 0:
        Ts, qg = thermal_forcing(p)
        th = O.OracleThermal(n, n, args.nz, 1500.0, p.H)
        th.initialize_from_temperature(Ts)
        th.set_surface_enthalpy_from_temperature(Ts)
        th.Q_geo = qg
        u, v = om.get("u"), om.get("v")
        phi = np.repeat(O.strain_heating(u, v, om.get("B"), 3.0, 1e-6, 1500.0)[:, :, None], args.nz, axis=2)
        t0 = time.perf_counter()
        O.thermal_step(th, u, v, om.get("H"), p.beta, args.dt * SEC_PER_YR, n_smooth=10, strain_heating=phi)
        t_th = time.perf_counter() - t0
    nf, na = cycles
    step_s = nf * t_fwd + na * t_adj + t_th
    info = cpu_info()
    return {"value": round(1.0 / step_s, 6), "unit": "steps/s", "cores": O.num_threads(), "kind": "port",
            "sample": (f"FP64 oracle at {n}x{n} ({args.levels} levels, reference counts): one forward V-cycle "
                       f"{t_fwd:.2f} s, one adjoint V-cycle + gradients {t_adj:.2f} s, one thermal step on "
                       f"{n}x{n}x{args.nz} columns {t_th:.2f} s; a step = {nf} forward + {na} adjoint V-cycles "
                       f"(the FP64 form's counts at equal error on the GPU) + the thermal step; the enthalpy "
                       f"adjoint is not timed on the host"),
            "sample_seconds": round(t_fwd + t_adj + t_th, 2), **info}


def inversion(args, torch):
    """--inversion: the adjoint-inversion workload.  512 x 512 dome, 5 levels,
    --inv-steps implicit steps of --inv-dt years; observations u_obs, v_obs at
    every step from a forward run with a true log(beta) = log(1e3 Pa a/m) +
    1.5 x a seeded Gaussian random field of 30 km correlation, plus Gaussian
    noise of 5 m/a; loss = sum over steps of 0.5 |u - u_obs|^2 +
    0.5 |v - v_obs|^2 plus 1e-3 |grad log beta|^2, from log(beta) =
    log(1e3).  The gradient with respect to log(beta) comes from the native
    adjoint solves chained by torch with a checkpoint every --inv-checkpoint
    steps (glide_amd.inversion), in the mixed form; it is Taylor-tested in
    --inv-directions seeded random directions (e = 1e-2 ... 1e-6) and compared
    with the gradient of the plain FP64 form.  Prints one JSON line."""
    import numpy as np
    from glide_amd import IceDynamics
    from glide_amd.inversion import gradient_penalty, misfit, rollout
    from glide_amd.synthetic import dome_problem, gaussian_random_field, install
    n, dt, steps, ck = args.n, args.inv_dt, args.inv_steps, args.inv_checkpoint
    p = dome_problem(n, 1500.0, seed=args.seed)
    if args.inv_grounded:  # bed + 0.917 H >= 250 m: the adjoint's fixed grounded fraction is exact
        p.bed = np.maximum(p.bed, 0.0) + 250.0
    rng = np.random.default_rng(args.seed + 101)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=torch.float64)
    true_lb = T(np.log(1e3) + 1.5 * gaussian_random_field(rng, n, n, 1500.0, 30e3))
    lb0 = torch.full_like(true_lb, float(np.log(1e3)))
    H0, bed, smb = T(p.H), T(p.bed), T(p.smb)

    def make(precision):
        m = IceDynamics(ny=n, nx=n, dx=1500.0, n_levels=args.levels, device="cuda", precision=precision)
        install(m.mg, p)
        for s in (m.forward_solver, m.adjoint_solver):
            s.fas_options.set(maximum_vcycles=64, relative_tolerance=1e-12, absolute_tolerance=0.0,
                              report_norms=False)
            s.mixed_options.relative_tolerance = 1e-12
            s.mixed_options.max_outer = 64
        return m

    model = make("mixed")
    obs = []

    def record(k, u, v, H):
        obs.append((u.detach().clone(), v.detach().clone()))
        return u.new_zeros(())

    with torch.no_grad():
        rollout(model, H0, bed, torch.exp(true_lb), smb, steps, dt, record, checkpoint=None)
    gen = torch.Generator(device="cuda").manual_seed(args.seed + 7)
    obs = [(u + 5.0 * torch.randn(u.shape, generator=gen, device="cuda", dtype=u.dtype),
            v + 5.0 * torch.randn(v.shape, generator=gen, device="cuda", dtype=v.dtype)) for u, v in obs]
    print(f"inversion: observations of {steps} steps recorded", flush=True)

    def J(mdl, lb, checkpoint=ck):
        loss, _ = rollout(mdl, H0, bed, torch.exp(lb), smb, steps, dt, lambda k, u, v, H: misfit(u, v, *obs[k]),
                          checkpoint=checkpoint)
        return loss + gradient_penalty(lb, 1e-3)

    def gradient(mdl):
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        base = torch.cuda.memory_allocated()
        t0 = time.perf_counter()
        m = lb0.clone().requires_grad_()
        val = J(mdl, m)
        val.backward()
        torch.cuda.synchronize()
        return float(val), m.grad.clone(), time.perf_counter() - t0, torch.cuda.max_memory_allocated() - base

    J0, g, t_grad, peak = gradient(model)
    print(f"inversion: mixed gradient in {t_grad:.1f} s, J = {J0:.6e}", flush=True)
    # cells whose grounded fraction is off its +-20 clamp (varies with H)
    g0 = model.mg[0]
    z = float(g0.geometry.flotation_reg_driving.value) * (g0.geometry.bed.data + 0.917 * g0.state.H.data)
    partial = int(((z > -20.0) & (z < 20.0)).sum())
    taylor, orders = [], []
    eps = [1e-2, 1e-3, 1e-4, 1e-5, 1e-6]
    floor = 1e-15 * abs(J0)  # below it the remainders are rounding of the loss sums
    dgen = torch.Generator(device="cuda").manual_seed(args.seed + 11)
    for q in range(args.inv_directions):
        d = torch.randn(lb0.shape, generator=dgen, device="cuda", dtype=lb0.dtype)
        dJ = float((g * d).sum())
        rem = []
        with torch.no_grad():
            for e in eps:
                rem.append(abs(float(J(model, lb0 + e * d, checkpoint=None)) - J0 - e * dJ))
        for k in range(len(eps) - 1):
            if rem[k + 1] > 10 * floor:
                orders.append(round(float(np.log10(rem[k] / rem[k + 1])), 3))
        taylor.append({"directional_derivative": dJ, "remainders": [float(f"{r:.3e}") for r in rem]})
        print(f"inversion: direction {q}: " + " ".join(f"{r:.2e}" for r in rem), flush=True)
    del model
    ref = make("fp64")
    J64, g64, t64, _ = gradient(ref)
    print(f"inversion: FP64 gradient in {t64:.1f} s", flush=True)
    line = {"workload": (f"adjoint inversion, dome on GRF bed{' raised to grounded ice' if args.inv_grounded else ''}, "
                         f"{n}x{n}, {args.levels} levels, {steps} steps of {dt:g} a, checkpoint every {ck} steps, "
                         "d loss / d log(beta)"),
            "loss": J0, "gradient_seconds_mixed": round(t_grad, 3), "gradient_seconds_fp64": round(t64, 3),
            "peak_extra_hbm_bytes_mixed": int(peak), "loss_fp64": J64,
            "gradient_rel_l2_mixed_vs_fp64": float(f"{rel_l2(g, g64):.3e}"),
            "taylor_eps": eps, "taylor": taylor, "taylor_orders_above_floor": orders,
            "taylor_floor": floor, "partly_grounded_cells": partial}
    print(json.dumps(line), flush=True)


def ensemble(args, torch):
    """--ensemble: the batched-ensemble workload.  --ens-domains valley
    glaciers (glide_amd.ensemble.valley_glacier, dx = 100 m) with side lengths
    drawn log-uniformly from 48 to 384 cells (seeded, multiples of 8), each
    taking one forward + adjoint step (3 iterations each: FP32 red-black
    multigrid with FP64 defect correction, no host wait) per ensemble step:
    batched (one launch per stage over a table of the domains taking part,
    V-cycles level by level, one graph per step, red-black tiles from
    autotune_tiles), then per domain over --ens-streams streams (each
    domain's solves from its own graphs) and the FP64 form (3 reference-form
    V-cycles); --warmup steps, then --steps timed.  Prints one JSON line with
    the summed steps/s of each and whether the batched results equal the
    per-domain ones bitwise."""
    from glide_amd.ensemble import Ensemble, autotune_tiles, ensemble_sizes, valley_glacier
    sizes = ensemble_sizes(args.ens_domains, seed=args.seed)
    probs = [valley_glacier(n, seed=args.seed + 1000 + k) for k, n in enumerate(sizes)]
    out = {"workload": (f"{len(probs)} valley glaciers, sides {min(sizes)}-{max(sizes)} cells "
                        f"(log-uniform, seeded), dx = 100 m, one forward + adjoint step each per ensemble step, "
                        "3 iterations per solve"),
           "domains": len(probs), "cells": int(sum(n * n for n in sizes)), "streams": args.ens_streams}
    forms = [f for f in ("batched", "mixed", "fp64") if f in args.ens_forms.split(",")]
    if args.ens_tiles:
        tiles = [int(x) for x in args.ens_tiles.split(",")]
        out["tiles"] = tiles
    else:
        t0 = time.perf_counter()
        tiles, timings = autotune_tiles(probs, iterations=3, dt=1.0)
        out["autotune"] = {"tiles": tiles, "seconds": round(time.perf_counter() - t0, 2),
                           "candidates_timed": len(timings),
                           "ms_default_tiles": round(timings[",".join(["4"] * len(tiles))], 2),
                           "ms_tuned_tiles": round(timings[",".join(str(x) for x in tiles)], 2)}
        print(f"ensemble autotune: {out['autotune']}", flush=True)
    final = {}
    for form in forms:
        prec = "fp64" if form == "fp64" else "mixed"
        ens = Ensemble(probs, precision=prec, iterations=3, streams=args.ens_streams, batched=form == "batched")
        if form == "batched":
            ens.set_tiles(tiles)
        t = 0.0
        for _ in range(args.warmup):
            ens.step(t, 1.0)
            t += 1.0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ens.step(t, 1.0)
            t += 1.0
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        finite = all(bool(torch.isfinite(m.mg[0].state.u.data).all()) for m in ens.models)
        out[form] = {"domain_steps_per_s": round(len(probs) * args.steps / el, 2),
                     "ensemble_step_ms": round(1000.0 * el / args.steps, 2), "finite": finite}
        if form == "batched":
            out[form]["launches_per_step"] = ens.launches_per_step
        if form != "fp64":
            final[form] = [torch.cat([m.mg[0].state.u.data.flatten(), m.mg[0].sliding.beta.grad.flatten()]).cpu()
                           for m in ens.models]
        print(f"ensemble {form}: {out[form]}", flush=True)
        if form == "batched":
            ens.close()
        del ens
    if "batched" in final and "mixed" in final:
        # the same steps from the same start: the batched step equals the per-domain solves bitwise
        out["batched"]["bitwise_equal_to_per_domain"] = all(
            torch.equal(a, b) for a, b in zip(final["batched"], final["mixed"]))
        out["batched"]["speedup_vs_per_domain"] = round(
            out["batched"]["domain_steps_per_s"] / out["mixed"]["domain_steps_per_s"], 2)
    print(json.dumps(out), flush=True)


PROFILE_TOLERANCES = (3e-10, 1e-8, 3e-10)  # forward, enthalpy, adjoint: the verified values at 512^2


def profile_kernels(args, torch):
    """--profile-kernels: a short run of the mixed step for profiler and
    counter passes: --warmup steps, then --steps coupled steps with the
    tolerances the equal-error calibration picks at 512^2 (no truth, no
    calibration, no FP64 form), then three element-wise copies of 2^25
    doubles (256 MiB read, 256 MiB written) that calibrate the byte
    counters.  Prints one line per step."""
    form = Form(args, torch, "mixed")
    form.set_tolerances(*PROFILE_TOLERANCES)
    t = 0.0
    for q in range(args.warmup + args.steps):
        form.step(t, args.dt)
        torch.cuda.synchronize()
        st = _solver_stats(form)
        print(f"step {q} ({'warm-up' if q < args.warmup else 'profiled'}): forward {st.get('forward_iterations')} "
              f"and adjoint {st.get('adjoint_iterations')} iterations", flush=True)
        t += args.dt
    src = torch.rand(1 << 25, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for q in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print("calibration copies: 3 x 256 MiB", flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); there is no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from glide_amd import _lib
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm device", file=sys.stderr)
        sys.exit(1)
    _lib.load()
    if args.profile_kernels:
        profile_kernels(args, torch)
        return
    if args.inversion:
        inversion(args, torch)
        return
    if args.ensemble:
        ensemble(args, torch)
        return
    truth = Form(args, torch, "truth")
    mixed = Form(args, torch, "mixed")
    res_m = run_form(mixed, truth, args, torch)
    tot, fine_t = kernel_times(mixed, res_m, args, torch, _lib)
    roofline = roofline_of(tot, fine_t, args)
    kernels_m = {k: {"launches": v[0], "total_ms": round(v[1], 3)} for k, v in tot.items() if v[0] > 0}
    fracs = hbm_fractions(fine_t, args)
    del mixed

    fp64 = None
    cycles = (0, 0)
    if args.fp64:
        f64 = Form(args, torch, "fp64")
        res_d = run_form(f64, truth, args, torch)
        st = res_d["stats"]
        cycles = (st.get("forward_iterations", 0), st.get("adjoint_iterations", 0))
        tot_d, fine_d = kernel_times(f64, res_d, args, torch, _lib)
        fp64 = {"value": round(args.steps / res_d["elapsed"], 4),
                "ms_per_step": round(1000.0 * res_d["elapsed"] / args.steps, 3),
                "tolerances": res_d["tolerances"], "target_met_on_calibration": res_d["calibrated"],
                "target_met": all(v <= args.target for v in res_d["errors"].values()),
                "timed_runs": res_d["verify_rounds"],
                "rel_l2_err_vs_truth": {k: float(f"{v:.3e}") for k, v in res_d["errors"].items()},
                "last_step": st,
                "kernels": {k: {"launches": v[0], "total_ms": round(v[1], 3)} for k, v in tot_d.items() if v[0] > 0},
                "form": "reference algorithm in FP64: additive Vanka (30 local Newton steps), counts 200/10/20/50"}
        del f64

    cpu = None
    if not args.no_cpu_baseline and args.fp64:
        try:
            cpu = cpu_baseline(args, cycles)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"error": repr(e)}

    value = args.steps / res_m["elapsed"]
    errs_m = {k: float(f"{v:.3e}") for k, v in res_m["errors"].items()}
    line = {"metric": METRIC, "value": round(value, 4), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(1000.0 * res_m["elapsed"] / args.steps, 3),
            "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 multigrid, f64 residual/accumulation/coarsest", "data": "synthetic",
            "config": {"workload": (f"dome on GRF bed, {args.n}x{args.n} MAC grid, dx={args.dx:g} m, {args.levels} levels, "
                                    + ("grounding line (beta = 0 where b < -300 m and the ice floats), " if args.grounding_line else "")
                                    + f"dt={args.dt} a, forward + adjoint solved to {args.target:g} relative L2 error"
                                    + (f", coupled enthalpy on {args.n}x{args.n}x{args.nz} columns "
                                       "(strain + frictional heating, rate-factor feedback)" if args.nz > 0 else "")),
                       "grid": f"{args.n}x{args.n}", "levels": args.levels, "sigma_levels": args.nz,
                       "dt_years": args.dt, "target_rel_l2": args.target, "parallelism": "single GPU"},
            "rel_l2_err_vs_truth": errs_m, "target_met": all(v <= args.target for v in res_m["errors"].values()),
            "tolerances": res_m["tolerances"], "timed_runs": res_m["verify_rounds"], "last_step": res_m["stats"],
            "speedup_vs_fp64": (round(value / fp64["value"], 3) if fp64 else None),
            "fp64": fp64, "roofline": roofline, "hbm_fractions": fracs, "cpu_baseline": cpu,
            "kernels": kernels_m,
            "truth": (f"FP64 solves to the residual floor from the same state: forward FP64 red-black V-cycles to "
                      f"{TRUTH_FORWARD:g}, adjoint FP64 reference-form V-cycles to {TRUTH_ADJOINT:g}, enthalpy "
                      f"sweeps to {TRUTH_ENTHALPY:g}, enthalpy adjoint sweeps to {TRUTH_ADJOINT:g} (relative "
                      f"residuals); the adjoint is the coupled step's (enthalpy adjoint, thermal input gradients "
                      f"into the momentum adjoint's seeds)")}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
